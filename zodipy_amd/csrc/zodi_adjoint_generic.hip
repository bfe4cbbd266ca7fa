// Adjoint of custom component layouts (gfx950, FP64).
//
// Models registered with any combination of the component kinds
// (zodipy/zodiacal_light_model.py:116-128, zodipy/number_density.py:408-420) that
// no registry layout kernel covers. For each ray, every group's components are
// taken one per pass over the group's nodes: the component's kind is dispatched
// once per pass, outside the node loop, to the same compile-time-typed code the
// registry kernels run for a one-component chunk (adjoint_group), so the
// partials, the accumulation order within a component and the Jacobian columns
// are those of the registry kernels. Node geometry, B(T) and dB/dT are formed
// once per component and node instead of once per chunk.

#include <hip/hip_runtime.h>

#include <utility>

#include "zodi_adjoint.hpp"

namespace zodi {

template <int INTEGRAND, int SPLIT, bool JAC, int KIND>
__device__ __forceinline__ void generic_component(const Ctx& cx, const AdjArgs& a, const Ray& y,
                                                  int g, int c, double gscale, double* row,
                                                  int lane, int sub, double (&sh)[6],
                                                  double& total) {
  adjoint_group<INTEGRAND, SPLIT, JAC, 1, 0, 0u, false>(Grp<KIND>{},
                                                        std::integer_sequence<int, 0>{}, cx, a, y,
                                                        g, c, gscale, row, lane, sub, sh, total,
                                                        nullptr);
}

// Every component of every group for one ray, in model order, then the
// shared-parameter sums (adjoint_ray_tail).
template <int INTEGRAND, int SPLIT, bool JAC>
__device__ __forceinline__ void generic_adjoint_ray(const Ctx& cx, const AdjArgs& a,
                                                    const Ray& y, double gscale, double* row,
                                                    int lane, int sub) {
  const ModelPtr M = cx.M;
  double sh[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double total = 0.0;
  const int n_groups = M->n_groups;
  for (int g = 0; g < n_groups; ++g) {
    const int first = M->group_first[g];
    const int cnt = M->group_count[g];
    for (int j = 0; j < cnt; ++j) {
      const int c = first + j;
      switch (M->comp_kind[c]) {
#define ZODI_GENERIC_CASE(K)                                                                   \
  case K:                                                                                      \
    generic_component<INTEGRAND, SPLIT, JAC, K>(cx, a, y, g, c, gscale, row, lane, sub, sh, \
                                                total);                                        \
    break;
        ZODI_GENERIC_CASE(ZODI_CLOUD)
        ZODI_GENERIC_CASE(ZODI_BAND)
        ZODI_GENERIC_CASE(ZODI_RING)
        ZODI_GENERIC_CASE(ZODI_FEATURE)
        ZODI_GENERIC_CASE(ZODI_FAN)
        ZODI_GENERIC_CASE(ZODI_COMET)
        ZODI_GENERIC_CASE(ZODI_INTERSTELLAR)
        ZODI_GENERIC_CASE(ZODI_NARROW_BAND)
        ZODI_GENERIC_CASE(ZODI_BROAD_BAND)
        ZODI_GENERIC_CASE(ZODI_RING_RRM)
        ZODI_GENERIC_CASE(ZODI_FEATURE_RRM)
#undef ZODI_GENERIC_CASE
        default:
          break;  // the host validates the kinds
      }
    }
  }
  adjoint_ray_tail<INTEGRAND, JAC>(cx, a, y, sh, total, row, lane);
}

template <int INTEGRAND, int SPLIT, bool JAC>
__global__ __launch_bounds__(kAdjBlock) __attribute__((amdgpu_waves_per_eu(2, 2))) void
adjoint_generic(AdjArgs a) {
  adjoint_body<SPLIT, JAC, 0>(
      a, [&](const Ctx& cx, const Ray& y, double gscale, double* row, int lane, int sub,
             double*) { generic_adjoint_ray<INTEGRAND, SPLIT, JAC>(cx, a, y, gscale, row, lane, sub); });
}

namespace {

template <int INTEGRAND>
AdjKernel pick_generic_split(int split, bool jac) {
  if (jac) return split == 1 ? adjoint_generic<INTEGRAND, 1, true> : nullptr;
  switch (split) {
    case 1:
      return adjoint_generic<INTEGRAND, 1, false>;
    case 2:
      return adjoint_generic<INTEGRAND, 2, false>;
    case 4:
      return adjoint_generic<INTEGRAND, 4, false>;
    default:
      return nullptr;
  }
}

}  // namespace

AdjKernel generic_adjoint_kernel(int integrand, int split, bool jac) {
  if (integrand == ZODI_INTEGRAND_KELSALL) return pick_generic_split<ZODI_INTEGRAND_KELSALL>(split, jac);
  if (integrand == ZODI_INTEGRAND_RRM) return pick_generic_split<ZODI_INTEGRAND_RRM>(split, jac);
  return nullptr;
}

}  // namespace zodi
