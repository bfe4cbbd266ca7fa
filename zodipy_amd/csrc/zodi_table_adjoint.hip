// Gradient of the line-of-sight integral with respect to the Planck table of a
// model block (gfx950, FP64).
//
// The reference tabulates the bandpass-weighted blackbody emission B_k = B(T_k)
// over 100 temperatures (zodipy/blackbody.py:33-49) and interpolates it at each
// node's temperature (zodipy/brightness.py:48, np.interp). On the table's index
// axis u = (T - T_0)/dT, clamped to [0, n - 1], a node with j = floor(u) < n - 1
// and f = u - j takes B = (1 - f) B_j + f B_{j+1}; past the last entry it takes
// B_{n-1}. The integral of component c is h sum_k w_k s_c n_c(X_k) B(T(X_k)) plus
// terms without B (scattering), s_c = (1 - albedo) emissivity (Kelsall) or the
// calibration (RRM), so for weights g
//
//   d/dB_j sum_r sum_c g_{c,r} I_{c,r} = sum over rays, groups and nodes of
//       h w_k (sum_c g_{c,r} s_c n_c) times (1 - f) for j, f for j + 1.
//
// A bandpass fit needs this and the effective-parameter gradient of the other
// adjoint kernels: both are linear in the normalised bandpass weights.
//
// One ray per lane (or SPLIT lanes sharing its nodes). The registry layouts run
// with their kinds fixed at compile time (table_layout), any other layout through
// the kind switch (table_adjoint). A lane sums the shares of consecutive nodes in
// one table interval in registers and adds them to the wave's row in LDS with
// wavefront-scope atomic adds when the interval changes (lanes of a wave may meet
// in one entry); the block adds its wave rows in wave order and a reduction
// kernel adds the blocks in block order.

#include <hip/hip_runtime.h>

#include <utility>

#include "zodi_adjoint.hpp"

namespace zodi {

namespace {

__device__ __forceinline__ void lds_add(double* p, double x) {
  __hip_atomic_fetch_add(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// A lane's pending shares of consecutive nodes in one table interval: entry j
// gets s0, entry j + 1 gets s1 (none past the last entry). Nodes along a ray
// change temperature slowly, so most of them only add to registers; the LDS
// atomics run when the interval changes and at the end of the group.
struct TablePending {
  int j = -1;
  double s0 = 0.0, s1 = 0.0;

  __device__ __forceinline__ void flush(double* row, int n_last) {
    if (j >= 0) {
      lds_add(row + j, s0);
      if (j < n_last) lds_add(row + j + 1, s1);
    }
  }
  // One node at interval jt, fraction f, weight v.
  __device__ __forceinline__ void add(double* row, int n_last, int jt, double f, double v) {
    const int e = jt < n_last ? jt : n_last;
    const double c0 = jt < n_last ? v - v * f : v;
    const double c1 = jt < n_last ? v * f : 0.0;
    if (e != j) {
      flush(row, n_last);
      j = e;
      s0 = c0;
      s1 = c1;
    } else {
      s0 += c0;
      s1 += c1;
    }
  }
};

template <int INTEGRAND, int SPLIT>
__device__ __forceinline__ void table_ray(const Ctx& cx, const AdjArgs& a, const Ray& y,
                                          double gscale, double* row, int sub) {
  const ModelPtr M = cx.M;
  const int n_last = M->n_temps - 1;
  const int n_groups = M->n_groups;
  for (int g = 0; g < n_groups; ++g) {
    const int first = M->group_first[g];
    const int cnt = M->group_count[g];
    double h, mid;
    group_range(M, g, y, h, mid);
    const double us = M->group_T0[g] * cx.tb.inv_step;
    const double uo = cx.tb.T_first * cx.tb.inv_step;
    const double mdelta = -M->group_delta[g];
    double cw[kMaxSlots];
    FeatureDir fd[kMaxSlots];
#pragma unroll
    for (int j = 0; j < kMaxSlots; ++j) {
      cw[j] = 0.0;
      fd[j] = FeatureDir{1.0, 0.0};
      if (j < cnt) {
        const int c = first + j;
        const ParamPtr P = M->comp[c];
        const double s = INTEGRAND == ZODI_INTEGRAND_KELSALL ? P[ZODI_P_SCALE] : cx.calibration;
        double w = gscale;
        if (a.mode == ZODI_ADJ_VJP_COMPS) w = y.active ? a.weights[(int64_t)c * a.f.n_rays + y.r] : 0.0;
        cw[j] = (h * s) * w;
        if (is_feature(M->comp_kind[c])) fd[j] = feature_dir(P, y.ex, y.ey);
      }
    }
    const int K = y.active ? cx.K : 0;
    TablePending pend;
    for (int k = sub; k < K; k += SPLIT) {
      const ModelPtr Mk = opaque(M);
      const double xk = Mk->gl_x[k];
      const double wk = Mk->gl_w[k];
      const double s_los = h * xk + mid;
      const double X0 = s_los * y.ux + y.ox;
      const double X1 = s_los * y.uy + y.oy;
      const double X2 = s_los * y.uz + y.oz;
      const double R2 = X0 * X0 + X1 * X1 + X2 * X2;
      double R, inv_R;
      norm_and_inverse_node(R2, R, inv_R);
      const double u = fma(pow_radius(R, mdelta), us, -uo);
      const double uc = fmin(fmax(u, 0.0), (double)(ZODI_N_TEMPS - 1));
      const int jt = (int)uc;
      const double f = __builtin_amdgcn_fract(uc);
      double S = 0.0;
#pragma unroll 1
      for (int q = 0; q < cnt; ++q) {
        const int c = first + q;
        S += cw[q] * density(Mk->comp_kind[c], Mk->comp[c], Mk->comp_flags[c], X0, X1, X2, R,
                             inv_R, R2, fd[q]);
      }
      pend.add(row, n_last, jt, f, wk * S);
    }
    pend.flush(row, n_last);
  }
}

// Registry layouts: the group's kinds (and, with PROPS, the registry forms of the
// forward kernels) fixed at compile time, the densities inlined.
template <int INTEGRAND, int SPLIT, uint32_t PROPS, int... Ks, int... J>
__device__ __forceinline__ void table_group(Grp<Ks...>, std::integer_sequence<int, J...>,
                                            const Ctx& cx, const AdjArgs& a, const Ray& y, int g,
                                            int first, double gscale, double* row, int sub) {
  constexpr int C = sizeof...(Ks);
  constexpr int kinds[C] = {Ks...};
  const ModelPtr M = cx.M;
  const int n_last = M->n_temps - 1;
  double h, mid;
  group_range(M, g, y, h, mid);
  const double us = M->group_T0[g] * cx.tb.inv_step;
  const double uo = cx.tb.T_first * cx.tb.inv_step;
  const double mdelta = -M->group_delta[g];
  FeatureDir fdir[C];
  ((fdir[J] = is_feature(kinds[J]) ? feature_dir(M->comp[first + J], y.ex, y.ey)
                                   : FeatureDir{1.0, 0.0}),
   ...);
  double cw[C];
  (
      [&] {
        const int c = first + J;
        const double s =
            INTEGRAND == ZODI_INTEGRAND_KELSALL ? M->comp[c][ZODI_P_SCALE] : cx.calibration;
        double w = gscale;
        if (a.mode == ZODI_ADJ_VJP_COMPS)
          w = y.active ? a.weights[(int64_t)c * a.f.n_rays + y.r] : 0.0;
        cw[J] = (h * s) * w;
      }(),
      ...);
  const int K = y.active ? cx.K : 0;
  TablePending pend;
  for (int k = sub; k < K; k += SPLIT) {
    const ModelPtr Mk = opaque(M);
    const double xk = Mk->gl_x[k];
    const double wk = Mk->gl_w[k];
    const double s_los = h * xk + mid;
    Node nd;
    nd.X0 = s_los * y.ux + y.ox;
    nd.X1 = s_los * y.uy + y.oy;
    nd.X2 = s_los * y.uz + y.oz;
    nd.R2 = nd.X0 * nd.X0 + nd.X1 * nd.X1 + nd.X2 * nd.X2;
    norm_and_inverse_node(nd.R2, nd.R, nd.inv_R);
    nd.B = 0.0;
    nd.phi = 0.0;
    const double u = fma(pow_radius(nd.R, mdelta), us, -uo);
    const double uc = fmin(fmax(u, 0.0), (double)(ZODI_N_TEMPS - 1));
    const int jt = (int)uc;
    const double f = __builtin_amdgcn_fract(uc);
    double cut = 0.0;
    double S = 0.0;
    ((S += cw[J] * density_k<kinds[J], (PROPS >> (4 * J)) & 15, cut_mode<PROPS, J, Ks...>()>(
               Mk->comp[first + J], Mk->comp_flags[first + J], nd, fdir[J], cut)),
     ...);
    pend.add(row, n_last, jt, f, wk * S);
  }
  pend.flush(row, n_last);
}

template <int SPLIT, uint32_t PROPS, int INTEGRAND, class... Gs, int... G>
__device__ __forceinline__ void table_groups(Layout<INTEGRAND, Gs...>,
                                             std::integer_sequence<int, G...>, const Ctx& cx,
                                             const AdjArgs& a, const Ray& y, double gscale,
                                             double* row, int sub) {
  (table_group<INTEGRAND, SPLIT, props_shift(PROPS, group_first<Gs...>(G))>(
       Gs{}, std::make_integer_sequence<int, Gs::size>{}, cx, a, y, G, group_first<Gs...>(G),
       gscale, row, sub),
   ...);
}

template <int SPLIT, uint32_t PROPS, int INTEGRAND, class... Gs>
__device__ __forceinline__ void run_table(Layout<INTEGRAND, Gs...> lay, const Ctx& cx,
                                          const AdjArgs& a, const Ray& y, double gscale,
                                          double* row, int sub) {
  table_groups<SPLIT, PROPS>(lay, std::make_integer_sequence<int, sizeof...(Gs)>{}, cx, a, y,
                             gscale, row, sub);
}

// SPEC: the registry forms (RegistryProps, with the shared band cut-off when the
// host finds the block allows it: SHARE).
template <class LAYOUT, int SPLIT, bool SPEC, bool SHARE>
__global__ __launch_bounds__(kAdjBlock) __attribute__((amdgpu_waves_per_eu(4))) void
table_layout(AdjArgs a) {
  constexpr uint32_t PROPS =
      !SPEC ? 0u
            : (SHARE ? RegistryProps<LAYOUT>::value
                     : props_without_share(RegistryProps<LAYOUT>::value));
  adjoint_body<SPLIT, false, 0>(
      a, [&](const Ctx& cx, const Ray& y, double gscale, double* row, int, int sub, double*) {
        run_table<SPLIT, PROPS>(LAYOUT{}, cx, a, y, gscale, row, sub);
      });
}

template <int INTEGRAND, int SPLIT>
__global__ __launch_bounds__(kAdjBlock) __attribute__((amdgpu_waves_per_eu(2))) void
table_adjoint(AdjArgs a) {
  adjoint_body<SPLIT, false, 0>(
      a, [&](const Ctx& cx, const Ray& y, double gscale, double* row, int, int sub, double*) {
        table_ray<INTEGRAND, SPLIT>(cx, a, y, gscale, row, sub);
      });
}

template <int INTEGRAND>
AdjKernel pick_table_split(int split) {
  switch (split) {
    case 1:
      return table_adjoint<INTEGRAND, 1>;
    case 2:
      return table_adjoint<INTEGRAND, 2>;
    case 4:
      return table_adjoint<INTEGRAND, 4>;
    default:
      return nullptr;
  }
}

template <class LAYOUT, bool SPEC, bool SHARE>
AdjKernel pick_table_layout_split(int split) {
  switch (split) {
    case 1:
      return table_layout<LAYOUT, 1, SPEC, SHARE>;
    case 2:
      return table_layout<LAYOUT, 2, SPEC, SHARE>;
    case 4:
      return table_layout<LAYOUT, 4, SPEC, SHARE>;
    default:
      return nullptr;
  }
}

template <class LAYOUT>
AdjKernel pick_table_layout(const zodi_model* m, int split) {
  constexpr uint32_t kProps = RegistryProps<LAYOUT>::value;
  if constexpr (kProps != props_without_share(kProps)) {
    if (props_match(m, kProps)) return pick_table_layout_split<LAYOUT, true, true>(split);
  }
  if (props_match(m, props_without_share(kProps)))
    return pick_table_layout_split<LAYOUT, true, false>(split);
  return pick_table_layout_split<LAYOUT, false, false>(split);
}

}  // namespace

bool table_adjoint_is_layout(const zodi_model* m) { return match_layout(m) != kLayoutGeneric; }

AdjKernel table_adjoint_kernel(const zodi_model* m, int split) {
  switch (match_layout(m)) {
    case kLayoutDirbe:
      return pick_table_layout<LayoutDirbe>(m, split);
    case kLayoutKelsall4:
      return pick_table_layout<LayoutKelsall4>(m, split);
    case kLayoutRRM:
      return pick_table_layout<LayoutRRM>(m, split);
    case kLayoutRRMSplit:
      return pick_table_layout<LayoutRRMSplit>(m, split);
    default:
      break;
  }
  if (m->integrand == ZODI_INTEGRAND_KELSALL) return pick_table_split<ZODI_INTEGRAND_KELSALL>(split);
  if (m->integrand == ZODI_INTEGRAND_RRM) return pick_table_split<ZODI_INTEGRAND_RRM>(split);
  return nullptr;
}

}  // namespace zodi
