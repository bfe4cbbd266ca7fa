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
// One ray per lane (or SPLIT lanes sharing its nodes); a group's components are
// evaluated through the kind switch (any layout). Each node adds its two shares
// to the wave's row in LDS with wavefront-scope atomic adds (lanes of a wave may
// meet in one entry); the block adds its wave rows in wave order and a reduction
// kernel adds the blocks in block order.

#include <hip/hip_runtime.h>

#include "zodi_adjoint.hpp"

namespace zodi {

namespace {

constexpr int kMaxGroupSlots = 4;  // components per group (packing.MAX_SLOTS_PER_GROUP)

__device__ __forceinline__ void lds_add(double* p, double x) {
  __hip_atomic_fetch_add(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

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
    double cw[kMaxGroupSlots];
    FeatureDir fd[kMaxGroupSlots];
#pragma unroll
    for (int j = 0; j < kMaxGroupSlots; ++j) {
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
      const double v = wk * S;
      if (jt < n_last) {
        lds_add(row + jt, v - v * f);
        lds_add(row + jt + 1, v * f);
      } else {
        lds_add(row + n_last, v);
      }
    }
  }
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

}  // namespace

AdjKernel table_adjoint_kernel(int integrand, int split) {
  if (integrand == ZODI_INTEGRAND_KELSALL) return pick_table_split<ZODI_INTEGRAND_KELSALL>(split);
  if (integrand == ZODI_INTEGRAND_RRM) return pick_table_split<ZODI_INTEGRAND_RRM>(split);
  return nullptr;
}

}  // namespace zodi
