// Gradient of the line-of-sight integral with respect to the Planck table of a
// model block and, with EFF, the effective spectral parameters (gfx950, FP64).
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
// EFF adds the effective spectral parameters of zodipy/brightness.py:21-83, in
// which the integrand is linear: with W = h g, per component the node sums
// E_c = sum_k w_k n_c B and S_c = sum_k w_k n_c F Phi / R^2 give
// d/d(emissivity_c) = W (1 - a_c) E_c and d/d(albedo_c) = W (S_c - eps_c E_c);
// the solar irradiance F and the phase coefficients C1..C3 take
// sum_c W a_c sum_k w_k n_c Phi / R^2 (times F and dPhi/dC_i / Phi), and the RRM
// calibration sum_c W E_c. These are the entries of zodi_adjoint's gradient for
// the same parameters (the same partials, summed in another order), without the
// density and geometry partials: a bandpass or spectral-table fit needs nothing
// else, at about the cost of a forward pass.
//
// One ray per lane (or SPLIT lanes sharing its nodes). The registry layouts run
// with their kinds fixed at compile time (table_layout), any other layout through
// the kind switch (table_adjoint). A lane sums the shares of consecutive nodes in
// one table interval in registers and adds them to the wave's row in LDS with
// wavefront-scope atomic adds when the interval changes (lanes of a wave may meet
// in one entry); the spectral sums are wave sums per group (per component) and
// per ray (shared parameters). The block adds its wave rows in wave order and a
// reduction kernel adds the blocks in block order.

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

// The weight of component c for this ray: g_{c,r} (VJP_COMPS) or the ray's g.
__device__ __forceinline__ double comp_weight(const AdjArgs& a, const Ray& y, int c,
                                              double gscale) {
  if (a.mode == ZODI_ADJ_VJP_COMPS) return y.active ? a.weights[(int64_t)c * a.f.n_rays + y.r] : 0.0;
  return gscale;
}

// Where the scattering partials need Phi: as zodi_adjoint decides (Phi enters
// d/d(albedo) when F != 0, and d/dF, d/dC_i when an albedo is non-zero; with
// C2 = C3 = 0 it does not depend on the scattering angle).
struct PhaseUse {
  bool phase, angle;
  double phi_const;
};

__device__ __forceinline__ PhaseUse phase_use(ModelPtr M, double solar) {
  PhaseUse p;
  p.phase = solar != 0.0 || (M->flags & 1) != 0;
  p.angle = p.phase && ((M->flags & 1) != 0 || M->shared[ZODI_S_C2] != 0.0 ||
                        M->shared[ZODI_S_C3] != 0.0);
  p.phi_const = M->shared[ZODI_S_PHASE_NORM] * (M->shared[ZODI_S_C1] + 1.0);
  return p;
}

__device__ __forceinline__ PhaseGrad node_phase(const PhaseUse& pu, const Ray& y, double X0,
                                                double X1, double X2, double inv_R, ModelPtr M) {
  PhaseGrad pg = {pu.phase ? pu.phi_const : 0.0, 0.0, 0.0, 0.0};
  if (pu.angle) pg = phase_grad(y.ux, y.uy, y.uz, X0, X1, X2, inv_R, M->shared);
  return pg;
}

// The emissivity and albedo entries of component c from its sums (lane 0 adds
// the wave's).
__device__ __forceinline__ void eff_comp_finish(ModelPtr M, int c, double W, double E, double S,
                                                double* row, int lane) {
  const ParamPtr P = M->comp[c];
  double pe = W * (P[ZODI_P_ONE_MINUS_ALBEDO] * E);
  double pa = W * (S - P[ZODI_P_EMISSIVITY] * E);
  pe = wave_sum(pe);
  pa = wave_sum(pa);
  if (lane == 0) {
    row[M->grad_spec_a[c]] += pe;
    row[M->grad_spec_b[c]] += pa;
  }
}

// The shared entries of one ray (Kelsall: F, C1..C3 in sh[2..5]; RRM: the
// calibration in sh[0]), then cleared for the next ray.
template <int INTEGRAND>
__device__ __forceinline__ void eff_ray_finish(ModelPtr M, double (&sh)[6], double* row,
                                               int lane) {
  constexpr int Q0 = INTEGRAND == ZODI_INTEGRAND_KELSALL ? 2 : 0;
  constexpr int Q1 = INTEGRAND == ZODI_INTEGRAND_KELSALL ? 6 : 1;
#pragma unroll
  for (int q = Q0; q < Q1; ++q) {
    const double s = wave_sum(sh[q]);
    if (lane == 0) row[M->grad_shared[q]] += s;
    sh[q] = 0.0;
  }
}

// Any layout: the kind switch per component. The spectral sums of the (at most
// kMaxSlots) components of a group are selected by unrolled compares, so they
// stay in registers.
template <int INTEGRAND, int SPLIT, bool EFF>
__device__ __forceinline__ void table_ray(const Ctx& cx, const AdjArgs& a, const Ray& y,
                                          double gscale, double* row, int lane, int sub) {
  constexpr bool kKelsall = INTEGRAND == ZODI_INTEGRAND_KELSALL;
  const ModelPtr M = cx.M;
  const int n_last = M->n_temps - 1;
  const int n_groups = M->n_groups;
  double* trow = EFF ? row + M->n_params : row;
  double sh[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  PhaseUse pu = {false, false, 0.0};
  if constexpr (EFF && kKelsall) pu = phase_use(M, cx.solar);
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
    double W[kMaxSlots], ca[kMaxSlots], E[kMaxSlots], S[kMaxSlots];
#pragma unroll
    for (int j = 0; j < kMaxSlots; ++j) {
      cw[j] = 0.0;
      fd[j] = FeatureDir{1.0, 0.0};
      W[j] = ca[j] = E[j] = S[j] = 0.0;
      if (j < cnt) {
        const int c = first + j;
        const ParamPtr P = M->comp[c];
        const double s = kKelsall ? P[ZODI_P_SCALE] : cx.calibration;
        const double w = comp_weight(a, y, c, gscale);
        cw[j] = (h * s) * w;
        if (is_feature(M->comp_kind[c])) fd[j] = feature_dir(P, y.ex, y.ey);
        if constexpr (EFF) {
          W[j] = y.active ? h * w : 0.0;
          if (kKelsall && (M->comp_flags[c] & ZODI_CF_SCATTER)) ca[j] = W[j] * P[ZODI_P_ALBEDO];
        }
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
      double wB = 0.0, wsu = 0.0, inv_R2 = 0.0;
      PhaseGrad pg = {0.0, 0.0, 0.0, 0.0};
      if constexpr (EFF) {
        const double* e = cx.tb.tab + 4 * jt;
        wB = wk * fma(f, e[3], e[2]);
        if constexpr (kKelsall) {
          if (pu.phase) {
            pg = node_phase(pu, y, X0, X1, X2, inv_R, Mk);
            inv_R2 = inv_R * inv_R;
            wsu = wk * ((cx.solar * inv_R2) * pg.phi);
          }
        }
      }
      double Sn = 0.0, Sa = 0.0;
#pragma unroll 1
      for (int q = 0; q < cnt; ++q) {
        const int c = first + q;
        const double n = density(Mk->comp_kind[c], Mk->comp[c], Mk->comp_flags[c], X0, X1, X2,
                                 R, inv_R, R2, fd[q]);
        Sn += cw[q] * n;
        if constexpr (EFF) {
#pragma unroll
          for (int j = 0; j < kMaxSlots; ++j) {
            if (j == q) {
              E[j] = fma(wB, n, E[j]);
              if (kKelsall) {
                S[j] = fma(wsu, n, S[j]);
                Sa = fma(ca[j], n, Sa);
              } else {
                Sa = fma(W[j], n, Sa);
              }
            }
          }
        }
      }
      pend.add(trow, n_last, jt, f, wk * Sn);
      if constexpr (EFF) {
        if constexpr (kKelsall) {
          if (pu.phase) {
            const double was = (wk * Sa) * inv_R2;
            sh[2] += was * pg.phi;
            const double wasf = was * cx.solar;
            sh[3] += wasf * pg.d1;
            sh[4] += wasf * pg.d2;
            sh[5] += wasf * pg.d3;
          }
        } else {
          sh[0] += wB * Sa;  // calibration
        }
      }
    }
    pend.flush(trow, n_last);
    if constexpr (EFF && kKelsall) {
#pragma unroll
      for (int j = 0; j < kMaxSlots; ++j)
        if (j < cnt) eff_comp_finish(M, first + j, W[j], E[j], S[j], row, lane);
    }
  }
  if constexpr (EFF) eff_ray_finish<INTEGRAND>(M, sh, row, lane);
}

// Registry layouts: the group's kinds (and, with PROPS, the registry forms of the
// forward kernels) fixed at compile time, the densities inlined. NOSCAT (EFF
// only): no component scatters and F = 0, so the scattering sums are exact zeros
// and left out.
template <int INTEGRAND, int SPLIT, uint32_t PROPS, bool EFF, bool NOSCAT, int... Ks, int... J>
__device__ __forceinline__ void table_group(Grp<Ks...>, std::integer_sequence<int, J...>,
                                            const Ctx& cx, const AdjArgs& a, const Ray& y, int g,
                                            int first, double gscale, double* row, int lane,
                                            int sub, const PhaseUse& pu, double (&sh)[6]) {
  constexpr int C = sizeof...(Ks);
  constexpr int kinds[C] = {Ks...};
  constexpr bool kKelsall = INTEGRAND == ZODI_INTEGRAND_KELSALL;
  constexpr bool kScat = EFF && kKelsall && !NOSCAT;
  const ModelPtr M = cx.M;
  const int n_last = M->n_temps - 1;
  double* trow = EFF ? row + M->n_params : row;
  double h, mid;
  group_range(M, g, y, h, mid);
  const double us = M->group_T0[g] * cx.tb.inv_step;
  const double uo = cx.tb.T_first * cx.tb.inv_step;
  const double mdelta = -M->group_delta[g];
  FeatureDir fdir[C];
  ((fdir[J] = is_feature(kinds[J]) ? feature_dir(M->comp[first + J], y.ex, y.ey)
                                   : FeatureDir{1.0, 0.0}),
   ...);
  double cw[C];  // h s_c g_c
  double ca[C];  // EFF: Kelsall h g_c a_c where c scatters (else 0); RRM h g_c
  double E[C], S[C];
  (
      [&] {
        const int c = first + J;
        const double s = kKelsall ? M->comp[c][ZODI_P_SCALE] : cx.calibration;
        const double w = comp_weight(a, y, c, gscale);
        cw[J] = (h * s) * w;
        if constexpr (EFF) {
          E[J] = 0.0;
          S[J] = 0.0;
          const double W = y.active ? h * w : 0.0;
          ca[J] = kKelsall ? ((M->comp_flags[c] & ZODI_CF_SCATTER) ? W * M->comp[c][ZODI_P_ALBEDO]
                                                                  : 0.0)
                           : W;
        }
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
    if constexpr (!EFF) {
      double Sn = 0.0;
      ((Sn += cw[J] * density_k<kinds[J], (PROPS >> (4 * J)) & 15, cut_mode<PROPS, J, Ks...>()>(
                Mk->comp[first + J], Mk->comp_flags[first + J], nd, fdir[J], cut)),
       ...);
      pend.add(trow, n_last, jt, f, wk * Sn);
    } else {
      // B and (with scattering) Phi first, so each density is summed as it is formed
      const double* e = cx.tb.tab + 4 * jt;
      const double wB = wk * fma(f, e[3], e[2]);
      PhaseGrad pg = {0.0, 0.0, 0.0, 0.0};
      double inv_R2 = 0.0, wsu = 0.0;
      if constexpr (kScat) {
        if (pu.phase) {
          pg = node_phase(pu, y, nd.X0, nd.X1, nd.X2, nd.inv_R, Mk);
          inv_R2 = nd.inv_R * nd.inv_R;
          wsu = wk * ((cx.solar * inv_R2) * pg.phi);
        }
      }
      double Sn = 0.0, Sa = 0.0;
      (
          [&] {
            const double n =
                density_k<kinds[J], (PROPS >> (4 * J)) & 15, cut_mode<PROPS, J, Ks...>()>(
                    Mk->comp[first + J], Mk->comp_flags[first + J], nd, fdir[J], cut);
            Sn += cw[J] * n;
            if constexpr (kKelsall) {
              E[J] = fma(wB, n, E[J]);
              if constexpr (kScat) {
                S[J] = fma(wsu, n, S[J]);
                Sa = fma(ca[J], n, Sa);
              }
            } else {
              Sa = fma(ca[J], n, Sa);
            }
          }(),
          ...);
      pend.add(trow, n_last, jt, f, wk * Sn);
      if constexpr (kScat) {
        if (pu.phase) {
          const double was = (wk * Sa) * inv_R2;
          sh[2] += was * pg.phi;
          const double wasf = was * cx.solar;
          sh[3] += wasf * pg.d1;
          sh[4] += wasf * pg.d2;
          sh[5] += wasf * pg.d3;
        }
      }
      if constexpr (!kKelsall) sh[0] += wB * Sa;  // calibration
    }
  }
  pend.flush(trow, n_last);
  if constexpr (EFF && kKelsall) {
    (eff_comp_finish(M, first + J, comp_weight(a, y, first + J, gscale) * (y.active ? h : 0.0),
                     E[J], S[J], row, lane),
     ...);
  }
}

template <int SPLIT, uint32_t PROPS, bool EFF, bool NOSCAT, int INTEGRAND, class... Gs, int... G>
__device__ __forceinline__ void table_groups(Layout<INTEGRAND, Gs...>,
                                             std::integer_sequence<int, G...>, const Ctx& cx,
                                             const AdjArgs& a, const Ray& y, double gscale,
                                             double* row, int lane, int sub, const PhaseUse& pu,
                                             double (&sh)[6]) {
  (table_group<INTEGRAND, SPLIT, props_shift(PROPS, group_first<Gs...>(G)), EFF, NOSCAT>(
       Gs{}, std::make_integer_sequence<int, Gs::size>{}, cx, a, y, G, group_first<Gs...>(G),
       gscale, row, lane, sub, pu, sh),
   ...);
}

template <int SPLIT, uint32_t PROPS, bool EFF, bool NOSCAT, int INTEGRAND, class... Gs>
__device__ __forceinline__ void run_table(Layout<INTEGRAND, Gs...> lay, const Ctx& cx,
                                          const AdjArgs& a, const Ray& y, double gscale,
                                          double* row, int lane, int sub) {
  double sh[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  PhaseUse pu = {false, false, 0.0};
  if constexpr (EFF && INTEGRAND == ZODI_INTEGRAND_KELSALL && !NOSCAT)
    pu = phase_use(cx.M, cx.solar);
  table_groups<SPLIT, PROPS, EFF, NOSCAT>(lay, std::make_integer_sequence<int, sizeof...(Gs)>{},
                                          cx, a, y, gscale, row, lane, sub, pu, sh);
  if constexpr (EFF && (INTEGRAND != ZODI_INTEGRAND_KELSALL || !NOSCAT))
    eff_ray_finish<INTEGRAND>(cx.M, sh, row, lane);
}

// Resident waves per SIMD of the registry-layout kernels (their register budget):
// four, or three for the Kelsall layouts' spectral sums with the scattering terms
// (at four those spill more and ran 20 % slower). Measured on MI355X, four waves
// with a small spill beat three without it for the RRM layouts (5 %) and for the
// DIRBE spectral sums without scattering (3.5 %).
template <class LAYOUT, bool EFF, bool NOSCAT>
constexpr int table_waves() {
  return layout_integrand<LAYOUT>() == ZODI_INTEGRAND_KELSALL && EFF && !NOSCAT
             ? kTableWavesWide
             : kTableWavesLayout;
}

// SPEC: the registry forms (RegistryProps, with the shared band cut-off when the
// host finds the block allows it: SHARE).
template <class LAYOUT, int SPLIT, bool SPEC, bool SHARE, bool EFF, bool NOSCAT>
__global__ __launch_bounds__(kAdjBlock) __attribute__((
    amdgpu_waves_per_eu(table_waves<LAYOUT, EFF, NOSCAT>()))) void table_layout(AdjArgs a) {
  constexpr uint32_t PROPS =
      !SPEC ? 0u
            : (SHARE ? RegistryProps<LAYOUT>::value
                     : props_without_share(RegistryProps<LAYOUT>::value));
  adjoint_body<SPLIT, false, 0>(
      a, [&](const Ctx& cx, const Ray& y, double gscale, double* row, int lane, int sub, double*) {
        run_table<SPLIT, PROPS, EFF, NOSCAT>(LAYOUT{}, cx, a, y, gscale, row, lane, sub);
      });
}

template <int INTEGRAND, int SPLIT, bool EFF>
__global__ __launch_bounds__(kAdjBlock) __attribute__((
    amdgpu_waves_per_eu(kTableWavesGeneric))) void table_adjoint(AdjArgs a) {
  adjoint_body<SPLIT, false, 0>(
      a, [&](const Ctx& cx, const Ray& y, double gscale, double* row, int lane, int sub, double*) {
        table_ray<INTEGRAND, SPLIT, EFF>(cx, a, y, gscale, row, lane, sub);
      });
}

template <int INTEGRAND, bool EFF>
AdjKernel pick_table_split(int split) {
  switch (split) {
    case 1:
      return table_adjoint<INTEGRAND, 1, EFF>;
    case 2:
      return table_adjoint<INTEGRAND, 2, EFF>;
    case 4:
      return table_adjoint<INTEGRAND, 4, EFF>;
    default:
      return nullptr;
  }
}

template <class LAYOUT, bool SPEC, bool SHARE, bool EFF, bool NOSCAT>
TableKernel pick_table_layout_split(int split) {
  constexpr int kWaves = table_waves<LAYOUT, EFF, NOSCAT>();
  switch (split) {
    case 1:
      return {table_layout<LAYOUT, 1, SPEC, SHARE, EFF, NOSCAT>, kWaves};
    case 2:
      return {table_layout<LAYOUT, 2, SPEC, SHARE, EFF, NOSCAT>, kWaves};
    case 4:
      return {table_layout<LAYOUT, 4, SPEC, SHARE, EFF, NOSCAT>, kWaves};
    default:
      return {nullptr, 0};
  }
}

// NOSCAT applies to the Kelsall layouts with EFF (the RRM integrand has no
// scattering term, and without EFF nothing is summed for it).
template <class LAYOUT, bool SPEC, bool SHARE, bool EFF>
TableKernel pick_table_noscat(int split, bool noscat) {
  if constexpr (EFF && layout_integrand<LAYOUT>() == ZODI_INTEGRAND_KELSALL) {
    if (noscat) return pick_table_layout_split<LAYOUT, SPEC, SHARE, true, true>(split);
  }
  return pick_table_layout_split<LAYOUT, SPEC, SHARE, EFF, false>(split);
}

template <class LAYOUT, bool EFF>
TableKernel pick_table_layout(const zodi_model* m, int split) {
  constexpr uint32_t kProps = RegistryProps<LAYOUT>::value;
  const bool noscat = block_noscat(m);
  if constexpr (kProps != props_without_share(kProps)) {
    if (props_match(m, kProps)) return pick_table_noscat<LAYOUT, true, true, EFF>(split, noscat);
  }
  if (props_match(m, props_without_share(kProps)))
    return pick_table_noscat<LAYOUT, true, false, EFF>(split, noscat);
  return pick_table_noscat<LAYOUT, false, false, EFF>(split, noscat);
}

template <bool EFF>
TableKernel pick_table(const zodi_model* m, int split) {
  switch (match_layout(m)) {
    case kLayoutDirbe:
      return pick_table_layout<LayoutDirbe, EFF>(m, split);
    case kLayoutKelsall4:
      return pick_table_layout<LayoutKelsall4, EFF>(m, split);
    case kLayoutRRM:
      return pick_table_layout<LayoutRRM, EFF>(m, split);
    case kLayoutRRMSplit:
      return pick_table_layout<LayoutRRMSplit, EFF>(m, split);
    default:
      break;
  }
  if (m->integrand == ZODI_INTEGRAND_KELSALL)
    return {pick_table_split<ZODI_INTEGRAND_KELSALL, EFF>(split), kTableWavesGeneric};
  if (m->integrand == ZODI_INTEGRAND_RRM)
    return {pick_table_split<ZODI_INTEGRAND_RRM, EFF>(split), kTableWavesGeneric};
  return {nullptr, 0};
}

}  // namespace

TableKernel table_adjoint_kernel(const zodi_model* m, int split, bool eff) {
  return eff ? pick_table<true>(m, split) : pick_table<false>(m, split);
}

}  // namespace zodi
