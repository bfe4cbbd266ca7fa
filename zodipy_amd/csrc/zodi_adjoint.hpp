// Adjoint device code shared by the adjoint kernels (zodi_adjoint.hip: the
// registry layouts; zodi_adjoint_generic.hip: any component layout). See
// zodi_adjoint.hip for the method.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <tuple>
#include <utility>

#include "zodi_common.hpp"
#include "zodi_grad.hpp"

namespace zodi {

constexpr int kWave = 64;
constexpr int kAdjWaves = 4;
constexpr int kAdjBlock = kWave * kAdjWaves;

struct AdjArgs {
  FwdArgs f;
  const double* weights;   // VJP_SUM: [n_rays]; VJP_COMPS: [n_comps][n_rays]
  const double* target;    // CHI2: data [n_rays]
  const double* inv_var;   // CHI2: 1/sigma^2 per ray, or null for inv_var_scalar
  double inv_var_scalar;
  const double* model_vals;  // CHI2: model values I [n_rays] (forward kernel output)
  double* partials;          // [gridDim.x][n_params + 1]
  double* jac;               // Jacobian pass: [n_params][ld_jac]
  int64_t ld_jac;
  double* model_vals_out;    // Jacobian pass: I [n_rays]
  int32_t n_params;
  int32_t mode;              // ZODI_ADJ_*
};

// Sum over the 64 lanes of a wave; a + b == b + a in IEEE arithmetic, so every
// lane ends with identical bits.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// B(T) of the Planck table on its index axis, u = (T - T_first)/dT, as
// planck_lookup_u forms it, and the table slope dB/dT of the interval (0 where
// the table clamps, u <= 0 or u >= ZODI_N_TEMPS - 1; the LDS entries past the
// last temperature carry slope 0). NaN passes.
__device__ __forceinline__ double planck_lookup_u_slope(double u, const double* __restrict__ tab,
                                                        double& dBdT) {
  const double uc = fmin(fmax(u, 0.0), (double)(ZODI_N_TEMPS - 1));
  int j = (int)uc;
  j = j > ZODI_N_TEMPS - 2 ? ZODI_N_TEMPS - 2 : j;
  const double f = uc - (double)j;
  const double* e = tab + 4 * j;
  dBdT = (u > 0.0 && u < (double)(ZODI_N_TEMPS - 1)) ? e[1] : 0.0;
  const double B = fma(f, e[3], e[2]);
  return u == u ? B : u;
}

// Accumulators of one component: the geometry moments (geo_moments), the kind's
// parameters (NP) and two more slots: the spectral parameters (emissivity and
// albedo, or RRM's T_0 and delta). In the Jacobian pass a component of a Kelsall
// model sums sum w n B and sum w n F Phi / R^2 in those two slots instead (its
// spectral partials follow from them after the node loop); a component of an RRM
// model (of any kind) sums sum w n B in a third slot.
template <int INTEGRAND, int KIND>
constexpr int acc_slots() {
  return geo_moments<is_feature(KIND)>() + kind_nparams<KIND>() + 2 +
         (INTEGRAND == ZODI_INTEGRAND_RRM ? 1 : 0);
}

template <int INTEGRAND, int KIND>
struct CompAcc {
  static constexpr int kSlots = acc_slots<INTEGRAND, KIND>();
  double v[kSlots];
  __device__ __forceinline__ void add(int q, double x) { v[q] += x; }
  __device__ __forceinline__ double get(int q) const { return v[q]; }
};

// The same accumulators in LDS, one slot per 64-lane row (slot q of a lane at
// p[64 q]): the components of a Jacobian chunk beyond the register ones keep theirs
// here (in_lds), so one pass over a group's nodes serves four components without
// spilling registers. Each lane owns its slots, and each sum is one ds_add_f64
// per node in program order, so the sums repeat bit for bit; they agree with the
// register sums to rounding (there the product may fuse into the addition).
template <int INTEGRAND, int KIND>
struct LdsAcc {
  static constexpr int kSlots = acc_slots<INTEGRAND, KIND>();
  double* p;
  __device__ __forceinline__ void add(int q, double x) const {
    __hip_atomic_fetch_add(p + q * kWave, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  }
  __device__ __forceinline__ double get(int q) const { return p[q * kWave]; }
};

template <int INTEGRAND, int KIND, bool IN_LDS>
struct AccOf {
  using type = CompAcc<INTEGRAND, KIND>;
};
template <int INTEGRAND, int KIND>
struct AccOf<INTEGRAND, KIND, true> {
  using type = LdsAcc<INTEGRAND, KIND>;
};

// Components per Jacobian chunk whose accumulators stay in registers by default
// (two components' sums with the scattering terms fill the registers that two
// waves per SIMD leave).
constexpr int kJacRegComps = 2;

template <bool JAC, int CHUNK>
constexpr int reg_comps() {
  return JAC && CHUNK > kJacRegComps ? kJacRegComps : CHUNK;
}

// Whether component i of a chunk of n, nreg of them in registers, sums in LDS:
// the first n - nreg. A wave's LDS operations complete in order, so the first
// table read after a component's ds_add_f64 waits for them; the next component's
// arithmetic before its first table read covers part of that wait, while after
// the chunk's last component only the next node's radius precedes its log.
constexpr bool in_lds(int i, int n, int nreg) { return i < n - nreg; }

// LDS slots per lane for the largest chunk of a layout (0 without LDS components).
template <int INTEGRAND, int CHUNK, int NREG, int... Ks>
constexpr int group_lds_slots(Grp<Ks...>) {
  constexpr int C = sizeof...(Ks);
  const int slots[C] = {acc_slots<INTEGRAND, Ks>()...};
  int best = 0;
  for (int s = 0; s < C; s += CHUNK) {
    const int n = C - s < CHUNK ? C - s : CHUNK;
    int t = 0;
    for (int i = 0; i < n; ++i) t += in_lds(i, n, NREG) ? slots[s + i] : 0;
    best = t > best ? t : best;
  }
  return best;
}

template <int CHUNK, int NREG, int INTEGRAND, class... Gs>
constexpr int layout_lds_slots(Layout<INTEGRAND, Gs...>) {
  int best = 0;
  ((best = group_lds_slots<INTEGRAND, CHUNK, NREG>(Gs{}) > best
               ? group_lds_slots<INTEGRAND, CHUNK, NREG>(Gs{})
               : best),
   ...);
  return best;
}

// Components J (consecutive, from S) of group g, whose first component is `first`.
// Lane `sub` of the SPLIT lanes sharing the ray takes nodes sub, sub + SPLIT, ...
// `sh` holds the shared-parameter sums of the ray (Kelsall: T0, delta, F, C1..C3;
// RRM: calibration) across all groups.
//
// JAC (one lane per ray): the weights are 1, and each lane stores its ray's sums
// as a column of the Jacobian, dI_r/dtheta_p -> jac[p][r], instead of reducing
// them over the wave; `total` receives I_r, each component's node sum finished
// as the forward kernel finishes it and added in component order.
//
// PROPS: 4 bits per component of the group (component J at bits 4J; see
// density_k): the frame mode and kFast, the registry forms of the band's
// partials (p = 4, the pre-scaled plane normal, the R^2 cut-off). NOSCAT: the
// block has no scattering component and no solar irradiance (the host checks),
// so every scattering sum is an exact zero and is left out.
template <int INTEGRAND, int SPLIT, bool JAC, int NREG, int S, uint32_t PROPS, bool NOSCAT,
          int... Ks, int... J>
__device__ __forceinline__ void adjoint_group(Grp<Ks...>, std::integer_sequence<int, J...>,
                                              const Ctx& cx, const AdjArgs& a, const Ray& y,
                                              int g, int first, double gscale, double* row,
                                              int lane, int sub, double (&sh)[6],
                                              double& total, double* lane_acc) {
  static_assert(!JAC || SPLIT == 1, "the Jacobian pass uses one lane per ray");
  static_assert(JAC || NREG >= (int)sizeof...(J), "LDS accumulators are for the Jacobian pass");
  constexpr int C = sizeof...(Ks);
  constexpr int N = sizeof...(J);
  constexpr int kinds[C] = {Ks...};
  // LDS slots of each chunk component (0 for those in registers) and their offsets.
  constexpr int lds_len[N] = {
      (in_lds(J - S, N, NREG) ? acc_slots<INTEGRAND, kinds[J]>() : 0)...};
  constexpr auto lds_off = [](const int(&len)[N], int i) {
    int o = 0;
    for (int t = 0; t < i; ++t) o += len[t];
    return o;
  };
  const ModelPtr M = cx.M;
  double h, mid;
  group_range(M, g, y, h, mid);
  const double T0 = M->group_T0[g];
  const double mdelta = -M->group_delta[g];
  const double us = T0 * cx.tb.inv_step;  // T = T0 R^-delta on the table's index axis
  const double uo = cx.tb.T_first * cx.tb.inv_step;

  double W[N];  // h times the weight of each component
  if constexpr (JAC) {
    ((W[J - S] = y.active ? h : 0.0), ...);
  } else if (a.mode == ZODI_ADJ_VJP_COMPS) {
    ((W[J - S] = y.active ? h * a.weights[(int64_t)(first + J) * a.f.n_rays + y.r] : 0.0), ...);
  } else {
    ((W[J - S] = h * gscale), ...);
  }
  FeatureDir fdir[N];
  ((fdir[J - S] = is_feature(kinds[J]) ? feature_dir(M->comp[first + J], y.ex, y.ey)
                                       : FeatureDir{1.0, 0.0}),
   ...);
  EarthTerm et[N];
  ((et[J - S] = is_feature(kinds[J]) ? earth_term(M->comp[first + J], y.ex, y.ey)
                                     : EarthTerm{0.0, 0.0}),
   ...);
  const double solar = cx.solar;
  // Phi enters d/d(albedo) when F != 0, and d/dF, d/dC_i when an albedo is non-zero.
  // With every albedo 0 the C_i partials are multiplied by 0, and with C2 = C3 = 0
  // Phi = N (C1 + 1) does not depend on the scattering angle: skip the angle.
  const bool need_phase = !NOSCAT && INTEGRAND == ZODI_INTEGRAND_KELSALL &&
                          (solar != 0.0 || (M->flags & 1) != 0);
  const bool need_angle = need_phase && ((M->flags & 1) != 0 || M->shared[ZODI_S_C2] != 0.0 ||
                                         M->shared[ZODI_S_C3] != 0.0);
  const double phi_const = M->shared[ZODI_S_PHASE_NORM] * (M->shared[ZODI_S_C1] + 1.0);

  std::tuple<typename AccOf<INTEGRAND, kinds[J], in_lds(J - S, N, NREG)>::type...> acc{};
  (
      [&] {
        if constexpr (in_lds(J - S, N, NREG)) {
          auto& A = std::get<J - S>(acc);
          A.p = lane_acc + lds_off(lds_len, J - S) * kWave;
#pragma unroll
          for (int q = 0; q < A.kSlots; ++q) A.p[q * kWave] = 0.0;
        }
      }(),
      ...);
  // JAC: node sums per component, sum w n B and (Kelsall) sum w n F Phi / R^2.
  // The component's integral, and its emissivity and albedo partials (and RRM's
  // calibration partial), are linear in them: they are formed after the loop.

  const int K = y.active ? cx.K : 0;
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
    const double logR = log_pos(nd.R);
    const double Rd = exp_fast_bounded(mdelta * logR);  // R^-delta = dT/dT0
    double dBdT;
    nd.B = planck_lookup_u_slope(fma(Rd, us, -uo), cx.tb.tab, dBdT);
    const double dT_dT0 = Rd;
    const double dT_ddelta = -(T0 * Rd) * logR;
    PhaseGrad pg = {need_phase ? phi_const : 0.0, 0.0, 0.0, 0.0};
    if (need_angle) {
      pg = phase_grad(y.ux, y.uy, y.uz, nd.X0, nd.X1, nd.X2, nd.inv_R, Mk->shared);
    }
    const double inv_R2 = nd.inv_R * nd.inv_R;
    double sw = 0.0;  // Kelsall: sum over the chunk of w n (1 - a) eps, for T_0 and delta
    CutShare cs = {0.0, 0.0};  // a band's cut-off, for the bands that share it (in order)
    (
        [&] {
          constexpr int KIND = kinds[J];
          constexpr int NP = kind_nparams<KIND>();
          constexpr int G = geo_moments<is_feature(KIND)>();
          constexpr int CUT = cut_mode<PROPS, J, Ks...>();
          static_assert(CUT != kCutReuse || first_band<Ks...>() >= S,
                        "a band that reuses a cut-off needs its publisher in the same pass");
          const ParamPtr P = Mk->comp[first + J];
          const uint32_t flags = Mk->comp_flags[first + J];
          constexpr uint32_t PC = (PROPS >> (4 * J)) & 15u;
          const CompFrame f =
              comp_frame<PC & 3>(P, flags, nd.X0, nd.X1, nd.X2, nd.R, nd.inv_R, nd.R2);
          const DensGrad<NP> dg =
              density_grad<KIND, (PC & kFast) != 0, CUT>(P, flags, f, fdir[J - S], logR, cs);
          auto& A = std::get<J - S>(acc);
          const double wn = wk * W[J - S];
          if constexpr (INTEGRAND == ZODI_INTEGRAND_KELSALL) {
            // f = [(1 - a) eps B + a F Phi / R^2] n   (zodipy/brightness.py:21-56)
            const double alb = P[ZODI_P_ALBEDO], eps = P[ZODI_P_EMISSIVITY];
            const double oma = P[ZODI_P_ONE_MINUS_ALBEDO];
            const double scat_unit = NOSCAT ? 0.0 : (solar * inv_R2) * pg.phi;  // F Phi / R^2
            double e = oma * (eps * nd.B);
            if (!NOSCAT && (flags & ZODI_CF_SCATTER)) e += alb * scat_unit;
            const double we = wn * e;
            geo_accumulate<is_feature(KIND)>(A, we, f, dg.dR, dg.dZ, dg.dth);
#pragma unroll
            for (int q = 0; q < NP; ++q) A.add(G + q, we * dg.dp[q]);
            const double wnn = wn * dg.n;
            if constexpr (JAC) {
              A.add(G + NP, (wk * nd.B) * dg.n);
              if constexpr (!NOSCAT) A.add(G + NP + 1, (wk * scat_unit) * dg.n);
            } else {
              A.add(G + NP, wnn * (oma * nd.B));                  // emissivity
              A.add(G + NP + 1, wnn * (scat_unit - eps * nd.B));  // albedo
            }
            // (1 - a) eps, packed with the same rounding as oma * eps
            sw += wnn * P[ZODI_P_SCALE];
            // F, C1..C3 enter through a Phi / R^2, which is 0 when a == 0
            if (!NOSCAT && (flags & ZODI_CF_SCATTER)) {
              const double was = wnn * alb * inv_R2;
              sh[2] += was * pg.phi;
              const double wasf = was * solar;
              sh[3] += wasf * pg.d1;
              sh[4] += wasf * pg.d2;
              sh[5] += wasf * pg.d3;
            }
          } else {
            // f = calibration B(T_c) n   (zodipy/brightness.py:59-83)
            if constexpr (JAC) A.add(G + NP + 2, (wk * nd.B) * dg.n);
            const double wc = wn * cx.calibration;
            const double wcb = wc * nd.B;
            geo_accumulate<is_feature(KIND)>(A, wcb, f, dg.dR, dg.dZ, dg.dth);
#pragma unroll
            for (int q = 0; q < NP; ++q) A.add(G + q, wcb * dg.dp[q]);
            const double wt = wc * dg.n * dBdT;
            A.add(G + NP, wt * dT_dT0);         // T_0 of this component
            A.add(G + NP + 1, wt * dT_ddelta);  // delta of this component
            if constexpr (!JAC) sh[0] += wn * nd.B * dg.n;  // calibration
          }
        }(),
        ...);
    if constexpr (INTEGRAND == ZODI_INTEGRAND_KELSALL) {
      // T = T_0 R^-delta is shared by the chunk's components
      const double wt = sw * dBdT;
      sh[0] += wt * dT_dT0;
      sh[1] += wt * dT_ddelta;
    }
  }

  if constexpr (JAC) {
    // The ray's Jacobian entries of these components (coalesced over the wave).
    (
        [&] {
          constexpr int KIND = kinds[J];
          constexpr int NP = kind_nparams<KIND>();
          constexpr int G = geo_moments<is_feature(KIND)>();
          const auto& AC = std::get<J - S>(acc);
          double A[acc_slots<INTEGRAND, KIND>()];
#pragma unroll
          for (int q = 0; q < acc_slots<INTEGRAND, KIND>(); ++q) A[q] = AC.get(q);
          const int c = first + J;
          const ParamPtr P = M->comp[c];
          const uint32_t flags = M->comp_flags[c];
          double Ib;  // sum w n B
          if constexpr (INTEGRAND == ZODI_INTEGRAND_KELSALL) {
            Ib = A[G + NP];
          } else {
            Ib = A[G + NP + 2];
          }
          const double hb = y.active ? h * Ib : 0.0;
          double node_sum = Ib;  // the forward's sum: w B n, or w e n with scattering
          if constexpr (INTEGRAND == ZODI_INTEGRAND_KELSALL) {
            const double Is = A[G + NP + 1];  // sum w n F Phi / R^2
            const double hs = y.active ? h * Is : 0.0;
            A[G + NP] = P[ZODI_P_ONE_MINUS_ALBEDO] * hb;      // emissivity
            A[G + NP + 1] = hs - P[ZODI_P_EMISSIVITY] * hb;  // albedo
            if (flags & ZODI_CF_SCATTER) node_sum = P[ZODI_P_SCALE] * Ib + P[ZODI_P_ALBEDO] * Is;
          } else {
            sh[0] += hb;  // calibration
          }
          total = __dadd_rn(total, finish<INTEGRAND>(P, flags, node_sum, h, cx.calibration));
          double geo[5];
          geo_finish<is_feature(kinds[J])>(M->comp[c], A, et[J - S], geo);
          if (y.active) {
            double* col = a.jac + y.r;
            const int64_t ld = a.ld_jac;
            const int off = M->grad_comp[c];
#pragma unroll
            for (int q = 0; q < 5; ++q) col[(off + q) * ld] = geo[q];
#pragma unroll
            for (int q = 0; q < NP; ++q) col[(off + 5 + q) * ld] = A[G + q];
            col[M->grad_spec_a[c] * ld] = A[G + NP];
            col[M->grad_spec_b[c] * ld] = A[G + NP + 1];
          }
        }(),
        ...);
  } else {
    // Wave sums into the wave's gradient row (lane 0 writes).
    (
        [&] {
          constexpr int NP = kind_nparams<kinds[J]>();
          constexpr int G = geo_moments<is_feature(kinds[J])>();
          auto& A = std::get<J - S>(acc).v;  // registers: NREG covers the chunk
          const int c = first + J;
          double geo[5];
          geo_finish<is_feature(kinds[J])>(M->comp[c], A, et[J - S], geo);
#pragma unroll
          for (int q = 0; q < 5; ++q) geo[q] = wave_sum(geo[q]);
#pragma unroll
          for (int q = G; q < G + NP + 2; ++q) A[q] = wave_sum(A[q]);
          if (lane == 0) {
            const int off = M->grad_comp[c];
#pragma unroll
            for (int q = 0; q < 5; ++q) row[off + q] += geo[q];
#pragma unroll
            for (int q = 0; q < NP; ++q) row[off + 5 + q] += A[G + q];
            row[M->grad_spec_a[c]] += A[G + NP];
            row[M->grad_spec_b[c]] += A[G + NP + 1];
          }
        }(),
        ...);
  }
}

template <int S, int... I>
constexpr std::integer_sequence<int, (S + I)...> shift_seq(std::integer_sequence<int, I...>) {
  return {};
}

// Group g in chunks of at most CHUNK components, from component S of the group.
template <int INTEGRAND, int SPLIT, int CHUNK, int NREG, bool JAC, int S, uint32_t PROPS,
          bool NOSCAT, int... Ks>
__device__ __forceinline__ void adjoint_chunks(Grp<Ks...> grp, const Ctx& cx, const AdjArgs& a,
                                               const Ray& y, int g, int first, double gscale,
                                               double* row, int lane, int sub, double (&sh)[6],
                                               double& total, double* lane_acc) {
  constexpr int C = sizeof...(Ks);
  if constexpr (S < C) {
    constexpr int N = (C - S) < CHUNK ? (C - S) : CHUNK;
    adjoint_group<INTEGRAND, SPLIT, JAC, NREG, S, PROPS, NOSCAT>(
        grp, shift_seq<S>(std::make_integer_sequence<int, N>{}), cx, a, y, g, first, gscale, row,
        lane, sub, sh, total, lane_acc);
    adjoint_chunks<INTEGRAND, SPLIT, CHUNK, NREG, JAC, S + N, PROPS, NOSCAT>(
        grp, cx, a, y, g, first, gscale, row, lane, sub, sh, total, lane_acc);
  }
}

// Every group of the layout, with its first component known at compile time.
template <int SPLIT, int CHUNK, int NREG, bool JAC, uint32_t PROPS, bool NOSCAT, int INTEGRAND,
          class... Gs, int... G>
__device__ __forceinline__ void adjoint_groups(Layout<INTEGRAND, Gs...>,
                                               std::integer_sequence<int, G...>, const Ctx& cx,
                                               const AdjArgs& a, const Ray& y, double gscale,
                                               double* row, int lane, int sub, double (&sh)[6],
                                               double& total, double* lane_acc) {
  (adjoint_chunks<INTEGRAND, SPLIT, CHUNK, NREG, JAC, 0,
                  props_shift(PROPS, group_first<Gs...>(G)), NOSCAT>(
       Gs{}, cx, a, y, G, group_first<Gs...>(G), gscale, row, lane, sub, sh, total, lane_acc),
   ...);
}

// After every group of a ray: the shared-parameter sums go to the wave's
// gradient row (wave sum, lane 0 writes) or, in the Jacobian pass, to the ray's
// Jacobian column together with I_r.
template <int INTEGRAND, bool JAC>
__device__ __forceinline__ void adjoint_ray_tail(const Ctx& cx, const AdjArgs& a, const Ray& y,
                                                 double (&sh)[6], double total, double* row,
                                                 int lane) {
  const ModelPtr M = cx.M;
  constexpr int NSH = INTEGRAND == ZODI_INTEGRAND_KELSALL ? 6 : 1;
  if constexpr (JAC) {
    if (y.active) {
      a.model_vals_out[y.r] = total;
#pragma unroll
      for (int q = 0; q < NSH; ++q) a.jac[M->grad_shared[q] * a.ld_jac + y.r] = sh[q];
    }
  } else {
#pragma unroll
    for (int q = 0; q < NSH; ++q) sh[q] = wave_sum(sh[q]);
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < NSH; ++q) row[M->grad_shared[q]] += sh[q];
    }
  }
}

// Every group of the layout for one ray, then adjoint_ray_tail.
template <int SPLIT, int CHUNK, int NREG, bool JAC, uint32_t PROPS, bool NOSCAT, int INTEGRAND,
          class... Gs>
__device__ __forceinline__ void run_adjoint(Layout<INTEGRAND, Gs...> lay, const Ctx& cx,
                                            const AdjArgs& a, const Ray& y, double gscale,
                                            double* row, int lane, int sub, double* lane_acc) {
  double sh[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double total = 0.0;
  adjoint_groups<SPLIT, CHUNK, NREG, JAC, PROPS, NOSCAT>(
      lay, std::make_integer_sequence<int, sizeof...(Gs)>{}, cx, a, y, gscale, row, lane, sub, sh,
      total, lane_acc);
  adjoint_ray_tail<INTEGRAND, JAC>(cx, a, y, sh, total, row, lane);
}

// The body of every adjoint kernel: one wave per tile of 64 / SPLIT rays (SPLIT
// lanes share a ray's nodes), the per-ray weight (chi^2 mode: 2 (I - d)/sigma^2,
// and chi^2 itself summed), `ray(cx, y, gscale, row, lane, sub, lane_acc)` per
// ray, and the block's partial row. JAC: the Jacobian pass (one lane per ray, no
// weights, no reduction). LDS_SLOTS: per-lane LDS accumulator slots.
template <int SPLIT, bool JAC, int LDS_SLOTS, class RAY>
__device__ __forceinline__ void adjoint_body(const AdjArgs& a, RAY ray) {
  constexpr int RAYS = kWave / SPLIT;
  extern __shared__ double sRows[];  // kAdjWaves rows of n_params + 1 (not used by JAC)
  __shared__ double sTab[4 * ZODI_N_TEMPS];
  __shared__ double sAcc[LDS_SLOTS > 0 ? kAdjWaves * kWave * LDS_SLOTS : 1];
  const ModelPtr M = as_const(a.f.m);
  const int tid = threadIdx.x;
  const int wave = tid / kWave;
  const int lane = tid % kWave;
  const int sub = lane % SPLIT;
  const int np1 = a.n_params + 1;
  if constexpr (!JAC) {
    for (int i = tid; i < np1 * kAdjWaves; i += kAdjBlock) sRows[i] = 0.0;
  }
  stage_math_tables();
  load_table(M, sTab);  // ends with a barrier
  const Ctx cx = make_ctx(M, sTab, 0);
  double* row = JAC ? nullptr : sRows + wave * np1;
  double* lane_acc = sAcc + wave * kWave * LDS_SLOTS + lane;

  const int64_t stride = (int64_t)gridDim.x * kAdjWaves;
  for (int64_t tile = (int64_t)blockIdx.x * kAdjWaves + wave; tile < a.f.n_tiles;
       tile += stride) {
    const Ray y = load_ray<RAYS>(a.f, tile, lane / SPLIT);
    if constexpr (JAC) {
      ray(cx, y, 1.0, row, lane, sub, lane_acc);
      continue;
    }
    double gscale = 0.0;
    if (a.mode == ZODI_ADJ_CHI2) {
      double chi = 0.0;
      if (y.active) {
        const double I = a.model_vals[y.r];
        const double iv = a.inv_var != nullptr ? a.inv_var[y.r] : a.inv_var_scalar;
        const double res = I - a.target[y.r];
        chi = sub == 0 ? res * res * iv : 0.0;
        gscale = 2.0 * res * iv;
      }
      chi = wave_sum(chi);
      if (lane == 0) row[a.n_params] += chi;
    } else if (a.mode == ZODI_ADJ_VJP_SUM) {
      gscale = y.active ? a.weights[y.r] : 0.0;
    }
    ray(cx, y, gscale, row, lane, sub, lane_acc);
  }
  if constexpr (JAC) return;
  __syncthreads();
  for (int p = tid; p < np1; p += kAdjBlock) {
    double s = sRows[p];
#pragma unroll
    for (int w = 1; w < kAdjWaves; ++w) s += sRows[w * np1 + p];
    a.partials[(int64_t)blockIdx.x * np1 + p] = s;
  }
}

// A Kelsall block whose components all skip scattering and whose solar
// irradiance is 0: every scattering sum of the adjoint is an exact zero.
inline bool block_noscat(const zodi_model* m) {
  return m->integrand == ZODI_INTEGRAND_KELSALL && (m->flags & 1) == 0 &&
         m->shared[ZODI_S_SOLAR] == 0.0;
}

// Kernel entry points of the other adjoint source files: custom component
// layouts, and the gradient with respect to the Planck table (with eff, also
// the effective spectral parameters).
using AdjKernel = void (*)(AdjArgs);
AdjKernel generic_adjoint_kernel(int integrand, int split, bool jac);

// Resident waves per SIMD the table kernels are compiled for (their register
// budgets): the registry layouts; the Kelsall ones with the spectral sums of
// scattering components; the kind switch (at three rather than its
// unconstrained two: 14-21 % faster for the table gradient, 5 % with the
// spectral sums, despite a small spill).
constexpr int kTableWavesLayout = 4;
constexpr int kTableWavesWide = 3;
constexpr int kTableWavesGeneric = 3;

struct TableKernel {
  AdjKernel k;
  int waves;  // resident blocks per CU of kAdjBlock threads
};
TableKernel table_adjoint_kernel(const zodi_model* m, int split, bool eff);

}  // namespace zodi
