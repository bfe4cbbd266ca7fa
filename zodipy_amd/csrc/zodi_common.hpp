// Shared device and host pieces of the zodiacal kernels (forward and adjoint).
#pragma once

#include <hip/hip_runtime.h>

#include <initializer_list>
#include <utility>

#include "zodi_device.hpp"
#include "zodi_hip.h"

namespace zodi {

// Host helpers (defined in zodi_forward.hip).
zodi_status fail(zodi_status code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
zodi_status check_hip(hipError_t e, const char* what);
int cu_count();
zodi_status validate_model(const zodi_model* m);

// Kernel arguments (passed by value in the kernarg segment).
struct FwdArgs {
  const zodi_model* m;
  const double* u;
  int64_t n_rays;
  const int64_t* seg;
  int32_t n_seg;
  int32_t obs_mode;
  const double* obs;
  const double* earth;
  int64_t obs_stride;  // row length of obs/earth: n_seg or n_rays
  const uint32_t* beyond;
  double* out;
  int32_t out_mode;
  int32_t pad;
  int64_t n_tiles;
};

// Index of the observer of ray r (largest s with seg[s] <= r); `s0` is that
// index for an earlier ray of the same tile.
__device__ __forceinline__ int64_t walk_segment(const int64_t* __restrict__ seg, int32_t n_seg,
                                                int64_t s0, int64_t r) {
  int64_t s = s0;
  while (s + 1 < n_seg && seg[s + 1] <= r) ++s;
  return s;
}

__device__ __forceinline__ int64_t find_segment(const int64_t* __restrict__ seg, int32_t n_seg,
                                                int64_t r) {
  int32_t lo = 0, hi = n_seg - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (seg[mid] <= r) {
      lo = mid;
    } else {
      hi = mid - 1;
    }
  }
  return lo;
}

// Planck table to LDS: entry i is (T_i, slope_i, B_i, B_{i+1} - B_i). Entries
// from the last temperature on to ZODI_N_TEMPS - 1 hold (T_last, 0, B_last, 0), so
// any index from n_temps - 1 to ZODI_N_TEMPS - 1 reads the table's end value.
__device__ __forceinline__ void load_table(ModelPtr M, double* sTab) {
  const int n = M->n_temps;
  for (int i = threadIdx.x; i < ZODI_N_TEMPS; i += blockDim.x) {
    const bool in = i < n - 1;
    const int k = in ? i : n - 1;
    sTab[4 * i] = M->tab_T[k];
    sTab[4 * i + 1] = in ? M->tab_slope[k] : 0.0;
    sTab[4 * i + 2] = M->tab_B[k];
    sTab[4 * i + 3] = in ? M->tab_B[k + 1] - M->tab_B[k] : 0.0;
  }
  __syncthreads();
}

// Per-ray state shared by every group.
struct Ray {
  double ux, uy, uz, ox, oy, oz, ex, ey;
  RayGeom G;
  uint32_t bits;
  int64_t r;
  bool active;
};

template <int RAYS>
__device__ __forceinline__ Ray load_ray(const FwdArgs& a, int64_t tile, int local) {
  Ray y;
  const int64_t n = a.n_rays;
  y.r = tile * RAYS + local;
  y.active = y.r < n;
  const int64_t rr = y.active ? y.r : n - 1;
  int64_t s;
  if (a.obs_mode == ZODI_OBS_PER_RAY) {
    s = rr;
  } else {
    const int64_t s0 = find_segment(a.seg, a.n_seg, tile * RAYS);
    s = walk_segment(a.seg, a.n_seg, s0, rr);
  }
  y.bits = a.beyond[a.obs_mode == ZODI_OBS_PER_RAY ? 0 : s];
  const double* __restrict__ U = a.u;
  y.ux = U[rr];
  y.uy = U[n + rr];
  y.uz = U[2 * n + rr];
  const int64_t os = a.obs_stride;
  y.ox = a.obs[s];
  y.oy = a.obs[os + s];
  y.oz = a.obs[2 * os + s];
  y.ex = a.earth[s];
  y.ey = a.earth[os + s];
  y.G = ray_geometry(y.ux, y.uy, y.uz, y.ox, y.oy, y.oz);
  return y;
}

// Line-of-sight range of group g for this ray.
__device__ __forceinline__ void group_range(ModelPtr M, int g, const Ray& y, double& h,
                                            double& mid) {
  const double start = ((y.bits >> (2 * g)) & 1u) ? kEps : sphere_distance(y.G, M->group_r_in[g]);
  const double stop =
      ((y.bits >> (2 * g + 1)) & 1u) ? kEps : sphere_distance(y.G, M->group_r_out[g]);
  h = 0.5 * (stop - start);
  mid = 0.5 * (stop + start);
}

// Node geometry of one group: heliocentric position, |X|, 1/|X|, Planck value.
struct Node {
  double X0, X1, X2, R, inv_R, R2, B, phi;
};

struct Tables {
  const double* tab;
  double T_first;
  double T_last;
  double inv_step;
  int n_t;
};

// us = T_0 / dT and uo = T_first / dT place T = T_0 R^-delta on the table's
// index axis: u = R^-delta us - uo.
template <bool NANFIX = true>
__device__ __forceinline__ Node node_geometry(const Ray& y, double h, double mid, double xk,
                                              double us, double uo, double mdelta,
                                              const Tables& tb) {
  Node nd;
  const double s_los = h * xk + mid;
  nd.X0 = s_los * y.ux + y.ox;
  nd.X1 = s_los * y.uy + y.oy;
  nd.X2 = s_los * y.uz + y.oz;
  nd.R2 = nd.X0 * nd.X0 + nd.X1 * nd.X1 + nd.X2 * nd.X2;
  norm_and_inverse_node(nd.R2, nd.R, nd.inv_R);
  // T = T0 R^-delta (zodipy/blackbody.py:16-30), B = interp(T) (zodipy/brightness.py:48)
  const double u = fma(pow_radius(nd.R, mdelta), us, -uo);
  nd.B = planck_lookup_u<NANFIX>(u, tb.tab);
  nd.phi = 0.0;
  return nd;
}

// Node contribution of one component to its accumulator, as two factors that
// node_add adds with one fused multiply-add, node by node in order.
// Kelsall (zodipy/brightness.py:21-56): (1 - a) eps B n, plus a F/R^2 Phi n with
// scattering; without scattering the constant (1 - a) eps is applied once to the
// node sum (finish). RRM (zodipy/brightness.py:59-83): B n, calibration applied
// in finish. NOSCAT: no component of the block scatters (no ZODI_CF_SCATTER flag,
// which the host checks), so the flag is not tested.
struct NodeTerm {
  double a, b;
};

template <int INTEGRAND, bool NOSCAT = false>
__device__ __forceinline__ NodeTerm node_term(ParamPtr P, uint32_t flags, const Node& nd,
                                              double wk, double wB, double dens, double solar) {
  if constexpr (INTEGRAND == ZODI_INTEGRAND_KELSALL && !NOSCAT) {
    if (flags & ZODI_CF_SCATTER) {
      double em = P[ZODI_P_ONE_MINUS_ALBEDO] * (P[ZODI_P_EMISSIVITY] * nd.B);
      em += (P[ZODI_P_ALBEDO] * (solar * (nd.inv_R * nd.inv_R))) * nd.phi;
      return {wk, em * dens};
    }
  }
  return {wB, dens};
}

__device__ __forceinline__ double node_add(double acc, NodeTerm t) { return fma(t.a, t.b, acc); }

// The terms of L consecutive nodes, one per lane of each L-lane segment, added to
// every lane's sums in node order (the first `left` of them; `left` is the same
// on all lanes).
template <int L, int C>
__device__ __forceinline__ void add_in_node_order(double (&acc)[C], const NodeTerm (&t)[C],
                                                  int left) {
#pragma unroll
  for (int s = 0; s < L; ++s) {
    if (s < left) {
#pragma unroll
      for (int j = 0; j < C; ++j)
        acc[j] = node_add(acc[j], NodeTerm{__shfl(t[j].a, s, L), __shfl(t[j].b, s, L)});
    }
  }
}

// Integrated value of one component from its node sum; h = (stop - start)/2.
template <int INTEGRAND, bool NOSCAT = false>
__device__ __forceinline__ double finish(ParamPtr P, uint32_t flags, double acc, double h,
                                         double calibration) {
  if constexpr (INTEGRAND == ZODI_INTEGRAND_KELSALL) {
    if (!NOSCAT && (flags & ZODI_CF_SCATTER)) return __dmul_rn(acc, h);
    return __dmul_rn(__dmul_rn(acc, P[ZODI_P_SCALE]), h);
  } else {
    return __dmul_rn(__dmul_rn(acc, calibration), h);
  }
}

// ------------------------------------------------------------ layout kernels
template <int... Ks>
struct Grp {
  static constexpr int size = sizeof...(Ks);
};

template <int INTEGRAND, class... Gs>
struct Layout {};

// Registry layouts.
using LayoutDirbe = Layout<ZODI_INTEGRAND_KELSALL, Grp<ZODI_CLOUD, ZODI_BAND, ZODI_BAND, ZODI_BAND>,
                           Grp<ZODI_RING>, Grp<ZODI_FEATURE>>;
using LayoutKelsall4 =
    Layout<ZODI_INTEGRAND_KELSALL, Grp<ZODI_CLOUD, ZODI_BAND, ZODI_BAND, ZODI_BAND>>;
using LayoutRRM =
    Layout<ZODI_INTEGRAND_RRM, Grp<ZODI_FAN>, Grp<ZODI_COMET>,
           Grp<ZODI_NARROW_BAND, ZODI_NARROW_BAND, ZODI_BROAD_BAND>, Grp<ZODI_INTERSTELLAR>,
           Grp<ZODI_RING_RRM>, Grp<ZODI_FEATURE_RRM>>;

// The RRM model with one component per group, as parameter fits pack it (the
// components' own temperature laws may then differ).
using LayoutRRMSplit =
    Layout<ZODI_INTEGRAND_RRM, Grp<ZODI_FAN>, Grp<ZODI_COMET>, Grp<ZODI_NARROW_BAND>,
           Grp<ZODI_NARROW_BAND>, Grp<ZODI_BROAD_BAND>, Grp<ZODI_INTERSTELLAR>, Grp<ZODI_RING_RRM>,
           Grp<ZODI_FEATURE_RRM>>;

enum LayoutId {
  kLayoutGeneric = 0,
  kLayoutDirbe = 1,
  kLayoutKelsall4 = 2,
  kLayoutRRM = 3,
  kLayoutRRMSplit = 4
};

inline bool layout_is(const zodi_model* m, int integrand, std::initializer_list<std::initializer_list<int>> groups) {
  if (m->integrand != integrand || m->n_groups != (int)groups.size()) return false;
  int g = 0, c = 0;
  for (const auto& grp : groups) {
    if (m->group_count[g] != (int)grp.size()) return false;
    for (int kind : grp) {
      if (m->comp_kind[c++] != kind) return false;
    }
    ++g;
  }
  return c == m->n_comps;
}

inline LayoutId match_layout(const zodi_model* m) {
  if (layout_is(m, ZODI_INTEGRAND_KELSALL,
                {{ZODI_CLOUD, ZODI_BAND, ZODI_BAND, ZODI_BAND}, {ZODI_RING}, {ZODI_FEATURE}}))
    return kLayoutDirbe;
  if (layout_is(m, ZODI_INTEGRAND_KELSALL, {{ZODI_CLOUD, ZODI_BAND, ZODI_BAND, ZODI_BAND}}))
    return kLayoutKelsall4;
  if (layout_is(m, ZODI_INTEGRAND_RRM,
                {{ZODI_FAN},
                 {ZODI_COMET},
                 {ZODI_NARROW_BAND, ZODI_NARROW_BAND, ZODI_BROAD_BAND},
                 {ZODI_INTERSTELLAR},
                 {ZODI_RING_RRM},
                 {ZODI_FEATURE_RRM}}))
    return kLayoutRRM;
  if (layout_is(m, ZODI_INTEGRAND_RRM,
                {{ZODI_FAN},
                 {ZODI_COMET},
                 {ZODI_NARROW_BAND},
                 {ZODI_NARROW_BAND},
                 {ZODI_BROAD_BAND},
                 {ZODI_INTERSTELLAR},
                 {ZODI_RING_RRM},
                 {ZODI_FEATURE_RRM}}))
    return kLayoutRRMSplit;
  return kLayoutGeneric;
}

// ---------------------------------------------------------- per-ray forward evaluation
// Static properties of a layout kernel's component, 4 bits each: the frame mode
// (FrameMode, bits 0-1), kFast (bit 2): a Kelsall kind evaluated with its
// registry-only shortcuts (the band's p = 4, the cloud's floored g; see their
// density functions), and kShareCut (bit 3, a band of the forward kernels): its
// radial cut-off 1 - exp(-(R/delta_r)^20) is that of the first band of its group
// (both centred and fast, the same 1/delta_r^2 bit for bit, so the same R^2 and the
// same x), which the band reuses instead of forming it again. The DIRBE bands 1
// and 3 share delta_r = 1.5.
constexpr uint32_t kFast = 4u;
constexpr uint32_t kShareCut = 8u;

// The property word without the kShareCut bits (the adjoint kernels form every
// band's cut-off and its partials themselves).
constexpr uint32_t props_without_share(uint32_t props) { return props & 0x77777777u; }

template <int PROPS>
constexpr bool fast_of() {
  return (PROPS & kFast) != 0;
}

// Density of a component whose kind and properties are known at compile time.
// CUT (bands): see band_density; `cut` carries the shared cut-off within a node.
template <int KIND, int PROPS, int CUT = kCutOwn>
__device__ __forceinline__ double density_k(ParamPtr P, uint32_t flags, const Node& nd,
                                            const FeatureDir& fd, double& cut) {
  if constexpr (KIND == ZODI_INTERSTELLAR) {
    return P[6];
  } else {
    constexpr bool FAST = fast_of<PROPS>();
    const CompFrame f =
        comp_frame<PROPS & 3>(P, flags, nd.X0, nd.X1, nd.X2, nd.R, nd.inv_R, nd.R2);
    if constexpr (KIND == ZODI_CLOUD) return cloud_density<FAST>(P, f);
    if constexpr (KIND == ZODI_BAND) return band_density<FAST, CUT>(P, flags, f, cut);
    if constexpr (KIND == ZODI_RING) return ring_density(P, f);
    if constexpr (KIND == ZODI_FEATURE) return feature_density(P, f, fd);
    if constexpr (KIND == ZODI_FAN) return fan_density(P, f);
    if constexpr (KIND == ZODI_COMET) return comet_density(P, f);
    if constexpr (KIND == ZODI_NARROW_BAND) return narrow_band_density(P, f);
    if constexpr (KIND == ZODI_BROAD_BAND) return broad_band_density(P, f);
    if constexpr (KIND == ZODI_RING_RRM) return P[10] * ring_density(P, f);
    if constexpr (KIND == ZODI_FEATURE_RRM) return P[12] * feature_density(P, f, fd);
  }
}

struct Ctx {
  ModelPtr M;
  Tables tb;
  int K;
  int sub;
  double solar;
  double calibration;
};

// Index of the first band among a group's kinds (-1 without one).
template <int... Ks>
constexpr int first_band() {
  constexpr int kinds[] = {Ks...};
  for (int j = 0; j < (int)sizeof...(Ks); ++j)
    if (kinds[j] == ZODI_BAND) return j;
  return -1;
}

// The cut-off mode (BandCut) of component J of a group with these properties.
template <uint32_t PROPS, int J, int... Ks>
constexpr int cut_mode() {
  constexpr int C = sizeof...(Ks);
  bool shared = false;
  for (int j = 0; j < C && j < 8; ++j) shared = shared || ((PROPS >> (4 * j)) & kShareCut);
  if ((PROPS >> (4 * J)) & kShareCut) return kCutReuse;
  return shared && J == first_band<Ks...>() ? kCutPublish : kCutOwn;
}

// PROPS: 4 bits per component of the group (see density_k), component J at bits 4J.
template <int INTEGRAND, int L, uint32_t PROPS, bool NOSCAT, int... Ks, int... J>
__device__ __forceinline__ void group_pass(Grp<Ks...>, std::integer_sequence<int, J...>,
                                           const Ctx& cx, const FwdArgs& a, const Ray& y, int g,
                                           int first, double& total) {
  constexpr int C = sizeof...(Ks);
  constexpr int kinds[C] = {Ks...};
  const ModelPtr M = cx.M;
  double h, mid;
  group_range(M, g, y, h, mid);
  const double T0 = M->group_T0[g];
  const double mdelta = -M->group_delta[g];
  const double us = T0 * cx.tb.inv_step;
  const double uo = cx.tb.T_first * cx.tb.inv_step;

  // The feature's centre line, once per ray.
  FeatureDir fdir[C];
  ((fdir[J] = is_feature(kinds[J]) ? feature_dir(M->comp[first + J], y.ex, y.ey)
                                   : FeatureDir{1.0, 0.0}),
   ...);
  bool scatter = false;
  if constexpr (INTEGRAND == ZODI_INTEGRAND_KELSALL && !NOSCAT) {
    ((scatter = scatter || (M->comp_flags[first + J] & ZODI_CF_SCATTER)), ...);
  }

  // A NaN temperature comes from a NaN position. The Kelsall densities are NaN
  // there too, so their groups skip the lookup's NaN pass-through; the interstellar
  // density is a constant and the other RRM kinds mask NaN radii to 0 (as the
  // reference's boolean masks do), so their groups keep it.
  constexpr bool kNanFix = ((kinds[J] >= ZODI_FAN && kinds[J] <= ZODI_BROAD_BAND) || ...);
  double acc[C] = {};
  if constexpr (L == 1) {
    for (int k = 0; k < cx.K; ++k) {
      const ModelPtr Mk = opaque(M);
      const double xk = Mk->gl_x[k];
      const double wk = Mk->gl_w[k];
      Node nd = node_geometry<kNanFix>(y, h, mid, xk, us, uo, mdelta, cx.tb);
      if constexpr (INTEGRAND == ZODI_INTEGRAND_KELSALL) {
        if (scatter) {
          nd.phi = phase_function(h * xk + mid, y.ux, y.uy, y.uz, nd.X0, nd.X1, nd.X2, nd.R,
                                  Mk->shared);
        }
      }
      const double wB = wk * nd.B;
      double cut = 0.0;  // a band's cut-off, for the bands that share it (in order)
      ((acc[J] = node_add(
            acc[J], node_term<INTEGRAND, NOSCAT>(
                        Mk->comp[first + J], Mk->comp_flags[first + J], nd, wk, wB,
                        density_k<kinds[J], (PROPS >> (4 * J)) & 15, cut_mode<PROPS, J, Ks...>()>(
                            Mk->comp[first + J], Mk->comp_flags[first + J], nd, fdir[J], cut),
                        Mk->shared[ZODI_S_SOLAR]))),
       ...);
    }
  } else {
    // L lanes evaluate L consecutive nodes at once; every lane then adds the L
    // terms in node order, so each lane's sums are those of one lane per ray, bit
    // for bit. Past the last node a lane repeats it, and its term is not added.
    for (int k0 = 0; k0 < cx.K; k0 += L) {
      const int k = k0 + cx.sub < cx.K ? k0 + cx.sub : cx.K - 1;
      const ModelPtr Mk = opaque(M);
      const double xk = Mk->gl_x[k];
      const double wk = Mk->gl_w[k];
      Node nd = node_geometry<kNanFix>(y, h, mid, xk, us, uo, mdelta, cx.tb);
      if constexpr (INTEGRAND == ZODI_INTEGRAND_KELSALL) {
        if (scatter) {
          nd.phi = phase_function(h * xk + mid, y.ux, y.uy, y.uz, nd.X0, nd.X1, nd.X2, nd.R,
                                  Mk->shared);
        }
      }
      const double wB = wk * nd.B;
      double cut = 0.0;
      NodeTerm t[C];
      ((t[J] = node_term<INTEGRAND, NOSCAT>(
            Mk->comp[first + J], Mk->comp_flags[first + J], nd, wk, wB,
            density_k<kinds[J], (PROPS >> (4 * J)) & 15, cut_mode<PROPS, J, Ks...>()>(
                Mk->comp[first + J], Mk->comp_flags[first + J], nd, fdir[J], cut),
            Mk->shared[ZODI_S_SOLAR])),
       ...);
      add_in_node_order<L>(acc, t, cx.K - k0);
    }
  }
  const bool write = y.active && cx.sub == 0;
  (
      [&] {
        const double e = finish<INTEGRAND, NOSCAT>(M->comp[first + J],
                                                   M->comp_flags[first + J], acc[J], h,
                                                   cx.calibration);
        if (a.out_mode == ZODI_OUT_COMPS) {
          if (write) a.out[(int64_t)(first + J) * a.n_rays + y.r] = e;
        } else {
          total = __dadd_rn(total, e);
        }
      }(),
      ...);
}

// First component of group G of a layout.
template <class... Gs>
constexpr int group_first(int G) {
  constexpr int sizes[] = {Gs::size...};
  int f = 0;
  for (int i = 0; i < G; ++i) f += sizes[i];
  return f;
}

// The property word of the components from `first` on (4 bits each).
constexpr uint32_t props_shift(uint32_t props, int first) {
  return first >= 8 ? 0u : props >> (4 * first);
}

template <class LAYOUT>
struct LayoutIntegrand;
template <int INTEGRAND, class... Gs>
struct LayoutIntegrand<Layout<INTEGRAND, Gs...>> {
  static constexpr int value = INTEGRAND;
};
template <class LAYOUT>
constexpr int layout_integrand() {
  return LayoutIntegrand<LAYOUT>::value;
}

// PROPS: 4 bits per component of the layout (see density_k), component c at bits 4c.
template <int L, uint32_t PROPS, bool NOSCAT, int INTEGRAND, class... Gs, int... G>
__device__ __forceinline__ void run_layout_groups(Layout<INTEGRAND, Gs...>,
                                                  std::integer_sequence<int, G...>, const Ctx& cx,
                                                  const FwdArgs& a, const Ray& y, double& total) {
  (group_pass<INTEGRAND, L, props_shift(PROPS, group_first<Gs...>(G)), NOSCAT>(
       Gs{}, std::make_integer_sequence<int, Gs::size>{}, cx, a, y, G, group_first<Gs...>(G),
       total),
   ...);
}

template <int L, uint32_t PROPS, bool NOSCAT, int INTEGRAND, class... Gs>
__device__ __forceinline__ void run_layout(Layout<INTEGRAND, Gs...> lay, const Ctx& cx,
                                           const FwdArgs& a, const Ray& y, double& total) {
  run_layout_groups<L, PROPS, NOSCAT>(lay, std::make_integer_sequence<int, sizeof...(Gs)>{}, cx,
                                      a, y, total);
}

// Component properties of a layout kernel's registry variant: every component is
// centred on the Sun except the Kelsall cloud (the first component of the DIRBE
// and Kelsall4 layouts), and the Kelsall components take their fast forms. At most
// 8 components (4 bits each).
template <class LAYOUT>
struct RegistryProps;
template <int INTEGRAND, class... Gs>
struct RegistryProps<Layout<INTEGRAND, Gs...>> {
  static constexpr int n = group_first<Gs...>(sizeof...(Gs));
  static_assert(n <= 8, "4 bits per component in 32 bits");
  static constexpr uint32_t value = [] {
    uint32_t m = 0;
    for (int c = 0; c < n; ++c) m |= uint32_t(kFrameCentred) << (4 * c);
    return m;
  }();
};
template <>
struct RegistryProps<LayoutDirbe> {  // cloud offset, 5 centred, all fast; band 3 shares band 1's cut-off
  static constexpr uint32_t value = (0x55555u << 4 | kFrameOffset | 0x444444u) | kShareCut << 12;
};
template <>
struct RegistryProps<LayoutKelsall4> {  // cloud offset, 3 centred, all fast; band 3 shares band 1's cut-off
  static constexpr uint32_t value = (0x555u << 4 | kFrameOffset | 0x4444u) | kShareCut << 12;
};

// Whether a model block meets the preconditions of a layout kernel with these
// properties: every component fixed as centred carries ZODI_CF_CENTERED (the
// offset path is exact for any component); a fast band has p = 4
// (ZODI_CF_P_IS_4) and a fast cloud gamma >= 1/2; a band with kShareCut and the
// first band of its group are both centred and fast and hold the same
// 1/delta_r^2.
inline bool props_match(const zodi_model* m, uint32_t props) {
  for (int c = 0; c < m->n_comps && c < 8; ++c) {
    const uint32_t pc = (props >> (4 * c)) & 15u;
    if ((pc & 3u) == kFrameCentred && !(m->comp_flags[c] & ZODI_CF_CENTERED)) return false;
    if (pc & kFast) {
      const int kind = m->comp_kind[c];
      if (kind == ZODI_BAND && !(m->comp_flags[c] & ZODI_CF_P_IS_4)) return false;
      if (kind == ZODI_CLOUD && !(m->comp[c][9] >= 0.5)) return false;
    }
    if (pc & kShareCut) {
      int g = 0;
      while (g < m->n_groups && c >= m->group_first[g] + m->group_count[g]) ++g;
      if (g == m->n_groups) return false;
      int b = m->group_first[g];
      while (b < c && m->comp_kind[b] != ZODI_BAND) ++b;
      if (b >= c || m->comp_kind[c] != ZODI_BAND) return false;
      const uint32_t pb = (props >> (4 * b)) & 15u;
      const uint32_t need = kFast | kFrameCentred;
      if ((pc & 7u) != need || (pb & 7u) != need) return false;
      if (!(m->comp_flags[b] & ZODI_CF_CENTERED)) return false;
      if (m->comp[c][ZODI_P_D0 + 6] != m->comp[b][ZODI_P_D0 + 6]) return false;
    }
  }
  return true;
}

__device__ __forceinline__ Ctx make_ctx(ModelPtr M, const double* sTab, int sub) {
  Ctx cx;
  cx.M = M;
  cx.tb.tab = sTab;
  cx.tb.T_first = M->tab_T[0];
  cx.tb.T_last = M->tab_T[M->n_temps - 1];
  cx.tb.inv_step = M->tab_inv_step;
  cx.tb.n_t = M->n_temps;
  cx.K = M->n_nodes;
  cx.sub = sub;
  cx.solar = M->shared[ZODI_S_SOLAR];
  cx.calibration = M->shared[ZODI_S_CALIBRATION];
  return cx;
}


}  // namespace zodi
