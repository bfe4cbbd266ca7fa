// Fused FP64 line-of-sight kernels for MI355X (gfx950) and their C ABI.
//
// One ray per lane (or per `L` adjacent lanes, which then split the
// Gauss-Legendre nodes and combine with cross-lane shuffles). For each ray the
// per-ray sphere-intersection setup runs once; then, per group of components
// sharing a line-of-sight range and temperature law, the node loop evaluates
// position, 1/R, temperature and the Planck lookup once and every component
// density of the group on top, accumulating in registers. The model block is
// read through the constant address space with wave-uniform addresses (scalar
// loads into SGPRs); the Planck table is staged in LDS.
//
// Two kernel families:
//  * forward_layout<Layout, BLOCK, L>: the component kinds of every group are
//    template parameters, so each slot carries only its own density code
//    (registry layouts: DIRBE/planck13, planck15/18/odegard, RRM);
//  * forward_generic<BLOCK, L>: any layout the model block can describe; one
//    density switch per slot, per-slot accumulators in LDS.
//
// Reference: zodipy/model.py:253-277 and the functions it calls (see zodi_hip.h).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <string>
#include <utility>

#include "zodi_device.hpp"
#include "zodi_hip.h"

namespace zodi {

namespace {
thread_local std::string g_last_error;

zodi_status fail(zodi_status code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
zodi_status fail(zodi_status code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

zodi_status check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    return fail(ZODI_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  }
  return ZODI_OK;
}

int cu_count() {
  static int counts[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (counts[dev] == 0) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        n <= 0) {
      n = 256;
    }
    counts[dev] = n;
  }
  return counts[dev];
}

zodi_status validate_model(const zodi_model* m) {
  if (m == nullptr) return fail(ZODI_ERR_ARG, "model_host is null");
  if (m->abi_version != ZODI_ABI_VERSION)
    return fail(ZODI_ERR_MODEL, "model abi_version %d, library %d", m->abi_version,
                ZODI_ABI_VERSION);
  if (m->integrand != ZODI_INTEGRAND_KELSALL && m->integrand != ZODI_INTEGRAND_RRM)
    return fail(ZODI_ERR_MODEL, "unknown integrand %d", m->integrand);
  if (m->n_comps < 1 || m->n_comps > ZODI_MAX_COMPS)
    return fail(ZODI_ERR_MODEL, "n_comps %d outside [1, %d]", m->n_comps, ZODI_MAX_COMPS);
  if (m->n_groups < 1 || m->n_groups > ZODI_MAX_GROUPS)
    return fail(ZODI_ERR_MODEL, "n_groups %d outside [1, %d]", m->n_groups, ZODI_MAX_GROUPS);
  if (m->n_nodes < 1 || m->n_nodes > ZODI_MAX_NODES)
    return fail(ZODI_ERR_MODEL, "n_nodes %d outside [1, %d]", m->n_nodes, ZODI_MAX_NODES);
  if (m->n_temps < 2 || m->n_temps > ZODI_N_TEMPS)
    return fail(ZODI_ERR_MODEL, "n_temps %d outside [2, %d]", m->n_temps, ZODI_N_TEMPS);
  int next = 0;
  for (int g = 0; g < m->n_groups; ++g) {
    if (m->group_first[g] != next)
      return fail(ZODI_ERR_MODEL, "group %d starts at %d, expected %d", g, m->group_first[g],
                  next);
    if (m->group_count[g] < 1 || m->group_count[g] > kMaxSlots)
      return fail(ZODI_ERR_MODEL, "group %d has %d components (1..%d)", g, m->group_count[g],
                  kMaxSlots);
    next += m->group_count[g];
  }
  if (next != m->n_comps)
    return fail(ZODI_ERR_MODEL, "groups cover %d components, model has %d", next, m->n_comps);
  for (int c = 0; c < m->n_comps; ++c) {
    if (m->comp_kind[c] < ZODI_CLOUD || m->comp_kind[c] > ZODI_FEATURE_RRM)
      return fail(ZODI_ERR_MODEL, "component %d has unknown kind %d", c, m->comp_kind[c]);
  }
  return ZODI_OK;
}

}  // namespace

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

// Planck table to LDS as (T_j, slope_j, B_j, 0) entries.
__device__ __forceinline__ void load_table(ModelPtr M, double* sTab) {
  const int n = M->n_temps;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    sTab[4 * i] = M->tab_T[i];
    sTab[4 * i + 1] = M->tab_slope[i];
    sTab[4 * i + 2] = M->tab_B[i];
    sTab[4 * i + 3] = 0.0;
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
  double X0, X1, X2, R, inv_R, B, phi;
};

struct Tables {
  const double* tab;
  double T_first;
  double T_last;
  double inv_step;
  int n_t;
};

__device__ __forceinline__ Node node_geometry(const Ray& y, double h, double mid, double xk,
                                              double T0, double mdelta, const Tables& tb) {
  Node nd;
  const double s_los = h * xk + mid;
  nd.X0 = s_los * y.ux + y.ox;
  nd.X1 = s_los * y.uy + y.oy;
  nd.X2 = s_los * y.uz + y.oz;
  norm_and_inverse(nd.X0 * nd.X0 + nd.X1 * nd.X1 + nd.X2 * nd.X2, nd.R, nd.inv_R);
  // T = T0 R^-delta (zodipy/blackbody.py:16-30), B = interp(T) (zodipy/brightness.py:48)
  const double T = T0 * pow_radius(nd.R, mdelta);
  nd.B = planck_lookup(T, tb.tab, tb.n_t, tb.T_first, tb.T_last, tb.inv_step);
  nd.phi = 0.0;
  return nd;
}

// Node contribution of one component to its accumulator.
// Kelsall (zodipy/brightness.py:21-56): (1 - a) eps B n, plus a F/R^2 Phi n with
// scattering; without scattering the constant (1 - a) eps is applied once to the
// node sum (finish). RRM (zodipy/brightness.py:59-83): B n, calibration applied
// in finish.
template <int INTEGRAND>
__device__ __forceinline__ double node_term(ParamPtr P, uint32_t flags, const Node& nd,
                                            double wk, double wB, double dens, double solar) {
  if constexpr (INTEGRAND == ZODI_INTEGRAND_KELSALL) {
    if (flags & ZODI_CF_SCATTER) {
      double em = P[ZODI_P_ONE_MINUS_ALBEDO] * (P[ZODI_P_EMISSIVITY] * nd.B);
      em += (P[ZODI_P_ALBEDO] * (solar * (nd.inv_R * nd.inv_R))) * nd.phi;
      return wk * (em * dens);
    }
  }
  return wB * dens;
}

// Integrated value of one component from its node sum; h = (stop - start)/2.
template <int INTEGRAND>
__device__ __forceinline__ double finish(ParamPtr P, uint32_t flags, double acc, double h,
                                         double calibration) {
  if constexpr (INTEGRAND == ZODI_INTEGRAND_KELSALL) {
    if (flags & ZODI_CF_SCATTER) return __dmul_rn(acc, h);
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

// Density of a component whose kind is known at compile time.
template <int KIND>
__device__ __forceinline__ double density_k(ParamPtr P, uint32_t flags, const Node& nd,
                                            double earth_lon) {
  if constexpr (KIND == ZODI_INTERSTELLAR) {
    return P[6];
  } else {
    const CompFrame f = comp_frame(P, flags, nd.X0, nd.X1, nd.X2, nd.R, nd.inv_R);
    if constexpr (KIND == ZODI_CLOUD) return cloud_density(P, f);
    if constexpr (KIND == ZODI_BAND) return band_density(P, flags, f);
    if constexpr (KIND == ZODI_RING) return ring_density(P, f);
    if constexpr (KIND == ZODI_FEATURE) return feature_density(P, f, earth_lon);
    if constexpr (KIND == ZODI_FAN) return fan_density(P, f);
    if constexpr (KIND == ZODI_COMET) return comet_density(P, f);
    if constexpr (KIND == ZODI_NARROW_BAND) return narrow_band_density(P, f);
    if constexpr (KIND == ZODI_BROAD_BAND) return broad_band_density(P, f);
    if constexpr (KIND == ZODI_RING_RRM) return P[10] * ring_density(P, f);
    if constexpr (KIND == ZODI_FEATURE_RRM) return P[12] * feature_density(P, f, earth_lon);
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

template <int INTEGRAND, int L, int... Ks, int... J>
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

  double elon[C];
  ((elon[J] = is_feature(kinds[J]) ? atan2(y.ey - M->comp[first + J][ZODI_P_Y0],
                                           y.ex - M->comp[first + J][ZODI_P_X0])
                                   : 0.0),
   ...);
  bool scatter = false;
  if constexpr (INTEGRAND == ZODI_INTEGRAND_KELSALL) {
    ((scatter = scatter || (M->comp_flags[first + J] & ZODI_CF_SCATTER)), ...);
  }

  double acc[C] = {};
  for (int k = cx.sub; k < cx.K; k += L) {
    const ModelPtr Mk = opaque(M);
    const double xk = Mk->gl_x[k];
    const double wk = Mk->gl_w[k];
    Node nd = node_geometry(y, h, mid, xk, T0, mdelta, cx.tb);
    if constexpr (INTEGRAND == ZODI_INTEGRAND_KELSALL) {
      if (scatter) {
        nd.phi = phase_function(h * xk + mid, y.ux, y.uy, y.uz, nd.X0, nd.X1, nd.X2, nd.R,
                                Mk->shared);
      }
    }
    const double wB = wk * nd.B;
    ((acc[J] += node_term<INTEGRAND>(
          Mk->comp[first + J], Mk->comp_flags[first + J], nd, wk, wB,
          density_k<kinds[J]>(Mk->comp[first + J], Mk->comp_flags[first + J], nd, elon[J]),
          Mk->shared[ZODI_S_SOLAR])),
     ...);
  }
  if constexpr (L > 1) {
#pragma unroll
    for (int j = 0; j < C; ++j) {
#pragma unroll
      for (int off = L / 2; off > 0; off >>= 1) acc[j] += __shfl_xor(acc[j], off, L);
    }
  }
  const bool write = y.active && cx.sub == 0;
#pragma unroll
  for (int j = 0; j < C; ++j) {
    const double e = finish<INTEGRAND>(M->comp[first + j], M->comp_flags[first + j], acc[j], h,
                                       cx.calibration);
    if (a.out_mode == ZODI_OUT_COMPS) {
      if (write) a.out[(int64_t)(first + j) * a.n_rays + y.r] = e;
    } else {
      total = __dadd_rn(total, e);
    }
  }
}

template <int L, int INTEGRAND, class... Gs>
__device__ __forceinline__ void run_layout(Layout<INTEGRAND, Gs...>, const Ctx& cx,
                                           const FwdArgs& a, const Ray& y, double& total) {
  int g = 0;
  int first = 0;
  ((group_pass<INTEGRAND, L>(Gs{}, std::make_integer_sequence<int, Gs::size>{}, cx, a, y, g,
                             first, total),
    ++g, first += Gs::size),
   ...);
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

template <class LAYOUT, int BLOCK, int L>
__global__ __launch_bounds__(BLOCK) void forward_layout(FwdArgs a) {
  __shared__ double sTab[4 * ZODI_N_TEMPS];
  const ModelPtr M = as_const(a.m);
  stage_math_tables();
  load_table(M, sTab);
  const Ctx cx = make_ctx(M, sTab, threadIdx.x % L);
  constexpr int RAYS = BLOCK / L;
  const int local = threadIdx.x / L;
  for (int64_t tile = blockIdx.x; tile < a.n_tiles; tile += gridDim.x) {
    const Ray y = load_ray<RAYS>(a, tile, local);
    double total = 0.0;
    run_layout<L>(LAYOUT{}, cx, a, y, total);
    if (a.out_mode == ZODI_OUT_SUM && y.active && cx.sub == 0) a.out[y.r] = total;
  }
}

// Registry layouts.
using LayoutDirbe = Layout<ZODI_INTEGRAND_KELSALL, Grp<ZODI_CLOUD, ZODI_BAND, ZODI_BAND, ZODI_BAND>,
                           Grp<ZODI_RING>, Grp<ZODI_FEATURE>>;
using LayoutKelsall4 =
    Layout<ZODI_INTEGRAND_KELSALL, Grp<ZODI_CLOUD, ZODI_BAND, ZODI_BAND, ZODI_BAND>>;
using LayoutRRM =
    Layout<ZODI_INTEGRAND_RRM, Grp<ZODI_FAN>, Grp<ZODI_COMET>,
           Grp<ZODI_NARROW_BAND, ZODI_NARROW_BAND, ZODI_BROAD_BAND>, Grp<ZODI_INTERSTELLAR>,
           Grp<ZODI_RING_RRM>, Grp<ZODI_FEATURE_RRM>>;

// ------------------------------------------------------------ generic kernel
template <int INTEGRAND>
__device__ __forceinline__ void generic_groups(const Ctx& cx, const FwdArgs& a, const Ray& y,
                                               double (*sAcc)[512], int tid, int L,
                                               double& total) {
  const ModelPtr M = cx.M;
  const int n_groups = M->n_groups;
  for (int g = 0; g < n_groups; ++g) {
    const int first = M->group_first[g];
    const int cnt = M->group_count[g];
    double h, mid;
    group_range(M, g, y, h, mid);
    const double T0 = M->group_T0[g];
    const double mdelta = -M->group_delta[g];
    bool scatter = false;
    for (int j = 0; j < cnt; ++j) {
      sAcc[j][tid] = 0.0;
      scatter = scatter || (M->comp_flags[first + j] & ZODI_CF_SCATTER);
    }
    for (int k = cx.sub; k < cx.K; k += L) {
      const double xk = M->gl_x[k];
      const double wk = M->gl_w[k];
      Node nd = node_geometry(y, h, mid, xk, T0, mdelta, cx.tb);
      if (INTEGRAND == ZODI_INTEGRAND_KELSALL && scatter) {
        nd.phi = phase_function(h * xk + mid, y.ux, y.uy, y.uz, nd.X0, nd.X1, nd.X2, nd.R,
                                M->shared);
      }
      const double wB = wk * nd.B;
#pragma unroll 1
      for (int j = 0; j < cnt; ++j) {
        const int c = first + j;
        const ParamPtr P = M->comp[c];
        const uint32_t flags = M->comp_flags[c];
        const int kind = M->comp_kind[c];
        const double elon =
            is_feature(kind) ? atan2(y.ey - P[ZODI_P_Y0], y.ex - P[ZODI_P_X0]) : 0.0;
        const double dens = density(kind, P, flags, nd.X0, nd.X1, nd.X2, nd.R, nd.inv_R, elon);
        sAcc[j][tid] += node_term<INTEGRAND>(P, flags, nd, wk, wB, dens, cx.solar);
      }
    }
    for (int j = 0; j < cnt; ++j) {
      double acc = sAcc[j][tid];
      for (int off = L / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, L);
      const double e = finish<INTEGRAND>(M->comp[first + j], M->comp_flags[first + j], acc, h,
                                         cx.calibration);
      if (a.out_mode == ZODI_OUT_COMPS) {
        if (y.active && cx.sub == 0) a.out[(int64_t)(first + j) * a.n_rays + y.r] = e;
      } else {
        total = __dadd_rn(total, e);
      }
    }
  }
}

template <int BLOCK, int L>
__global__ __launch_bounds__(BLOCK) void forward_generic(FwdArgs a) {
  __shared__ double sTab[4 * ZODI_N_TEMPS];
  __shared__ double sAcc[kMaxSlots][512];
  const ModelPtr M = as_const(a.m);
  stage_math_tables();
  load_table(M, sTab);
  const int tid = threadIdx.x;
  const Ctx cx = make_ctx(M, sTab, tid % L);
  constexpr int RAYS = BLOCK / L;
  const int local = tid / L;
  const bool kelsall = M->integrand == ZODI_INTEGRAND_KELSALL;
  for (int64_t tile = blockIdx.x; tile < a.n_tiles; tile += gridDim.x) {
    const Ray y = load_ray<RAYS>(a, tile, local);
    double total = 0.0;
    if (kelsall) {
      generic_groups<ZODI_INTEGRAND_KELSALL>(cx, a, y, sAcc, tid, L, total);
    } else {
      generic_groups<ZODI_INTEGRAND_RRM>(cx, a, y, sAcc, tid, L, total);
    }
    if (a.out_mode == ZODI_OUT_SUM && y.active && cx.sub == 0) a.out[y.r] = total;
  }
}

// Per-group start/stop distances (zodipy/line_of_sight.py:88-105).
__global__ __launch_bounds__(256) void line_of_sight_kernel(FwdArgs a, double* __restrict__ start,
                                                            double* __restrict__ stop) {
  const ModelPtr M = as_const(a.m);
  const int64_t n = a.n_rays;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int64_t s = a.obs_mode == ZODI_OBS_PER_RAY ? r : find_segment(a.seg, a.n_seg, r);
  const uint32_t bits = a.beyond[a.obs_mode == ZODI_OBS_PER_RAY ? 0 : s];
  const int64_t os = a.obs_stride;
  const RayGeom G = ray_geometry(a.u[r], a.u[n + r], a.u[2 * n + r], a.obs[s], a.obs[os + s],
                                 a.obs[2 * os + s]);
  for (int g = 0; g < M->n_groups; ++g) {
    start[(int64_t)g * n + r] =
        ((bits >> (2 * g)) & 1u) ? kEps : sphere_distance(G, M->group_r_in[g]);
    stop[(int64_t)g * n + r] =
        ((bits >> (2 * g + 1)) & 1u) ? kEps : sphere_distance(G, M->group_r_out[g]);
  }
}

// Component densities at heliocentric points (zodipy/number_density.py:47-420).
__global__ __launch_bounds__(256) void number_density_kernel(const zodi_model* model,
                                                             const double* __restrict__ X,
                                                             int64_t n, const double* earth,
                                                             double* __restrict__ out) {
  const ModelPtr M = as_const(model);
  stage_math_tables();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x0 = X[i], x1 = X[n + i], x2 = X[2 * n + i];
  double R, inv_R;
  norm_and_inverse(x0 * x0 + x1 * x1 + x2 * x2, R, inv_R);
  const double ex = earth[0], ey = earth[1];
  for (int c = 0; c < M->n_comps; ++c) {
    const ParamPtr P = M->comp[c];
    const int kind = M->comp_kind[c];
    const double elon = is_feature(kind) ? atan2(ey - P[ZODI_P_Y0], ex - P[ZODI_P_X0]) : 0.0;
    out[(int64_t)c * n + i] = density(kind, P, M->comp_flags[c], x0, x1, x2, R, inv_R, elon);
  }
}

namespace {

using FwdKernel = void (*)(FwdArgs);

enum LayoutId { kLayoutGeneric = 0, kLayoutDirbe = 1, kLayoutKelsall4 = 2, kLayoutRRM = 3 };

bool layout_is(const zodi_model* m, int integrand, std::initializer_list<std::initializer_list<int>> groups) {
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

LayoutId match_layout(const zodi_model* m) {
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
  return kLayoutGeneric;
}

template <class LAYOUT, int BLOCK>
FwdKernel pick_layout_lanes(int L) {
  switch (L) {
    case 1:
      return forward_layout<LAYOUT, BLOCK, 1>;
    case 2:
      return forward_layout<LAYOUT, BLOCK, 2>;
    case 4:
      return forward_layout<LAYOUT, BLOCK, 4>;
    case 8:
      return forward_layout<LAYOUT, BLOCK, 8>;
    default:
      return nullptr;
  }
}

template <class LAYOUT>
FwdKernel pick_layout(int block, int L) {
  switch (block) {
    case 64:
      return pick_layout_lanes<LAYOUT, 64>(L);
    case 128:
      return pick_layout_lanes<LAYOUT, 128>(L);
    case 256:
      return pick_layout_lanes<LAYOUT, 256>(L);
    case 512:
      return pick_layout_lanes<LAYOUT, 512>(L);
    default:
      return nullptr;
  }
}

template <int BLOCK>
FwdKernel pick_generic_lanes(int L) {
  switch (L) {
    case 1:
      return forward_generic<BLOCK, 1>;
    case 2:
      return forward_generic<BLOCK, 2>;
    case 4:
      return forward_generic<BLOCK, 4>;
    case 8:
      return forward_generic<BLOCK, 8>;
    default:
      return nullptr;
  }
}

FwdKernel pick_generic(int block, int L) {
  switch (block) {
    case 64:
      return pick_generic_lanes<64>(L);
    case 128:
      return pick_generic_lanes<128>(L);
    case 256:
      return pick_generic_lanes<256>(L);
    case 512:
      return pick_generic_lanes<512>(L);
    default:
      return nullptr;
  }
}

FwdKernel pick_forward(const zodi_model* m, int block, int L, int variant) {
  const LayoutId id = variant == 0 ? kLayoutGeneric : match_layout(m);
  switch (id) {
    case kLayoutDirbe:
      return pick_layout<LayoutDirbe>(block, L);
    case kLayoutKelsall4:
      return pick_layout<LayoutKelsall4>(block, L);
    case kLayoutRRM:
      return pick_layout<LayoutRRM>(block, L);
    default:
      return pick_generic(block, L);
  }
}

zodi_status fill_args(FwdArgs& a, const zodi_model* model_host, const zodi_model* model_dev,
                      const double* u, int64_t n_rays, const int64_t* seg_offsets, int32_t n_seg,
                      const double* obs, const double* earth, int32_t obs_mode,
                      const uint32_t* beyond) {
  zodi_status st = validate_model(model_host);
  if (st != ZODI_OK) return st;
  if (model_dev == nullptr) return fail(ZODI_ERR_ARG, "model_dev is null");
  if (n_rays < 0) return fail(ZODI_ERR_ARG, "n_rays %lld < 0", (long long)n_rays);
  if (u == nullptr || obs == nullptr || earth == nullptr || beyond == nullptr)
    return fail(ZODI_ERR_ARG, "null input pointer");
  if (obs_mode == ZODI_OBS_PER_SEGMENT) {
    if (seg_offsets == nullptr || n_seg < 1)
      return fail(ZODI_ERR_ARG, "per-segment mode needs seg_offsets and n_seg >= 1");
  } else if (obs_mode != ZODI_OBS_PER_RAY) {
    return fail(ZODI_ERR_ARG, "unknown obs_mode %d", obs_mode);
  }
  a.m = model_dev;
  a.u = u;
  a.n_rays = n_rays;
  a.seg = seg_offsets;
  a.n_seg = n_seg;
  a.obs_mode = obs_mode;
  a.obs = obs;
  a.earth = earth;
  a.obs_stride = obs_mode == ZODI_OBS_PER_RAY ? n_rays : n_seg;
  a.beyond = beyond;
  a.out = nullptr;
  a.out_mode = 0;
  a.pad = 0;
  a.n_tiles = 0;
  return ZODI_OK;
}

}  // namespace
}  // namespace zodi

using namespace zodi;

extern "C" {

int32_t zodi_abi_version(void) { return ZODI_ABI_VERSION; }

int64_t zodi_model_size(void) { return (int64_t)sizeof(zodi_model); }

const char* zodi_last_error(void) { return g_last_error.c_str(); }

zodi_status zodi_forward(const zodi_model* model_host, const zodi_model* model_dev,
                         const double* u, int64_t n_rays, const int64_t* seg_offsets,
                         int32_t n_seg, const double* obs, const double* earth, int32_t obs_mode,
                         const uint32_t* beyond, double* out, int32_t out_mode,
                         const zodi_launch* launch, void* stream) {
  FwdArgs a;
  zodi_status st = fill_args(a, model_host, model_dev, u, n_rays, seg_offsets, n_seg, obs, earth,
                             obs_mode, beyond);
  if (st != ZODI_OK) return st;
  if (out == nullptr) return fail(ZODI_ERR_ARG, "out is null");
  if (out_mode != ZODI_OUT_COMPS && out_mode != ZODI_OUT_SUM)
    return fail(ZODI_ERR_ARG, "unknown out_mode %d", out_mode);
  if (n_rays == 0) return ZODI_OK;

  int block = 256, L = 1, per_cu = 0, variant = 1;
  if (launch != nullptr) {
    if (launch->block) block = launch->block;
    if (launch->lanes_per_ray) L = launch->lanes_per_ray;
    per_cu = launch->blocks_per_cu;
    variant = launch->variant;
  }
  if (variant < 0 || variant > 1) return fail(ZODI_ERR_ARG, "unknown variant %d", variant);
  if (L > block) return fail(ZODI_ERR_ARG, "lanes_per_ray %d > block %d", L, block);
  FwdKernel k = pick_forward(model_host, block, L, variant);
  if (k == nullptr)
    return fail(ZODI_ERR_ARG, "unsupported launch: block %d, lanes_per_ray %d", block, L);
  if (per_cu < 0 || per_cu > 64)
    return fail(ZODI_ERR_ARG, "blocks_per_cu %d outside [0, 64]", per_cu);

  const int64_t rays_per_block = block / L;
  const int64_t n_tiles = (n_rays + rays_per_block - 1) / rays_per_block;
  int64_t grid = n_tiles;
  if (per_cu > 0) grid = std::min<int64_t>(grid, (int64_t)per_cu * cu_count());
  if (grid > 0x7fffffff) return fail(ZODI_ERR_TOO_LARGE, "grid of %lld blocks", (long long)grid);
  a.out = out;
  a.out_mode = out_mode;
  a.n_tiles = n_tiles;
  hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(block), 0, (hipStream_t)stream, a);
  return check_hip(hipGetLastError(), "forward launch");
}

zodi_status zodi_line_of_sight(const zodi_model* model_host, const zodi_model* model_dev,
                               const double* u, int64_t n_rays, const int64_t* seg_offsets,
                               int32_t n_seg, const double* obs, int32_t obs_mode,
                               const uint32_t* beyond, double* start, double* stop,
                               void* stream) {
  FwdArgs a;
  zodi_status st = fill_args(a, model_host, model_dev, u, n_rays, seg_offsets, n_seg, obs, obs,
                             obs_mode, beyond);
  if (st != ZODI_OK) return st;
  if (start == nullptr || stop == nullptr) return fail(ZODI_ERR_ARG, "null output pointer");
  if (n_rays == 0) return ZODI_OK;
  const int64_t grid = (n_rays + 255) / 256;
  if (grid > 0x7fffffff) return fail(ZODI_ERR_TOO_LARGE, "too many rays");
  hipLaunchKernelGGL(line_of_sight_kernel, dim3((unsigned)grid), dim3(256), 0,
                     (hipStream_t)stream, a, start, stop);
  return check_hip(hipGetLastError(), "line_of_sight launch");
}

zodi_status zodi_number_density(const zodi_model* model_host, const zodi_model* model_dev,
                                const double* X, int64_t n_points, const double* earth,
                                double* out, void* stream) {
  zodi_status st = validate_model(model_host);
  if (st != ZODI_OK) return st;
  if (model_dev == nullptr || X == nullptr || earth == nullptr || out == nullptr)
    return fail(ZODI_ERR_ARG, "null pointer");
  if (n_points < 0) return fail(ZODI_ERR_ARG, "n_points < 0");
  if (n_points == 0) return ZODI_OK;
  const int64_t grid = (n_points + 255) / 256;
  if (grid > 0x7fffffff) return fail(ZODI_ERR_TOO_LARGE, "too many points");
  hipLaunchKernelGGL(number_density_kernel, dim3((unsigned)grid), dim3(256), 0,
                     (hipStream_t)stream, model_dev, X, n_points, earth, out);
  return check_hip(hipGetLastError(), "number_density launch");
}

}  // extern "C"
