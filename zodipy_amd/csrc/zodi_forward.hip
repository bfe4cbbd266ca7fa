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
#include <string>
#include <utility>

#include "zodi_common.hpp"

namespace zodi {

namespace {
thread_local std::string g_last_error;
}  // namespace




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


// Minimum waves per SIMD the register allocator must leave room for, per layout
// (1 = no constraint). Measured on MI355X: DIRBE and Kelsall4 run at 5 (93 and
// 84 VGPRs in their registry variants, 89 and 79 without scattering code, no
// spill; DIRBE 2.7 % faster than at 4, and 6 waves with a spill gains nothing,
// 0.8 % slower without scattering code); RRM at 4 (128 VGPRs with a small
// spill), 15 % faster than its unconstrained 190 VGPRs / 2 waves and 3.7 %
// faster than 3 waves without the spill. The kind switch (forward_generic)
// runs at 3. The split-lane variants (L > 1) run only where the
// launch is small, so they ask for 2 and keep their registers (DIRBE: 98 VGPRs
// instead of 96 with a 16-byte spill; 3072 rays at 8 lanes 36.2 -> 34.6 us).
template <class LAYOUT>
struct ForwardOccupancy {
  static constexpr int value = 1;
};
template <>
struct ForwardOccupancy<LayoutDirbe> {
  static constexpr int value = 5;
};
template <>
struct ForwardOccupancy<LayoutRRM> {
  static constexpr int value = 4;
};
template <>
struct ForwardOccupancy<LayoutRRMSplit> {
  static constexpr int value = 4;
};
template <>
struct ForwardOccupancy<LayoutKelsall4> {
  static constexpr int value = 5;
};

// SPEC: every component's frame path and form fixed at compile time to the
// registry models' (RegistryProps); the host runs it only for blocks that meet
// its preconditions (props_match). SHARE (with SPEC): the bands that share a
// radial cut-off reuse it (kShareCut), for blocks whose delta_r values allow it
// (the registry parameters; a fit that moves them runs SPEC alone). NOSCAT (with
// SHARE): no component scatters, so no scattering code is compiled in (no flag
// test per component and node; 1.5-2 % faster on the DIRBE 12 um, planck18 and
// six-component 857 GHz batches).
template <class LAYOUT, int BLOCK, int L, bool SPEC, bool SHARE = false, bool NOSCAT = false>
__global__ __launch_bounds__(BLOCK)
__attribute__((amdgpu_waves_per_eu(L == 1 ? ForwardOccupancy<LAYOUT>::value : 2))) void
forward_layout(FwdArgs a) {
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
    constexpr uint32_t kProps =
        !SPEC ? 0u
              : (SHARE ? RegistryProps<LAYOUT>::value
                       : props_without_share(RegistryProps<LAYOUT>::value));
    run_layout<L, kProps, NOSCAT>(LAYOUT{}, cx, a, y, total);
    if (a.out_mode == ZODI_OUT_SUM && y.active && cx.sub == 0) a.out[y.r] = total;
  }
}

// ------------------------------------------------------------ generic kernel
template <int INTEGRAND, int BLOCK>
__device__ __forceinline__ void generic_groups(const Ctx& cx, const FwdArgs& a, const Ray& y,
                                               double (*sAcc)[BLOCK], int tid, int L,
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
    const double us = T0 * cx.tb.inv_step;
    const double uo = cx.tb.T_first * cx.tb.inv_step;
    bool scatter = false;
    for (int j = 0; j < cnt; ++j) {
      sAcc[j][tid] = 0.0;
      scatter = scatter || (M->comp_flags[first + j] & ZODI_CF_SCATTER);
    }
    // L lanes evaluate L consecutive nodes; each lane adds the L terms in node
    // order (as forward_layout does), past the last node repeating it unadded.
    for (int k0 = 0; k0 < cx.K; k0 += L) {
      const int k = k0 + cx.sub < cx.K ? k0 + cx.sub : cx.K - 1;
      const int left = cx.K - k0;
      const double xk = M->gl_x[k];
      const double wk = M->gl_w[k];
      Node nd = node_geometry(y, h, mid, xk, us, uo, mdelta, cx.tb);
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
        const FeatureDir fd = is_feature(kind) ? feature_dir(P, y.ex, y.ey) : FeatureDir{1.0, 0.0};
        const double dens =
            density(kind, P, flags, nd.X0, nd.X1, nd.X2, nd.R, nd.inv_R, nd.R2, fd);
        const NodeTerm t = node_term<INTEGRAND>(P, flags, nd, wk, wB, dens, cx.solar);
        double acc = sAcc[j][tid];
        if (L == 1) {
          acc = node_add(acc, t);
        } else {
          for (int s = 0; s < L && s < left; ++s)
            acc = node_add(acc, NodeTerm{__shfl(t.a, s, L), __shfl(t.b, s, L)});
        }
        sAcc[j][tid] = acc;
      }
    }
    for (int j = 0; j < cnt; ++j) {
      const double acc = sAcc[j][tid];
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

// Three waves per SIMD (167-168 VGPRs, up to a 16-byte spill with split lanes):
// 22-30 % faster than the unconstrained 169-175 VGPRs at two (DIRBE through the
// kind switch on 196 608 to 3.1 M rays; the same on 49 152).
template <int BLOCK, int L>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3))) void
forward_generic(FwdArgs a) {
  __shared__ double sTab[4 * ZODI_N_TEMPS];
  __shared__ double sAcc[kMaxSlots][BLOCK];  // one node sum per lane and component slot
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
  const double R2 = x0 * x0 + x1 * x1 + x2 * x2;
  double R, inv_R;
  norm_and_inverse(R2, R, inv_R);
  if (R2 == 0.0) {  // the Sun's centre (a grid may hold it): R = 0 as sqrt gives it
    R = 0.0;
    inv_R = __builtin_inf();
  }
  const double ex = earth[0], ey = earth[1];
  for (int c = 0; c < M->n_comps; ++c) {
    const ParamPtr P = M->comp[c];
    const int kind = M->comp_kind[c];
    const FeatureDir fd = is_feature(kind) ? feature_dir(P, ex, ey) : FeatureDir{1.0, 0.0};
    out[(int64_t)c * n + i] = density(kind, P, M->comp_flags[c], x0, x1, x2, R, inv_R, R2, fd);
  }
}

namespace {

using FwdKernel = void (*)(FwdArgs);

template <class LAYOUT, int BLOCK, bool SPEC>
FwdKernel pick_layout_lanes(int L) {
  switch (L) {
    case 1:
      return forward_layout<LAYOUT, BLOCK, 1, SPEC>;
    case 2:
      return forward_layout<LAYOUT, BLOCK, 2, SPEC>;
    case 4:
      return forward_layout<LAYOUT, BLOCK, 4, SPEC>;
    case 8:
      return forward_layout<LAYOUT, BLOCK, 8, SPEC>;
    default:
      return nullptr;
  }
}

template <class LAYOUT, bool SPEC>
FwdKernel pick_layout_block(int block, int L) {
  switch (block) {
    case 64:
      return pick_layout_lanes<LAYOUT, 64, SPEC>(L);
    case 128:
      return pick_layout_lanes<LAYOUT, 128, SPEC>(L);
    case 256:
      return pick_layout_lanes<LAYOUT, 256, SPEC>(L);
    case 512:
      return pick_layout_lanes<LAYOUT, 512, SPEC>(L);
    default:
      return nullptr;
  }
}

// Whether no component of the block has its scattering term (ZODI_CF_SCATTER).
inline bool no_scattering(const zodi_model* m) {
  for (int c = 0; c < m->n_comps; ++c)
    if (m->comp_flags[c] & ZODI_CF_SCATTER) return false;
  return true;
}

template <class LAYOUT, int BLOCK>
FwdKernel pick_shared_cut(const zodi_model* m) {
  if (no_scattering(m)) return forward_layout<LAYOUT, BLOCK, 1, true, true, true>;
  return forward_layout<LAYOUT, BLOCK, 1, true, true>;
}

// The registry variant when the block meets its preconditions (with the shared
// cut-offs for one lane per ray, the launch the tuner picks for large batches,
// and without scattering code when no component scatters), else the one that
// tests each component's flags.
template <class LAYOUT>
FwdKernel pick_layout(const zodi_model* m, int block, int L) {
  constexpr uint32_t kProps = RegistryProps<LAYOUT>::value;
  if constexpr (kProps != props_without_share(kProps)) {
    if (L == 1 && props_match(m, kProps)) {
      switch (block) {
        case 64:
          return pick_shared_cut<LAYOUT, 64>(m);
        case 128:
          return pick_shared_cut<LAYOUT, 128>(m);
        case 256:
          return pick_shared_cut<LAYOUT, 256>(m);
        case 512:
          return pick_shared_cut<LAYOUT, 512>(m);
        default:
          return nullptr;
      }
    }
  }
  if (props_match(m, props_without_share(kProps))) return pick_layout_block<LAYOUT, true>(block, L);
  return pick_layout_block<LAYOUT, false>(block, L);
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

constexpr int kGenericWaves = 3;

// Resident waves per SIMD of the kernel pick_forward chooses.
int forward_waves(const zodi_model* m, int variant) {
  switch (variant == 0 ? kLayoutGeneric : match_layout(m)) {
    case kLayoutDirbe:
      return ForwardOccupancy<LayoutDirbe>::value;
    case kLayoutKelsall4:
      return ForwardOccupancy<LayoutKelsall4>::value;
    case kLayoutRRM:
      return ForwardOccupancy<LayoutRRM>::value;
    case kLayoutRRMSplit:
      return ForwardOccupancy<LayoutRRMSplit>::value;
    default:
      return kGenericWaves;
  }
}

// Lanes per ray when the caller leaves it open: the node terms are added in node
// order whatever the split, so it only sets the speed. A small launch whose one
// lane per ray leaves SIMDs idle, or one wave on a SIMD with nothing to hide its
// latency, splits each ray's nodes over up to 8 lanes while the waves still fit
// the resident slots (kernel waves per SIMD x 4 SIMDs x CUs). Measured (DIRBE,
// one observer): 3072 rays 102 -> 36 us at 8 lanes, 49 152 rays 107 -> 76 us at
// 4, and 196 608 rays fastest at 1.
int default_lanes(const zodi_model* m, int variant, int64_t n_rays) {
  const int64_t slots = (int64_t)forward_waves(m, variant) * 4 * cu_count();
  for (int L : {8, 4, 2})
    if ((n_rays * L + 63) / 64 <= slots) return L;
  return 1;
}

FwdKernel pick_forward(const zodi_model* m, int block, int L, int variant) {
  const LayoutId id = variant == 0 ? kLayoutGeneric : match_layout(m);
  switch (id) {
    case kLayoutDirbe:
      return pick_layout<LayoutDirbe>(m, block, L);
    case kLayoutKelsall4:
      return pick_layout<LayoutKelsall4>(m, block, L);
    case kLayoutRRM:
      return pick_layout<LayoutRRM>(m, block, L);
    case kLayoutRRMSplit:
      return pick_layout<LayoutRRMSplit>(m, block, L);
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

  int block = 256, L = 0, per_cu = 0, variant = 1;
  if (launch != nullptr) {
    if (launch->block) block = launch->block;
    L = launch->lanes_per_ray;
    per_cu = launch->blocks_per_cu;
    variant = launch->variant;
  }
  if (variant < 0 || variant > 1) return fail(ZODI_ERR_ARG, "unknown variant %d", variant);
  if (L < 0) return fail(ZODI_ERR_ARG, "lanes_per_ray %d < 0", L);
  if (L == 0) L = default_lanes(model_host, variant, n_rays);
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

zodi_status zodi_host_device_pointer(const void* host, void** device) {
  if (host == nullptr || device == nullptr) return fail(ZODI_ERR_ARG, "null pointer");
  *device = nullptr;
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, host);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(ZODI_ERR_ARG, "%p is not memory known to HIP (%s)", host, hipGetErrorString(e));
  }
  if (attr.type != hipMemoryTypeHost)
    return fail(ZODI_ERR_ARG, "%p is not page-locked host memory (memory type %d)", host,
                (int)attr.type);
  void* d = nullptr;
  e = hipHostGetDevicePointer(&d, const_cast<void*>(host), 0);
  if (e != hipSuccess || d == nullptr) {
    (void)hipGetLastError();
    return fail(ZODI_ERR_ARG, "no device mapping for host memory %p (%s)", host,
                hipGetErrorString(e));
  }
  *device = d;
  return ZODI_OK;
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

