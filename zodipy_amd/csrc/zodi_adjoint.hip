// Adjoint (vector-Jacobian product) of the line-of-sight integral, gfx950 FP64.
//
// For weights g (per ray, or per component and ray) the kernels form
//   grad_p = sum_rays sum_c g_c,r dI_c,r / dtheta_p
// for every model parameter theta_p, recomputing the integrand (no stored node
// values). The integral of component c along ray r is I = h sum_k w_k f(s_k)
// with h = (stop - start)/2 fixed by the cutoffs, so dI/dtheta = h sum_k w_k
// df/dtheta. In the chi^2 mode the forward kernel first writes the model values
// I; the adjoint kernel then weights each ray by g = 2 (I - d) / sigma^2 and
// also sums chi^2.
//
// One wave owns a tile of 64 rays, one ray per lane. Per chunk of components
// of a group (components sharing a range; the chunk size is a compile-time
// choice that trades recomputed node geometry for register pressure), each
// node's geometry, B(T), dB/dT and the phase function with its C1..C3 partials
// are formed once; every component of the chunk then adds its contributions to
// accumulators sized at compile time (layout templates): registers, or in the
// Jacobian pass's four-component chunks LDS rows for the first two. After a
// chunk's node loop the wave sums each accumulator over its lanes with an xor
// butterfly (every lane obtains the same bits) and lane 0 adds the sums to the
// wave's gradient row in LDS. At the end the block
// adds its wave rows in wave order into one partial vector and a second kernel
// adds the partials in a fixed order: the result is reproducible bit for bit.
//
// The parameter vector follows zodipy/zodiacal_light_model.py:38-55 (component
// init fields in declaration order) followed by the effective spectral
// parameters; offsets come from the model block (grad_* fields).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <tuple>
#include <utility>

#include "zodi_adjoint.hpp"

namespace zodi {

// The registry layouts' adjoint kernel. SPEC: the registry variant
// (RegistryProps; the host checks props_match). NOSCAT: see adjoint_group. NREG:
// components of a chunk whose accumulators stay in registers (the rest sum in
// LDS). SHARE (with SPEC, four-component passes): bands with a shared radial
// cut-off (kShareCut) reuse it.
template <class LAYOUT, int CHUNK, int SPLIT, bool JAC, bool SPEC = false, bool NOSCAT = false,
          int NREG = reg_comps<JAC, CHUNK>(), bool SHARE = false>
__global__ __launch_bounds__(kAdjBlock) __attribute__((amdgpu_waves_per_eu(2, 2))) void
adjoint_layout(AdjArgs a) {
  constexpr uint32_t PROPS =
      !SPEC ? 0u
            : (SHARE ? RegistryProps<LAYOUT>::value
                     : props_without_share(RegistryProps<LAYOUT>::value));
  // Components of a chunk beyond NREG: LDS accumulators, per wave kLdsSlots rows
  // of 64 lanes.
  constexpr int kLdsSlots = layout_lds_slots<CHUNK, NREG>(LAYOUT{});
  adjoint_body<SPLIT, JAC, kLdsSlots>(
      a, [&](const Ctx& cx, const Ray& y, double gscale, double* row, int lane, int sub,
             double* lane_acc) {
        run_adjoint<SPLIT, CHUNK, NREG, JAC, PROPS, NOSCAT>(LAYOUT{}, cx, a, y, gscale, row, lane,
                                                            sub, lane_acc);
      });
}

// One block per gradient entry: thread t sums block partials t, t + 256, ... in
// order, then a fixed LDS tree adds the 256 sums. The order of every addition
// is fixed by the launch shape alone, so results repeat bit for bit.
__global__ __launch_bounds__(256) void adjoint_reduce(const double* __restrict__ partials,
                                                      int n_blocks, int np1,
                                                      double* __restrict__ grad) {
  __shared__ double red[256];
  const int p = blockIdx.x;
  const int t = threadIdx.x;
  double s = 0.0;
  for (int b = t; b < n_blocks; b += 256) s += partials[(int64_t)b * np1 + p];
  red[t] = s;
  __syncthreads();
#pragma unroll
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) grad[p] = red[0];
}

// chi^2 and its gradient from a stored Jacobian: per ray g_r = 2 (I_r - d_r) w_r
// with w_r = 1/sigma_r^2, grad_p = sum_r g_r J[p][r], chi^2 = sum_r (I_r - d_r)^2 w_r.
// Block (x, y) takes rays [x R, x R + R) (R from jacobian_chunk_rays) and
// parameters [y P, y P + P). Thread t takes the ray pairs (r, r + 1) with
// r = x R + 2 t + 512 i over the largest multiple of 512 rays of the range, then
// single rays t, t + 256, ... of the rest, adding r before r + 1. Each wave sums its
// lanes with the xor butterfly, and the waves' sums are added in wave order into
// partials[x][p] (entry n_params: chi^2, from the y = 0 blocks). adjoint_reduce
// then adds the partials of all x in a fixed order.
//
// VEC reads each pair's two Jacobian entries with one 16-byte load (J rows 16-byte
// aligned: even leading dimension and base); otherwise two 8-byte loads. The
// additions are the same either way, so the result depends on n_rays alone.
constexpr int kJacParams = 32;
constexpr int kJacPairStep = 512;  // rays per block iteration of the pair loop

// Rays per block of jacobian_chi2: about 1024 blocks per parameter tile (enough
// loads in flight to stream J at full rate), 256 to 8192 rays each. A function of
// n_rays alone, so the order of the additions does not depend on the device.
__host__ __device__ inline int64_t jacobian_chunk_rays(int64_t n_rays) {
  int64_t r = (n_rays + 1023) / 1024;
  r = (r + 255) / 256 * 256;
  return r < 256 ? 256 : (r > 8192 ? 8192 : r);
}

template <bool VEC>
__global__ __launch_bounds__(256) void jacobian_chi2(const double* __restrict__ jac, int64_t ld,
                                                     int n_params, int64_t n_rays,
                                                     int64_t chunk_rays,
                                                     const double* __restrict__ vals,
                                                     const double* __restrict__ target,
                                                     const double* __restrict__ inv_var,
                                                     double inv_var_scalar,
                                                     double* __restrict__ partials) {
  __shared__ double red[kAdjWaves][kJacParams + 1];
  const int t = threadIdx.x;
  const int wave = t / kWave, lane = t % kWave;
  const int p0 = blockIdx.y * kJacParams;
  const int np = min(kJacParams, n_params - p0);
  const int64_t r0 = (int64_t)blockIdx.x * chunk_rays;
  const int64_t r1 = min(n_rays, r0 + chunk_rays);
  const int64_t rv = r0 + (r1 - r0) / kJacPairStep * kJacPairStep;
  const double* rows = jac + (int64_t)p0 * ld;
  double acc[kJacParams];
#pragma unroll
  for (int q = 0; q < kJacParams; ++q) acc[q] = 0.0;
  double chi = 0.0;
  for (int64_t r = r0 + 2 * t; r < rv; r += kJacPairStep) {
    const double ra = vals[r] - target[r];
    const double rb = vals[r + 1] - target[r + 1];
    const double iva = inv_var != nullptr ? inv_var[r] : inv_var_scalar;
    const double ivb = inv_var != nullptr ? inv_var[r + 1] : inv_var_scalar;
    chi += ra * ra * iva;
    chi += rb * rb * ivb;
    const double ga = 2.0 * ra * iva;
    const double gb = 2.0 * rb * ivb;
    const double* col = rows + r;
#pragma unroll
    for (int q = 0; q < kJacParams; ++q) {
      if (q < np) {
        double ja, jb;
        if constexpr (VEC) {
          const double2 j = *reinterpret_cast<const double2*>(col + q * ld);
          ja = j.x;
          jb = j.y;
        } else {
          ja = col[q * ld];
          jb = col[q * ld + 1];
        }
        acc[q] += ga * ja;
        acc[q] += gb * jb;
      }
    }
  }
  for (int64_t r = rv + t; r < r1; r += 256) {
    const double res = vals[r] - target[r];
    const double iv = inv_var != nullptr ? inv_var[r] : inv_var_scalar;
    chi += res * res * iv;
    const double g = 2.0 * res * iv;
    const double* col = rows + r;
#pragma unroll
    for (int q = 0; q < kJacParams; ++q) {
      if (q < np) acc[q] += g * col[q * ld];
    }
  }
#pragma unroll
  for (int q = 0; q < kJacParams; ++q) {
    const double v = wave_sum(acc[q]);
    if (lane == 0) red[wave][q] = v;
  }
  chi = wave_sum(chi);
  if (lane == 0) red[wave][kJacParams] = chi;
  __syncthreads();
  double* out = partials + (int64_t)blockIdx.x * (n_params + 1);
  if (t < np) {
    double s = red[0][t];
#pragma unroll
    for (int w = 1; w < kAdjWaves; ++w) s += red[w][t];
    out[p0 + t] = s;
  }
  if (blockIdx.y == 0 && t == 0) {
    double s = red[0][kJacParams];
#pragma unroll
    for (int w = 1; w < kAdjWaves; ++w) s += red[w][kJacParams];
    out[n_params] = s;
  }
}

// Block partials of the allocation-free adjoint (zodi_adjoint_noalloc): one
// file-scope array per GPU, kept until the process ends. Its host-side name is
// not a device address; hipGetSymbolAddress gives each device's copy.
constexpr int64_t kScratchDoubles = int64_t(1) << 19;  // 4 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {


template <class LAYOUT, int CHUNK, bool SPEC, bool NOSCAT>
AdjKernel pick_adj_split(int split, bool jac) {
  if (jac) return split == 1 ? adjoint_layout<LAYOUT, CHUNK, 1, true, SPEC, NOSCAT> : nullptr;
  switch (split) {
    case 1:
      return adjoint_layout<LAYOUT, CHUNK, 1, false, SPEC, NOSCAT>;
    case 2:
      return adjoint_layout<LAYOUT, CHUNK, 2, false, SPEC, NOSCAT>;
    case 4:
      return adjoint_layout<LAYOUT, CHUNK, 4, false, SPEC, NOSCAT>;
    default:
      return nullptr;
  }
}

// Components of a four-component Jacobian chunk whose accumulators stay in
// registers in the NOSCAT variants: without the scattering sums a third one fits
// (250 VGPRs for the DIRBE layout, 237 for Kelsall4, no spill), which halves the
// LDS adds per node; with all four in registers the kernel spills. With the
// shared band cut-off as well (SHARE) all four fit (Kelsall4 in 256 VGPRs; DIRBE
// spills 12, 64 bytes of scratch per lane), and no sum goes through LDS: 2.7 %
// (DIRBE) and 5 % (Kelsall4) faster than three.
constexpr int kJacRegCompsNoscat = 3;
constexpr int kJacRegCompsShare = 4;

// The registry variants (SPEC, and NOSCAT for Kelsall layouts) exist for the
// default plans: four components per pass in the Jacobian pass, two in the
// weighted pass; SHARE (with NOSCAT) for the Jacobian pass of the layouts whose
// registry properties share a band cut-off. `spec` and `share` say whether the
// block meets the preconditions without and with the sharing.
template <class LAYOUT>
AdjKernel pick_adj_chunk(int chunk, int split, bool jac, bool spec, bool share, bool noscat) {
  constexpr bool kKelsall = layout_integrand<LAYOUT>() == ZODI_INTEGRAND_KELSALL;
  constexpr uint32_t kProps = RegistryProps<LAYOUT>::value;
  const bool default_plan = jac ? chunk == 4 : chunk == 2;
  if (spec && default_plan) {
    if (jac) {
      if (split != 1) return nullptr;
      if constexpr (kKelsall && kProps != props_without_share(kProps)) {
        if (noscat && share)
          return adjoint_layout<LAYOUT, 4, 1, true, true, true, kJacRegCompsShare, true>;
      }
      if constexpr (kKelsall) {
        if (noscat) return adjoint_layout<LAYOUT, 4, 1, true, true, true, kJacRegCompsNoscat>;
      }
      return adjoint_layout<LAYOUT, 4, 1, true, true, false>;
    }
    if (kKelsall && noscat) return pick_adj_split<LAYOUT, 2, true, kKelsall>(split, false);
    return pick_adj_split<LAYOUT, 2, true, false>(split, false);
  }
  switch (chunk) {
    case 1:
      return pick_adj_split<LAYOUT, 1, false, false>(split, jac);
    case 2:
      return pick_adj_split<LAYOUT, 2, false, false>(split, jac);
    case 4:  // the Jacobian pass (adjoint_plan maps 4 to 2 elsewhere)
      return jac && split == 1 ? adjoint_layout<LAYOUT, 4, 1, true> : nullptr;
    default:
      return nullptr;
  }
}

template <class LAYOUT>
AdjKernel pick_adj_layout(const zodi_model* m, int chunk, int split, bool jac) {
  constexpr uint32_t kProps = RegistryProps<LAYOUT>::value;
  return pick_adj_chunk<LAYOUT>(chunk, split, jac, props_match(m, props_without_share(kProps)),
                                props_match(m, kProps), block_noscat(m));
}

AdjKernel pick_adjoint(const zodi_model* m, int chunk, int split, bool jac) {
  switch (match_layout(m)) {
    case kLayoutDirbe:
      return pick_adj_layout<LayoutDirbe>(m, chunk, split, jac);
    case kLayoutKelsall4:
      return pick_adj_layout<LayoutKelsall4>(m, chunk, split, jac);
    case kLayoutRRM:
      return pick_adj_layout<LayoutRRM>(m, chunk, split, jac);
    case kLayoutRRMSplit:
      return pick_adj_layout<LayoutRRMSplit>(m, chunk, split, jac);
    default:
      return nullptr;
  }
}

// Kind-specific parameter count of the gradient layout (kind_nparams on the host).
int host_kind_nparams(int kind) {
  switch (kind) {
    case ZODI_CLOUD:
      return kind_nparams<ZODI_CLOUD>();
    case ZODI_BAND:
      return kind_nparams<ZODI_BAND>();
    case ZODI_RING:
      return kind_nparams<ZODI_RING>();
    case ZODI_FEATURE:
      return kind_nparams<ZODI_FEATURE>();
    case ZODI_FAN:
      return kind_nparams<ZODI_FAN>();
    case ZODI_COMET:
      return kind_nparams<ZODI_COMET>();
    case ZODI_NARROW_BAND:
      return kind_nparams<ZODI_NARROW_BAND>();
    case ZODI_BROAD_BAND:
      return kind_nparams<ZODI_BROAD_BAND>();
    case ZODI_INTERSTELLAR:
      return kind_nparams<ZODI_INTERSTELLAR>();
    case ZODI_RING_RRM:
      return kind_nparams<ZODI_RING_RRM>();
    case ZODI_FEATURE_RRM:
      return kind_nparams<ZODI_FEATURE_RRM>();
    default:
      return -1;
  }
}

// The kernels index the LDS gradient rows with the grad_* offsets: every slot a
// component or the shared parameters write must lie inside [0, n_params).
zodi_status validate_grad_layout(const zodi_model* m) {
  const int np = m->n_params;
  for (int c = 0; c < m->n_comps; ++c) {
    const int k = host_kind_nparams(m->comp_kind[c]);
    if (k < 0) return fail(ZODI_ERR_MODEL, "component %d kind %d has no adjoint", c,
                           m->comp_kind[c]);
    if (m->grad_comp[c] < 0 || m->grad_comp[c] + 5 + k > np)
      return fail(ZODI_ERR_MODEL, "grad_comp[%d] = %d out of range for %d parameters", c,
                  m->grad_comp[c], np);
    if (m->grad_spec_a[c] < 0 || m->grad_spec_a[c] >= np || m->grad_spec_b[c] < 0 ||
        m->grad_spec_b[c] >= np)
      return fail(ZODI_ERR_MODEL, "grad_spec_a/b[%d] out of range for %d parameters", c, np);
  }
  const int nsh = m->integrand == ZODI_INTEGRAND_KELSALL ? 6 : 1;
  for (int q = 0; q < nsh; ++q) {
    if (m->grad_shared[q] < 0 || m->grad_shared[q] >= np)
      return fail(ZODI_ERR_MODEL, "grad_shared[%d] = %d out of range for %d parameters", q,
                  m->grad_shared[q], np);
  }
  return ZODI_OK;
}

constexpr int kDefaultChunk = 2;
constexpr int kDefaultJacChunk = 4;
constexpr int kDefaultBlocksPerCu = 2;

struct AdjPlan {
  int chunk;    // components per pass over a group's nodes
  int split;    // lanes per ray
  int per_cu;   // resident blocks per CU (grid-stride loop)
  int64_t n_tiles;
  int64_t grid;
};

// Launch plan. With lanes_per_ray = 0 one lane owns a ray unless the rays do not
// fill every resident wave once; then the smallest split (2 or 4 lanes per ray)
// that does. Each extra lane per ray repeats the per-ray set-up and the wave
// reductions, which costs more than a partly idle last round. The Jacobian pass
// (jac) always uses one lane per ray.
zodi_status adjoint_plan(const zodi_launch* launch, int64_t n_rays, bool jac, AdjPlan& p) {
  p.chunk = jac ? kDefaultJacChunk : kDefaultChunk;
  p.split = 0;
  p.per_cu = kDefaultBlocksPerCu;
  if (launch != nullptr) {
    if (launch->block != 0 && launch->block != kAdjBlock)
      return fail(ZODI_ERR_ARG, "adjoint block must be %d, got %d", kAdjBlock, launch->block);
    if (launch->lanes_per_ray != 0 && launch->lanes_per_ray != 1 && launch->lanes_per_ray != 2 &&
        launch->lanes_per_ray != 4)
      return fail(ZODI_ERR_ARG, "adjoint lanes_per_ray must be 0, 1, 2 or 4, got %d",
                  launch->lanes_per_ray);
    if (launch->blocks_per_cu < 0 || launch->blocks_per_cu > 64)
      return fail(ZODI_ERR_ARG, "blocks_per_cu %d outside [0, 64]", launch->blocks_per_cu);
    if (launch->variant != 0 && launch->variant != 1 && launch->variant != 2 &&
        launch->variant != 4)
      return fail(ZODI_ERR_ARG, "adjoint variant (components per pass) must be 0, 1, 2 or 4, got %d",
                  launch->variant);
    // 4 keeps two components' accumulators in LDS, which only the Jacobian pass
    // does; the weighted (recompute) pass runs two components per pass instead.
    if (launch->variant != 0) p.chunk = launch->variant == 4 && !jac ? 2 : launch->variant;
    p.split = launch->lanes_per_ray;
    if (launch->blocks_per_cu > 0) p.per_cu = launch->blocks_per_cu;
  }
  if (jac) {
    if (p.split > 1)
      return fail(ZODI_ERR_ARG, "the Jacobian pass uses one lane per ray, got lanes_per_ray %d",
                  p.split);
    p.split = 1;
  }
  const int64_t waves = (int64_t)p.per_cu * cu_count() * kAdjWaves;
  if (p.split == 0) {
    p.split = 4;
    for (int s : {1, 2}) {
      if ((n_rays * s + kWave - 1) / kWave >= waves) {
        p.split = s;
        break;
      }
    }
  }
  p.n_tiles = (n_rays * p.split + kWave - 1) / kWave;
  const int64_t blocks = (p.n_tiles + kAdjWaves - 1) / kAdjWaves;
  p.grid = std::max<int64_t>(1, std::min<int64_t>(blocks, waves / kAdjWaves));
  return ZODI_OK;
}

// Rows of the Planck-table gradient: one entry per table temperature, then chi^2;
// with the effective spectral parameters (eff), the parameter vector first.
constexpr int kTableRow = ZODI_N_TEMPS + 1;

int table_row(const zodi_model* m, bool eff) { return kTableRow + (eff ? m->n_params : 0); }

// Launch plan and kernel of the table gradient: the weighted pass's plan, with as
// many resident blocks per CU as the kernel is compiled for unless the caller
// sets blocks_per_cu; the components-per-pass variant does not apply.
zodi_status table_plan(const zodi_model* m, const zodi_launch* launch, int64_t n_rays, bool eff,
                       AdjPlan& p, AdjKernel& k) {
  if (eff) {
    const zodi_status st = validate_grad_layout(m);
    if (st != ZODI_OK) return st;
  }
  const TableKernel any = table_adjoint_kernel(m, 1, eff);
  if (any.k == nullptr) return fail(ZODI_ERR_MODEL, "no table adjoint kernel for this model");
  zodi_launch l = launch != nullptr ? *launch : zodi_launch{0, 0, 0, 0};
  if (l.blocks_per_cu == 0) l.blocks_per_cu = any.waves;
  l.variant = 0;
  const zodi_status st = adjoint_plan(&l, n_rays, false, p);
  if (st != ZODI_OK) return st;
  k = table_adjoint_kernel(m, p.split, eff).k;
  if (k == nullptr)
    return fail(ZODI_ERR_ARG, "no table adjoint kernel for %d lanes per ray", p.split);
  return ZODI_OK;
}

// Largest Jacobian the chi^2 mode stores in its workspace (a quarter of one
// MI355X's 288 GB of HBM); larger problems recompute the integrand instead.
constexpr int64_t kJacobianMaxBytes = int64_t(72) << 30;

// Workspace of the recompute path: block partials, then I.
int64_t recompute_bytes(const AdjPlan& p, int np1, int64_t n_rays) {
  return (p.grid * np1 + n_rays) * (int64_t)sizeof(double);
}

// Workspace of the chi^2 mode's Jacobian path: J [n_params][n_rays], the
// partials of jacobian_chi2, then I.
int64_t jacobian_bytes(int np, int64_t n_rays) {
  const int64_t R = jacobian_chunk_rays(n_rays);
  const int64_t chunks = (n_rays + R - 1) / R;
  return ((int64_t)np * n_rays + chunks * (np + 1) + n_rays) * (int64_t)sizeof(double);
}

// The chi^2 mode can take the Jacobian path: one lane per ray allowed, J small enough.
bool jacobian_eligible(const zodi_launch* launch, int np, int64_t n_rays) {
  if (launch != nullptr && launch->lanes_per_ray > 1) return false;
  return (int64_t)np * n_rays * (int64_t)sizeof(double) <= kJacobianMaxBytes;
}

AdjArgs make_args(const zodi_model* model_dev, const double* u, int64_t n_rays,
                  const int64_t* seg_offsets, int32_t n_seg, const double* obs,
                  const double* earth, int32_t obs_mode, const uint32_t* beyond,
                  const AdjPlan& plan, int n_params) {
  AdjArgs a;
  memset(&a, 0, sizeof(a));
  a.f.m = model_dev;
  a.f.u = u;
  a.f.n_rays = n_rays;
  a.f.seg = seg_offsets;
  a.f.n_seg = n_seg;
  a.f.obs_mode = obs_mode;
  a.f.obs = obs;
  a.f.earth = earth;
  a.f.obs_stride = obs_mode == ZODI_OBS_PER_RAY ? n_rays : n_seg;
  a.f.beyond = beyond;
  a.f.out_mode = ZODI_OUT_SUM;
  a.f.n_tiles = plan.n_tiles;
  a.n_params = n_params;
  return a;
}

// Checks shared by zodi_adjoint and zodi_jacobian.
zodi_status check_rays(const zodi_model* model_host, const zodi_model* model_dev,
                       const double* u, int64_t n_rays, const int64_t* seg_offsets,
                       int32_t n_seg, const double* obs, const double* earth, int32_t obs_mode,
                       const uint32_t* beyond) {
  if (model_host == nullptr) return fail(ZODI_ERR_ARG, "null model");
  zodi_status st = validate_model(model_host);
  if (st != ZODI_OK) return st;
  if (n_rays < 0) return fail(ZODI_ERR_ARG, "n_rays < 0");
  if (model_dev == nullptr) return fail(ZODI_ERR_ARG, "null pointer (model_dev)");
  if (n_rays > 0 && (u == nullptr || obs == nullptr || earth == nullptr || beyond == nullptr))
    return fail(ZODI_ERR_ARG, "null pointer (u, obs, earth or beyond)");
  if (obs_mode == ZODI_OBS_PER_SEGMENT) {
    if (n_rays > 0 && (seg_offsets == nullptr || n_seg < 1))
      return fail(ZODI_ERR_ARG, "per-segment mode needs seg_offsets and n_seg >= 1");
  } else if (obs_mode != ZODI_OBS_PER_RAY) {
    return fail(ZODI_ERR_ARG, "unknown obs_mode %d", obs_mode);
  }
  if (model_host->n_params < 1 || model_host->n_params > ZODI_MAX_PARAMS)
    return fail(ZODI_ERR_MODEL, "n_params %d outside [1, %d]", model_host->n_params,
                ZODI_MAX_PARAMS);
  return ZODI_OK;
}

AdjKernel pick_checked(const zodi_model* model_host, const AdjPlan& plan, bool jac,
                       zodi_status& st) {
  st = validate_grad_layout(model_host);
  if (st != ZODI_OK) return nullptr;
  // Registry layouts run their layout kernels; any other layout the generic one
  // (one component per pass; the components-per-pass choice does not apply).
  AdjKernel k = match_layout(model_host) == kLayoutGeneric
                    ? generic_adjoint_kernel(model_host->integrand, plan.split, jac)
                    : pick_adjoint(model_host, plan.chunk, plan.split, jac);
  if (k == nullptr) {
    st = fail(ZODI_ERR_ARG, "no adjoint kernel for %d components per pass and %d lanes per ray",
              plan.chunk, plan.split);
    return nullptr;
  }
  return k;
}

// Ray addressing shared by the entry points.
struct RayArgs {
  const double* u;
  int64_t n_rays;
  const int64_t* seg_offsets;
  int32_t n_seg;
  const double* obs;
  const double* earth;
  int32_t obs_mode;
  const uint32_t* beyond;
};

// The recompute path: in chi^2 mode the forward kernel first writes I to vals;
// the adjoint kernel then writes one partial row per block ([plan.grid][n_params
// + 1], every entry written) and adjoint_reduce adds them in block order.
zodi_status adjoint_recompute(const zodi_model* model_host, const zodi_model* model_dev,
                              const RayArgs& r, int32_t mode, const double* weights,
                              const double* target, const double* inv_var,
                              double inv_var_scalar, const AdjPlan& plan, AdjKernel k,
                              double* partials, double* vals, double* grad, void* stream) {
  const int np = model_host->n_params;
  const int np1 = np + 1;
  if (mode == ZODI_ADJ_CHI2 && r.n_rays > 0) {
    const zodi_status st =
        zodi_forward(model_host, model_dev, r.u, r.n_rays, r.seg_offsets, r.n_seg, r.obs,
                     r.earth, r.obs_mode, r.beyond, vals, ZODI_OUT_SUM, nullptr, stream);
    if (st != ZODI_OK) return st;
  }
  AdjArgs a = make_args(model_dev, r.u, r.n_rays, r.seg_offsets, r.n_seg, r.obs, r.earth,
                        r.obs_mode, r.beyond, plan, np);
  a.weights = weights;
  a.target = target;
  a.inv_var = inv_var;
  a.inv_var_scalar = inv_var_scalar;
  a.model_vals = vals;
  a.partials = partials;
  a.mode = mode;
  const size_t lds = (size_t)np1 * kAdjWaves * sizeof(double);
  hipLaunchKernelGGL(k, dim3((unsigned)plan.grid), dim3(kAdjBlock), lds, (hipStream_t)stream, a);
  const zodi_status st = check_hip(hipGetLastError(), "adjoint launch");
  if (st != ZODI_OK) return st;
  hipLaunchKernelGGL(adjoint_reduce, dim3(np1), dim3(256), 0, (hipStream_t)stream,
                     (const double*)partials, (int)plan.grid, np1, grad);
  return check_hip(hipGetLastError(), "adjoint reduce launch");
}

int64_t table_workspace(const zodi_model* m, int64_t n_rays, const zodi_launch* launch,
                        bool eff) {
  if (m == nullptr || validate_model(m) != ZODI_OK || n_rays < 0) return -1;
  AdjPlan p;
  AdjKernel k = nullptr;
  if (table_plan(m, launch, n_rays, eff, p, k) != ZODI_OK) return -1;
  return recompute_bytes(p, table_row(m, eff), n_rays);
}

// zodi_table_vjp (eff false) and zodi_spectral_vjp (eff true): the recompute path
// with the table kernel, whose rows take the place of the parameter rows.
zodi_status table_vjp(bool eff, const zodi_model* model_host, const zodi_model* model_dev,
                      const RayArgs& r, int32_t mode, const double* weights,
                      const double* target, const double* inv_var, double inv_var_scalar,
                      double* model_out, double* grad, double* workspace,
                      int64_t workspace_bytes, const zodi_launch* launch, void* stream) {
  zodi_status st = check_rays(model_host, model_dev, r.u, r.n_rays, r.seg_offsets, r.n_seg,
                              r.obs, r.earth, r.obs_mode, r.beyond);
  if (st != ZODI_OK) return st;
  if (grad == nullptr || workspace == nullptr)
    return fail(ZODI_ERR_ARG, "null pointer (grad or workspace)");
  if (mode == ZODI_ADJ_CHI2) {
    if (r.n_rays > 0 && target == nullptr) return fail(ZODI_ERR_ARG, "chi2 mode needs target");
  } else if (mode == ZODI_ADJ_VJP_SUM || mode == ZODI_ADJ_VJP_COMPS) {
    if (r.n_rays > 0 && weights == nullptr) return fail(ZODI_ERR_ARG, "vjp mode needs weights");
  } else {
    return fail(ZODI_ERR_ARG, "unknown adjoint mode %d", mode);
  }
  AdjPlan plan;
  AdjKernel k = nullptr;
  st = table_plan(model_host, launch, r.n_rays, eff, plan, k);
  if (st != ZODI_OK) return st;
  const int row = table_row(model_host, eff);
  const int64_t need = recompute_bytes(plan, row, r.n_rays);
  if (workspace_bytes < need)
    return fail(ZODI_ERR_ARG, "workspace of %lld bytes needed, %lld given", (long long)need,
                (long long)workspace_bytes);
  double* vals = model_out != nullptr ? model_out : workspace + plan.grid * row;
  zodi_model shape = *model_host;
  shape.n_params = row - 1;
  return adjoint_recompute(&shape, model_dev, r, mode, weights, target, inv_var, inv_var_scalar,
                           plan, k, workspace, vals, grad, stream);
}

// Device address of this GPU's copy of `scratch`, looked up once per device with
// that device current (the current device when first asked).
zodi_status scratch_address(double** out) {
  static std::mutex mu;
  static void* addr[64] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return check_hip(e, "hipGetDevice");
  if (dev < 0 || dev >= 64) return fail(ZODI_ERR_ARG, "device %d outside [0, 64)", dev);
  std::lock_guard<std::mutex> lock(mu);
  if (addr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return check_hip(e, "hipGetSymbolAddress(scratch)");
    if (p == nullptr) return fail(ZODI_ERR_HIP, "hipGetSymbolAddress(scratch) gave null");
    addr[dev] = p;
  }
  *out = static_cast<double*>(addr[dev]);
  return ZODI_OK;
}

}  // namespace
}  // namespace zodi

using namespace zodi;

extern "C" {

int64_t zodi_adjoint_workspace(const zodi_model* model_host, int64_t n_rays, int32_t mode,
                               const zodi_launch* launch) {
  if (model_host == nullptr || validate_model(model_host) != ZODI_OK || n_rays < 0) return -1;
  if (mode != ZODI_ADJ_VJP_SUM && mode != ZODI_ADJ_VJP_COMPS && mode != ZODI_ADJ_CHI2) {
    fail(ZODI_ERR_ARG, "unknown adjoint mode %d", mode);
    return -1;
  }
  const int np = model_host->n_params;
  if (mode == ZODI_ADJ_CHI2 && jacobian_eligible(launch, np, n_rays)) {
    AdjPlan p;
    if (adjoint_plan(launch, n_rays, true, p) != ZODI_OK) return -1;
    return jacobian_bytes(np, n_rays);
  }
  AdjPlan p;
  if (adjoint_plan(launch, n_rays, false, p) != ZODI_OK) return -1;
  return recompute_bytes(p, np + 1, n_rays);
}

zodi_status zodi_jacobian(const zodi_model* model_host, const zodi_model* model_dev,
                          const double* u, int64_t n_rays, const int64_t* seg_offsets,
                          int32_t n_seg, const double* obs, const double* earth,
                          int32_t obs_mode, const uint32_t* beyond, double* model_out,
                          double* jac, int64_t ld_jac, const zodi_launch* launch, void* stream) {
  zodi_status st = check_rays(model_host, model_dev, u, n_rays, seg_offsets, n_seg, obs, earth,
                              obs_mode, beyond);
  if (st != ZODI_OK) return st;
  if (n_rays > 0 && (model_out == nullptr || jac == nullptr))
    return fail(ZODI_ERR_ARG, "null pointer (model_out or jac)");
  if (ld_jac < n_rays)
    return fail(ZODI_ERR_ARG, "ld_jac %lld < n_rays %lld", (long long)ld_jac, (long long)n_rays);
  AdjPlan plan;
  st = adjoint_plan(launch, n_rays, true, plan);
  if (st != ZODI_OK) return st;
  AdjKernel k = pick_checked(model_host, plan, true, st);
  if (k == nullptr) return st;
  if (n_rays == 0) return ZODI_OK;
  AdjArgs a = make_args(model_dev, u, n_rays, seg_offsets, n_seg, obs, earth, obs_mode, beyond,
                        plan, model_host->n_params);
  a.jac = jac;
  a.ld_jac = ld_jac;
  a.model_vals_out = model_out;
  hipLaunchKernelGGL(k, dim3((unsigned)plan.grid), dim3(kAdjBlock), 0, (hipStream_t)stream, a);
  return check_hip(hipGetLastError(), "jacobian launch");
}

zodi_status zodi_adjoint(const zodi_model* model_host, const zodi_model* model_dev,
                         const double* u, int64_t n_rays, const int64_t* seg_offsets,
                         int32_t n_seg, const double* obs, const double* earth, int32_t obs_mode,
                         const uint32_t* beyond, int32_t mode, const double* weights,
                         const double* target, const double* inv_var, double inv_var_scalar,
                         double* model_out, double* grad, double* workspace,
                         int64_t workspace_bytes, const zodi_launch* launch, void* stream) {
  zodi_status st = check_rays(model_host, model_dev, u, n_rays, seg_offsets, n_seg, obs, earth,
                              obs_mode, beyond);
  if (st != ZODI_OK) return st;
  if (grad == nullptr || workspace == nullptr)
    return fail(ZODI_ERR_ARG, "null pointer (grad or workspace)");
  if (mode == ZODI_ADJ_CHI2) {
    if (n_rays > 0 && target == nullptr) return fail(ZODI_ERR_ARG, "chi2 mode needs target");
  } else if (mode == ZODI_ADJ_VJP_SUM || mode == ZODI_ADJ_VJP_COMPS) {
    if (n_rays > 0 && weights == nullptr) return fail(ZODI_ERR_ARG, "vjp mode needs weights");
  } else {
    return fail(ZODI_ERR_ARG, "unknown adjoint mode %d", mode);
  }
  const int np = model_host->n_params;
  const int np1 = np + 1;

  // chi^2 with room for the Jacobian: one pass writes J and I, then J^T g.
  if (mode == ZODI_ADJ_CHI2 && jacobian_eligible(launch, np, n_rays) &&
      workspace_bytes >= jacobian_bytes(np, n_rays)) {
    AdjPlan plan;
    st = adjoint_plan(launch, n_rays, true, plan);
    if (st != ZODI_OK) return st;
    AdjKernel k = pick_checked(model_host, plan, true, st);
    if (k == nullptr) return st;
    const int64_t chunk_rays = jacobian_chunk_rays(n_rays);
    const int64_t chunks = (n_rays + chunk_rays - 1) / chunk_rays;
    double* partials = workspace + (int64_t)np * n_rays;
    double* vals = model_out != nullptr ? model_out : partials + chunks * np1;
    if (n_rays > 0) {
      AdjArgs a = make_args(model_dev, u, n_rays, seg_offsets, n_seg, obs, earth, obs_mode,
                            beyond, plan, np);
      a.jac = workspace;
      a.ld_jac = n_rays;
      a.model_vals_out = vals;
      hipLaunchKernelGGL(k, dim3((unsigned)plan.grid), dim3(kAdjBlock), 0, (hipStream_t)stream,
                         a);
      st = check_hip(hipGetLastError(), "jacobian launch");
      if (st != ZODI_OK) return st;
      const dim3 g2((unsigned)chunks, (unsigned)((np + kJacParams - 1) / kJacParams));
      // 16-byte loads of J when its rows are 16-byte aligned (same additions either way)
      const bool vec = (n_rays % 2) == 0 && ((uintptr_t)workspace % 16) == 0;
      hipLaunchKernelGGL(vec ? jacobian_chi2<true> : jacobian_chi2<false>, g2, dim3(256), 0,
                         (hipStream_t)stream, (const double*)workspace, n_rays, np, n_rays,
                         chunk_rays, (const double*)vals, target, inv_var, inv_var_scalar,
                         partials);
      st = check_hip(hipGetLastError(), "jacobian chi2 launch");
      if (st != ZODI_OK) return st;
    }
    hipLaunchKernelGGL(adjoint_reduce, dim3(np1), dim3(256), 0, (hipStream_t)stream,
                       (const double*)partials, (int)chunks, np1, grad);
    return check_hip(hipGetLastError(), "adjoint reduce launch");
  }

  AdjPlan plan;
  st = adjoint_plan(launch, n_rays, false, plan);
  if (st != ZODI_OK) return st;
  AdjKernel k = pick_checked(model_host, plan, false, st);
  if (k == nullptr) return st;
  const int64_t need = recompute_bytes(plan, np1, n_rays);
  if (workspace_bytes < need)
    return fail(ZODI_ERR_ARG, "workspace of %lld bytes needed, %lld given", (long long)need,
                (long long)workspace_bytes);
  double* vals = model_out != nullptr ? model_out : workspace + plan.grid * np1;
  const RayArgs r = {u, n_rays, seg_offsets, n_seg, obs, earth, obs_mode, beyond};
  return adjoint_recompute(model_host, model_dev, r, mode, weights, target, inv_var,
                           inv_var_scalar, plan, k, workspace, vals, grad, stream);
}

int64_t zodi_table_vjp_workspace(const zodi_model* model_host, int64_t n_rays,
                                 const zodi_launch* launch) {
  return table_workspace(model_host, n_rays, launch, false);
}

int64_t zodi_spectral_vjp_workspace(const zodi_model* model_host, int64_t n_rays,
                                    const zodi_launch* launch) {
  return table_workspace(model_host, n_rays, launch, true);
}

zodi_status zodi_table_vjp(const zodi_model* model_host, const zodi_model* model_dev,
                           const double* u, int64_t n_rays, const int64_t* seg_offsets,
                           int32_t n_seg, const double* obs, const double* earth,
                           int32_t obs_mode, const uint32_t* beyond, int32_t mode,
                           const double* weights, const double* target, const double* inv_var,
                           double inv_var_scalar, double* model_out, double* grad_table,
                           double* workspace, int64_t workspace_bytes, const zodi_launch* launch,
                           void* stream) {
  const RayArgs r = {u, n_rays, seg_offsets, n_seg, obs, earth, obs_mode, beyond};
  return table_vjp(false, model_host, model_dev, r, mode, weights, target, inv_var,
                   inv_var_scalar, model_out, grad_table, workspace, workspace_bytes, launch,
                   stream);
}

zodi_status zodi_spectral_vjp(const zodi_model* model_host, const zodi_model* model_dev,
                              const double* u, int64_t n_rays, const int64_t* seg_offsets,
                              int32_t n_seg, const double* obs, const double* earth,
                              int32_t obs_mode, const uint32_t* beyond, int32_t mode,
                              const double* weights, const double* target, const double* inv_var,
                              double inv_var_scalar, double* model_out, double* grad,
                              double* workspace, int64_t workspace_bytes,
                              const zodi_launch* launch, void* stream) {
  const RayArgs r = {u, n_rays, seg_offsets, n_seg, obs, earth, obs_mode, beyond};
  return table_vjp(true, model_host, model_dev, r, mode, weights, target, inv_var,
                   inv_var_scalar, model_out, grad, workspace, workspace_bytes, launch, stream);
}

zodi_status zodi_adjoint_noalloc(const zodi_model* model_host, const zodi_model* model_dev,
                                 const double* u, int64_t n_rays, const int64_t* seg_offsets,
                                 int32_t n_seg, const double* obs, const double* earth,
                                 int32_t obs_mode, const uint32_t* beyond, int32_t mode,
                                 const double* weights, const double* target,
                                 const double* inv_var, double inv_var_scalar, double* model_out,
                                 double* grad, const zodi_launch* launch, void* stream) {
  zodi_status st = check_rays(model_host, model_dev, u, n_rays, seg_offsets, n_seg, obs, earth,
                              obs_mode, beyond);
  if (st != ZODI_OK) return st;
  if (grad == nullptr) return fail(ZODI_ERR_ARG, "null pointer (grad)");
  if (mode == ZODI_ADJ_CHI2) {
    if (n_rays > 0 && (target == nullptr || model_out == nullptr))
      return fail(ZODI_ERR_ARG, "chi2 mode needs target and model_out");
  } else if (mode == ZODI_ADJ_VJP_SUM || mode == ZODI_ADJ_VJP_COMPS) {
    if (n_rays > 0 && weights == nullptr) return fail(ZODI_ERR_ARG, "vjp mode needs weights");
  } else {
    return fail(ZODI_ERR_ARG, "unknown adjoint mode %d", mode);
  }
  const int np1 = model_host->n_params + 1;
  AdjPlan plan;
  st = adjoint_plan(launch, n_rays, false, plan);
  if (st != ZODI_OK) return st;
  AdjKernel k = pick_checked(model_host, plan, false, st);
  if (k == nullptr) return st;
  const int64_t need = plan.grid * np1;
  if (need > kScratchDoubles)
    return fail(ZODI_ERR_TOO_LARGE,
                "the allocation-free adjoint holds %lld block partials, this launch needs %lld "
                "(%lld blocks x %d entries): use fewer blocks_per_cu or zodi_adjoint",
                (long long)kScratchDoubles, (long long)need, (long long)plan.grid, np1);
  double* partials = nullptr;
  st = scratch_address(&partials);
  if (st != ZODI_OK) return st;
  const RayArgs r = {u, n_rays, seg_offsets, n_seg, obs, earth, obs_mode, beyond};
  return adjoint_recompute(model_host, model_dev, r, mode, weights, target, inv_var,
                           inv_var_scalar, plan, k, partials, model_out, grad, stream);
}

}  // extern "C"
