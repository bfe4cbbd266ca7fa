// Spectral set-up on the device (gfx950, FP64): the effective spectral
// parameters and the Planck table of a model block, written in place.
//
// The reference unpacks the spectral tables once per model on the host
// (zodipy/unpack_model.py:23-173: linear or nearest interpolation at the
// model's frequencies, then the trapezoid over the bandpass) and tabulates the
// blackbody over 100 temperatures (zodipy/blackbody.py:33-49). For fits that
// move the tables, both steps are linear or elementwise and small, so they run
// here, on the block the kernels read, without a host round trip:
//
//   effective value of table t = sum_j A[t][j] tables[t][j]
//
// (A: the interpolation and trapezoid weights of the model's frequencies, fixed
// for a fit), followed by the derived slots the kernels use (1 - albedo,
// (1 - albedo) emissivity, the phase-function normalisation and its C3
// derivative), and B(T_k) = sum_i q_i B_nu_i(T_k) with the slopes of the linear
// table interpolation. One block of 256 threads; every sum runs in index order.

#include <hip/hip_runtime.h>

#include <cstring>

#include "zodi_common.hpp"

namespace zodi {

constexpr int kMaxSpecTargets = 64;
constexpr int kMaxSpecGrid = 256;
constexpr int kMaxFreq = 1 << 16;

struct SpecArgs {
  zodi_model* m;
  int n_tab, n_grid, n_freq, n_comps;
  int kelsall;
  const double* A;       // [n_tab][n_grid]
  const double* tables;  // [n_tab][n_grid]
  const double* nu;      // [n_freq] Hz
  const double* q;       // [n_freq]
  int target[kMaxSpecTargets];
};

// B_nu(T) in MJy/sr: 2 h nu^3 / (c^2 expm1(h nu / (k T))) / 1e-20, with the exact
// SI constants of CODATA 2018.
__device__ __forceinline__ double planck_mjysr(double nu, double T) {
  constexpr double h = 6.62607015e-34, k = 1.380649e-23, c = 299792458.0, mjy = 1e-20;
  const double a = h * nu / (k * T);
  return (2.0 * h * (nu * nu * nu) / (c * c * expm1(a))) / mjy;
}

__global__ __launch_bounds__(256) void spectral_setup_kernel(SpecArgs a) {
  zodi_model* m = a.m;
  const int t = threadIdx.x;
  if (t < a.n_tab) {
    const double* row = a.A + (int64_t)t * a.n_grid;
    const double* tab = a.tables + (int64_t)t * a.n_grid;
    double eff = 0.0;
    for (int j = 0; j < a.n_grid; ++j) eff += row[j] * tab[j];
    const int code = a.target[t];
    const int c = code & 0xff;
    switch (code & ~0xff) {
      case ZODI_SPEC_EMISSIVITY:
        m->comp[c][ZODI_P_EMISSIVITY] = eff;
        break;
      case ZODI_SPEC_ALBEDO:
        m->comp[c][ZODI_P_ALBEDO] = eff;
        break;
      default:
        if (code == ZODI_SPEC_SOLAR) m->shared[ZODI_S_SOLAR] = eff;
        if (code == ZODI_SPEC_C1) m->shared[ZODI_S_C1] = eff;
        if (code == ZODI_SPEC_C2) m->shared[ZODI_S_C2] = eff;
        if (code == ZODI_SPEC_C3) m->shared[ZODI_S_C3] = eff;
        if (code == ZODI_SPEC_CALIBRATION) m->shared[ZODI_S_CALIBRATION] = eff;
        break;
    }
  }
  if (a.n_freq > 0 && t < m->n_temps) {
    const double T = m->tab_T[t];
    double B = 0.0;
    for (int i = 0; i < a.n_freq; ++i) B += a.q[i] * planck_mjysr(a.nu[i], T);
    m->tab_B[t] = B;
  }
  __syncthreads();
  if (a.kelsall && t < a.n_comps) {
    const double alb = m->comp[t][ZODI_P_ALBEDO];
    m->comp[t][ZODI_P_ONE_MINUS_ALBEDO] = 1.0 - alb;
    m->comp[t][ZODI_P_SCALE] = (1.0 - alb) * m->comp[t][ZODI_P_EMISSIVITY];
  }
  if (a.kelsall && t == 0) {
    // N = 1 / (2 pi D), D = 2 C1 + pi C2 + (exp(pi C3) + 1) / (C3^2 + 1)
    // (zodipy/scattering.py:53-59), and dD/dC3 for the adjoint.
    const double C1 = m->shared[ZODI_S_C1], C2 = m->shared[ZODI_S_C2];
    const double C3 = m->shared[ZODI_S_C3];
    const double e3 = exp(C3 * kPi);
    const double q1 = C3 * C3 + 1.0;
    m->shared[ZODI_S_PHASE_NORM] = 1.0 / (2.0 * kPi * (2.0 * C1 + kPi * C2 + (e3 + 1.0) / q1));
    m->shared[ZODI_S_DD_DC3] = (kPi * e3 * q1 - (e3 + 1.0) * 2.0 * C3) / (q1 * q1);
  }
  if (a.n_freq > 0 && t < m->n_temps) {
    const int last = m->n_temps - 1;
    m->tab_slope[t] = t < last ? (m->tab_B[t + 1] - m->tab_B[t]) / (m->tab_T[t + 1] - m->tab_T[t])
                               : 0.0;
  }
}

}  // namespace zodi

using namespace zodi;

extern "C" {

zodi_status zodi_spectral_setup(const zodi_model* model_host, zodi_model* model_dev,
                                int32_t n_tab, int32_t n_grid, const double* A,
                                const double* tables, const int32_t* target, int32_t n_freq,
                                const double* nu_hz, const double* q, void* stream) {
  if (model_host == nullptr || model_dev == nullptr) return fail(ZODI_ERR_ARG, "null model");
  zodi_status st = validate_model(model_host);
  if (st != ZODI_OK) return st;
  if (n_tab < 0 || n_tab > kMaxSpecTargets)
    return fail(ZODI_ERR_ARG, "n_tab %d outside [0, %d]", n_tab, kMaxSpecTargets);
  if (n_tab > 0 && (n_grid < 1 || n_grid > kMaxSpecGrid))
    return fail(ZODI_ERR_ARG, "n_grid %d outside [1, %d]", n_grid, kMaxSpecGrid);
  if (n_freq < 0 || n_freq > kMaxFreq)
    return fail(ZODI_ERR_ARG, "n_freq %d outside [0, %d]", n_freq, kMaxFreq);
  if (n_tab > 0 && (A == nullptr || tables == nullptr || target == nullptr))
    return fail(ZODI_ERR_ARG, "null pointer (A, tables or target)");
  if (n_freq > 0 && (nu_hz == nullptr || q == nullptr))
    return fail(ZODI_ERR_ARG, "null pointer (nu_hz or q)");
  const bool kelsall = model_host->integrand == ZODI_INTEGRAND_KELSALL;
  SpecArgs a;
  memset(&a, 0, sizeof(a));
  for (int t = 0; t < n_tab; ++t) {
    const int code = target[t];
    const int kind = code & ~0xff, c = code & 0xff;
    if (kind == ZODI_SPEC_EMISSIVITY || kind == ZODI_SPEC_ALBEDO) {
      if (!kelsall || c >= model_host->n_comps)
        return fail(ZODI_ERR_ARG, "target 0x%x: no such Kelsall component", code);
      // The scattering branch must be present for an albedo that may leave 0.
      if (kind == ZODI_SPEC_ALBEDO && !(model_host->comp_flags[c] & ZODI_CF_SCATTER))
        return fail(ZODI_ERR_MODEL,
                    "target 0x%x: component %d was packed without its scattering branch", code,
                    c);
    } else if (code == ZODI_SPEC_SOLAR || code == ZODI_SPEC_C1 || code == ZODI_SPEC_C2 ||
               code == ZODI_SPEC_C3) {
      if (!kelsall) return fail(ZODI_ERR_ARG, "target 0x%x needs a Kelsall model", code);
    } else if (code == ZODI_SPEC_CALIBRATION) {
      if (kelsall) return fail(ZODI_ERR_ARG, "target 0x%x needs an RRM model", code);
    } else {
      return fail(ZODI_ERR_ARG, "unknown spectral target 0x%x", code);
    }
    a.target[t] = code;
  }
  if (kelsall && (model_host->flags & 1) == 0) {
    for (int t = 0; t < n_tab; ++t) {
      if ((target[t] & ~0xff) == ZODI_SPEC_ALBEDO)
        return fail(ZODI_ERR_MODEL, "model flags lack the scattering bit");
    }
  }
  if (n_tab == 0 && n_freq == 0) return ZODI_OK;  // nothing to write
  a.m = model_dev;
  a.n_tab = n_tab;
  a.n_grid = n_grid;
  a.n_freq = n_freq;
  a.n_comps = model_host->n_comps;
  a.kelsall = kelsall ? 1 : 0;
  a.A = A;
  a.tables = tables;
  a.nu = nu_hz;
  a.q = q;
  hipLaunchKernelGGL(spectral_setup_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, a);
  return check_hip(hipGetLastError(), "spectral setup launch");
}

}  // extern "C"
