// Pointing ingestion on the device (gfx950, FP64): ecliptic unit vectors for
// the line-of-sight kernels, written SoA [3][n] straight into HBM.
//
// * zodi_pix2vec: centres of HEALPix RING-scheme pixels (Gorski et al. 2005,
//   ApJ 622, 759, eqs. 4-5), the pointings of the reference's HEALPix map
//   examples (docs/examples/healpy_map.py:16 builds them with healpy's pix2ang);
// * zodi_lonlat_to_vec: longitude/latitude pairs, as SkyCoord holds them;
// * zodi_spline_positions: observer/Earth positions per ray for time-ordered
//   data, from the cubic spline of an hourly ephemeris grid.
//
// Both apply a 3x3 frame rotation to the mean ecliptic and equinox of J2000,
// the work of skycoord.transform_to(BarycentricMeanEcliptic) and .cartesian in
// the reference (zodipy/model.py:247-251); the rotation matrices themselves are
// built on the host. Each element is independent: grid-stride loops, coalesced
// loads and stores, HBM bound (24 bytes written per element, plus 8 or 16 read).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "zodi_common.hpp"

namespace zodi {

struct Rot3 {
  double m[9];  // row-major
};

__device__ __forceinline__ void store_rotated(const Rot3& R, bool identity, double x, double y,
                                              double z, double* __restrict__ out, int64_t n,
                                              int64_t i) {
  if (identity) {
    out[i] = x;
    out[n + i] = y;
    out[2 * n + i] = z;
  } else {
    out[i] = R.m[0] * x + R.m[1] * y + R.m[2] * z;
    out[n + i] = R.m[3] * x + R.m[4] * y + R.m[5] * z;
    out[2 * n + i] = R.m[6] * x + R.m[7] * y + R.m[8] * z;
  }
}

// Per-launch constants of a RING-scheme map.
struct RingGeometry {
  int64_t ns, ncap, npix;
  double fact2;    // 1 / (3 ns^2)
  double inv_4ns;  // 1 / (4 ns), for the belt's ring estimate only
};

__device__ __forceinline__ RingGeometry ring_geometry(int64_t ns) {
  RingGeometry g;
  g.ns = ns;
  g.ncap = 2 * ns * (ns - 1);
  g.npix = 12 * ns * ns;
  g.fact2 = 1.0 / (3.0 * (double)ns * (double)ns);
  g.inv_4ns = 1.0 / (4.0 * (double)ns);
  return g;
}

// (z, sin(theta), phi) of RING pixel p: rings 1 .. nside - 1 form the north
// cap (4 i pixels each), rings nside .. 3 nside the belt (4 nside pixels each),
// the rest the south cap. In the caps 1 - |z| = i^2 / (3 nside^2) gives
// sin(theta) without the cancellation of 1 - z^2. The ring index from the
// square root, and the belt's ring from a floating-point quotient (exact to
// within one for every map of at most 12 * 2^58 pixels), are corrected by one
// where rounding put them off; no 64-bit integer division is needed. z and phi
// are formed with the same operations, in the same order, as the host's
// zodipy_amd.healpix.pix2ang_ring.
__device__ __forceinline__ void ring_pixel(const RingGeometry& g, int64_t p, double& z,
                                           double& sth, double& phi) {
  const double half_pi = 1.5707963267948966;
  if (p < g.ncap) {
    int64_t i = (1 + (int64_t)floor(sqrt(1.0 + 2.0 * (double)p))) / 2;
    if (2 * i * (i - 1) > p) --i;
    if (2 * (i + 1) * i <= p) ++i;
    const int64_t j = p + 1 - 2 * i * (i - 1);
    const double t = (double)i * (double)i * g.fact2;
    z = 1.0 - t;
    sth = sqrt(t * (2.0 - t));
    phi = ((double)j - 0.5) * half_pi / (double)i;
  } else if (p >= g.npix - g.ncap) {
    const int64_t q = g.npix - p;
    int64_t i = (1 + (int64_t)floor(sqrt(2.0 * (double)q - 1.0))) / 2;
    if (2 * i * (i - 1) >= q) --i;
    if (2 * (i + 1) * i < q) ++i;
    const int64_t j = 4 * i + 1 - (q - 2 * i * (i - 1));
    const double t = (double)i * (double)i * g.fact2;
    z = -(1.0 - t);
    sth = sqrt(t * (2.0 - t));
    phi = ((double)j - 0.5) * half_pi / (double)i;
  } else {
    const int64_t b = p - g.ncap;
    const int64_t w = 4 * g.ns;
    int64_t r = (int64_t)((double)b * g.inv_4ns);  // ring within the belt
    int64_t c = b - r * w;                          // pixel within the ring
    if (c < 0) {
      --r;
      c += w;
    } else if (c >= w) {
      ++r;
      c -= w;
    }
    const int64_t i = r + g.ns;
    const double fodd = ((i + g.ns) & 1) ? 1.0 : 0.5;
    z = (double)(2 * g.ns - i) * (2.0 / 3.0) / (double)g.ns;
    sth = sqrt((1.0 - z) * (1.0 + z));
    phi = ((double)(c + 1) - fodd) * half_pi / (double)g.ns;
  }
}

__global__ __launch_bounds__(256) void pix2vec_kernel(int64_t nside, const int64_t* __restrict__ pix,
                                                      int64_t first, int64_t n, Rot3 R,
                                                      int identity, double* __restrict__ out) {
  const RingGeometry g = ring_geometry(nside);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t p = pix != nullptr ? pix[i] : first + i;
    double x, y, z;
    if (p < 0 || p >= g.npix) {
      x = y = z = __builtin_nan("");
    } else {
      double sth, phi, s, c;
      ring_pixel(g, p, z, sth, phi);
      zodi_trig::sincos_moderate(phi, &s, &c);  // phi in [0, 2 pi)
      x = sth * c;
      y = sth * s;
    }
    store_rotated(R, identity != 0, x, y, z, out, n, i);
  }
}

// sin and cos of an angle in radians: the moderate-argument form for |x| <= 64
// (any longitude or latitude given in [-360, 720] degrees), the library's
// otherwise (and for non-finite x).
__device__ __forceinline__ void sincos_angle(double x, double* s, double* c) {
  if (fabs(x) <= 64.0) {
    zodi_trig::sincos_moderate(x, s, c);
  } else {
    sincos(x, s, c);
  }
}

// Degrees are scaled by the double nearest pi/180 before sincos, the rounding
// of the reference's degree-to-radian conversion.
__global__ __launch_bounds__(256) void lonlat_kernel(const double* __restrict__ lon,
                                                     const double* __restrict__ lat, int64_t n,
                                                     double scale, Rot3 R, int identity,
                                                     double* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double sl, cl, sb, cb;
    sincos_angle(lon[i] * scale, &sl, &cl);
    sincos_angle(lat[i] * scale, &sb, &cb);
    store_rotated(R, identity != 0, cb * cl, cb * sl, sb, out, n, i);
  }
}

// Positions along a tabulated track: the cubic spline of an hourly ephemeris
// grid (zodipy/bodies.py:22-35 builds scipy's CubicSpline over it), in scipy's
// piecewise-polynomial form. The interval is scipy's (the last breakpoint at or
// below t, clipped to the first and last intervals, which then extrapolate) and
// the polynomial is summed in increasing powers of t - x_k as scipy does.
//
// Each interval is one 128-byte record (kSplineRecord doubles): x_k, the
// constant, linear, quadratic and cubic coefficients of x, y, z, then x_{k+1}
// and two zeros. Times in random order then gather one cache line per ray
// instead of six. The interval is first guessed as if the grid were uniform
// (the hourly grids of time-ordered data are), checked against the record's
// own two breakpoints, and searched by bisection only when the guess misses.
constexpr int kSplineRecord = 16;

__device__ __forceinline__ void load_record(const double* __restrict__ rec, int64_t k,
                                            double (&v)[14]) {
  const double2* r2 = reinterpret_cast<const double2*>(rec + k * kSplineRecord);
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const double2 w = r2[q];
    v[2 * q] = w.x;
    v[2 * q + 1] = w.y;
  }
}

__global__ __launch_bounds__(256) void spline_kernel(const double* __restrict__ rec, int64_t ni,
                                                     const double* __restrict__ t, int64_t n,
                                                     double* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const double x0 = rec[0];
  const double xl = rec[(ni - 1) * kSplineRecord + 13];
  // Reached only with x0 <= t < xl, so the spacing is positive there; a guess
  // outside [0, ni - 1] (an infinite or NaN quotient) is clipped below.
  const double inv_step = (double)ni / (xl - x0);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double ti = t[i];
    int64_t k;
    if (!(ti >= x0)) {
      k = 0;  // below the grid, or NaN (which then propagates)
    } else if (ti >= xl) {
      k = ni - 1;
    } else {
      const double g = (ti - x0) * inv_step;
      k = (g >= 0.0 && g < (double)(ni - 1)) ? (int64_t)g : ni - 1;
    }
    // The guessed record is loaded whole at once (one round trip to memory) and
    // kept unless its breakpoints show the guess missed.
    double v[14];
    load_record(rec, k, v);
    if (ti >= x0 && ti < xl && !(v[0] <= ti && ti < v[13])) {
      int64_t lo = 0, hi = ni;  // breakpoints: x_lo <= ti < x_hi
      while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;  // lo < mid < hi <= ni: a record's x_k
        if (rec[mid * kSplineRecord] <= ti) {
          lo = mid;
        } else {
          hi = mid;
        }
      }
      load_record(rec, lo, v);
    }
    {
      // Separately rounded products and sums (no fused multiply-add), term by
      // term in increasing powers, as scipy does.
#pragma clang fp contract(off)
      const double s = ti - v[0];
      const double s2 = s * s;
      const double s3 = s2 * s;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double res = v[1 + d];
        res = res + v[4 + d] * s;
        res = res + v[7 + d] * s2;
        res = res + v[10 + d] * s3;
        // streamed out without allocating in L2, which keeps the lines for the
        // track records the next rays read (sorted times 0.52 -> 0.45 ms,
        // random 0.94 -> 0.88 ms on 50.7 M times)
        __builtin_nontemporal_store(res, out + d * n + i);
      }
    }
  }
}

// Largest observer distance to the Sun over per-ray positions [3][n]: the
// reference tests (r_obs > cutoff).any() over all observers of one call
// (zodipy/line_of_sight.py:71-73). Block b reduces r^2 = x^2 + y^2 + z^2 over a
// grid-stride share (NaN compares false and is skipped) into part[b]; one block
// then reduces the partials. sqrt is monotone and correctly rounded, so
// sqrt(max r^2) is the maximum of the rounded distances.
constexpr int kMaxRadiusBlocks = 1024;

__global__ __launch_bounds__(256) void max_sq_radius_kernel(const double* __restrict__ xyz,
                                                            int64_t n,
                                                            double* __restrict__ part) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  double m = -1.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + t; i < n; i += (int64_t)gridDim.x * 256) {
    const double x = xyz[i], y = xyz[n + i], z = xyz[2 * n + i];
    double r2;
    {
      // rounded term by term like the reference's x**2 + y**2 + z**2 (no FMA), so
      // the cutoff comparison sees the reference's distance bit for bit
#pragma clang fp contract(off)
      r2 = x * x + y * y + z * z;
    }
    m = r2 > m ? r2 : m;
  }
  red[t] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] = red[t + w] > red[t] ? red[t + w] : red[t];
    __syncthreads();
  }
  if (t == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void max_radius_finish(const double* __restrict__ part,
                                                         int n_part, double* __restrict__ out) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  double m = -1.0;
  for (int b = t; b < n_part; b += 256) m = part[b] > m ? part[b] : m;
  red[t] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] = red[t + w] > red[t] ? red[t + w] : red[t];
    __syncthreads();
  }
  if (t == 0) out[0] = red[0] >= 0.0 ? sqrt(red[0]) : -1.0;
}

namespace {

// Up to 256 blocks of 256 threads per CU, each thread looping over a few
// elements: on 50M-element inputs this beats both 16 blocks per CU (too few
// loads in flight) and one element per thread (per-thread set-up no longer
// amortised) by 5-15%.
int64_t elementwise_grid(int64_t n) {
  const int64_t blocks = (n + 255) / 256;
  return std::max<int64_t>(1, std::min<int64_t>(blocks, 256 * (int64_t)cu_count()));
}

void load_rotation(const double* rot, Rot3& R, int& identity) {
  identity = rot == nullptr ? 1 : 0;
  for (int k = 0; k < 9; ++k) R.m[k] = rot != nullptr ? rot[k] : (k % 4 == 0 ? 1.0 : 0.0);
}

}  // namespace
}  // namespace zodi

using namespace zodi;

extern "C" {

zodi_status zodi_pix2vec(int64_t nside, const int64_t* pix, int64_t first, int64_t n,
                         const double* rot, double* out, void* stream) {
  if (nside < 1 || nside > (int64_t(1) << 29))
    return fail(ZODI_ERR_ARG, "nside %lld outside [1, 2^29]", (long long)nside);
  if (n < 0) return fail(ZODI_ERR_ARG, "n < 0");
  const int64_t npix = 12 * nside * nside;
  if (pix == nullptr && (first < 0 || first > npix || n > npix - first))
    return fail(ZODI_ERR_ARG, "pixel range [%lld, %lld) outside [0, %lld)", (long long)first,
                (long long)(first + n), (long long)npix);
  if (n == 0) return ZODI_OK;
  if (out == nullptr) return fail(ZODI_ERR_ARG, "out is null");
  Rot3 R;
  int identity;
  load_rotation(rot, R, identity);
  hipLaunchKernelGGL(pix2vec_kernel, dim3((unsigned)elementwise_grid(n)), dim3(256), 0,
                     (hipStream_t)stream, nside, pix, first, n, R, identity, out);
  return check_hip(hipGetLastError(), "pix2vec launch");
}

zodi_status zodi_lonlat_to_vec(const double* lon, const double* lat, int64_t n, int32_t degrees,
                               const double* rot, double* out, void* stream) {
  if (n < 0) return fail(ZODI_ERR_ARG, "n < 0");
  if (n > 0 && (lon == nullptr || lat == nullptr || out == nullptr))
    return fail(ZODI_ERR_ARG, "null pointer");
  if (n == 0) return ZODI_OK;
  Rot3 R;
  int identity;
  load_rotation(rot, R, identity);
  const double scale = degrees ? 0.017453292519943295 : 1.0;
  hipLaunchKernelGGL(lonlat_kernel, dim3((unsigned)elementwise_grid(n)), dim3(256), 0,
                     (hipStream_t)stream, lon, lat, n, scale, R, identity, out);
  return check_hip(hipGetLastError(), "lonlat launch");
}

zodi_status zodi_spline_positions(const double* records, int64_t n_intervals, const double* t,
                                  int64_t n, double* out, void* stream) {
  if (n_intervals < 1) return fail(ZODI_ERR_ARG, "a track needs at least 1 interval, got %lld",
                                   (long long)n_intervals);
  if (n < 0) return fail(ZODI_ERR_ARG, "n < 0");
  if (n == 0) return ZODI_OK;
  if (records == nullptr || t == nullptr || out == nullptr)
    return fail(ZODI_ERR_ARG, "null pointer");
  if (reinterpret_cast<uintptr_t>(records) % 16 != 0)
    return fail(ZODI_ERR_ARG, "track records must be 16-byte aligned");
  hipLaunchKernelGGL(spline_kernel, dim3((unsigned)elementwise_grid(n)), dim3(256), 0,
                     (hipStream_t)stream, records, n_intervals, t, n, out);
  return check_hip(hipGetLastError(), "spline launch");
}

zodi_status zodi_max_radius(const double* xyz, int64_t n, double* out, void* stream) {
  if (n < 0) return fail(ZODI_ERR_ARG, "n < 0");
  if (out == nullptr) return fail(ZODI_ERR_ARG, "out is null");
  if (n > 0 && xyz == nullptr) return fail(ZODI_ERR_ARG, "xyz is null");
  const int blocks =
      (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, kMaxRadiusBlocks));
  if (n > 0) {
    hipLaunchKernelGGL(max_sq_radius_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       xyz, n, out + 1);
  }
  hipLaunchKernelGGL(max_radius_finish, dim3(1), dim3(256), 0, (hipStream_t)stream,
                     (const double*)(out + 1), n > 0 ? blocks : 0, out);
  return check_hip(hipGetLastError(), "max radius launch");
}

}  // extern "C"
