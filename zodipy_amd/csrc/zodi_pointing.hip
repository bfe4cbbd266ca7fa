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

// (z, sin(theta), phi) of RING pixel p: rings 1 .. nside - 1 form the north cap
// (4 i pixels each), rings nside .. 3 nside the belt (4 nside pixels each), the
// rest the south cap. In the caps 1 - |z| = i^2 / (3 nside^2) gives sin(theta)
// without the cancellation of 1 - z^2. The ring index from the square root is
// corrected by one where rounding put it off.
__device__ __forceinline__ void ring_pixel(int64_t ns, int64_t p, double& z, double& sth,
                                           double& phi) {
  const int64_t ncap = 2 * ns * (ns - 1);
  const int64_t npix = 12 * ns * ns;
  const double fact2 = 1.0 / (3.0 * (double)ns * (double)ns);
  const double half_pi = 1.5707963267948966;
  if (p < ncap) {
    int64_t i = (1 + (int64_t)floor(sqrt(1.0 + 2.0 * (double)p))) / 2;
    if (2 * i * (i - 1) > p) --i;
    if (2 * (i + 1) * i <= p) ++i;
    const int64_t j = p + 1 - 2 * i * (i - 1);
    const double t = (double)i * (double)i * fact2;
    z = 1.0 - t;
    sth = sqrt(t * (2.0 - t));
    phi = ((double)j - 0.5) * half_pi / (double)i;
  } else if (p >= npix - ncap) {
    const int64_t q = npix - p;
    int64_t i = (1 + (int64_t)floor(sqrt(2.0 * (double)q - 1.0))) / 2;
    if (2 * i * (i - 1) >= q) --i;
    if (2 * (i + 1) * i < q) ++i;
    const int64_t j = 4 * i + 1 - (q - 2 * i * (i - 1));
    const double t = (double)i * (double)i * fact2;
    z = -(1.0 - t);
    sth = sqrt(t * (2.0 - t));
    phi = ((double)j - 0.5) * half_pi / (double)i;
  } else {
    const int64_t b = p - ncap;
    const int64_t i = b / (4 * ns) + ns;
    const int64_t j = b % (4 * ns) + 1;
    const double fodd = ((i + ns) & 1) ? 1.0 : 0.5;
    z = (double)(2 * ns - i) * (2.0 / 3.0) / (double)ns;
    sth = sqrt((1.0 - z) * (1.0 + z));
    phi = ((double)j - fodd) * half_pi / (double)ns;
  }
}

__global__ __launch_bounds__(256) void pix2vec_kernel(int64_t nside, const int64_t* __restrict__ pix,
                                                      int64_t first, int64_t n, Rot3 R,
                                                      int identity, double* __restrict__ out) {
  const int64_t npix = 12 * nside * nside;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t p = pix != nullptr ? pix[i] : first + i;
    double x, y, z;
    if (p < 0 || p >= npix) {
      x = y = z = __builtin_nan("");
    } else {
      double sth, phi, s, c;
      ring_pixel(nside, p, z, sth, phi);
      sincos(phi, &s, &c);
      x = sth * c;
      y = sth * s;
    }
    store_rotated(R, identity != 0, x, y, z, out, n, i);
  }
}

__global__ __launch_bounds__(256) void lonlat_kernel(const double* __restrict__ lon,
                                                     const double* __restrict__ lat, int64_t n,
                                                     double scale, Rot3 R, int identity,
                                                     double* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double sl, cl, sb, cb;
    sincos(lon[i] * scale, &sl, &cl);
    sincos(lat[i] * scale, &sb, &cb);
    store_rotated(R, identity != 0, cb * cl, cb * sl, sb, out, n, i);
  }
}

// Positions along a tabulated track: the cubic spline of an hourly ephemeris
// grid (zodipy/bodies.py:22-35 builds scipy's CubicSpline over it), in scipy's
// piecewise-polynomial form. The interval is scipy's (the last breakpoint at or
// below t, clipped to the first and last intervals, which then extrapolate) and
// the polynomial is summed in increasing powers of t - x_k as scipy does.
// The interval is first guessed as if the grid were uniform (the hourly grids
// of time-ordered data are), checked against its two breakpoints, and only
// searched by bisection when the guess misses.
__global__ __launch_bounds__(256) void spline_kernel(const double* __restrict__ x, int64_t m,
                                                     const double* __restrict__ c,
                                                     const double* __restrict__ t, int64_t n,
                                                     double* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t ni = m - 1;  // intervals
  const double x0 = x[0];
  const double xl = x[m - 1];
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
      if (!(x[k] <= ti && ti < x[k + 1])) {
        int64_t lo = 0, hi = m - 1;  // x[lo] <= ti < x[hi]
        while (hi - lo > 1) {
          const int64_t mid = (lo + hi) >> 1;
          if (x[mid] <= ti) {
            lo = mid;
          } else {
            hi = mid;
          }
        }
        k = lo;
      }
    }
    {
      // c is [4][ni][3]; c[0] multiplies s^3. Separately rounded products and sums
      // (no fused multiply-add), term by term in increasing powers, as scipy does.
#pragma clang fp contract(off)
      const double s = ti - x[k];
      const double s2 = s * s;
      const double s3 = s2 * s;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double res = c[(3 * ni + k) * 3 + d];
        res = res + c[(2 * ni + k) * 3 + d] * s;
        res = res + c[(1 * ni + k) * 3 + d] * s2;
        res = res + c[k * 3 + d] * s3;
        out[d * n + i] = res;
      }
    }
  }
}

namespace {

int64_t elementwise_grid(int64_t n) {
  const int64_t blocks = (n + 255) / 256;
  return std::max<int64_t>(1, std::min<int64_t>(blocks, 16 * (int64_t)cu_count()));
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

zodi_status zodi_spline_positions(const double* x, int64_t m, const double* c, const double* t,
                                  int64_t n, double* out, void* stream) {
  if (m < 2) return fail(ZODI_ERR_ARG, "a track needs at least 2 breakpoints, got %lld",
                         (long long)m);
  if (n < 0) return fail(ZODI_ERR_ARG, "n < 0");
  if (n == 0) return ZODI_OK;
  if (x == nullptr || c == nullptr || t == nullptr || out == nullptr)
    return fail(ZODI_ERR_ARG, "null pointer");
  hipLaunchKernelGGL(spline_kernel, dim3((unsigned)elementwise_grid(n)), dim3(256), 0,
                     (hipStream_t)stream, x, m, c, t, n, out);
  return check_hip(hipGetLastError(), "spline launch");
}

}  // extern "C"
