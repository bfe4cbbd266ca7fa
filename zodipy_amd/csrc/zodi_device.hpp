// Device functions of the zodiacal line-of-sight integrand (gfx950, FP64 VALU).
//
// Everything here is per-lane scalar FP64 code; model parameters are read from
// the device model block with wave-uniform addresses, so they are fetched by
// scalar loads into SGPRs. Each function cites the reference expression it
// evaluates (paths relative to the zodipy 1.1.3 tree).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zodi_hip.h"

namespace zodi {

// np.finfo(np.float64).eps (zodipy/line_of_sight.py:14, 73)
constexpr double kEps = 2.220446049250313080847e-16;
constexpr double kPi = 3.141592653589793116;        // np.pi
constexpr double kTwoPi = 6.283185307179586232;     // 2 * np.pi
constexpr double kRad2Deg = 57.295779513082322865;  // 180.0 / np.pi (np.rad2deg)

constexpr int kMaxSlots = 4;  // components per group held in registers

// np.maximum: returns the first argument when it is NaN, propagates NaN.
__device__ __forceinline__ double np_maximum(double a, double b) {
  return (a >= b || a != a) ? a : b;
}

// Python/NumPy float remainder a % b for b > 0: fmod, then shift negative
// results into [0, b) (numpy npy_divmod). fmod is exact.
__device__ __forceinline__ double floor_mod_pos(double a, double b) {
  double m = fmod(a, b);
  if (m < 0.0) {
    m += b;
  } else if (m == 0.0) {
    m = 0.0;
  }
  return m;
}

// Per-ray quantities shared by every cutoff sphere: the linear coefficient b of
// the ray/sphere quadratic and the observer's heliocentric distance.
// zodipy/line_of_sight.py:70-80: b = 2 (x0 cos(lat) cos(lon) + y0 cos(lat) sin(lon)),
// with lon = atan2(u_y, u_x), lat = asin(u_z); z0 does not enter b.
struct RayGeom {
  double b;
  double r_obs;
};

__device__ __forceinline__ RayGeom ray_geometry(double ux, double uy, double uz, double ox,
                                                double oy, double oz) {
  const double lon = atan2(uy, ux);
  const double lat = asin(uz);
  const double cos_lat = cos(lat);
  double sl, cl;
  sincos(lon, &sl, &cl);
  RayGeom g;
  g.b = 2.0 * (ox * cos_lat * cl + oy * cos_lat * sl);
  g.r_obs = sqrt(ox * ox + oy * oy + oz * oz);
  return g;
}

// Distance from the observer to the heliocentric sphere of radius `cutoff`
// (zodipy/line_of_sight.py:81-85): c = r_obs^2 - cutoff^2,
// q = -b/2 (1 + sqrt(b^2 - 4c)/|b|), result max(q, c/q). b == 0 gives NaN.
__device__ __forceinline__ double sphere_distance(const RayGeom& g, double cutoff) {
  const double c = g.r_obs * g.r_obs - cutoff * cutoff;
  const double q = -0.5 * g.b * (1.0 + sqrt(g.b * g.b - 4.0 * c) / fabs(g.b));
  return np_maximum(q, c / q);
}

// Linear interpolation in the Planck table with np.interp semantics: NaN in,
// NaN out; clamped to the end values outside [T0, T_last]; exact table values
// at the nodes. Tables live in LDS.
__device__ __forceinline__ double planck_lookup(double T, const double* __restrict__ tT,
                                                const double* __restrict__ tB,
                                                const double* __restrict__ tS, int n,
                                                double inv_step) {
  if (!(T == T)) return T;
  if (T < tT[0]) return tB[0];
  if (T >= tT[n - 1]) return tB[n - 1];
  int j = (int)((T - tT[0]) * inv_step);
  j = j > n - 2 ? n - 2 : j;
  if (T < tT[j]) {
    --j;
  } else if (T >= tT[j + 1]) {
    ++j;
  }
  if (j >= n - 1) return tB[n - 1];
  const double Tj = tT[j];
  if (T == Tj) return tB[j];
  return tS[j] * (T - Tj) + tB[j];
}

// Heliocentric position in a component's frame: offset by X_0 unless the
// component is centred on the Sun (then it is exactly the heliocentric one).
struct CompFrame {
  double x, y, z, r, zc;  // offset position, its norm, height above the symmetry plane
};

__device__ __forceinline__ CompFrame comp_frame(const double* __restrict__ P, uint32_t flags,
                                                double X0, double X1, double X2, double R) {
  CompFrame f;
  if (flags & ZODI_CF_CENTERED) {
    f.x = X0;
    f.y = X1;
    f.z = X2;
    f.r = R;
  } else {
    f.x = X0 - P[ZODI_P_X0];
    f.y = X1 - P[ZODI_P_Y0];
    f.z = X2 - P[ZODI_P_Z0];
    f.r = sqrt(f.x * f.x + f.y * f.y + f.z * f.z);
  }
  // Z = X'_x sin(Omega) sin(i) - X'_y cos(Omega) sin(i) + X'_z cos(i)
  f.zc = f.x * P[ZODI_P_NX] - f.y * P[ZODI_P_NY] + f.z * P[ZODI_P_NZ];
  return f;
}

// ---------------------------------------------------------------- densities
// zodipy/number_density.py:47-73
__device__ __forceinline__ double cloud_density(const double* __restrict__ P, const CompFrame& f) {
  const double n0 = P[6], alpha = P[7], beta = P[8], gamma = P[9], mu = P[10];
  const double zeta = fabs(f.zc / f.r);
  const double g = zeta < mu ? (zeta * zeta) / (2.0 * mu) : zeta - mu / 2.0;
  return n0 * pow(f.r, -alpha) * exp(-beta * pow(g, gamma));
}

__device__ __forceinline__ double pow20(double x) {
  const double x2 = x * x;
  const double x4 = x2 * x2;
  const double x8 = x4 * x4;
  const double x16 = x8 * x8;
  return x16 * x4;
}

// zodipy/number_density.py:76-110 (P[6] holds 3*n_0; the 1/v factor as in the reference)
__device__ __forceinline__ double band_density(const double* __restrict__ P, uint32_t flags,
                                               const CompFrame& f) {
  const double three_n0 = P[6], dzeta = P[7], p = P[8], v = P[9], dr = P[10];
  const double zeta = fabs(f.zc / f.r);
  const double zd = zeta / dzeta;
  const double zd2 = zd * zd;
  const double zd4 = zd2 * zd2;
  const double zd6 = zd4 * zd2;
  const double term1 = three_n0 / f.r;
  const double term2 = exp(-zd6);
  const double zdp = (flags & ZODI_CF_P_IS_4) ? zd4 : pow(zd, p);
  const double term3 = 1.0 + zdp / v;
  const double term4 = 1.0 - exp(-pow20(f.r / dr));
  return term1 * term2 * term3 * term4;
}

// zodipy/number_density.py:113-139
__device__ __forceinline__ double ring_density(const double* __restrict__ P, const CompFrame& f) {
  const double n0 = P[6], Rr = P[7], sr = P[8], sz = P[9];
  const double d = f.r - Rr;
  const double term1 = -(d * d) / (sr * sr);
  const double term2 = fabs(f.zc) / sz;
  return n0 * exp(term1 - term2);
}

// zodipy/number_density.py:142-181; `earth_lon` = atan2(E'_y, E'_x) with E' = X_earth - X_0.
__device__ __forceinline__ double feature_density(const double* __restrict__ P, const CompFrame& f,
                                                  double earth_lon) {
  const double n0 = P[6], Rf = P[7], sr = P[8], sz = P[9], th = P[10], sth = P[11];
  const double theta_comp = atan2(f.y, f.x) - earth_lon;
  double dth = theta_comp - th;
  dth = floor_mod_pos(dth + kPi, kTwoPi) - kPi;
  const double d = f.r - Rf;
  double e = (d * d) / (sr * sr);
  e += fabs(f.zc) / sz;
  e += (dth * dth) / (sth * sth);
  return n0 * exp(-e);
}

// zodipy/number_density.py:184-218 (zero outside R <= R_outer)
__device__ __forceinline__ double fan_density(const double* __restrict__ P, const CompFrame& f) {
  const double Q = P[6], Pp = P[7], gamma = P[8], Z0 = P[9], Rout = P[10];
  if (!(f.r <= Rout)) return 0.0;
  const double beta = asin(f.zc / f.r);
  const double za = fabs(f.zc);
  const double eps = za < Z0 ? 2.0 - za / Z0 : 1.0;
  const double fv = pow(cos(beta), Q) * exp(-Pp * sin(pow(fabs(beta), eps)));
  return pow(f.r, -gamma) * fv;
}

// zodipy/number_density.py:221-256 (zero outside R_inner <= R <= R_outer)
__device__ __forceinline__ double comet_density(const double* __restrict__ P, const CompFrame& f) {
  const double gamma = P[6], Z0 = P[7], Pp = P[8], amp = P[9], Rin = P[10], Rout = P[11];
  if (!(f.r >= Rin && f.r <= Rout)) return 0.0;
  const double beta = asin(f.zc / f.r);
  const double za = fabs(f.zc);
  const double eps = za < Z0 ? 2.0 - za / Z0 : 1.0;
  const double fv = exp(-Pp * sin(pow(fabs(beta), eps)));
  return amp * pow(f.r, -gamma) * fv;
}

// zodipy/number_density.py:267-304 (latitude in degrees)
__device__ __forceinline__ double narrow_band_density(const double* __restrict__ P,
                                                      const CompFrame& f) {
  const double bnb = P[6], G = P[7], gamma = P[8], A = P[9], Rin = P[10], Rout = P[11];
  if (!(f.r >= Rin && f.r <= Rout)) return 0.0;
  const double babs = fabs(asin(f.zc / f.r) * kRad2Deg);
  const double fv = babs < bnb ? exp(G * (babs - bnb)) : 0.0;
  return A * pow(f.r / Rout, -gamma) * fv;
}

// zodipy/number_density.py:307-342 (latitude in degrees)
__device__ __forceinline__ double broad_band_density(const double* __restrict__ P,
                                                     const CompFrame& f) {
  const double bbb = P[6], sbb = P[7], gamma = P[8], A = P[9], Rin = P[10], Rout = P[11];
  if (!(f.r >= Rin && f.r <= Rout)) return 0.0;
  const double beta = asin(f.zc / f.r) * kRad2Deg;
  const double a = (beta - bbb) / sbb;
  const double c = (beta + bbb) / sbb;
  const double fv = exp(-0.5 * (a * a)) + exp(-0.5 * (c * c));
  return A * pow(f.r / Rout, -gamma) * fv;
}

// Density of component `kind` at heliocentric X (|X| = R).
__device__ __forceinline__ double density(int kind, const double* __restrict__ P, uint32_t flags,
                                          double X0, double X1, double X2, double R,
                                          double earth_lon) {
  if (kind == ZODI_INTERSTELLAR) return P[6];  // zodipy/number_density.py:259-264
  const CompFrame f = comp_frame(P, flags, X0, X1, X2, R);
  switch (kind) {
    case ZODI_CLOUD:
      return cloud_density(P, f);
    case ZODI_BAND:
      return band_density(P, flags, f);
    case ZODI_RING:
      return ring_density(P, f);
    case ZODI_FEATURE:
      return feature_density(P, f, earth_lon);
    case ZODI_FAN:
      return fan_density(P, f);
    case ZODI_COMET:
      return comet_density(P, f);
    case ZODI_NARROW_BAND:
      return narrow_band_density(P, f);
    case ZODI_BROAD_BAND:
      return broad_band_density(P, f);
    case ZODI_RING_RRM:  // zodipy/number_density.py:345-370
      return P[10] * ring_density(P, f);
    case ZODI_FEATURE_RRM:  // zodipy/number_density.py:373-404
      return P[12] * feature_density(P, f, earth_lon);
    default:
      return __builtin_nan("");
  }
}

__device__ __forceinline__ bool is_feature(int kind) {
  return kind == ZODI_FEATURE || kind == ZODI_FEATURE_RRM;
}

// Scattering phase function at the node (zodipy/scattering.py:11-50):
// Theta = arccos(-clip(X_los . X_helio / (R_los R_helio), -1, 1)),
// Phi = N (C1 + C2 Theta + exp(C3 Theta)).
__device__ __forceinline__ double phase_function(double R_los, double ux, double uy, double uz,
                                                 double X0, double X1, double X2, double R,
                                                 const double* __restrict__ S) {
  const double lx = R_los * ux, ly = R_los * uy, lz = R_los * uz;
  double ct = (lx * X0 + ly * X1 + lz * X2) / (R_los * R);
  ct = ct != ct ? ct : (ct < -1.0 ? -1.0 : (ct > 1.0 ? 1.0 : ct));
  const double theta = acos(-ct);
  return S[ZODI_S_PHASE_NORM] * (S[ZODI_S_C1] + S[ZODI_S_C2] * theta + exp(S[ZODI_S_C3] * theta));
}

}  // namespace zodi
