// Device functions of the zodiacal line-of-sight integrand (gfx950, FP64 VALU).
//
// Everything here is per-lane scalar FP64 code; model parameters are read from
// the device model block with wave-uniform addresses (scalar loads into SGPRs).
// Divisions by parameters use reciprocals precomputed on the host, and each
// node forms one 1/R shared by its components, so the node loop is free of
// IEEE FP64 division sequences except for that one. Each function cites the
// reference expression it evaluates (paths relative to the zodipy 1.1.3 tree);
// reassociations change results by a few ulp, far below the 1e-10 parity bound.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zodi_hip.h"
#include "zodi_math.hpp"
#include "zodi_trig.hpp"

namespace zodi {

// np.finfo(np.float64).eps (zodipy/line_of_sight.py:14, 73)
constexpr double kEps = 2.220446049250313080847e-16;
constexpr double kPi = 3.141592653589793116;        // np.pi
constexpr double kTwoPi = 6.283185307179586232;     // 2 * np.pi
constexpr double kRad2Deg = 57.295779513082322865;  // 180.0 / np.pi (np.rad2deg)

constexpr int kMaxSlots = 4;  // components per group held in registers

// Parameter block accessor: loads through the constant address space are
// known not to be written by the kernel, so wave-uniform loads become scalar.
typedef const __attribute__((address_space(4))) zodi_model* ModelPtr;
typedef const __attribute__((address_space(4))) double* ParamPtr;

__device__ __forceinline__ ModelPtr as_const(const zodi_model* m) { return (ModelPtr)m; }

// The same pointer, made opaque to the optimizer (an empty asm statement, no
// instruction): loads through it inside a loop are not hoisted out of it, so
// the many uniform parameters are re-read by scalar loads each iteration
// instead of being held across the loop, which would overflow the SGPR budget
// into VGPRs.
__device__ __forceinline__ ModelPtr opaque(ModelPtr m) {
  uint64_t v = (uint64_t)m;
  asm volatile("" : "+s"(v));
  return (ModelPtr)v;
}

// np.maximum: returns the first argument when it is NaN, propagates NaN.
__device__ __forceinline__ double np_maximum(double a, double b) {
  return (a >= b || a != a) ? a : b;
}

// x^y for x >= 0 (every power in the model has a non-negative base) as
// exp_fast(y log x). The result differs from a correctly rounded pow by about
// |y log x| ulp; x = 0 gives 0 for y > 0; y = 0 gives 1 like pow.
__device__ __forceinline__ double pow_pos(double x, double y) {
  return y == 0.0 ? 1.0 : exp_fast(y * log_nonneg(x));
}

// x^y for a strictly positive finite base (radii): |y log x| <= 745 |y|.
__device__ __forceinline__ double pow_radius(double x, double y) {
  return y == 0.0 ? 1.0 : exp_fast_bounded(y * log_pos(x));
}

// |X| and 1/|X| from |X|^2: a hardware reciprocal square root refined by two
// Newton steps (y <- y (3 - a y^2) / 2), then |X| = a y. Both within ~1 ulp.
// For per-ray and per-point quantities.
__device__ __forceinline__ void norm_and_inverse(double a, double& r, double& inv_r) {
  double y = __builtin_amdgcn_rsq(a);
  double t = a * y;
  double h = 0.5 * y;
  double e = fma(-t, h, 0.5);
  y = fma(y, e, y);  // refined 1/sqrt(a)
  t = fma(t, e, t);  // refined sqrt(a)
  h = 0.5 * y;
  const double d = fma(-t, t, a);
  t = fma(d, h, t);  // sqrt(a), last Newton correction
  e = fma(-t, y, 1.0);
  y = fma(y, e, y);
  r = t;
  inv_r = y;
}

// The same for the quadrature nodes, with one Newton step: the hardware
// estimate's relative error is below 2^-24 (measured on gfx950 over 2^20
// arguments spanning 2^-20..2^20), and one coupled step y <- y (1 + e),
// t <- t (1 + e), e = (1 - a y^2)/2, t = a y, leaves 1.5 eps^2 < 5e-15 for both.
// Every node quantity then moves by about 1e-14 relative or less, which the
// integrals' 1e-10 parity bound leaves untouched.
__device__ __forceinline__ void norm_and_inverse_node(double a, double& r, double& inv_r) {
  const double y = __builtin_amdgcn_rsq(a);
  const double t = a * y;
  const double e = fma(-t, 0.5 * y, 0.5);
  inv_r = fma(y, e, y);
  r = fma(t, e, t);
}

// Per-ray quantities shared by every cutoff sphere: the linear coefficient b of
// the ray/sphere quadratic and the observer's heliocentric distance.
// zodipy/line_of_sight.py:70-80: b = 2 (x0 cos(lat) cos(lon) + y0 cos(lat) sin(lon)),
// with lon = atan2(u_y, u_x), lat = asin(u_z); z0 does not enter b.
struct RayGeom {
  double b;
  double r_obs;
};

// Not inlined: the trigonometric routines' literal coefficients then live only
// inside this once-per-ray call instead of in vector registers across the
// kernels' node loops (where they were spilled to scratch); the call costs a
// few instructions per ray.
__device__ __attribute__((noinline)) RayGeom ray_geometry(double ux, double uy, double uz,
                                                          double ox, double oy, double oz) {
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

// Linear interpolation in the Planck table with np.interp semantics, on the
// table's index axis (the forward kernels): u = (T - T_first)/dT is clamped to
// [0, ZODI_N_TEMPS - 1] (np.interp's end values outside the table), j = floor(u)
// and B = B_j + f (B_{j+1} - B_j) with f = u - j = fract(u). The LDS entries from
// the last temperature on repeat it with a zero difference, so j = ZODI_N_TEMPS - 1
// (u at the clamp) reads the end value and j needs no clamp of its own. The
// clamps are against constants, so no operand needs canonicalizing, and the
// temperature itself is never formed. The grid is
// uniform (T_j = T_first + j dT up to rounding), so this is np.interp's
// interpolation to ~1e-14 relative. NANFIX passes a NaN argument through
// (np.interp returns NaN for a NaN temperature); a NaN T comes from a NaN
// position, which also makes the Kelsall densities NaN, so callers whose
// densities are all Kelsall kinds may leave the check out (the masked RRM
// densities are 0 at a NaN position, so theirs need it).
template <bool NANFIX = true>
__device__ __forceinline__ double planck_lookup_u(double u, const double* __restrict__ tab) {
  const double uc = fmin(fmax(u, 0.0), (double)(ZODI_N_TEMPS - 1));
  const int j = (int)uc;
  const double f = __builtin_amdgcn_fract(uc);  // uc - j for uc in [0, 99]
  const double* e = tab + 4 * j;
  const double B = fma(f, e[3], e[2]);
  if constexpr (NANFIX) return u == u ? B : u;
  return B;
}

// Position in a component's frame: offset by X_0 unless the component is
// centred on the Sun (then it is exactly the heliocentric one, and so are |X|
// and 1/|X|). zc is the height above the symmetry plane:
// Z = X'_x sin(Omega) sin(i) - X'_y cos(Omega) sin(i) + X'_z cos(i).
struct CompFrame {
  double x, y, z, r, inv_r, zc, r2;
};

// Frame modes: decided per call from ZODI_CF_CENTERED, or fixed at compile time
// by a kernel whose host side checked the flags. The offset path is exact for a
// centred component too (X - 0 = X, and |X| is formed the same way), only slower.
enum FrameMode { kFrameRuntime = 0, kFrameCentred = 1, kFrameOffset = 2 };

// R2 = |X|^2, the node's own sum of squares (r2 of a centred frame).
template <int MODE = kFrameRuntime>
__device__ __forceinline__ CompFrame comp_frame(ParamPtr P, uint32_t flags, double X0, double X1,
                                                double X2, double R, double inv_R, double R2) {
  CompFrame f;
  const bool centred =
      MODE == kFrameCentred || (MODE == kFrameRuntime && (flags & ZODI_CF_CENTERED));
  if (centred) {
    f.x = X0;
    f.y = X1;
    f.z = X2;
    f.r = R;
    f.inv_r = inv_R;
    f.r2 = R2;
  } else {
    f.x = X0 - P[ZODI_P_X0];
    f.y = X1 - P[ZODI_P_Y0];
    f.z = X2 - P[ZODI_P_Z0];
    f.r2 = f.x * f.x + f.y * f.y + f.z * f.z;
    norm_and_inverse_node(f.r2, f.r, f.inv_r);
  }
  f.zc = f.x * P[ZODI_P_NX] - f.y * P[ZODI_P_NY] + f.z * P[ZODI_P_NZ];
  return f;
}

// ---------------------------------------------------------------- densities
// Cloud, zodipy/number_density.py:47-73:
// n_0 R^-alpha exp(-beta g^gamma), g = zeta^2/(2 mu) if zeta < mu else zeta - mu/2,
// evaluated as n_0 exp(-alpha log R - beta g^gamma).
//
// FAST (the registry kernels, for blocks whose cloud has gamma >= 1/2, which the
// host checks): g floored at the smallest normal double instead of selecting
// log 0 = -inf. For
// g below the floor the exact g^gamma < 1e-150 and the floored one < 1e-154 are
// both absorbed in alpha log R (or, at R = 1 exactly, leave exp(-beta g^gamma)
// = 1): the result is unchanged.
template <bool FAST = false>
__device__ __forceinline__ double cloud_density(ParamPtr P, const CompFrame& f) {
  const double n0 = P[6], alpha = P[7], beta = P[8], gamma = P[9], mu = P[10];
  const double inv_2mu = P[13], half_mu = P[14];
  const double zeta = fabs(f.zc) * f.inv_r;
  const double g = zeta < mu ? (zeta * zeta) * inv_2mu : zeta - half_mu;
  if constexpr (FAST) {
    const double gf = fmax(g, 0x1p-1022);
    const double gg = exp_fast_bounded(gamma * log_pos(gf));
    return n0 * exp_fast_bounded(-alpha * log_pos(f.r) - beta * gg);
  }
  // |log R| <= 745 and g^gamma <= 1 (zeta <= 1): a bounded exponent
  return n0 * exp_fast_bounded(-alpha * log_pos(f.r) - beta * pow_pos(g, gamma));
}

__device__ __forceinline__ double pow20(double x) {
  const double x2 = x * x;
  const double x4 = x2 * x2;
  const double x8 = x4 * x4;
  const double x16 = x8 * x8;
  return x16 * x4;
}

// exp(-x) for the bands' radial cut-off 1 - exp(-(R/delta_r)^20). For x >= 38,
// exp(-x) < 2^-54 and 1 - exp(-x) rounds to 1 exactly, so 0 stands in for it; the
// branch skips the exponential for the waves whose points all lie that far out.
__device__ __forceinline__ double band_cutoff_exp(double x) {
  double ex = 0.0;
  if (x < 38.0) ex = exp_fast_bounded(-x);
  return ex;
}

// Band, zodipy/number_density.py:76-110 (P[6] holds 3 n_0; the 1/v factor as in
// the reference): 3 n_0/R exp(-zd^6) (1 + zd^p / v) (1 - exp(-(R/delta_r)^20)),
// zd = zeta / delta_zeta.
//
// FAST (the registry kernels, for blocks whose bands have p = 4, which the host
// checks): the plane normal pre-divided by delta_zeta (derived slots 16-18, so
// zd = |Z'|/R), and (R/delta_r)^20 = (R^2/delta_r^2)^10 from R^2 (slot 19 =
// 1/delta_r^2). The amplitude stays inside each density, where the reference
// applies it: when the density underflows to a subnormal number its rounding
// then matches the reference's.
//
// CUT: kCutOwn forms the cut-off term 1 - exp(-x); kCutPublish also stores it in
// `cut`; kCutReuse takes it from `cut` (a band whose cut-off the host has checked to
// be bit for bit that of a band published earlier in the same node).
enum BandCut { kCutOwn = 0, kCutPublish = 1, kCutReuse = 2 };

template <bool FAST = false, int CUT = kCutOwn>
__device__ __forceinline__ double band_density(ParamPtr P, uint32_t flags, const CompFrame& f,
                                               double& cut) {
  const double three_n0 = P[6], p = P[8];
  const double inv_dzeta = P[13], inv_v = P[14], inv_dr = P[15];
  double zd, x = 0.0;
  if constexpr (FAST) {
    const double zs = f.x * P[ZODI_P_D0 + 3] - f.y * P[ZODI_P_D0 + 4] + f.z * P[ZODI_P_D0 + 5];
    zd = fabs(zs) * f.inv_r;
    if constexpr (CUT != kCutReuse) {
      const double y = f.r2 * P[ZODI_P_D0 + 6];
      const double y2 = y * y;
      const double y4 = y2 * y2;
      x = (y4 * y4) * y2;
    }
  } else {
    zd = (fabs(f.zc) * f.inv_r) * inv_dzeta;
    if constexpr (CUT != kCutReuse) x = pow20(f.r * inv_dr);
  }
  const double zd2 = zd * zd;
  const double zd4 = zd2 * zd2;
  const double zd6 = zd4 * zd2;
  const double zdp = (FAST || (flags & ZODI_CF_P_IS_4)) ? zd4 : pow_pos(zd, p);
  const double term3 = 1.0 + zdp * inv_v;
  double term4;
  if constexpr (CUT == kCutReuse) {
    term4 = cut;
  } else {
    term4 = 1.0 - band_cutoff_exp(x);
    if constexpr (CUT == kCutPublish) cut = term4;
  }
  return (three_n0 * f.inv_r) * exp_fast(-zd6) * term3 * term4;
}

template <bool FAST = false>
__device__ __forceinline__ double band_density(ParamPtr P, uint32_t flags, const CompFrame& f) {
  double cut = 0.0;
  return band_density<FAST, kCutOwn>(P, flags, f, cut);
}

// Ring, zodipy/number_density.py:113-139: n_0 exp(-(R - R_r)^2/sigma_r^2 - |Z|/sigma_z).
__device__ __forceinline__ double ring_density(ParamPtr P, const CompFrame& f) {
  const double n0 = P[6], Rr = P[7];
  const double d = f.r - Rr;
  return n0 * exp_fast_bounded(-(d * d) * P[13] - fabs(f.zc) * P[14]);
}

// The feature's centre line for one ray: (cos psi, sin psi) with
// psi = theta_E + theta_f, theta_E = atan2(E'_y, E'_x) the Earth's longitude about
// the component centre (E' = X_earth - X_0) and theta_f the feature's offset
// (cos and sin in derived slots 18 and 19).
struct FeatureDir {
  double c, s;
};

__device__ __forceinline__ FeatureDir feature_dir(ParamPtr P, double ex, double ey) {
  const double epx = ex - P[ZODI_P_X0], epy = ey - P[ZODI_P_Y0];
  const double rho2 = epx * epx + epy * epy;
  double ce, se;
  if (rho2 == 0.0) {  // atan2(+-0, +-0) = 0 or +-pi
    ce = __builtin_signbit(epx) ? -1.0 : 1.0;
    se = 0.0;
  } else {
    double rho, inv_rho;
    norm_and_inverse(rho2, rho, inv_rho);
    ce = epx * inv_rho;
    se = epy * inv_rho;
  }
  const double ct = P[ZODI_P_D0 + 5], st = P[ZODI_P_D0 + 6];
  return {ce * ct - se * st, se * ct + ce * st};
}

// dtheta = theta_comp - theta_f wrapped into [-pi, pi) with a floor modulo, as
// zodipy/number_density.py:170-173 forms it, theta_comp = atan2(y', x') - theta_E.
// It is the longitude of (x', y') about the centre line: the atan2 of the point
// rotated by -psi, which lies in (-pi, pi] and agrees up to rounding (and the
// sign at +-pi, where dtheta^2 is the same). On the axis (x' = y' = 0) the
// reference's atan2(+-0, +-0) is 0 (x' = +0) or +-pi (x' = -0), so dtheta is the
// angle of (cos psi, -sin psi) or of (-cos psi, sin psi).
__device__ __forceinline__ double feature_dtheta(const CompFrame& f, const FeatureDir& d) {
  double xr = f.x * d.c + f.y * d.s;
  double yr = f.y * d.c - f.x * d.s;
  if (f.x == 0.0 && f.y == 0.0) {
    const double sg = __builtin_signbit(f.x) ? -1.0 : 1.0;
    xr = sg * d.c;
    yr = -sg * d.s;
  }
  return zodi_trig::atan2_fast(yr, xr);
}

// Feature, zodipy/number_density.py:142-181: n_0 exp(-((R - R_f)^2/sigma_r^2 +
// |Z|/sigma_z + dtheta^2/sigma_theta^2)).
__device__ __forceinline__ double feature_density(ParamPtr P, const CompFrame& f,
                                                  const FeatureDir& fd) {
  const double n0 = P[6], Rf = P[7];
  const double dth = feature_dtheta(f, fd);
  const double d = f.r - Rf;
  double e = (d * d) * P[13];
  e += fabs(f.zc) * P[14];
  e += (dth * dth) * P[15];
  return n0 * exp_fast_bounded(-e);
}

// Fan, zodipy/number_density.py:184-218 (zero outside R <= R_outer).
__device__ __forceinline__ double fan_density(ParamPtr P, const CompFrame& f) {
  const double Q = P[6], Pp = P[7], gamma = P[8], Z0 = P[9], Rout = P[10];
  if (!(f.r <= Rout)) return 0.0;
  const double beta = asin(f.zc * f.inv_r);
  const double za = fabs(f.zc);
  const double eps = za < Z0 ? 2.0 - za * P[13] : 1.0;
  const double fv = pow_pos(cos(beta), Q) * exp_fast(-Pp * sin(pow_pos(fabs(beta), eps)));
  return pow_radius(f.r, -gamma) * fv;
}

// Comet, zodipy/number_density.py:221-256 (zero outside R_inner <= R <= R_outer).
__device__ __forceinline__ double comet_density(ParamPtr P, const CompFrame& f) {
  const double gamma = P[6], Z0 = P[7], Pp = P[8], amp = P[9], Rin = P[10], Rout = P[11];
  if (!(f.r >= Rin && f.r <= Rout)) return 0.0;
  const double beta = asin(f.zc * f.inv_r);
  const double za = fabs(f.zc);
  const double eps = za < Z0 ? 2.0 - za * P[13] : 1.0;
  const double fv = exp_fast(-Pp * sin(pow_pos(fabs(beta), eps)));
  return amp * pow_radius(f.r, -gamma) * fv;
}

// Narrow band, zodipy/number_density.py:267-304 (latitude in degrees).
__device__ __forceinline__ double narrow_band_density(ParamPtr P, const CompFrame& f) {
  const double bnb = P[6], G = P[7], gamma = P[8], A = P[9], Rin = P[10], Rout = P[11];
  if (!(f.r >= Rin && f.r <= Rout)) return 0.0;
  const double babs = fabs(asin(f.zc * f.inv_r) * kRad2Deg);
  const double fv = babs < bnb ? exp_fast(G * (babs - bnb)) : 0.0;
  return A * pow_radius(f.r * P[13], -gamma) * fv;
}

// Broad band, zodipy/number_density.py:307-342 (latitude in degrees).
__device__ __forceinline__ double broad_band_density(ParamPtr P, const CompFrame& f) {
  const double bbb = P[6], gamma = P[8], A = P[9], Rin = P[10], Rout = P[11];
  if (!(f.r >= Rin && f.r <= Rout)) return 0.0;
  const double beta = asin(f.zc * f.inv_r) * kRad2Deg;
  const double a = (beta - bbb) * P[13];
  const double c = (beta + bbb) * P[13];
  const double fv = exp_fast(-0.5 * (a * a)) + exp_fast(-0.5 * (c * c));
  return A * pow_radius(f.r * P[14], -gamma) * fv;
}

__host__ __device__ constexpr bool is_feature(int kind) {
  return kind == ZODI_FEATURE || kind == ZODI_FEATURE_RRM;
}

// Density of component `kind` at heliocentric X (|X| = R, 1/|X| = inv_R, |X|^2 = R2).
__device__ __forceinline__ double density(int kind, ParamPtr P, uint32_t flags, double X0,
                                          double X1, double X2, double R, double inv_R,
                                          double R2, const FeatureDir& fd) {
  if (kind == ZODI_INTERSTELLAR) return P[6];  // zodipy/number_density.py:259-264
  const CompFrame f = comp_frame(P, flags, X0, X1, X2, R, inv_R, R2);
  switch (kind) {
    case ZODI_CLOUD:
      return cloud_density(P, f);
    case ZODI_BAND:
      return band_density(P, flags, f);
    case ZODI_RING:
      return ring_density(P, f);
    case ZODI_FEATURE:
      return feature_density(P, f, fd);
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
      return P[12] * feature_density(P, f, fd);
    default:
      return __builtin_nan("");
  }
}

// Scattering phase function at the node (zodipy/scattering.py:11-50):
// Theta = arccos(-clip(X_los . X_helio / (R_los R_helio), -1, 1)),
// Phi = N (C1 + C2 Theta + exp(C3 Theta)).
__device__ __forceinline__ double phase_function(double R_los, double ux, double uy, double uz,
                                                 double X0, double X1, double X2, double R,
                                                 ParamPtr S) {
  const double lx = R_los * ux, ly = R_los * uy, lz = R_los * uz;
  double ct = (lx * X0 + ly * X1 + lz * X2) / (R_los * R);
  ct = ct != ct ? ct : (ct < -1.0 ? -1.0 : (ct > 1.0 ? 1.0 : ct));
  const double theta = zodi_trig::acos_fast(-ct);
  return S[ZODI_S_PHASE_NORM] * (S[ZODI_S_C1] + S[ZODI_S_C2] * theta + exp_fast(S[ZODI_S_C3] * theta));
}

}  // namespace zodi
