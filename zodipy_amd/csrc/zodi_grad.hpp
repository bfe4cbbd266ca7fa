// Analytic partial derivatives of the component densities (gfx950, FP64).
//
// Each function returns the density n at a point together with
//   dR  = dn/dR'   (R' = |X - X_0|, other arguments held fixed),
//   dZ  = dn/dZ    (Z = height above the component's symmetry plane),
//   dth = dn/dtheta_comp (the Earth-trailing feature's longitude, else 0),
// and the partials with respect to the component's own parameters, in the
// declaration order of its parameter dataclass after (x_0, y_0, z_0, i, Omega).
// Angles given in degrees in the parameter dictionary (delta_zeta, theta,
// sigma_theta) are differentiated per degree. geo_accumulate() and geo_finish()
// turn dR, dZ, dth into the partials with respect to x_0, y_0, z_0, i and Omega.
//
// The densities are those of zodipy/number_density.py:47-404; |Z| is
// differentiated as sign(Z) (0 at Z = 0).
#pragma once

#include "zodi_device.hpp"

namespace zodi {

constexpr double kDegToRad = 0.017453292519943295;  // np.pi / 180 (np.radians)

// Number of kind-specific parameters (after the five geometric ones).
template <int KIND>
constexpr int kind_nparams() {
  if constexpr (KIND == ZODI_CLOUD || KIND == ZODI_BAND) return 5;
  if constexpr (KIND == ZODI_RING) return 4;
  if constexpr (KIND == ZODI_FEATURE) return 6;
  if constexpr (KIND == ZODI_FAN) return 5;
  if constexpr (KIND == ZODI_COMET || KIND == ZODI_NARROW_BAND || KIND == ZODI_BROAD_BAND)
    return 6;
  if constexpr (KIND == ZODI_INTERSTELLAR) return 1;
  if constexpr (KIND == ZODI_RING_RRM) return 5;
  if constexpr (KIND == ZODI_FEATURE_RRM) return 7;
  return -1;  // not differentiable here
}

template <int NP>
struct DensGrad {
  double n, dR, dZ, dth;
  double dp[NP];
};

__device__ __forceinline__ double sign_of(double z) { return z > 0.0 ? 1.0 : (z < 0.0 ? -1.0 : 0.0); }

// 1/x for the derivative terms from the hardware estimate (relative error below
// 2^-24 on gfx950, measured) and one Newton step, which squares it: below 1e-14
// relative, far inside what the gradients resolve. NaN for x = 0.
__device__ __forceinline__ double rcp_fast(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
}

// Cloud: n_0, alpha, beta, gamma, mu. `logR` is log of the heliocentric
// distance, which equals log r for a Sun-centred cloud.
__device__ __forceinline__ DensGrad<5> cloud_grad(ParamPtr P, uint32_t flags, const CompFrame& f,
                                                  double logR) {
  const double n0 = P[6], alpha = P[7], beta = P[8], gamma = P[9], mu = P[10];
  const double inv_2mu = P[13], half_mu = P[14];
  const double za = fabs(f.zc);
  const double zeta = za * f.inv_r;
  const bool inner = zeta < mu;
  const double g = inner ? (zeta * zeta) * inv_2mu : zeta - half_mu;
  const double logr = (flags & ZODI_CF_CENTERED) ? logR : log_pos(f.r);
  const double logg = log_nonneg(g);
  const double G = gamma == 0.0 ? 1.0 : exp_fast(gamma * logg);  // g^gamma
  DensGrad<5> o;
  const double E = exp_fast_bounded(-alpha * logr - beta * G);  // n / n_0
  o.n = n0 * E;
  const double dg_dzeta = inner ? zeta * (2.0 * inv_2mu) : 1.0;
  // gamma g^(gamma-1) written as gamma G / g; its limit at g = 0 is 0 for gamma > 1/2.
  const double gG = g > 0.0 ? gamma * G * rcp_fast(g) : 0.0;
  const double nbg = o.n * beta * gG;  // -dn/dg
  const double dn_dzeta = -nbg * dg_dzeta;
  o.dR = -(alpha * o.n + dn_dzeta * zeta) * f.inv_r;
  o.dZ = dn_dzeta * sign_of(f.zc) * f.inv_r;
  o.dth = 0.0;
  o.dp[0] = E;
  o.dp[1] = -o.n * logr;
  o.dp[2] = -o.n * G;
  o.dp[3] = g > 0.0 ? -(o.n * beta) * G * logg : 0.0;
  const double dg_dmu = inner ? -(zeta * zeta) * (2.0 * inv_2mu * inv_2mu) : -0.5;
  o.dp[4] = -nbg * dg_dmu;
  return o;
}

// A band's radial cut-off argument x = (R/delta_r)^20 and exp(-x) (0 where
// 1 - exp(-x) rounds to 1), shared within a node by bands with the same cut-off.
struct CutShare {
  double x, ex;
};

// Band: n_0, delta_zeta [deg], v, p, delta_r. FAST (the registry kernels, for
// blocks whose bands have p = 4): the band density's registry form (the plane
// normal pre-divided by delta_zeta, the cut-off from R^2), p = 4 without a test.
// CUT (BandCut): kCutPublish stores x and exp(-x) in `cs`, kCutReuse takes them
// from it (a band whose cut-off the host has checked to equal the publisher's).
template <bool FAST = false, int CUT = kCutOwn>
__device__ __forceinline__ DensGrad<5> band_grad(ParamPtr P, uint32_t flags, const CompFrame& f,
                                                 CutShare& cs) {
  const double three_n0 = P[6], dzeta = P[7], p = P[8], v = P[9], dr = P[10];
  const double inv_dzeta = P[13], inv_v = P[14], inv_dr = P[15];
  const bool p4 = FAST || (flags & ZODI_CF_P_IS_4);
  double zd, zsign, x = 0.0;
  if constexpr (FAST) {
    const double zs = f.x * P[ZODI_P_D0 + 3] - f.y * P[ZODI_P_D0 + 4] + f.z * P[ZODI_P_D0 + 5];
    zd = fabs(zs) * f.inv_r;
    zsign = sign_of(zs);  // the sign of Z
    if constexpr (CUT != kCutReuse) {
      const double y = f.r2 * P[ZODI_P_D0 + 6];
      const double y2 = y * y;
      const double y4 = y2 * y2;
      x = (y4 * y4) * y2;
    }
  } else {
    zd = (fabs(f.zc) * f.inv_r) * inv_dzeta;
    zsign = sign_of(f.zc);
    if constexpr (CUT != kCutReuse) x = pow20(f.r * inv_dr);
  }
  const double zd2 = zd * zd;
  const double zd4 = zd2 * zd2;
  const double zd5 = zd4 * zd;
  const double zd6 = zd4 * zd2;
  // zd^p = exp(p log zd) needs log 0 = -inf; with p = 4 the log only enters the
  // p partial, which zd = 0 masks, so the special cases can be left out.
  const double logzd = p4 ? log_pos(zd) : log_nonneg(zd);
  const double zdp = p4 ? zd4 : (p == 0.0 ? 1.0 : exp_fast(p * logzd));
  const double t1 = three_n0 * f.inv_r;
  const double t2 = exp_fast(-zd6);
  const double t3 = 1.0 + zdp * inv_v;
  double ex;
  if constexpr (CUT == kCutReuse) {
    x = cs.x;
    ex = cs.ex;
  } else {
    ex = band_cutoff_exp(x);  // 0 where 1 - exp(-x) == 1
    if constexpr (CUT == kCutPublish) cs = {x, ex};
  }
  const double t4 = 1.0 - ex;
  DensGrad<5> o;
  const double t24 = t2 * t4;
  const double t124 = t1 * t24;
  o.n = t124 * t3;
  // d(zd^p)/dzd = p zd^(p-1) (taken as 0 at zd = 0 for p > 1)
  const double dzdp = p4 ? 4.0 * (zd2 * zd) : (zd > 0.0 ? p * zdp * rcp_fast(zd) : 0.0);
  const double dn_dzd = t124 * (-6.0 * zd5 * t3 + dzdp * inv_v);
  // R' d/dR' of the radial cut-off factor: 20 x exp(-x) / (1 - exp(-x)) times n
  const double rdr = (t1 * t2) * t3 * (20.0 * x) * ex;
  o.dR = (rdr - o.n - dn_dzd * zd) * f.inv_r;
  o.dZ = dn_dzd * zsign * (f.inv_r * inv_dzeta);
  o.dth = 0.0;
  o.dp[0] = 3.0 * (t24 * t3) * f.inv_r;
  o.dp[1] = -dn_dzd * zd * (inv_dzeta * kDegToRad);
  const double nzv = t124 * zdp * inv_v;  // dn/dv = -nzv / v, dn/dp = nzv log zd
  o.dp[2] = -nzv * inv_v;
  o.dp[3] = zd > 0.0 ? nzv * logzd : 0.0;
  o.dp[4] = -rdr * inv_dr;
  (void)dzeta;
  (void)v;
  (void)dr;
  return o;
}

// Ring: n_0, R, sigma_r, sigma_z.
__device__ __forceinline__ DensGrad<4> ring_grad(ParamPtr P, const CompFrame& f) {
  const double n0 = P[6], Rr = P[7];
  const double inv_sr2 = P[13], inv_sz = P[14], inv_sr = P[15];
  const double d = f.r - Rr;
  const double za = fabs(f.zc);
  DensGrad<4> o;
  const double E = exp_fast_bounded(-(d * d) * inv_sr2 - za * inv_sz);
  o.n = n0 * E;
  o.dR = -2.0 * o.n * d * inv_sr2;
  o.dZ = -o.n * sign_of(f.zc) * inv_sz;
  o.dth = 0.0;
  o.dp[0] = E;
  o.dp[1] = -o.dR;
  o.dp[2] = 2.0 * o.n * (d * d) * inv_sr2 * inv_sr;
  o.dp[3] = o.n * za * (inv_sz * inv_sz);
  return o;
}

// Feature: n_0, R, sigma_r, sigma_z, theta [deg], sigma_theta [deg].
__device__ __forceinline__ DensGrad<6> feature_grad(ParamPtr P, const CompFrame& f,
                                                    const FeatureDir& fd) {
  const double n0 = P[6], Rf = P[7];
  const double inv_sr2 = P[13], inv_sz = P[14], inv_sth2 = P[15], inv_sr = P[16], inv_sth = P[17];
  const double dth = feature_dtheta(f, fd);
  const double d = f.r - Rf;
  const double za = fabs(f.zc);
  double e = (d * d) * inv_sr2;
  e += za * inv_sz;
  e += (dth * dth) * inv_sth2;
  DensGrad<6> o;
  const double E = exp_fast_bounded(-e);
  o.n = n0 * E;
  o.dR = -2.0 * o.n * d * inv_sr2;
  o.dZ = -o.n * sign_of(f.zc) * inv_sz;
  o.dth = -2.0 * o.n * dth * inv_sth2;
  o.dp[0] = E;
  o.dp[1] = -o.dR;
  o.dp[2] = 2.0 * o.n * (d * d) * inv_sr2 * inv_sr;
  o.dp[3] = o.n * za * (inv_sz * inv_sz);
  o.dp[4] = -o.dth * kDegToRad;
  o.dp[5] = 2.0 * o.n * (dth * dth) * inv_sth2 * inv_sth * kDegToRad;
  return o;
}

// ------------------------------------------------------------ RRM components
// The RRM densities (zodipy/number_density.py:184-342) depend on the ecliptic
// latitude beta = asin(s), s = Z/R. At fixed R, d beta/dZ = 1/(R cos beta); at
// fixed Z, d beta/dR = -s/(R cos beta). Mask edges (R_inner, and R_outer where it
// only masks) move the quadrature's value in jumps, so their derivative is 0.
// Points where the density is 0 (masked, or cos(beta)^Q = 0) get all-zero
// partials.

template <int NP>
__device__ __forceinline__ DensGrad<NP> zero_grad() {
  DensGrad<NP> o;
  o.n = o.dR = o.dZ = o.dth = 0.0;
#pragma unroll
  for (int k = 0; k < NP; ++k) o.dp[k] = 0.0;
  return o;
}

// Latitude exponent shared by the fan and the comet: u = |beta|^eps with
// eps = 2 - |Z|/Z_0 for |Z| < Z_0 and 1 otherwise; returns u and fills
// du/dbeta, du/deps (u log|beta|, 0 at beta = 0) and deps/dZ.
struct LatTerm {
  double beta, cb, u, du_db, du_deps, deps_dz, za;
  bool inner;
};

__device__ __forceinline__ LatTerm lat_term(const CompFrame& f, double inv_Z0, double Z0) {
  LatTerm t;
  const double s = f.zc * f.inv_r;
  t.beta = asin(s);
  t.cb = cos(t.beta);
  t.za = fabs(f.zc);
  t.inner = t.za < Z0;
  const double eps = t.inner ? 2.0 - t.za * inv_Z0 : 1.0;
  const double L = fabs(t.beta);
  t.u = pow_pos(L, eps);
  const bool pos = L > 0.0;
  t.du_db = pos ? eps * t.u * rcp_fast(L) * sign_of(t.beta) : 0.0;
  t.du_deps = pos ? t.u * log_pos(L) : 0.0;
  t.deps_dz = t.inner ? -sign_of(f.zc) * inv_Z0 : 0.0;
  return t;
}

// Fan: cos(beta)^Q exp(-P sin(u)) R^-gamma for R <= R_outer. Partials in the
// declaration order gamma, Z_0, Q, P, R_outer.
__device__ __forceinline__ DensGrad<5> fan_grad(ParamPtr P, const CompFrame& f) {
  const double Q = P[6], Pp = P[7], gamma = P[8], Z0 = P[9], Rout = P[10];
  if (!(f.r <= Rout)) return zero_grad<5>();
  const LatTerm t = lat_term(f, P[13], Z0);
  const double su = sin(t.u), cu = cos(t.u);
  const double fv = pow_pos(t.cb, Q) * exp_fast(-Pp * su);
  DensGrad<5> o;
  o.n = pow_radius(f.r, -gamma) * fv;
  if (o.n == 0.0) return zero_grad<5>();
  const double tanb = (f.zc * f.inv_r) * rcp_fast(t.cb);
  const double dl_db = -Q * tanb - Pp * cu * t.du_db;  // d ln n / d beta
  const double dl_deps = -Pp * cu * t.du_deps;
  const double db_dz = f.inv_r * rcp_fast(t.cb);
  o.dZ = o.n * (dl_db * db_dz + dl_deps * t.deps_dz);
  o.dR = o.n * (-gamma * f.inv_r - dl_db * tanb * f.inv_r);
  o.dth = 0.0;
  o.dp[0] = -o.n * log_pos(f.r);
  o.dp[1] = t.inner ? o.n * dl_deps * t.za * (P[13] * P[13]) : 0.0;
  o.dp[2] = o.n * log_nonneg(t.cb);
  o.dp[3] = -o.n * su;
  o.dp[4] = 0.0;
  return o;
}

// Comet: amp R^-gamma exp(-P sin(u)) for R_inner <= R <= R_outer. Partials in the
// declaration order gamma, Z_0, P, amp, R_inner, R_outer.
__device__ __forceinline__ DensGrad<6> comet_grad(ParamPtr P, const CompFrame& f) {
  const double gamma = P[6], Z0 = P[7], Pp = P[8], amp = P[9], Rin = P[10], Rout = P[11];
  if (!(f.r >= Rin && f.r <= Rout)) return zero_grad<6>();
  const LatTerm t = lat_term(f, P[13], Z0);
  const double su = sin(t.u), cu = cos(t.u);
  const double rg = pow_radius(f.r, -gamma);
  const double fv = exp_fast(-Pp * su);
  DensGrad<6> o;
  o.n = amp * rg * fv;
  const double tanb = (f.zc * f.inv_r) * rcp_fast(t.cb);
  const double dl_db = -Pp * cu * t.du_db;
  const double dl_deps = -Pp * cu * t.du_deps;
  const double db_dz = f.inv_r * rcp_fast(t.cb);
  o.dZ = o.n * (dl_db * db_dz + dl_deps * t.deps_dz);
  o.dR = o.n * (-gamma * f.inv_r - dl_db * tanb * f.inv_r);
  o.dth = 0.0;
  o.dp[0] = -o.n * log_pos(f.r);
  o.dp[1] = t.inner ? o.n * dl_deps * t.za * (P[13] * P[13]) : 0.0;
  o.dp[2] = -o.n * su;
  o.dp[3] = rg * fv;
  o.dp[4] = 0.0;
  o.dp[5] = 0.0;
  return o;
}

// Narrow band: A (R/R_outer)^-gamma exp(G (|beta_deg| - beta_nb)) for |beta_deg| <
// beta_nb and R_inner <= R <= R_outer. Partials in the declaration order gamma, A,
// G, R_inner, R_outer, beta_nb (degrees).
__device__ __forceinline__ DensGrad<6> narrow_band_grad(ParamPtr P, const CompFrame& f) {
  const double bnb = P[6], G = P[7], gamma = P[8], A = P[9], Rin = P[10], Rout = P[11];
  if (!(f.r >= Rin && f.r <= Rout)) return zero_grad<6>();
  const double s = f.zc * f.inv_r;
  const double beta = asin(s);
  const double babs = fabs(beta * kRad2Deg);
  if (!(babs < bnb)) return zero_grad<6>();
  const double fv = exp_fast(G * (babs - bnb));
  const double rho = pow_radius(f.r * P[13], -gamma);
  DensGrad<6> o;
  o.n = A * rho * fv;
  const double inv_cb = rcp_fast(cos(beta));
  const double dl_db = G * sign_of(beta) * kRad2Deg;  // per radian
  o.dZ = o.n * dl_db * f.inv_r * inv_cb;
  o.dR = o.n * (-gamma * f.inv_r - dl_db * s * f.inv_r * inv_cb);
  o.dth = 0.0;
  o.dp[0] = -o.n * log_pos(f.r * P[13]);
  o.dp[1] = rho * fv;
  o.dp[2] = o.n * (babs - bnb);
  o.dp[3] = 0.0;
  o.dp[4] = o.n * gamma * P[13];
  o.dp[5] = -G * o.n;
  return o;
}

// Broad band: A (R/R_outer)^-gamma [exp(-a^2/2) + exp(-c^2/2)], a = (beta_deg -
// beta_bb)/sigma_bb, c = (beta_deg + beta_bb)/sigma_bb, for R_inner <= R <= R_outer.
// Partials in the declaration order gamma, A, R_inner, R_outer, beta_bb, sigma_bb.
__device__ __forceinline__ DensGrad<6> broad_band_grad(ParamPtr P, const CompFrame& f) {
  const double bbb = P[6], gamma = P[8], A = P[9], Rin = P[10], Rout = P[11];
  const double inv_sig = P[13];
  if (!(f.r >= Rin && f.r <= Rout)) return zero_grad<6>();
  const double s = f.zc * f.inv_r;
  const double beta = asin(s);
  const double bdeg = beta * kRad2Deg;
  const double a = (bdeg - bbb) * inv_sig;
  const double c = (bdeg + bbb) * inv_sig;
  const double ea = exp_fast(-0.5 * (a * a));
  const double ec = exp_fast(-0.5 * (c * c));
  const double fv = ea + ec;
  const double rho = pow_radius(f.r * P[14], -gamma);
  const double Ar = A * rho;
  DensGrad<6> o;
  o.n = Ar * fv;
  const double inv_cb = rcp_fast(cos(beta));
  const double dn_db = -Ar * (a * ea + c * ec) * inv_sig * kRad2Deg;  // per radian
  o.dZ = dn_db * f.inv_r * inv_cb;
  o.dR = -gamma * f.inv_r * o.n - dn_db * s * f.inv_r * inv_cb;
  o.dth = 0.0;
  o.dp[0] = -o.n * log_pos(f.r * P[14]);
  o.dp[1] = rho * fv;
  o.dp[2] = 0.0;
  o.dp[3] = o.n * gamma * P[14];
  o.dp[4] = Ar * (a * ea - c * ec) * inv_sig;
  o.dp[5] = Ar * (a * a * ea + c * c * ec) * inv_sig;
  return o;
}

// Value and parameter partials of a component whose kind is known at compile time.
// `logR` = log |X| (heliocentric), reused where the component frame is the Sun's.
// FAST: the registry forms (band_grad); CUT and `cs`: a band's shared cut-off.
template <int KIND, bool FAST = false, int CUT = kCutOwn>
__device__ __forceinline__ DensGrad<kind_nparams<KIND>()> density_grad(ParamPtr P, uint32_t flags,
                                                                        const CompFrame& f,
                                                                        const FeatureDir& fd,
                                                                        double logR, CutShare& cs) {
  if constexpr (KIND == ZODI_CLOUD) {
    return cloud_grad(P, flags, f, logR);
  } else if constexpr (KIND == ZODI_BAND) {
    return band_grad<FAST, CUT>(P, flags, f, cs);
  } else if constexpr (KIND == ZODI_RING) {
    return ring_grad(P, f);
  } else if constexpr (KIND == ZODI_FEATURE) {
    return feature_grad(P, f, fd);
  } else if constexpr (KIND == ZODI_FAN) {
    return fan_grad(P, f);
  } else if constexpr (KIND == ZODI_COMET) {
    return comet_grad(P, f);
  } else if constexpr (KIND == ZODI_NARROW_BAND) {
    return narrow_band_grad(P, f);
  } else if constexpr (KIND == ZODI_BROAD_BAND) {
    return broad_band_grad(P, f);
  } else if constexpr (KIND == ZODI_INTERSTELLAR) {
    DensGrad<1> o;
    o.n = P[6];
    o.dR = o.dZ = o.dth = 0.0;
    o.dp[0] = 1.0;
    return o;
  } else if constexpr (KIND == ZODI_RING_RRM) {
    const DensGrad<4> r = ring_grad(P, f);
    const double A = P[10];
    DensGrad<5> o;
    o.n = A * r.n;
    o.dR = A * r.dR;
    o.dZ = A * r.dZ;
    o.dth = 0.0;
    for (int k = 0; k < 4; ++k) o.dp[k] = A * r.dp[k];
    o.dp[4] = r.n;
    return o;
  } else {
    static_assert(KIND == ZODI_FEATURE_RRM, "no derivatives for this component kind");
    const DensGrad<6> r = feature_grad(P, f, fd);
    const double A = P[12];
    DensGrad<7> o;
    o.n = A * r.n;
    o.dR = A * r.dR;
    o.dZ = A * r.dZ;
    o.dth = A * r.dth;
    for (int k = 0; k < 6; ++k) o.dp[k] = A * r.dp[k];
    o.dp[6] = r.n;
    return o;
  }
}

// Earth-direction terms of the feature's longitude for one ray: E' = X_earth - X_0,
// (E'_y, E'_x) / |E'_xy|^2 (the partials of atan2(E'_y, E'_x) up to sign).
struct EarthTerm {
  double gx, gy;
};

__device__ __forceinline__ EarthTerm earth_term(ParamPtr P, double ex, double ey) {
  const double epx = ex - P[ZODI_P_X0], epy = ey - P[ZODI_P_Y0];
  const double inv = 1.0 / (epx * epx + epy * epy);
  return {epy * inv, epx * inv};
}

// Partials with respect to x_0, y_0, z_0, i [deg], Omega [deg] from dR, dZ, dth.
// X' = X - X_0; Z = x' sin(Omega) sin(i) - y' cos(Omega) sin(i) + z' cos(i); for the
// feature theta_comp = atan2(y', x') - atan2(E'_y, E'_x) with E' = X_earth - X_0.
// Per point,
//   d/dx_0 = -(dR x'/R' + dZ n_x) [+ dth y'/rho^2 - dth E'_y/|E'_xy|^2],
//   d/dy_0 = -(dR y'/R' - dZ n_y) [+ dth E'_x/|E'_xy|^2 - dth x'/rho^2],
//   d/dz_0 = -(dR z'/R' + dZ n_z),
//   d/di   = dZ (x' sin(Omega) cos(i) - y' cos(Omega) cos(i) - z' sin(i)) pi/180,
//   d/dOmega = dZ (x' cos(Omega) sin(i) + y' sin(Omega) sin(i)) pi/180,
// with (n_x, n_y, n_z) the plane normal and rho^2 = x'^2 + y'^2. The component's
// constants and the ray's Earth terms factor out of the node sums, so the node
// loop accumulates the weighted moments (geo_accumulate) and geo_finish combines
// them once per ray.
template <bool FEATURE>
constexpr int geo_moments() {
  return FEATURE ? 10 : 7;
}

// Moments: sum w dR x'/R', y', z' terms (0-2); sum w dZ x', y', z' (3-5); sum w dZ (6);
// feature: sum w dth y'/rho^2, x'/rho^2 (7, 8) and sum w dth (9).
// ACC: any accumulator with add(slot, value).
template <bool FEATURE, class ACC>
__device__ __forceinline__ void geo_accumulate(ACC& A, double w, const CompFrame& f, double dR,
                                               double dZ, double dth) {
  const double wr = w * (dR * f.inv_r);
  const double wz = w * dZ;
  A.add(0, wr * f.x);
  A.add(1, wr * f.y);
  A.add(2, wr * f.z);
  A.add(3, wz * f.x);
  A.add(4, wz * f.y);
  A.add(5, wz * f.z);
  A.add(6, wz);
  if constexpr (FEATURE) {
    const double wt = w * dth;
    const double wtr = wt * rcp_fast(f.x * f.x + f.y * f.y);
    A.add(7, wtr * f.y);
    A.add(8, wtr * f.x);
    A.add(9, wt);
  }
}

template <bool FEATURE>
__device__ __forceinline__ void geo_finish(ParamPtr P, const double* A, const EarthTerm& et,
                                           double out[5]) {
  const double si = P[ZODI_P_SIN_I], ci = P[ZODI_P_COS_I];
  const double so = P[ZODI_P_SIN_OMEGA], co = P[ZODI_P_COS_OMEGA];
  out[0] = -(A[0] + A[6] * P[ZODI_P_NX]);
  out[1] = -(A[1] - A[6] * P[ZODI_P_NY]);
  out[2] = -(A[2] + A[6] * P[ZODI_P_NZ]);
  if constexpr (FEATURE) {
    out[0] += A[7] - A[9] * et.gx;
    out[1] += A[9] * et.gy - A[8];
  }
  out[3] = (A[3] * (so * ci) - A[4] * (co * ci) - A[5] * si) * kDegToRad;
  out[4] = (A[3] * (co * si) + A[4] * (so * si)) * kDegToRad;
}

// Scattering phase function and its partials with respect to C1, C2, C3
// (zodipy/scattering.py:34-59): Phi = N (C1 + C2 Theta + exp(C3 Theta)),
// N = 1/(2 pi D), D = 2 C1 + pi C2 + (exp(pi C3) + 1)/(C3^2 + 1).
struct PhaseGrad {
  double phi, d1, d2, d3;
};

// cos(Theta) = (X_los . X) / (R_los R) is formed as (u . X) / R with the unit
// vector u = X_los / R_los, which agrees to rounding.
__device__ __forceinline__ PhaseGrad phase_grad(double ux, double uy, double uz, double X0,
                                                double X1, double X2, double inv_R, ParamPtr S) {
  double ct = (ux * X0 + uy * X1 + uz * X2) * inv_R;
  ct = ct != ct ? ct : (ct < -1.0 ? -1.0 : (ct > 1.0 ? 1.0 : ct));
  const double theta = zodi_trig::acos_fast(-ct);
  const double C1 = S[ZODI_S_C1], C2 = S[ZODI_S_C2], C3 = S[ZODI_S_C3];
  const double N = S[ZODI_S_PHASE_NORM];
  const double e3 = exp_fast(C3 * theta);
  const double Pv = C1 + C2 * theta + e3;
  const double inv_D = kTwoPi * N;  // 1/D
  PhaseGrad o;
  o.phi = N * Pv;
  o.d1 = N - o.phi * 2.0 * inv_D;
  o.d2 = N * theta - o.phi * kPi * inv_D;
  o.d3 = N * theta * e3 - o.phi * S[ZODI_S_DD_DC3] * inv_D;
  return o;
}

}  // namespace zodi
