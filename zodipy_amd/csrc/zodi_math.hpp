// Table-driven FP64 exp and log for the zodiacal kernels (gfx950).
//
// exp: Tang's table-driven method (ACM TOMS 15(2):144-157, 1989). x = (256 m + j) ln2/256 + r,
//   |r| <= ln2/512; exp(x) = 2^m 2^(j/256) (1 + expm1(r)) with expm1(r) from its Taylor
//   series to r^4 (truncation < 4e-17), evaluated as r + r^2 (1/2 + r (r + 4)/24): the
//   4 is an inline constant of the FMA unit, so no step reads two scalar constants.
//   Exactly rounded 1 + expm1(r) for |x| < ln2/512, which keeps 1 - exp(-x) faithful
//   for small x.
// log: Tang's table-driven method (ACM TOMS 16(4):378-400, 1990). x = 2^e m, m in [1/2, 1),
//   c_j = 1/2 + j/512 with j the top 8 mantissa bits, r = m/c_j - 1 in [0, 1/256);
//   log(x) = e ln2 + log(c_j) + log1p(r) with log1p from its Taylor series to r^6
//   (truncation < 3e-18). r is formed as fma(m, 1/c_j, -1) with the rounded table value
//   1/c_j: one instruction instead of forming c_j and an exact m - c_j, at the cost of
//   about 1.1e-16 absolute in r. The kernels only use log inside exp(y log x), where the
//   absolute error of the log is what matters.
//
// The accuracy of both sequences is pinned on the host (tests/test_math_accuracy.py
// builds this header with g++ and compares with mpmath). On the device the tables are
// staged in LDS by every kernel that calls these functions (stage_math_tables);
// polynomial coefficients sit in constant memory so they are fetched by scalar loads
// and used as SGPR operands.
#pragma once

#include <stdint.h>

#include <cmath>

#include "zodi_math_tables.inc"

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define ZODI_MATH_FN __host__ __device__ __forceinline__
#else
#define ZODI_MATH_FN inline
#endif

namespace zodi {

#ifdef __HIPCC__
__constant__ double kMathTablesConst[256 + 256 + 256] = {ZODI_EXP2_TABLE, ZODI_LOG_INV_TABLE,
                                                         ZODI_LOG_C_TABLE};
// Taylor coefficients of log1p, r^6..r^3.
__constant__ double kLogPoly[4] = {-1.0 / 6.0, 1.0 / 5.0, -1.0 / 4.0, 1.0 / 3.0};

// The three tables in one LDS array: 2^(j/256) at [0, 256), 1/c_j at [256, 512),
// log(c_j) at [512, 768). Aligned to 2 KiB, so every table starts at a multiple of
// 512 bytes and a log reads its two entries with one ds_read2st64 whose address is
// j * 8 (immediate offsets, no add).
__shared__ __attribute__((aligned(2048))) double g_math[768];

// Copy the tables to LDS; every thread of the block must call this before use.
__device__ __forceinline__ void stage_math_tables() {
  for (int i = threadIdx.x; i < 768; i += blockDim.x) g_math[i] = kMathTablesConst[i];
  __syncthreads();
}
#endif

#if defined(__HIP_DEVICE_COMPILE__)
#define ZODI_EXP2_AT(j) g_math[j]
#define ZODI_LOG_INV_AT(j) g_math[256 + (j)]
#define ZODI_LOG_C_AT(j) g_math[512 + (j)]
#define ZODI_LOG_POLY(k) kLogPoly[k]
#else
// Host build (accuracy tests): the same tables as plain arrays.
inline constexpr double kHostExp2[256] = {ZODI_EXP2_TABLE};
inline constexpr double kHostLogInv[256] = {ZODI_LOG_INV_TABLE};
inline constexpr double kHostLogC[256] = {ZODI_LOG_C_TABLE};
inline constexpr double kHostLogPoly[4] = {-1.0 / 6.0, 1.0 / 5.0, -1.0 / 4.0, 1.0 / 3.0};
#define ZODI_EXP2_AT(j) kHostExp2[j]
#define ZODI_LOG_INV_AT(j) kHostLogInv[j]
#define ZODI_LOG_C_AT(j) kHostLogC[j]
#define ZODI_LOG_POLY(k) kHostLogPoly[k]
#endif

ZODI_MATH_FN double math_ldexp(double x, int e) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_ldexp(x, e);
#else
  return std::ldexp(x, e);
#endif
}

// 2^m 2^(j/256) (1 + q), the last step of both exp forms.
ZODI_MATH_FN double exp_tail(double r, int k) {
  const int j = k & 255;
  const int m = k >> 8;
  const double p = fma((r + 4.0) * (1.0 / 24.0), r, 0.5);
  const double q = fma(p * r, r, r);  // expm1(r)
  const double t = ZODI_EXP2_AT(j);
  return math_ldexp(fma(t, q, t), m);
}

constexpr double kExpInvL = 0x1.71547652b82fep+8;  // 256 / ln 2
constexpr double kExpLHi = 0x1.62e42fee00000p-9;   // ln2/256, 32 significant bits
constexpr double kExpLLo = 0x1.a39ef35793c76p-41;  // ln2/256 - kExpLHi

// exp for arguments above -9e15 (or NaN). Below -746 the result rounds to 0 by
// itself: the reduced argument stays small (its rounding error grows like
// |x| 1e-16) and ldexp underflows (below about -5.8e6 the integer conversion of
// x 256/ln 2 saturates, which still gives 0). -inf gives NaN. Callers whose
// arguments are bounded that way (the temperature law, the densities' Gaussian and
// power-law factors) call it directly; exp_fast clamps first.
ZODI_MATH_FN double exp_fast_bounded(double x) {
  const double kd = __builtin_rint(x * kExpInvL);
  double r = fma(-kd, kExpLHi, x);
  r = fma(-kd, kExpLLo, r);
  return exp_tail(r, (int)kd);
}

// exp for arguments below 5.8e6 (above 709.8 the result overflows to +inf): below
// -746 (-inf included) x gets the high word of -746, which leaves it in
// (-746.0005, -746] where the result rounds to 0. With |x 256/ln 2| < 2^31 the
// integer n = rint(x 256/ln 2) is then the low word of x 256/ln 2 + 1.5 2^52 (one
// FMA; the sum's ulp is 1, so it rounds to the nearest integer) instead of a round
// and a conversion. Every exponent in the model is below ~50. NaN passes.
ZODI_MATH_FN double exp_fast(double x) {
  constexpr double kShift = 0x1.8p52;
  const uint64_t xb = __builtin_bit_cast(uint64_t, x);
  const uint32_t hi = x < -746.0 ? 0xC0875000u : (uint32_t)(xb >> 32);  // -746.0
  const double xc = __builtin_bit_cast(double, ((uint64_t)hi << 32) | (uint32_t)xb);
  const double sh = fma(xc, kExpInvL, kShift);
  const double kd = sh - kShift;
  double r = fma(-kd, kExpLHi, xc);
  r = fma(-kd, kExpLLo, r);
  return exp_tail(r, (int)(uint32_t)__builtin_bit_cast(uint64_t, sh));
}

// log for positive finite (or NaN) arguments: no special-case handling.
ZODI_MATH_FN double log_pos(double x) {
  constexpr double kLn2Hi = 0x1.62e42fee00000p-1;  // ln 2, 32 significant bits
  constexpr double kLn2Lo = 0x1.a39ef35793c76p-33;  // ln 2 - kLn2Hi
#if defined(__HIP_DEVICE_COMPILE__)
  const int e = __builtin_amdgcn_frexp_exp(x);
  const double m = __builtin_amdgcn_frexp_mant(x);  // [1/2, 1) for finite x != 0
#else
  int e = 0;
  const double m = std::frexp(x, &e);
#endif
  const uint32_t hi = (uint32_t)(__builtin_bit_cast(uint64_t, m) >> 32);
  const int j = (int)((hi >> 12) & 255u);
  const double r = fma(m, ZODI_LOG_INV_AT(j), -1.0);  // m/c_j - 1
  double p = fma(ZODI_LOG_POLY(0), r, ZODI_LOG_POLY(1));
  p = fma(p, r, ZODI_LOG_POLY(2));
  p = fma(p, r, ZODI_LOG_POLY(3));
  p = fma(p, r, -0.5);
  const double l1p = fma(p * r, r, r);  // log1p(r)
  const double de = (double)e;
  return fma(de, kLn2Hi, ZODI_LOG_C_AT(j) + fma(de, kLn2Lo, l1p));
}

// log for finite non-negative (or NaN) arguments: 0 -> -inf.
ZODI_MATH_FN double log_nonneg(double x) {
  const double res = log_pos(x);
  return x == 0.0 ? -__builtin_inf() : res;
}

}  // namespace zodi
