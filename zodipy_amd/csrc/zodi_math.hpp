// Table-driven FP64 exp and log for the zodiacal kernels (gfx950).
//
// exp: Tang's table-driven method (ACM TOMS 15(2):144-157, 1989). x = (256 m + j) ln2/256 + r,
//   |r| <= ln2/512; exp(x) = 2^m 2^(j/256) (1 + expm1(r)) with expm1(r) from its Taylor
//   series to r^4 (truncation < 4e-17), evaluated as r + r^2 (1/2 + r (r + 4)/24): the
//   4 is an inline constant of the FMA unit, so no step reads two scalar constants.
//   Within 1 ulp of the exact value (checked against mpmath on 6e5 arguments); exactly
//   rounded 1 + expm1(r) for |x| < ln2/512, which keeps 1 - exp(-x) faithful for small x.
// log: Tang's table-driven method (ACM TOMS 16(4):378-400, 1990). x = 2^e m, m in [1/2, 1),
//   c_j = 1/2 + j/512 with j the top 8 mantissa bits, r = (m - c_j)/c_j in [0, 1/256);
//   log(x) = e ln2 + log(c_j) + log1p(r) with log1p from its Taylor series to r^6
//   (truncation < 3e-18). Absolute error ~1e-16 + 0.5 ulp: the kernels only use log
//   inside exp(y log x), where the absolute error is what matters.
//
// Tables (correctly rounded, gen_math_tables.py) are staged in LDS by every kernel
// that calls these functions (stage_math_tables); polynomial coefficients sit in
// constant memory so they are fetched by scalar loads and used as SGPR operands.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zodi_math_tables.inc"

namespace zodi {

__constant__ double kMathTablesConst[256 + 256 + 256] = {ZODI_EXP2_TABLE, ZODI_LOG_INV_TABLE,
                                                         ZODI_LOG_C_TABLE};
// Taylor coefficients of log1p, r^6..r^3.
__constant__ double kLogPoly[4] = {-1.0 / 6.0, 1.0 / 5.0, -1.0 / 4.0, 1.0 / 3.0};

__shared__ double g_exp2[256];
__shared__ double g_log_inv[256];
__shared__ double g_log_c[256];

// Copy the tables to LDS; every thread of the block must call this before use.
__device__ __forceinline__ void stage_math_tables() {
  for (int i = threadIdx.x; i < 256; i += blockDim.x) {
    g_exp2[i] = kMathTablesConst[i];
    g_log_inv[i] = kMathTablesConst[256 + i];
    g_log_c[i] = kMathTablesConst[512 + i];
  }
  __syncthreads();
}

// exp for arguments above -9e15 (or NaN). Below -746 the result rounds to 0 by
// itself: the reduced argument stays small (its rounding error grows like
// |x| 1e-16) and ldexp underflows (below about -5.8e6 the integer conversion of
// x 256/ln 2 saturates, which still gives 0). -inf gives NaN. Callers whose arguments are bounded that way (the
// temperature law, the densities' Gaussian and power-law factors) call it
// directly; exp_fast clamps first.
__device__ __forceinline__ double exp_fast_bounded(double x) {
  constexpr double kInvL = 0x1.71547652b82fep+8;        // 256 / ln 2
  constexpr double kLHi = 0x1.62e42fee00000p-9;         // ln2/256, 32 significant bits
  constexpr double kLLo = 0x1.a39ef35793c76p-41;        // ln2/256 - kLHi
  const double kd = __builtin_rint(x * kInvL);
  double r = fma(-kd, kLHi, x);
  r = fma(-kd, kLLo, r);
  const int k = (int)kd;
  const int j = k & 255;
  const int m = k >> 8;
  const double p = fma((r + 4.0) * (1.0 / 24.0), r, 0.5);
  const double q = fma(p * r, r, r);  // expm1(r)
  const double t = g_exp2[j];
  return __builtin_amdgcn_ldexp(fma(t, q, t), m);
}

__device__ __forceinline__ double exp_fast(double x) {
  // Every exponent in the model is below ~50 (the positive ones are small
  // temperature and power-law exponents). Arguments below -746 (-inf included)
  // get the high word of -746, which leaves them in (-746.0005, -746]: the result
  // rounds to 0. One select of a 32-bit word instead of two. NaN passes.
  const uint64_t xb = __builtin_bit_cast(uint64_t, x);
  const uint32_t hi = x < -746.0 ? 0xC0875000u : (uint32_t)(xb >> 32);  // -746.0
  return exp_fast_bounded(__builtin_bit_cast(double, ((uint64_t)hi << 32) | (uint32_t)xb));
}

// log for positive finite (or NaN) arguments: no special-case handling.
__device__ __forceinline__ double log_pos(double x) {
  constexpr double kLn2Hi = 0x1.62e42fee00000p-1;  // ln 2, 32 significant bits
  constexpr double kLn2Lo = 0x1.a39ef35793c76p-33;  // ln 2 - kLn2Hi
  const int e = __builtin_amdgcn_frexp_exp(x);
  const double m = __builtin_amdgcn_frexp_mant(x);  // [1/2, 1) for finite x != 0
  const uint32_t hi = (uint32_t)(__builtin_bit_cast(uint64_t, m) >> 32);
  const int j = (int)((hi >> 12) & 255u);
  const double c = fma((double)j, 1.0 / 512.0, 0.5);
  const double r = (m - c) * g_log_inv[j];  // m - c is exact
  double p = fma(kLogPoly[0], r, kLogPoly[1]);
  p = fma(p, r, kLogPoly[2]);
  p = fma(p, r, kLogPoly[3]);
  p = fma(p, r, -0.5);
  const double l1p = fma(p * r, r, r);  // log1p(r)
  const double de = (double)e;
  return fma(de, kLn2Hi, g_log_c[j] + fma(de, kLn2Lo, l1p));
}

// log for finite non-negative (or NaN) arguments: 0 -> -inf.
__device__ __forceinline__ double log_nonneg(double x) {
  const double res = log_pos(x);
  return x == 0.0 ? -__builtin_inf() : res;
}

}  // namespace zodi
