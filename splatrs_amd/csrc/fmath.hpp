// Scalar math shared by the kernels and the host code of libsplat.
//
// The reference computes in binary32 and calls glibc's expf/logf (Rust's f32::exp and
// f32::ln on Linux).  To composite bit-identically on the GPU, `splat_expf` restates the
// table-driven algorithm of glibc's expf (glibc >= 2.27, from ARM's optimized-routines,
// Szabolcs Nagy 2017): x*32/ln2 = k + r with k rounded by the 0x1.8p52 shift trick,
// 2^(k/32) from a 32-entry table of correctly rounded doubles, a cubic in r, all in
// binary64, then one rounding to binary32.  Over every binary32 input in (-87, 88) it
// agrees with glibc 2.35 expf on all but 2 of 2.24e9 values (with or without FMA).
// `splat_logf` rounds a binary64 logarithm to binary32 (correctly rounded except in rare
// hard cases); it differs from glibc logf in about 5e-4 of inputs by one ulp and is used
// only for the per-Gaussian alpha-threshold exponent.
#pragma once

#include <stdint.h>
#include <string.h>

#include <hip/hip_runtime.h>

#ifndef SPLAT_HD
#define SPLAT_HD __host__ __device__ __forceinline__
#endif

namespace splat {

// 2^(i/32) as correctly rounded binary64, with i << 47 subtracted from the bit pattern so
// that adding k << 47 for the integer k = 32 e + i yields 2^(k/32).
#define SPLAT_EXP2_TABLE                                                                  \
  {0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full,                   \
   0x3fef9301d0125b51ull, 0x3fef72b83c7d517bull, 0x3fef54873168b9aaull,                   \
   0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull, 0x3fef06fe0a31b715ull,                   \
   0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,                   \
   0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull,                   \
   0x3feea47eb03a5585ull, 0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull,                   \
   0x3feea11473eb0187ull, 0x3feea589994cce13ull, 0x3feeace5422aa0dbull,                   \
   0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,                   \
   0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull,                   \
   0x3fef3720dcef9069ull, 0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full,                   \
   0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull}

__device__ const uint64_t kExp2Table[32] = SPLAT_EXP2_TABLE;

SPLAT_HD uint32_t f2bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
SPLAT_HD float bits2f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
SPLAT_HD uint64_t d2bits(double d) {
  uint64_t u;
  memcpy(&u, &d, 8);
  return u;
}
SPLAT_HD double bits2d(uint64_t u) {
  double d;
  memcpy(&d, &u, 8);
  return d;
}

// expf with the table passed in (global or LDS copy of SPLAT_EXP2_TABLE).
SPLAT_HD float expf_tab(float x, const uint64_t* tab) {
  const uint32_t abstop = (f2bits(x) >> 20) & 0x7ff;
  if (abstop >= (f2bits(88.0f) >> 20)) {
    if (f2bits(x) == f2bits(-__builtin_inff())) return 0.0f;
    if (abstop >= (f2bits(__builtin_inff()) >> 20)) return x + x;
    if (x > 0x1.62e42ep6f) return __builtin_inff();
    if (x < -0x1.9fe368p6f) return 0.0f;
  }
  const double inv_ln2_n = 0x1.71547652b82fep+0 * 32.0;
  const double shift = 0x1.8p+52;
  const double c0 = 0x1.c6af84b912394p-5 / 32.0 / 32.0 / 32.0;
  const double c1 = 0x1.ebfce50fac4f3p-3 / 32.0 / 32.0;
  const double c2 = 0x1.62e42ff0c52d6p-1 / 32.0;
  const double xd = (double)x;
  const double z = inv_ln2_n * xd;
  double kd = z + shift;
  const uint64_t ki = d2bits(kd);
  kd -= shift;
  const double r = z - kd;
  uint64_t t = tab[ki % 32];
  t += ki << 47;
  const double s = bits2d(t);
  const double zz = c0 * r + c1;
  const double r2 = r * r;
  double y = c2 * r + 1.0;
  y = zz * r2 + y;
  y = y * s;
  return (float)y;
}

// expf_tab for x <= 0 or NaN without branches: glibc's special cases reduce to the
// underflow cut (x < log(0x1p-150) -> +0, which also covers -inf); NaN propagates.  The
// cubic is evaluated with binary64 FMAs: over every binary32 x <= 0 (2.14e9 inputs) the
// rounded result equals that of the separate multiplies and adds of expf_tab (checked
// exhaustively against both; the two agree with glibc's expf except at one input).
SPLAT_HD float expf_nonpos_tab(float x, const uint64_t* tab) {
  const double inv_ln2_n = 0x1.71547652b82fep+0 * 32.0;
  const double shift = 0x1.8p+52;
  const double c0 = 0x1.c6af84b912394p-5 / 32.0 / 32.0 / 32.0;
  const double c1 = 0x1.ebfce50fac4f3p-3 / 32.0 / 32.0;
  const double c2 = 0x1.62e42ff0c52d6p-1 / 32.0;
  const double z = inv_ln2_n * (double)x;
  double kd = z + shift;
  const uint64_t ki = d2bits(kd);
  kd -= shift;
  const double r = z - kd;
  const uint64_t t = tab[ki % 32] + (ki << 47);
  const double s = bits2d(t);
  const double zz = __builtin_fma(c0, r, c1);
  const double r2 = r * r;
  double y = __builtin_fma(c2, r, 1.0);
  y = __builtin_fma(zz, r2, y);
  y = y * s;
  return x < -0x1.9fe368p6f ? 0.0f : (float)y;
}

// expf_nonpos_tab on the compositing walks' arguments: x = -sigma with 0 <= sigma <=
// sigma_max = ln(255 o) on every pixel that uses the result; sigma_max < 94.5 for any
// finite opacity (ln 255 for a sigmoid one), above the underflow cut at -103.97, so the
// cut's compare and select are left out.  Other lanes' results are discarded by their
// predicates.
SPLAT_HD float expf_walk_tab(float x, const uint64_t* tab) {
  const double inv_ln2_n = 0x1.71547652b82fep+0 * 32.0;
  const double shift = 0x1.8p+52;
  const double c0 = 0x1.c6af84b912394p-5 / 32.0 / 32.0 / 32.0;
  const double c1 = 0x1.ebfce50fac4f3p-3 / 32.0 / 32.0;
  const double c2 = 0x1.62e42ff0c52d6p-1 / 32.0;
  const double z = inv_ln2_n * (double)x;
  double kd = z + shift;
  const uint64_t ki = d2bits(kd);
  kd -= shift;
  const double r = z - kd;
  const uint64_t t = tab[ki % 32] + (ki << 47);
  const double s = bits2d(t);
  const double zz = __builtin_fma(c0, r, c1);
  const double r2 = r * r;
  double y = __builtin_fma(c2, r, 1.0);
  y = __builtin_fma(zz, r2, y);
  y = y * s;
  return (float)y;
}

__device__ __forceinline__ float splat_expf(float x) { return expf_tab(x, kExp2Table); }
__device__ __forceinline__ float splat_logf(float x) { return (float)log((double)x); }

// The same restatement evaluated on the host (host code otherwise calls glibc directly).
inline float host_expf(float x) {
  static const uint64_t tab[32] = SPLAT_EXP2_TABLE;
  return expf_tab(x, tab);
}

// Rust `f as u32`: saturating, NaN -> 0.
SPLAT_HD uint32_t as_u32(float v) {
  if (!(v > 0.0f)) return 0u;
  if (v >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)v;
}
// Rust f32::clamp (NaN passes through).
SPLAT_HD float rclamp(float v, float lo, float hi) {
  if (v < lo) v = lo;
  if (v > hi) v = hi;
  return v;
}
// Rust f32::max / f32::min (IEEE maxNum / minNum).
SPLAT_HD float rmax(float a, float b) { return fmaxf(a, b); }
SPLAT_HD float rmin(float a, float b) { return fminf(a, b); }

// Real::powi by square-and-multiply (the compiler runtime's __powisf2).
SPLAT_HD float powi(float a, int b) {
  const bool recip = b < 0;
  unsigned e = recip ? 0u - (unsigned)b : (unsigned)b;
  float r = 1.0f;
  while (true) {
    if (e & 1u) r *= a;
    e >>= 1;
    if (e == 0) break;
    a *= a;
  }
  return recip ? 1.0f / r : r;
}

}  // namespace splat
