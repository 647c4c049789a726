// adsr_logf.hpp -- natural logarithm of an f32, bit-identical to glibc's
// logf (the libm the reference's f32::ln calls on x86-64 Linux,
// src/score/scalar.rs:19).
//
// Method: ARM optimized-routines logf (Szabolcs Nagy, 2017; shipped as
// glibc's logf since 2.28).  x = 2^k z with z in [0x1.6p-1, 0x1.6p0) split
// into 16 subintervals; log(x) = k ln2 + log(c) + log1p(z/c - 1), with
// log1p evaluated by a degree-4 polynomial in double precision and a single
// rounding to f32 at the end.  The table holds 1/c and log(c) for each
// subinterval.  Evaluated with fused multiply-adds, the result matches
// glibc logf on every positive normal float (checked exhaustively).
//
// The 16-entry table and the three polynomial coefficients below are those
// of that routine, which bit-identical results require.  Their origin:
//   Copyright (c) 2017-2018, Arm Limited.
//   SPDX-License-Identifier: MIT OR Apache-2.0 WITH LLVM-exception
// (ARM optimized-routines, math/logf.c and math/logf_data.c; the same
// constants are in glibc's sysdeps/ieee754/flt-32/e_logf_data.c.)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace adsr {

struct LogfEntry {
  double invc, logc;
};

#define ADSR_LOGF_TABLE                                 \
  {                                                     \
    {0x1.661ec79f8f3bep+0, -0x1.57bf7808caadep-2},      \
    {0x1.571ed4aaf883dp+0, -0x1.2bef0a7c06ddbp-2},      \
    {0x1.49539f0f010bp+0, -0x1.01eae7f513a67p-2},       \
    {0x1.3c995b0b80385p+0, -0x1.b31d8a68224e9p-3},      \
    {0x1.30d190c8864a5p+0, -0x1.6574f0ac07758p-3},      \
    {0x1.25e227b0b8eap+0, -0x1.1aa2bc79c81p-3},         \
    {0x1.1bb4a4a1a343fp+0, -0x1.a4e76ce8c0e5ep-4},      \
    {0x1.12358f08ae5bap+0, -0x1.1973c5a611cccp-4},      \
    {0x1.0953f419900a7p+0, -0x1.252f438e10c1ep-5},      \
    {0x1p+0, 0x0p+0},                                   \
    {0x1.e608cfd9a47acp-1, 0x1.aa5aa5df25984p-5},       \
    {0x1.ca4b31f026aap-1, 0x1.c5e53aa362eb4p-4},        \
    {0x1.b2036576afce6p-1, 0x1.526e57720db08p-3},       \
    {0x1.9c2d163a1aa2dp-1, 0x1.bc2860d22477p-3},        \
    {0x1.886e6037841edp-1, 0x1.1058bc8a07ee1p-2},       \
    {0x1.767dcf5534862p-1, 0x1.4043057b6ee09p-2},       \
  }

constexpr double kLogfLn2 = 0x1.62e42fefa39efp-1;
constexpr double kLogfA0 = -0x1.00ea348b88334p-2;
constexpr double kLogfA1 = 0x1.5575b0be00b6ap-2;
constexpr double kLogfA2 = -0x1.ffffef20a4123p-2;

// The table is passed in so that device code can read it from LDS or
// constant memory and host code from a static array.  Branch-free: the main
// path runs on a sanitised input (subnormals normalised as in the reference
// algorithm) and zero, negative, infinite and NaN inputs are selected
// afterwards.  glibc's early return for x == 1 is not needed: the main path
// yields +0 for it.  Checked against glibc logf on all 2^32 inputs.
__host__ __device__ __forceinline__ float logf_exact(float x,
                                                     const LogfEntry *tab) {
  const uint32_t ix = __builtin_bit_cast(uint32_t, x);
  const bool special = (ix - 0x00800000u) >= (0x7f800000u - 0x00800000u);
  const bool sub = ix != 0u && ix < 0x00800000u;  // positive subnormal
  const uint32_t ixn = __builtin_bit_cast(uint32_t, x * 0x1p23f) - (23u << 23);
  const uint32_t iu = sub ? ixn : (special ? 0x3f800000u : ix);
  const uint32_t tmp = iu - 0x3f330000u;
  const int i = (int)((tmp >> 19) & 15u);
  const int k = (int32_t)tmp >> 23;
  const uint32_t iz = iu - (tmp & 0xff800000u);
  const double invc = tab[i].invc, logc = tab[i].logc;
  const double z = (double)__builtin_bit_cast(float, iz);
  const double r = __builtin_fma(z, invc, -1.0);
  const double y0 = __builtin_fma((double)k, kLogfLn2, logc);
  const double r2 = r * r;
  double y = __builtin_fma(kLogfA1, r, kLogfA2);
  y = __builtin_fma(kLogfA0, r2, y);
  y = __builtin_fma(y, r2, y0 + r);
  float res = (float)y;
  if (special && !sub) {
    res = (ix << 1) == 0u ? -__builtin_inff()
          : ix == 0x7f800000u ? x
                             : __builtin_nanf("");
  }
  return res;
}

// The main path of logf_exact alone: valid for positive normal x
// (0x00800000 <= bits(x) < 0x7f800000), where it gives the same bits.
__host__ __device__ __forceinline__ float logf_normal(float x,
                                                      const LogfEntry *tab) {
  const uint32_t ix = __builtin_bit_cast(uint32_t, x);
  const uint32_t tmp = ix - 0x3f330000u;
  const int i = (int)((tmp >> 19) & 15u);
  const int k = (int32_t)tmp >> 23;
  const uint32_t iz = ix - (tmp & 0xff800000u);
  const double invc = tab[i].invc, logc = tab[i].logc;
  const double z = (double)__builtin_bit_cast(float, iz);
  const double r = __builtin_fma(z, invc, -1.0);
  const double y0 = __builtin_fma((double)k, kLogfLn2, logc);
  const double r2 = r * r;
  double y = __builtin_fma(kLogfA1, r, kLogfA2);
  y = __builtin_fma(kLogfA0, r2, y);
  y = __builtin_fma(y, r2, y0 + r);
  return (float)y;
}

__host__ __device__ __forceinline__ bool logf_is_normal(float x) {
  const uint32_t ix = __builtin_bit_cast(uint32_t, x);
  return (ix - 0x00800000u) < (0x7f800000u - 0x00800000u);
}

}  // namespace adsr
