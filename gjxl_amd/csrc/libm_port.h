// Bit-exact device evaluations of the libm calls the reference makes on its
// decision path.  Each is written out for the device from a published
// algorithm and checked bit for bit against the host libm (glibc 2.35) over
// every non-negative float it can receive.
//
// FdlibmLog1pf.  The reference calls std::log1p(float) per pixel when it
// builds the initial pixel mask (src/codec/adaptive_quantization.cpp:656).
// On the reference's platform that is glibc's log1pf, the classic fdlibm
// algorithm (SunPro fdlibm s_log1p.c, float conversion by Ian Lance Taylor,
// Cygnus Support, as distributed in glibc sysdeps/ieee754/flt-32): argument
// reduction to f in [sqrt(2)/2-1, sqrt(2)-1] with a correction term c, then
// log(1+f) = f - s*(f - R) with s = f/(2+f) and a degree-7 minimax polynomial
// R(s^2).  It is not correctly rounded, so a correctly rounded device log1p
// would differ from the reference in about 0.55 % of pixels.  fdlibm's
// notice, which its terms require to be kept with the code:
//
//   Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
//
//   Developed at SunPro, a Sun Microsystems, Inc. business.
//   Permission to use, copy, modify, and distribute this
//   software is freely granted, provided that this notice
//   is preserved.
//
// GlibcPowf.  The reference calls std::pow(float, float) for the RD-proxy
// cost root loss^(1/8) (src/codec/ac_strategy.cpp:979) and the Butteraugli
// 16-norm root (src/codec/butteraugli.cpp:188-190).  glibc's powf
// (sysdeps/ieee754/flt-32/e_powf.c) is Szabolcs Nagy's method from Arm's
// optimized-routines (math/powf.c, powf_log2_data.c, exp2f_data.c):
// log2(x) in binary64 from a 16-interval table of (1/c, log2 c) and a
// degree-5 polynomial in r = z/c - 1, y*log2(x) split as k/32 + r with a
// 32-entry table of 2^(i/32) and a cubic for 2^r, one final rounding to
// binary32.  It is not correctly rounded either: a correctly rounded root
// differs from it in about 0.06 % of arguments.  Arm's notice for that code:
//
//   Copyright (c) 2017-2018, Arm Limited.
//   SPDX-License-Identifier: MIT
//
//   Permission is hereby granted, free of charge, to any person obtaining a
//   copy of this software and associated documentation files (the
//   "Software"), to deal in the Software without restriction, including
//   without limitation the rights to use, copy, modify, merge, publish,
//   distribute, sublicense, and/or sell copies of the Software, and to
//   permit persons to whom the Software is furnished to do so, subject to
//   the following conditions: The above copyright notice and this
//   permission notice shall be included in all copies or substantial
//   portions of the Software.  THE SOFTWARE IS PROVIDED "AS IS", WITHOUT
//   WARRANTY OF ANY KIND, EXPRESS OR IMPLIED.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#define GJXL_AMD_HD __host__ __device__
#else
#define GJXL_AMD_HD
#endif

namespace gjxl_amd {

GJXL_AMD_HD inline int32_t LibmFloatWord(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __float_as_int(v);
#else
  int32_t i;
  std::memcpy(&i, &v, 4);
  return i;
#endif
}

GJXL_AMD_HD inline float LibmWordFloat(int32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __int_as_float(i);
#else
  float v;
  std::memcpy(&v, &i, 4);
  return v;
#endif
}

// fdlibm log1pf, for x >= 0 finite or any other float (NaN/inf/-1 handled
// as fdlibm does).  Every operation is a separate binary32 rounding.
GJXL_AMD_HD inline float FdlibmLog1pf(float x) {
  const float ln2_hi = 6.9313812256e-01f;  // 0x3f317180
  const float ln2_lo = 9.0580006145e-06f;  // 0x3717f7d1
  const float Lp1 = 6.6666668653e-01f;     // 0x3F2AAAAB
  const float Lp2 = 4.0000000596e-01f;     // 0x3ECCCCCD
  const float Lp3 = 2.8571429849e-01f;     // 0x3E924925
  const float Lp4 = 2.2222198546e-01f;     // 0x3E638E29
  const float Lp5 = 1.8183572590e-01f;     // 0x3E3A3325
  const float Lp6 = 1.5313838422e-01f;     // 0x3E1CD04F
  const float Lp7 = 1.4798198640e-01f;     // 0x3E178897

  const int32_t hx = LibmFloatWord(x);
  const int32_t ax = hx & 0x7fffffff;
  float f = 0.0f, c = 0.0f;
  int32_t k = 1, hu = 0;
  if (hx < 0x3ed413d7) {            // x < 0.41422
    if (ax >= 0x3f800000) {         // x <= -1
      if (x == -1.0f) return -__builtin_inff();
      return __builtin_nanf("");
    }
    if (ax < 0x31000000) {          // |x| < 2^-29
      if (ax < 0x24800000) return x;  // |x| < 2^-54
      return x - x * x * 0.5f;
    }
    if (hx > 0 || hx <= static_cast<int32_t>(0xbe95f61fu)) {
      k = 0;                        // -0.2929 < x < 0.41422
      f = x;
      hu = 1;
    }
  } else if (hx >= 0x7f800000) {
    return x + x;
  }
  if (k != 0) {
    float u;
    if (hx < 0x5a000000) {
      u = 1.0f + x;
      hu = LibmFloatWord(u);
      k = (hu >> 23) - 127;
      c = (k > 0) ? 1.0f - (u - x) : x - (u - 1.0f);  // correction term
      c /= u;
    } else {
      u = x;
      hu = LibmFloatWord(u);
      k = (hu >> 23) - 127;
      c = 0.0f;
    }
    hu &= 0x007fffff;
    if (hu < 0x3504f7) {
      u = LibmWordFloat(hu | 0x3f800000);  // normalise u
    } else {
      k += 1;
      u = LibmWordFloat(hu | 0x3f000000);  // normalise u/2
      hu = (0x00800000 - hu) >> 2;
    }
    f = u - 1.0f;
  }
  const float hfsq = 0.5f * f * f;
  const float fk = static_cast<float>(k);
  if (hu == 0) {                    // |f| < 2^-20
    if (f == 0.0f) {
      if (k == 0) return 0.0f;
      c += fk * ln2_lo;
      return fk * ln2_hi + c;
    }
    const float R = hfsq * (1.0f - 0.66666666666666666f * f);
    if (k == 0) return f - R;
    return fk * ln2_hi - ((R - (fk * ln2_lo + c)) - f);
  }
  const float s = f / (2.0f + f);
  const float z = s * s;
  const float R =
    z * (Lp1 + z * (Lp2 + z * (Lp3 + z * (Lp4 + z * (Lp5 + z * (Lp6 + z * Lp7))))));
  if (k == 0) return f - (hfsq - s * (hfsq + R));
  return fk * ln2_hi - ((hfsq - (s * (hfsq + R) + (fk * ln2_lo + c))) - f);
}

GJXL_AMD_HD inline uint64_t LibmDoubleWord(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return static_cast<uint64_t>(__double_as_longlong(v));
#else
  uint64_t u;
  std::memcpy(&u, &v, 8);
  return u;
#endif
}

GJXL_AMD_HD inline double LibmWordDouble(uint64_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __longlong_as_double(static_cast<long long>(u));
#else
  double v;
  std::memcpy(&v, &u, 8);
  return v;
#endif
}

GJXL_AMD_HD inline double LibmFma(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __fma_rn(a, b, c);
#else
  return std::fma(a, b, c);
#endif
}

// glibc powf for y > 0 finite and not an integer (the reference's 1/8 and
// 1/16).  x <= 0 gives +0 for x = +-0 and NaN below zero; +inf gives +inf;
// NaN gives NaN.  For these y, |y log2 x| < 20, so the overflow and
// underflow branches of the general routine are never taken.  The binary64
// multiply-adds are fused, as in the FMA build that glibc selects on x86-64
// processors with FMA (the unfused build gives the same results for these
// exponents: checked over every non-negative float).
GJXL_AMD_HD inline float GlibcPowf(float x, float y) {
  // (1/c, log2 c) for the 16 subintervals of [0x3f330000, 2 * 0x3f330000)
  constexpr double kInvC[16] = {
    0x1.661ec79f8f3bep+0, 0x1.571ed4aaf883dp+0, 0x1.49539f0f010bp+0,
    0x1.3c995b0b80385p+0, 0x1.30d190c8864a5p+0, 0x1.25e227b0b8eap+0,
    0x1.1bb4a4a1a343fp+0, 0x1.12358f08ae5bap+0, 0x1.0953f419900a7p+0,
    0x1p+0,               0x1.e608cfd9a47acp-1, 0x1.ca4b31f026aap-1,
    0x1.b2036576afce6p-1, 0x1.9c2d163a1aa2dp-1, 0x1.886e6037841edp-1,
    0x1.767dcf5534862p-1};
  constexpr double kLog2C[16] = {
    -0x1.efec65b963019p-2, -0x1.b0b6832d4fca4p-2, -0x1.7418b0a1fb77bp-2,
    -0x1.39de91a6dcf7bp-2, -0x1.01d9bf3f2b631p-2, -0x1.97c1d1b3b7afp-3,
    -0x1.2f9e393af3c9fp-3, -0x1.960cbbf788d5cp-4, -0x1.a6f9db6475fcep-5,
    0x0p+0,                0x1.338ca9f24f53dp-4,  0x1.476a9543891bap-3,
    0x1.e840b4ac4e4d2p-3,  0x1.40645f0c6651cp-2,  0x1.88e9c2c1b9ff8p-2,
    0x1.ce0a44eb17bccp-2};
  // log2(1 + r) ~ A0 r^5 + A1 r^4 + A2 r^3 + A3 r^2 + A4 r
  constexpr double kA[5] = {0x1.27616c9496e0bp-2, -0x1.71969a075c67ap-2,
                            0x1.ec70a6ca7baddp-2, -0x1.7154748bef6c8p-1,
                            0x1.71547652ab82bp0};
  // bits of RN(2^(i/32)) minus i << 47 (so that adding k << 47 scales by 2^k)
  constexpr uint64_t kExp2[32] = {
    0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full,
    0x3fef9301d0125b51ull, 0x3fef72b83c7d517bull, 0x3fef54873168b9aaull,
    0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull, 0x3fef06fe0a31b715ull,
    0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
    0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull,
    0x3feea47eb03a5585ull, 0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull,
    0x3feea11473eb0187ull, 0x3feea589994cce13ull, 0x3feeace5422aa0dbull,
    0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
    0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull,
    0x3fef3720dcef9069ull, 0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full,
    0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull};
  // 2^r ~ C0 r^3 + C1 r^2 + C2 r + 1 for |r| <= 1/64
  constexpr double kC[3] = {0x1.c6af84b912394p-5, 0x1.ebfce50fac4f3p-3,
                            0x1.62e42ff0c52d6p-1};
  uint32_t ix = static_cast<uint32_t>(LibmFloatWord(x));
  if (ix - 0x00800000u >= 0x7f800000u - 0x00800000u) {
    // zero, subnormal, negative, infinite or NaN
    if ((ix & 0x7fffffffu) == 0) return 0.0f;
    if (ix > 0x7f800000u && ix < 0x80000000u) return x + y;  // NaN
    if (ix == 0x7f800000u) return x;
    if (ix & 0x80000000u) return __builtin_nanf("");
    // subnormal: normalise so that the exponent field goes negative
    ix = static_cast<uint32_t>(LibmFloatWord(x * 0x1p23f)) & 0x7fffffffu;
    ix -= 23u << 23;
  }
  // x = 2^k z with z in [0x3f330000, 2 * 0x3f330000) and exact
  const uint32_t tmp = ix - 0x3f330000u;
  const int i = static_cast<int>((tmp >> 19) & 15u);
  const uint32_t top = tmp & 0xff800000u;
  const uint32_t iz = ix - top;
  const int k = static_cast<int32_t>(top) >> 23;
  const double z = static_cast<double>(LibmWordFloat(static_cast<int32_t>(iz)));
  const double r = LibmFma(z, kInvC[i], -1.0);
  const double y0 = kLog2C[i] + static_cast<double>(k);
  const double r2 = r * r;
  double l = LibmFma(kA[0], r, kA[1]);
  const double p = LibmFma(kA[2], r, kA[3]);
  const double r4 = r2 * r2;
  double q = LibmFma(kA[4], r, y0);
  q = LibmFma(p, r2, q);
  l = LibmFma(l, r4, q);
  const double xd = static_cast<double>(y) * l;
  // xd = j/32 + u with |u| <= 1/64: adding 1.5 * 2^47 rounds xd to 1/32
  constexpr double kShift = 0x1.8p+47;
  double kd = xd + kShift;
  const uint64_t ki = LibmDoubleWord(kd);
  kd -= kShift;
  const double u = xd - kd;
  const uint64_t t = kExp2[ki & 31u] + (ki << 47);
  const double s = LibmWordDouble(t);
  const double zz = LibmFma(kC[0], u, kC[1]);
  const double u2 = u * u;
  double e = LibmFma(kC[2], u, 1.0);
  e = LibmFma(zz, u2, e);
  return static_cast<float>(e * s);
}

}  // namespace gjxl_amd
