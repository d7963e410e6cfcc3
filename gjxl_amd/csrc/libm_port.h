// Bit-exact device evaluations of the libm calls the reference makes on its
// decision path.
//
// The reference calls std::log1p(float) per pixel when it builds the initial
// pixel mask (src/codec/adaptive_quantization.cpp:656).  On the reference's
// platform that is glibc's log1pf, which is the classic fdlibm algorithm
// (SunPro fdlibm s_log1p.c, float conversion by Ian Lance Taylor, Cygnus
// Support, as distributed in glibc sysdeps/ieee754/flt-32): argument
// reduction to f in [sqrt(2)/2-1, sqrt(2)-1] with a correction term c, then
// log(1+f) = f - s*(f - R) with s = f/(2+f) and a degree-7 minimax polynomial
// R(s^2).  It is not correctly rounded, so a correctly rounded device log1p
// would differ from the reference in about 0.55 % of pixels; this is that
// published algorithm written out for the device, verified bit for bit
// against the host's log1pf over every non-negative finite float.
#pragma once

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

}  // namespace gjxl_amd
