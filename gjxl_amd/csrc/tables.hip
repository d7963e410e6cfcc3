// Host generation of the quantization matrices and DCT bases, uploaded once
// per device handle.
#include <cmath>
#include <cstring>

#include "common.h"
#include "tables.h"

namespace gjxl_amd {
namespace {

uint32_t Bits(float v) { uint32_t u; std::memcpy(&u, &v, 4); return u; }
float FromBits(uint32_t u) { float v; std::memcpy(&v, &u, 4); return v; }

// libjxl FastLog2f / FastPow2f (src/util/fast_math.h:14-79), host copies.
float HostFastLog2(float value) {
  const uint32_t bits = Bits(value);
  const int32_t e = static_cast<int32_t>(bits - 0x3f2aaaabu) >> 23;
  const float x = FromBits(bits - (static_cast<uint32_t>(e) << 23)) - 1.0f;
  float num = std::fma(0.74245873327820566f, x, 1.4287160470083755f);
  num = std::fma(num, x, -1.8503833400518310e-06f);
  float den = std::fma(0.17409343003366853f, x, 1.0096718572241148f);
  den = std::fma(den, x, 0.99032814277590719f);
  return num / den + static_cast<float>(e);
}

float HostFastPow2(float value) {
  const float fl = std::floor(value);
  const int32_t e = static_cast<int32_t>(fl) + 127;
  const float ev = FromBits(static_cast<uint32_t>(e) << 23);
  const float f = value - fl;
  float num = f + 1.01749063e+01f;
  num = std::fma(num, f, 4.88687798e+01f);
  num = std::fma(num, f, 9.85506591e+01f);
  num *= ev;
  float den = std::fma(f, 2.10242958e-01f, -2.22328856e-02f);
  den = std::fma(den, f, -1.94414990e+01f);
  den = std::fma(den, f, 9.85506633e+01f);
  return num / den;
}

// libjxl default DCT quantization parameters (the values the reference's
// scripts/generate_quantization_tables.py lists from libjxl e8ff0976):
// per family and channel, the first band weight followed by the relative
// band steps.
struct FamilyParams {
  int bands;
  double p[3][8];
};

const FamilyParams kParams[kFamilyCount] = {
  {6, {{3150.0, 0.0, -0.4, -0.4, -0.4, -2.0},
       {560.0, 0.0, -0.3, -0.3, -0.3, -0.3},
       {512.0, -2.0, -1.0, 0.0, -1.0, -2.0}}},
  {7, {{8996.8725711814115328, -1.3000777393353804, -0.49424529824571225,
        -0.439093774457103443, -0.6350101832695744, -0.90177264050827612,
        -1.6162099239887414},
       {3191.48366296844234752, -0.67424582104194355, -0.80745813428471001,
        -0.44925837484843441, -0.35865440981033403, -0.31322389111877305,
        -0.37615025315725483},
       {1157.50408145487200256, -2.0531423165804414, -1.4,
        -0.50687130033378396, -0.42708730624733904, -1.4856834539296244,
        -4.9209142884401604}}},
  {8, {{15718.40830982518931456, -1.025, -0.98, -0.9012, -0.4,
        -0.48819395464, -0.421064, -0.27},
       {7305.7636810695983104, -0.8041958212306401, -0.7633036457487539,
        -0.55660379990111464, -0.49785304658857626, -0.43699592683512467,
        -0.40180866526242109, -0.27321683125358037},
       {3803.53173721215041536, -3.060733579805728, -2.0413270132490346,
        -2.0235650159727417, -0.5495389509954993, -0.4, -0.4, -0.3}}},
  {7, {{7240.7734393502, -0.7, -0.7, -0.2, -0.2, -0.2, -0.5},
       {1448.15468787004, -0.5, -0.5, -0.5, -0.2, -0.2, -0.2},
       {506.854140754517, -1.4, -0.2, -0.5, -0.5, -1.5, -3.6}}},
  {8, {{13844.97076442300573, -0.97113799999999995, -0.658, -0.42026,
        -0.22712, -0.2206, -0.226, -0.6},
       {4798.964084220744293, -0.61125308982767057, -0.83770786552491361,
        -0.79014862079498627, -0.2692727459704829, -0.38272769465388551,
        -0.22924222653091453, -0.20719098826199578},
       {1807.236946760964614, -1.2, -1.2, -0.7, -0.7, -0.7, -0.4, -0.5}}},
};

constexpr int kLowFrequencyRows[kFamilyCount] = {1, 2, 4, 1, 2};
constexpr int kLowFrequencyCols[kFamilyCount] = {1, 2, 4, 2, 4};

}  // namespace

void BuildQuantTables(float* out) {
  for (int f = 0; f < kFamilyCount; ++f) {
    const FamilyParams& fp = kParams[f];
    const int rows = kFamilyRows[f], cols = kFamilyCols[f];
    for (int c = 0; c < 3; ++c) {
      float bands[8];
      bands[0] = static_cast<float>(fp.p[c][0]);
      for (int i = 1; i < fp.bands; ++i) {
        const float v = static_cast<float>(fp.p[c][i]);
        bands[i] = bands[i - 1] * (v > 0.0f ? 1.0f + v : 1.0f / (1.0f - v));
      }
      const float sqrt_two = 1.41421356237f;
      const float scale = static_cast<float>(fp.bands - 1) / (sqrt_two + 1.0e-6f);
      const float row_scale = scale / static_cast<float>(rows - 1);
      const float col_scale = scale / static_cast<float>(cols - 1);
      float* dq = out + kFamilyOffset[f] * 3 + c * kFamilySize[f];
      float* inv = dq + 3 * kFamilyTotal;
      for (int y = 0; y < rows; ++y) {
        const float dy = static_cast<float>(y) * row_scale;
        const float dy2 = dy * dy;
        for (int x = 0; x < cols; ++x) {
          const float dx = static_cast<float>(x) * col_scale;
          const float position = std::sqrt(std::fma(dx, dx, dy2));
          int index = static_cast<int>(position);
          if (index > fp.bands - 2) index = fp.bands - 2;
          const float fraction = position - static_cast<float>(index);
          const float a = bands[index];
          const float ratio = bands[index + 1] / a;
          const float weight = a * HostFastPow2(HostFastLog2(ratio) * fraction);
          dq[y * cols + x] = 1.0f / weight;
          inv[y * cols + x] =
            (y < kLowFrequencyRows[f] && x < kLowFrequencyCols[f]) ? 0.0f : weight;
        }
      }
    }
  }
}

// src/codec/dct.cpp:22-45: orthonormal DCT bases in binary64 (host cos).
void BuildDctBases(double* out) {
  const double sqrt_two = 1.4142135623730951;
  const double pi = 3.141592653589793;
  for (int forward = 1; forward >= 0; --forward) {
    for (int n : {8, 16, 32}) {
      double* b = out + BasisOffset(n, forward != 0);
      const double scale = forward ? sqrt_two / static_cast<double>(n) : sqrt_two;
      for (int f = 0; f < n; ++f) {
        const double alpha = f == 0 ? 1.0 / sqrt_two : 1.0;
        for (int s = 0; s < n; ++s) {
          volatile double angle = (static_cast<double>(s) + 0.5) *
                               static_cast<double>(f) * pi / static_cast<double>(n);
          b[f * n + s] = scale * alpha * std::cos(angle);
        }
      }
    }
  }
  for (int n : {8, 16, 32}) {
    const double* b = out + BasisOffset(n, true);
    double* t = out + BasisOffsetT(n);
    for (int f = 0; f < n; ++f)
      for (int s = 0; s < n; ++s) t[s * n + f] = b[f * n + s];
  }
}

}  // namespace gjxl_amd
