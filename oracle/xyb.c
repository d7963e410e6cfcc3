/*
 * Oracle: sRGB8 -> linear RGB and linear RGB <-> XYB.  Test infrastructure
 * only; see oracle.h.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

static uint32_t f2u(float v) { uint32_t u; memcpy(&u, &v, 4); return u; }
static float u2f(uint32_t u) { float v; memcpy(&v, &u, 4); return v; }

/* src/c_api/image_conversion.cpp:18-31: the EOTF is evaluated in binary64
 * and rounded once to binary32. */
void oracle_srgb_to_linear_table(float table[256]) {
  for (int i = 0; i < 256; ++i) {
    const double encoded = (double)i / 255.0;
    const double linear = encoded <= 0.04045
      ? encoded / 12.92
      : pow((encoded + 0.055) / 1.055, 2.4);
    table[i] = (float)linear;
  }
}

/* src/c_api/image_conversion.cpp:123-145: planar R, G, B, each w*h. */
void oracle_srgb8_to_linear(const uint8_t* pixels, size_t row_stride_bytes,
                            uint32_t bytes_per_pixel, uint32_t width,
                            uint32_t height, float* linear) {
  float table[256];
  oracle_srgb_to_linear_table(table);
  const size_t plane = (size_t)width * height;
  for (uint32_t y = 0; y < height; ++y) {
    const uint8_t* row = pixels + (size_t)y * row_stride_bytes;
    for (uint32_t x = 0; x < width; ++x) {
      for (int c = 0; c < 3; ++c) {
        linear[c * plane + (size_t)y * width + x] =
          table[row[(size_t)x * bytes_per_pixel + c]];
      }
    }
  }
}

/* src/codec/color_transform.cpp:54-79: reciprocal-cube-root seed from the
 * exponent bits, three Newton steps and one refinement, then r^2*v + add. */
float oracle_fast_cube_root_and_add(float value, float add) {
  const float one_third = 1.0f / 3.0f;
  const float four_thirds = 4.0f / 3.0f;
  const uint32_t bits = f2u(value);
  const uint32_t seed = bits == 0 ? 0u : 0x54800000u - (bits >> 23) * 0x002AAAAAu;
  float r = u2f(seed);
  const float divided = one_third * value;
  for (int i = 0; i < 3; ++i) {
    const float sq = r * r;
    r = fmaf(-divided, sq * sq, four_thirds * r);
  }
  float sq = r * r;
  r = fmaf(one_third, fmaf(-value, sq * sq, r), r);
  sq = r * r;
  return fmaf(sq, value, add);
}

static const float kOpsinBias = 0.0037930732552754493f;
static const float kOpsinMatrix[3][3] = {
  {0.30f, 0.622f, 0.078f},
  {0.23f, 0.692f, 0.078f},
  {0.24342268924547819f, 0.20476744424496821f, 0.55180986650955360f},
};
static const float kInverseOpsinMatrix[3][3] = {
  {11.031566901960783f, -9.866943921568629f, -0.16462299647058826f},
  {-3.254147380392157f, 4.418770392156863f, -0.16462299647058826f},
  {-3.6588512862745097f, 2.7129230470588235f, 1.9459282392156863f},
};

/* src/codec/color_transform.cpp:318-341 (scalar path) and :343-351 / :413-421
 * (right and bottom edge replication of the transformed values). */
void oracle_linear_to_padded_opsin(const float* linear, uint32_t width,
                                   uint32_t height, float intensity_target,
                                   uint32_t padded_width,
                                   uint32_t padded_height, float* opsin) {
  const float scale = intensity_target / 255.0f;
  const float bias_cbrt = oracle_fast_cube_root_and_add(kOpsinBias, 0.0f);
  const size_t in_plane = (size_t)width * height;
  const size_t out_plane = (size_t)padded_width * padded_height;
  for (uint32_t y = 0; y < height; ++y) {
    for (uint32_t x = 0; x < width; ++x) {
      const size_t i = (size_t)y * width + x;
      const float r = linear[i], g = linear[in_plane + i],
                  b = linear[2 * in_plane + i];
      float gamma[3];
      for (int row = 0; row < 3; ++row) {
        float v = fmaf(scale * kOpsinMatrix[row][2], b, kOpsinBias);
        v = fmaf(scale * kOpsinMatrix[row][1], g, v);
        v = fmaf(scale * kOpsinMatrix[row][0], r, v);
        gamma[row] = oracle_fast_cube_root_and_add(v > 0.0f ? v : 0.0f,
                                                   -bias_cbrt);
      }
      const size_t o = (size_t)y * padded_width + x;
      opsin[o] = 0.5f * (gamma[0] - gamma[1]);
      opsin[out_plane + o] = 0.5f * (gamma[0] + gamma[1]);
      opsin[2 * out_plane + o] = gamma[2];
    }
    for (int c = 0; c < 3; ++c) {
      float* row = opsin + c * out_plane + (size_t)y * padded_width;
      for (uint32_t x = width; x < padded_width; ++x) row[x] = row[width - 1];
    }
  }
  for (int c = 0; c < 3; ++c) {
    const float* edge = opsin + c * out_plane + (size_t)(height - 1) * padded_width;
    for (uint32_t y = height; y < padded_height; ++y) {
      memcpy(opsin + c * out_plane + (size_t)y * padded_width, edge,
             sizeof(float) * padded_width);
    }
  }
}

/* src/codec/color_transform.cpp:433-471 (planar, `count` samples/plane). */
void oracle_opsin_to_linear(const float* opsin, size_t count,
                            float intensity_target, float* linear) {
  const float bias_cbrt = cbrtf(kOpsinBias);
  const float scale = 255.0f / intensity_target;
  for (size_t i = 0; i < count; ++i) {
    const float x = opsin[i], y = opsin[count + i], b = opsin[2 * count + i];
    const float gamma[3] = {y + x + bias_cbrt, y - x + bias_cbrt, b + bias_cbrt};
    float mixed[3];
    for (int c = 0; c < 3; ++c) {
      mixed[c] = gamma[c] * gamma[c] * gamma[c] - kOpsinBias;
    }
    for (int row = 0; row < 3; ++row) {
      float v = scale * kInverseOpsinMatrix[row][0] * mixed[0];
      v = fmaf(scale * kInverseOpsinMatrix[row][1], mixed[1], v);
      v = fmaf(scale * kInverseOpsinMatrix[row][2], mixed[2], v);
      linear[row * count + i] = v;
    }
  }
}

/* src/util/fast_math.h:14-38 */
float oracle_fast_log2(float value) {
  if (!isfinite(value) || value <= 0.0f) return NAN;
  const float p0 = -1.8503833400518310e-06f, p1 = 1.4287160470083755f,
              p2 = 0.74245873327820566f;
  const float q0 = 0.99032814277590719f, q1 = 1.0096718572241148f,
              q2 = 0.17409343003366853f;
  const uint32_t bits = f2u(value);
  const int32_t e = (int32_t)(bits - 0x3f2aaaabu) >> 23;
  const float x = u2f(bits - ((uint32_t)e << 23)) - 1.0f;
  float num = fmaf(p2, x, p1);
  num = fmaf(num, x, p0);
  float den = fmaf(q2, x, q1);
  den = fmaf(den, x, q0);
  return num / den + (float)e;
}

/* src/util/fast_math.h:40-70 */
float oracle_fast_pow2(float value) {
  if (!isfinite(value) || value < -126.0f || value > 127.0f) return NAN;
  const float fl = floorf(value);
  const int32_t e = (int32_t)fl + 127;
  const float ev = u2f((uint32_t)e << 23);
  const float f = value - fl;
  float num = f + 1.01749063e+01f;
  num = fmaf(num, f, 4.88687798e+01f);
  num = fmaf(num, f, 9.85506591e+01f);
  num *= ev;
  float den = fmaf(f, 2.10242958e-01f, -2.22328856e-02f);
  den = fmaf(den, f, -1.94414990e+01f);
  den = fmaf(den, f, 9.85506633e+01f);
  return num / den;
}

/* src/util/fast_math.h:72-79 */
float oracle_fast_pow(float base, float exponent) {
  if (!isfinite(exponent)) return NAN;
  return oracle_fast_pow2(oracle_fast_log2(base) * exponent);
}
