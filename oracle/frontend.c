/*
 * Oracle: Gaborish inverse (Symmetric5 convolution) and the initial quant
 * field with its masks.  Test infrastructure only; see oracle.h.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* std::max / std::min argument order and NaN behaviour. */
static inline float std_max(float a, float b) { return a < b ? b : a; }
static inline float std_min(float a, float b) { return b < a ? b : a; }

/* src/core/image_ops.h:20-47: mirror into [0, size), repeating edges. */
size_t oracle_mirror(long long c, size_t size) {
  if (c >= 0 && (size_t)c < size) return (size_t)c;
  const unsigned long long period = 2ull * size;
  unsigned long long phase;
  if (c >= 0) {
    phase = (unsigned long long)c % period;
  } else {
    const unsigned long long mag = (unsigned long long)(-(c + 1)) + 1;
    const unsigned long long rem = mag % period;
    phase = rem == 0 ? 0 : period - rem;
  }
  return (size_t)(phase < size ? phase : period - 1 - phase);
}

/* src/codec/convolution.cpp:16-35 */
static float weighted_row(const float* plane, size_t w, size_t h, long long x,
                          long long y, float center, float near, float far) {
  const float* row = plane + oracle_mirror(y, h) * w;
#define S(dx) row[oracle_mirror(x + (dx), w)]
  const float f = far * (S(-2) + S(2));
  const float n = near * (S(-1) + S(1));
  const float c = center * S(0);
#undef S
  return f + (n + c);
}

/* src/codec/convolution.cpp:38-119.  weights: d0, d1, d2, d4, d8, d5. */
void oracle_convolve_symmetric5(const float* in, size_t w, size_t h,
                                const float weights[6], float* out) {
  const float d0 = weights[0], d1 = weights[1], d2 = weights[2],
              d4 = weights[3], d8 = weights[4], d5 = weights[5];
  for (size_t y = 0; y < h; ++y) {
    for (size_t x = 0; x < w; ++x) {
      const long long sx = (long long)x, sy = (long long)y;
      float sum0 = weighted_row(in, w, h, sx, sy, d0, d1, d2);
      sum0 += weighted_row(in, w, h, sx, sy - 2, d2, d5, d8);
      float sum1 = weighted_row(in, w, h, sx, sy + 2, d2, d5, d8);
      sum0 += weighted_row(in, w, h, sx, sy - 1, d1, d4, d5);
      sum1 += weighted_row(in, w, h, sx, sy + 1, d1, d4, d5);
      out[y * w + x] = sum0 + sum1;
    }
  }
}

/* src/codec/gaborish.cpp:22-47 */
void oracle_gaborish_inverse_weights(float multiplier, float weights[6]) {
  const float k[5] = {
    -0.09495815671340026f, -0.041031725066768575f, 0.013710004822696948f,
    0.006510206083837737f, -0.0014789063378272242f,
  };
  const float partial = k[0] + k[1] + k[2] + k[4];
  double sum = 1.0 + (double)multiplier * 4.0 * ((double)partial + 2.0 * (double)k[3]);
  if (sum < 1.0e-5) sum = 1.0e-5;
  const float normalize = (float)(1.0 / sum);
  const float nm = multiplier * normalize;
  weights[0] = normalize;
  weights[1] = nm * k[0];
  weights[2] = nm * k[2];
  weights[3] = nm * k[1];
  weights[4] = nm * k[4];
  weights[5] = nm * k[3];
}

/* src/codec/gaborish.cpp:49-99 (planar 3-channel image, stride == w). */
void oracle_gaborish_inverse(const float* in, size_t w, size_t h,
                             const float multipliers[3], float* out) {
  for (int c = 0; c < 3; ++c) {
    float weights[6];
    oracle_gaborish_inverse_weights(multipliers[c], weights);
    oracle_convolve_symmetric5(in + c * w * h, w, h, weights, out + c * w * h);
  }
}

/* ---- initial quant field: src/codec/adaptive_quantization.cpp ---------- */

static const float kInverseLog2E = 0.6931471805599453f;

/* :117-143 GammaDerivativeRatio<Invert> */
static float gamma_derivative_ratio(float value, int invert) {
  const float eps = 1.0e-2f;
  const float sg_mul = 226.77216153508914f;
  const float sg_mul2 = 1.0f / 73.377132366608819f;
  const float sg_return_mul = sg_mul2 * 18.6580932135f * kInverseLog2E;
  const float sg_offset = 7.7825991679894591f;
  value = std_max(value, 0.0f);
  const float sq = value * value;
  const float num = fmaf(sg_return_mul * 3.0f * sg_mul, sq, eps);
  const float den = fmaf(kInverseLog2E * sg_mul * value, sq,
                         sg_offset * kInverseLog2E + eps);
  return invert ? num / den : den / num;
}

/* :145-162 */
static float compute_mask(float value) {
  const float base = -0.7647f, mul4 = 9.4708735624378946f,
              mul2 = 17.35036561631863f, offset2 = 302.59587815579727f,
              mul3 = 6.7943250517376494f, offset3 = 3.7179635626140772f,
              mul0 = 0.80061762862741759f;
  const float offset4 = 0.25f * offset3;
  const float v1 = std_max(value * mul0, 1.0e-3f);
  const float v2 = 1.0f / (v1 + offset2);
  const float v3 = 1.0f / fmaf(v1, v1, offset3);
  const float v4 = 1.0f / fmaf(v1, v1, offset4);
  return base + fmaf(mul4, v4, fmaf(mul2, v2, mul3 * v3));
}

/* :164-170 */
static float masking_sqrt(float value) {
  const float log_offset = 27.505837037000106f;
  const float mul = 211.66567973503678f;
  const float inner = sqrtf(mul * 1.0e8f);
  return 0.25f * sqrtf(fmaf(value, inner, log_offset));
}

/* :172-196 */
static void store_min4(float v, float* m0, float* m1, float* m2, float* m3) {
  if (v >= *m3) return;
  if (v < *m0) { *m3 = *m2; *m2 = *m1; *m1 = *m0; *m0 = v; }
  else if (v < *m1) { *m3 = *m2; *m2 = *m1; *m1 = v; }
  else if (v < *m2) { *m3 = *m2; *m2 = v; }
  else { *m3 = v; }
}

/* :198-217 */
static void sort4(float v[4]) {
#define SW(a, b) if (v[a] > v[b]) { float t = v[a]; v[a] = v[b]; v[b] = t; }
  SW(0, 1) SW(0, 2) SW(0, 3) SW(1, 2) SW(1, 3) SW(2, 3)
#undef SW
}

/* :219-287 FuzzyErosion: pre-erosion (w4 x h4) -> block field (w4/2 x h4/2) */
static void fuzzy_erosion(float target, size_t sw, size_t sh, const float* src,
                          size_t dw, float* dst) {
  const float mul_base[4] = {0.125f, 0.1f, 0.09f, 0.06f};
  const float mul_add[4] = {0.0f, -0.1f, -0.09f, -0.06f};
  const float total = 0.29959705784054957f;
  float mix = 0.0f;
  if (target < 2.0f) mix = (2.0f - target) * 0.5f;
  float weights[4], wsum = 0.0f;
  for (int i = 0; i < 4; ++i) {
    weights[i] = mul_base[i] + mix * mul_add[i];
    wsum += weights[i];
  }
  for (int i = 0; i < 4; ++i) weights[i] *= total / wsum;
  for (size_t y = 0; y < sh; ++y) {
    const size_t top = y == 0 ? y : y - 1;
    const size_t bottom = y + 1 < sh ? y + 1 : y;
    for (size_t x = 0; x < sw; ++x) {
      const size_t left = x == 0 ? x : x - 1;
      const size_t right = x + 1 < sw ? x + 1 : x;
#define AT(xx, yy) src[(yy) * sw + (xx)]
      float m[4] = {AT(x, y), AT(left, y), AT(right, y), AT(left, top)};
      sort4(m);
      store_min4(AT(x, top), &m[0], &m[1], &m[2], &m[3]);
      store_min4(AT(right, top), &m[0], &m[1], &m[2], &m[3]);
      store_min4(AT(left, bottom), &m[0], &m[1], &m[2], &m[3]);
      store_min4(AT(x, bottom), &m[0], &m[1], &m[2], &m[3]);
      store_min4(AT(right, bottom), &m[0], &m[1], &m[2], &m[3]);
#undef AT
      const float value = weights[0] * m[0] + weights[1] * m[1] +
                          weights[2] * m[2] + weights[3] * m[3];
      const size_t di = (y / 2) * dw + x / 2;
      if ((x & 1u) == 0 && (y & 1u) == 0) dst[di] = value;
      else dst[di] += value;
    }
  }
}

/* :332-362 */
static float gamma_modulation(const float* X, const float* Y, size_t w,
                              size_t bx, size_t by, float value) {
  float lane[4] = {0, 0, 0, 0};
  for (size_t dy = 0; dy < 8; ++dy) {
    for (size_t dx = 0; dx < 8; ++dx) {
      const size_t i = (by + dy) * w + bx + dx;
      const float in_y = Y[i] + 0.16f;
      const float in_x = X[i];
      lane[dx & 3] += gamma_derivative_ratio(in_y - in_x, 1);
      lane[dx & 3] += gamma_derivative_ratio(in_y + in_x, 1);
    }
  }
  const float overall = ((lane[0] + lane[1]) + (lane[2] + lane[3])) *
                        (0.5f / 64.0f);
  if (!isfinite(overall) || overall <= 0.0f) return NAN;
  return fmaf(0.1005613337192697f, oracle_fast_log2(overall), value);
}

/* :364-392 */
static float high_frequency_modulation(const float* Y, size_t w, size_t bx,
                                       size_t by, float value) {
  const float limit = 0.0206f;
  float lane[4] = {0, 0, 0, 0};
  for (size_t dy = 0; dy < 8; ++dy) {
    const float* row = Y + (by + dy) * w + bx;
    const float* next = dy + 1 < 8 ? Y + (by + dy + 1) * w + bx : row;
    for (size_t dx = 0; dx < 8; ++dx) {
      if (dx + 1 < 8) {
        lane[dx & 3] += std_min(limit, fabsf(row[dx] - row[dx + 1]));
      }
      lane[dx & 3] += std_min(limit, fabsf(row[dx] - next[dx]));
    }
  }
  const float sum = (lane[0] + lane[1]) + (lane[2] + lane[3]);
  return value + (sum * -0.38f + 0.42f);
}

/* :394-424 */
static float blue_modulation(const float* X, const float* Y, const float* B,
                             size_t w, size_t bx, size_t by, float value) {
  const float limit = 0.010474084867598155f;
  const float offset = 0.0031994768654636393f;
  float lane[4] = {0, 0, 0, 0};
  for (size_t dy = 0; dy < 8; ++dy) {
    for (size_t dx = 0; dx < 8; ++dx) {
      const size_t i = (by + dy) * w + bx + dx;
      const float ey = Y[i] + offset + fabsf(X[i]);
      if (B[i] > ey) lane[dx & 3] += std_min(B[i] - ey, limit);
    }
  }
  float sum = (lane[0] + lane[1]) + (lane[2] + lane[3]);
  if (sum >= 32.0f * limit) sum = 64.0f * limit - sum;
  sum = std_min(sum, 15.463398341612438f * limit);
  return value + sum * 0.90590804735610064f;
}

/* :583-738 ComputeInitialQuantField (non-uniform branch).  opsin: padded
 * planar 3 x h x w, w and h multiples of 8.  Outputs: quant_field and
 * strategy_mask (w/8 x h/8), pixel_mask (w x h, blurred). */
int oracle_initial_quant_field(const float* opsin, size_t w, size_t h,
                               float target, float rescale, float* quant_field,
                               float* strategy_mask, float* pixel_mask) {
  const size_t n = w * h, bw = w / 8, bh = h / 8, pw4 = w / 4, ph4 = h / 4;
  const float* X = opsin;
  const float* Y = opsin + n;
  const float* B = opsin + 2 * n;
  float* raw_mask = malloc(sizeof(float) * n);
  float* pre = malloc(sizeof(float) * pw4 * ph4);
  float* rowd = malloc(sizeof(float) * w);
  if (!raw_mask || !pre || !rowd) { free(raw_mask); free(pre); free(rowd); return -1; }
  const float match_gamma_offset = 0.019f, diff_limit = 0.2f;
  for (size_t gy = 0; gy < ph4; ++gy) {
    for (size_t r = 0; r < 4; ++r) {
      const size_t y = gy * 4 + r;
      const size_t top = y == 0 ? y : y - 1;
      const size_t bottom = y + 1 < h ? y + 1 : y;
      const float* row = Y + y * w;
      const float* trow = Y + top * w;
      const float* brow = Y + bottom * w;
      for (size_t x = 0; x < w; ++x) {
        const size_t left = x == 0 ? x : x - 1;
        const size_t right = x + 1 < w ? x + 1 : x;
        const float base = 0.25f * (brow[x] + trow[x] + row[left] + row[right]);
        const float gamma = gamma_derivative_ratio(row[x] + match_gamma_offset, 0);
        float pd = fabsf(gamma * (row[x] - base));
        pd = log1pf(pd);
        raw_mask[y * w + x] = 1.0f / (pd + 0.01f);
        float bd = gamma * (row[x] - base);
        bd *= bd;
        bd = std_min(bd, diff_limit);
        bd = masking_sqrt(bd);
        if (r == 0) rowd[x] = bd; else rowd[x] += bd;
      }
    }
    for (size_t x = 0; x < pw4; ++x) {
      pre[gy * pw4 + x] = (rowd[4 * x] + rowd[4 * x + 1] + rowd[4 * x + 2] +
                           rowd[4 * x + 3]) * 0.25f;
    }
  }
  fuzzy_erosion(target, pw4, ph4, pre, bw, quant_field);
  for (size_t i = 0; i < bw * bh; ++i) {
    strategy_mask[i] = 1.0f / (quant_field[i] + 0.001f);
  }
  /* :426-467 PerBlockModulations */
  const float scale = 0.765f / target * rescale;
  const float base_level = 0.48f * scale;
  float dampen = 1.0f;
  if (target >= 2.0f) {
    dampen = std_max(1.0f - (target - 2.0f) / 12.0f, 0.0f);
  }
  const float multiplier = scale * dampen;
  const float addend = (1.0f - dampen) * base_level;
  for (size_t by = 0; by < bh; ++by) {
    for (size_t bx = 0; bx < bw; ++bx) {
      const size_t i = by * bw + bx;
      const float mask = compute_mask(quant_field[i]);
      const float g = gamma_modulation(X, Y, w, bx * 8, by * 8, mask);
      const float hf = high_frequency_modulation(Y, w, bx * 8, by * 8, g);
      const float bl = blue_modulation(X, Y, B, w, bx * 8, by * 8, g);
      const float e = std_min(hf, bl);
      quant_field[i] = oracle_fast_pow2(e * 1.442695041f) * multiplier + addend;
    }
  }
  /* :263-292 BlurPixelMask */
  const float f[5] = {0.364911248f, 0.05f, 0.1688888021f, 0.221069183f,
                      0.306563504f};
  const float partial = f[0] + f[1] + f[2] + f[4];
  double ws = 1.0 + 4.0 * ((double)partial + 2.0 * (double)f[3]);
  if (ws < 1.0e-5) ws = 1.0e-5;
  const float norm = (float)(1.0 / ws);
  const float weights[6] = {norm, norm * f[0], norm * f[2], norm * f[1],
                            norm * f[4], norm * f[3]};
  oracle_convolve_symmetric5(raw_mask, w, h, weights, pixel_mask);
  free(raw_mask);
  free(pre);
  free(rowd);
  for (size_t i = 0; i < bw * bh; ++i) {
    if (!isfinite(quant_field[i]) || !(quant_field[i] > 0.0f)) return 1;
  }
  return 0;
}

/* src/codec/adaptive_quantization.cpp:556-580 */
float oracle_initial_quant_dc(float target) {
  const float dc_mul = 0.3f;
  const float nonlinear = dc_mul * powf(target / dc_mul, 0.83f);
  const float target_dc = std_max(0.5f * target, std_min(target, nonlinear));
  return std_min(1.095924047623553f / target_dc, 50.0f);
}
