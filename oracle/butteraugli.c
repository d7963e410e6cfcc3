/*
 * Oracle: the native scalar Butteraugli of the reference (opsin dynamics,
 * frequency separation, Malta / L2 / masking difference stages, two-scale
 * composition), its per-transform 16-norm reduction and the AQ field update.
 * Test infrastructure only; see oracle.h.
 *
 * Images are planar binary32 with stride == width.  Every product-then-add
 * the reference writes through its volatile UnfusedMultiplyAdd is a product
 * rounded to binary32 followed by the add (this file is compiled with
 * -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static inline float bmax(float a, float b) { return a < b ? b : a; }

/* Butteraugli's Mirror (src/codec/butteraugli_internal.cpp:235-244). */
static int64_t bmirror(int64_t c, int64_t size) {
  while (c < 0 || c >= size) c = c < 0 ? -c - 1 : 2 * size - 1 - c;
  return c;
}

/* ---- Gaussian blur: src/codec/butteraugli_internal.cpp:245-620 ---------- */

int oracle_butteraugli_kernel(float sigma, float* kernel) {
  const float scaled = 2.25f * fabsf(sigma);
  int radius = (int)scaled;
  if (radius < 1) radius = 1;
  const int size = 2 * radius + 1;
  if (size > 33) return 0;
  const double exponent_scale = -1.0 / (2.0 * sigma * sigma);
  for (int i = -radius; i <= radius; ++i) {
    kernel[i + radius] = (float)exp(exponent_scale * i * i);
  }
  return size;
}

static float horizontal5(const float* row, int64_t x, int64_t w, float w0,
                         float w1, float w2) {
  float r = row[x] * w0;
  r = (row[bmirror(x - 1, w)] + row[bmirror(x + 1, w)]) * w1 + r;
  r = (row[bmirror(x - 2, w)] + row[bmirror(x + 2, w)]) * w2 + r;
  return r;
}

static void separable5(const float* in, size_t w, size_t h, const float* k,
                       float* out) {
  float sum = 0.0f;
  for (int i = 0; i < 5; ++i) sum += k[i];
  const float scale = 1.0f / sum;
  const float w0 = k[2] * scale, w1 = k[1] * scale, w2 = k[0] * scale;
  const int64_t W = (int64_t)w, H = (int64_t)h;
  for (int64_t y = 0; y < H; ++y) {
    for (int64_t x = 0; x < W; ++x) {
      const float c = horizontal5(in + y * W, x, W, w0, w1, w2);
      const float t1 = horizontal5(in + bmirror(y - 1, H) * W, x, W, w0, w1, w2);
      const float b1 = horizontal5(in + bmirror(y + 1, H) * W, x, W, w0, w1, w2);
      const float t2 = horizontal5(in + bmirror(y - 2, H) * W, x, W, w0, w1, w2);
      const float b2 = horizontal5(in + bmirror(y + 2, H) * W, x, W, w0, w1, w2);
      float r = c * w0;
      r = (t1 + b1) * w1 + r;
      r = (t2 + b2) * w2 + r;
      out[y * W + x] = r;
    }
  }
}

/* One pass: each row of `in` (w x h) convolved along x; the result is
 * written transposed (out is h wide, w tall), as the reference's
 * ConvolutionWithTranspose does. */
static void convolve_transpose(const float* in, size_t w, size_t h,
                               const float* kernel, int size, float* out) {
  const int radius = size / 2;
  float full = 0.0f;
  for (int i = 0; i < size; ++i) full += kernel[i];
  const float scale = 1.0f / full;
  float sk[17];
  for (int i = 0; i <= radius; ++i) sk[i] = kernel[i] * scale;
  const size_t first = w < (size_t)radius ? w : (size_t)radius;
  const size_t last = w > (size_t)radius ? w - radius : 0;
  if (first < last) {
    for (size_t y = 0; y < h; ++y) {
      for (size_t x = first; x < last; ++x) {
        const float* s = in + y * w + x - radius;
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int i = 0; i < radius; ++i) {
          const float pair = (s[i] + s[2 * radius - i]) * sk[i];
          acc[i % 4] = i < 4 ? pair : acc[i % 4] + pair;
        }
        float r = s[radius] * sk[radius];
        for (int i = 0; i < 4 && i < radius; ++i) r = r + acc[i];
        out[x * h + y] = r;
      }
    }
  }
  for (size_t x = 0; x < w; ++x) {
    if (x >= first && x < last) continue;
    const size_t lo = x < (size_t)radius ? 0 : x - radius;
    const size_t hi = (w - 1 < x + radius) ? w - 1 : x + radius;
    float weight = 0.0f;
    for (size_t sx = lo; sx <= hi; ++sx) weight += kernel[sx + radius - x];
    const float s = 1.0f / weight;
    for (size_t y = 0; y < h; ++y) {
      float sum = 0.0f;
      for (size_t sx = lo; sx <= hi; ++sx) sum += in[y * w + sx] * kernel[sx + radius - x];
      out[x * h + y] = sum * s;
    }
  }
}

int oracle_butteraugli_blur(const float* in, size_t w, size_t h, float sigma,
                            float* out) {
  float kernel[33];
  const int size = oracle_butteraugli_kernel(sigma, kernel);
  if (size == 5) {
    separable5(in, w, h, kernel, out);
    return 0;
  }
  if (size != 7 && size != 13 && size != 15 && size != 33) return -1;
  float* t = malloc(sizeof(float) * w * h);
  if (!t) return -1;
  convolve_transpose(in, w, h, kernel, size, t);
  convolve_transpose(t, h, w, kernel, size, out);
  free(t);
  return 0;
}

/* ---- opsin dynamics: :94-180, :622-700 ---------------------------------- */

static float butteraugli_fast_log2(float v) {
  uint32_t bits;
  memcpy(&bits, &v, 4);
  const int32_t e = (int32_t)(bits - 0x3f2aaaabu) >> 23;
  const uint32_t mbits = bits - ((uint32_t)e << 23);
  float m;
  memcpy(&m, &mbits, 4);
  const float x = m - 1.0f;
  float num = 0.74245873327820566f * x + 1.4287160470083755f;
  num = num * x + -1.8503833400518310e-06f;
  float den = 0.17409343003366853f * x + 1.0096718572241148f;
  den = den * x + 0.99032814277590719f;
  return num / den + (float)e;
}

static float butteraugli_gamma(float v) {
  const float ret_mul = 19.245013259874995f * 0.6931471805599453f;
  v = bmax(v, 0.0f);
  return ret_mul * butteraugli_fast_log2(v + 9.9710635769299145f) +
         -23.16046239805755f;
}

static void opsin_absorbance(float r, float g, float b, int clamp, float out[3]) {
  out[0] = 0.29956550340058319f * r +
           (0.63373087833825936f * g + (0.077705617820981968f * b + 1.7557483643287353f));
  out[1] = 0.22158691104574774f * r +
           (0.69391388044116142f * g + (0.0987313588422f * b + 1.7557483643287353f));
  out[2] = 0.02f * r + (0.02f * g + (0.20480129041026129f * b + 12.226454707163354f));
  if (clamp) {
    out[0] = bmax(out[0], 1.7557483643287353f);
    out[1] = bmax(out[1], 1.7557483643287353f);
    out[2] = bmax(out[2], 12.226454707163354f);
  }
}

int oracle_butteraugli_opsin(const float* rgb, size_t w, size_t h,
                             float intensity, float* xyb) {
  const size_t n = w * h;
  float* blurred = malloc(sizeof(float) * 3 * n);
  if (!blurred) return -1;
  for (int c = 0; c < 3; ++c) {
    if (oracle_butteraugli_blur(rgb + c * n, w, h, 1.2f, blurred + c * n)) {
      free(blurred);
      return -1;
    }
  }
  for (size_t i = 0; i < n; ++i) {
    float pre[3], cur[3], sens[3];
    opsin_absorbance(blurred[i] * intensity, blurred[n + i] * intensity,
                     blurred[2 * n + i] * intensity, 1, pre);
    for (int c = 0; c < 3; ++c) {
      pre[c] = bmax(pre[c], 1.0e-4f);
      sens[c] = bmax(butteraugli_gamma(pre[c]) / pre[c], 1.0e-4f);
    }
    opsin_absorbance(rgb[i] * intensity, rgb[n + i] * intensity,
                     rgb[2 * n + i] * intensity, 0, cur);
    cur[0] = bmax(cur[0] * sens[0], 1.7557483643287353f);
    cur[1] = bmax(cur[1] * sens[1], 1.7557483643287353f);
    cur[2] = bmax(cur[2] * sens[2], 12.226454707163354f);
    xyb[i] = cur[0] - cur[1];
    xyb[n + i] = cur[0] + cur[1];
    xyb[2 * n + i] = cur[2];
  }
  free(blurred);
  return 0;
}

/* ---- frequency separation: :702-846 ------------------------------------- */

static float maximum_clamp(float v, float maximum) {
  const float k = 0.724216145665f;
  if (v >= maximum) return (v - maximum) * k + maximum;
  if (v < -maximum) return (v + maximum) * k + -maximum;
  return v;
}
static float remove_range(float v, float width) {
  if (v > width) return v - width;
  if (v < -width) return v + width;
  return 0.0f;
}
static float amplify_range(float v, float width) {
  if (v > width) return v + width;
  if (v < -width) return v - width;
  return v + v;
}
static float suppress_x_by_y(float y, float x) {
  const float s = 46.0f, min_scale = 0.653020556257f;
  const float scaler = (s / (y * y + s)) * (1.0f - min_scale) + min_scale;
  return scaler * x;
}

/* psycho: 10 planes -- LF X,Y,B; MF X,Y,B; HF X,Y; UHF X,Y. */
int oracle_butteraugli_frequencies(const float* xyb, size_t w, size_t h,
                                   float* psycho) {
  const size_t n = w * h;
  float* lf = psycho;
  float* mf = psycho + 3 * n;
  float* hf = psycho + 6 * n;
  float* uhf = psycho + 8 * n;
  float* blurred = malloc(sizeof(float) * n);
  if (!blurred) return -1;
  for (int c = 0; c < 3; ++c) {
    oracle_butteraugli_blur(xyb + c * n, w, h, 7.15593339443f, lf + c * n);
    for (size_t i = 0; i < n; ++i) mf[c * n + i] = xyb[c * n + i] - lf[c * n + i];
  }
  for (size_t i = 0; i < n; ++i) {
    const float ly = lf[n + i];
    lf[2 * n + i] = (-0.362267051518f * ly + lf[2 * n + i]) * 49.87984651440f;
    lf[i] *= 33.832837186260f;
    lf[n + i] = ly * 14.458268100570f;
  }
  for (int c = 0; c < 3; ++c) {
    if (c < 2) memcpy(hf + c * n, mf + c * n, sizeof(float) * n);
    oracle_butteraugli_blur(mf + c * n, w, h, 3.22489901262f, blurred);
    if (c == 2) {
      memcpy(mf + 2 * n, blurred, sizeof(float) * n);
      continue;
    }
    for (size_t i = 0; i < n; ++i) {
      hf[c * n + i] -= blurred[i];
      mf[c * n + i] = c == 0 ? remove_range(blurred[i], 0.29f)
                             : amplify_range(blurred[i], 0.1f);
    }
  }
  for (size_t i = 0; i < n; ++i) hf[i] = suppress_x_by_y(hf[n + i], hf[i]);
  for (int c = 0; c < 2; ++c) {
    memcpy(uhf + c * n, hf + c * n, sizeof(float) * n);
    oracle_butteraugli_blur(hf + c * n, w, h, 1.56416327805f, blurred);
    for (size_t i = 0; i < n; ++i) {
      float hv = blurred[i];
      if (c == 0) {
        uhf[i] = remove_range(uhf[i] - hv, 0.04f);
        hf[i] = remove_range(hv, 1.5f);
      } else {
        hv = maximum_clamp(hv, 28.4691806922f);
        float uv = uhf[n + i] - hv;
        uv = maximum_clamp(uv, 5.19175294647f);
        uhf[n + i] = uv * 2.69313763794f;
        hf[n + i] = amplify_range(hv * 2.155f, 0.132f);
      }
    }
  }
  free(blurred);
  return 0;
}

/* ---- difference stages: src/codec/butteraugli_distance_internal.cpp ----- */

static const double kMaltaWeights[6] = {37.0819870399, 8246.75321353, 18.7237414387,
                                        6923.99476109, 1.10039032555, 173.5};
static const double kMaltaNorms[6] = {130262059.556, 1009002.70582, 4498534.45232,
                                      8051.15833247, 71.7800275169, 5.0};
static const float kL2Weights[9] = {400.0f, 1.50815703118f, 0.0f,
                                    2150.0f, 10.6195433239f, 16.2176043152f,
                                    29.2353797994f, 0.844626970982f,
                                    0.703646627719f};

static float zb(const float* p, size_t w, size_t h, size_t x, size_t y, int dx,
                int dy) {
  const int64_t sx = (int64_t)x + dx, sy = (int64_t)y + dy;
  if (sx < 0 || sy < 0 || sx >= (int64_t)w || sy >= (int64_t)h) return 0.0f;
  return p[sy * (int64_t)w + sx];
}

static float sum4(float a, float b, float c, float d) { return (a + b) + (c + d); }
static float sum5(float a, float b, float c, float d, float e) { return sum4(a, b, c, d + e); }
static float sum7(float a, float b, float c, float d, float e, float f, float g) {
  return sum4(a, b, c, sum4(d, e, f, g));
}
static float sum9(float a, float b, float c, float d, float e, float f, float g,
                  float hh, float i) {
  return (sum4(a, b, c, d) + sum4(e, f, g, hh)) + i;
}

/* Malta line patterns (:208-333): 16 lines of 5 taps (LF) or 7/9 taps. */
static const int8_t kLf[16][5][2] = {
  {{-4, 0}, {-2, 0}, {0, 0}, {2, 0}, {4, 0}},
  {{0, -4}, {0, -2}, {0, 0}, {0, 2}, {0, 4}},
  {{-3, -3}, {-2, -2}, {0, 0}, {2, 2}, {3, 3}},
  {{3, -3}, {2, -2}, {0, 0}, {-2, 2}, {-3, 3}},
  {{1, -4}, {1, -2}, {0, 0}, {-1, 2}, {-1, 4}},
  {{-1, -4}, {-1, -2}, {0, 0}, {1, 2}, {1, 4}},
  {{-4, -1}, {-2, -1}, {0, 0}, {2, 1}, {4, 1}},
  {{-4, 1}, {-2, 1}, {0, 0}, {2, -1}, {4, -1}},
  {{-2, -3}, {-1, -2}, {0, 0}, {1, 2}, {2, 3}},
  {{2, -3}, {1, -2}, {0, 0}, {-1, 2}, {-2, 3}},
  {{-3, -2}, {-2, -1}, {0, 0}, {2, 1}, {3, 2}},
  {{3, -2}, {2, -1}, {0, 0}, {-2, 1}, {-3, 2}},
  {{-4, 2}, {-2, 1}, {0, 0}, {2, -1}, {4, -2}},
  {{-4, -2}, {-2, -1}, {0, 0}, {2, 1}, {4, 2}},
  {{-2, -4}, {-1, -2}, {0, 0}, {1, 2}, {2, 4}},
  {{2, -4}, {1, -2}, {0, 0}, {-1, 2}, {-2, 4}},
};
/* Full patterns: taps per line (9 or 7), unused entries ignored. */
static const int kFullTaps[16] = {9, 9, 7, 7, 9, 9, 9, 9, 7, 7, 7, 7, 9, 9, 9, 9};
static const int8_t kFull[16][9][2] = {
  {{-4, 0}, {-3, 0}, {-2, 0}, {-1, 0}, {0, 0}, {1, 0}, {2, 0}, {3, 0}, {4, 0}},
  {{0, -4}, {0, -3}, {0, -2}, {0, -1}, {0, 0}, {0, 1}, {0, 2}, {0, 3}, {0, 4}},
  {{-3, -3}, {-2, -2}, {-1, -1}, {0, 0}, {1, 1}, {2, 2}, {3, 3}},
  {{3, -3}, {2, -2}, {1, -1}, {0, 0}, {-1, 1}, {-2, 2}, {-3, 3}},
  {{1, -4}, {1, -3}, {1, -2}, {0, -1}, {0, 0}, {0, 1}, {-1, 2}, {-1, 3}, {-1, 4}},
  {{-1, -4}, {-1, -3}, {-1, -2}, {0, -1}, {0, 0}, {0, 1}, {1, 2}, {1, 3}, {1, 4}},
  {{-4, -1}, {-3, -1}, {-2, -1}, {-1, 0}, {0, 0}, {1, 0}, {2, 1}, {3, 1}, {4, 1}},
  {{-4, 1}, {-3, 1}, {-2, 1}, {-1, 0}, {0, 0}, {1, 0}, {2, -1}, {3, -1}, {4, -1}},
  {{-2, -3}, {-1, -2}, {-1, -1}, {0, 0}, {1, 1}, {1, 2}, {2, 3}},
  {{2, -3}, {1, -2}, {1, -1}, {0, 0}, {-1, 1}, {-1, 2}, {-2, 3}},
  {{-3, -2}, {-2, -1}, {-1, -1}, {0, 0}, {1, 1}, {2, 1}, {3, 2}},
  {{3, -2}, {2, -1}, {1, -1}, {0, 0}, {-1, 1}, {-2, 1}, {-3, 2}},
  {{-4, 1}, {-3, 1}, {-2, 1}, {-1, 0}, {0, 0}, {1, 0}, {2, -1}, {3, -1}, {4, -1}},
  {{-4, -1}, {-3, -1}, {-2, -1}, {-1, 0}, {0, 0}, {1, 0}, {2, 1}, {3, 1}, {4, 1}},
  {{-1, -4}, {-1, -3}, {-1, -2}, {0, -1}, {0, 0}, {0, 1}, {1, 2}, {1, 3}, {1, 4}},
  {{1, -4}, {1, -3}, {1, -2}, {0, -1}, {0, 0}, {0, 1}, {-1, 2}, {-1, 3}, {-1, 4}},
};

static float malta_unit(const float* d, size_t w, size_t h, size_t x, size_t y,
                        int lf) {
  float result = 0.0f;
  for (int l = 0; l < 16; ++l) {
    float v[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float s;
    if (lf) {
      for (int t = 0; t < 5; ++t) v[t] = zb(d, w, h, x, y, kLf[l][t][0], kLf[l][t][1]);
      s = sum5(v[0], v[1], v[2], v[3], v[4]);
    } else {
      const int taps = kFullTaps[l];
      for (int t = 0; t < taps; ++t) v[t] = zb(d, w, h, x, y, kFull[l][t][0], kFull[l][t][1]);
      s = taps == 9 ? sum9(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8])
                    : sum7(v[0], v[1], v[2], v[3], v[4], v[5], v[6]);
    }
    result = l == 0 ? s * s : s * s + result;
  }
  return result;
}

/* MaltaDiffMap (:482-577) into out (overwritten: out = 0 + malta). */
static void malta_diff_map(const float* p0, const float* p1, size_t w, size_t h,
                           int lf, double w_gt, double w_lt, double norm,
                           float* scratch, float* out) {
  const double multiplier = lf ? 0.611612573796 : 0.39905817637;
  const double pre_gt = multiplier * sqrt(0.5f * w_gt) / (3.75 * 2.0 + 1.0);
  const double pre_lt = multiplier * sqrt(0.33f * w_lt) / (3.75 * 2.0 + 1.0);
  const float n2_gt = (float)(pre_gt * norm), n2_lt = (float)(pre_lt * norm);
  const size_t n = w * h;
  for (size_t i = 0; i < n; ++i) {
    const float v0 = p0[i], v1 = p1[i];
    const float absolute = 0.5f * (fabsf(v0) + fabsf(v1));
    const float diff = v0 - v1;
    const float scaler = n2_gt / ((float)norm + absolute);
    scratch[i] = scaler * diff;
    const float scaler2 = n2_lt / ((float)norm + absolute);
    const double magnitude = fabs((double)v0);
    const double too_small = 0.55 * magnitude, too_big = 1.05 * magnitude;
    if (v0 < 0.0f) {
      if (v1 > -too_small) {
        const double impact = scaler2 * (v1 + too_small);
        scratch[i] = (float)(scratch[i] - impact);
      } else if (v1 < -too_big) {
        const double impact = scaler2 * (-v1 - too_big);
        scratch[i] = (float)(scratch[i] + impact);
      }
    } else if (v1 < too_small) {
      const double impact = scaler2 * (too_small - v1);
      scratch[i] = (float)(scratch[i] + impact);
    } else if (v1 > too_big) {
      const double impact = scaler2 * (v1 - too_big);
      scratch[i] = (float)(scratch[i] - impact);
    }
  }
  for (size_t y = 0; y < h; ++y)
    for (size_t x = 0; x < w; ++x) out[y * w + x] = 0.0f + malta_unit(scratch, w, h, x, y, lf);
}

static void store_min3(float v, float* m0, float* m1, float* m2) {
  if (v < *m2) {
    if (v < *m0) {
      *m2 = *m1;
      *m1 = *m0;
      *m0 = v;
    } else if (v < *m1) {
      *m2 = *m1;
      *m1 = v;
    } else {
      *m2 = v;
    }
  }
}

/* FuzzyErosion (:579-630), step 3. */
void oracle_butteraugli_erosion(const float* in, size_t w, size_t h, float* out) {
  const size_t s = 3;
  for (size_t y = 0; y < h; ++y) {
    for (size_t x = 0; x < w; ++x) {
      float m0 = in[y * w + x];
      float m1 = 2.0f * m0, m2 = m1;
      if (x >= s) {
        store_min3(in[y * w + x - s], &m0, &m1, &m2);
        if (y >= s) store_min3(in[(y - s) * w + x - s], &m0, &m1, &m2);
        if (y + s < h) store_min3(in[(y + s) * w + x - s], &m0, &m1, &m2);
      }
      if (x + s < w) {
        store_min3(in[y * w + x + s], &m0, &m1, &m2);
        if (y >= s) store_min3(in[(y - s) * w + x + s], &m0, &m1, &m2);
        if (y + s < h) store_min3(in[(y + s) * w + x + s], &m0, &m1, &m2);
      }
      if (y >= s) store_min3(in[(y - s) * w + x], &m0, &m1, &m2);
      if (y + s < h) store_min3(in[(y + s) * w + x], &m0, &m1, &m2);
      out[y * w + x] = 0.45f * m0 + 0.3f * m1 + 0.25f * m2;
    }
  }
}

/* MaskY / MaskDcY (:632-655) */
static float mask_y(float delta) {
  const double gs = 1.0 / (17.83 * 0.79079917404);
  const double v = gs * (1.0 + 2.5485944793 / (0.451936922203 * delta + 0.829591754942));
  return (float)(v * v);
}
static float mask_dc_y(float delta) {
  const double gs = 1.0 / (17.83 * 0.79079917404);
  const double v = gs * (1.0 + 0.505054525019 / (3.87449418804 * delta + 0.20025578522));
  return (float)(v * v);
}

/* CombineChannelsForMasking + DiffPrecompute (:335-360) + blur 2.7 */
static void mask_source(const float* ps, size_t w, size_t h, float* tmp,
                        float* out) {
  const size_t n = w * h;
  const float* hf = ps + 6 * n;
  const float* uhf = ps + 8 * n;
  const float mul = 6.19424080439f;
  const float bias = mul * 12.61050594197f;
  const float sqrt_bias = sqrtf(bias);
  for (size_t i = 0; i < n; ++i) {
    const float xd = (uhf[i] + hf[i]) * 2.5f;
    const float yd = uhf[n + i] * 0.4f + hf[n + i] * 0.4f;
    const float a = sqrtf(xd * xd + yd * yd);
    tmp[i] = sqrtf(mul * fabsf(a) + bias) - sqrt_bias;
  }
  oracle_butteraugli_blur(tmp, w, h, 2.7f, out);
}

/* ComputeDifferenceStages (:657-870).  stages: 9 planes in the order of
 * DifferenceStage (6 Malta maps, mask, masked AC Y, final composition). */
int oracle_butteraugli_differences(const float* ps0, const float* ps1, size_t w,
                                   size_t h, float hf_asymmetry,
                                   float x_multiplier, float* stages) {
  const size_t n = w * h;
  float* scratch = malloc(sizeof(float) * n * 10);
  if (!scratch) return -1;
  float* ac = scratch + n;       /* 3 planes */
  float* dc = scratch + 4 * n;   /* 3 planes */
  float* b0 = scratch + 7 * n;
  float* b1 = scratch + 8 * n;
  float* tmp = scratch + 9 * n;
  const float asym = hf_asymmetry;
  const float sqrt_asym = sqrtf(asym);
  const int plane_of[6] = {4, 3, 7, 6, 9, 8}; /* MF Y, MF X, HF Y, HF X, UHF Y, UHF X */
  const double up[6] = {kMaltaWeights[0], kMaltaWeights[1],
                        kMaltaWeights[2] * sqrt_asym, kMaltaWeights[3] * sqrt_asym,
                        kMaltaWeights[4] * asym, kMaltaWeights[5] * asym};
  const double down[6] = {kMaltaWeights[0], kMaltaWeights[1],
                          kMaltaWeights[2] / sqrt_asym, kMaltaWeights[3] / sqrt_asym,
                          kMaltaWeights[4] / asym, kMaltaWeights[5] / asym};
  for (int k = 0; k < 6; ++k) {
    malta_diff_map(ps0 + plane_of[k] * n, ps1 + plane_of[k] * n, w, h, k < 4,
                   up[k], down[k], kMaltaNorms[k], scratch, stages + k * n);
  }
  memset(ac, 0, sizeof(float) * 6 * n);
  const int order[6] = {4, 5, 2, 3, 0, 1};
  for (int o = 0; o < 6; ++o) {
    const int k = order[o];
    float* a = ac + (k % 2 == 0 ? 1 : 0) * n;
    for (size_t i = 0; i < n; ++i) a[i] += stages[k * n + i];
  }
  for (int c = 0; c < 3; ++c) {
    float* a = ac + c * n;
    if (c < 2) {
      /* L2DiffAsymmetric (:421-470) */
      const float wg = kL2Weights[c] * asym * 0.8f;
      const float wl = kL2Weights[c] / asym * 0.8f;
      const float* h0 = ps0 + (6 + c) * n;
      const float* h1 = ps1 + (6 + c) * n;
      if (!(kL2Weights[c] * asym == 0.0f && kL2Weights[c] / asym == 0.0f)) {
        for (size_t i = 0; i < n; ++i) {
          const float v0 = h0[i], v1 = h1[i];
          const float d = v0 - v1;
          const float total = (d * d) * wg + a[i];
          const float mag = fabsf(v0);
          const float small = 0.4f * mag;
          float sec = 0.0f;
          if (v0 < 0.0f) {
            if (v1 > -small) sec = v1 + small;
            else if (v1 < -mag) sec = -v1 - mag;
          } else if (v1 < small) {
            sec = small - v1;
          } else if (v1 > mag) {
            sec = v1 - mag;
          }
          a[i] = wl * (sec * sec) + total;
        }
      }
    }
    /* L2Diff (:379-400) */
    const float wm = kL2Weights[3 + c];
    if (wm != 0.0f) {
      const float* m0 = ps0 + (3 + c) * n;
      const float* m1 = ps1 + (3 + c) * n;
      for (size_t i = 0; i < n; ++i) {
        const float d = m0[i] - m1[i];
        a[i] = (d * d) * wm + a[i];
      }
    }
    /* SetL2Diff (:402-419) */
    const float wd = kL2Weights[6 + c];
    const float* l0 = ps0 + c * n;
    const float* l1 = ps1 + c * n;
    for (size_t i = 0; i < n; ++i) {
      const float d = l0[i] - l1[i];
      dc[c * n + i] = wd == 0.0f ? 0.0f : d * d * wd;
    }
  }
  mask_source(ps0, w, h, tmp, b0);
  oracle_butteraugli_erosion(b0, w, h, stages + 6 * n);
  mask_source(ps1, w, h, tmp, b1);
  for (size_t i = 0; i < n; ++i) {
    const float d = b0[i] - b1[i];
    ac[n + i] += 10.0f * d * d;
  }
  memcpy(stages + 7 * n, ac + n, sizeof(float) * n);
  for (size_t i = 0; i < n; ++i) {
    const float m = stages[6 * n + i];
    const float my = mask_y(m), mdc = mask_dc_y(m);
    const float dc0 = dc[i] * x_multiplier;
    const float ac0 = ac[i] * x_multiplier;
    const float masked_dc = dc0 * mdc + dc[n + i] * mdc + dc[2 * n + i] * mdc;
    const float masked_ac = ac0 * my + ac[n + i] * my + ac[2 * n + i] * my;
    stages[8 * n + i] = sqrtf(masked_dc + masked_ac);
  }
  free(scratch);
  return 0;
}

/* ---- complete distance: :872-1003 --------------------------------------- */

static void expand_image(const float* in, size_t w, size_t h, size_t xb,
                         size_t yb, size_t ew, size_t eh, float* out) {
  for (int c = 0; c < 3; ++c) {
    for (size_t y = 0; y < eh; ++y) {
      size_t sy = y > yb ? y - yb : 0;
      if (sy > h - 1) sy = h - 1;
      for (size_t x = 0; x < ew; ++x) {
        size_t sx = x > xb ? x - xb : 0;
        if (sx > w - 1) sx = w - 1;
        out[c * ew * eh + y * ew + x] = in[c * w * h + sy * w + sx];
      }
    }
  }
}

void oracle_butteraugli_subsample(const float* in, size_t w, size_t h, float* out) {
  const size_t sw = (w + 1) / 2, sh = (h + 1) / 2;
  for (int c = 0; c < 3; ++c) {
    float* o = out + c * sw * sh;
    for (size_t i = 0; i < sw * sh; ++i) o[i] = 0.0f;
    for (size_t y = 0; y < h; ++y)
      for (size_t x = 0; x < w; ++x) o[(y / 2) * sw + x / 2] += 0.25f * in[c * w * h + y * w + x];
    if (w & 1)
      for (size_t y = 0; y < sh; ++y) o[y * sw + sw - 1] *= 2.0f;
    if (h & 1)
      for (size_t x = 0; x < sw; ++x) o[(sh - 1) * sw + x] *= 2.0f;
  }
}

/* One scale: both images to psycho planes, then the difference stages. */
static int single_scale(const float* ref, const float* dist, size_t w, size_t h,
                        float hf_asymmetry, float x_multiplier, float intensity,
                        float* stages) {
  const size_t n = w * h;
  float* buf = malloc(sizeof(float) * n * 26);
  if (!buf) return -1;
  float* xyb0 = buf;
  float* xyb1 = buf + 3 * n;
  float* ps0 = buf + 6 * n;
  float* ps1 = buf + 16 * n;
  int rc = oracle_butteraugli_opsin(ref, w, h, intensity, xyb0);
  rc |= oracle_butteraugli_opsin(dist, w, h, intensity, xyb1);
  rc |= oracle_butteraugli_frequencies(xyb0, w, h, ps0);
  rc |= oracle_butteraugli_frequencies(xyb1, w, h, ps1);
  rc |= oracle_butteraugli_differences(ps0, ps1, w, h, hf_asymmetry, x_multiplier, stages);
  free(buf);
  return rc;
}

int oracle_butteraugli_distance(const float* ref, const float* dist, size_t w,
                                size_t h, float hf_asymmetry,
                                float x_multiplier, float intensity, float* map,
                                double* score) {
  const int expanded = w < 8 || h < 8;
  const size_t ew = w < 8 ? 8 : w, eh = h < 8 ? 8 : h;
  const size_t xb = w < 8 ? (8 - w) / 2 : 0, yb = h < 8 ? (8 - h) / 2 : 0;
  float* main_stages = malloc(sizeof(float) * 9 * ew * eh);
  if (!main_stages) return -1;
  int rc = 0;
  if (expanded) {
    float* e0 = malloc(sizeof(float) * 3 * ew * eh);
    float* e1 = malloc(sizeof(float) * 3 * ew * eh);
    expand_image(ref, w, h, xb, yb, ew, eh, e0);
    expand_image(dist, w, h, xb, yb, ew, eh, e1);
    rc = single_scale(e0, e1, ew, eh, hf_asymmetry, x_multiplier, intensity, main_stages);
    free(e0);
    free(e1);
    const float* fm = main_stages + 8 * ew * eh;
    for (size_t y = 0; y < h; ++y)
      for (size_t x = 0; x < w; ++x) map[y * w + x] = fm[(y + yb) * ew + x + xb];
  } else {
    rc = single_scale(ref, dist, w, h, hf_asymmetry, x_multiplier, intensity, main_stages);
    const float* fm = main_stages + 8 * w * h;
    if (w >= 15 && h >= 15) {
      const size_t sw = (w + 1) / 2, sh = (h + 1) / 2;
      float* s0 = malloc(sizeof(float) * 3 * sw * sh);
      float* s1 = malloc(sizeof(float) * 3 * sw * sh);
      float* sub = malloc(sizeof(float) * 9 * sw * sh);
      oracle_butteraugli_subsample(ref, w, h, s0);
      oracle_butteraugli_subsample(dist, w, h, s1);
      rc |= single_scale(s0, s1, sw, sh, hf_asymmetry, x_multiplier, intensity, sub);
      const float* sm = sub + 8 * sw * sh;
      for (size_t y = 0; y < h; ++y)
        for (size_t x = 0; x < w; ++x)
          map[y * w + x] = fm[y * w + x] * 0.85f + 0.5f * sm[(y / 2) * sw + x / 2];
      free(s0);
      free(s1);
      free(sub);
    } else {
      memcpy(map, fm, sizeof(float) * w * h);
    }
  }
  free(main_stages);
  float maximum = 0.0f;
  for (size_t i = 0; i < w * h; ++i) maximum = bmax(maximum, map[i]);
  *score = maximum;
  return rc;
}

/* ---- a17: ReduceButteraugliDistanceMap (src/codec/butteraugli.cpp:116-222) */
void oracle_reduce_distance_map(const float* map, size_t w, size_t h,
                                const uint8_t* cells, size_t bw, size_t bh,
                                float* out) {
  for (size_t by = 0; by < bh; ++by) {
    for (size_t bx = 0; bx < bw; ++bx) {
      const uint8_t c = cells[by * bw + bx];
      if (!(c & 1)) continue;
      int cw, ch;
      oracle_strategy_covered(c >> 1, &cw, &ch);
      const size_t x0 = bx * 8, y0 = by * 8;
      const size_t x1 = x0 + cw * 8 < w ? x0 + cw * 8 : w;
      const size_t y1 = y0 + ch * 8 < h ? y0 + ch * 8 : h;
      float norm = 0.0f;
      size_t count = 0;
      for (size_t y = y0; y < y1; ++y) {
        for (size_t x = x0; x < x1; ++x) {
          float v = map[y * w + x];
          v *= v;
          v *= v;
          v *= v;
          v *= v;
          norm += v;
          ++count;
        }
      }
      const float d = 1.2f * powf(norm / (float)count, 1.0f / 16.0f);
      for (int dy = 0; dy < ch; ++dy)
        for (int dx = 0; dx < cw; ++dx) out[(by + dy) * bw + bx + dx] = d;
    }
  }
}

/* ---- a18: one field update of RunAdaptiveQuantizationPolicyImpl
 * (src/codec/adaptive_quantization.cpp:1686-1733), after evaluation
 * `iteration` (< iterations).  scale / inverse_global_scale are the
 * evaluation's quantizer's. */
void oracle_aq_update(float* field, const float* adjusted_initial,
                      const float* block_distance, size_t n, float target,
                      int iteration, float lower, float upper, float scale,
                      float inverse_global_scale) {
  if (iteration == 1) {
    for (size_t i = 0; i < n; ++i) {
      const float clamp = 0.4f * field[i] + 0.6f * adjusted_initial[i];
      if (field[i] < clamp) field[i] = clamp < lower ? lower : (upper < clamp ? upper : clamp);
    }
  }
  const double power = iteration < 2 ? 0.2 : 0.0;
  for (size_t i = 0; i < n; ++i) {
    const float difference = block_distance[i] / target;
    if (difference <= 1.0f) {
      if (power != 0.0) field[i] *= (float)pow((double)difference, power);
    } else {
      const float old = field[i];
      field[i] *= difference;
      const long old_raw = lroundf(old * inverse_global_scale);
      const long new_raw = lroundf(field[i] * inverse_global_scale);
      if (old_raw == new_raw) field[i] = old + scale;
    }
    field[i] = field[i] < lower ? lower : (upper < field[i] ? upper : field[i]);
  }
}
