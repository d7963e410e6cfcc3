/*
 * Oracle: one adaptive-quantization evaluation without Butteraugli --
 * quantizer from the field, final CfL, coefficient coding (adjusted raw
 * quant, AC thresholds, prediction-aware DC), EPF sigma, reconstruction,
 * Gaborish + EPF loop filters -- plus the field adjustment, policy bounds and
 * matrix-scale selection around it.  Test infrastructure only; see oracle.h.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static inline float std_max(float a, float b) { return a < b ? b : a; }
static inline float std_min(float a, float b) { return b < a ? b : a; }
static inline float std_clamp(float v, float lo, float hi) {
  return v < lo ? lo : (hi < v ? hi : v);
}

static size_t coefficient_index(int pw, int ph, int v, int u) {
  if (ph < pw) return (size_t)v * pw + u;
  return (size_t)u * ph + v;
}

/* ---- matrix scales: src/codestream/workflow.cpp:1318-1484 ---------------- */

void oracle_matrix_scale_stats(const float* opsin, size_t w, size_t h,
                               size_t stride, size_t plane, float stats[3]) {
  float x_edge = 0.0f, b_edge = 0.0f, exposed = 0.0f;
  const float* X = opsin;
  const float* Y = opsin + plane;
  const float* B = opsin + 2 * plane;
  for (size_t y = 1; y < h; ++y) {
    for (size_t x = 1; x < w; ++x) {
      const size_t i = y * stride + x, l = i - 1, u = i - stride;
      x_edge = std_max(x_edge, std_max(fabsf(X[i] - X[l]), fabsf(X[i] - X[u])));
      const float d = B[i] - Y[i];
      b_edge = std_max(b_edge, std_max(fabsf(d - (B[l] - Y[l])), fabsf(d - (B[u] - Y[u]))));
      float e = B[i] - 1.2f * Y[i];
      if (e >= 0.0f) {
        e *= fabsf(B[i] - B[l]) + fabsf(B[i] - B[u]);
        exposed = std_max(exposed, e);
      }
    }
  }
  stats[0] = x_edge;
  stats[1] = b_edge;
  stats[2] = exposed;
}

void oracle_select_matrix_scales(const float stats[3], float target,
                                 int* x_scale, int* b_scale) {
  int x = 3, b = 2;
  if (target > 2.5f) ++x;
  if (target > 5.5f) ++x;
  if (target > 9.5f) ++x;
  int xa = 0;
  if (stats[0] >= 0.026f) xa = 3;
  else if (stats[0] >= 0.022f) xa = 2;
  else if (stats[0] >= 0.015f) xa = 1;
  if (x < 2 + xa) x = 2 + xa;
  const int eb = stats[2] >= 0.13f ? 1 : 0;
  int ba = 0;
  if (stats[1] > 0.38f) ba = 2 + eb;
  else if (stats[1] > 0.33f) ba = 1 + eb;
  else if (stats[1] > 0.28f) ba = eb;
  *x_scale = x;
  *b_scale = b + ba;
}

/* src/codec/codestream.cpp:92-96 */
float oracle_matrix_multiplier(int scale) {
  return (float)pow(1.25, (double)(scale - 2));
}

/* ---- field adjustment and policy bounds --------------------------------- */

/* src/codec/adaptive_quantization.cpp:740-835 */
void oracle_adjust_quant_field(const uint8_t* cells, size_t bw, size_t bh,
                               float target, const float* in, float* out) {
  memcpy(out, in, sizeof(float) * bw * bh);
  float mixer = 1.0f;
  if (target > 1.54138f) mixer = std_max(0.0f, mixer - (target - 1.54138f) * 0.56391f);
  for (size_t y = 0; y < bh; ++y) {
    for (size_t x = 0; x < bw; ++x) {
      const uint8_t c = cells[y * bw + x];
      if (!(c & 1)) continue;
      int cw, ch;
      oracle_strategy_covered(c >> 1, &cw, &ch);
      float maximum = out[y * bw + x];
      float mean = 0.0f;
      for (int dy = 0; dy < ch; ++dy) {
        for (int dx = 0; dx < cw; ++dx) {
          const float v = out[(y + dy) * bw + x + dx];
          maximum = std_max(maximum, v);
          mean += v;
        }
      }
      const int count = cw * ch;
      mean /= (float)count;
      float result = maximum;
      if (count >= 4) result = maximum * mixer + mean * (1.0f - mixer);
      for (int dy = 0; dy < ch; ++dy)
        for (int dx = 0; dx < cw; ++dx) out[(y + dy) * bw + x + dx] = result;
    }
  }
}

/* src/codec/adaptive_quantization.cpp:1178-1242 */
int oracle_policy_setup(const float* adjusted, size_t n, float target,
                        float* lower, float* upper, float* quant_dc) {
  float mn = INFINITY, mx = 0.0f;
  for (size_t i = 0; i < n; ++i) {
    mn = std_min(mn, adjusted[i]);
    mx = std_max(mx, adjusted[i]);
  }
  const float ratio = mx / mn;
  const float max_dev = sqrtf(250.0f / ratio);
  const float asym = std_min(2.0f, max_dev);
  *lower = mn / (asym * max_dev);
  *upper = mx * (max_dev / asym);
  *quant_dc = oracle_initial_quant_dc(target);
  return (*upper / *lower >= 253.0f) ? 1 : 0;
}

/* ---- quantizer: src/codec/quantization.cpp:536-639, core/quantizer.h ----- */

static int cmp_float(const void* a, const void* b) {
  const float x = *(const float*)a, y = *(const float*)b;
  return (x > y) - (x < y);
}

/* Fills raw quant; q[0..5] = scale, inverse_global_scale, global_scale,
 * quant_dc (integer), dc_step[X], dc_step[Y] and q[6] = dc_step[B];
 * q[7..9] = inverse dc steps. */
int oracle_quantizer_from_field(float quant_dc, const float* field, size_t n,
                                int32_t* raw, float q[10]) {
  float* v = malloc(sizeof(float) * n);
  if (!v) return -1;
  memcpy(v, field, sizeof(float) * n);
  qsort(v, n, sizeof(float), cmp_float);
  const float median = v[n / 2];
  for (size_t i = 0; i < n; ++i) v[i] = fabsf(field[i] - median);
  qsort(v, n, sizeof(float), cmp_float);
  const float mad = v[n / 2];
  free(v);
  /* ComputeQuantizerParams (:59-115) */
  float scale = (float)65536u * (median - mad) / 5.0f;
  scale = std_clamp(scale, 1.0f, (float)(1u << 15));
  int32_t global_scale = (int32_t)scale;
  const int32_t scaled_dc = (int32_t)((double)(quant_dc * 4096.0f) * 1.6);
  if (global_scale > scaled_dc) global_scale = scaled_dc > 1 ? scaled_dc : 1;
  const float inv_gs = (float)65536u / (float)global_scale;
  const float qdc_value = std_min((float)(1u << 16), quant_dc * inv_gs + 0.5f);
  const uint32_t qdc = (uint32_t)qdc_value;
  /* Quantizer::Create */
  const float s = (float)(uint32_t)global_scale / (float)65536u;
  const float igs = (float)65536u / (float)(uint32_t)global_scale;
  const float inv_qdc = igs / (float)qdc;
  const float inv_dc_quant[3] = {4096.0f, 512.0f, 256.0f};
  q[0] = s;
  q[1] = igs;
  q[2] = (float)global_scale;
  q[3] = (float)qdc;
  for (int c = 0; c < 3; ++c) {
    q[4 + c] = inv_qdc / inv_dc_quant[c];
    q[7 + c] = inv_dc_quant[c] * s * (float)qdc;
  }
  for (size_t i = 0; i < n; ++i) {
    float r = field[i] * igs + 0.5f;
    r = std_clamp(r, 1.0f, 256.0f);
    raw[i] = (int32_t)r;
  }
  return 0;
}

/* ---- final CfL (fast): src/codec/chroma_from_luma.cpp:543-670 ------------ */

/* coefs: 3 planes of per-anchor forward coefficients stored in anchor order
 * at anchor_offset[] (anchors in raster order).  Uses the fast LS. */
void oracle_final_cfl_fast(const float* const coefs[3], const uint8_t* cells,
                           size_t bw, size_t bh, const size_t* anchor_offset,
                           const int32_t* raw, float scale, int8_t* y_to_x,
                           int8_t* y_to_b) {
  const size_t tw = (bw + 7) / 8, th = (bh + 7) / 8;
  float* v = malloc(sizeof(float) * 4 * 4096);
  for (size_t ty = 0; ty < th; ++ty) {
    for (size_t tx = 0; tx < tw; ++tx) {
      size_t n = 0;
      for (size_t by = ty * 8; by < ty * 8 + 8 && by < bh; ++by) {
        for (size_t bx = tx * 8; bx < tx * 8 + 8 && bx < bw; ++bx) {
          const uint8_t c = cells[by * bw + bx];
          if (!(c & 1)) continue;
          const int st = c >> 1;
          int cw, ch;
          oracle_strategy_covered(st, &cw, &ch);
          const int W = cw * 8, H = ch * 8, N = W * H;
          float dqx[1024], invx[1024], dqb[1024], invb[1024];
          oracle_quant_matrix(st, 0, dqx, invx);
          oracle_quant_matrix(st, 2, dqb, invb);
          const float qs = scale * 128.0f * (float)raw[by * bw + bx];
          const size_t off = anchor_offset[by * bw + bx];
          const int lw = W > H ? cw : ch, lh = W > H ? ch : cw; /* LLF extent */
          const int cwidth = W > H ? W : H;                      /* coef width */
          for (int i = 0; i < N; ++i) {
            const int xx = i % cwidth, yy = i / cwidth;
            const int llf = xx < lw && yy < lh;
            const float cy = llf ? 0.0f : coefs[1][off + i];
            const float cx = llf ? 0.0f : coefs[0][off + i];
            const float cb = llf ? 0.0f : coefs[2][off + i];
            v[0 * 4096 + n + i] = cy * invx[i] * qs;
            v[1 * 4096 + n + i] = cx * invx[i] * qs;
            v[2 * 4096 + n + i] = cy * invb[i] * qs;
            v[3 * 4096 + n + i] = cb * invb[i] * qs;
          }
          n += (size_t)N;
        }
      }
      y_to_x[ty * tw + tx] = oracle_find_best_multiplier(v, v + 4096, n, 0.0f, 1, 20);
      y_to_b[ty * tw + tx] =
        oracle_find_best_multiplier(v + 2 * 4096, v + 3 * 4096, n, 1.0f, 1, 20);
    }
  }
  free(v);
}

/* ---- coefficient coding ------------------------------------------------ */

static const float kQuantBias[4] = {
  1.0f - 0.05465007330715401f, 1.0f - 0.07005449891748593f,
  1.0f - 0.049935103337343655f, 0.145f,
};

/* src/codec/quantization.cpp:300-533 AdjustQuantForChannel.  Matrix layout:
 * rows = smaller dimension (height), cols = larger (width). */
static void adjust_quant_for_channel(float qscale, int channel, float mult,
                                     int strategy, int cols, int rows,
                                     const float* inv, const float* coef,
                                     float thr[4], int32_t* raw_quant) {
  const int xsize = cols / 8, ysize = rows / 8;
  const float qac = qscale * (float)*raw_quant;
  if (xsize > 1 || ysize > 1) {
    const float reduction = std_clamp(0.003f * (float)(xsize * ysize), 0.0f, 0.08f);
    for (int i = 0; i < 4; ++i) thr[i] = std_max(0.54f, thr[i] - reduction);
  }
  float border = 0.0f, err_sum = 0.0f, val_sum = 0.0f;
  float hf_nz[4] = {0, 0, 0, 0}, hf_max_err[4] = {0, 0, 0, 0};
  for (int y = 0; y < rows; ++y) {
    for (int x = 0; x < cols; ++x) {
      const int index = y * cols + x;
      if (x < xsize && y < ysize) continue;
      const int quadrant = (y >= rows / 2) * 2 + (x >= cols / 2);
      const float value = coef[index] * (inv[index] * qac * mult);
      const float quantized = fabsf(value) < thr[quadrant] ? 0.0f : rintf(value);
      const float error = fabsf(value - quantized);
      err_sum += error;
      val_sum += fabsf(quantized);
      if (channel == 1 && quantized == 0.0f) {
        hf_max_err[quadrant] = std_max(hf_max_err[quadrant], error);
      }
      if (quantized != 0.0f) {
        hf_nz[quadrant] += fabsf(quantized);
        const int in_corner = y >= 7 * ysize && x >= 7 * xsize;
        const int on_border = y + 1 == rows || x + 1 == cols;
        const int in_larger = x >= 4 * xsize && y >= 4 * ysize;
        if (in_corner || (on_border && in_larger)) border += fabsf(value);
      }
    }
  }
  if (channel == 1 && val_sum * 8.0f < (float)(xsize * ysize)) {
    const double limit = 0.46, mul = 0.9999;
    const int32_t original = *raw_quant;
    int32_t new_quant = *raw_quant;
    for (int q = 1; q < 4; ++q) {
      if (hf_nz[q] == 0.0f && hf_max_err[q] > limit) {
        new_quant = original + 1;
        break;
      }
    }
    *raw_quant = new_quant;
    if (hf_nz[3] == 0.0f && hf_max_err[3] > limit) {
      thr[3] = (float)(mul * hf_max_err[3] * new_quant / original);
    } else if ((hf_nz[1] == 0.0f && hf_max_err[1] > limit) ||
               (hf_nz[2] == 0.0f && hf_max_err[2] > limit)) {
      thr[1] = (float)(mul * std_max(hf_max_err[1], hf_max_err[2]) * new_quant / original);
      thr[2] = thr[1];
    } else if (hf_nz[0] == 0.0f && hf_max_err[0] > limit) {
      thr[0] = (float)(mul * hf_max_err[0] * new_quant / original);
    }
  }
  const float all_nz = hf_nz[0] + hf_nz[1] + hf_nz[2] + hf_nz[3] + 1.0f;
  const float border_mult[3] = {70.0f, 30.0f, 60.0f};
  if (border_mult[channel] * border >= all_nz) {
    *raw_quant += border_mult[channel] * border / all_nz;
    if (*raw_quant >= 256) *raw_quant = 255;
  }
  if (strategy == 0 && hf_nz[0] + hf_nz[1] + hf_nz[2] + hf_nz[3] < 11.0f) {
    ++*raw_quant;
    if (*raw_quant >= 256) *raw_quant = 255;
  }
  static const double first[4][3] = {
    {0.22080615753848404, 0.45797479824262011, 0.29859235095977965},
    {0.70109486510286834, 0.16185281305512639, 0.14387691730035473},
    {0.114985964456218638, 0.44656840441027695, 0.10587658215149048},
    {0.46849665264409396, 0.41239077937781954, 0.088667407767185444},
  };
  static const double second[4][3] = {
    {0.27450281941822197, 1.1255766549984996, 0.98950459134128388},
    {0.4652168675598285, 0.40945807983455818, 0.36581899811751367},
    {0.28034972424715715, 0.9182653201929738, 1.5581531543057416},
    {0.26873118114033728, 0.68863712390392484, 1.2082185408666786},
  };
  const double normalizer = 2.2942708343284721;
  err_sum *= normalizer;
  val_sum *= normalizer;
  if (strategy >= 4) {
    int cls = 3;
    if (strategy == 10 || strategy == 11) cls = 1;
    else if (strategy == 4) cls = 0;
    else if (strategy == 5) cls = 2;
    const double threshold = first[cls][channel] * xsize * ysize * 8 * 8 +
                             second[cls][channel] * val_sum;
    int32_t step = (int32_t)(err_sum / threshold);
    step = step < 0 ? 0 : (step > 2 ? 2 : step);
    if (err_sum > threshold) {
      *raw_quant += step;
      if (*raw_quant >= 256) *raw_quant = 255;
    }
  }
  const int32_t divisor = xsize * ysize;
  float min_nz = hf_nz[0];
  for (int q = 1; q < 4; ++q) if (hf_nz[q] < min_nz) min_nz = hf_nz[q];
  int32_t activity = 15;
  if (min_nz < 15.0f * divisor) activity = ((int32_t)min_nz + divisor / 2) / divisor;
  int32_t adjusted = *raw_quant - activity;
  if (channel == 1) for (int q = 1; q < 4; ++q) thr[q] += 0.01f * activity;
  const int32_t limit = (*raw_quant / 2) > 4 ? *raw_quant / 2 : 4;
  if (adjusted < limit) adjusted = limit;
  *raw_quant = adjusted;
}

/* src/codec/quantization.cpp:220-290 QuantizeAcBlockWithThresholds */
static void quantize_ac(float qscale, int32_t raw, float mult, int cols,
                        int rows, const float* inv, const float thr[4],
                        const float* coef, int32_t* out) {
  const float quantization_scale = qscale * (float)raw * mult;
  const int n = cols * rows;
  for (int i = 0; i < n; ++i) {
    const int x = i % cols, y = i / cols;
    const int t = (y >= rows / 2) * 2 + (x >= cols / 2);
    const float value = inv[i] * quantization_scale * coef[i];
    if (fabsf(value) < thr[t]) { out[i] = 0; continue; }
    out[i] = (int32_t)nearbyint((double)value);
  }
}

/* src/codec/quantization.cpp:183-218, :808-835 */
void oracle_dequantize_ac(float igs, int32_t raw, float mult, int channel,
                          int n, const float* dq, const int32_t* q, float* out) {
  const float scale = igs / (float)raw / mult;
  for (int i = 0; i < n; ++i) {
    const float v = (float)q[i];
    const float a = fabsf(v);
    float b;
    if (a < 0.125f) b = 0.0f;
    else if (a < 1.125f) b = copysignf(kQuantBias[channel], v);
    else b = v - kQuantBias[3] / v;
    out[i] = b * (dq[i] * scale);
  }
}

/* ---- DC conversion: src/codec/dc_conversion.cpp ------------------------- */

static float downsample_scale(int length, int f) {
  static const float d2[2] = {1.0f, 0.901764195028874394f};
  static const float d4[4] = {1.0f, 0.974886821136879522f, 0.901764195028874394f,
                              0.787054918159101335f};
  return length == 1 ? 1.0f : length == 2 ? d2[f] : d4[f];
}
static float upsample_scale(int length, int f) {
  static const float u2[2] = {1.0f, 1.108937353592731823f};
  static const float u4[4] = {1.0f, 1.025760096781116015f, 1.108937353592731823f,
                              1.270559368765487251f};
  return length == 1 ? 1.0f : length == 2 ? u2[f] : u4[f];
}
static double dc_basis(int length, int f, int s, int forward) {
  const double sqrt2 = 1.4142135623730951, pi = 3.141592653589793;
  const double alpha = f == 0 ? sqrt2 / 2.0 : 1.0;
  const double angle = ((double)s + 0.5) * (double)f * pi / (double)length;
  return forward ? sqrt2 * alpha * cos(angle) / (double)length
                 : sqrt2 * alpha * cos(angle);
}
static size_t small_index(int w, int h, int v, int u) {
  if (h < w) return (size_t)v * w + u;
  return (size_t)u * h + v;
}

/* :170-230 ConvertLowFrequenciesToDc: dc[cw*ch] in raster order. */
void oracle_llf_to_dc(int cw, int ch, const float* coef, float* dc) {
  const int W = cw * 8, H = ch * 8;
  double scaled[16];
  for (int v = 0; v < ch; ++v)
    for (int u = 0; u < cw; ++u)
      scaled[small_index(cw, ch, v, u)] =
        (double)coef[coefficient_index(W, H, v, u)] *
        (double)downsample_scale(ch, v) * (double)downsample_scale(cw, u);
  for (int y = 0; y < ch; ++y) {
    for (int x = 0; x < cw; ++x) {
      double value = 0.0;
      for (int v = 0; v < ch; ++v)
        for (int u = 0; u < cw; ++u)
          value += scaled[small_index(cw, ch, v, u)] * dc_basis(ch, v, y, 0) *
                   dc_basis(cw, u, x, 0);
      dc[y * cw + x] = (float)value;
    }
  }
}

/* :252-320 ConvertDcToLowFrequencies */
void oracle_dc_to_llf(int cw, int ch, const float* dc, float* coef) {
  const int W = cw * 8, H = ch * 8;
  float low[16];
  for (int v = 0; v < ch; ++v) {
    for (int u = 0; u < cw; ++u) {
      double value = 0.0;
      for (int y = 0; y < ch; ++y)
        for (int x = 0; x < cw; ++x)
          value += (double)dc[y * cw + x] * dc_basis(ch, v, y, 1) * dc_basis(cw, u, x, 1);
      low[small_index(cw, ch, v, u)] =
        (float)(value * (double)upsample_scale(ch, v) * (double)upsample_scale(cw, u));
    }
  }
  for (int v = 0; v < ch; ++v)
    for (int u = 0; u < cw; ++u)
      coef[coefficient_index(W, H, v, u)] = low[small_index(cw, ch, v, u)];
}

/* ---- prediction-aware DC: src/codec/vardct/dc_quantization.cpp ---------- */
/* Weighted predictor: src/codec/modular/prediction.h:59-153 (libjxl's
 * self-correcting predictor with its default parameters). */
typedef struct {
  size_t width;
  int64_t predictions[4];
  int64_t prediction;
  uint32_t* pe[4];
  int32_t* err;
} Weighted;

static const uint8_t kWpCoeff[7] = {16, 10, 7, 7, 7, 0, 0};
static const uint8_t kWpWeights[4] = {13, 12, 12, 12};

static uint32_t reciprocal(size_t i) { return (1u << 24) / (uint32_t)(i + 1); }
static int bit_width64(uint64_t v) { int b = 0; while (v) { ++b; v >>= 1; } return b; }

static int64_t wp_predict(Weighted* s, size_t x, size_t y, int64_t n, int64_t w,
                          int64_t ne, int64_t nw, int64_t nn) {
  const size_t W = s->width;
  const size_t current = (y & 1) ? 0 : W + 2;
  const size_t previous = (y & 1) ? W + 2 : 0;
  const size_t north = previous + x;
  const size_t northeast = x + 1 < W ? north + 1 : north;
  const size_t northwest = x ? north - 1 : north;
  uint32_t weights[4];
  uint32_t weight_sum = 0;
  for (int i = 0; i < 4; ++i) {
    const uint32_t e = s->pe[i][north] + s->pe[i][northeast] + s->pe[i][northwest];
    int shift = bit_width64((uint64_t)e + 1) - 6;
    if (shift < 0) shift = 0;
    weights[i] = 4 + ((kWpWeights[i] * reciprocal(e >> shift)) >> shift);
    weight_sum += weights[i];
  }
  const unsigned log_weight = (unsigned)bit_width64(weight_sum) - 1;
  weight_sum = 0;
  for (int i = 0; i < 4; ++i) {
    weights[i] >>= log_weight - 4;
    weight_sum += weights[i];
  }
  const int64_t ew = x ? s->err[current + x - 1] : 0;
  const int64_t en = s->err[north], enw = s->err[northwest], ene = s->err[northeast];
  n *= 8; w *= 8; ne *= 8; nw *= 8; nn *= 8;
  s->predictions[0] = w + ne - n;
  s->predictions[1] = n - (((en + ew + ene) * kWpCoeff[0]) >> 5);
  s->predictions[2] = w - (((en + ew + enw) * kWpCoeff[1]) >> 5);
  s->predictions[3] = n - ((enw * kWpCoeff[2] + en * kWpCoeff[3] + ene * kWpCoeff[4] +
                            (nn - n) * kWpCoeff[5] + (nw - w) * kWpCoeff[6]) >> 5);
  int64_t sum = (int64_t)(uint32_t)((weight_sum >> 1) - 1);
  for (int i = 0; i < 4; ++i) sum += s->predictions[i] * (int64_t)weights[i];
  s->prediction = (sum * (int64_t)reciprocal(weight_sum - 1)) >> 24;
  if (((en ^ ew) | (en ^ enw)) <= 0) {
    int64_t lo = w < ne ? w : ne; lo = lo < n ? lo : n;
    int64_t hi = w > ne ? w : ne; hi = hi > n ? hi : n;
    if (s->prediction < lo) s->prediction = lo;
    if (s->prediction > hi) s->prediction = hi;
  }
  return (s->prediction + 3) >> 3;
}

static void wp_update(Weighted* s, int32_t value, size_t x, size_t y) {
  const size_t W = s->width;
  const size_t current = (y & 1) ? 0 : W + 2;
  const size_t previous = (y & 1) ? W + 2 : 0;
  const int64_t scaled = (int64_t)value * 8;
  const int64_t error = s->prediction - scaled;
  uint32_t errors[4];
  for (int i = 0; i < 4; ++i) {
    int64_t d = s->predictions[i] - scaled;
    if (d < 0) d = -d;
    errors[i] = (uint32_t)((uint64_t)(d + 3) >> 3);
  }
  s->err[current + x] = (int32_t)error;
  for (int i = 0; i < 4; ++i) {
    s->pe[i][current + x] = errors[i];
    s->pe[i][previous + x + 1] += errors[i];
  }
}

static float round_half_away(float v) { return roundf(v); }

/* :121-240 (prediction-aware branch, weighted predictor, extra precision
 * 1).  dc: 3 planes bw*bh (X, Y, B) in; quantized and reconstructed out. */
void oracle_quantize_dc(const float* dc, size_t bw, size_t bh,
                        const float dc_steps[3], const float inv_dc_steps[3],
                        int extra_precision, int32_t* quantized,
                        float* reconstructed) {
  const size_t n = bw * bh;
  float inv[3], steps[3];
  const float precision = (float)(1u << extra_precision);
  for (int c = 0; c < 3; ++c) {
    inv[c] = inv_dc_steps[c] * precision;
    steps[c] = dc_steps[c] / precision;
  }
  for (size_t gy = 0; gy < bh; gy += 256) {
    const size_t h = bh - gy < 256 ? bh - gy : 256;
    for (size_t gx = 0; gx < bw; gx += 256) {
      const size_t w = bw - gx < 256 ? bw - gx : 256;
      Weighted s;
      s.width = w;
      for (int i = 0; i < 4; ++i) s.pe[i] = calloc((w + 2) * 2, sizeof(uint32_t));
      s.err = calloc((w + 2) * 2, sizeof(int32_t));
      const int order[3] = {1, 0, 2};
      for (int oi = 0; oi < 3; ++oi) {
        const int c = order[oi];
        for (int i = 0; i < 4; ++i) memset(s.pe[i], 0, (w + 2) * 2 * sizeof(uint32_t));
        memset(s.err, 0, (w + 2) * 2 * sizeof(int32_t));
        memset(s.predictions, 0, sizeof(s.predictions));
        s.prediction = 0;
        int32_t* plane = quantized + c * n;
        for (size_t y = 0; y < h; ++y) {
          for (size_t x = 0; x < w; ++x) {
            const size_t index = (gy + y) * bw + gx + x;
            const int64_t W_ = x ? plane[index - 1] : y ? plane[index - bw] : 0;
            const int64_t N_ = y ? plane[index - bw] : W_;
            const int64_t NW_ = x && y ? plane[index - bw - 1] : W_;
            const int64_t NE_ = y && x + 1 < w ? plane[index - bw + 1] : N_;
            const int64_t NN_ = y > 1 ? plane[index - 2 * bw] : N_;
            const int64_t prediction = wp_predict(&s, x, y, N_, W_, NE_, NW_, NN_);
            float value = dc[c * n + index];
            if (c == 2) value = fmaf(-(float)quantized[1 * n + index], steps[1], value);
            float residual = value * inv[c];
            residual -= (float)prediction;
            if (residual > -0.62f && residual < 0.62f) residual = 0;
            float rounded = round_half_away(residual);
            if (rounded > 2 || rounded < -2) rounded = round_half_away(residual * 0.5f) * 2;
            const int32_t ir = (int32_t)(double)round_half_away(rounded);
            plane[index] = (int32_t)(prediction + ir);
            wp_update(&s, plane[index], x, y);
          }
        }
      }
      for (int i = 0; i < 4; ++i) free(s.pe[i]);
      free(s.err);
    }
  }
  for (size_t i = 0; i < n; ++i) {
    const float ry = (float)quantized[1 * n + i] * steps[1];
    for (int c = 0; c < 3; ++c) {
      float r = (float)quantized[c * n + i] * steps[c];
      if (c == 2) r = fmaf(ry, 1.0f, r);
      reconstructed[c * n + i] = r;
    }
  }
}

/* ---- DC smoothing: src/codec/dc_smoothing.cpp --------------------------- */

void oracle_smooth_dc(const float* dc, size_t bw, size_t bh,
                      const float dc_steps[3], float* out) {
  const size_t n = bw * bh;
  memcpy(out, dc, sizeof(float) * 3 * n);
  const float side = 0.20345139757231578f, corner = 0.0334829185968739f;
  const float center_w = 1.0f - 4.0f * (side + corner);
  for (size_t y = 1; y + 1 < bh; ++y) {
    for (size_t x = 1; x + 1 < bw; ++x) {
      float center[3], smoothed[3];
      float gap = 0.5f;
      for (int c = 0; c < 3; ++c) {
        const float* top = dc + c * n + (y - 1) * bw;
        const float* row = dc + c * n + y * bw;
        const float* bottom = dc + c * n + (y + 1) * bw;
        const float corners = (top[x - 1] + top[x + 1]) + (bottom[x - 1] + bottom[x + 1]);
        const float sides = (row[x - 1] + row[x + 1]) + (top[x] + bottom[x]);
        center[c] = row[x];
        smoothed[c] = fmaf(corners, corner, fmaf(sides, side, row[x] * center_w));
        gap = std_max(gap, fabsf((row[x] - smoothed[c]) / dc_steps[c]));
      }
      const float factor = std_max(0.0f, fmaf(-4.0f, gap, 3.0f));
      for (int c = 0; c < 3; ++c) {
        out[c * n + y * bw + x] = fmaf(smoothed[c] - center[c], factor, center[c]);
      }
    }
  }
}

/* ---- anchors and AC groups: src/codec/reconstruction.cpp:386-760 ---------- */

/* Fills, for every anchor cell, the offset of its coefficients inside its
 * 256x256 AC group (anchors of a group in raster order), and returns the
 * per-anchor offset into a dense per-image coefficient store laid out group
 * by group (group g starts at g * 65536). */
void oracle_anchor_offsets(const uint8_t* cells, size_t bw, size_t bh,
                           size_t* offsets) {
  const size_t gw = (bw + 31) / 32;
  const size_t gh = (bh + 31) / 32;
  size_t* used = calloc(gw * gh, sizeof(size_t));
  for (size_t y = 0; y < bh; ++y) {
    for (size_t x = 0; x < bw; ++x) {
      const uint8_t c = cells[y * bw + x];
      offsets[y * bw + x] = (size_t)-1;
      if (!(c & 1)) continue;
      int cw, ch;
      oracle_strategy_covered(c >> 1, &cw, &ch);
      const size_t g = (y / 32) * gw + x / 32;
      offsets[y * bw + x] = g * 65536 + used[g];
      used[g] += (size_t)cw * ch * 64;
    }
  }
  free(used);
}

/* Forward DCT of every anchor into 3 planes of gw*gh*65536 floats. */
void oracle_anchor_coefficients(const float* opsin, size_t w, size_t h,
                                const uint8_t* cells, const size_t* offsets,
                                float* coefs) {
  const size_t bw = w / 8, bh = h / 8, plane = w * h;
  const size_t store = ((bw + 31) / 32) * ((bh + 31) / 32) * 65536;
  float pix[1024];
  for (size_t y = 0; y < bh; ++y) {
    for (size_t x = 0; x < bw; ++x) {
      const uint8_t c = cells[y * bw + x];
      if (!(c & 1)) continue;
      int cw, ch;
      oracle_strategy_covered(c >> 1, &cw, &ch);
      const int W = cw * 8, H = ch * 8;
      for (int k = 0; k < 3; ++k) {
        for (int r = 0; r < H; ++r)
          memcpy(pix + r * W, opsin + k * plane + (y * 8 + r) * w + x * 8, sizeof(float) * W);
        oracle_forward_dct(c >> 1, pix, coefs + k * store + offsets[y * bw + x]);
      }
    }
  }
}

/* One transform of ComputeQuantizedCoefficients (src/codec/reconstruction.cpp:
 * 565-700): adjusted shared raw quant (quantization.cpp:741-806), Y dead
 * zone, Y round trip, CfL removal from X/B, X/B quantisation, and the
 * per-block DC values of the three channels.  coef: 3*n, modified in place. */
void oracle_code_transform(int st, const float q[10], int32_t initial,
                           float x_mult, float b_mult, float fx, float fb,
                           float* coef, int32_t* quantized, float* dc,
                           int32_t* decided_out) {
  int cw, ch;
  oracle_strategy_covered(st, &cw, &ch);
  const int W = cw * 8, H = ch * 8, N = W * H;
  const int cols = W > H ? W : H, rows = W > H ? H : W;
  const float mults[3] = {x_mult, 1.0f, b_mult};
  float dq[3][1024], inv[3][1024];
  for (int k = 0; k < 3; ++k) oracle_quant_matrix(st, k, dq[k], inv[k]);
  float* c0 = coef;
  float* c1 = coef + N;
  float* c2 = coef + 2 * N;
  float* cs[3] = {c0, c1, c2};
  int32_t decided = 0;
  float y_thr[4] = {0.58f, 0.64f, 0.64f, 0.64f};
  const int order[3] = {1, 0, 2};
  for (int oi = 0; oi < 3; ++oi) {
    const int k = order[oi];
    float thr[4] = {0.58f, 0.64f, 0.64f, 0.64f};
    int32_t cand = initial;
    adjust_quant_for_channel(q[0], k, mults[k], st, cols, rows, inv[k], cs[k],
                             thr, &cand);
    if (k == 1) memcpy(y_thr, thr, sizeof(thr));
    if (cand > decided) decided = cand;
  }
  *decided_out = decided;
  oracle_llf_to_dc(cw, ch, c1, dc + 16);
  quantize_ac(q[0], decided, 1.0f, cols, rows, inv[1], y_thr, c1, quantized + N);
  oracle_dequantize_ac(q[1], decided, 1.0f, 1, N, dq[1], quantized + N, c1);
  for (int i = 0; i < N; ++i) {
    c0[i] -= fx * c1[i];
    c2[i] -= fb * c1[i];
  }
  const int covered = cw * ch;
  for (int k = 0; k < 3; k += 2) {
    oracle_llf_to_dc(cw, ch, cs[k], dc + 16 * k);
    float thr[4] = {0.58f, 0.62f, 0.62f, 0.62f};
    if (covered >= 4)
      for (int t = 0; t < 4; ++t) thr[t] = std_max(0.5f, thr[t] - 0.00744f * (float)covered);
    quantize_ac(q[0], decided, mults[k], cols, rows, inv[k], thr, cs[k],
                quantized + k * N);
  }
}

/* ComputeQuantizedCoefficients (:386-760) with the adjusted shared quant,
 * prediction-aware weighted DC with one extra precision bit.  q: quantizer
 * as filled by oracle_quantizer_from_field.  raw: in = field raw quant, out
 * = adjusted at anchors.  quantized: 3 planes of gw*gh*65536 int32 in the
 * dense per-group layout.  dc_q, dc_rec: 3 planes bw*bh. */
void oracle_code_coefficients(const float* coefs, const uint8_t* cells,
                              size_t bw, size_t bh, const size_t* offsets,
                              const float q[10], int32_t* raw,
                              const int8_t* y_to_x, const int8_t* y_to_b,
                              float x_mult, float b_mult, int32_t* quantized,
                              int32_t* dc_q, float* dc_rec) {
  const size_t n = bw * bh, tw = (bw + 7) / 8;
  const size_t store = ((bw + 31) / 32) * ((bh + 31) / 32) * 65536;
  float* dc = calloc(3 * n, sizeof(float));
  for (size_t y = 0; y < bh; ++y) {
    for (size_t x = 0; x < bw; ++x) {
      const uint8_t cell = cells[y * bw + x];
      if (!(cell & 1)) continue;
      const int st = cell >> 1;
      int cw, ch;
      oracle_strategy_covered(st, &cw, &ch);
      const int N = cw * ch * 64;
      const size_t off = offsets[y * bw + x];
      float coef[3 * 1024];
      for (int k = 0; k < 3; ++k) memcpy(coef + k * N, coefs + k * store + off, sizeof(float) * N);
      const size_t tile = (y / 8) * tw + x / 8;
      const float fx = (float)y_to_x[tile] / 84.0f;
      const float fb = 1.0f + (float)y_to_b[tile] / 84.0f;
      int32_t qt[3 * 1024];
      float dcv[3][16];
      int32_t decided;
      oracle_code_transform(st, q, raw[y * bw + x], x_mult, b_mult, fx, fb,
                            coef, qt, &dcv[0][0], &decided);
      raw[y * bw + x] = decided;
      for (int k = 0; k < 3; ++k) {
        memcpy(quantized + k * store + off, qt + k * N, sizeof(int32_t) * N);
        for (int dy = 0; dy < ch; ++dy)
          for (int dx = 0; dx < cw; ++dx)
            dc[k * n + (y + dy) * bw + x + dx] = dcv[k][dy * cw + dx];
      }
    }
  }
  oracle_quantize_dc(dc, bw, bh, q + 4, q + 7, 1, dc_q, dc_rec);
  free(dc);
}

/* ComputeEpfInverseSigma (src/codec/epf.cpp:207-290) with neutral
 * sharpness 4 everywhere (FillDefaultEpfSharpness). */
void oracle_epf_inverse_sigma(const uint8_t* cells, size_t bw, size_t bh,
                              const int32_t* raw, float scale, float* out) {
  const float k_inv_sigma_numerator = -1.1715728752538099024f;
  const float lut4 = 4.0f / 7.0f;
  for (size_t y = 0; y < bh; ++y) {
    for (size_t x = 0; x < bw; ++x) {
      const uint8_t c = cells[y * bw + x];
      if (!(c & 1)) continue;
      int cw, ch;
      oracle_strategy_covered(c >> 1, &cw, &ch);
      const float sigma_quant =
        0.46f / (scale * (float)raw[y * bw + x] * k_inv_sigma_numerator);
      for (int dy = 0; dy < ch; ++dy) {
        for (int dx = 0; dx < cw; ++dx) {
          float sigma = sigma_quant * lut4;
          sigma = std_min(-1.0e-4f, sigma);
          out[(y + dy) * bw + x + dx] = 1.0f / sigma;
        }
      }
    }
  }
}

/* ReconstructQuantizedCoefficients (src/codec/reconstruction.cpp:790-936)
 * with adaptive DC smoothing: padded XYB image out (3 planes w*h). */
void oracle_reconstruct(const int32_t* quantized, const uint8_t* cells,
                        size_t w, size_t h, const size_t* offsets,
                        const float q[10], const int32_t* raw,
                        const int8_t* y_to_x, const int8_t* y_to_b,
                        float x_mult, float b_mult, const float* dc_rec,
                        float* out) {
  const size_t bw = w / 8, bh = h / 8, n = bw * bh, tw = (bw + 7) / 8;
  const size_t store = ((bw + 31) / 32) * ((bh + 31) / 32) * 65536;
  const size_t plane = w * h;
  float* smooth = malloc(sizeof(float) * 3 * n);
  oracle_smooth_dc(dc_rec, bw, bh, q + 4, smooth);
  const float mults[3] = {x_mult, 1.0f, b_mult};
  for (size_t y = 0; y < bh; ++y) {
    for (size_t x = 0; x < bw; ++x) {
      const uint8_t cell = cells[y * bw + x];
      if (!(cell & 1)) continue;
      const int st = cell >> 1;
      int cw, ch;
      oracle_strategy_covered(st, &cw, &ch);
      const int W = cw * 8, H = ch * 8, N = W * H;
      const size_t off = offsets[y * bw + x];
      const int32_t r = raw[y * bw + x];
      float coef[3][1024];
      for (int k = 0; k < 3; ++k) {
        float dq[1024], inv[1024];
        oracle_quant_matrix(st, k, dq, inv);
        oracle_dequantize_ac(q[1], r, mults[k], k, N, dq, quantized + k * store + off, coef[k]);
      }
      const size_t tile = (y / 8) * tw + x / 8;
      const float fx = (float)y_to_x[tile] / 84.0f;
      const float fb = 1.0f + (float)y_to_b[tile] / 84.0f;
      for (int i = 0; i < N; ++i) {
        coef[0][i] += fx * coef[1][i];
        coef[2][i] += fb * coef[1][i];
      }
      for (int k = 0; k < 3; ++k) {
        float dcv[16], pix[1024];
        for (int dy = 0; dy < ch; ++dy)
          for (int dx = 0; dx < cw; ++dx) dcv[dy * cw + dx] = smooth[k * n + (y + dy) * bw + x + dx];
        oracle_dc_to_llf(cw, ch, dcv, coef[k]);
        oracle_inverse_dct(st, coef[k], pix);
        for (int rr = 0; rr < H; ++rr)
          memcpy(out + k * plane + (y * 8 + rr) * w + x * 8, pix + rr * W, sizeof(float) * W);
      }
    }
  }
  free(smooth);
}

/* ---- loop filters: src/codec/gaborish.cpp:101-190, src/codec/epf.cpp ---- */

/* ApplyGaborish on an unpadded image (3 planes w*h), default weights. */
void oracle_gaborish(const float* in, size_t w, size_t h, float* out) {
  const float w1 = 1.1f * 0.104699568f, w2 = 1.1f * 0.055680538f;
  const float divisor = 1.0f + 4.0f * (w1 + w2);
  const float cw = 1.0f / divisor, aw = w1 / divisor, dw = w2 / divisor;
  for (int c = 0; c < 3; ++c) {
    const float* p = in + c * w * h;
    for (size_t y = 0; y < h; ++y) {
      for (size_t x = 0; x < w; ++x) {
        const long long sx = (long long)x, sy = (long long)y;
#define S(dx, dy) p[oracle_mirror(sy + (dy), h) * w + oracle_mirror(sx + (dx), w)]
        const float axes = (S(-1, 0) + S(1, 0)) + (S(0, -1) + S(0, 1));
        const float diag = (S(-1, -1) + S(1, -1)) + (S(-1, 1) + S(1, 1));
        out[c * w * h + y * w + x] = cw * S(0, 0) + aw * axes + dw * diag;
#undef S
      }
    }
  }
}

static float epf_sample(const float* p, size_t w, size_t h, long long x, long long y) {
  return p[oracle_mirror(y, h) * w + oracle_mirror(x, w)];
}

/* ApplyEpfPass (src/codec/epf.cpp:61-205), pass 1 or 2. */
static void epf_pass(const float* in, size_t w, size_t h, const float* inv_sigma,
                     int pass, float* out) {
  const float channel_scale[3] = {40.0f, 5.0f, 3.5f};
  const float min_inv_sigma = -1.1715728752538099024f / 0.3f;
  const float sigma_scale = 1.65f * (pass == 2 ? 6.5f : 1.0f);
  const size_t plane = w * h, bw = (w + 7) / 8;
  static const int card[4][2] = {{0, -1}, {-1, 0}, {1, 0}, {0, 1}};
  static const int plus[5][2] = {{0, 0}, {0, -1}, {-1, 0}, {0, 1}, {1, 0}};
  for (size_t y = 0; y < h; ++y) {
    for (size_t x = 0; x < w; ++x) {
      const float bis = inv_sigma[(y / 8) * bw + x / 8];
      if (bis < min_inv_sigma) {
        for (int c = 0; c < 3; ++c) out[c * plane + y * w + x] = in[c * plane + y * w + x];
        continue;
      }
      const int border = x % 8 == 0 || x % 8 == 7 || y % 8 == 0 || y % 8 == 7;
      const float sis = bis * sigma_scale * (border ? 2.0f / 3.0f : 1.0f);
      const long long sx = (long long)x, sy = (long long)y;
      float sum[3] = {in[y * w + x], in[plane + y * w + x], in[2 * plane + y * w + x]};
      float wsum = 1.0f;
      for (int k = 0; k < 4; ++k) {
        const int dx = card[k][0], dy = card[k][1];
        float sad = 0.0f;
        for (int c = 0; c < 3; ++c) {
          const float* p = in + c * plane;
          if (pass == 1) {
            float cs = 0.0f;
            for (int o = 0; o < 5; ++o) {
              cs += fabsf(epf_sample(p, w, h, sx + plus[o][0], sy + plus[o][1]) -
                          epf_sample(p, w, h, sx + dx + plus[o][0], sy + dy + plus[o][1]));
            }
            sad = fmaf(cs, channel_scale[c], sad);
          } else {
            sad = fmaf(fabsf(epf_sample(p, w, h, sx, sy) -
                             epf_sample(p, w, h, sx + dx, sy + dy)),
                       channel_scale[c], sad);
          }
        }
        const float weight = std_max(0.0f, fmaf(sad, sis, 1.0f));
        wsum += weight;
        for (int c = 0; c < 3; ++c)
          sum[c] = fmaf(weight, epf_sample(in + c * plane, w, h, sx + dx, sy + dy), sum[c]);
      }
      for (int c = 0; c < 3; ++c) out[c * plane + y * w + x] = sum[c] / wsum;
    }
  }
}

/* ApplyLoopFilters (src/codec/loop_filter.cpp:14-65): Gaborish then EPF with
 * two iterations (passes 1, 2) on the unpadded image. */
void oracle_loop_filters(const float* in, size_t w, size_t h,
                         const float* inv_sigma, float* out) {
  float* a = malloc(sizeof(float) * 3 * w * h);
  float* b = malloc(sizeof(float) * 3 * w * h);
  oracle_gaborish(in, w, h, a);
  epf_pass(a, w, h, inv_sigma, 1, b);
  epf_pass(b, w, h, inv_sigma, 2, out);
  free(a);
  free(b);
}
