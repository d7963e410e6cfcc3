// Shared device pieces of the adaptive-quantization evaluation: the
// separable binary64 DCTs of one transform in LDS (one thread per row or
// column, basis reads uniform across the wave), DC <-> LLF conversion,
// dequantisation and the anchor layout of the AC groups.
#pragma once

#include "common.h"
#include "tables.h"

namespace gjxl_amd {

// Anchors: one entry per transform, listed per 256x256 AC group in raster
// order.  anchor_list[g * kGroupAnchors + k] = by << 16 | bx.
constexpr int kGroupBlocks = 32;
constexpr int kGroupAnchors = kGroupBlocks * kGroupBlocks;
constexpr int kGroupCoefficients = 65536;
constexpr uint32_t kNoAnchor = 0xffffffffu;

// Per-strategy anchor lists, in the stage order of the candidate search:
// DCT8, 16x8, 8x16, 16x16, 32x16, 16x32, 32x32.
constexpr int kStrategyStages = 7;
__host__ __device__ constexpr int StrategyStage(int strategy) {
  return strategy == kDct8 ? 0 : strategy == kDct16x8 ? 1 : strategy == kDct8x16 ? 2
       : strategy == kDct16x16 ? 3 : strategy == kDct32x16 ? 4
       : strategy == kDct16x32 ? 5 : 6;
}

// The strategy lists of a layout: entries[stage * capacity + k] = by << 16 |
// bx, counts[stage] on the host.
struct StrategyLists {
  const uint32_t* entries;
  size_t capacity;
  uint32_t counts[kStrategyStages];
};

// Re-targets the session to an extent whose buffers fit the ones it
// allocated; the session must be prepared again.  Clears the coefficient
// stores of the new extent and waits.
GJXLResult ResizeAqSession(GJXLAmdAqSession* s, uint32_t width, uint32_t height);
// The quantizing part of gjxl_amd_aq_evaluate (quantizer, final CfL,
// coefficient coding, DC) without reconstruction and loop filters; the
// session is validated and prepared.
GJXLResult AqQuantize(GJXLAmdAqSession* s, const float* quant_field);

// The split-precision forward DCT of every transform of a frame into the
// per-group coefficient store (split_dct.hip).
GJXLResult LaunchSplitAnchorDct(GJXLAmdDevice* d, const float* opsin, int width, int height,
                                const uint32_t* offsets, const StrategyLists& lists,
                                size_t store, float* coefs);

__device__ __forceinline__ float DStdMax(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float DStdMin(float a, float b) { return b < a ? b : a; }

// src/core/ac_strategy.h:95-108 for runtime extents.
__device__ __forceinline__ int CoefIndex(int W, int H, int v, int u) {
  return H < W ? v * W + u : u * H + v;
}

// Forward DCT of `channels` planes of W x H pixels held in LDS `pix`
// (row-major, channel stride N) into `coef` (libjxl layout), with the
// binary64 scratch `tmp` (channel stride N).  Sequential sums per output in
// the reference's order (src/codec/dct.cpp:135-182).
__device__ inline void ForwardDctLds(const double* __restrict__ bases, int W,
                                     int H, int channels, const float* pix,
                                     double* tmp, float* coef) {
  const int N = W * H;
  const double* fh = bases + BasisOffset(W, true);
  const double* fv = bases + BasisOffset(H, true);
  // horizontal: task = (channel, row y); outputs u = 0..W-1
  for (int task = threadIdx.x; task < channels * H; task += blockDim.x) {
    const int c = task / H, y = task % H;
    const float* row = pix + c * N + y * W;
    double* out = tmp + c * N + y * W;
    for (int u = 0; u < W; ++u) {
      double sum = 0.0;
      for (int x = 0; x < W; ++x) sum += static_cast<double>(row[x]) * fh[u * W + x];
      out[u] = sum;
    }
  }
  __syncthreads();
  // vertical: task = (channel, column u); outputs v = 0..H-1
  for (int task = threadIdx.x; task < channels * W; task += blockDim.x) {
    const int c = task / W, u = task % W;
    const double* col = tmp + c * N + u;
    float* out = coef + c * N;
    for (int v = 0; v < H; ++v) {
      double sum = 0.0;
      for (int y = 0; y < H; ++y) sum += fv[v * H + y] * col[y * W];
      out[CoefIndex(W, H, v, u)] = static_cast<float>(sum);
    }
  }
  __syncthreads();
}

// Inverse DCT (src/codec/dct.cpp:184-233): coef (libjxl layout) -> pix.
__device__ inline void InverseDctLds(const double* __restrict__ bases, int W,
                                     int H, int channels, const float* coef,
                                     double* tmp, float* pix) {
  const int N = W * H;
  const double* ih = bases + BasisOffset(W, false);
  const double* iv = bases + BasisOffset(H, false);
  for (int task = threadIdx.x; task < channels * H; task += blockDim.x) {
    const int c = task / H, v = task % H;
    const float* in = coef + c * N;
    double* out = tmp + c * N + v * W;
    for (int x = 0; x < W; ++x) {
      double sum = 0.0;
      for (int u = 0; u < W; ++u) {
        sum += static_cast<double>(in[CoefIndex(W, H, v, u)]) * ih[u * W + x];
      }
      out[x] = sum;
    }
  }
  __syncthreads();
  for (int task = threadIdx.x; task < channels * W; task += blockDim.x) {
    const int c = task / W, x = task % W;
    const double* col = tmp + c * N + x;
    float* out = pix + c * N;
    for (int y = 0; y < H; ++y) {
      double sum = 0.0;
      for (int v = 0; v < H; ++v) sum += iv[v * H + y] * col[v * W];
      out[y * W + x] = static_cast<float>(sum);
    }
  }
  __syncthreads();
}

// src/codec/dc_conversion.cpp: the small DCT bases of the DC <-> LLF
// conversion, evaluated on the device in binary64 (cos is only evaluated
// at the handful of (length, frequency, sample) points of lengths 1, 2, 4
// and is taken from a host-computed table).
struct DcBasisTable {
  // [forward?][length index 0:1,1:2,2:4][f][s]
  double v[2][3][4][4];
};

__device__ __forceinline__ int LengthIndex(int n) { return n == 1 ? 0 : n == 2 ? 1 : 2; }

__device__ __forceinline__ float DownsampleScale(int length, int f) {
  if (length == 1) return 1.0f;
  if (length == 2) return f == 0 ? 1.0f : 0.901764195028874394f;
  constexpr float d4[4] = {1.0f, 0.974886821136879522f, 0.901764195028874394f,
                           0.787054918159101335f};
  return d4[f];
}

__device__ __forceinline__ float UpsampleScale(int length, int f) {
  if (length == 1) return 1.0f;
  if (length == 2) return f == 0 ? 1.0f : 1.108937353592731823f;
  constexpr float u4[4] = {1.0f, 1.025760096781116015f, 1.108937353592731823f,
                           1.270559368765487251f};
  return u4[f];
}

__device__ __forceinline__ int SmallIndex(int w, int h, int v, int u) {
  return h < w ? v * w + u : u * h + v;
}

// ConvertLowFrequenciesToDc (src/codec/dc_conversion.cpp:170-230) by one
// thread: dc[cw*ch] in raster order.
__device__ inline void LlfToDc(const DcBasisTable& t, int cw, int ch,
                               const float* coef, float* dc) {
  const int W = cw * 8, H = ch * 8;
  double scaled[16];
  for (int v = 0; v < ch; ++v)
    for (int u = 0; u < cw; ++u)
      scaled[SmallIndex(cw, ch, v, u)] =
        static_cast<double>(coef[CoefIndex(W, H, v, u)]) *
        static_cast<double>(DownsampleScale(ch, v)) *
        static_cast<double>(DownsampleScale(cw, u));
  const int lh = LengthIndex(ch), lw = LengthIndex(cw);
  for (int y = 0; y < ch; ++y) {
    for (int x = 0; x < cw; ++x) {
      double value = 0.0;
      for (int v = 0; v < ch; ++v)
        for (int u = 0; u < cw; ++u)
          value += scaled[SmallIndex(cw, ch, v, u)] * t.v[0][lh][v][y] * t.v[0][lw][u][x];
      dc[y * cw + x] = static_cast<float>(value);
    }
  }
}

// ConvertDcToLowFrequencies (:252-320) by one thread.
__device__ inline void DcToLlf(const DcBasisTable& t, int cw, int ch,
                               const float* dc, float* coef) {
  const int W = cw * 8, H = ch * 8;
  const int lh = LengthIndex(ch), lw = LengthIndex(cw);
  float low[16];
  for (int v = 0; v < ch; ++v) {
    for (int u = 0; u < cw; ++u) {
      double value = 0.0;
      for (int y = 0; y < ch; ++y)
        for (int x = 0; x < cw; ++x)
          value += static_cast<double>(dc[y * cw + x]) * t.v[1][lh][v][y] * t.v[1][lw][u][x];
      low[SmallIndex(cw, ch, v, u)] = static_cast<float>(
        value * static_cast<double>(UpsampleScale(ch, v)) *
        static_cast<double>(UpsampleScale(cw, u)));
    }
  }
  for (int v = 0; v < ch; ++v)
    for (int u = 0; u < cw; ++u)
      coef[CoefIndex(W, H, v, u)] = low[SmallIndex(cw, ch, v, u)];
}

// AdjustQuantizationBias (src/codec/quantization.cpp:203-218).
__device__ __forceinline__ float QuantBias(int32_t q, int channel) {
  constexpr float kBias[4] = {1.0f - 0.05465007330715401f,
                              1.0f - 0.07005449891748593f,
                              1.0f - 0.049935103337343655f, 0.145f};
  const float v = static_cast<float>(q);
  const float a = fabsf(v);
  if (a < 0.125f) return 0.0f;
  if (a < 1.125f) return copysignf(kBias[channel], v);
  return v - kBias[3] / v;
}

// Quantizer state as the device sees it (core/quantizer.h).
struct QuantizerState {
  float scale;
  float inverse_global_scale;
  uint32_t global_scale;
  uint32_t quant_dc;
  float dc_steps[3];
  float inverse_dc_steps[3];
};

// ---- host launchers of the evaluation (aq.hip, filters.hip) ----

struct AqLayout {
  int bw, bh, groups_w, groups_h, groups;
  size_t store;
};
AqLayout MakeAqLayout(int width, int height);
GJXLResult LaunchAnchorLayout(hipStream_t s, const uint8_t* cells, int width,
                              int height, uint32_t* offsets, uint32_t* list,
                              uint32_t* count, uint32_t* error,
                              uint32_t* strategy_lists, size_t capacity,
                              uint32_t* strategy_counts);
GJXLResult LaunchAnchorDct(GJXLAmdDevice* d, const float* opsin, int width,
                           int height, const uint32_t* offsets,
                           const StrategyLists& lists, float* coefs);
GJXLResult LaunchAdjustField(hipStream_t s, const uint8_t* cells, int width,
                             int height, float target, const float* in,
                             float* out);
GJXLResult LaunchMatrixStats(hipStream_t s, const float* opsin, int width,
                             int height, int padded_w, int padded_h,
                             uint32_t* stats);
GJXLResult LaunchRange(hipStream_t s, const float* v, size_t n, uint32_t* range);
GJXLResult LaunchQuantizer(hipStream_t s, const float* field, size_t n,
                           float quant_dc, uint32_t* select_state,
                           uint32_t* hist, QuantizerState* q, int32_t* raw);
GJXLResult LaunchFinalCfl(GJXLAmdDevice* d, const float* coefs, int width,
                          int height, const uint8_t* cells,
                          const uint32_t* offsets, const int32_t* raw,
                          const QuantizerState* q, int8_t* y_to_x,
                          int8_t* y_to_b);
GJXLResult LaunchCodeCoefficients(GJXLAmdDevice* d, const float* coefs,
                                  int width, int height, const uint8_t* cells,
                                  const uint32_t* offsets,
                                  const StrategyLists& lists, int32_t* raw,
                                  const int8_t* y_to_x, const int8_t* y_to_b,
                                  const QuantizerState* q, float x_mult,
                                  float b_mult, int32_t* quantized, float* dc,
                                  int32_t* dc_quantized, float* dc_rec);
GJXLResult LaunchReconstruct(GJXLAmdDevice* d, const int32_t* quantized,
                             int width, int height, const uint8_t* cells,
                             const uint32_t* offsets, const StrategyLists& lists,
                             const int32_t* raw,
                             const int8_t* y_to_x, const int8_t* y_to_b,
                             const QuantizerState* q, float x_mult, float b_mult,
                             const float* dc_rec, float* smoothed,
                             float* inv_sigma, float* out);
GJXLResult LaunchLoopFiltersAndColor(hipStream_t s, const float* reconstructed,
                                     int pw, int ph, int w, int h,
                                     const float* inverse_sigma,
                                     float* filtered, float* linear,
                                     float intensity_target);


}  // namespace gjxl_amd
