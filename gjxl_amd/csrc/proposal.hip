// The learned AC-strategy proposal: a chosen form beside the plain path,
// with no counterpart in the reference.  A six-layer 3x3 convolutional
// network over the 8x8-block grid predicts the plain search's strategy per
// block (src/codec/ac_strategy.cpp:1101 FindAcStrategyGrid) and the log
// ratio of the Butteraugli loop's final quant field to the adjusted initial
// field (src/codec/adaptive_quantization.cpp:1600-1738).
//
// Kernels:
//   FeaturesKernel     16 bfloat16 features per block (means, DCT8 band
//                      energies, log initial field)
//   ConvKernel         one 3x3 convolution on v_mfma_f32_16x16x16_bf16: the
//                      forward (bias, ReLU, bfloat16 or binary32 out) and
//                      the input gradient (ReLU mask on the incoming
//                      gradient, transposed tap-flipped weights)
//   WeightGradKernel   weight and bias gradients of one layer, partial sums
//                      per work group on MFMA, then ReduceKernel
//   LossKernel         softmax cross-entropy and field error per position,
//                      output gradient in bfloat16
//   AdamKernel         Adam on the binary32 master parameters, repacking
//                      the bfloat16 forward and gradient weights
//   DecideKernel       allowed stages per block
//   WarmStartKernel    the loop's starting field
//   SampleKernel       labels and targets of a plain run
//   Norm3*Kernel       the 3-norm of the block distances (the judge)
//
// MFMA operands (16x16x16, 64 lanes): lane l holds A[l % 16][4 (l / 16) + i],
// B[4 (l / 16) + i][l % 16] and D[4 (l / 16) + i][l % 16], i = 0..3.  In the
// convolutions A is the weights (output channel x input channel of one tap),
// B the inputs (input channel x 16 positions of one row) and D the outputs
// (output channel x position): each lane ends with four consecutive output
// channels of one position, stored as one 8- or 16-byte write.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "proposal.h"

namespace gjxl_amd {
namespace {

typedef short Bf16x4 __attribute__((ext_vector_type(4)));
typedef float Float4 __attribute__((ext_vector_type(4)));

constexpr int kFeatures = GJXL_AMD_PROPOSAL_FEATURES;
constexpr int kChannels = GJXL_AMD_PROPOSAL_CHANNELS;
constexpr int kLayers = GJXL_AMD_PROPOSAL_LAYERS;
constexpr int kClasses = GJXL_AMD_PROPOSAL_CLASSES;
constexpr int kFieldChannel = GJXL_AMD_PROPOSAL_FIELD_CHANNEL;
constexpr int kTaps = 9;
constexpr int kIgnore = GJXL_AMD_PROPOSAL_IGNORE;

__host__ __device__ constexpr int LayerInputs(int layer) {
  return layer == 0 ? kFeatures : kChannels;
}
__host__ __device__ constexpr int LayerWeights(int layer) {
  return kTaps * kChannels * LayerInputs(layer);
}
__host__ __device__ constexpr int ParamOffset(int layer) {
  int at = 0;
  for (int i = 0; i < layer; ++i) at += LayerWeights(i) + kChannels;
  return at;
}
__host__ __device__ constexpr int WeightOffset(int layer) {
  int at = 0;
  for (int i = 0; i < layer; ++i) at += LayerWeights(i);
  return at;
}
constexpr int kParams = ParamOffset(kLayers);
constexpr int kWeights = WeightOffset(kLayers);
static_assert(kParams == 50880, "parameter layout");

// strategy code of each candidate stage (DCT8, 16x8, 8x16, 16x16, 32x16,
// 16x32, 32x32)
__constant__ int kStageCode[7] = {0, 6, 7, 4, 10, 11, 5};

__device__ __forceinline__ uint16_t ToBf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // nearest, ties to even
  return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float FromBf16(uint16_t h) {
  return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

// ---- features ---------------------------------------------------------

struct Dct8Table {
  float d[8][8];  // orthonormal DCT-II, d[k][n]
};

Dct8Table HostDct8() {
  Dct8Table t;
  for (int k = 0; k < 8; ++k) {
    for (int n = 0; n < 8; ++n) {
      double v = std::cos(M_PI * (2 * n + 1) * k / 16.0) * std::sqrt(2.0 / 8.0);
      if (k == 0) v *= std::sqrt(0.5);
      t.d[k][n] = static_cast<float>(v);
    }
  }
  return t;
}

// One thread per block.
__global__ __launch_bounds__(256) void FeaturesKernel(
  const float* __restrict__ image, int pw, int ph, const float* __restrict__ qf,
  Dct8Table t, uint16_t* __restrict__ out) {
  const int bw = pw / 8, bh = ph / 8;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= bw * bh) return;
  const int bx = i % bw, by = i / bw;
  const size_t plane = static_cast<size_t>(pw) * ph;
  constexpr float kMeanScale[3] = {16.0f, 1.0f, 1.0f};
  constexpr float kGain[3] = {1.0e4f, 1.0e2f, 1.0e2f};
  float f[kFeatures];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float x[8][8];
#pragma unroll
    for (int y = 0; y < 8; ++y) {
      const float4* row = reinterpret_cast<const float4*>(
        image + c * plane + static_cast<size_t>(by * 8 + y) * pw + bx * 8);
      const float4 a = row[0], b = row[1];
      x[y][0] = a.x; x[y][1] = a.y; x[y][2] = a.z; x[y][3] = a.w;
      x[y][4] = b.x; x[y][5] = b.y; x[y][6] = b.z; x[y][7] = b.w;
    }
    float sum = 0.0f;
#pragma unroll
    for (int y = 0; y < 8; ++y)
#pragma unroll
      for (int xx = 0; xx < 8; ++xx) sum += x[y][xx];
    f[c] = sum / 64.0f * kMeanScale[c];
    // rows: r[y][u] = sum_x d[u][x] x[y][x]
    float r[8][8];
#pragma unroll
    for (int y = 0; y < 8; ++y)
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        float acc = 0.0f;
#pragma unroll
        for (int xx = 0; xx < 8; ++xx) acc += t.d[u][xx] * x[y][xx];
        r[y][u] = acc;
      }
    float e[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int v = 0; v < 8; ++v)
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        float acc = 0.0f;
#pragma unroll
        for (int y = 0; y < 8; ++y) acc += t.d[v][y] * r[y][u];
        const int s = u + v;
        if (s >= 1 && s <= 2) e[0] += acc * acc;
        else if (s >= 3 && s <= 5) e[1] += acc * acc;
        else if (s >= 6) e[2] += acc * acc;
      }
#pragma unroll
    for (int b = 0; b < 3; ++b) f[3 + 3 * c + b] = log1pf(kGain[c] * e[b]);
  }
  f[12] = logf(qf[i]);
  f[13] = f[14] = f[15] = 0.0f;
  uint32_t packed[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    packed[k] = static_cast<uint32_t>(ToBf16(f[2 * k])) |
                (static_cast<uint32_t>(ToBf16(f[2 * k + 1])) << 16);
  }
  uint4* dst = reinterpret_cast<uint4*>(out + static_cast<size_t>(i) * kFeatures);
  dst[0] = make_uint4(packed[0], packed[1], packed[2], packed[3]);
  dst[1] = make_uint4(packed[4], packed[5], packed[6], packed[7]);
}

// ---- convolution --------------------------------------------------------

constexpr int kConvWaves = 4;

__device__ __forceinline__ Bf16x4 LoadBf16x4(const uint16_t* p) {
  return *reinterpret_cast<const Bf16x4*>(p);
}

// One wave per run of 16 positions of one row (grid-stride); the weights
// of all nine taps live in the wave's registers as A operands.
template <int kCin, bool kBias, bool kRelu, bool kMask, bool kHeadOut>
__global__ __launch_bounds__(64 * kConvWaves) void ConvKernel(
  const uint16_t* __restrict__ in, const uint16_t* __restrict__ mask,
  const uint16_t* __restrict__ weights, const float* __restrict__ bias, int n, int h,
  int w, uint16_t* __restrict__ out16, float* __restrict__ out32) {
  constexpr int KC = kCin / 16;
  const int lane = threadIdx.x & 63;
  const int r16 = lane & 15, q4 = (lane >> 4) * 4;
  Bf16x4 a[kTaps][KC][2];
#pragma unroll
  for (int t = 0; t < kTaps; ++t)
#pragma unroll
    for (int kc = 0; kc < KC; ++kc)
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
        a[t][kc][mt] = LoadBf16x4(weights + (t * kChannels + 16 * mt + r16) * kCin + 16 * kc + q4);
  float b_out[2][4];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int i = 0; i < 4; ++i) b_out[mt][i] = kBias ? bias[16 * mt + q4 + i] : 0.0f;
  const int tiles_w = (w + 15) / 16;
  const int64_t tiles = static_cast<int64_t>(n) * h * tiles_w;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kConvWaves;
  const Bf16x4 zero = {0, 0, 0, 0};
  for (int64_t tile = static_cast<int64_t>(blockIdx.x) * kConvWaves + (threadIdx.x >> 6);
       tile < tiles; tile += stride) {
    const int tx = static_cast<int>(tile % tiles_w);
    const int64_t row = tile / tiles_w;  // frame * h + y
    const int y = static_cast<int>(row % h);
    const int64_t frame_row0 = row - y;  // frame * h
    const int x = tx * 16 + r16;
    Float4 acc[2] = {{0.0f, 0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f, 0.0f}};
#pragma unroll
    for (int t = 0; t < kTaps; ++t) {
      const int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
      const bool ok = x < w && yy >= 0 && yy < h && xx >= 0 && xx < w;
      const size_t at = ok ? (static_cast<size_t>(frame_row0 + yy) * w + xx) * kCin : 0;
#pragma unroll
      for (int kc = 0; kc < KC; ++kc) {
        Bf16x4 b = LoadBf16x4(in + at + 16 * kc + q4);
        if constexpr (kMask) {
          const Bf16x4 m = LoadBf16x4(mask + at + 16 * kc + q4);
#pragma unroll
          for (int i = 0; i < 4; ++i) b[i] = m[i] != 0 ? b[i] : static_cast<short>(0);
        }
        b = ok ? b : zero;
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
          acc[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t][kc][mt], b, acc[mt], 0, 0, 0);
      }
    }
    if (x < w) {
      const size_t o = (static_cast<size_t>(row) * w + x) * kChannels;
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) {
        float v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          v[i] = acc[mt][i] + b_out[mt][i];
          if (kRelu) v[i] = v[i] > 0.0f ? v[i] : 0.0f;
        }
        if constexpr (kHeadOut) {
          *reinterpret_cast<float4*>(out32 + o + 16 * mt + q4) = make_float4(v[0], v[1], v[2], v[3]);
        } else {
          const uint32_t lo = static_cast<uint32_t>(ToBf16(v[0])) |
                              (static_cast<uint32_t>(ToBf16(v[1])) << 16);
          const uint32_t hi = static_cast<uint32_t>(ToBf16(v[2])) |
                              (static_cast<uint32_t>(ToBf16(v[3])) << 16);
          *reinterpret_cast<uint2*>(out16 + o + 16 * mt + q4) = make_uint2(lo, hi);
        }
      }
    }
  }
}

// ---- weight gradients -----------------------------------------------------

constexpr int kGradGroups = 256;

// One wave per work group over tiles_per_group runs of 16 positions: the
// positions are the MFMA's K dimension.  A = the masked output gradient
// (output channel x position), B = the shifted inputs (position x input
// channel); D accumulates dW of every tap in registers.  Partial sums per
// group go to `partial` ([group][9][32][kCin] then 32 bias sums).
template <int kCin, bool kMask>
__global__ __launch_bounds__(64) void WeightGradKernel(
  const uint16_t* __restrict__ grad, const uint16_t* __restrict__ mask,
  const uint16_t* __restrict__ in, int n, int h, int w, int64_t tiles_per_group,
  float* __restrict__ partial) {
  constexpr int KC = kCin / 16;
  constexpr int kStride = kTaps * kChannels * kCin + kChannels;
  const int lane = threadIdx.x;
  const int r16 = lane & 15, q4 = (lane >> 4) * 4;
  Float4 acc[2][kTaps][KC];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int t = 0; t < kTaps; ++t)
#pragma unroll
      for (int kc = 0; kc < KC; ++kc) acc[mt][t][kc] = Float4{0.0f, 0.0f, 0.0f, 0.0f};
  float bsum[2] = {0.0f, 0.0f};
  const int tiles_w = (w + 15) / 16;
  const int64_t tiles = static_cast<int64_t>(n) * h * tiles_w;
  const int64_t first = static_cast<int64_t>(blockIdx.x) * tiles_per_group;
  const int64_t last = first + tiles_per_group < tiles ? first + tiles_per_group : tiles;
  for (int64_t tile = first; tile < last; ++tile) {
    const int tx = static_cast<int>(tile % tiles_w);
    const int64_t row = tile / tiles_w;
    const int y = static_cast<int>(row % h);
    const int64_t frame_row0 = row - y;
    Bf16x4 a[2];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int x = tx * 16 + q4 + i;
      const bool ok = x < w;
      const size_t at = ok ? (static_cast<size_t>(row) * w + x) * kChannels : 0;
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) {
        uint16_t g = grad[at + 16 * mt + r16];
        if constexpr (kMask) g = mask[at + 16 * mt + r16] != 0 ? g : 0;
        g = ok ? g : 0;
        a[mt][i] = static_cast<short>(g);
        bsum[mt] += FromBf16(g);
      }
    }
#pragma unroll
    for (int t = 0; t < kTaps; ++t) {
      const int yy = y + t / 3 - 1;
#pragma unroll
      for (int kc = 0; kc < KC; ++kc) {
        Bf16x4 b;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int x = tx * 16 + q4 + i, xx = x + t % 3 - 1;
          const bool ok = x < w && yy >= 0 && yy < h && xx >= 0 && xx < w;
          const size_t at = ok ? (static_cast<size_t>(frame_row0 + yy) * w + xx) * kCin : 0;
          const uint16_t v = in[at + 16 * kc + r16];
          b[i] = static_cast<short>(ok ? v : 0);
        }
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
          acc[mt][t][kc] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[mt], b, acc[mt][t][kc], 0, 0, 0);
      }
    }
  }
  // D[4 (l / 16) + i][l % 16]: output channel 16 mt + q4 + i, input channel
  // 16 kc + r16
  float* p = partial + static_cast<size_t>(blockIdx.x) * kStride;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int t = 0; t < kTaps; ++t)
#pragma unroll
      for (int kc = 0; kc < KC; ++kc)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          p[(t * kChannels + 16 * mt + q4 + i) * kCin + 16 * kc + r16] = acc[mt][t][kc][i];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    float v = bsum[mt];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    if (lane < 16) p[kTaps * kChannels * kCin + 16 * mt + lane] = v;
  }
}

// out[j] = sum over groups of partial[g][j], groups in order.
__global__ __launch_bounds__(256) void ReduceKernel(const float* __restrict__ partial,
                                                    int groups, int stride,
                                                    float* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= stride) return;
  float s = 0.0f;
  for (int g = 0; g < groups; ++g) s += partial[static_cast<size_t>(g) * stride + j];
  out[j] = s;
}

// ---- loss ------------------------------------------------------------------

constexpr int kLossThreads = 256;

__global__ __launch_bounds__(kLossThreads) void LossKernel(
  const float* __restrict__ head, const uint8_t* __restrict__ labels,
  const float* __restrict__ targets, size_t count, float qf_weight,
  uint16_t* __restrict__ grad, float* __restrict__ partial) {
  __shared__ float red[3][kLossThreads];
  const size_t i = static_cast<size_t>(blockIdx.x) * kLossThreads + threadIdx.x;
  float ce = 0.0f, se = 0.0f, correct = 0.0f;
  if (i < count) {
    float z[kChannels];
    const float4* src = reinterpret_cast<const float4*>(head + i * kChannels);
#pragma unroll
    for (int k = 0; k < kChannels / 4; ++k) {
      const float4 v = src[k];
      z[4 * k] = v.x; z[4 * k + 1] = v.y; z[4 * k + 2] = v.z; z[4 * k + 3] = v.w;
    }
    const int lab = labels[i];
    const bool live = lab != kIgnore && lab < kClasses;
    const int y = live ? lab : 0;
    float m = z[0];
    int arg = 0;
#pragma unroll
    for (int c = 1; c < kClasses; ++c) {
      if (z[c] > m) { m = z[c]; arg = c; }
    }
    float e[kClasses];
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < kClasses; ++c) {
      e[c] = expf(z[c] - m);
      s += e[c];
    }
    float zy = z[0];
#pragma unroll
    for (int c = 1; c < kClasses; ++c) zy = c == y ? z[c] : zy;
    ce = live ? logf(s) + m - zy : 0.0f;
    const float err = live ? z[kFieldChannel] - targets[i] : 0.0f;
    se = err * err;
    correct = live && arg == y ? 1.0f : 0.0f;
    uint32_t packed[kChannels / 2];
#pragma unroll
    for (int k = 0; k < kChannels / 2; ++k) {
      float g[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int c = 2 * k + j;
        float v = 0.0f;
        if (c < kClasses) v = live ? e[c] / s - (c == y ? 1.0f : 0.0f) : 0.0f;
        if (c == kFieldChannel) v = 2.0f * qf_weight * err;
        g[j] = v;
      }
      packed[k] = static_cast<uint32_t>(ToBf16(g[0])) | (static_cast<uint32_t>(ToBf16(g[1])) << 16);
    }
    uint4* dst = reinterpret_cast<uint4*>(grad + i * kChannels);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      dst[k] = make_uint4(packed[4 * k], packed[4 * k + 1], packed[4 * k + 2], packed[4 * k + 3]);
  }
  red[0][threadIdx.x] = ce;
  red[1][threadIdx.x] = se;
  red[2][threadIdx.x] = correct;
  __syncthreads();
  for (int half = kLossThreads / 2; half > 0; half /= 2) {
    if (threadIdx.x < half) {
#pragma unroll
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + half];
    }
    __syncthreads();
  }
  if (threadIdx.x < 3) partial[blockIdx.x * 3 + threadIdx.x] = red[threadIdx.x][0];
}

// ---- optimizer ------------------------------------------------------------

// Writes parameter j's bfloat16 copies: the forward weights [9][32][inputs]
// and the input-gradient weights [9][inputs][32] with the taps flipped
// (tap 8 - t), which make the input gradient the same convolution.
__device__ __forceinline__ void PackParameter(int j, float p, uint16_t* __restrict__ wf,
                                              uint16_t* __restrict__ wd) {
  int layer = 0;
#pragma unroll
  for (int L = 1; L < kLayers; ++L) layer = j >= ParamOffset(L) ? L : layer;
  const int local = j - ParamOffset(layer);
  const int cin = LayerInputs(layer);
  if (local >= LayerWeights(layer)) return;  // a bias
  const int t = local / (kChannels * cin), oc = (local / cin) % kChannels, ic = local % cin;
  const uint16_t b = ToBf16(p);
  wf[WeightOffset(layer) + local] = b;
  wd[WeightOffset(layer) + ((kTaps - 1 - t) * cin + ic) * kChannels + oc] = b;
}

__global__ __launch_bounds__(256) void PackKernel(const float* __restrict__ params,
                                                  uint16_t* __restrict__ wf,
                                                  uint16_t* __restrict__ wd) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < kParams) PackParameter(j, params[j], wf, wd);
}

struct AdamParams {
  float lr, beta1, one_minus_beta1, beta2, one_minus_beta2, eps, bc1, bc2, scale;
};

__global__ __launch_bounds__(256) void AdamKernel(float* __restrict__ params,
                                                  float* __restrict__ m,
                                                  float* __restrict__ v,
                                                  const float* __restrict__ grads,
                                                  AdamParams a, uint16_t* __restrict__ wf,
                                                  uint16_t* __restrict__ wd) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= kParams) return;
  const float g = grads[j] * a.scale;
  const float mj = a.beta1 * m[j] + a.one_minus_beta1 * g;
  const float vj = a.beta2 * v[j] + a.one_minus_beta2 * g * g;
  m[j] = mj;
  v[j] = vj;
  const float update = (mj / a.bc1) / (sqrtf(vj / a.bc2) + a.eps);
  const float p = params[j] - a.lr * update;
  params[j] = p;
  PackParameter(j, p, wf, wd);
}

// ---- decisions --------------------------------------------------------------

__global__ __launch_bounds__(64) void DecideKernel(const float* __restrict__ head, int bw,
                                                   int bh, float threshold, int top_k,
                                                   float p_min, uint8_t* __restrict__ allow) {
  const int tiles_w = (bw + 7) / 8;
  const int tx = blockIdx.x % tiles_w, ty = blockIdx.x / tiles_w;
  const int lane = threadIdx.x;
  const int bx = tx * 8 + lane % 8, by = ty * 8 + lane / 8;
  const bool inside = bx < bw && by < bh;
  const size_t at = inside ? static_cast<size_t>(by) * bw + bx : 0;
  float z[7];
#pragma unroll
  for (int s = 0; s < 7; ++s) z[s] = head[at * kChannels + kStageCode[s]];
  float m = z[0];
#pragma unroll
  for (int s = 1; s < 7; ++s) m = z[s] > m ? z[s] : m;
  float p[7], sum = 0.0f;
#pragma unroll
  for (int s = 0; s < 7; ++s) {
    p[s] = expf(z[s] - m);
    sum += p[s];
  }
  float conf = 0.0f;
#pragma unroll
  for (int s = 0; s < 7; ++s) {
    p[s] = p[s] / sum;
    conf = p[s] > conf ? p[s] : conf;
  }
  uint32_t bits = 1;
#pragma unroll
  for (int s = 1; s < 7; ++s) {
    int rank = 0;
#pragma unroll
    for (int o = 0; o < 7; ++o) {
      if (o != s) rank += (p[o] > p[s]) || (p[o] == p[s] && o < s);
    }
    if (rank < top_k || p[s] >= p_min) bits |= 1u << s;
  }
  if (conf < threshold) bits = 0x7f;  // an unsure block keeps the plain search
  if (inside) allow[at] = static_cast<uint8_t>(bits);
}

__global__ __launch_bounds__(256) void WarmStartKernel(const float* __restrict__ head,
                                                       const float* __restrict__ adjusted,
                                                       size_t n, float limit, float offset,
                                                       float lower, float upper,
                                                       float* __restrict__ field) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float r = head[i * kChannels + kFieldChannel] + offset;
  r = r < -limit ? -limit : (limit < r ? limit : r);
  const float f = adjusted[i] * expf(r);
  field[i] = f < lower ? lower : (upper < f ? upper : f);
}

__global__ __launch_bounds__(256) void SampleKernel(const uint8_t* __restrict__ cells,
                                                    const float* __restrict__ field,
                                                    const float* __restrict__ adjusted, size_t n,
                                                    uint8_t* __restrict__ labels,
                                                    float* __restrict__ targets) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  labels[i] = static_cast<uint8_t>(cells[i] >> 1);
  targets[i] = logf(field[i] / adjusted[i]);
}

// (mean of d^3)^(1/3) over n block distances: kNormGroups work groups each
// sum a contiguous range in binary64 (fixed strides, a fixed tree), then one
// group adds the partial sums in order.
constexpr int kNormThreads = 256;
constexpr int kNormGroups = 256;

__device__ __forceinline__ double BlockSum(double s, double* part) {
  part[threadIdx.x] = s;
  __syncthreads();
  for (int half = kNormThreads / 2; half > 0; half /= 2) {
    if (threadIdx.x < half) part[threadIdx.x] += part[threadIdx.x + half];
    __syncthreads();
  }
  return part[0];
}

__global__ __launch_bounds__(kNormThreads) void Norm3PartialKernel(const float* __restrict__ d,
                                                                   size_t n,
                                                                   double* __restrict__ partial) {
  __shared__ double part[kNormThreads];
  const size_t per = (n + kNormGroups - 1) / kNormGroups;
  const size_t begin = blockIdx.x * per, end = begin + per < n ? begin + per : n;
  double s = 0.0;
  for (size_t i = begin + threadIdx.x; i < end; i += kNormThreads) {
    const double v = d[i];
    s += v * v * v;
  }
  const double total = BlockSum(s, part);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

__global__ __launch_bounds__(kNormThreads) void Norm3FinalKernel(const double* __restrict__ partial,
                                                                 size_t n, float* __restrict__ out) {
  __shared__ double part[kNormThreads];
  const double total = BlockSum(threadIdx.x < kNormGroups ? partial[threadIdx.x] : 0.0, part);
  if (threadIdx.x == 0) *out = static_cast<float>(cbrt(total / static_cast<double>(n)));
}

static_assert(kNormGroups <= kNormThreads, "one final group sums the partials");
static_assert(kNormGroups == kProposalNormPartials, "the caller's partial buffer");

}  // namespace

GJXLResult LaunchProposalNorm3(GJXLAmdDevice* device, const float* distances, size_t n,
                               double* partial, float* out) {
  hipLaunchKernelGGL(Norm3PartialKernel, dim3(kNormGroups), dim3(kNormThreads), 0,
                     device->stream, distances, n, partial);
  GJXL_AMD_TRY(CheckLaunch("Norm3PartialKernel"));
  hipLaunchKernelGGL(Norm3FinalKernel, dim3(1), dim3(kNormThreads), 0, device->stream, partial,
                     n, out);
  return CheckLaunch("Norm3FinalKernel");
}

GJXLResult LaunchProposalFeatures(GJXLAmdDevice* device, const float* opsin, int width,
                                  int height, const float* quant_field, uint16_t* features) {
  static const Dct8Table table = HostDct8();
  const uint32_t blocks = static_cast<uint32_t>(width / 8) * (height / 8);
  hipLaunchKernelGGL(FeaturesKernel, dim3(CeilDiv(blocks, 256)), dim3(256), 0, device->stream,
                     opsin, width, height, quant_field, table, features);
  return CheckLaunch("FeaturesKernel");
}

GJXLResult LaunchProposalDecide(GJXLAmdDevice* device, const float* head, int bw, int bh,
                                float threshold, int top_k, float p_min, uint8_t* allow) {
  const uint32_t tiles = CeilDiv(bw, 8) * CeilDiv(bh, 8);
  hipLaunchKernelGGL(DecideKernel, dim3(tiles), dim3(64), 0, device->stream, head, bw, bh,
                     threshold, top_k, p_min, allow);
  return CheckLaunch("DecideKernel");
}

GJXLResult LaunchProposalWarmStart(GJXLAmdDevice* device, const float* head,
                                   const float* adjusted, size_t n, float limit, float offset,
                                   float lower, float upper, float* field) {
  hipLaunchKernelGGL(WarmStartKernel, dim3(CeilDiv(static_cast<uint32_t>(n), 256)), dim3(256),
                     0, device->stream, head, adjusted, n, limit, offset, lower, upper, field);
  return CheckLaunch("WarmStartKernel");
}

GJXLResult LaunchProposalSample(GJXLAmdDevice* device, const uint8_t* cells, const float* field,
                                const float* adjusted, size_t n, uint8_t* labels,
                                float* targets) {
  hipLaunchKernelGGL(SampleKernel, dim3(CeilDiv(static_cast<uint32_t>(n), 256)), dim3(256), 0,
                     device->stream, cells, field, adjusted, n, labels, targets);
  return CheckLaunch("SampleKernel");
}

}  // namespace gjxl_amd

// The network's device state.  Activation planes: act[L] is layer L's
// stored ReLU output (L = 0..4); grad[2] ping-pong input gradients.
struct GJXLAmdProposal {
  GJXLAmdDevice* device = nullptr;
  size_t max_positions = 0;
  float* params = nullptr;   // kParams
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  float* grads = nullptr;
  uint16_t* wf = nullptr;    // kWeights bfloat16, forward
  uint16_t* wd = nullptr;    // kWeights bfloat16, input gradient
  float* head = nullptr;     // max_positions * 32
  uint16_t* act = nullptr;   // 5 * max_positions * 32
  uint16_t* grad = nullptr;  // 2 * max_positions * 32
  float* partial = nullptr;  // kGradGroups * (9 * 32 * 32 + 32)
  float* loss_partial = nullptr;
  double* host_sums = nullptr;
  const uint16_t* features = nullptr;  // of the last forward
  uint32_t n = 0, h = 0, w = 0;
  int step = 0;
  void* block = nullptr;
};

namespace {

using gjxl_amd::Fail;

constexpr size_t kPartialStride = gjxl_amd::kTaps * gjxl_amd::kChannels * gjxl_amd::kChannels +
                                  gjxl_amd::kChannels;

GJXLResult ValidateNet(const GJXLAmdProposal* net) {
  if (net == nullptr || net->block == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "proposal network is null or destroyed");
  }
  return GJXL_OK;
}

size_t LossBlocks(size_t positions) {
  return (positions + gjxl_amd::kLossThreads - 1) / gjxl_amd::kLossThreads;
}

}  // namespace

namespace gjxl_amd {

GJXLResult ProposalForward(GJXLAmdProposal* net, hipStream_t st, const uint16_t* features,
                           uint32_t n, uint32_t h, uint32_t w, const float** head) {
  const size_t positions = static_cast<size_t>(n) * h * w;
  if (features == nullptr || positions == 0 || positions > net->max_positions) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "proposal forward: no input or too many positions");
  }
  const size_t plane = kChannels * net->max_positions;
  const uint32_t tiles = static_cast<uint32_t>(
    std::min<size_t>(static_cast<size_t>(n) * h * ((w + 15) / 16), 1u << 30));
  const dim3 grid(std::min<uint32_t>(CeilDiv(tiles, kConvWaves), 2048));
  const dim3 block(64 * kConvWaves);
  const int ni = static_cast<int>(n), hi = static_cast<int>(h), wi = static_cast<int>(w);
  hipLaunchKernelGGL((ConvKernel<kFeatures, true, true, false, false>), grid, block, 0, st,
                     features, nullptr, net->wf + WeightOffset(0),
                     net->params + ParamOffset(0) + LayerWeights(0), ni, hi, wi, net->act,
                     nullptr);
  GJXL_AMD_TRY(CheckLaunch("ConvKernel"));
  for (int L = 1; L < kLayers - 1; ++L) {
    hipLaunchKernelGGL((ConvKernel<kChannels, true, true, false, false>), grid, block, 0, st,
                       net->act + (L - 1) * plane, nullptr, net->wf + WeightOffset(L),
                       net->params + ParamOffset(L) + LayerWeights(L), ni, hi, wi,
                       net->act + L * plane, nullptr);
    GJXL_AMD_TRY(CheckLaunch("ConvKernel"));
  }
  hipLaunchKernelGGL((ConvKernel<kChannels, true, false, false, true>), grid, block, 0, st,
                     net->act + (kLayers - 2) * plane, nullptr, net->wf + WeightOffset(kLayers - 1),
                     net->params + ParamOffset(kLayers - 1) + LayerWeights(kLayers - 1), ni, hi,
                     wi, nullptr, net->head);
  GJXL_AMD_TRY(CheckLaunch("ConvKernel"));
  net->features = features;
  net->n = n;
  net->h = h;
  net->w = w;
  if (head != nullptr) *head = net->head;
  return GJXL_OK;
}

}  // namespace gjxl_amd

extern "C" {

size_t gjxl_amd_proposal_parameter_count(void) { return gjxl_amd::kParams; }

GJXLResult gjxl_amd_proposal_create(GJXLAmdDevice* device, size_t max_positions,
                                    GJXLAmdProposal** out) {
  using namespace gjxl_amd;
  GJXL_AMD_TRY(ValidateDevice(device));
  if (out == nullptr || max_positions == 0 || max_positions > (size_t{1} << 28)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "proposal positions must be in [1, 2^28]");
  }
  *out = nullptr;
  auto align = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t sizes[] = {
    4 * size_t{kParams}, 4 * size_t{kParams}, 4 * size_t{kParams}, 4 * size_t{kParams},
    2 * size_t{kWeights}, 2 * size_t{kWeights}, 4 * kChannels * max_positions,
    2 * 5 * kChannels * max_positions, 2 * 2 * kChannels * max_positions,
    4 * kGradGroups * kPartialStride, 4 * 3 * LossBlocks(max_positions)};
  size_t total = 0;
  for (size_t s : sizes) total += align(s);
  auto* net = new GJXLAmdProposal();
  net->device = device;
  net->max_positions = max_positions;
  void* block = nullptr;
  const GJXLResult r = DeviceAlloc(device, total, &block);
  if (r != GJXL_OK) {
    delete net;
    return r;
  }
  net->block = block;
  char* at = static_cast<char*>(block);
  void* parts[11];
  for (int i = 0; i < 11; ++i) {
    parts[i] = at;
    at += align(sizes[i]);
  }
  net->params = static_cast<float*>(parts[0]);
  net->adam_m = static_cast<float*>(parts[1]);
  net->adam_v = static_cast<float*>(parts[2]);
  net->grads = static_cast<float*>(parts[3]);
  net->wf = static_cast<uint16_t*>(parts[4]);
  net->wd = static_cast<uint16_t*>(parts[5]);
  net->head = static_cast<float*>(parts[6]);
  net->act = static_cast<uint16_t*>(parts[7]);
  net->grad = static_cast<uint16_t*>(parts[8]);
  net->partial = static_cast<float*>(parts[9]);
  net->loss_partial = static_cast<float*>(parts[10]);
  GJXLResult status = Check(hipHostMalloc(reinterpret_cast<void**>(&net->host_sums),
                                          3 * LossBlocks(max_positions) * sizeof(float) + 64),
                            "hipHostMalloc(proposal sums)");
  if (status == GJXL_OK) {
    status = Check(hipMemsetAsync(block, 0, total, device->stream), "hipMemsetAsync(proposal)");
  }
  if (status == GJXL_OK) status = Check(hipStreamSynchronize(device->stream), "hipStreamSynchronize");
  if (status != GJXL_OK) {
    gjxl_amd_proposal_destroy(net);
    return status;
  }
  *out = net;
  return GJXL_OK;
}

void gjxl_amd_proposal_destroy(GJXLAmdProposal* net) {
  if (net == nullptr) return;
  if (net->block != nullptr) {
    (void)hipStreamSynchronize(net->device->stream);
    gjxl_amd::DeviceFree(net->device, net->block);
  }
  if (net->host_sums != nullptr) (void)hipHostFree(net->host_sums);
  net->block = nullptr;
  delete net;
}

GJXLResult gjxl_amd_proposal_set_parameters(GJXLAmdProposal* net, const float* host,
                                            size_t count) {
  using namespace gjxl_amd;
  GJXL_AMD_TRY(ValidateNet(net));
  if (host == nullptr || count != static_cast<size_t>(kParams)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "proposal parameters: wrong count");
  }
  for (size_t i = 0; i < count; ++i) {
    if (!std::isfinite(host[i])) {
      return Fail(GJXL_ERROR_INVALID_ARGUMENT, "proposal parameters must be finite");
    }
  }
  hipStream_t st = net->device->stream;
  GJXL_AMD_TRY(Check(hipMemcpyAsync(net->params, host, 4 * count, hipMemcpyHostToDevice, st),
                     "hipMemcpyAsync(parameters)"));
  GJXL_AMD_TRY(Check(hipMemsetAsync(net->adam_m, 0, 4 * count, st), "hipMemsetAsync"));
  GJXL_AMD_TRY(Check(hipMemsetAsync(net->adam_v, 0, 4 * count, st), "hipMemsetAsync"));
  hipLaunchKernelGGL(PackKernel, dim3(CeilDiv(kParams, 256)), dim3(256), 0, st, net->params,
                     net->wf, net->wd);
  GJXL_AMD_TRY(CheckLaunch("PackKernel"));
  net->step = 0;
  return Check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

GJXLResult gjxl_amd_proposal_get_parameters(GJXLAmdProposal* net, float* host, size_t count) {
  using namespace gjxl_amd;
  GJXL_AMD_TRY(ValidateNet(net));
  if (host == nullptr || count != static_cast<size_t>(kParams)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "proposal parameters: wrong count");
  }
  GJXL_AMD_TRY(Check(hipMemcpyAsync(host, net->params, 4 * count, hipMemcpyDeviceToHost,
                                    net->device->stream), "hipMemcpyAsync(parameters)"));
  return Check(hipStreamSynchronize(net->device->stream), "hipStreamSynchronize");
}

GJXLResult gjxl_amd_proposal_buffer(GJXLAmdProposal* net, int32_t id, void** pointer,
                                    size_t* bytes) {
  using namespace gjxl_amd;
  GJXL_AMD_TRY(ValidateNet(net));
  if (pointer == nullptr) return Fail(GJXL_ERROR_INVALID_ARGUMENT, "pointer is null");
  size_t b = 0;
  switch (id) {
    case GJXL_AMD_PROPOSAL_HEAD: *pointer = net->head; b = 4 * kChannels * net->max_positions; break;
    case GJXL_AMD_PROPOSAL_PARAMETERS: *pointer = net->params; b = 4 * size_t{kParams}; break;
    case GJXL_AMD_PROPOSAL_GRADIENTS: *pointer = net->grads; b = 4 * size_t{kParams}; break;
    case GJXL_AMD_PROPOSAL_ACTIVATIONS:
      *pointer = net->act;
      b = 2 * 5 * kChannels * net->max_positions;
      break;
    default: return Fail(GJXL_ERROR_INVALID_ARGUMENT, "unknown proposal buffer");
  }
  if (bytes != nullptr) *bytes = b;
  return GJXL_OK;
}

GJXLResult gjxl_amd_proposal_features(GJXLAmdDevice* device, const float* opsin,
                                      uint32_t width, uint32_t height,
                                      const float* quant_field, uint16_t* features) {
  using namespace gjxl_amd;
  GJXL_AMD_TRY(ValidateDevice(device));
  if (width == 0 || height == 0 || width % 8 != 0 || height % 8 != 0 || width > (1u << 16) ||
      height > (1u << 16)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "padded extent must be multiples of 8 up to 65536");
  }
  if (opsin == nullptr || quant_field == nullptr || features == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "feature buffers must be non-null");
  }
  return LaunchProposalFeatures(device, opsin, static_cast<int>(width),
                                static_cast<int>(height), quant_field, features);
}

GJXLResult gjxl_amd_proposal_forward(GJXLAmdProposal* net, const uint16_t* features,
                                     uint32_t n, uint32_t h, uint32_t w) {
  GJXL_AMD_TRY(ValidateNet(net));
  return gjxl_amd::ProposalForward(net, net->device->stream, features, n, h, w, nullptr);
}

GJXLResult gjxl_amd_proposal_backward(GJXLAmdProposal* net, const uint8_t* labels,
                                      const float* targets, float qf_weight, double* sums) {
  using namespace gjxl_amd;
  GJXL_AMD_TRY(ValidateNet(net));
  if (net->features == nullptr) return Fail(GJXL_ERROR_INVALID_ARGUMENT, "no forward to differentiate");
  if (labels == nullptr || targets == nullptr || !std::isfinite(qf_weight) || qf_weight < 0.0f) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "proposal backward: labels, targets, weight");
  }
  hipStream_t st = net->device->stream;
  const size_t positions = static_cast<size_t>(net->n) * net->h * net->w;
  const size_t plane = kChannels * net->max_positions;
  const int ni = static_cast<int>(net->n), hi = static_cast<int>(net->h),
            wi = static_cast<int>(net->w);
  const size_t loss_blocks = LossBlocks(positions);
  uint16_t* g_cur = net->grad;
  uint16_t* g_next = net->grad + plane;
  hipLaunchKernelGGL(LossKernel, dim3(static_cast<uint32_t>(loss_blocks)), dim3(kLossThreads), 0,
                     st, net->head, labels, targets, positions, qf_weight, g_cur,
                     net->loss_partial);
  GJXL_AMD_TRY(CheckLaunch("LossKernel"));
  const int64_t tiles = static_cast<int64_t>(net->n) * net->h * ((net->w + 15) / 16);
  const int groups = static_cast<int>(std::min<int64_t>(kGradGroups, tiles));
  const int64_t per_group = (tiles + groups - 1) / groups;
  const uint32_t conv_tiles = static_cast<uint32_t>(std::min<int64_t>(tiles, 1 << 30));
  const dim3 grid(std::min<uint32_t>(CeilDiv(conv_tiles, kConvWaves), 2048));
  for (int L = kLayers - 1; L >= 0; --L) {
    const uint16_t* mask = L == kLayers - 1 ? nullptr : net->act + L * plane;
    const uint16_t* in = L == 0 ? net->features : net->act + (L - 1) * plane;
    const int cin = LayerInputs(L);
    const int stride = LayerWeights(L) + kChannels;
    if (L == 0) {
      hipLaunchKernelGGL((WeightGradKernel<kFeatures, true>), dim3(groups), dim3(64), 0, st,
                         g_cur, mask, in, ni, hi, wi, per_group, net->partial);
    } else if (mask != nullptr) {
      hipLaunchKernelGGL((WeightGradKernel<kChannels, true>), dim3(groups), dim3(64), 0, st,
                         g_cur, mask, in, ni, hi, wi, per_group, net->partial);
    } else {
      hipLaunchKernelGGL((WeightGradKernel<kChannels, false>), dim3(groups), dim3(64), 0, st,
                         g_cur, mask, in, ni, hi, wi, per_group, net->partial);
    }
    GJXL_AMD_TRY(CheckLaunch("WeightGradKernel"));
    hipLaunchKernelGGL(ReduceKernel, dim3(CeilDiv(stride, 256)), dim3(256), 0, st,
                       net->partial, groups, stride, net->grads + ParamOffset(L));
    GJXL_AMD_TRY(CheckLaunch("ReduceKernel"));
    (void)cin;
    if (L > 0) {
      if (mask != nullptr) {
        hipLaunchKernelGGL((ConvKernel<kChannels, false, false, true, false>), grid, dim3(64 * kConvWaves),
                           0, st, g_cur, mask, net->wd + WeightOffset(L), nullptr, ni, hi, wi,
                           g_next, nullptr);
      } else {
        hipLaunchKernelGGL((ConvKernel<kChannels, false, false, false, false>), grid,
                           dim3(64 * kConvWaves), 0, st, g_cur, nullptr, net->wd + WeightOffset(L),
                           nullptr, ni, hi, wi, g_next, nullptr);
      }
      GJXL_AMD_TRY(CheckLaunch("ConvKernel"));
      std::swap(g_cur, g_next);
    }
  }
  if (sums != nullptr) {
    GJXL_AMD_TRY(Check(hipMemcpyAsync(net->host_sums, net->loss_partial,
                                      3 * loss_blocks * sizeof(float), hipMemcpyDeviceToHost, st),
                       "hipMemcpyAsync(loss)"));
    GJXL_AMD_TRY(Check(hipStreamSynchronize(st), "hipStreamSynchronize"));
    const float* p = reinterpret_cast<const float*>(net->host_sums);
    double s[3] = {0.0, 0.0, 0.0};
    for (size_t b = 0; b < loss_blocks; ++b)
      for (int k = 0; k < 3; ++k) s[k] += p[3 * b + k];
    for (int k = 0; k < 3; ++k) sums[k] = s[k];
  }
  return GJXL_OK;
}

GJXLResult gjxl_amd_proposal_adam(GJXLAmdProposal* net, float learning_rate, float beta1,
                                  float beta2, float epsilon, float grad_scale) {
  using namespace gjxl_amd;
  GJXL_AMD_TRY(ValidateNet(net));
  if (!(learning_rate >= 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) ||
      !(beta2 >= 0.0f && beta2 < 1.0f) || !(epsilon > 0.0f) || !std::isfinite(grad_scale)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Adam hyper-parameters out of range");
  }
  ++net->step;
  AdamParams a;
  a.lr = learning_rate;
  a.beta1 = beta1;
  a.one_minus_beta1 = static_cast<float>(1.0 - static_cast<double>(beta1));
  a.beta2 = beta2;
  a.one_minus_beta2 = static_cast<float>(1.0 - static_cast<double>(beta2));
  a.eps = epsilon;
  a.bc1 = static_cast<float>(1.0 - std::pow(static_cast<double>(beta1), net->step));
  a.bc2 = static_cast<float>(1.0 - std::pow(static_cast<double>(beta2), net->step));
  a.scale = grad_scale;
  hipLaunchKernelGGL(AdamKernel, dim3(CeilDiv(kParams, 256)), dim3(256), 0, net->device->stream,
                     net->params, net->adam_m, net->adam_v, net->grads, a, net->wf, net->wd);
  return CheckLaunch("AdamKernel");
}

GJXLResult gjxl_amd_proposal_decide(GJXLAmdDevice* device, const float* head,
                                    uint32_t block_width, uint32_t block_height,
                                    float threshold, uint32_t top_k, float p_min,
                                    uint8_t* allow) {
  using namespace gjxl_amd;
  GJXL_AMD_TRY(ValidateDevice(device));
  if (head == nullptr || allow == nullptr || block_width == 0 || block_height == 0 ||
      block_width > 8192 || block_height > 8192 || top_k > 7 || !std::isfinite(threshold) ||
      !std::isfinite(p_min)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "proposal decide: buffers, extent or options");
  }
  return LaunchProposalDecide(device, head, static_cast<int>(block_width),
                              static_cast<int>(block_height), threshold,
                              static_cast<int>(top_k), p_min, allow);
}

GJXLResult gjxl_amd_proposal_warm_start(GJXLAmdDevice* device, const float* head,
                                        const float* adjusted, size_t count, float limit,
                                        float offset, float lower, float upper, float* field) {
  using namespace gjxl_amd;
  GJXL_AMD_TRY(ValidateDevice(device));
  if (head == nullptr || adjusted == nullptr || field == nullptr || count == 0 ||
      count > (size_t{1} << 31) || !(limit >= 0.0f) || !std::isfinite(offset) ||
      !(lower > 0.0f) || !(upper >= lower)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "proposal warm start: buffers or bounds");
  }
  return LaunchProposalWarmStart(device, head, adjusted, count, limit, offset, lower, upper,
                                 field);
}

GJXLResult gjxl_amd_proposal_options_init(GJXLAmdProposalOptions* options, size_t struct_size) {
  if (options == nullptr || struct_size < sizeof(GJXLAmdProposalOptions)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "proposal options are missing or too small");
  }
  std::memset(options, 0, sizeof(*options));
  options->struct_size = sizeof(*options);
  options->threshold = 0.3f;
  options->top_k = 2;
  options->p_min = 0.1f;
  options->final_compare = 1;
  options->evaluations = 2;
  options->field_limit = 2.0f;
  options->field_offset = 0.0f;
  options->judge_ratio = 0.7f;
  return GJXL_OK;
}

}  // extern "C"
