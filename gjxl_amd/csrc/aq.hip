// a11-a13: adaptive-quantization setup and one evaluation up to the
// reconstructed, Gaborish-inverted-domain XYB image.
//
//   AnchorLayoutKernel      per 256x256 AC group: anchors in raster order,
//                           their coefficient offsets and the group lists
//                           (src/codec/reconstruction.cpp:386-760 layout)
//   AnchorDctKernel         forward DCT of every selected transform, once
//   AdjustFieldKernel       AdjustQuantField (adaptive_quantization.cpp:740)
//   MatrixStatsKernel       X/B matrix-scale statistics (workflow.cpp:1318)
//   RangeKernel             min/max of the adjusted field
//   Radix*Kernel            exact median and MAD of the quant field
//   RawQuantKernel          CreateQuantizerFromField's raw quant
//   FinalCflKernel          ComputeFinalColorCorrelationMap (fast LS)
//   CodeTransformKernel     one transform of ComputeQuantizedCoefficients
//   DcQuantKernel           prediction-aware DC, weighted predictor,
//                           evaluated as a wavefront over each DC group
//   EpfSigmaKernel          ComputeEpfInverseSigma
//   SmoothDcKernel          SmoothDcCoefficients
//   ReconstructKernel       ReconstructQuantizedCoefficients
#include <cmath>
#include <cstring>
#include <type_traits>

#include "aq_common.h"
#include "common.h"
#include "tables.h"

namespace gjxl_amd {
namespace {

// Fills words with a value (a kernel, so that the sequence stays capturable
// in a HIP graph without memset nodes).
__global__ void FillWordsKernel(uint32_t* __restrict__ p, int n, uint32_t value) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = value;
}

// ---- anchor layout ---------------------------------------------------------

__device__ __forceinline__ bool SearchedStrategy(int st) {
  return st == kDct8 || st == kDct16x16 || st == kDct32x32 || st == kDct16x8 ||
         st == kDct8x16 || st == kDct32x16 || st == kDct16x32;
}

// One work group per AC group, one thread per cell of its 32x32 blocks.
// The grid is validated on the way (every cell covered by exactly one
// transform of a searched strategy, footprints inside the group): a group
// that fails gets no anchors and raises *error, so that no later kernel
// indexes past the group's 65536 coefficient slots.  Anchors are numbered in
// raster order and their coefficient offsets are the prefix sums of the
// transform sizes in that order (an exclusive scan over the group).
// strategy_lists may be null; otherwise each group appends its anchors to
// the per-strategy lists (one reservation per strategy and group; the order
// within a list is not specified).
__global__ __launch_bounds__(1024) void AnchorLayoutKernel(
  const uint8_t* __restrict__ cells, int bw, int bh,
  uint32_t* __restrict__ offsets, uint32_t* __restrict__ anchor_list,
  uint32_t* __restrict__ anchor_count, uint32_t* __restrict__ error,
  uint32_t* __restrict__ strategy_lists, size_t capacity,
  uint32_t* __restrict__ strategy_counts) {
  __shared__ uint32_t cover[kGroupAnchors];
  __shared__ uint32_t per[kStrategyStages], base[kStrategyStages];
  __shared__ int bad;
  const int gw = (bw + kGroupBlocks - 1) / kGroupBlocks;
  const int g = blockIdx.x;
  const int gx0 = (g % gw) * kGroupBlocks, gy0 = (g / gw) * kGroupBlocks;
  const int ew = min(kGroupBlocks, bw - gx0), eh = min(kGroupBlocks, bh - gy0);
  const int t = threadIdx.x, lx = t & (kGroupBlocks - 1), ly = t / kGroupBlocks;
  const bool inside = lx < ew && ly < eh;
  const size_t cell_index = static_cast<size_t>(gy0 + ly) * bw + gx0 + lx;
  cover[t] = 0;
  if (t < kStrategyStages) per[t] = 0;
  if (t == 0) bad = 0;
  __syncthreads();
  const uint8_t c = inside ? cells[cell_index] : 0;
  const bool anchor = inside && (c & 1);
  const int st = c >> 1;
  int cw = 1, ch = 1;
  if (anchor) {
    if (!SearchedStrategy(st)) {
      bad = 1;
    } else {
      cw = CoveredW(st);
      ch = CoveredH(st);
      if (lx + cw > ew || ly + ch > eh) {
        bad = 1;
      } else {
        for (int dy = 0; dy < ch; ++dy)
          for (int dx = 0; dx < cw; ++dx) atomicAdd(&cover[(ly + dy) * kGroupBlocks + lx + dx], 1u);
      }
    }
  }
  __syncthreads();
  if (inside && cover[t] != 1) bad = 1;
  __syncthreads();
  if (bad) {
    if (inside) offsets[cell_index] = kNoAnchor;
    if (t == 0) {
      anchor_count[g] = 0;
      atomicOr(error, 1u);
    }
    return;
  }
  // anchors in the high bits, coefficient slots (at most 65536) in the low 20
  uint32_t total;
  const uint32_t mine = anchor ? (1u << 20) | static_cast<uint32_t>(cw * ch * 64) : 0u;
  const uint32_t before = BlockExclusiveScanU32(mine, &total);
  const uint32_t packed = (static_cast<uint32_t>(gy0 + ly) << 16) | static_cast<uint32_t>(gx0 + lx);
  if (inside) {
    offsets[cell_index] = anchor ? static_cast<uint32_t>(g) * kGroupCoefficients + (before & 0xfffffu)
                                 : kNoAnchor;
  }
  if (anchor) anchor_list[static_cast<size_t>(g) * kGroupAnchors + (before >> 20)] = packed;
  if (t == 0) anchor_count[g] = total >> 20;
  if (strategy_lists == nullptr) return;
  uint32_t rank = 0;
  const int stage = StrategyStage(st);
  if (anchor) rank = atomicAdd(&per[stage], 1u);
  __syncthreads();
  if (t < kStrategyStages) base[t] = per[t] ? atomicAdd(&strategy_counts[t], per[t]) : 0u;
  __syncthreads();
  if (anchor) strategy_lists[static_cast<size_t>(stage) * capacity + base[stage] + rank] = packed;
}

// ---- per-strategy transform kernels -----------------------------------------
//
// Each launch handles the transforms of one strategy (from its list), kK of
// them per work group, with compile-time extents.

constexpr int kAqThreads = 256;
constexpr int kSlots = 3072;

template <int kStrategy>
struct AnchorShape {
  static constexpr int kCw = CoveredW(kStrategy), kCh = CoveredH(kStrategy);
  static constexpr int kW = kCw * 8, kH = kCh * 8, kN = kW * kH;
  static constexpr int kK = kSlots / (3 * kN);
};

struct TransformSlot {
  int bx, by;
  uint32_t offset;
  int ok;
};

// Thread k < kK decodes this group's k-th list entry.
template <int kK>
__device__ __forceinline__ void LoadSlots(const uint32_t* __restrict__ entries,
                                          uint32_t count,
                                          const uint32_t* __restrict__ offsets, int bw,
                                          TransformSlot* slots) {
  if (threadIdx.x < kK) {
    const uint32_t idx = blockIdx.x * kK + threadIdx.x;
    TransformSlot t{0, 0, 0u, 0};
    if (idx < count) {
      const uint32_t packed = entries[idx];
      t.bx = packed & 0xffff;
      t.by = packed >> 16;
      t.offset = offsets[static_cast<size_t>(t.by) * bw + t.bx];
      t.ok = 1;
    }
    slots[threadIdx.x] = t;
  }
}

// The separable binary64 passes.  A wave takes 64 rows (or columns) of the
// kPlanes planes and one group of kOut outputs, so that the basis entries,
// read from global memory, are the same for all its lanes (scalar loads);
// each output is one chain in the reference's order (src/codec/dct.cpp:
// 135-233).  Row passes write binary64 rows of stride W + 1 (padded against
// LDS bank conflicts).
constexpr int kOut = 8;

__device__ __forceinline__ int WaveIndex() {
  return __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
}

// Forward rows: dst[p][y][u] = sum_x src[p][x][y] * fh[u][x]; src holds the
// pixels transposed ([p][x][y]); fhT is the transposed forward basis.
template <int W, int H, int kPlanes>
__device__ __forceinline__ void ForwardRows(const double* __restrict__ fhT, const float* src,
                                            double* dst) {
  constexpr int N = W * H, TS = W + 1, kRows = kPlanes * H, kBlocks = (kRows + 63) / 64;
  const int lane = threadIdx.x & 63, waves = blockDim.x >> 6;
  for (int task = WaveIndex(); task < kBlocks * (W / kOut); task += waves) {
    const int g = task % (W / kOut), row = (task / (W / kOut)) * 64 + lane;
    if (row >= kRows) continue;
    const int p = row / H, y = row % H;
    const float* s = src + p * N + y;
    double acc[kOut] = {};
#pragma unroll 4
    for (int x = 0; x < W; ++x) {
      const double v = static_cast<double>(s[x * H]);
#pragma unroll
      for (int j = 0; j < kOut; ++j) acc[j] += v * fhT[x * W + g * kOut + j];
    }
    double* d = dst + row * TS + g * kOut;
#pragma unroll
    for (int j = 0; j < kOut; ++j) d[j] = acc[j];
  }
}

// Forward columns: coefficient (v, u) of plane p = sum_y fv[v][y] *
// src[p][y][u], handed to store(p, libjxl index, value); fvT is the
// transposed forward basis.
template <int W, int H, int kPlanes, typename Store>
__device__ __forceinline__ void ForwardCols(const double* __restrict__ fvT, const double* src,
                                            Store store) {
  constexpr int TS = W + 1, kCols = kPlanes * W, kBlocks = (kCols + 63) / 64;
  const int lane = threadIdx.x & 63, waves = blockDim.x >> 6;
  for (int task = WaveIndex(); task < kBlocks * (H / kOut); task += waves) {
    const int g = task % (H / kOut), col = (task / (H / kOut)) * 64 + lane;
    if (col >= kCols) continue;
    const int p = col / W, u = col % W;
    const double* s = src + p * H * TS + u;
    double acc[kOut] = {};
#pragma unroll 4
    for (int y = 0; y < H; ++y) {
      const double v = s[y * TS];
#pragma unroll
      for (int j = 0; j < kOut; ++j) acc[j] += fvT[y * H + g * kOut + j] * v;
    }
#pragma unroll
    for (int j = 0; j < kOut; ++j) {
      store(p, CoefIndex(W, H, g * kOut + j, u), static_cast<float>(acc[j]));
    }
  }
}

// Inverse rows: dst[p][v][x] = sum_u src[p][idx(v, u)] * ih[u][x].
template <int W, int H, int kPlanes>
__device__ __forceinline__ void InverseRows(const double* __restrict__ ih, const float* src,
                                            double* dst) {
  constexpr int N = W * H, TS = W + 1, kRows = kPlanes * H, kBlocks = (kRows + 63) / 64;
  const int lane = threadIdx.x & 63, waves = blockDim.x >> 6;
  for (int task = WaveIndex(); task < kBlocks * (W / kOut); task += waves) {
    const int g = task % (W / kOut), row = (task / (W / kOut)) * 64 + lane;
    if (row >= kRows) continue;
    const int p = row / H, v = row % H;
    const float* s = src + p * N;
    double acc[kOut] = {};
#pragma unroll 4
    for (int u = 0; u < W; ++u) {
      const double c = static_cast<double>(s[CoefIndex(W, H, v, u)]);
#pragma unroll
      for (int j = 0; j < kOut; ++j) acc[j] += c * ih[u * W + g * kOut + j];
    }
    double* d = dst + row * TS + g * kOut;
#pragma unroll
    for (int j = 0; j < kOut; ++j) d[j] = acc[j];
  }
}

// Inverse columns: pixel (y, x) of plane p = sum_v iv[v][y] * src[p][v][x],
// handed to store(p, y, x, value).
template <int W, int H, int kPlanes, typename Store>
__device__ __forceinline__ void InverseCols(const double* __restrict__ iv, const double* src,
                                            Store store) {
  constexpr int TS = W + 1, kCols = kPlanes * W, kBlocks = (kCols + 63) / 64;
  const int lane = threadIdx.x & 63, waves = blockDim.x >> 6;
  for (int task = WaveIndex(); task < kBlocks * (H / kOut); task += waves) {
    const int g = task % (H / kOut), col = (task / (H / kOut)) * 64 + lane;
    if (col >= kCols) continue;
    const int p = col / W, x = col % W;
    const double* s = src + p * H * TS + x;
    double acc[kOut] = {};
#pragma unroll 4
    for (int v = 0; v < H; ++v) {
      const double t = s[v * TS];
#pragma unroll
      for (int j = 0; j < kOut; ++j) acc[j] += iv[v * H + g * kOut + j] * t;
    }
#pragma unroll
    for (int j = 0; j < kOut; ++j) store(p, g * kOut + j, x, static_cast<float>(acc[j]));
  }
}

// Forward DCT of the three channels of every transform of one strategy into
// the coefficient store (3 planes of groups * 65536 floats).
template <int kStrategy>
__global__ __launch_bounds__(kAqThreads) void AnchorDctKernel(
  const float* __restrict__ opsin, int width, int height,
  const uint32_t* __restrict__ entries, uint32_t count,
  const uint32_t* __restrict__ offsets, const double* __restrict__ bases,
  size_t store, float* __restrict__ coefs) {
  using S = AnchorShape<kStrategy>;
  constexpr int W = S::kW, H = S::kH, N = S::kN, K = S::kK;
  __shared__ float pix[K * 3 * N];            // transposed: [p][x][y]
  __shared__ double tmp[K * 3 * H * (W + 1)];
  __shared__ TransformSlot slots[K];
  LoadSlots<K>(entries, count, offsets, width / 8, slots);
  __syncthreads();
  const size_t plane = static_cast<size_t>(width) * height;
  for (int e = threadIdx.x; e < K * 3 * N; e += kAqThreads) {
    const int p = e / N, k = p / 3, c = p % 3, r = e % N;
    const int y = r / W, x = r % W;
    const TransformSlot& t = slots[k];
    pix[p * N + x * H + y] =
      t.ok ? opsin[c * plane + static_cast<size_t>(t.by * 8 + y) * width + t.bx * 8 + x] : 0.0f;
  }
  __syncthreads();
  ForwardRows<W, H, 3 * K>(bases + BasisOffsetT(W), pix, tmp);
  __syncthreads();
  ForwardCols<W, H, 3 * K>(bases + BasisOffsetT(H), tmp, [&](int p, int i, float v) {
    const TransformSlot& t = slots[p / 3];
    if (t.ok) coefs[(p % 3) * store + t.offset + i] = v;
  });
}

// AdjustQuantField (src/codec/adaptive_quantization.cpp:740-835), one
// thread per anchor.
__global__ __launch_bounds__(256) void AdjustFieldKernel(
  const uint8_t* __restrict__ cells, int bw, int bh, float mixer,
  const float* __restrict__ in, float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= bw) return;
  const uint8_t c = cells[static_cast<size_t>(y) * bw + x];
  if (!(c & 1)) return;
  const int cw = CoveredW(c >> 1), ch = CoveredH(c >> 1);
  float maximum = in[static_cast<size_t>(y) * bw + x];
  float mean = 0.0f;
  for (int dy = 0; dy < ch; ++dy) {
    for (int dx = 0; dx < cw; ++dx) {
      const float v = in[static_cast<size_t>(y + dy) * bw + x + dx];
      maximum = DStdMax(maximum, v);
      mean += v;
    }
  }
  const int n = cw * ch;
  mean /= static_cast<float>(n);
  float result = maximum;
  if (n >= 4) result = maximum * mixer + mean * (1.0f - mixer);
  for (int dy = 0; dy < ch; ++dy)
    for (int dx = 0; dx < cw; ++dx) out[static_cast<size_t>(y + dy) * bw + x + dx] = result;
}

// Non-negative float maxima through their bit patterns (order-free), one
// atomic per work group.
__global__ __launch_bounds__(256) void MatrixStatsKernel(
  const float* __restrict__ opsin, int width, int height, int stride,
  size_t plane, uint32_t* __restrict__ stats) {
  const float* X = opsin;
  const float* Y = opsin + plane;
  const float* B = opsin + 2 * plane;
  float xe = 0.0f, be = 0.0f, ex = 0.0f;
  const size_t n = static_cast<size_t>(width) * height;
  for (size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int x = static_cast<int>(k % width), y = static_cast<int>(k / width);
    if (x < 1 || y < 1) continue;
    const size_t i = static_cast<size_t>(y) * stride + x, l = i - 1, u = i - stride;
    xe = DStdMax(xe, DStdMax(fabsf(X[i] - X[l]), fabsf(X[i] - X[u])));
    const float d = B[i] - Y[i];
    be = DStdMax(be, DStdMax(fabsf(d - (B[l] - Y[l])), fabsf(d - (B[u] - Y[u]))));
    float e = B[i] - 1.2f * Y[i];
    if (e >= 0.0f) {
      e *= fabsf(B[i] - B[l]) + fabsf(B[i] - B[u]);
      ex = DStdMax(ex, e);
    }
  }
  uint32_t v[3] = {__float_as_uint(xe), __float_as_uint(be), __float_as_uint(ex)};
  BlockReduceU32(v, U32Max());
  if (threadIdx.x == 0) {
    atomicMax(&stats[0], v[0]);
    atomicMax(&stats[1], v[1]);
    atomicMax(&stats[2], v[2]);
  }
}

// min / max of a positive field via bit patterns, one atomic pair per group.
__global__ __launch_bounds__(256) void RangeKernel(const float* __restrict__ v,
                                                   size_t n,
                                                   uint32_t* __restrict__ range) {
  uint32_t lo[1] = {0xffffffffu}, hi[1] = {0u};
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const uint32_t b = __float_as_uint(v[i]);
    lo[0] = b < lo[0] ? b : lo[0];
    hi[0] = b > hi[0] ? b : hi[0];
  }
  BlockReduceU32(lo, U32Min());
  __syncthreads();
  BlockReduceU32(hi, U32Max());
  if (threadIdx.x == 0) {
    atomicMin(&range[0], lo[0]);
    atomicMax(&range[1], hi[0]);
  }
}

// ---- exact median / MAD (radix select on non-negative float bits) ----------

// state: [0] prefix, [1] k (rank within the prefix), [2] median bits,
// [3..] unused; hist: 256 counters.
__device__ __forceinline__ uint32_t SelectKey(const float* v, size_t i,
                                              int pass_kind, uint32_t median) {
  if (pass_kind == 0) return __float_as_uint(v[i]);
  return __float_as_uint(fabsf(v[i] - __uint_as_float(median)));
}

__global__ __launch_bounds__(256) void RadixHistogramKernel(
  const float* __restrict__ v, size_t n, int shift, int pass_kind,
  const uint32_t* __restrict__ state, uint32_t* __restrict__ hist) {
  __shared__ uint32_t local[256];
  local[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t prefix = state[0];
  const uint32_t mask = shift == 24 ? 0u : (0xffffffffu << (shift + 8));
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const uint32_t key = SelectKey(v, i, pass_kind, state[2]);
    if ((key & mask) == (prefix & mask)) atomicAdd(&local[(key >> shift) & 255], 1u);
  }
  __syncthreads();
  if (local[threadIdx.x]) atomicAdd(&hist[threadIdx.x], local[threadIdx.x]);
}

__global__ void RadixPickKernel(uint32_t* __restrict__ hist, int shift,
                                int pass_kind, uint32_t* __restrict__ state) {
  if (threadIdx.x != 0) return;
  uint32_t k = state[1];
  uint32_t bin = 0;
  for (; bin < 256; ++bin) {
    if (k < hist[bin]) break;
    k -= hist[bin];
  }
  state[0] |= bin << shift;
  state[1] = k;
  for (int i = 0; i < 256; ++i) hist[i] = 0;
  if (shift == 0 && pass_kind == 0) state[2] = state[0];  // median found
  if (shift == 0 && pass_kind == 1) state[3] = state[0];  // MAD found
}

__global__ void RadixResetKernel(uint32_t* state, uint32_t k, int pass_kind) {
  state[0] = 0;
  state[1] = k;
  (void)pass_kind;
}

// ComputeQuantizerParams + Quantizer::Create (src/codec/quantization.cpp:
// 59-115, core/quantizer.h:33-82), one thread.
__global__ void QuantizerParamsKernel(const uint32_t* __restrict__ state,
                                      float quant_dc,
                                      QuantizerState* __restrict__ out) {
  if (threadIdx.x != 0) return;
  const float median = __uint_as_float(state[2]);
  const float mad = __uint_as_float(state[3]);
  float scale = 65536.0f * (median - mad) / 5.0f;
  scale = scale < 1.0f ? 1.0f : (32768.0f < scale ? 32768.0f : scale);
  int32_t global_scale = static_cast<int32_t>(scale);
  const int32_t scaled_dc =
    static_cast<int32_t>(static_cast<double>(quant_dc * 4096.0f) * 1.6);
  if (global_scale > scaled_dc) global_scale = scaled_dc > 1 ? scaled_dc : 1;
  const float inv_gs = 65536.0f / static_cast<float>(global_scale);
  const float qdv = DStdMin(65536.0f, quant_dc * inv_gs + 0.5f);
  const uint32_t qdc = static_cast<uint32_t>(qdv);
  QuantizerState q;
  q.global_scale = static_cast<uint32_t>(global_scale);
  q.quant_dc = qdc;
  q.scale = static_cast<float>(q.global_scale) / 65536.0f;
  q.inverse_global_scale = 65536.0f / static_cast<float>(q.global_scale);
  const float inv_qdc = q.inverse_global_scale / static_cast<float>(qdc);
  constexpr float kInvDc[3] = {4096.0f, 512.0f, 256.0f};
  for (int c = 0; c < 3; ++c) {
    q.dc_steps[c] = inv_qdc / kInvDc[c];
    q.inverse_dc_steps[c] = kInvDc[c] * q.scale * static_cast<float>(qdc);
  }
  *out = q;
}

__global__ __launch_bounds__(256) void RawQuantKernel(
  const float* __restrict__ field, size_t n, const QuantizerState* __restrict__ q,
  int32_t* __restrict__ raw) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float r = field[i] * q->inverse_global_scale + 0.5f;
  r = r < 1.0f ? 1.0f : (256.0f < r ? 256.0f : r);
  raw[i] = static_cast<int32_t>(r);
}

// ---- final CfL, fast least squares (chroma_from_luma.cpp:543-670, :112-
// 134), one work group per 64x64 colour tile.  The tile's terms (its anchors
// in raster order, each transform's coefficients in storage order) are
// formed by the whole work group kCflChunk at a time in LDS; sixteen threads
// carry the reference's four-lane fma chains (2 maps x {quadratic, linear}
// x 4 lanes) through the chunks in order.

constexpr int kCflTerms = 64 * 64;  // coefficients of one tile
constexpr int kCflChunk = 1024;

__global__ __launch_bounds__(256) void FinalCflKernel(
  const float* __restrict__ coefs, size_t store,
  const uint8_t* __restrict__ cells, const uint32_t* __restrict__ offsets,
  int bw, int bh, const int32_t* __restrict__ raw,
  const QuantizerState* __restrict__ q, const float* __restrict__ tables,
  int8_t* __restrict__ y_to_x, int8_t* __restrict__ y_to_b) {
  // [map][a, b][position in chunk], padded so that the chains read
  // distinct banks
  __shared__ float terms[2][2][kCflChunk + 4];
  __shared__ float lanes[2][2][4];
  __shared__ uint32_t anchor_base[64], anchor_offset[64];
  __shared__ float anchor_qs[64];
  __shared__ uint8_t anchor_strategy[64];
  __shared__ int anchor_inverse[64][2];  // X and B inverse-table offsets
  __shared__ int anchors, total;
  const int tx = blockIdx.x, ty = blockIdx.y;
  const int t = threadIdx.x;
  const float scale128 = q->scale * 128.0f;
  // the tile's anchors in raster order with their term offsets, strategy,
  // coefficient offset and quantisation scale
  if (t < 64) {
    const int bx = tx * 8 + (t & 7), by = ty * 8 + (t >> 3);
    uint32_t n = 0;
    int strategy = 0;
    uint32_t offset = 0;
    float qs = 0.0f;
    if (bx < bw && by < bh) {
      const size_t b = static_cast<size_t>(by) * bw + bx;
      const uint8_t c = cells[b];
      const int32_t r = raw[b];
      offset = offsets[b];
      strategy = c >> 1;
      qs = scale128 * static_cast<float>(r);
      if (c & 1) n = static_cast<uint32_t>(CoveredW(c >> 1) * CoveredH(c >> 1) * 64);
    }
    uint32_t base = n;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t other = __shfl_up(base, off);
      if (t >= off) base += other;
    }
    base -= n;  // exclusive
    // anchors whose terms fit (all of them in a valid grid) form a prefix
    const bool fits = n != 0 && base + n <= kCflTerms;
    const uint64_t kept = __ballot(fits);
    const int rank = __popcll(kept & ((uint64_t{1} << t) - 1));
    if (fits) {
      anchor_base[rank] = base;
      anchor_offset[rank] = offset;
      anchor_qs[rank] = qs;
      anchor_strategy[rank] = static_cast<uint8_t>(strategy);
      anchor_inverse[rank][0] = InverseOffset(strategy, 0);
      anchor_inverse[rank][1] = InverseOffset(strategy, 2);
    }
    if (t == 63) {
      anchors = __popcll(kept);
      total = static_cast<int>(base + n);
    }
  }
  __syncthreads();
  const int count = anchors, terms_used = min(total, kCflTerms);
  const int problem_t = t / 8, kind_t = (t / 4) % 2, lane_t = t % 4;
  float acc = 0.0f;
  constexpr int kPerThread = kCflChunk / 256;
  for (int first = 0; first < terms_used; first += kCflChunk) {
    const int chunk = min(kCflChunk, terms_used - first);
    // this thread's elements: their anchor (binary search over the LDS
    // list), then every global read, then the terms
    int idx[kPerThread], strategy[kPerThread];
    float qs_e[kPerThread], cy[kPerThread], cx[kPerThread], cb[kPerThread];
    float inv_x[kPerThread], inv_b[kPerThread];
    bool live[kPerThread], llf[kPerThread];
#pragma unroll
    for (int m = 0; m < kPerThread; ++m) {
      const int j = t + m * 256;
      const int p = first + j;
      live[m] = j < chunk && count > 0;
      int lo = 0, hi = count - 1;  // the last anchor starting at or before p
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (anchor_base[mid] <= static_cast<uint32_t>(p)) lo = mid; else hi = mid - 1;
      }
      const int st = anchor_strategy[lo];
      const int cw = CoveredW(st), ch = CoveredH(st);
      const int W = cw * 8, H = ch * 8, N = W * H;
      const int i = p - static_cast<int>(anchor_base[lo]);
      live[m] = live[m] && i < N;
      const int cols = W > H ? W : H;
      const int lw = W > H ? cw : ch, lh = W > H ? ch : cw;
      const int xx = i % cols, yy = i / cols;
      llf[m] = xx < lw && yy < lh;
      const size_t off = anchor_offset[lo];
      const int ii = live[m] ? i : 0;
      idx[m] = ii;
      strategy[m] = st;
      qs_e[m] = anchor_qs[lo];
      // LLF coefficients count as zero (their reads are kept in bounds and
      // zeroed only when the terms are formed, after every read is issued)
      cy[m] = coefs[store + off + ii];
      cx[m] = coefs[off + ii];
      cb[m] = coefs[2 * store + off + ii];
      inv_x[m] = tables[anchor_inverse[lo][0] + ii];
      inv_b[m] = tables[anchor_inverse[lo][1] + ii];
    }
#pragma unroll
    for (int m = 0; m < kPerThread; ++m) {
      if (!live[m]) continue;
      const int j = t + m * 256;
      const float y = llf[m] ? 0.0f : cy[m];
#pragma unroll
      for (int problem = 0; problem < 2; ++problem) {
        const float inv = problem == 0 ? inv_x[m] : inv_b[m];
        const float cc = llf[m] ? 0.0f : problem == 0 ? cx[m] : cb[m];
        const float luma = y * inv * qs_e[m];
        const float chroma = cc * inv * qs_e[m];
        terms[problem][0][j] = luma / 84.0f;
        terms[problem][1][j] = (problem == 0 ? 0.0f : 1.0f) * luma - chroma;
      }
    }
    __syncthreads();
    if (t < 16) {
      const float* a = terms[problem_t][0];
      const float* b = terms[problem_t][kind_t == 0 ? 0 : 1];
      // chunk is a multiple of 64: batches of 16 terms per lane, each
      // batch's loads issued before its ordered fmas
      for (int base = 0; base < chunk; base += 64) {
        float va[16], vb[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) {
          va[m] = a[base + lane_t + 4 * m];
          vb[m] = b[base + lane_t + 4 * m];
        }
#pragma unroll
        for (int m = 0; m < 16; ++m) acc = __builtin_fmaf(va[m], vb[m], acc);
      }
    }
    __syncthreads();
  }
  if (t < 16) lanes[problem_t][kind_t][lane_t] = acc;
  __syncthreads();
  if (t < 2) {
    const float (*l)[4] = lanes[t];
    const float quad = (l[0][0] + l[0][1]) + (l[0][2] + l[0][3]);
    const float lin = (l[1][0] + l[1][1]) + (l[1][2] + l[1][3]);
    int8_t r = 0;
    if (total > 0) {
      const float x = -lin / (quad + static_cast<float>(total) * 1.0e-9f * 0.5f);
      float v = x;
      if (v >= 2.6f) v -= 2.6f;
      else if (v <= -2.6f) v += 2.6f;
      else v = 0.0f;
      float rr = roundf(v);
      rr = rr < -128.0f ? -128.0f : (127.0f < rr ? 127.0f : rr);
      r = static_cast<int8_t>(rr);
    }
    const int tile = ty * ((bw + 7) / 8) + tx;
    if (t == 0) y_to_x[tile] = r; else y_to_b[tile] = r;
  }
}

// ---- coefficient coding, one wave per transform ---------------------------

struct CodingParams {
  float x_mult;
  float b_mult;
  DcBasisTable dc_basis;
};

// AdjustQuantForChannel (src/codec/quantization.cpp:292-533) in three
// parts: the quadrant thresholds, the per-coefficient terms of its scan
// (formed in parallel: the quantisation error and |quantized|, zero in the
// LLF corner the scan skips), and the ordered sums of the scan with the
// decision that follows (one thread per channel, in the scan's raster order;
// adding a zero term to these non-negative float sums changes nothing).

// Quadrant thresholds before the scan.
__device__ __forceinline__ void AdjustThresholds(int cols, int rows, float thr[4]) {
  thr[0] = 0.58f;
  thr[1] = thr[2] = thr[3] = 0.64f;
  const int xsize = cols / 8, ysize = rows / 8;
  if (xsize > 1 || ysize > 1) {
    float reduction = 0.003f * static_cast<float>(xsize * ysize);
    reduction = reduction < 0.0f ? 0.0f : (0.08f < reduction ? 0.08f : reduction);
    for (int i = 0; i < 4; ++i) thr[i] = DStdMax(0.54f, thr[i] - reduction);
  }
}

__device__ __forceinline__ float AdjustQuantized(float value, float threshold) {
  return fabsf(value) < threshold ? 0.0f : rintf(value);
}

// The scan's running sums are independent chains: the error and |quantized|
// sums over all terms and, per quadrant, the |quantized| sum and (Y) the
// largest error of a zero coefficient, each over its terms in the scan's
// raster order.  Chain `role` of one channel: 0 error sum, 1 |quantized|
// sum, 2 + q quadrant q's sum and maximum.
constexpr int kScanChains = 6;

template <int cols, int rows>
__device__ void ScanChain(const float2* __restrict__ terms, int role, int channel,
                          float* sum_out, float* max_out) {
  if (role < 2) {
    // .x (error) or .y (|quantized|) of every term, in raster order; the
    // loads of a batch issue before its ordered adds
    const float* src = reinterpret_cast<const float*>(terms) + role;
    float sum = 0.0f;
    for (int base = 0; base < cols * rows; base += 16) {
      float v[16];
#pragma unroll
      for (int m = 0; m < 16; ++m) v[m] = src[2 * (base + m)];
#pragma unroll
      for (int m = 0; m < 16; ++m) sum += v[m];
    }
    *sum_out = sum;
    return;
  }
  const int qd = role - 2, yh = qd / 2, xh = qd % 2;
  float sum = 0.0f, mx = 0.0f;
#pragma unroll 1
  for (int y = yh * rows / 2; y < (yh + 1) * rows / 2; ++y) {
    const float2* row = terms + y * cols + xh * (cols / 2);
#pragma unroll
    for (int x = 0; x < cols / 2; ++x) {
      const float2 t = row[x];
      sum += t.y;
      if (channel == 1) mx = DStdMax(mx, t.y == 0.0f ? t.x : 0.0f);
    }
  }
  *sum_out = sum;
  *max_out = mx;
}

// The decision of one channel from its scan sums.  coef and inv give the
// border terms again.
template <int cols, int rows>
__device__ void AdjustQuantDecide(float err_sum, float val_sum, const float (&nz)[4],
                                  const float (&mx)[4], const float* __restrict__ coef,
                                  const float* __restrict__ inv, float qac,
                                  float mult, int channel, int strategy, float thr[4],
                                  int32_t* raw_quant) {
  constexpr int xsize = cols / 8, ysize = rows / 8;
  // border terms in raster order: x = cols - 1 on rows [4 ysize, 7 ysize),
  // x >= 7 xsize on rows [7 ysize, rows - 1), x >= 4 xsize on the last row
  float border = 0.0f;
  const float thr_q[4] = {thr[0], thr[1], thr[2], thr[3]};
  auto add_border = [&](int y, int x) {
    const int index = y * cols + x;
    const bool lower = y >= rows / 2, right = x >= cols / 2;
    const float threshold = lower ? (right ? thr_q[3] : thr_q[2]) : (right ? thr_q[1] : thr_q[0]);
    const float value = coef[index] * (inv[index] * qac * mult);
    if (AdjustQuantized(value, threshold) != 0.0f) border += fabsf(value);
  };
  for (int y = 4 * ysize; y < 7 * ysize; ++y) add_border(y, cols - 1);
  for (int y = 7 * ysize; y < rows - 1; ++y)
    for (int x = 7 * xsize; x < cols; ++x) add_border(y, x);
  for (int x = 4 * xsize; x < cols; ++x) add_border(rows - 1, x);

  if (channel == 1 && val_sum * 8.0f < static_cast<float>(xsize * ysize)) {
    const double limit = 0.46, multiplier = 0.9999;
    const int32_t original = *raw_quant;
    int32_t new_quant = *raw_quant;
    for (int qd = 1; qd < 4; ++qd) {
      if (nz[qd] == 0.0f && mx[qd] > limit) {
        new_quant = original + 1;
        break;
      }
    }
    *raw_quant = new_quant;
    if (nz[3] == 0.0f && mx[3] > limit) {
      thr[3] = static_cast<float>(multiplier * mx[3] * new_quant / original);
    } else if ((nz[1] == 0.0f && mx[1] > limit) || (nz[2] == 0.0f && mx[2] > limit)) {
      thr[1] = static_cast<float>(multiplier * DStdMax(mx[1], mx[2]) * new_quant / original);
      thr[2] = thr[1];
    } else if (nz[0] == 0.0f && mx[0] > limit) {
      thr[0] = static_cast<float>(multiplier * mx[0] * new_quant / original);
    }
  }
  const float all_nz = nz[0] + nz[1] + nz[2] + nz[3] + 1.0f;
  const float border_weight = channel == 0 ? 70.0f : channel == 1 ? 30.0f : 60.0f;
  if (border_weight * border >= all_nz) {
    *raw_quant = static_cast<int32_t>(static_cast<float>(*raw_quant) +
                                      border_weight * border / all_nz);
    if (*raw_quant >= 256) *raw_quant = 255;
  }
  if (strategy == kDct8 && nz[0] + nz[1] + nz[2] + nz[3] < 11.0f) {
    ++*raw_quant;
    if (*raw_quant >= 256) *raw_quant = 255;
  }
  constexpr double kFirst[4][3] = {
    {0.22080615753848404, 0.45797479824262011, 0.29859235095977965},
    {0.70109486510286834, 0.16185281305512639, 0.14387691730035473},
    {0.114985964456218638, 0.44656840441027695, 0.10587658215149048},
    {0.46849665264409396, 0.41239077937781954, 0.088667407767185444},
  };
  constexpr double kSecond[4][3] = {
    {0.27450281941822197, 1.1255766549984996, 0.98950459134128388},
    {0.4652168675598285, 0.40945807983455818, 0.36581899811751367},
    {0.28034972424715715, 0.9182653201929738, 1.5581531543057416},
    {0.26873118114033728, 0.68863712390392484, 1.2082185408666786},
  };
  err_sum = static_cast<float>(static_cast<double>(err_sum) * 2.2942708343284721);
  val_sum = static_cast<float>(static_cast<double>(val_sum) * 2.2942708343284721);
  if (strategy >= kDct16x16) {
    int cls = 3;
    if (strategy == kDct32x16 || strategy == kDct16x32) cls = 1;
    else if (strategy == kDct16x16) cls = 0;
    else if (strategy == kDct32x32) cls = 2;
    const double threshold =
      kFirst[cls][channel] * xsize * ysize * 8 * 8 + kSecond[cls][channel] * val_sum;
    int32_t step = static_cast<int32_t>(err_sum / threshold);
    step = step < 0 ? 0 : (step > 2 ? 2 : step);
    if (err_sum > threshold) {
      *raw_quant += step;
      if (*raw_quant >= 256) *raw_quant = 255;
    }
  }
  const int32_t divisor = xsize * ysize;
  float min_nz = nz[0];
  for (int qd = 1; qd < 4; ++qd) if (nz[qd] < min_nz) min_nz = nz[qd];
  int32_t activity = 15;
  if (min_nz < 15.0f * divisor) activity = (static_cast<int32_t>(min_nz) + divisor / 2) / divisor;
  int32_t adjusted = *raw_quant - activity;
  if (channel == 1) for (int qd = 1; qd < 4; ++qd) thr[qd] += 0.01f * activity;
  const int32_t floor_quant = max(4, *raw_quant / 2);
  if (adjusted < floor_quant) adjusted = floor_quant;
  *raw_quant = adjusted;
}

// One launch per strategy: kK transforms per work group.  All threads form
// the terms of the AdjustQuantForChannel scans, three threads per transform
// run their ordered sums and decisions, one thread per transform decides the
// shared raw quant and converts the DC values, and all threads quantise.
template <int kStrategy>
__global__ __launch_bounds__(kAqThreads) void CodeTransformKernel(
  const float* __restrict__ coefs, size_t store, const uint32_t* __restrict__ entries,
  uint32_t count, const uint32_t* __restrict__ offsets, int bw,
  int32_t* __restrict__ raw, const int8_t* __restrict__ y_to_x,
  const int8_t* __restrict__ y_to_b, const QuantizerState* __restrict__ qstate,
  const float* __restrict__ tables, CodingParams params,
  int32_t* __restrict__ quantized, float* __restrict__ dc, size_t dc_plane) {
  using S = AnchorShape<kStrategy>;
  constexpr int W = S::kW, H = S::kH, N = S::kN, K = S::kK;
  constexpr int cols = W > H ? W : H, rows = W > H ? H : W;
  __shared__ float coef[K][3][N];
  __shared__ float2 terms[K][3][N];
  __shared__ float chain_sum[K][3][kScanChains];
  __shared__ float chain_max[K][3][4];
  __shared__ float thr_y[K][4];
  __shared__ int32_t cand[K][3];
  __shared__ int32_t decided[K];
  __shared__ float cfl[K][2];
  __shared__ TransformSlot slots[K];
  __shared__ int32_t raw_k[K];
  LoadSlots<K>(entries, count, offsets, bw, slots);
  __syncthreads();
  const QuantizerState q = *qstate;
  // every global read that depends only on the slots, issued together: the
  // coefficients, the inverse-quantisation entries of the same elements and
  // the raw quant of each transform.  The reads are unconditional (an unused
  // slot reads element 0 and discards it): a read under a branch makes the
  // compiler wait for it, and for every earlier read, at the branch.
  constexpr int kPer = (K * 3 * N + kAqThreads - 1) / kAqThreads;
  float cv[kPer], invv[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int e = threadIdx.x + j * kAqThreads;
    const bool in_range = e < K * 3 * N;
    const int ec = in_range ? e : 0;
    const int k = ec / (3 * N), c = (ec / N) % 3, i = ec % N;
    const TransformSlot& t = slots[k];
    cv[j] = coefs[in_range && t.ok ? c * store + t.offset + i : 0];
    invv[j] = tables[InverseOffset(kStrategy, c) + i];
  }
  {
    const TransformSlot& t = slots[threadIdx.x < K ? threadIdx.x : 0];
    const int32_t r = raw[t.ok ? static_cast<size_t>(t.by) * bw + t.bx : 0];
    if (threadIdx.x < K) raw_k[threadIdx.x] = t.ok ? r : 0;  // after the other reads
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int e = threadIdx.x + j * kAqThreads;
    if (e < K * 3 * N) (&coef[0][0][0])[e] = slots[e / (3 * N)].ok ? cv[j] : 0.0f;
  }
  __syncthreads();
  // terms of the AdjustQuantForChannel scans
  {
    float thr0[4];
    AdjustThresholds(cols, rows, thr0);
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int e = threadIdx.x + j * kAqThreads;
      if (e >= K * 3 * N) continue;
      const int k = e / (3 * N), c = (e / N) % 3, i = e % N;
      const TransformSlot& t = slots[k];
      const int x = i % cols, y = i / cols;
      float2 term = make_float2(0.0f, 0.0f);
      if (t.ok && !(x < cols / 8 && y < rows / 8)) {
        const float qac = q.scale * static_cast<float>(raw_k[k]);
        // thresholds 1-3 start equal; quadrant 0 has its own
        const bool q0 = y < rows / 2 && x < cols / 2;
        const float mult = c == 0 ? params.x_mult : c == 2 ? params.b_mult : 1.0f;
        const float value = cv[j] * (invv[j] * qac * mult);
        const float quantized = AdjustQuantized(value, q0 ? thr0[0] : thr0[1]);
        term = make_float2(fabsf(value - quantized), fabsf(quantized));
      }
      terms[k][c][i] = term;
    }
  }
  __syncthreads();
  // the scan's chains, one thread each
  for (int item = threadIdx.x; item < K * 3 * kScanChains; item += kAqThreads) {
    const int k = item / (3 * kScanChains), c = (item / kScanChains) % 3;
    const int role = item % kScanChains;
    if (slots[k].ok) {
      ScanChain<cols, rows>(terms[k][c], role, c, &chain_sum[k][c][role],
                            &chain_max[k][c][role < 2 ? 0 : role - 2]);
    }
  }
  __syncthreads();
  if (threadIdx.x < 3 * K) {
    const int k = threadIdx.x / 3, c = threadIdx.x % 3;
    const TransformSlot& t = slots[k];
    if (t.ok) {
      float thr[4];
      AdjustThresholds(cols, rows, thr);
      int32_t r = raw_k[k];
      const float qac = q.scale * static_cast<float>(r);
      const float nz[4] = {chain_sum[k][c][2], chain_sum[k][c][3], chain_sum[k][c][4],
                           chain_sum[k][c][5]};
      const float mx[4] = {c == 1 ? chain_max[k][c][0] : 0.0f, c == 1 ? chain_max[k][c][1] : 0.0f,
                           c == 1 ? chain_max[k][c][2] : 0.0f, c == 1 ? chain_max[k][c][3] : 0.0f};
      AdjustQuantDecide<cols, rows>(chain_sum[k][c][0], chain_sum[k][c][1], nz, mx, coef[k][c],
                                    tables + InverseOffset(kStrategy, c), qac,
                                    c == 0 ? params.x_mult : c == 2 ? params.b_mult : 1.0f,
                                    c, kStrategy, thr, &r);
      cand[k][c] = r;
      if (c == 1) {
        for (int i = 0; i < 4; ++i) thr_y[k][i] = thr[i];
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int k = threadIdx.x;
    const TransformSlot& t = slots[k];
    if (t.ok) {
      // max over Y, X, B (starting from zero) as SelectAdjustedAcQuantization
      int32_t d = 0;
      const int order[3] = {1, 0, 2};
      for (int i = 0; i < 3; ++i) d = cand[k][order[i]] > d ? cand[k][order[i]] : d;
      decided[k] = d;
      raw[static_cast<size_t>(t.by) * bw + t.bx] = d;
      const int tile = (t.by / 8) * ((bw + 7) / 8) + t.bx / 8;
      cfl[k][0] = static_cast<float>(y_to_x[tile]) / 84.0f;
      cfl[k][1] = 1.0f + static_cast<float>(y_to_b[tile]) / 84.0f;
      // Y DC from the unmodified Y coefficients
      float dcv[16];
      LlfToDc(params.dc_basis, S::kCw, S::kCh, coef[k][1], dcv);
      for (int dy = 0; dy < S::kCh; ++dy)
        for (int dx = 0; dx < S::kCw; ++dx)
          dc[dc_plane + static_cast<size_t>(t.by + dy) * bw + t.bx + dx] = dcv[dy * S::kCw + dx];
    }
  }
  __syncthreads();
  // Y: quantise with the adjusted thresholds, store, dequantise in place
  {
    const float* inv = tables + InverseOffset(kStrategy, 1);
    const float* dq = tables + DequantOffset(kStrategy, 1);
    for (int e = threadIdx.x; e < K * N; e += kAqThreads) {
      const int k = e / N, i = e % N;
      const TransformSlot& t = slots[k];
      if (!t.ok) continue;
      const int32_t d = decided[k];
      const float qscale = q.scale * static_cast<float>(d) * 1.0f;
      const float dscale = q.inverse_global_scale / static_cast<float>(d) / 1.0f;
      const int x = i % cols, y = i / cols;
      const int tq = (y >= rows / 2) * 2 + (x >= cols / 2);
      const float value = inv[i] * qscale * coef[k][1][i];
      int32_t qv = 0;
      if (!(fabsf(value) < thr_y[k][tq])) qv = static_cast<int32_t>(rint(static_cast<double>(value)));
      quantized[store + t.offset + i] = qv;
      coef[k][1][i] = QuantBias(qv, 1) * (dq[i] * dscale);
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < K * N; e += kAqThreads) {
    const int k = e / N, i = e % N;
    coef[k][0][i] -= cfl[k][0] * coef[k][1][i];
    coef[k][2][i] -= cfl[k][1] * coef[k][1][i];
  }
  __syncthreads();
  if (threadIdx.x < 2 * K) {
    const int k = threadIdx.x / 2, c = (threadIdx.x % 2) * 2;
    const TransformSlot& t = slots[k];
    if (t.ok) {
      float dcv[16];
      LlfToDc(params.dc_basis, S::kCw, S::kCh, coef[k][c], dcv);
      for (int dy = 0; dy < S::kCh; ++dy)
        for (int dx = 0; dx < S::kCw; ++dx)
          dc[c * dc_plane + static_cast<size_t>(t.by + dy) * bw + t.bx + dx] =
            dcv[dy * S::kCw + dx];
    }
  }
  constexpr int covered = S::kCw * S::kCh;
  for (int e = threadIdx.x; e < K * 2 * N; e += kAqThreads) {
    const int k = e / (2 * N), c = ((e / N) % 2) * 2, i = e % N;
    const TransformSlot& t = slots[k];
    if (!t.ok) continue;
    // quadrant 0 starts at 0.58, quadrants 1-3 at 0.62
    const float thr0 = covered >= 4 ? DStdMax(0.5f, 0.58f - 0.00744f * static_cast<float>(covered))
                                    : 0.58f;
    const float thr1 = covered >= 4 ? DStdMax(0.5f, 0.62f - 0.00744f * static_cast<float>(covered))
                                    : 0.62f;
    const float* inv = tables + InverseOffset(kStrategy, c);
    const float qscale = q.scale * static_cast<float>(decided[k]) *
                         (c == 0 ? params.x_mult : params.b_mult);
    const int x = i % cols, y = i / cols;
    const int tq = (y >= rows / 2) * 2 + (x >= cols / 2);
    const float value = inv[i] * qscale * coef[k][c][i];
    int32_t qv = 0;
    if (!(fabsf(value) < (tq == 0 ? thr0 : thr1))) qv = static_cast<int32_t>(rint(static_cast<double>(value)));
    quantized[c * store + t.offset + i] = qv;
  }
}

// ---- prediction-aware DC with the weighted predictor ----------------------
//
// src/codec/vardct/dc_quantization.cpp:121-240 with
// src/codec/modular/prediction.h:59-153.  The predictor's two-row buffers
// are unrolled into full per-group planes: the "north" error it reads at
// (x, y) is e(x, y-1) + e(x-1, y) because the update of (x-1, y) adds its
// errors into the previous row's slot x; "north-west" is
// e(x-1, y-1) + e(x-2, y); "north-east" e(x+1, y-1).  Every input of (x, y)
// is therefore final once (x+1, y-1) and (x-1, y) are.
//
// A work group walks the wavefront of one DC group (<= 256 x 256 blocks) for
// two channels.  Each channel has two waves; lane l of wave h owns rows
// 64h + l and 128 + 64h + l, so the row above a lane's row is the previous
// lane's, except on lane 0.  Row y processes sample x at step
// off(y) + x with off(y) = 2y + D * (y / 64) (+ D for B, below): sample
// (x+1, y-1) is one step old when (x, y) needs it.  Within a wave, lane l
// takes the outputs of lane l - 1's last sample by a lane shift of the
// registers (DPP wave_shr:1); its north and north-west cells are the
// north-east cells of the two previous steps, and north-north is the north
// of its north.  Lane 0 takes them from a ring in LDS that lane 63 of the
// other wave of its channel writes; the extra D steps of lag at every band
// of 64 rows make that data at least D + 1 steps old, so a barrier every D
// steps orders it.  B subtracts the quantized Y of the same sample, which
// the Y wave of its row wrote into a per-row LDS ring D steps earlier.

constexpr int kDcGroup = 256;
constexpr int kDcBatch = 8;            // steps per barrier (D)
constexpr int kDcEdgeRing = 32;        // > 2D + 2 slots of a band edge
constexpr int kDcYRing = 16;           // >= 2D slots of quantized Y per row
constexpr int kDcRowThreads = kDcGroup / 2;
constexpr int kDcWorkGroup = 2 * kDcRowThreads;

// (1 << 24) / (i + 1) for the arguments the predictor forms: e >> shift
// is below 64 and the final weight sum minus one below 32, so one table of
// 128 entries covers both (filled per work group).
constexpr int kReciprocals = 128;

// Bit width of e + 1 for a 32-bit e (33 for 2^32 - 1).
__device__ __forceinline__ int BitWidthPlusOne(uint32_t e) {
  return e == 0xffffffffu ? 33 : 32 - __clz(static_cast<int>(e + 1));
}
__device__ __forceinline__ int BitWidth32(uint32_t v) {
  return 32 - __clz(static_cast<int>(v));
}

// What the row below reads of a sample (x, y): its quantized value,
// prediction error and sub-predictor errors, and q(x, y - 1).
struct DcCell {
  int32_t q = 0, err = 0;
  uint32_t e[4] = {0, 0, 0, 0};
  int32_t q_up = 0;
};

// The predictor state a row keeps in registers: its own last two samples.
struct DcRowState {
  int32_t q1 = 0, err1 = 0;
  uint32_t e1[4] = {0, 0, 0, 0}, e2[4] = {0, 0, 0, 0};
};

// Bounds of the 32-bit evaluation: while every quantized value stays below
// 2^18 and every stored prediction error below 2^21 in magnitude, each sum
// of three errors fits the signed 24-bit multiplier, each sub-prediction is
// below 2^23, the weighted sum below 2^28 and the prediction below 2^24, so
// 32-bit arithmetic gives the 64-bit results.
constexpr int32_t kDcNarrowQ = 1 << 18;
constexpr int32_t kDcNarrowErr = 1 << 21;

// One sample (x, y) of one channel from its north-east, north and
// north-west cells (the north-north value is n.q_up) and the row state:
// weighted prediction, quantisation of the residual, update of the row
// state.  Returns the sample's cell.  Row 0 reads all-zero north cells, so
// the reference's zero errors and zero "up" value above the first row need
// no selects; the row state's errors read zero at x = 0 and x = 1 because
// every step off the row stores zero errors (every row starts at least two
// steps after the thread's last sample).  The sample is evaluated on every
// step; off the row (`active` false) its result is never read.  kWide
// evaluates the predictor in 64-bit integers as the reference does;
// otherwise in 32-bit integers, exact while the inputs keep the bounds
// above, and `overflow` is set when an active sample's outputs leave them
// (later samples would then read out-of-bound inputs).  Cells of absent
// neighbours hold values that the selects discard.
template <bool kWide>
__device__ __forceinline__ DcCell DcSample(const DcCell& ne, const DcCell& n, const DcCell& nw,
                                           DcRowState& st, const uint32_t* reciprocal, int x,
                                           int y, bool has_ne, bool active, float value,
                                           float inv, bool& overflow) {
  using I = std::conditional_t<kWide, int64_t, int32_t>;
  using U = std::conditional_t<kWide, uint64_t, uint32_t>;
  constexpr uint8_t kCoeff[7] = {16, 10, 7, 7, 7, 0, 0};
  constexpr uint8_t kMaxWeights[4] = {13, 12, 12, 12};
  // quantized neighbours
  const I up = n.q;
  const I W_ = x ? st.q1 : up;
  const I N_ = y ? up : W_;
  const I NW_ = x && y ? nw.q : W_;
  const I NE_ = y && has_ne ? ne.q : N_;
  const I NN_ = y > 1 ? n.q_up : N_;
  // Predict (prediction.h:75-123)
  uint32_t weights[4];
  uint32_t weight_sum = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t north = n.e[i] + st.e1[i];
    const uint32_t e_ne = has_ne ? ne.e[i] : north;
    const uint32_t e_nw = x ? nw.e[i] + st.e2[i] : north;
    const uint32_t e = north + e_ne + e_nw;
    int shift = BitWidthPlusOne(e) - 6;
    if (shift < 0) shift = 0;
    // k * r as k * (r - 1) + k: r - 1 < 2^24 fits the 24-bit multiplier
    const uint32_t k = kMaxWeights[i];
    weights[i] = 4 + ((__umul24(reciprocal[e >> shift] - 1, k) + k) >> shift);
    weight_sum += weights[i];
  }
  const unsigned log_weight = BitWidth32(weight_sum) - 1;
  weight_sum = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    weights[i] >>= log_weight - 4;
    weight_sum += weights[i];
  }
  const I ew = st.err1;
  const I en = n.err;
  const I enw = x ? nw.err : en;
  const I ene = has_ne ? ne.err : en;
  const I n8 = N_ * 8, w8 = W_ * 8, ne8 = NE_ * 8, nw8 = NW_ * 8, nn8 = NN_ * 8;
  // the reference's sums of products, with the three equal error
  // coefficients of the last one factored (the same integer)
  static_assert(kCoeff[2] == kCoeff[3] && kCoeff[3] == kCoeff[4], "factored coefficients");
  I c1, c2, c3;
  if constexpr (kWide) {
    c1 = (en + ew + ene) * kCoeff[0];
    c2 = (en + ew + enw) * kCoeff[1];
    c3 = (enw + en + ene) * kCoeff[2];
  } else {
    c1 = __mul24(en + ew + ene, static_cast<int32_t>(kCoeff[0]));
    c2 = __mul24(en + ew + enw, static_cast<int32_t>(kCoeff[1]));
    c3 = __mul24(enw + en + ene, static_cast<int32_t>(kCoeff[2]));
  }
  I pred[4];
  pred[0] = w8 + ne8 - n8;
  pred[1] = n8 - (c1 >> 5);
  pred[2] = w8 - (c2 >> 5);
  pred[3] = n8 - ((c3 + (nn8 - n8) * kCoeff[5] + (nw8 - w8) * kCoeff[6]) >> 5);
  I sum = static_cast<I>((weight_sum >> 1) - 1);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    // within the 32-bit bounds |pred| < 2^23 fits the signed 24-bit multiplier
    if constexpr (kWide) {
      sum += pred[i] * static_cast<I>(weights[i]);
    } else {
      sum += __mul24(pred[i], static_cast<int32_t>(weights[i]));
    }
  }
  // the table entry is below 2^24: a signed 32-bit operand
  I prediction = static_cast<I>(
    (static_cast<int64_t>(sum) *
     static_cast<int64_t>(static_cast<int32_t>(reciprocal[weight_sum - 1]))) >> 24);
  if (((en ^ ew) | (en ^ enw)) <= 0) {
    I lo = w8 < ne8 ? w8 : ne8;
    lo = lo < n8 ? lo : n8;
    I hi = w8 > ne8 ? w8 : ne8;
    hi = hi > n8 ? hi : n8;
    prediction = prediction < lo ? lo : (prediction > hi ? hi : prediction);
  }
  const I predicted = (prediction + 3) >> 3;
  // quantise the residual (dc_quantization.cpp:176-203)
  float residual = value * inv;
  residual -= static_cast<float>(predicted);
  if (residual > -0.62f && residual < 0.62f) residual = 0;
  float rounded = roundf(residual);
  if (rounded > 2 || rounded < -2) rounded = roundf(residual * 0.5f) * 2;
  // rounded holds an integer
  const int32_t ir = static_cast<int32_t>(rounded);
  DcCell out;
  // the low 32 bits of predicted + ir
  out.q = static_cast<int32_t>(static_cast<uint32_t>(predicted) + static_cast<uint32_t>(ir));
  out.q_up = n.q;
  // Update (prediction.h:124-144); unsigned arithmetic wraps where a 32-bit
  // evaluation leaves its bounds (the sample is then flagged)
  const I scaled = static_cast<I>(static_cast<U>(static_cast<I>(out.q)) << 3);
  out.err = static_cast<int32_t>(static_cast<U>(prediction) - static_cast<U>(scaled));
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if constexpr (kWide) {
      I dd = static_cast<I>(static_cast<U>(pred[i]) - static_cast<U>(scaled));
      const U magnitude = dd < 0 ? U(0) - static_cast<U>(dd) : static_cast<U>(dd);
      out.e[i] = static_cast<uint32_t>((magnitude + 3) >> 3);
    } else {
      // |pred - scaled| + 3 on sign-flipped words (the order of signed words
      // as unsigned ones); within the bounds the difference fits 32 bits
      constexpr uint32_t kFlip = 0x80000000u;
      out.e[i] = __usad(static_cast<uint32_t>(pred[i]) ^ kFlip,
                        static_cast<uint32_t>(scaled) ^ kFlip, 3u) >> 3;
    }
    st.e2[i] = st.e1[i];
    st.e1[i] = active ? out.e[i] : 0u;
  }
  st.q1 = out.q;
  st.err1 = active ? out.err : 0;
  if (!kWide) {
    overflow |= active && (static_cast<uint32_t>(out.q) + kDcNarrowQ >= 2u * kDcNarrowQ ||
                           static_cast<uint32_t>(out.err) + kDcNarrowErr >= 2u * kDcNarrowErr);
  }
  return out;
}

// Lane l receives lane l - 1's v, lane 0 receives edge.
__device__ __forceinline__ int32_t FromLaneAbove(int32_t edge, int32_t v) {
  constexpr int kWaveShr1 = 0x138;
  return __builtin_amdgcn_update_dpp(edge, v, kWaveShr1, 0xf, 0xf, false);
}

__device__ __forceinline__ DcCell CellFromLaneAbove(const int32_t (&edge)[8], const DcCell& c) {
  DcCell r;
  r.q = FromLaneAbove(edge[0], c.q);
  r.err = FromLaneAbove(edge[1], c.err);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    r.e[i] = static_cast<uint32_t>(FromLaneAbove(edge[2 + i], static_cast<int32_t>(c.e[i])));
  }
  r.q_up = FromLaneAbove(edge[6], c.q_up);
  return r;
}

struct DcShared {
  uint32_t reciprocal[kReciprocals];
  // [channel slot][writing wave][x % kDcEdgeRing]: lane 63's cells; slot
  // kDcEdgeRing of every ring holds zeros (row 0's north)
  int32_t edge[2][2][kDcEdgeRing + 1][8];
  // quantized Y per row of the group, x % kDcYRing (B's work groups)
  int32_t y_ring[kDcGroup][kDcYRing];
};

// grid: DC groups x 2.  Work group z = 0 quantises Y and X, z = 1 Y again
// (without storing it) and B, which subtracts the sample's quantized Y
// (times the Y step) from its DC first and so runs D steps behind Y (the
// reference quantizes B after Y's whole plane).  Each thread reads its DC
// samples a batch of D steps ahead.  A group is evaluated in 32-bit
// arithmetic first and again in 64-bit arithmetic if any of its samples
// left the 32-bit bounds.
struct DcWavefront {
  const float* src;      // this channel's DC plane
  int32_t* dst;          // this channel's quantized plane
  uint32_t row_a, row_b;  // offsets of rows ya and yb
  int ya, yb, off_a, off_b, w, h, slot, channel, lane, wave;
  int lane0_off_a, lane0_off_b;   // off() of lane 0's two rows
  bool store, write_y;
  float inv, step_y;

  // the row this thread works on at step t (the second from two steps
  // before its first sample) and its column there
  __device__ __forceinline__ void Locate(int t, int& y, int& x) const {
    const bool second = t >= off_b - 2;
    y = second ? yb : ya;
    x = t - (second ? off_b : off_a);
  }
  __device__ __forceinline__ bool Active(int y, int x) const {
    return y < h && x >= 0 && x < w;
  }

  __device__ __forceinline__ void Fetch(float (&out)[kDcBatch], int t0) const {
#pragma unroll
    for (int j = 0; j < kDcBatch; ++j) {
      int y, x;
      Locate(t0 + j, y, x);
      out[j] = Active(y, x) ? src[(y == ya ? row_a : row_b) + x] : 0.0f;
    }
  }

  // The edge ring slot lane 0 reads at step t: the north-east column of
  // its sample, x + 1, or the zero slot on row 0 (wave-uniform).
  __device__ __forceinline__ int EdgeSlot(int t) const {
    const bool second = t >= lane0_off_b - 2;
    const int x = t - (second ? lane0_off_b : lane0_off_a);
    const int slot = wave == 0 && !second ? kDcEdgeRing : (x + 1) & (kDcEdgeRing - 1);
    return __builtin_amdgcn_readfirstlane(slot);
  }

  // Walks the group's wavefront; returns whether a 32-bit evaluation left
  // its bounds on this thread.  kB: this thread quantises B.  A batch of D
  // steps is computed without branches (the scheduler interleaves the
  // steps' independent work with the dependent chain); its stores and ring
  // writes follow it, before the barrier, since every reader of those rings
  // takes data at least D steps old.
  template <bool kWide, bool kB>
  __device__ __forceinline__ bool Run(DcShared& sh, int steps) const {
    float cur[kDcBatch], next[kDcBatch];
    Fetch(cur, 0);
    Fetch(next, kDcBatch);
    DcRowState st;
    DcCell out, n, nw;  // this lane's last sample; north-east cells 1 and 2 steps ago
    int32_t edge[8];
    const int32_t(*ring_in)[8] = sh.edge[slot][wave ^ 1];
    int32_t(*ring_out)[8] = sh.edge[slot][wave];
    {
      const int32_t* e = ring_in[EdgeSlot(0)];
#pragma unroll
      for (int k = 0; k < 8; ++k) edge[k] = e[k];
    }
    bool overflow = false;
    for (int t0 = 0; t0 < steps; t0 += kDcBatch) {
      DcCell cells[kDcBatch];
#pragma unroll
      for (int j = 0; j < kDcBatch; ++j) {
        const int t = t0 + j;
        const DcCell ne = CellFromLaneAbove(edge, out);
        {
          const int32_t* e = ring_in[EdgeSlot(t + 1)];
#pragma unroll
          for (int k = 0; k < 8; ++k) edge[k] = e[k];
        }
        int y, x;
        Locate(t, y, x);
        const bool active = Active(y, x);
        float value = cur[j];
        if constexpr (kB) {
          // (every row and slot index is inside the ring, also off the row)
          const int32_t qy = sh.y_ring[y][x & (kDcYRing - 1)];
          value = __builtin_fmaf(-static_cast<float>(qy), step_y, value);
        }
        out = DcSample<kWide>(ne, n, nw, st, sh.reciprocal, x, y, x + 1 < w, active, value,
                              inv, overflow);
        cells[j] = out;
        nw = n;
        n = ne;
      }
#pragma unroll
      for (int j = 0; j < kDcBatch; ++j) {
        int y, x;
        Locate(t0 + j, y, x);
        const bool active = Active(y, x);
        // Y's values for B: off-row steps write their slots too, which the
        // ring's size keeps away from every slot B still reads
        if (write_y) sh.y_ring[y][x & (kDcYRing - 1)] = cells[j].q;
        if (active && store) dst[(y == ya ? row_a : row_b) + x] = cells[j].q;
        if (active && lane == 63) {
          int32_t* e = ring_out[x & (kDcEdgeRing - 1)];
          e[0] = cells[j].q;
          e[1] = cells[j].err;
#pragma unroll
          for (int i = 0; i < 4; ++i) e[2 + i] = static_cast<int32_t>(cells[j].e[i]);
          e[6] = cells[j].q_up;
        }
      }
      // orders the edge and Y rings between waves
      LdsBarrier();
#pragma unroll
      for (int j = 0; j < kDcBatch; ++j) cur[j] = next[j];
      Fetch(next, t0 + 2 * kDcBatch);
    }
    return overflow;
  }
};

__device__ __forceinline__ int DcOffset(int y, int lag) {
  return 2 * y + kDcBatch * (y >> 6) + lag;
}

__global__ __launch_bounds__(kDcWorkGroup) void DcQuantKernel(
  const float* __restrict__ dc, int bw, int bh,
  const QuantizerState* __restrict__ qstate, int extra_precision,
  int32_t* __restrict__ quantized) {
  __shared__ DcShared sh;
  if (threadIdx.x < kReciprocals) sh.reciprocal[threadIdx.x] = (1u << 24) / (threadIdx.x + 1);
  if (threadIdx.x < 2 * 2 * 8) {
    sh.edge[threadIdx.x / 16][(threadIdx.x / 8) % 2][kDcEdgeRing][threadIdx.x % 8] = 0;
  }
  __syncthreads();
  const int gx0 = blockIdx.x * kDcGroup, gy0 = blockIdx.y * kDcGroup;
  const size_t n = static_cast<size_t>(bw) * bh;
  const bool with_b = blockIdx.z == 1;
  DcWavefront f;
  f.w = min(kDcGroup, bw - gx0);
  f.h = min(kDcGroup, bh - gy0);
  // waves 0-1: Y, waves 2-3: X (z = 0) or B (z = 1); the wave's index
  // (from the first lane) keeps every per-wave quantity scalar
  const int wave_id = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x)) / 64;
  f.slot = wave_id / 2;
  f.channel = f.slot == 0 ? 1 : with_b ? 2 : 0;
  const int lag = f.channel == 2 ? kDcBatch : 0;
  f.store = f.slot == 1 || !with_b;
  f.write_y = f.slot == 0 && with_b;
  const QuantizerState q = *qstate;
  const float precision = static_cast<float>(1u << extra_precision);
  const float inverse_step = f.channel == 0 ? q.inverse_dc_steps[0]
                           : f.channel == 1 ? q.inverse_dc_steps[1] : q.inverse_dc_steps[2];
  f.inv = inverse_step * precision;
  f.step_y = q.dc_steps[1] / precision;
  const int r = threadIdx.x % kDcRowThreads;
  f.wave = wave_id % 2;
  f.lane = r % 64;
  f.ya = r;
  f.yb = r + kDcRowThreads;
  f.off_a = DcOffset(f.ya, lag);
  f.off_b = DcOffset(f.yb, lag);
  f.lane0_off_a = DcOffset(64 * f.wave, lag);
  f.lane0_off_b = DcOffset(64 * f.wave + kDcRowThreads, lag);
  // rows ya and yb, clamped into the group
  // (n < 2^32: the extent is at most 8192 x 8192 blocks)
  f.row_a = static_cast<uint32_t>(gy0 + min(f.ya, f.h - 1)) * bw + gx0;
  f.row_b = static_cast<uint32_t>(gy0 + min(f.yb, f.h - 1)) * bw + gx0;
  f.src = dc + f.channel * n;
  f.dst = quantized + f.channel * n;
  // the last sample of the group, (w - 1, h - 1), of the later channel
  const int steps = DcOffset(f.h - 1, with_b ? kDcBatch : 0) + f.w;
  const bool overflow =
    f.channel == 2 ? f.Run<false, true>(sh, steps) : f.Run<false, false>(sh, steps);
  // the rerun rewrites every sample the first walk stored
  if (__syncthreads_or(overflow)) {
    if (f.channel == 2) {
      f.Run<true, true>(sh, steps);
    } else {
      f.Run<true, false>(sh, steps);
    }
  }
}

__global__ __launch_bounds__(256) void DcReconstructKernel(
  const int32_t* __restrict__ quantized, size_t n,
  const QuantizerState* __restrict__ qstate, int extra_precision,
  float* __restrict__ reconstructed) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float precision = static_cast<float>(1u << extra_precision);
  const QuantizerState q = *qstate;
  const float ry = static_cast<float>(quantized[n + i]) * (q.dc_steps[1] / precision);
  for (int c = 0; c < 3; ++c) {
    float r = static_cast<float>(quantized[c * n + i]) * (q.dc_steps[c] / precision);
    if (c == 2) r = __builtin_fmaf(ry, 1.0f, r);
    reconstructed[c * n + i] = r;
  }
}

// ComputeEpfInverseSigma (src/codec/epf.cpp:207-290), neutral sharpness.
__global__ __launch_bounds__(256) void EpfSigmaKernel(
  const uint8_t* __restrict__ cells, int bw, int bh,
  const int32_t* __restrict__ raw, const QuantizerState* __restrict__ q,
  float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= bw) return;
  const uint8_t c = cells[static_cast<size_t>(y) * bw + x];
  if (!(c & 1)) return;
  const float sigma_quant =
    0.46f / (q->scale * static_cast<float>(raw[static_cast<size_t>(y) * bw + x]) *
             -1.1715728752538099024f);
  float sigma = sigma_quant * (4.0f / 7.0f);
  sigma = DStdMin(-1.0e-4f, sigma);
  const float value = 1.0f / sigma;
  const int cw = CoveredW(c >> 1), ch = CoveredH(c >> 1);
  for (int dy = 0; dy < ch; ++dy)
    for (int dx = 0; dx < cw; ++dx) out[static_cast<size_t>(y + dy) * bw + x + dx] = value;
}

// SmoothDcCoefficients (src/codec/dc_smoothing.cpp).
__global__ __launch_bounds__(256) void SmoothDcKernel(
  const float* __restrict__ dc, int bw, int bh,
  const QuantizerState* __restrict__ qstate, float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= bw) return;
  const size_t n = static_cast<size_t>(bw) * bh;
  const size_t i = static_cast<size_t>(y) * bw + x;
  if (y == 0 || x == 0 || y + 1 >= bh || x + 1 >= bw) {
    for (int c = 0; c < 3; ++c) out[c * n + i] = dc[c * n + i];
    return;
  }
  constexpr float kSide = 0.20345139757231578f, kCorner = 0.0334829185968739f;
  constexpr float kCenter = 1.0f - 4.0f * (kSide + kCorner);
  float center[3], smoothed[3];
  float gap = 0.5f;
  for (int c = 0; c < 3; ++c) {
    const float* top = dc + c * n + static_cast<size_t>(y - 1) * bw;
    const float* row = dc + c * n + static_cast<size_t>(y) * bw;
    const float* bottom = dc + c * n + static_cast<size_t>(y + 1) * bw;
    const float corners = (top[x - 1] + top[x + 1]) + (bottom[x - 1] + bottom[x + 1]);
    const float sides = (row[x - 1] + row[x + 1]) + (top[x] + bottom[x]);
    center[c] = row[x];
    smoothed[c] = __builtin_fmaf(corners, kCorner,
                                 __builtin_fmaf(sides, kSide, row[x] * kCenter));
    gap = DStdMax(gap, fabsf((row[x] - smoothed[c]) / qstate->dc_steps[c]));
  }
  const float factor = DStdMax(0.0f, __builtin_fmaf(-4.0f, gap, 3.0f));
  for (int c = 0; c < 3; ++c) {
    out[c * n + i] = __builtin_fmaf(smoothed[c] - center[c], factor, center[c]);
  }
}

// ReconstructQuantizedCoefficients (src/codec/reconstruction.cpp:790-936),
// one launch per strategy: dequantise, add CfL, restore the LLF from the
// smoothed DC, inverse DCT straight into the padded XYB image.
template <int kStrategy>
__global__ __launch_bounds__(kAqThreads) void ReconstructKernel(
  const int32_t* __restrict__ quantized, size_t store,
  const uint32_t* __restrict__ entries, uint32_t count,
  const uint32_t* __restrict__ offsets, int width, int height,
  const int32_t* __restrict__ raw, const int8_t* __restrict__ y_to_x,
  const int8_t* __restrict__ y_to_b, const QuantizerState* __restrict__ qstate,
  const float* __restrict__ tables, const double* __restrict__ bases,
  CodingParams params, const float* __restrict__ smoothed_dc, float* __restrict__ out) {
  using S = AnchorShape<kStrategy>;
  constexpr int W = S::kW, H = S::kH, N = S::kN, K = S::kK;
  __shared__ float coef[K * 3 * N];
  __shared__ double tmp[K * 3 * H * (W + 1)];
  __shared__ TransformSlot slots[K];
  __shared__ float cfl[K][2];
  __shared__ float scale[K][3];
  const int bw = width / 8, bh = height / 8;
  const double* ih = bases + BasisOffset(W, false);
  const double* iv = bases + BasisOffset(H, false);
  LoadSlots<K>(entries, count, offsets, bw, slots);
  __syncthreads();
  const QuantizerState q = *qstate;
  // every global read that depends only on the slots, issued together:
  // quantized coefficients and dequantisation entries, the raw quant and
  // CfL multipliers, the smoothed DC samples.  The reads are unconditional
  // (an unused slot reads element 0 and discards it): a read under a branch
  // makes the compiler wait for it, and for every earlier read, at the
  // branch.
  constexpr int kPer = (K * 3 * N + kAqThreads - 1) / kAqThreads;
  int32_t qv[kPer];
  float dqv[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int e = threadIdx.x + j * kAqThreads;
    const bool in_range = e < K * 3 * N;
    const int ec = in_range ? e : 0;
    const int k = ec / (3 * N), c = (ec / N) % 3, i = ec % N;
    const TransformSlot& t = slots[k];
    qv[j] = quantized[in_range && t.ok ? c * store + t.offset + i : 0];
    dqv[j] = tables[DequantOffset(kStrategy, c) + i];
  }
  // thread k < K: slot k's raw quant and CfL multipliers
  const TransformSlot& tk = slots[threadIdx.x < K ? threadIdx.x : 0];
  const size_t raw_at = tk.ok ? static_cast<size_t>(tk.by) * bw + tk.bx : 0;
  const int tile = tk.ok ? (tk.by / 8) * ((bw + 7) / 8) + tk.bx / 8 : 0;
  const int32_t r = raw[raw_at];
  const int8_t yx = y_to_x[tile], yb = y_to_b[tile];
  // thread 3k + c < 3K: the DC samples under slot k, channel c
  float dcv[16];
  if (threadIdx.x < 3 * K) {
    const int k = threadIdx.x / 3, c = threadIdx.x % 3;
    const TransformSlot& t = slots[k];
    const size_t n = static_cast<size_t>(bw) * bh;
    const size_t base = t.ok ? c * n + static_cast<size_t>(t.by) * bw + t.bx : 0;
#pragma unroll
    for (int dy = 0; dy < S::kCh; ++dy)
#pragma unroll
      for (int dx = 0; dx < S::kCw; ++dx)
        dcv[dy * S::kCw + dx] = smoothed_dc[t.ok ? base + static_cast<size_t>(dy) * bw + dx : 0];
  }
  if (threadIdx.x < K) {
    const int k = threadIdx.x;
    if (slots[k].ok) {
      const float mults[3] = {params.x_mult, 1.0f, params.b_mult};
      for (int c = 0; c < 3; ++c) scale[k][c] = q.inverse_global_scale / static_cast<float>(r) / mults[c];
      cfl[k][0] = static_cast<float>(yx) / 84.0f;
      cfl[k][1] = 1.0f + static_cast<float>(yb) / 84.0f;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int e = threadIdx.x + j * kAqThreads;
    if (e < K * 3 * N) {
      const int k = e / (3 * N), c = (e / N) % 3;
      coef[e] = slots[k].ok ? QuantBias(qv[j], c) * (dqv[j] * scale[k][c]) : 0.0f;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < K * N; e += kAqThreads) {
    const int k = e / N, i = e % N;
    float* ck = coef + k * 3 * N;
    ck[i] += cfl[k][0] * ck[N + i];
    ck[2 * N + i] += cfl[k][1] * ck[N + i];
  }
  __syncthreads();
  if (threadIdx.x < 3 * K) {
    const int k = threadIdx.x / 3, c = threadIdx.x % 3;
    if (slots[k].ok) DcToLlf(params.dc_basis, S::kCw, S::kCh, dcv, coef + (k * 3 + c) * N);
  }
  __syncthreads();
  InverseRows<W, H, 3 * K>(ih, coef, tmp);
  __syncthreads();
  const size_t plane = static_cast<size_t>(width) * height;
  InverseCols<W, H, 3 * K>(iv, tmp, [&](int p, int y, int x, float v) {
    const TransformSlot& t = slots[p / 3];
    if (t.ok) out[(p % 3) * plane + static_cast<size_t>(t.by * 8 + y) * width + t.bx * 8 + x] = v;
  });
}

DcBasisTable HostDcBasis() {
  DcBasisTable t{};
  const double sqrt2 = 1.4142135623730951, pi = 3.141592653589793;
  const int lengths[3] = {1, 2, 4};
  for (int fwd = 0; fwd < 2; ++fwd) {
    for (int li = 0; li < 3; ++li) {
      const int n = lengths[li];
      for (int f = 0; f < n; ++f) {
        for (int s = 0; s < n; ++s) {
          const double alpha = f == 0 ? sqrt2 / 2.0 : 1.0;
          volatile double angle = (static_cast<double>(s) + 0.5) * static_cast<double>(f) *
                               pi / static_cast<double>(n);
          // index 0: InverseBasis (used by LLF -> DC), 1: ForwardBasis
          t.v[fwd][li][f][s] = fwd == 0 ? sqrt2 * alpha * std::cos(angle)
                                        : sqrt2 * alpha * std::cos(angle) /
                                            static_cast<double>(n);
        }
      }
    }
  }
  return t;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host entry points (used by the C ABI in aq_api.hip).

AqLayout MakeAqLayout(int width, int height) {
  AqLayout l;
  l.bw = width / 8;
  l.bh = height / 8;
  l.groups_w = (l.bw + kGroupBlocks - 1) / kGroupBlocks;
  l.groups_h = (l.bh + kGroupBlocks - 1) / kGroupBlocks;
  l.groups = l.groups_w * l.groups_h;
  l.store = static_cast<size_t>(l.groups) * kGroupCoefficients;
  return l;
}

GJXLResult LaunchAnchorLayout(hipStream_t s, const uint8_t* cells, int width,
                              int height, uint32_t* offsets, uint32_t* list,
                              uint32_t* count, uint32_t* error,
                              uint32_t* strategy_lists, size_t capacity,
                              uint32_t* strategy_counts) {
  const AqLayout l = MakeAqLayout(width, height);
  hipLaunchKernelGGL(FillWordsKernel, dim3(1), dim3(64), 0, s, error, 1, 0u);
  if (strategy_lists != nullptr) {
    hipLaunchKernelGGL(FillWordsKernel, dim3(1), dim3(64), 0, s, strategy_counts,
                       kStrategyStages, 0u);
  }
  hipLaunchKernelGGL(AnchorLayoutKernel, dim3(l.groups), dim3(kGroupAnchors), 0, s, cells,
                     l.bw, l.bh, offsets, list, count, error, strategy_lists, capacity,
                     strategy_counts);
  return CheckLaunch("AnchorLayoutKernel");
}

namespace {

// Calls launch.template operator()<strategy>(entries, count) for every
// strategy with transforms in the lists.
template <typename Launch>
GJXLResult ForEachStrategy(const StrategyLists& lists, Launch&& launch) {
  const uint32_t* e = lists.entries;
  const size_t c = lists.capacity;
  if (lists.counts[0]) GJXL_AMD_TRY(launch.template operator()<kDct8>(e, lists.counts[0]));
  if (lists.counts[1]) GJXL_AMD_TRY(launch.template operator()<kDct16x8>(e + c, lists.counts[1]));
  if (lists.counts[2]) GJXL_AMD_TRY(launch.template operator()<kDct8x16>(e + 2 * c, lists.counts[2]));
  if (lists.counts[3]) GJXL_AMD_TRY(launch.template operator()<kDct16x16>(e + 3 * c, lists.counts[3]));
  if (lists.counts[4]) GJXL_AMD_TRY(launch.template operator()<kDct32x16>(e + 4 * c, lists.counts[4]));
  if (lists.counts[5]) GJXL_AMD_TRY(launch.template operator()<kDct16x32>(e + 5 * c, lists.counts[5]));
  if (lists.counts[6]) GJXL_AMD_TRY(launch.template operator()<kDct32x32>(e + 6 * c, lists.counts[6]));
  return GJXL_OK;
}

template <int kStrategy>
dim3 StrategyGrid(uint32_t count) {
  return dim3((count + AnchorShape<kStrategy>::kK - 1) / AnchorShape<kStrategy>::kK);
}

}  // namespace

GJXLResult LaunchAnchorDct(GJXLAmdDevice* d, const float* opsin, int width,
                           int height, const uint32_t* offsets,
                           const StrategyLists& lists, float* coefs) {
  const AqLayout l = MakeAqLayout(width, height);
  return ForEachStrategy(lists, [&]<int kStrategy>(const uint32_t* entries, uint32_t count) {
    hipLaunchKernelGGL(AnchorDctKernel<kStrategy>, StrategyGrid<kStrategy>(count),
                       dim3(kAqThreads), 0, d->stream, opsin, width, height, entries, count,
                       offsets, d->dct_bases, l.store, coefs);
    return CheckLaunch("AnchorDctKernel");
  });
}

GJXLResult LaunchAdjustField(hipStream_t s, const uint8_t* cells, int width,
                             int height, float target, const float* in,
                             float* out) {
  const AqLayout l = MakeAqLayout(width, height);
  float mixer = 1.0f;
  if (target > 1.54138f) {
    mixer = mixer - (target - 1.54138f) * 0.56391f;
    mixer = mixer < 0.0f ? 0.0f : mixer;
  }
  hipLaunchKernelGGL(AdjustFieldKernel, dim3(CeilDiv(l.bw, 256), l.bh), dim3(256),
                     0, s, cells, l.bw, l.bh, mixer, in, out);
  return CheckLaunch("AdjustFieldKernel");
}

GJXLResult LaunchMatrixStats(hipStream_t s, const float* opsin, int width,
                             int height, int padded_w, int padded_h,
                             uint32_t* stats) {
  hipLaunchKernelGGL(FillWordsKernel, dim3(1), dim3(64), 0, s, stats, 3, 0u);
  hipLaunchKernelGGL(MatrixStatsKernel,
                     dim3(ReductionGroups(static_cast<size_t>(width) * height)), dim3(256), 0,
                     s, opsin, width, height, padded_w,
                     static_cast<size_t>(padded_w) * padded_h, stats);
  return CheckLaunch("MatrixStatsKernel");
}

GJXLResult LaunchRange(hipStream_t s, const float* v, size_t n, uint32_t* range) {
  hipLaunchKernelGGL(FillWordsKernel, dim3(1), dim3(64), 0, s, range, 1, 0xffffffffu);
  hipLaunchKernelGGL(FillWordsKernel, dim3(1), dim3(64), 0, s, range + 1, 1, 0u);
  hipLaunchKernelGGL(RangeKernel, dim3(ReductionGroups(n)), dim3(256),
                     0, s, v, n, range);
  return CheckLaunch("RangeKernel");
}

// select: state (4 words), hist (256 words) of device scratch.
GJXLResult LaunchQuantizer(hipStream_t s, const float* field, size_t n,
                           float quant_dc, uint32_t* select_state,
                           uint32_t* hist, QuantizerState* q, int32_t* raw) {
  hipLaunchKernelGGL(FillWordsKernel, dim3(1), dim3(256), 0, s, hist, 256, 0u);
  const uint32_t k = static_cast<uint32_t>(n / 2);
  const int blocks = static_cast<int>(std::min<size_t>(CeilDiv(static_cast<uint32_t>(n), 256), 1024));
  for (int kind = 0; kind < 2; ++kind) {
    hipLaunchKernelGGL(RadixResetKernel, dim3(1), dim3(1), 0, s, select_state, k, kind);
    for (int shift = 24; shift >= 0; shift -= 8) {
      hipLaunchKernelGGL(RadixHistogramKernel, dim3(blocks), dim3(256), 0, s, field, n,
                         shift, kind, select_state, hist);
      hipLaunchKernelGGL(RadixPickKernel, dim3(1), dim3(64), 0, s, hist, shift, kind,
                         select_state);
    }
  }
  hipLaunchKernelGGL(QuantizerParamsKernel, dim3(1), dim3(64), 0, s, select_state,
                     quant_dc, q);
  hipLaunchKernelGGL(RawQuantKernel, dim3(CeilDiv(static_cast<uint32_t>(n), 256)), dim3(256),
                     0, s, field, n, q, raw);
  return CheckLaunch("RawQuantKernel");
}

GJXLResult LaunchFinalCfl(GJXLAmdDevice* d, const float* coefs, int width,
                          int height, const uint8_t* cells,
                          const uint32_t* offsets, const int32_t* raw,
                          const QuantizerState* q, int8_t* y_to_x,
                          int8_t* y_to_b) {
  const AqLayout l = MakeAqLayout(width, height);
  hipLaunchKernelGGL(FinalCflKernel, dim3(CeilDiv(l.bw, 8), CeilDiv(l.bh, 8)), dim3(256),
                     0, d->stream, coefs, l.store, cells, offsets, l.bw, l.bh, raw, q,
                     d->quant_tables, y_to_x, y_to_b);
  return CheckLaunch("FinalCflKernel");
}

GJXLResult LaunchCodeCoefficients(GJXLAmdDevice* d, const float* coefs,
                                  int width, int height, const uint8_t* cells,
                                  const uint32_t* offsets,
                                  const StrategyLists& lists, int32_t* raw,
                                  const int8_t* y_to_x, const int8_t* y_to_b,
                                  const QuantizerState* q, float x_mult,
                                  float b_mult, int32_t* quantized, float* dc,
                                  int32_t* dc_quantized, float* dc_rec) {
  (void)cells;
  const AqLayout l = MakeAqLayout(width, height);
  CodingParams p;
  p.x_mult = x_mult;
  p.b_mult = b_mult;
  static const DcBasisTable table = HostDcBasis();
  p.dc_basis = table;
  const size_t n = static_cast<size_t>(l.bw) * l.bh;
  GJXL_AMD_TRY(ForEachStrategy(lists, [&]<int kStrategy>(const uint32_t* entries,
                                                         uint32_t count) {
    hipLaunchKernelGGL(CodeTransformKernel<kStrategy>, StrategyGrid<kStrategy>(count),
                       dim3(kAqThreads), 0, d->stream, coefs, l.store, entries, count, offsets,
                       l.bw, raw, y_to_x, y_to_b, q, d->quant_tables, p, quantized, dc, n);
    return CheckLaunch("CodeTransformKernel");
  }));
  const dim3 groups(CeilDiv(l.bw, kDcGroup), CeilDiv(l.bh, kDcGroup), 2);
  const bool probed = ProbeBegin(d, GJXL_AMD_KERNEL_DC_QUANT);
  hipLaunchKernelGGL(DcQuantKernel, groups, dim3(kDcWorkGroup), 0, d->stream, dc, l.bw,
                     l.bh, q, 1, dc_quantized);
  ProbeEnd(d, probed);
  GJXL_AMD_TRY(CheckLaunch("DcQuantKernel"));
  hipLaunchKernelGGL(DcReconstructKernel, dim3(CeilDiv(static_cast<uint32_t>(n), 256)),
                     dim3(256), 0, d->stream, dc_quantized, n, q, 1, dc_rec);
  return CheckLaunch("DcReconstructKernel");
}

GJXLResult LaunchReconstruct(GJXLAmdDevice* d, const int32_t* quantized,
                             int width, int height, const uint8_t* cells,
                             const uint32_t* offsets, const StrategyLists& lists,
                             const int32_t* raw,
                             const int8_t* y_to_x, const int8_t* y_to_b,
                             const QuantizerState* q, float x_mult, float b_mult,
                             const float* dc_rec, float* smoothed,
                             float* inv_sigma, float* out) {
  const AqLayout l = MakeAqLayout(width, height);
  hipLaunchKernelGGL(EpfSigmaKernel, dim3(CeilDiv(l.bw, 256), l.bh), dim3(256), 0,
                     d->stream, cells, l.bw, l.bh, raw, q, inv_sigma);
  GJXL_AMD_TRY(CheckLaunch("EpfSigmaKernel"));
  hipLaunchKernelGGL(SmoothDcKernel, dim3(CeilDiv(l.bw, 256), l.bh), dim3(256), 0,
                     d->stream, dc_rec, l.bw, l.bh, q, smoothed);
  GJXL_AMD_TRY(CheckLaunch("SmoothDcKernel"));
  CodingParams p;
  p.x_mult = x_mult;
  p.b_mult = b_mult;
  static const DcBasisTable table = HostDcBasis();
  p.dc_basis = table;
  return ForEachStrategy(lists, [&]<int kStrategy>(const uint32_t* entries, uint32_t count) {
    hipLaunchKernelGGL(ReconstructKernel<kStrategy>, StrategyGrid<kStrategy>(count),
                       dim3(kAqThreads), 0, d->stream, quantized, l.store, entries, count,
                       offsets, width, height, raw, y_to_x, y_to_b, q, d->quant_tables,
                       d->dct_bases, p, smoothed, out);
    return CheckLaunch("ReconstructKernel");
  });
}

}  // namespace gjxl_amd
