// a6 initial chroma-from-luma, a7-a9 AC-strategy candidate costs, a10 the
// greedy merge.
//
// Arithmetic follows the reference's CPU code term for term:
//   * DCTs use the binary64 bases with one sequential multiply-add chain per
//     output in the reference's index order (src/codec/dct.cpp:135-233), so
//     every coefficient is the same binary32 value;
//   * every floating-point reduction whose order the reference fixes (the
//     candidate's sqrt-magnitude and 8th-power loss sums, CfL lane sums) is
//     evaluated by one thread in that order, after all lanes of the work
//     group have produced the terms in parallel.
// CandidateCostKernel: one work group evaluates kCands candidates of one
// strategy (3072 coefficient slots: 1x32x32, 2x32x16, 4x16x16, 8x16x8 or
// 16x8x8), three channels each, entirely in LDS.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <type_traits>

#include "common.h"
#include "libm_port.h"
#include "tables.h"

namespace gjxl_amd {
namespace {

__device__ __forceinline__ float StdMax(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float StdMin(float a, float b) { return b < a ? b : a; }

// src/core/ac_strategy.h:95-108
template <int W, int H>
__device__ __forceinline__ int CoefficientIndex(int v, int u) {
  if constexpr (H < W) return v * W + u;
  return u * H + v;
}

// ---- a6: initial chroma from luma --------------------------------------

constexpr float kDistanceMultiplierAc = 1.0e-9f;

// src/codec/chroma_from_luma.cpp:96-110
__device__ __forceinline__ int8_t QuantizeMultiplier(float x) {
  if (x >= 2.6f) x -= 2.6f;
  else if (x <= -2.6f) x += 2.6f;
  else x = 0.0f;
  float r = roundf(x);
  r = r < -128.0f ? -128.0f : (127.0f < r ? 127.0f : r);
  return static_cast<int8_t>(r);
}

__device__ __forceinline__ float CflDerivative(float a, float residual) {
  float v = (2.0f / 3.0f) * a * (fabsf(residual) + 1.0f);
  if (residual < 0.0f) v = -v;
  return v;
}

constexpr int kCflThreads = 256;
// Newton phase: waves 1-3 produce the derivative terms of kCflChunk elements
// at a time (kCflPer each) into one of two LDS buffers while wave 0's 24
// summing lanes add up the previous chunk.
constexpr int kCflProducers = kCflThreads - 64;
constexpr int kCflPer = 4;
constexpr int kCflChunk = kCflProducers * kCflPer;
constexpr int kCflChunks = (4096 + kCflChunk - 1) / kCflChunk;
// terms[buffer][problem][kind][lane][m] holds the term of chunk element
// 4m + lane, so that each summing lane reads its elements with 16-byte
// reads; the stride (196 = 4 mod 32 words) puts the eight lanes of each
// 16-byte read group on distinct banks
constexpr int kCflLaneCount = kCflChunk / 4;
constexpr int kCflTermStride = kCflLaneCount + 4;
constexpr int kCflSumBatch = 16;  // divides every chunk's per-lane count (n is a multiple of 64)
constexpr int kCflSumBatches = kCflLaneCount / kCflSumBatch;
// the read-ahead past a lane's last batch (at most four batches) stays
// inside the LDS array
static_assert((2 * 2 * 3 * 4 - 1) * (kCflLaneCount + 4) +
              (kCflSumBatches + 4) * kCflSumBatch <= 3 * 4096, "read-ahead in bounds");

// One work group per 64x64 colour tile (src/codec/chroma_from_luma.cpp:
// 434-529): DCT8 of each block and channel, LLF cleared, scaled by the DCT8
// X/B inverse-dequant tables into the reference's luma/chroma value arrays
// (block-raster then coefficient order), kept as the per-element a = luma /
// 84 and b = base * luma - chroma of both maps; then up to 20 damped Newton
// steps per map (:40-94, :112-165).  Each step's derivative terms are formed
// by the producer threads from values held in registers, a chunk at a time,
// and summed by one thread per lane of the reference's 4-lane sums, in
// element order, one chunk behind.
// Three work groups share a CU (48 KB of LDS and at most 168 VGPRs each).
__global__ __launch_bounds__(kCflThreads, 3) void InitialCflKernel(
  const float* __restrict__ opsin, int width, int height,
  const float* __restrict__ tables, const double* __restrict__ bases,
  int8_t* __restrict__ y_to_x, int8_t* __restrict__ y_to_b) {
  // DCT phase: X, Y, B coefficients in values; Newton phase: two buffers
  // of terms[problem][kind][element] in the same memory
  __shared__ __align__(16) float lds[3 * 4096];
  float (*values)[4096] = reinterpret_cast<float (*)[4096]>(lds);
  float (*terms)[2][3][4][kCflTermStride] =
    reinterpret_cast<float (*)[2][3][4][kCflTermStride]>(lds);
  static_assert(2 * 2 * 3 * 4 * kCflTermStride <= 3 * 4096, "term buffers fit");
  static_assert(kCflProducers % 4 == 0, "a producer's elements share one lane");
  __shared__ float lane_sum[2][3][4];
  __shared__ float state_x[2];
  __shared__ int state_done[2];
  const int tid = threadIdx.x;
  const int tiles_w = (width + 63) / 64;
  const int bx0 = blockIdx.x * 8, by0 = blockIdx.y * 8;
  const int nbx = min(8, width / 8 - bx0), nby = min(8, height / 8 - by0);
  const int nblocks = nbx * nby;
  const size_t plane = static_cast<size_t>(width) * height;
  const double* basis = bases + BasisOffset(8, true);
  const float* inv_x = tables + InverseOffset(kDct8, 0);
  const float* inv_b = tables + InverseOffset(kDct8, 2);

  for (int job = tid; job < nblocks * 3; job += kCflThreads) {
    const int k = job / 3, c = job % 3;
    const int bx = bx0 + k % nbx, by = by0 + k / nbx;
    const float* src = opsin + c * plane + static_cast<size_t>(by * 8) * width + bx * 8;
    float p[64];
#pragma unroll
    for (int y = 0; y < 8; ++y)
#pragma unroll
      for (int x = 0; x < 8; ++x) p[y * 8 + x] = src[static_cast<size_t>(y) * width + x];
    // one column u at a time: its eight horizontal sums, then its eight
    // coefficients (the same sums in the same order as row-wise passes)
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      double hz[8];
#pragma unroll
      for (int y = 0; y < 8; ++y) {
        double sum = 0.0;
#pragma unroll
        for (int x = 0; x < 8; ++x) sum += static_cast<double>(p[y * 8 + x]) * basis[u * 8 + x];
        hz[y] = sum;
      }
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        double sum = 0.0;
#pragma unroll
        for (int y = 0; y < 8; ++y) sum += basis[v * 8 + y] * hz[y];
        // the LLF coefficient is cleared
        values[c][k * 64 + CoefficientIndex<8, 8>(v, u)] =
          v == 0 && u == 0 ? 0.0f : static_cast<float>(sum);
      }
    }
  }
  if (tid < 2) {
    state_x[tid] = 0.0f;
    state_done[tid] = 0;
  }
  __syncthreads();
  const int n = nblocks * 64;
  // the producers' elements: chunk c, slot m -> c * kCflChunk + q + m *
  // kCflProducers; a and b of both problems
  const bool producer = tid >= 64;
  const int q = tid - 64;
  float pa[kCflChunks][kCflPer][2], pb[kCflChunks][kCflPer][2];
#pragma unroll
  for (int c = 0; c < kCflChunks; ++c) {
#pragma unroll
    for (int m = 0; m < kCflPer; ++m) {
      const int i = c * kCflChunk + q + m * kCflProducers;
      float ax = 0.0f, bxv = 0.0f, ab = 0.0f, bbv = 0.0f;
      if (producer && i < n) {
        const float ix = inv_x[i & 63], ib = inv_b[i & 63];
        const float yv = values[1][i], xv = values[0][i], bv = values[2][i];
        const float luma_x = yv * ix, chroma_x = xv * ix;
        const float luma_b = yv * ib, chroma_b = bv * ib;
        ax = luma_x / 84.0f;
        bxv = 0.0f * luma_x - chroma_x;
        ab = luma_b / 84.0f;
        bbv = 1.0f * luma_b - chroma_b;
      }
      pa[c][m][0] = ax;
      pb[c][m][0] = bxv;
      pa[c][m][1] = ab;
      pb[c][m][1] = bbv;
    }
  }
  __syncthreads();  // the values memory becomes the term buffers

  const float count = static_cast<float>(n);
  const int chunks = (n + kCflChunk - 1) / kCflChunk;
  constexpr float kEps = 100.0f;
  // summing threads: problem p, derivative kind, lane
  const int sp = tid / 12, skind = (tid / 4) % 3, slane = tid % 4;
  for (int it = 0; it < 20; ++it) {
    const int done0 = state_done[0], done1 = state_done[1];
    if (done0 && done1) break;
    const float x0 = state_x[0], x1 = state_x[1];
    // terms of chunk c into buffer c & 1 (c is a constant after unrolling)
    auto produce = [&](int c) {
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        if (p == 0 ? done0 : done1) continue;
        const float x = p == 0 ? x0 : x1;
#pragma unroll
        for (int m = 0; m < kCflPer; ++m) {
          const float a = pa[c][m][p], b = pb[c][m][p];
          float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f;
          const float cr = __builtin_fmaf(a, x, b);
          if (!(fabsf(cr) >= 100.0f)) {
            t0 = CflDerivative(a, cr);
            t1 = CflDerivative(a, __builtin_fmaf(a, x + kEps, b));
            t2 = CflDerivative(a, __builtin_fmaf(a, x - kEps, b));
          }
          // element q + m * kCflProducers: lane q & 3, slot q / 4 + m *
          // kCflProducers / 4
          const int slot = (q >> 2) + m * (kCflProducers / 4);
          terms[c & 1][p][0][q & 3][slot] = t0;
          terms[c & 1][p][1][q & 3][slot] = t1;
          terms[c & 1][p][2][q & 3][slot] = t2;
        }
      }
    };
    if (producer) produce(0);
    __syncthreads();
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < kCflChunks; ++c) {
      if (c < chunks) {
        if (tid < 24 && !(sp == 0 ? done0 : done1)) {
          // n is a multiple of 64, so the lane's count is a whole number of
          // batches.  Batches go in groups of three through a ring of three
          // register sets, each read two batches ahead of its ordered adds
          // (reads past the lane's count stay inside the term buffers); the
          // one or two batches after the last whole group are already in
          // the first two sets.
          const int batches = min(kCflChunk, n - c * kCflChunk) / 4 / kCflSumBatch;
          const float4* src =
            reinterpret_cast<const float4*>(&terms[c & 1][sp][skind][slane][0]);
          constexpr int kQuads = kCflSumBatch / 4;
          float4 r0[kQuads], r1[kQuads], r2[kQuads];
          auto load = [&](int b, float4 (&dst)[kQuads]) {
#pragma unroll
            for (int m = 0; m < kQuads; ++m) dst[m] = src[b * kQuads + m];
          };
          auto add = [&](const float4 (&v)[kQuads]) {
#pragma unroll
            for (int m = 0; m < kQuads; ++m) {
              acc = acc + v[m].x;
              acc = acc + v[m].y;
              acc = acc + v[m].z;
              acc = acc + v[m].w;
            }
          };
          load(0, r0);
          load(1, r1);
          int b = 0;
#pragma unroll 1
          for (; b + 3 <= batches; b += 3) {
            load(b + 2, r2);
            add(r0);
            load(b + 3, r0);
            add(r1);
            load(b + 4, r1);
            add(r2);
          }
          if (b < batches) add(r0);
          if (b + 1 < batches) add(r1);
        }
        if (c + 1 < kCflChunks && producer && c + 1 < chunks) produce(c + 1);
        __syncthreads();
      }
    }
    if (tid < 24) lane_sum[sp][skind][slane] = acc;
    __syncthreads();
    if (tid < 2 && !state_done[tid]) {
      const float x = state_x[tid];
      float center = 2.0f * kDistanceMultiplierAc * count * x;
      float plus = 2.0f * kDistanceMultiplierAc * count * (x + kEps);
      float minus = 2.0f * kDistanceMultiplierAc * count * (x - kEps);
      const float (*l)[4] = lane_sum[tid];
      center += (l[0][0] + l[0][1]) + (l[0][2] + l[0][3]);
      plus += (l[1][0] + l[1][1]) + (l[1][2] + l[1][3]);
      minus += (l[2][0] + l[2][1]) + (l[2][2] + l[2][3]);
      const float second = (plus - minus) / (2.0f * kEps);
      const float step = center / (second + 0.85f);
      const float clamped = step < -20.0f ? -20.0f : (20.0f < step ? 20.0f : step);
      state_x[tid] = x - clamped;
      if (fabsf(step) < 3.0e-3f) state_done[tid] = 1;
    }
    __syncthreads();
  }
  if (tid == 0) {
    const int t = blockIdx.y * tiles_w + blockIdx.x;
    y_to_x[t] = n == 0 ? 0 : QuantizeMultiplier(state_x[0]);
    y_to_b[t] = n == 0 ? 0 : QuantizeMultiplier(state_x[1]);
  }
}

// ---- a7-a9: candidate cost ---------------------------------------------

struct CostParams {
  float info_loss_multiplier;
  float zeros_multiplier;
  float cost_delta;
  float entropy_multiplier;
};

__device__ __forceinline__ int CeilLog2Nonzero(int v) {
  return v <= 1 ? 0 : 32 - __clz(v - 1);
}

// cost plane (stage) of each searched strategy
__host__ __device__ constexpr int StageOf(int strategy) {
  return strategy == kDct8 ? 0 : strategy == kDct16x8 ? 1 : strategy == kDct8x16 ? 2
       : strategy == kDct16x16 ? 3 : strategy == kDct32x16 ? 4 : strategy == kDct16x32 ? 5 : 6;
}

constexpr int kCostThreads = 512;
constexpr int kCostSlots = 3072;

template <int kStrategy>
struct CostShape {
  static constexpr int kCw = CoveredW(kStrategy);
  static constexpr int kCh = CoveredH(kStrategy);
  static constexpr int kW = kCw * 8;
  static constexpr int kH = kCh * 8;
  static constexpr int kN = kW * kH;
  static constexpr int kCands = kCostSlots / (3 * kN);
  static constexpr int kStep =
    (kStrategy == kDct32x16 || kStrategy == kDct16x32 || kStrategy == kDct32x32) ? 2 : 1;
  // anchor positions inside a 64x64 tile, in raster order
  static constexpr int kNx = (8 - kCw) / kStep + 1;
  static constexpr int kNy = (8 - kCh) / kStep + 1;
  static constexpr int kPositions = kNx * kNy;
};

// One work group's work: kCands anchor positions of one 64x64 tile
// (group.y), positions group.x * kCands + k of the tile's raster list, or
// (kListed) entries group.x * kCands + k of a list of anchors (x | y << 16,
// *list_count of them) from any tiles.  Only positions whose footprint lies
// inside their tile are evaluated; the rest of the cost plane is NaN
// (filled beforehand).  With allow, a candidate whose stage bit is clear at
// its anchor costs +infinity, which the greedy merge never selects.
template <int kStrategy, bool kListed = false>
__device__ __forceinline__ void CostWork(
  const TileIndex group, const float* __restrict__ opsin, int width, int height,
  const float* __restrict__ quant_field, const float* __restrict__ pixel_mask,
  const int8_t* __restrict__ y_to_x, const int8_t* __restrict__ y_to_b,
  const float* __restrict__ tables, const double* __restrict__ bases,
  CostParams params, const uint8_t* __restrict__ allow, float* __restrict__ costs,
  const int tid, const uint32_t* __restrict__ list = nullptr,
  const uint32_t* __restrict__ list_count = nullptr) {
  using S = CostShape<kStrategy>;
  constexpr int W = S::kW, H = S::kH, N = S::kN, K = S::kCands;
  constexpr int KC = 3 * K;     // candidate-channel planes
  constexpr int TS = W + 1;     // padded row stride of the binary64 passes
  // For 32-tall transforms, transposed planes have columns of H + 1 words
  // (HP): the lanes that write one transposed row then hit distinct banks
  // instead of sharing one per 32 words (measured: 32x32 12 % and 16x32 5 %
  // faster; the shorter transforms ran 5-10 % slower padded and keep
  // HP = H).  A plane's slot count is NP = W * HP.
  constexpr bool kPad = H == 32;
  constexpr int HP = kPad ? H + 1 : H, NP = W * HP;
  // Storage order (coefficient index i = u * H + v when H >= W) is padded
  // the same way: slot i + i / H; for H < W (i = v * W + u) it is unpadded.
  constexpr bool kPadStorage = kPad && H >= W;
  auto storage_slot = [](int i) { return kPadStorage ? i + i / H : i; };
  // A: pixels (transposed, [x][y], column stride HP); then sqrt(|rounded|)
  //    in (padded) storage order
  // B: coefficients ([v][u]); residuals (transposed, [u][v], column stride
  //    HP); loss terms ([y][x])
  // T: binary64 rows of the separable passes, stride TS
  __shared__ float A[KC * NP];
  __shared__ float B[KC * NP];
  __shared__ double T[KC * H * TS];
  __shared__ float mask[K * N];  // pixel mask of each candidate, raster order
  __shared__ int nnz[K][3];
  __shared__ float sums[K][3][2];
  __shared__ float qn[K];
  __shared__ float cfl[K][3];
  __shared__ int valid[K];
  __shared__ int bxs[K], bys[K];
  __shared__ uint32_t listed[kListed ? K : 1];

  // the thread index's range, which a caller's opaque copy hides
  __builtin_assume(tid >= 0);
  __builtin_assume(tid < kCostThreads);
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr int kWaves = kCostThreads / 64;
  const int bw = width / 8, bh = height / 8;
  const int tiles_w = (bw + 7) / 8;
  const int tx = group.y % tiles_w, ty = group.y / tiles_w;
  const int tbw = min(8, bw - tx * 8), tbh = min(8, bh - ty * 8);
  const size_t plane = static_cast<size_t>(width) * height;
  // binary64 bases, read with uniform (scalar) loads
  const double* fhT = bases + BasisOffsetT(W);        // [x][u]
  const double* fvT = bases + BasisOffsetT(H);        // [y][v]
  const double* ih = bases + BasisOffset(W, false);   // [u][x]
  const double* iv = bases + BasisOffset(H, false);   // [v][y]
  constexpr uint32_t kNoEntry = 0xffffffffu;
  if constexpr (kListed) {
    const uint32_t count = *list_count;
    if (static_cast<uint32_t>(group.x) * K >= count) return;  // past the list
    if (tid < K) {
      const uint32_t at = static_cast<uint32_t>(group.x) * K + tid;
      listed[tid] = at < count ? list[at] : kNoEntry;
    }
    __syncthreads();
  }
  // candidate k of this work group: anchor block and whether it is launched
  auto candidate = [&](int k, int* bx, int* by) {
    if constexpr (kListed) {
      const uint32_t e = listed[k];
      *bx = e == kNoEntry ? 0 : static_cast<int>(e & 0xffffu);
      *by = e == kNoEntry ? 0 : static_cast<int>(e >> 16);
      return e != kNoEntry;
    }
    const int p = group.x * K + k;
    const int lx = (p % S::kNx) * S::kStep, ly = (p / S::kNx) * S::kStep;
    *bx = tx * 8 + lx;
    *by = ty * 8 + ly;
    return p < S::kPositions && lx + S::kCw <= tbw && ly + S::kCh <= tbh;
  };
  // The learned form's pruning: a group none of whose candidates is allowed
  // at its anchor writes +infinity for them (never merged) and does no work.
  // (Listed groups hold allowed candidates only.)
  if (!kListed && allow != nullptr) {
    constexpr int kStage = StageOf(kStrategy);
    bool wanted = false;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      int bx, by;
      if (candidate(k, &bx, &by) && ((allow[static_cast<size_t>(by) * bw + bx] >> kStage) & 1)) {
        wanted = true;
      }
    }
    if (!wanted) {
      int bx, by;
      if (tid < K && candidate(tid, &bx, &by)) {
        costs[static_cast<size_t>(by) * bw + bx] = __builtin_inff();
      }
      return;
    }
  }

  // Every global read of the work group is issued here, in the order the
  // results are needed (reads complete in issue order): the pixels, which
  // the first barrier waits for, then the mask, the quantisation table
  // entries of the elements this thread quantises and the quant-field and
  // CfL inputs, which reach LDS behind the forward passes.
  // Each strategy's slots hold a whole number of elements per thread, so
  // every read below is unconditional: a candidate that is not launched
  // reads element 0 of the plane and discards it.  (A read under a branch
  // makes the compiler wait for the earlier reads at the branch's join.)
  constexpr int kPer = KC * N / kCostThreads;
  constexpr int kMaskPer = K * N / kCostThreads;
  static_assert(N % 64 == 0, "a wave's elements lie in one candidate channel");
  static_assert(KC * N == kPer * kCostThreads && K * N == kMaskPer * kCostThreads,
                "whole elements per thread");
  float pixel[kPer], dq_e[kPer], inv_e[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int e = tid + j * kCostThreads;
    const int kc = e / N, r = e % N;
    const int k = kc / 3, c = kc % 3;
    int bx, by;
    const bool live = candidate(k, &bx, &by);
    const size_t at = c * plane + static_cast<size_t>(by * 8 + r / W) * width + bx * 8 + r % W;
    const float v = opsin[live ? at : 0];
    pixel[j] = live ? v : 0.0f;
  }
  float mask_e[kMaskPer];
#pragma unroll
  for (int j = 0; j < kMaskPer; ++j) {
    const int e = tid + j * kCostThreads;
    const int k = e / N, r = e % N;
    int bx, by;
    const bool live = candidate(k, &bx, &by);
    const size_t at = static_cast<size_t>(by * 8 + r / W) * width + bx * 8 + r % W;
    const float v = pixel_mask[live ? at : 0];
    mask_e[j] = live ? v : 0.0f;
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int e = tid + j * kCostThreads;
    const int r = e % N, c = (e / N) % 3;
    // element e is coefficient (v, u) = (r / W, r % W) in the quantise step
    const int i = CoefficientIndex<W, H>(r / W, r % W);
    dq_e[j] = tables[DequantOffset(kStrategy, c) + i];
    inv_e[j] = tables[InverseOffset(kStrategy, c) + i];
  }
  // thread k < K: candidate k's quant-field blocks and the tile's CfL
  // multipliers (ComputeQuantNormUnchecked, src/codec/ac_strategy.cpp:103-137)
  constexpr int kBlocks = S::kCw * S::kCh;
  float qv[kBlocks];
  int8_t yx = 0, yb = 0;
  if constexpr (!kListed) {
    const int tile = ty * tiles_w + tx;
    yx = y_to_x[tile];
    yb = y_to_b[tile];
  }
  bool ok = false;
  if (tid < K) {
    int bx, by;
    ok = candidate(tid, &bx, &by);
    valid[tid] = ok;
    bxs[tid] = bx;
    bys[tid] = by;
    if constexpr (kListed) {
      // each listed candidate's own tile
      const int tile = ok ? (by / 8) * tiles_w + bx / 8 : 0;
      yx = y_to_x[tile];
      yb = y_to_b[tile];
    }
    const float* q = quant_field + static_cast<size_t>(ok ? by : 0) * bw + (ok ? bx : 0);
#pragma unroll
    for (int dy = 0; dy < S::kCh; ++dy) {
#pragma unroll
      for (int dx = 0; dx < S::kCw; ++dx) qv[dy * S::kCw + dx] = q[static_cast<size_t>(dy) * bw + dx];
    }
    for (int c = 0; c < 3; ++c) nnz[tid][c] = 0;
  }
  // pixels transposed ([x][y]) into LDS
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int e = tid + j * kCostThreads;
    if (e < KC * N) {
      const int kc = e / N, r = e % N;
      A[kc * NP + (r % W) * HP + r / W] = pixel[j];
    }
  }
  __syncthreads();

  // The separable passes: a wave takes 64 rows (or columns) and one group of
  // kOut outputs, so that the basis entries are the same for all its lanes;
  // each output is one binary64 chain in the reference's order
  // (src/codec/dct.cpp:135-233).
  constexpr int kOut = 8;
  constexpr int kRows = KC * H, kCols = KC * W;
  constexpr int kRowBlocks = (kRows + 63) / 64, kColBlocks = (kCols + 63) / 64;
  // Forward horizontal: T[y][u] = sum_x pix[y][x] * fh[u][x].
  for (int task = wave; task < kRowBlocks * (W / kOut); task += kWaves) {
    const int g = task % (W / kOut), r = (task / (W / kOut)) * 64 + lane;
    if (r < kRows) {
      const int kc = r / H, y = r % H;
      const float* src = A + kc * NP + y;
      double acc[kOut] = {};
#pragma unroll 4
      for (int x = 0; x < W; ++x) {
        const double v = static_cast<double>(src[x * HP]);
#pragma unroll
        for (int j = 0; j < kOut; ++j) acc[j] += v * fhT[x * W + g * kOut + j];
      }
      double* dst = T + (kc * H + y) * TS + g * kOut;
#pragma unroll
      for (int j = 0; j < kOut; ++j) dst[j] = acc[j];
    }
  }
  __syncthreads();
  // Forward vertical: coef[v][u] = sum_y fv[v][y] * T[y][u].
  for (int task = wave; task < kColBlocks * (H / kOut); task += kWaves) {
    const int g = task % (H / kOut), col = (task / (H / kOut)) * 64 + lane;
    if (col < kCols) {
      const int kc = col / W, u = col % W;
      const double* src = T + kc * H * TS + u;
      double acc[kOut] = {};
#pragma unroll 4
      for (int y = 0; y < H; ++y) {
        const double v = src[y * TS];
#pragma unroll
        for (int j = 0; j < kOut; ++j) acc[j] += fvT[y * H + g * kOut + j] * v;
      }
#pragma unroll
      for (int j = 0; j < kOut; ++j) B[kc * N + (g * kOut + j) * W + u] = static_cast<float>(acc[j]);
    }
  }
  // the reads issued at the start: mask into LDS, quant norms and CfL
#pragma unroll
  for (int j = 0; j < kMaskPer; ++j) {
    const int e = tid + j * kCostThreads;
    if (e < K * N) mask[e] = mask_e[j];
  }
  if (tid < K && ok) {
    float norm;
    if constexpr (kBlocks == 1) {
      norm = qv[0];
    } else if constexpr (kBlocks == 2) {
      norm = StdMax(qv[0], qv[1]);
    } else {
      float sum = 0.0f;
#pragma unroll
      for (int b = 0; b < kBlocks; ++b) {
        float v = qv[b];
        v *= v;
        v *= v;
        v *= v;
        sum += v * v;
      }
      sum /= static_cast<float>(kBlocks);
      norm = FastPow(sum, 1.0f / 16.0f);
    }
    qn[tid] = norm;
    cfl[tid][0] = static_cast<float>(yx) / 84.0f;
    cfl[tid][1] = 0.0f;
    cfl[tid][2] = 1.0f + static_cast<float>(yb) / 84.0f;
  }
  __syncthreads();
  // Quantise (src/codec/ac_strategy.cpp:914-926): sqrt(|rounded|) into A in
  // storage order; residuals staged in registers, then into B transposed.
  float residual[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int e = tid + j * kCostThreads;
    residual[j] = 0.0f;
    if (e < KC * N) {
      const int kc = e / N, rem = e % N;
      const int k = kc / 3, c = kc % 3;
      const int v = rem / W, u = rem % W;
      const int i = CoefficientIndex<W, H>(v, u);
      const float dec = B[e] - B[(k * 3 + 1) * N + rem] * cfl[k][c];
      const float scaled = dec * inv_e[j] * qn[k];
      const float rounded = roundf(scaled);
      const float diff = scaled - rounded;
      residual[j] = dq_e[j] * diff;
      A[kc * NP + storage_slot(i)] = sqrtf(fabsf(rounded));
      // a wave's 64 elements lie in one candidate channel (N is a multiple
      // of 64): one LDS add of the wave's nonzero count
      const int count = __popcll(__ballot(rounded != 0.0f));
      if (lane == 0 && count != 0) atomicAdd(&nnz[k][c], count);
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int e = tid + j * kCostThreads;
    if (e < KC * N) {
      const int kc = e / N, rem = e % N;
      const int v = rem / W, u = rem % W;
      B[kc * NP + u * HP + v] = residual[j];
    }
  }
  __syncthreads();
  // Inverse horizontal: T[v][x] = sum_u res(v, u) * ih[u][x].
  for (int task = wave; task < kRowBlocks * (W / kOut); task += kWaves) {
    const int g = task % (W / kOut), r = (task / (W / kOut)) * 64 + lane;
    if (r < kRows) {
      const int kc = r / H, v = r % H;
      const float* src = B + kc * NP + v;
      double acc[kOut] = {};
#pragma unroll 4
      for (int u = 0; u < W; ++u) {
        const double t = static_cast<double>(src[u * HP]);
#pragma unroll
        for (int j = 0; j < kOut; ++j) acc[j] += t * ih[u * W + g * kOut + j];
      }
      double* dst = T + (kc * H + v) * TS + g * kOut;
#pragma unroll
      for (int j = 0; j < kOut; ++j) dst[j] = acc[j];
    }
  }
  __syncthreads();
  // Inverse vertical and the masked 8th power (:935-952) into B, raster.
  // 12, 0, 4 for X, Y, B (selects, not a table: no memory access)
  for (int task = wave; task < kColBlocks * (H / kOut); task += kWaves) {
    const int g = task % (H / kOut), col = (task / (H / kOut)) * 64 + lane;
    if (col < kCols) {
      const int kc = col / W, x = col % W;
      const int k = kc / 3, c = kc % 3;
      const float mask_offset = c == 0 ? 12.0f : c == 2 ? 4.0f : 0.0f;
      const double* src = T + kc * H * TS + x;
      double acc[kOut] = {};
#pragma unroll 4
      for (int v = 0; v < H; ++v) {
        const double t = src[v * TS];
#pragma unroll
        for (int j = 0; j < kOut; ++j) acc[j] += iv[v * H + g * kOut + j] * t;
      }
#pragma unroll
      for (int j = 0; j < kOut; ++j) {
        const int y = g * kOut + j;
        float weighted = 0.0f;
        if (valid[k]) {
          const float m = mask[k * N + y * W + x];
          weighted = (m + mask_offset) * static_cast<float>(acc[j]);
          weighted *= weighted;
          weighted *= weighted;
          weighted *= weighted;
        }
        B[kc * N + y * W + x] = weighted;
      }
    }
  }
  __syncthreads();
  // Ordered sums, one thread each.
  if (tid < K * 6) {
    const int k = tid / 6, c = (tid / 2) % 3, kind = tid % 2;
    // kind 0: A in (padded) storage order; kind 1: B in raster order
    const float* src = kind == 0 ? A + (k * 3 + c) * NP : B + (k * 3 + c) * N;
    const bool padded = kind == 0 && kPadStorage;
    float s = 0.0f;
    // batches of 16: the loads of a batch issue before its ordered adds
    for (int base = 0; base < N; base += 16) {
      // kind 0 reads the padded slots i + i / H: a batch lies in one padded
      // column for H >= 16 and in two for H = 8 (the second one word on)
      const float* lo = src + base + (padded ? base / H : 0);
      const float* hi = lo + (padded && H < 16 ? 1 : 0);
      float v[16];
#pragma unroll
      for (int m = 0; m < 16; ++m) v[m] = (m >= 8 && H < 16 ? hi : lo)[m];
#pragma unroll
      for (int m = 0; m < 16; ++m) s += v[m];
    }
    sums[k][c][kind] = s;
  }
  __syncthreads();
  if (tid < K && valid[tid]) {
    const int k = tid;
    constexpr float kChannelMultiplier[3] = {2.0441408586549744e7f, 1.0f,
                                             1.266770081387616f};
    float entropy = 0.0f, loss = 0.0f;
    for (int c = 0; c < 3; ++c) {
      loss += sums[k][c][1] * kChannelMultiplier[c];
      entropy += params.cost_delta * sums[k][c][0];
      const int nz_bits = CeilLog2Nonzero(nnz[k][c] + 1) + 1;
      entropy += params.zeros_multiplier *
                 static_cast<float>(CeilLog2Nonzero(nz_bits + 17) + nz_bits);
      if (c == 0 && S::kCw * S::kCh >= 2) {
        const float weight =
          1.0f + StdMin(3.0f, static_cast<float>(S::kCw * S::kCh) / 8.0f);
        entropy *= weight;
        loss *= weight;
      }
    }
    const float normalized = loss / static_cast<float>(N);
    // std::pow(float, float) as the reference's libm evaluates it
    const float root8 = GlibcPowf(normalized, 1.0f / 8.0f);
    const float loss_cost = root8 * static_cast<float>(N) / qn[k];
    const size_t at = static_cast<size_t>(bys[k]) * bw + bxs[k];
    const bool allowed = kListed || allow == nullptr ||
                         ((allow[at] >> StageOf(kStrategy)) & 1) != 0;
    costs[at] = allowed ? entropy * params.entropy_multiplier +
                            params.info_loss_multiplier * loss_cost
                        : __builtin_inff();
  }
}

// One work group per (chunk, tile); the groups of one 64x64 tile
// (consecutive in dispatch order) share its pixels, and the XCD-aware order
// keeps them on one L2.
template <int kStrategy>
__global__ __launch_bounds__(kCostThreads) void CandidateCostKernel(
  const float* __restrict__ opsin, int width, int height,
  const float* __restrict__ quant_field, const float* __restrict__ pixel_mask,
  const int8_t* __restrict__ y_to_x, const int8_t* __restrict__ y_to_b,
  const float* __restrict__ tables, const double* __restrict__ bases,
  CostParams params, const uint8_t* __restrict__ allow, float* __restrict__ costs) {
  CostWork<kStrategy>(XcdTile(), opsin, width, height, quant_field, pixel_mask, y_to_x, y_to_b,
                      tables, bases, params, allow, costs, static_cast<int>(threadIdx.x));
}

// The allowed candidates of one strategy as a list of anchors (x | y << 16):
// one thread per (tile, raster position); a fitting candidate whose stage
// bit is clear at its anchor costs +infinity here.  Entries are appended in
// wave order through one counter (any order: each candidate has its own
// cost entry).
template <int kStrategy>
__global__ __launch_bounds__(256) void CandidateListKernel(
  const uint8_t* __restrict__ allow, int bw, int bh, uint32_t* __restrict__ list,
  uint32_t* __restrict__ count, float* __restrict__ costs) {
  using S = CostShape<kStrategy>;
  const int tiles_w = (bw + 7) / 8, tiles = tiles_w * ((bh + 7) / 8);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int tile = i / S::kPositions, p = i % S::kPositions;
  const int tx = tile % tiles_w, ty = tile / tiles_w;
  const int lx = (p % S::kNx) * S::kStep, ly = (p / S::kNx) * S::kStep;
  const int bx = tx * 8 + lx, by = ty * 8 + ly;
  const bool fits = tile < tiles && lx + S::kCw <= min(8, bw - tx * 8) &&
                    ly + S::kCh <= min(8, bh - ty * 8);
  const size_t at = fits ? static_cast<size_t>(by) * bw + bx : 0;
  const bool allowed = fits && ((allow[at] >> StageOf(kStrategy)) & 1) != 0;
  if (fits && !allowed) costs[at] = __builtin_inff();
  const uint64_t mask = __ballot(allowed);
  const int lane = threadIdx.x & 63;
  uint32_t base = 0;
  if (lane == 0 && mask != 0) base = atomicAdd(count, static_cast<uint32_t>(__popcll(mask)));
  base = __shfl(base, 0);
  if (allowed) {
    const uint64_t below = lane == 0 ? 0 : mask & (~uint64_t{0} >> (64 - lane));
    list[base + __popcll(below)] = static_cast<uint32_t>(bx) | (static_cast<uint32_t>(by) << 16);
  }
}

// Listed candidates: group x takes entries x * kCands ... of the list.
template <int kStrategy>
__global__ __launch_bounds__(kCostThreads) void CandidateCostListKernel(
  const float* __restrict__ opsin, int width, int height,
  const float* __restrict__ quant_field, const float* __restrict__ pixel_mask,
  const int8_t* __restrict__ y_to_x, const int8_t* __restrict__ y_to_b,
  const float* __restrict__ tables, const double* __restrict__ bases,
  CostParams params, const uint32_t* __restrict__ list, const uint32_t* __restrict__ count,
  float* __restrict__ costs) {
  CostWork<kStrategy, true>(TileIndex{static_cast<int>(blockIdx.x), 0}, opsin, width, height,
                            quant_field, pixel_mask, y_to_x, y_to_b, tables, bases, params,
                            nullptr, costs, static_cast<int>(threadIdx.x), list, count);
}

// ---- the tile queue: candidate costs of a batch of frames -----------------
//
// Persistent work groups pull work items -- (64x64 tile, chunk of anchor
// positions) of one strategy -- from counters in device memory until the
// batch's items run out.  The batch's tiles (frames in order) are split
// over the eight XCDs by tile index mod 8, one counter each, with a tile's
// chunks consecutive: a work group (dispatched round-robin over the XCDs)
// drains its own XCD's queue, where the tile's pixels stay in that XCD's
// L2, then helps the others.  Frames are described in a table in device
// memory (so one captured graph serves every batch); a plan kernel derives
// the tile offsets of the frames and clears the counters.

constexpr int kQueueMaxFrames = 256;

struct QueueFrame {
  const float* opsin;
  const float* quant_field;
  const float* pixel_mask;
  const int8_t* y_to_x;
  const int8_t* y_to_b;
  const uint8_t* allow;
  float* costs;
  int32_t width, height;
  CostParams params;  // entropy_multiplier set per strategy
};

constexpr int kQueueXcds = 8;

struct QueueTable {
  uint32_t count;
  uint32_t counters[7][kQueueXcds];
  uint32_t tile_offsets[kQueueMaxFrames + 1];  // first 64x64 tile of each frame
  QueueFrame frames[kQueueMaxFrames];
};

template <int kStrategy>
constexpr uint32_t QueueChunks() {
  return static_cast<uint32_t>((CostShape<kStrategy>::kPositions + CostShape<kStrategy>::kCands - 1) /
                               CostShape<kStrategy>::kCands);
}

__constant__ uint32_t kQueueStageChunks[7] = {
  QueueChunks<kDct8>(), QueueChunks<kDct16x8>(), QueueChunks<kDct8x16>(),
  QueueChunks<kDct16x16>(), QueueChunks<kDct32x16>(), QueueChunks<kDct16x32>(),
  QueueChunks<kDct32x32>()};

// tile offsets of the frames; counters cleared
__global__ __launch_bounds__(64) void QueuePlanKernel(QueueTable* __restrict__ t) {
  if (threadIdx.x < 7 * kQueueXcds) (&t->counters[0][0])[threadIdx.x] = 0;
  if (threadIdx.x != 0) return;
  const uint32_t count = t->count;
  uint32_t at = 0;
  for (uint32_t f = 0; f < count; ++f) {
    t->tile_offsets[f] = at;
    const int bw = t->frames[f].width / 8, bh = t->frames[f].height / 8;
    at += static_cast<uint32_t>(((bw + 7) / 8) * ((bh + 7) / 8));
  }
  t->tile_offsets[count] = at;
}

// NaN over every frame's seven cost planes
__global__ __launch_bounds__(256) void QueueFillKernel(const QueueTable* __restrict__ t) {
  const uint32_t count = t->count;
  const size_t step = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (uint32_t f = 0; f < count; ++f) {
    const QueueFrame& q = t->frames[f];
    const size_t n = 7 * static_cast<size_t>(q.width / 8) * (q.height / 8);
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) {
      q.costs[i] = __builtin_nanf("");
    }
  }
}

// Two groups per CU (four waves per SIMD), which the LDS allows.
template <int kStrategy>
__global__ __launch_bounds__(kCostThreads, 4) void CostQueueKernel(
  QueueTable* __restrict__ t, const float* __restrict__ tables,
  const double* __restrict__ bases, float entropy_multiplier) {
  constexpr int kStage = StageOf(kStrategy);
  constexpr uint32_t kChunks = QueueChunks<kStrategy>();
  __shared__ uint32_t next;
  const uint32_t count = t->count;
  const uint32_t tiles = t->tile_offsets[count];
  const int home = static_cast<int>(blockIdx.x % kQueueXcds);
  for (int k = 0; k < kQueueXcds; ++k) {
    const int x = (home + k) % kQueueXcds;
    // XCD x's queue: tiles x, x + 8, ... with their chunks consecutive
    const uint32_t length =
      tiles > static_cast<uint32_t>(x) ? (tiles - x + kQueueXcds - 1) / kQueueXcds * kChunks : 0;
    // thread 0 claims the next item while the group works on the current
    // one: the counter's device-scope atomic (a few microseconds) overlaps
    // the item's own loads instead of preceding them
    uint32_t claimed = 0;
    // the frame of the last item and where the next frame's tiles begin:
    // tiles grow along one queue, so the frame only moves forward, and its
    // fields stay in registers until it does (no chain of dependent loads
    // before each item's pixel reads)
    int frame = 0;
    uint32_t frame_end = count > 1 ? t->tile_offsets[1] : tiles;
    QueueFrame f = t->frames[0];
    uint32_t first_tile = 0;
    __syncthreads();  // every lane has read `next` and the last item's LDS
    if (threadIdx.x == 0) next = atomicAdd(&t->counters[kStage][x], 1u);
    for (;;) {
      __syncthreads();
      // uniform: the frame's fields become scalar loads
      const uint32_t q = __builtin_amdgcn_readfirstlane(next);
      if (q >= length) break;  // the counter only grows: every group ends here
      if (threadIdx.x == 0) claimed = atomicAdd(&t->counters[kStage][x], 1u);
      const uint32_t tile = static_cast<uint32_t>(x) + kQueueXcds * (q / kChunks);
      if (tile >= frame_end) {
        do {
          ++frame;
          frame_end = frame + 1 < static_cast<int>(count) ? t->tile_offsets[frame + 1] : tiles;
        } while (tile >= frame_end);
        f = t->frames[frame];
        first_tile = t->tile_offsets[frame];
      }
      const TileIndex g{static_cast<int>(q % kChunks), static_cast<int>(tile - first_tile)};
      CostParams params = f.params;
      params.entropy_multiplier = entropy_multiplier;
      const size_t stage = static_cast<size_t>(f.width / 8) * (f.height / 8);
      // the thread index through a register move the compiler cannot hoist:
      // the index arithmetic derived from it is redone per item instead of
      // staying live across the whole loop (which took 160+ registers)
      int tid;
      asm volatile("v_mov_b32 %0, %1" : "=v"(tid) : "v"(static_cast<int>(threadIdx.x)));
      CostWork<kStrategy>(g, f.opsin, f.width, f.height, f.quant_field, f.pixel_mask, f.y_to_x,
                          f.y_to_b, tables, bases, params, f.allow, f.costs + kStage * stage, tid);
      __syncthreads();  // every lane has read `next` and the item's LDS
      if (threadIdx.x == 0) next = claimed;
    }
  }
}

__global__ void FillNanKernel(float* __restrict__ p, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = __builtin_nanf("");
}

// ---- a10: greedy merge, one work group per 64x64 tile ----------------------

constexpr int kStageStrategy[7] = {kDct8, kDct16x8, kDct8x16, kDct16x16,
                                   kDct32x16, kDct16x32, kDct32x32};

constexpr int kSearchThreads = 64;

// The tile's search state lives across the wave: lane i keeps block i's
// cell (strategy << 1 | anchor), cost, priority and seven candidate costs in
// registers.  Every lane runs the search with the same (uniform) values and
// reads a block's state with readlane, writes it with a per-lane select: a few cycles
// where an LDS round trip costs about a hundred, and the search is one long
// chain of such accesses.
struct TileSearch {
  int bw, bh;
  int tbx, tby;  // tile origin in blocks
  int tw, th;
  int cell;      // this lane's block
  float cost;
  int prio;
  float table[7];  // [stage]

  __device__ __forceinline__ int Cell(int lx, int ly) const {
    return __builtin_amdgcn_readlane(cell, ly * 8 + lx);
  }
  __device__ __forceinline__ float Cost(int lx, int ly) const {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cost), ly * 8 + lx));
  }
  __device__ __forceinline__ int Prio(int lx, int ly) const {
    return __builtin_amdgcn_readlane(prio, ly * 8 + lx);
  }
  int lane;
  // a uniform value into one lane: a compare and a select
  __device__ __forceinline__ void SetCell(int lx, int ly, int v) {
    cell = lane == ly * 8 + lx ? v : cell;
  }
  __device__ __forceinline__ void SetCost(int lx, int ly, float v) {
    cost = lane == ly * 8 + lx ? v : cost;
  }
  __device__ __forceinline__ void SetPrio(int lx, int ly, int v) {
    prio = lane == ly * 8 + lx ? v : prio;
  }
  // strategy is a constant at every call site, so the stage folds
  __device__ __forceinline__ float Table(int strategy, int lx, int ly) const {
    return __int_as_float(
      __builtin_amdgcn_readlane(__float_as_int(table[StageOf(strategy)]), ly * 8 + lx));
  }

  __device__ __forceinline__ void Set(int lx, int ly, int strategy) {
    const int cw = CoveredW(strategy), ch = CoveredH(strategy);
    for (int dy = 0; dy < ch; ++dy)
      for (int dx = 0; dx < cw; ++dx)
        SetCell(lx + dx, ly + dy, (strategy << 1) | (dx == 0 && dy == 0));
  }

  // src/codec/ac_strategy.cpp:172-191 (global block coordinates)
  __device__ __forceinline__ bool CrossesHorizontal(int sx, int y, int ex) const {
    if (sx >= bw || y >= bh || y % 8 == 0) return false;
    ex = min(ex, bw);
    const int limit = sx & ~7;
    while (sx != limit && !(Cell(sx - tbx, y - tby) & 1)) --sx;
    for (int x = sx; x < ex;) {
      const int c = Cell(x - tbx, y - tby);
      if (!(c & 1)) return true;
      x += CoveredW(c >> 1);
    }
    return false;
  }

  // :193-213
  __device__ __forceinline__ bool CrossesVertical(int x, int sy, int ey) const {
    if (x >= bw || sy >= bh || x % 8 == 0) return false;
    ey = min(ey, bh);
    const int limit = sy & ~7;
    while (sy != limit && !(Cell(x - tbx, sy - tby) & 1)) --sy;
    for (int y = sy; y < ey;) {
      const int c = Cell(x - tbx, y - tby);
      if (!(c & 1)) return true;
      y += CoveredH(c >> 1);
    }
    return false;
  }

  __device__ __forceinline__ void SetCostForTransform(int lx, int ly, int strategy, float c) {
    const int cw = CoveredW(strategy), ch = CoveredH(strategy);
    for (int dy = 0; dy < ch; ++dy)
      for (int dx = 0; dx < cw; ++dx) SetCost(lx + dx, ly + dy, 0.0f);
    SetCost(lx, ly, c);
  }

  // :340-384
  __device__ __forceinline__ void TryMerge(int strategy, int lx, int ly, int priority) {
    const int cw = CoveredW(strategy), ch = CoveredH(strategy);
    float current = 0.0f;
    for (int dy = 0; dy < ch; ++dy) {
      for (int dx = 0; dx < cw; ++dx) {
        if (Prio(lx + dx, ly + dy) >= priority) return;
        current += Cost(lx + dx, ly + dy);
      }
    }
    const float candidate = Table(strategy, lx, ly);
    if (candidate >= current) return;
    for (int dy = 0; dy < ch; ++dy) {
      for (int dx = 0; dx < cw; ++dx) {
        SetCost(lx + dx, ly + dy, 0.0f);
        SetPrio(lx + dx, ly + dy, priority);
      }
    }
    Set(lx, ly, strategy);
    SetCost(lx, ly, candidate);
  }

  // :404-589
  __device__ __forceinline__ void FirstLevelDivision(int blocks, int lx, int ly) {
    const int half = blocks / 2;
    const int bx = tbx + lx, by = tby + ly;
    if (CrossesHorizontal(bx, by, bx + blocks) ||
        CrossesHorizontal(bx, by + blocks, bx + blocks) ||
        CrossesVertical(bx, by, by + blocks) ||
        CrossesVertical(bx + blocks, by, by + blocks)) {
      return;
    }
    const bool allow_v = !CrossesVertical(bx + half, by, by + blocks);
    const bool allow_h = !CrossesHorizontal(bx, by + half, bx + blocks);
    float q00 = 0.0f, q01 = 0.0f, q10 = 0.0f, q11 = 0.0f;  // [row][col]
    for (int dy = 0; dy < blocks; ++dy) {
      for (int dx = 0; dx < blocks; ++dx) {
        const float c = Cost(lx + dx, ly + dy);
        if (dy < half) { if (dx < half) q00 += c; else q01 += c; }
        else { if (dx < half) q10 += c; else q11 += c; }
      }
    }
    constexpr float kUnavailable = 3.40282347e+38f;
    float vl = kUnavailable, vr = kUnavailable, ht = kUnavailable, hb = kUnavailable;
    const int vs = blocks == 2 ? kDct16x8 : kDct32x16;
    const int hs = blocks == 2 ? kDct8x16 : kDct16x32;
    const int ss = blocks == 2 ? kDct16x16 : kDct32x32;
    if (allow_v) {
      if ((Cell(lx, ly) >> 1) != vs) vl = Table(vs, lx, ly);
      if ((Cell(lx + half, ly) >> 1) != vs) vr = Table(vs, lx + half, ly);
    }
    if (allow_h) {
      if ((Cell(lx, ly) >> 1) != hs) ht = Table(hs, lx, ly);
      if ((Cell(lx, ly + half) >> 1) != hs) hb = Table(hs, lx, ly + half);
    }
    const float sq = Table(ss, lx, ly);
    const float vcost = StdMin(vl, q00 + q10) + StdMin(vr, q01 + q11);
    const float hcost = StdMin(ht, q00 + q01) + StdMin(hb, q10 + q11);
    if (sq < vcost && sq < hcost) {
      Set(lx, ly, ss);
      SetCostForTransform(lx, ly, ss, sq);
    } else if (vcost < hcost) {
      if (vl < q00 + q10) { Set(lx, ly, vs); SetCostForTransform(lx, ly, vs, vl); }
      if (vr < q01 + q11) { Set(lx + half, ly, vs); SetCostForTransform(lx + half, ly, vs, vr); }
    } else {
      if (ht < q00 + q01) { Set(lx, ly, hs); SetCostForTransform(lx, ly, hs, ht); }
      if (hb < q10 + q11) { Set(lx, ly + half, hs); SetCostForTransform(lx, ly + half, hs, hb); }
    }
  }
};

// grid: one work group of one wave per tile.  Lane i loads block i's seven
// candidate costs and initialises its state; the wave runs the search
// (src/codec/ac_strategy.cpp:591-773 SearchTile); lane i stores block i's
// cell.
__global__ __launch_bounds__(kSearchThreads) void SearchKernel(
  const float* __restrict__ costs, int bw, int bh, float target,
  uint8_t* __restrict__ out_cells) {
  const int tiles_w = (bw + 7) / 8;
  const int tile = blockIdx.x;
  TileSearch s;
  s.bw = bw;
  s.bh = bh;
  s.tbx = (tile % tiles_w) * 8;
  s.tby = (tile / tiles_w) * 8;
  s.tw = min(8, bw - s.tbx);
  s.th = min(8, bh - s.tby);
  const int i = threadIdx.x, lx = i % 8, ly = i / 8;
  s.lane = i;
  const bool inside = lx < s.tw && ly < s.th;
  const size_t plane = static_cast<size_t>(bw) * bh;
  const size_t block = static_cast<size_t>(s.tby + (inside ? ly : 0)) * bw + s.tbx + (inside ? lx : 0);
  float v[7];
#pragma unroll
  for (int st = 0; st < 7; ++st) v[st] = costs[st * plane + block];
#pragma unroll
  for (int st = 0; st < 7; ++st) s.table[st] = inside ? v[st] : 0.0f;
  s.cell = (kDct8 << 1) | 1;
  s.prio = 0;
  const float baseline = 1.0f + -0.4f / (target + 1.4f);
  s.cost = inside ? s.table[StageOf(kDct8)] * baseline : 0.0f;
  constexpr int kMergeStrategy[4] = {kDct16x8, kDct8x16, kDct16x32, kDct32x16};
  constexpr int kMergePriority[4] = {2, 2, 4, 4};
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int st = kMergeStrategy[m];
    const int cw = CoveredW(st), ch = CoveredH(st);
    for (int y = 0; y + ch <= s.th; y += ch) {
      for (int x = 0; x + cw <= s.tw; x += cw) {
        if (y + 3 < s.th && x + 3 < s.tw) {
          if (st == kDct16x32) {
            if ((y | x) % 4 == 0) s.FirstLevelDivision(4, x, y);
            continue;
          }
          if (st == kDct32x16) continue;
        }
        if ((st == kDct16x32 && y % 4 != 0) || (st == kDct32x16 && x % 4 != 0)) continue;
        if (y + 1 < s.th && x + 1 < s.tw) {
          if (st == kDct8x16) {
            if ((y | x) % 2 == 0) s.FirstLevelDivision(2, x, y);
            continue;
          }
          if (st == kDct16x8) continue;
        }
        if ((st == kDct8x16 && y % 2 == 1) || (st == kDct16x8 && x % 2 == 1)) continue;
        s.TryMerge(st, x, y, kMergePriority[m]);
      }
    }
  }
  for (int y = 0; y + 1 < s.th; ++y)
    for (int x = 0; x + 1 < s.tw; ++x)
      if ((y | x) % 2 != 0) s.FirstLevelDivision(2, x, y);
  for (int y = 0; y + 3 < s.th; y += 2)
    for (int x = 0; x + 3 < s.tw; x += 2)
      if ((y | x) % 4 != 0) s.FirstLevelDivision(4, x, y);
  if (inside) out_cells[block] = static_cast<uint8_t>(s.cell);
}

template <int kStrategy>
GJXLResult LaunchCost(GJXLAmdDevice* device, const float* opsin, int width,
                      int height, const float* qf, const float* mask,
                      const int8_t* yx, const int8_t* yb,
                      const CostParams& params, const uint8_t* allow, float* stage_costs) {
  using S = CostShape<kStrategy>;
  const int bw = width / 8, bh = height / 8;
  const dim3 grid(CeilDiv(S::kPositions, S::kCands), CeilDiv(bw, 8) * CeilDiv(bh, 8));
  const bool probed = ProbeBegin(device, GJXL_AMD_KERNEL_CANDIDATE_COST);
  hipLaunchKernelGGL(CandidateCostKernel<kStrategy>, grid, dim3(kCostThreads),
                     0, device->stream, opsin, width, height, qf, mask, yx, yb,
                     device->quant_tables, device->dct_bases, params, allow,
                     stage_costs);
  ProbeEnd(device, probed);
  return CheckLaunch("CandidateCostKernel");
}

// Entries of a strategy's candidate list: every raster position of every
// tile.
template <int kStrategy>
size_t ListCapacity(int width, int height) {
  return static_cast<size_t>(CeilDiv(width / 8, 8)) * CeilDiv(height / 8, 8) *
         CostShape<kStrategy>::kPositions;
}

// The pruned costs of one strategy by list: the allowed anchors are listed
// (CandidateListKernel, which also writes +infinity for the others), then a
// grid as large as the list can be takes them kCands at a time; the groups
// past the list's count exit at once.  *count is zero beforehand.
template <int kStrategy>
GJXLResult LaunchListedCost(GJXLAmdDevice* device, const float* opsin, int width, int height,
                            const float* qf, const float* mask, const int8_t* yx,
                            const int8_t* yb, const CostParams& params, const uint8_t* allow,
                            uint32_t* list, uint32_t* count, float* stage_costs) {
  using S = CostShape<kStrategy>;
  const size_t capacity = ListCapacity<kStrategy>(width, height);
  hipLaunchKernelGGL(CandidateListKernel<kStrategy>,
                     dim3(CeilDiv(static_cast<uint32_t>(capacity), 256u)), dim3(256), 0,
                     device->stream, allow, width / 8, height / 8, list, count, stage_costs);
  GJXL_AMD_TRY(CheckLaunch("CandidateListKernel"));
  const dim3 grid(CeilDiv(static_cast<uint32_t>(capacity), static_cast<uint32_t>(S::kCands)));
  const bool probed = ProbeBegin(device, GJXL_AMD_KERNEL_CANDIDATE_COST);
  hipLaunchKernelGGL(CandidateCostListKernel<kStrategy>, grid, dim3(kCostThreads), 0,
                     device->stream, opsin, width, height, qf, mask, yx, yb,
                     device->quant_tables, device->dct_bases, params, list, count, stage_costs);
  ProbeEnd(device, probed);
  return CheckLaunch("CandidateCostListKernel");
}

__global__ void ClearCountsKernel(uint32_t* __restrict__ counts, int n) {
  if (static_cast<int>(threadIdx.x) < n) counts[threadIdx.x] = 0;
}

}  // namespace

GJXLResult LaunchInitialCfl(GJXLAmdDevice* device, const float* opsin,
                            int width, int height, int8_t* y_to_x,
                            int8_t* y_to_b) {
  const dim3 grid(CeilDiv(width, 64), CeilDiv(height, 64));
  hipLaunchKernelGGL(InitialCflKernel, grid, dim3(kCflThreads), 0,
                     device->stream, opsin, width, height,
                     device->quant_tables, device->dct_bases, y_to_x, y_to_b);
  return CheckLaunch("InitialCflKernel");
}

GJXLResult LaunchCandidateCosts(GJXLAmdDevice* device, const float* opsin,
                                int width, int height, const float* qf,
                                const float* mask, const int8_t* yx,
                                const int8_t* yb, float target, const uint8_t* allow,
                                float* costs) {
  // src/codec/ac_strategy.cpp:884-891 (host libm, once per call)
  const float kBias = 0.13731742964354549f;
  const float ratio = (target + kBias) / (1.0f + kBias);
  CostParams p;
  p.info_loss_multiplier = 1.2f * std::pow(ratio, 0.33677806662454718f);
  p.zeros_multiplier = 9.3089059022677905f * std::pow(ratio, 0.50990926717963703f);
  p.cost_delta = 10.833273317067883f * std::pow(ratio, 0.36702940662370243f);
  const size_t stage = static_cast<size_t>(width / 8) * (height / 8);
  hipLaunchKernelGGL(FillNanKernel, dim3(static_cast<uint32_t>((7 * stage + 255) / 256)),
                     dim3(256), 0, device->stream, costs, 7 * stage);
  GJXL_AMD_TRY(CheckLaunch("FillNanKernel"));
  constexpr float kEntropy[7] = {1.0f, 1.21f, 1.21f, 1.34f, 1.49f, 1.49f, 1.48f};
  const size_t capacities[6] = {
    ListCapacity<kDct16x8>(width, height),  ListCapacity<kDct8x16>(width, height),
    ListCapacity<kDct16x16>(width, height), ListCapacity<kDct32x16>(width, height),
    ListCapacity<kDct16x32>(width, height), ListCapacity<kDct32x32>(width, height)};
  size_t words = 8;
  for (size_t c : capacities) words += c;
  // List entries hold 16-bit block coordinates and the list grids 32-bit
  // thread counts; larger frames (beyond any device's memory) take the
  // pruned tile form below, which gives the same costs.
  const bool listed = width / 8 <= 0xffff && height / 8 <= 0xffff && words < (size_t{1} << 31);
  if (allow != nullptr && listed) {
    // Pruned: DCT8 (allowed everywhere by the proposal's fallback) by tile,
    // the larger transforms by list.  One stream-ordered scratch holds the
    // six lists and their counts.
    void* scratch = nullptr;
    GJXL_AMD_TRY(DeviceAlloc(device, words * sizeof(uint32_t), &scratch));
    uint32_t* counts = static_cast<uint32_t*>(scratch);
    uint32_t* lists[6];
    lists[0] = counts + 8;
    for (int i = 1; i < 6; ++i) lists[i] = lists[i - 1] + capacities[i - 1];
    hipLaunchKernelGGL(ClearCountsKernel, dim3(1), dim3(64), 0, device->stream, counts, 8);
    GJXLResult r = CheckLaunch("ClearCountsKernel");
    p.entropy_multiplier = kEntropy[0];
    if (r == GJXL_OK) r = LaunchCost<kDct8>(device, opsin, width, height, qf, mask, yx, yb, p, allow, costs);
    p.entropy_multiplier = kEntropy[1];
    if (r == GJXL_OK) r = LaunchListedCost<kDct16x8>(device, opsin, width, height, qf, mask, yx, yb, p, allow, lists[0], counts + 0, costs + 1 * stage);
    p.entropy_multiplier = kEntropy[2];
    if (r == GJXL_OK) r = LaunchListedCost<kDct8x16>(device, opsin, width, height, qf, mask, yx, yb, p, allow, lists[1], counts + 1, costs + 2 * stage);
    p.entropy_multiplier = kEntropy[3];
    if (r == GJXL_OK) r = LaunchListedCost<kDct16x16>(device, opsin, width, height, qf, mask, yx, yb, p, allow, lists[2], counts + 2, costs + 3 * stage);
    p.entropy_multiplier = kEntropy[4];
    if (r == GJXL_OK) r = LaunchListedCost<kDct32x16>(device, opsin, width, height, qf, mask, yx, yb, p, allow, lists[3], counts + 3, costs + 4 * stage);
    p.entropy_multiplier = kEntropy[5];
    if (r == GJXL_OK) r = LaunchListedCost<kDct16x32>(device, opsin, width, height, qf, mask, yx, yb, p, allow, lists[4], counts + 4, costs + 5 * stage);
    p.entropy_multiplier = kEntropy[6];
    if (r == GJXL_OK) r = LaunchListedCost<kDct32x32>(device, opsin, width, height, qf, mask, yx, yb, p, allow, lists[5], counts + 5, costs + 6 * stage);
    DeviceFree(device, scratch);
    return r;
  }
  p.entropy_multiplier = kEntropy[0];
  GJXL_AMD_TRY(LaunchCost<kDct8>(device, opsin, width, height, qf, mask, yx, yb, p, allow, costs + 0 * stage));
  p.entropy_multiplier = kEntropy[1];
  GJXL_AMD_TRY(LaunchCost<kDct16x8>(device, opsin, width, height, qf, mask, yx, yb, p, allow, costs + 1 * stage));
  p.entropy_multiplier = kEntropy[2];
  GJXL_AMD_TRY(LaunchCost<kDct8x16>(device, opsin, width, height, qf, mask, yx, yb, p, allow, costs + 2 * stage));
  p.entropy_multiplier = kEntropy[3];
  GJXL_AMD_TRY(LaunchCost<kDct16x16>(device, opsin, width, height, qf, mask, yx, yb, p, allow, costs + 3 * stage));
  p.entropy_multiplier = kEntropy[4];
  GJXL_AMD_TRY(LaunchCost<kDct32x16>(device, opsin, width, height, qf, mask, yx, yb, p, allow, costs + 4 * stage));
  p.entropy_multiplier = kEntropy[5];
  GJXL_AMD_TRY(LaunchCost<kDct16x32>(device, opsin, width, height, qf, mask, yx, yb, p, allow, costs + 5 * stage));
  p.entropy_multiplier = kEntropy[6];
  return LaunchCost<kDct32x32>(device, opsin, width, height, qf, mask, yx, yb, p, allow,
                               costs + 6 * stage);
}

GJXLResult LaunchSearch(GJXLAmdDevice* device, const float* costs, int width,
                        int height, float target, uint8_t* cells) {
  const int bw = width / 8, bh = height / 8;
  const int tiles = CeilDiv(bw, 8) * CeilDiv(bh, 8);
  hipLaunchKernelGGL(SearchKernel, dim3(tiles),
                     dim3(kSearchThreads), 0, device->stream, costs, bw, bh,
                     target, cells);
  return CheckLaunch("SearchKernel");
}

// The cost parameters of a Butteraugli target
// (src/codec/ac_strategy.cpp:884-891, host libm).
CostParams TargetCostParams(float target) {
  const float kBias = 0.13731742964354549f;
  const float ratio = (target + kBias) / (1.0f + kBias);
  CostParams p;
  p.info_loss_multiplier = 1.2f * std::pow(ratio, 0.33677806662454718f);
  p.zeros_multiplier = 9.3089059022677905f * std::pow(ratio, 0.50990926717963703f);
  p.cost_delta = 10.833273317067883f * std::pow(ratio, 0.36702940662370243f);
  p.entropy_multiplier = 1.0f;
  return p;
}

constexpr float kStageEntropy[7] = {1.0f, 1.21f, 1.21f, 1.34f, 1.49f, 1.49f, 1.48f};

__global__ void QueueResetKernel(QueueTable* __restrict__ t, int stage) {
  if (threadIdx.x < kQueueXcds) t->counters[stage][threadIdx.x] = 0;
}

GJXLResult LaunchQueueStage(GJXLAmdDevice* d, hipStream_t st, QueueTable* t, int stage,
                            uint32_t grid) {
  const dim3 g(grid), b(kCostThreads);
  const float e = kStageEntropy[stage];
  switch (stage) {
    case 0: hipLaunchKernelGGL(CostQueueKernel<kDct8>, g, b, 0, st, t, d->quant_tables, d->dct_bases, e); break;
    case 1: hipLaunchKernelGGL(CostQueueKernel<kDct16x8>, g, b, 0, st, t, d->quant_tables, d->dct_bases, e); break;
    case 2: hipLaunchKernelGGL(CostQueueKernel<kDct8x16>, g, b, 0, st, t, d->quant_tables, d->dct_bases, e); break;
    case 3: hipLaunchKernelGGL(CostQueueKernel<kDct16x16>, g, b, 0, st, t, d->quant_tables, d->dct_bases, e); break;
    case 4: hipLaunchKernelGGL(CostQueueKernel<kDct32x16>, g, b, 0, st, t, d->quant_tables, d->dct_bases, e); break;
    case 5: hipLaunchKernelGGL(CostQueueKernel<kDct16x32>, g, b, 0, st, t, d->quant_tables, d->dct_bases, e); break;
    default: hipLaunchKernelGGL(CostQueueKernel<kDct32x32>, g, b, 0, st, t, d->quant_tables, d->dct_bases, e); break;
  }
  return CheckLaunch("CostQueueKernel");
}

// plan, NaN fill, then the seven strategies' queues (kernels only: the
// sequence is what the graph captures)
GJXLResult LaunchQueue(GJXLAmdDevice* d, hipStream_t st, QueueTable* t, const uint32_t* grids) {
  hipLaunchKernelGGL(QueuePlanKernel, dim3(1), dim3(64), 0, st, t);
  GJXL_AMD_TRY(CheckLaunch("QueuePlanKernel"));
  hipLaunchKernelGGL(QueueFillKernel, dim3(1024), dim3(256), 0, st, t);
  GJXL_AMD_TRY(CheckLaunch("QueueFillKernel"));
  // the largest transforms first: their items are the longest
  constexpr int kOrder[7] = {6, 3, 4, 5, 1, 2, 0};
  for (int i = 0; i < 7; ++i) GJXL_AMD_TRY(LaunchQueueStage(d, st, t, kOrder[i], grids[kOrder[i]]));
  return GJXL_OK;
}

}  // namespace gjxl_amd

// A tile queue for the candidate costs of batches of frames: its device
// table, a pinned host copy, the persistent grid of each strategy and the
// captured launch sequence.
struct GJXLAmdCostQueue {
  GJXLAmdDevice* device = nullptr;
  uint32_t max_frames = 0;
  gjxl_amd::QueueTable* table = nullptr;
  gjxl_amd::QueueTable* host = nullptr;
  hipEvent_t copied = nullptr;
  bool copy_pending = false;
  uint32_t grids[7] = {};
  float tuned_ms[7] = {};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};


using gjxl_amd::Fail;

namespace {
GJXLResult ValidateGeometry(uint32_t width, uint32_t height) {
  if (width == 0 || height == 0 || width % 8 != 0 || height % 8 != 0 ||
      width > (1u << 28) || height > (1u << 28)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                "Padded image dimensions must be non-zero multiples of 8");
  }
  return GJXL_OK;
}

GJXLResult ValidateTarget(float target) {
  if (!std::isfinite(target) || target <= 0.0f) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                "Butteraugli target must be finite and positive");
  }
  return GJXL_OK;
}
}  // namespace

namespace gjxl_amd {
namespace {

void DropGraph(GJXLAmdCostQueue* q) {
  if (q->exec != nullptr) (void)hipGraphExecDestroy(q->exec);
  if (q->graph != nullptr) (void)hipGraphDestroy(q->graph);
  q->exec = nullptr;
  q->graph = nullptr;
}

GJXLResult CaptureQueue(GJXLAmdCostQueue* q) {
  DropGraph(q);
  hipStream_t st = q->device->stream;
  GJXL_AMD_TRY(Check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal),
                     "hipStreamBeginCapture"));
  const GJXLResult launched = LaunchQueue(q->device, st, q->table, q->grids);
  hipGraph_t graph = nullptr;
  const GJXLResult ended = Check(hipStreamEndCapture(st, &graph), "hipStreamEndCapture");
  GJXL_AMD_TRY(launched);
  GJXL_AMD_TRY(ended);
  q->graph = graph;
  return Check(hipGraphInstantiate(&q->exec, q->graph, nullptr, nullptr, 0),
               "hipGraphInstantiate");
}

// The batch into the host table (validated), then to the device table.
GJXLResult UploadFrames(GJXLAmdCostQueue* q, const GJXLAmdCostFrame* frames, uint32_t count) {
  if (frames == nullptr || count == 0 || count > q->max_frames) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "cost queue: 1 to max_frames frames");
  }
  uint64_t items = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const GJXLAmdCostFrame& f = frames[i];
    if (f.opsin == nullptr || f.quant_field == nullptr || f.pixel_mask == nullptr ||
        f.y_to_x == nullptr || f.y_to_b == nullptr || f.costs == nullptr) {
      return Fail(GJXL_ERROR_INVALID_ARGUMENT, "cost queue: frame buffers must be non-null");
    }
    GJXL_AMD_TRY(ValidateGeometry(f.width, f.height));
    if (f.width > 65536 || f.height > 65536) {
      return Fail(GJXL_ERROR_INVALID_ARGUMENT, "cost queue: frames up to 65536 pixels a side");
    }
    GJXL_AMD_TRY(ValidateTarget(f.butteraugli_target));
    items += 16ull * CeilDiv(f.width / 8, 8) * CeilDiv(f.height / 8, 8);
  }
  if (items >= (1ull << 31)) return Fail(GJXL_ERROR_INVALID_ARGUMENT, "cost queue: batch too large");
  if (q->copy_pending) {
    GJXL_AMD_TRY(Check(hipEventSynchronize(q->copied), "hipEventSynchronize"));
    q->copy_pending = false;
  }
  QueueTable* h = q->host;
  h->count = count;
  for (uint32_t i = 0; i < count; ++i) {
    const GJXLAmdCostFrame& f = frames[i];
    QueueFrame& o = h->frames[i];
    o.opsin = f.opsin;
    o.quant_field = f.quant_field;
    o.pixel_mask = f.pixel_mask;
    o.y_to_x = f.y_to_x;
    o.y_to_b = f.y_to_b;
    o.allow = f.allow;
    o.costs = f.costs;
    o.width = static_cast<int32_t>(f.width);
    o.height = static_cast<int32_t>(f.height);
    o.params = TargetCostParams(f.butteraugli_target);
  }
  hipStream_t st = q->device->stream;
  const size_t bytes = offsetof(QueueTable, frames) + count * sizeof(QueueFrame);
  GJXL_AMD_TRY(Check(hipMemcpyAsync(q->table, h, bytes, hipMemcpyHostToDevice, st),
                     "hipMemcpyAsync(cost queue)"));
  GJXL_AMD_TRY(Check(hipEventRecord(q->copied, st), "hipEventRecord"));
  q->copy_pending = true;
  return GJXL_OK;
}

}  // namespace
}  // namespace gjxl_amd

extern "C" {

GJXLResult gjxl_amd_initial_cfl(GJXLAmdDevice* device, const float* opsin,
                                uint32_t width, uint32_t height,
                                int8_t* y_to_x, int8_t* y_to_b) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  GJXL_AMD_TRY(ValidateGeometry(width, height));
  if (opsin == nullptr || y_to_x == nullptr || y_to_b == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "CfL buffers must be non-null");
  }
  return gjxl_amd::LaunchInitialCfl(device, opsin, width, height, y_to_x, y_to_b);
}

GJXLResult gjxl_amd_ac_strategy_costs(GJXLAmdDevice* device,
                                      const float* opsin, uint32_t width,
                                      uint32_t height, const float* quant_field,
                                      const float* pixel_mask,
                                      const int8_t* y_to_x,
                                      const int8_t* y_to_b,
                                      float butteraugli_target, float* costs) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  GJXL_AMD_TRY(ValidateGeometry(width, height));
  GJXL_AMD_TRY(ValidateTarget(butteraugli_target));
  if (opsin == nullptr || quant_field == nullptr || pixel_mask == nullptr ||
      y_to_x == nullptr || y_to_b == nullptr || costs == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "AC cost buffers must be non-null");
  }
  return gjxl_amd::LaunchCandidateCosts(device, opsin, width, height,
                                        quant_field, pixel_mask, y_to_x,
                                        y_to_b, butteraugli_target, nullptr, costs);
}

GJXLResult gjxl_amd_ac_strategy_costs_pruned(GJXLAmdDevice* device, const float* opsin,
                                             uint32_t width, uint32_t height,
                                             const float* quant_field,
                                             const float* pixel_mask, const int8_t* y_to_x,
                                             const int8_t* y_to_b, float butteraugli_target,
                                             const uint8_t* allow, float* costs) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  GJXL_AMD_TRY(ValidateGeometry(width, height));
  GJXL_AMD_TRY(ValidateTarget(butteraugli_target));
  if (opsin == nullptr || quant_field == nullptr || pixel_mask == nullptr ||
      y_to_x == nullptr || y_to_b == nullptr || costs == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "AC cost buffers must be non-null");
  }
  return gjxl_amd::LaunchCandidateCosts(device, opsin, width, height, quant_field, pixel_mask,
                                        y_to_x, y_to_b, butteraugli_target, allow, costs);
}

GJXLResult gjxl_amd_ac_strategy_search(GJXLAmdDevice* device,
                                       const float* costs, uint32_t width,
                                       uint32_t height,
                                       float butteraugli_target,
                                       uint8_t* cells) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  GJXL_AMD_TRY(ValidateGeometry(width, height));
  GJXL_AMD_TRY(ValidateTarget(butteraugli_target));
  if (costs == nullptr || cells == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "AC search buffers must be non-null");
  }
  return gjxl_amd::LaunchSearch(device, costs, width, height,
                                butteraugli_target, cells);
}

GJXLResult gjxl_amd_cost_queue_create(GJXLAmdDevice* device, uint32_t max_frames,
                                      GJXLAmdCostQueue** out) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  if (out == nullptr || max_frames == 0 ||
      max_frames > static_cast<uint32_t>(gjxl_amd::kQueueMaxFrames)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "cost queue: 1 to 256 frames");
  }
  *out = nullptr;
  auto* q = new GJXLAmdCostQueue();
  q->device = device;
  q->max_frames = max_frames;
  int cus = 0;
  GJXLResult r = gjxl_amd::Check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                                       device->ordinal),
                                 "hipDeviceGetAttribute");
  for (int i = 0; i < 7; ++i) q->grids[i] = 2 * static_cast<uint32_t>(cus > 0 ? cus : 256);
  void* table = nullptr;
  if (r == GJXL_OK) r = gjxl_amd::DeviceAlloc(device, sizeof(gjxl_amd::QueueTable), &table);
  q->table = static_cast<gjxl_amd::QueueTable*>(table);
  if (r == GJXL_OK) {
    r = gjxl_amd::Check(hipHostMalloc(reinterpret_cast<void**>(&q->host),
                                      sizeof(gjxl_amd::QueueTable), hipHostMallocDefault),
                        "hipHostMalloc(cost queue)");
  }
  if (r == GJXL_OK) r = gjxl_amd::Check(hipEventCreateWithFlags(&q->copied, hipEventDisableTiming),
                                        "hipEventCreate");
  if (r == GJXL_OK) r = gjxl_amd::CaptureQueue(q);
  if (r != GJXL_OK) {
    gjxl_amd_cost_queue_destroy(q);
    return r;
  }
  *out = q;
  return GJXL_OK;
}

void gjxl_amd_cost_queue_destroy(GJXLAmdCostQueue* q) {
  if (q == nullptr) return;
  (void)hipStreamSynchronize(q->device->stream);
  gjxl_amd::DropGraph(q);
  if (q->copied != nullptr) (void)hipEventDestroy(q->copied);
  if (q->host != nullptr) (void)hipHostFree(q->host);
  gjxl_amd::DeviceFree(q->device, q->table);
  delete q;
}

GJXLResult gjxl_amd_cost_queue_run(GJXLAmdCostQueue* q, const GJXLAmdCostFrame* frames,
                                   uint32_t count) {
  if (q == nullptr || q->exec == nullptr) return Fail(GJXL_ERROR_INVALID_ARGUMENT, "cost queue is null");
  GJXL_AMD_TRY(gjxl_amd::UploadFrames(q, frames, count));
  return gjxl_amd::Check(hipGraphLaunch(q->exec, q->device->stream), "hipGraphLaunch");
}

GJXLResult gjxl_amd_cost_queue_tune(GJXLAmdCostQueue* q, const GJXLAmdCostFrame* frames,
                                    uint32_t count, uint32_t repetitions) {
  using namespace gjxl_amd;
  if (q == nullptr) return Fail(GJXL_ERROR_INVALID_ARGUMENT, "cost queue is null");
  if (repetitions == 0 || repetitions > 100) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "cost queue tune: 1 to 100 repetitions");
  }
  GJXL_AMD_TRY(UploadFrames(q, frames, count));
  GJXLAmdDevice* d = q->device;
  hipStream_t st = d->stream;
  hipLaunchKernelGGL(QueuePlanKernel, dim3(1), dim3(64), 0, st, q->table);
  GJXL_AMD_TRY(CheckLaunch("QueuePlanKernel"));
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d->ordinal);
  uint64_t tiles = 0;
  for (uint32_t i = 0; i < count; ++i) {
    tiles += static_cast<uint64_t>(CeilDiv(frames[i].width / 8, 8)) * CeilDiv(frames[i].height / 8, 8);
  }
  hipEvent_t begin = nullptr, end = nullptr;
  GJXL_AMD_TRY(Check(hipEventCreate(&begin), "hipEventCreate"));
  GJXLResult r = Check(hipEventCreate(&end), "hipEventCreate");
  constexpr uint32_t kChunks[7] = {QueueChunks<kDct8>(), QueueChunks<kDct16x8>(),
                                   QueueChunks<kDct8x16>(), QueueChunks<kDct16x16>(),
                                   QueueChunks<kDct32x16>(), QueueChunks<kDct16x32>(),
                                   QueueChunks<kDct32x32>()};
  for (int s = 0; r == GJXL_OK && s < 7; ++s) {
    // persistent grids of one to four groups per CU, and one group per item
    const uint64_t items = kChunks[s] * tiles;
    const uint64_t variants[5] = {static_cast<uint64_t>(cus), 2ull * cus, 3ull * cus,
                                  4ull * cus, items};
    float best = 0.0f;
    for (uint64_t grid : variants) {
      grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, items));
      float total = 0.0f;
      for (uint32_t k = 0; r == GJXL_OK && k <= repetitions; ++k) {
        hipLaunchKernelGGL(QueueResetKernel, dim3(1), dim3(64), 0, st, q->table, s);
        r = CheckLaunch("QueueResetKernel");
        if (r == GJXL_OK) r = Check(hipEventRecord(begin, st), "hipEventRecord");
        if (r == GJXL_OK) r = LaunchQueueStage(d, st, q->table, s, static_cast<uint32_t>(grid));
        if (r == GJXL_OK) r = Check(hipEventRecord(end, st), "hipEventRecord");
        if (r == GJXL_OK) r = Check(hipEventSynchronize(end), "hipEventSynchronize");
        float ms = 0.0f;
        if (r == GJXL_OK) r = Check(hipEventElapsedTime(&ms, begin, end), "hipEventElapsedTime");
        if (k > 0) total += ms;  // the first run warms up
      }
      const float mean = total / static_cast<float>(repetitions);
      if (r == GJXL_OK && (best == 0.0f || mean < best)) {
        best = mean;
        q->grids[s] = static_cast<uint32_t>(grid);
        q->tuned_ms[s] = mean;
      }
    }
  }
  (void)hipEventDestroy(begin);
  (void)hipEventDestroy(end);
  GJXL_AMD_TRY(r);
  return CaptureQueue(q);
}

GJXLResult gjxl_amd_cost_queue_grids(GJXLAmdCostQueue* q, uint32_t* grids, float* milliseconds) {
  if (q == nullptr || grids == nullptr) return Fail(GJXL_ERROR_INVALID_ARGUMENT, "cost queue is null");
  for (int i = 0; i < 7; ++i) {
    grids[i] = q->grids[i];
    if (milliseconds != nullptr) milliseconds[i] = q->tuned_ms[i];
  }
  return GJXL_OK;
}

}  // extern "C"
