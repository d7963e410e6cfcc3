// AC tokenization of a quantized frame (src/codestream/ac_group.cpp:740-932,
// TokenizeSimpleAcGroupDirectValidated), as four passes over the per-group
// coefficient store:
//
//   count  one wave per (transform, channel): the non-LLF nonzero count and
//          the last nonzero scan position (coalesced reads of the transform,
//          positions through the inverse order), the scaled nonzero map;
//   scan   per AC group, the exclusive prefix of the token counts of its
//          (transform, channel) tasks in stream order; then over the groups;
//   emit   one wave per task: the nonzero token, then 64 scan positions per
//          step, each lane's remaining-nonzero count from a ballot prefix.
//
// Within a group the stream order is anchors in raster order, channels Y, X,
// B.  The nonzero prediction of a transform reads the map of its left and
// upper neighbours, which belong to earlier transforms of the group and are
// final once the count pass is done, so the emit pass sees the values the
// sequential tokenizer sees.
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "aq_common.h"
#include "tokenize.h"
#include "common.h"

namespace gjxl_amd {
namespace {

constexpr int kOrderFamilies = 13;
constexpr int kNonzeroBuckets = 37;       // NonzeroBucket range
constexpr int kZeroDensityContexts = 458;  // ZeroDensityContext range
constexpr int kSymbols = 128;              // prefix alphabet
constexpr int kMaxThresholds = 15;
constexpr int kMaxBlockContexts = 16;
constexpr int kMapEntries = 3 * kOrderFamilies * (kMaxThresholds + 1);
constexpr int kGroupTasks = 3 * kGroupAnchors;
constexpr int kTaskWaves = 4;              // waves per work group
constexpr int kOrderStride = 2624;         // one channel of all seven orders

// Stage s of the order table holds the strategy StageStrategy(s); its orders
// start at 3 * StageOffset(s), channel c at + c * size.
__host__ __device__ constexpr int StageStrategy(int stage) {
  return stage == 0 ? kDct8 : stage == 1 ? kDct16x8 : stage == 2 ? kDct8x16
       : stage == 3 ? kDct16x16 : stage == 4 ? kDct32x16 : stage == 5 ? kDct16x32
       : kDct32x32;
}
__host__ __device__ constexpr int StageOffset(int stage) {
  return stage == 0 ? 0 : stage == 1 ? 64 : stage == 2 ? 192 : stage == 3 ? 320
       : stage == 4 ? 576 : stage == 5 ? 1088 : 1600;
}
// src/codestream/simple_ac_context.h:20-23 for the searched strategies.
__host__ __device__ constexpr int OrderFamily(int strategy) {
  return strategy == kDct8 ? 0 : strategy == kDct16x16 ? 2 : strategy == kDct32x32 ? 3
       : (strategy == kDct16x8 || strategy == kDct8x16) ? 4 : 6;
}

// src/codestream/ac_group.cpp:40-53
__constant__ uint16_t kFrequencyContext[64] = {
  0xBAD, 0,  1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14,
  15,    15, 16, 16, 17, 17, 18, 18, 19, 19, 20, 20, 21, 21, 22, 22,
  23,    23, 23, 23, 24, 24, 24, 24, 25, 25, 25, 25, 26, 26, 26, 26,
  27,    27, 27, 27, 28, 28, 28, 28, 29, 29, 29, 29, 30, 30, 30, 30,
};
__constant__ uint16_t kNonzeroContext[64] = {
  0xBAD, 0,   31,  62,  62,  93,  93,  93,  93,  123, 123, 123, 123,
  152,   152, 152, 152, 152, 152, 152, 152, 180, 180, 180, 180, 180,
  180,   180, 180, 180, 180, 180, 180, 206, 206, 206, 206, 206, 206,
  206,   206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206,
  206,   206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206,
};

// The block-context map, passed by value (kernel arguments).
struct ContextModel {
  uint8_t map[kMapEntries];
  uint32_t thresholds[kMaxThresholds + 1];
  uint32_t threshold_count;
  uint32_t num_contexts;
};

constexpr uint32_t kGridError = 1u;
constexpr uint32_t kRawQuantError = 2u;

__device__ __forceinline__ int TaskChannel(int task) {
  const int k = task % 3;
  return k == 0 ? 1 : k == 1 ? 0 : 2;
}

__device__ __forceinline__ int Log2Covered(int covered) {
  return covered == 1 ? 0 : covered == 2 ? 1 : covered == 4 ? 2 : covered == 8 ? 3 : 4;
}

// src/codestream/entropy.h: zig-zag of a signed value.
__device__ __forceinline__ uint32_t PackSigned(int32_t v) {
  return (static_cast<uint32_t>(v) << 1) ^ ((static_cast<uint32_t>(~v) >> 31) - 1u);
}

// AppendDirectAcToken (:688-720): the HybridUint(4, 2, 0) symbol.
__device__ __forceinline__ uint32_t TokenSymbol(uint32_t value) {
  if (value < 16) return value;
  const uint32_t e = 31u - static_cast<uint32_t>(__clz(value));
  return 16u + ((e - 4u) << 2) + ((value - (1u << e)) >> (e - 2u));
}

__global__ void FillKernel(uint32_t* __restrict__ p, size_t n, uint32_t value) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    p[i] = value;
  }
}

// count pass: grid (kGroupTasks / kTaskWaves, groups), kTaskWaves waves.
__global__ __launch_bounds__(64 * kTaskWaves) void TokenCountKernel(
  const int32_t* __restrict__ quantized, size_t store,
  const uint8_t* __restrict__ cells, int bw, int bh,
  const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ anchor_list,
  const uint32_t* __restrict__ anchor_count,
  const uint16_t* __restrict__ inverse_orders, const int32_t* __restrict__ raw,
  uint32_t* __restrict__ status, uint32_t* __restrict__ tasks,
  uint8_t* __restrict__ nonzero_map) {
  const int g = blockIdx.y;
  const int task = blockIdx.x * kTaskWaves + static_cast<int>(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (task >= 3 * static_cast<int>(anchor_count[g])) return;  // whole waves
  const uint32_t packed = anchor_list[static_cast<size_t>(g) * kGroupAnchors + task / 3];
  const int bx = packed & 0xffff, by = packed >> 16;
  const size_t block = static_cast<size_t>(by) * bw + bx;
  const int st = cells[block] >> 1;
  const int cw = CoveredW(st), ch = CoveredH(st), covered = cw * ch;
  const int n = covered * 64;
  const int c = TaskChannel(task);
  const int32_t* coef = quantized + c * store + offsets[block];
  const uint16_t* inverse =
    inverse_orders + 3 * StageOffset(StrategyStage(st)) + c * n;
  uint32_t nonzeros = 0, last = 0;
  for (int i = lane; i < n; i += 64) {
    const int32_t v = coef[i];
    const uint32_t scan = inverse[i];
    if (v != 0 && scan >= static_cast<uint32_t>(covered)) {
      ++nonzeros;
      last = max(last, scan);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    nonzeros += __shfl_xor(nonzeros, off);
    last = max(last, __shfl_xor(last, off));
  }
  if (lane == 0) {
    const uint32_t tokens = nonzeros ? last - covered + 2 : 1u;
    tasks[static_cast<size_t>(g) * kGroupTasks + task] = nonzeros | (tokens << 16);
    if (c == 1) {
      const int32_t q = raw[block];
      if (q < 1 || q > 256) atomicOr(&status[1], kRawQuantError);
    }
  }
  if (lane < covered) {
    const uint32_t scaled = (nonzeros + covered - 1) >> Log2Covered(covered);
    nonzero_map[static_cast<size_t>(c) * bw * bh +
                static_cast<size_t>(by + lane / cw) * bw + bx + lane % cw] =
      static_cast<uint8_t>(scaled);
  }
}

// scan pass 1: one work group per AC group.
__global__ __launch_bounds__(kGroupAnchors) void TokenGroupScanKernel(
  const uint32_t* __restrict__ anchor_count, const uint32_t* __restrict__ tasks,
  uint32_t* __restrict__ task_offsets, uint32_t* __restrict__ group_tokens) {
  const int g = blockIdx.x, t = threadIdx.x;
  const int count = 3 * static_cast<int>(anchor_count[g]);
  const size_t base = static_cast<size_t>(g) * kGroupTasks + 3 * t;
  uint32_t own[3] = {0, 0, 0};
  for (int j = 0; j < 3; ++j) {
    if (3 * t + j < count) own[j] = tasks[base + j] >> 16;
  }
  uint32_t total;
  uint32_t before = BlockExclusiveScanU32(own[0] + own[1] + own[2], &total);
  for (int j = 0; j < 3; ++j) {
    if (3 * t + j < count) task_offsets[base + j] = before;
    before += own[j];
  }
  if (t == 0) group_tokens[g] = total;
}

// scan pass 2: one work group over all groups.
__global__ __launch_bounds__(1024) void TokenFrameScanKernel(
  const uint32_t* __restrict__ group_tokens, int groups,
  uint32_t* __restrict__ group_offsets, uint32_t* __restrict__ status) {
  uint32_t carry = 0;
  for (int first = 0; first < groups; first += blockDim.x) {
    const int g = first + threadIdx.x;
    const uint32_t v = g < groups ? group_tokens[g] : 0u;
    uint32_t total;
    const uint32_t before = BlockExclusiveScanU32(v, &total);
    if (g < groups) group_offsets[g] = carry + before;
    carry += total;
  }
  if (threadIdx.x == 0) {
    group_offsets[groups] = carry;
    status[0] = carry;
  }
}

// emit pass: grid (kGroupTasks / kTaskWaves, groups), kTaskWaves waves.
__global__ __launch_bounds__(64 * kTaskWaves) void TokenEmitKernel(
  const int32_t* __restrict__ quantized, size_t store,
  const uint8_t* __restrict__ cells, int bw, int bh,
  const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ anchor_list,
  const uint32_t* __restrict__ anchor_count, const uint32_t* __restrict__ orders,
  const int32_t* __restrict__ raw, const uint32_t* __restrict__ tasks,
  const uint32_t* __restrict__ task_offsets,
  const uint32_t* __restrict__ group_offsets,
  const uint8_t* __restrict__ nonzero_map, const ContextModel model,
  uint32_t* __restrict__ values, uint16_t* __restrict__ contexts,
  uint32_t* __restrict__ histogram) {
  const int g = blockIdx.y;
  const int task = blockIdx.x * kTaskWaves + static_cast<int>(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (task >= 3 * static_cast<int>(anchor_count[g])) return;  // whole waves
  const uint32_t packed = anchor_list[static_cast<size_t>(g) * kGroupAnchors + task / 3];
  const int bx = packed & 0xffff, by = packed >> 16;
  const size_t block = static_cast<size_t>(by) * bw + bx;
  const int st = cells[block] >> 1;
  const int cw = CoveredW(st), ch = CoveredH(st), covered = cw * ch;
  const int log2c = Log2Covered(covered);
  const int n = covered * 64;
  const int c = TaskChannel(task);
  const size_t slot = static_cast<size_t>(g) * kGroupTasks + task;
  const uint32_t meta = tasks[slot];
  const uint32_t nonzeros = meta & 0xffff, tokens = meta >> 16;
  const size_t out = static_cast<size_t>(group_offsets[g]) + task_offsets[slot];
  // PredictNonzeros (:102-113) in group coordinates
  const uint8_t* map = nonzero_map + static_cast<size_t>(c) * bw * bh;
  const bool left_edge = (bx & (kGroupBlocks - 1)) == 0;
  const bool top_edge = (by & (kGroupBlocks - 1)) == 0;
  uint32_t prediction;
  if (left_edge) {
    prediction = top_edge ? 32u : map[block - bw];
  } else if (top_edge) {
    prediction = map[block - 1];
  } else {
    prediction = (map[block - bw] + map[block - 1] + 1u) / 2u;
  }
  // BlockContextValidated (:115-131)
  const uint32_t q = static_cast<uint32_t>(raw[block]);
  uint32_t segment = 0;
  while (segment < model.threshold_count && q > model.thresholds[segment]) ++segment;
  const int row = c < 2 ? (c ^ 1) : 2;
  const uint32_t block_context =
    model.map[(row * kOrderFamilies + OrderFamily(st)) * (model.threshold_count + 1) + segment];
  if (lane == 0) {
    const uint32_t bucket = prediction < 8 ? prediction : prediction >= 64 ? 36u : 4u + prediction / 2u;
    const uint32_t ctx = bucket * model.num_contexts + block_context;
    values[out] = nonzeros;
    contexts[out] = static_cast<uint16_t>(ctx);
    if (histogram) atomicAdd(&histogram[ctx * kSymbols + TokenSymbol(nonzeros)], 1u);
  }
  if (nonzeros == 0) return;
  const int last = covered + static_cast<int>(tokens) - 2;
  const uint32_t* order = orders + 3 * StageOffset(StrategyStage(st)) + c * n;
  const int32_t* coef = quantized + c * store + offsets[block];
  const uint32_t context_base =
    model.num_contexts * kNonzeroBuckets + kZeroDensityContexts * block_context;
  uint32_t remaining = nonzeros;
  uint32_t previous = nonzeros > static_cast<uint32_t>(n / 16) ? 0u : 1u;
  const uint64_t below_mask = (uint64_t{1} << lane) - 1u;
  for (int first = covered; first <= last; first += 64) {
    const int scan = first + lane;
    const bool active = scan <= last;
    const int32_t v = active ? coef[order[scan]] : 0;
    const uint64_t nz = __ballot(v != 0);
    const uint32_t rem = remaining - static_cast<uint32_t>(__popcll(nz & below_mask));
    const uint32_t prev =
      lane == 0 ? previous : static_cast<uint32_t>((nz >> ((lane - 1) & 63)) & 1u);
    if (active) {
      const uint32_t local =
        (kNonzeroContext[(rem + covered - 1) >> log2c] + kFrequencyContext[scan >> log2c]) * 2u + prev;
      const uint32_t value = PackSigned(v);
      const uint32_t ctx = context_base + local;
      const size_t at = out + 1 + static_cast<size_t>(scan - covered);
      values[at] = value;
      contexts[at] = static_cast<uint16_t>(ctx);
      if (histogram) atomicAdd(&histogram[ctx * kSymbols + TokenSymbol(value)], 1u);
    }
    remaining -= static_cast<uint32_t>(__popcll(nz));
    previous = static_cast<uint32_t>(nz >> 63);
  }
}

// ---- frame statistics (CountGroupZeros, CountRawQuantAndOrders) ------------

// Word layout of GJXL_AMD_TOKENS_STATISTICS, the fields of
// GJXLAmdFrameStatistics in order, then the present mask and error bits.
constexpr int kZeroCounters = GJXL_AMD_ZERO_COUNTS;
constexpr int kStatSampled = 3 * kZeroCounters;
constexpr int kStatRaw = kStatSampled + 3 * 64;
constexpr int kStatPresent = kStatRaw + kOrderFamilies * 256;
constexpr int kStatError = kStatPresent + 1;
constexpr int kStatWords = kStatError + 1;
constexpr int kSearchedFamilies = 5;

// The counters of the searched families within one channel's 1984: DCT8,
// the 16x8 pair, 16x16, the 32x16 pair, 32x32.
__host__ __device__ constexpr int ZeroCounterOffset(int family) {
  return family == 0 ? 0 : family == 4 ? 64 : family == 2 ? 192 : family == 6 ? 448 : 960;
}
__host__ __device__ constexpr int CompactFamily(int family) {
  return family == 0 ? 0 : family == 4 ? 1 : family == 2 ? 2 : family == 6 ? 3 : 4;
}
__host__ __device__ constexpr int FullFamily(int compact) {
  return compact == 0 ? 0 : compact == 1 ? 4 : compact == 2 ? 2 : compact == 3 ? 6 : 3;
}

// Work group (AC group g, chunk of its anchors): three waves, wave c counts
// channel c.  A lane owns positions lane + 64 m of every family and keeps
// their zero counts in registers (31 counters: 1, 2, 4, 8 and 16 per
// family), so a transform costs its coalesced loads and no atomics; the
// counters go to the frame's words once per wave.  The work group's threads
// also count the chunk's blocks per order family and raw quant (LDS), and a
// DCT8 transform whose sample decision is set counts into the sampled
// counters too.  The chunk's transforms are first listed in LDS by order
// family (offset and sample decision, read by the first wave), so that each
// wave walks them family by family with several transforms' loads in
// flight.
constexpr int kStatChunks = 16;
constexpr int kStatThreads = 3 * 64;

constexpr int kStatMaxChunk = (kGroupAnchors + kStatChunks - 1) / kStatChunks;
static_assert(kStatMaxChunk <= 64, "one lane per transform of a chunk");

__global__ __launch_bounds__(kStatThreads) void FrameStatisticsKernel(
  const int32_t* __restrict__ quantized, size_t store,
  const uint8_t* __restrict__ cells, int bw, int bh, int groups_w,
  const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ anchor_list,
  const uint32_t* __restrict__ anchor_count, const int32_t* __restrict__ raw,
  const uint8_t* __restrict__ samples, uint32_t* __restrict__ stats) {
  __shared__ uint32_t raw_counts[kSearchedFamilies * 256];
  __shared__ uint32_t list_offset[kStatMaxChunk];
  __shared__ uint32_t list_sampled[kStatMaxChunk];
  __shared__ int family_start[kSearchedFamilies + 1];
  const int g = blockIdx.x, chunk = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63;
  const int c = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int i = tid; i < kSearchedFamilies * 256; i += kStatThreads) raw_counts[i] = 0;
  const int anchors = static_cast<int>(anchor_count[g]);
  const int per = (anchors + kStatChunks - 1) / kStatChunks;
  const int first = chunk * per, last = min(anchors, first + per);
  if (c == 0) {
    // lane k: transform first + k, bucketed by compact family (ballots in
    // family order give each transform its slot)
    const int k = first + lane;
    const bool live = k < last;
    int family = -1;
    uint32_t offset = 0, sampled = 0;
    if (live) {
      const uint32_t packed = anchor_list[static_cast<size_t>(g) * kGroupAnchors + k];
      const size_t block = static_cast<size_t>(packed >> 16) * bw + (packed & 0xffff);
      family = CompactFamily(OrderFamily(cells[block] >> 1));
      offset = offsets[block];
      sampled = samples[block] != 0;
    }
    int base = 0;
#pragma unroll
    for (int f = 0; f < kSearchedFamilies; ++f) {
      const uint64_t mask = __ballot(family == f);
      if (family == f) {
        const int slot = base + __popcll(mask & ((uint64_t{1} << lane) - 1));
        list_offset[slot] = offset;
        list_sampled[slot] = sampled;
      }
      if (lane == 0) family_start[f] = base;
      base += __popcll(mask);
    }
    if (lane == 0) family_start[kSearchedFamilies] = base;
  }
  __syncthreads();
  // zero counts of channel c over this chunk's transforms
  uint32_t z8 = 0, z16x8[2] = {}, z16[4] = {}, z32x16[8] = {}, z32[16] = {}, sampled = 0;
  uint32_t present = 0;
  const int32_t* plane = quantized + c * store;
  // family f's transforms, U at a time: the loads of a batch issue
  // together, unconditionally (a batch's tail repeats the family's last
  // transform and counts nothing)
  auto walk = [&](int f, auto per_transform, auto unroll) {
    constexpr int U = decltype(unroll)::value;
    const int begin = family_start[f], end = family_start[f + 1];
    if (end > begin) present |= 1u << FullFamily(f);
    for (int k = begin; k < end; k += U) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int at = min(k + u, end - 1);
        per_transform(plane + list_offset[at], at, k + u < end);
      }
    }
  };
  walk(CompactFamily(0), [&](const int32_t* coef, int k, bool live) {
    const uint32_t zero = (coef[lane] == 0) & live;
    z8 += zero;
    sampled += list_sampled[k] ? zero : 0u;
  }, std::integral_constant<int, 8>{});
  walk(CompactFamily(4), [&](const int32_t* coef, int, bool live) {
#pragma unroll
    for (int m = 0; m < 2; ++m) z16x8[m] += (coef[lane + 64 * m] == 0) & live;
  }, std::integral_constant<int, 4>{});
  walk(CompactFamily(2), [&](const int32_t* coef, int, bool live) {
#pragma unroll
    for (int m = 0; m < 4; ++m) z16[m] += (coef[lane + 64 * m] == 0) & live;
  }, std::integral_constant<int, 4>{});
  walk(CompactFamily(6), [&](const int32_t* coef, int, bool live) {
#pragma unroll
    for (int m = 0; m < 8; ++m) z32x16[m] += (coef[lane + 64 * m] == 0) & live;
  }, std::integral_constant<int, 2>{});
  walk(CompactFamily(3), [&](const int32_t* coef, int, bool live) {
#pragma unroll
    for (int m = 0; m < 16; ++m) z32[m] += (coef[lane + 64 * m] == 0) & live;
  }, std::integral_constant<int, 1>{});
  uint32_t* zc = stats + c * kZeroCounters;
  auto flush = [&](int family, const uint32_t* counts, int n) {
    for (int m = 0; m < n; ++m) {
      if (counts[m] != 0) atomicAdd(&zc[ZeroCounterOffset(family) + lane + 64 * m], counts[m]);
    }
  };
  flush(0, &z8, 1);
  flush(4, z16x8, 2);
  flush(2, z16, 4);
  flush(6, z32x16, 8);
  flush(3, z32, 16);
  if (sampled != 0) atomicAdd(&stats[kStatSampled + c * 64 + lane], sampled);
  if (lane == 0 && present != 0) atomicOr(&stats[kStatPresent], present);
  // blocks per order family and raw quant: this chunk's share of the
  // group's blocks
  const int x0 = (g % groups_w) * kGroupBlocks, y0 = (g / groups_w) * kGroupBlocks;
  const int w = min(kGroupBlocks, bw - x0), h = min(kGroupBlocks, bh - y0);
  const int count = w * h;
  const int bper = (count + kStatChunks - 1) / kStatChunks;
  for (int i = chunk * bper + tid; i < min(count, (chunk + 1) * bper); i += kStatThreads) {
    const size_t block = static_cast<size_t>(y0 + i / w) * bw + x0 + i % w;
    const int32_t q = raw[block];
    if (q < 1 || q > 256) {
      atomicOr(&stats[kStatError], kRawQuantError);
    } else {
      const int family = OrderFamily(cells[block] >> 1);
      atomicAdd(&raw_counts[CompactFamily(family) * 256 + q - 1], 1u);
    }
  }
  __syncthreads();
  for (int i = tid; i < kSearchedFamilies * 256; i += kStatThreads) {
    if (raw_counts[i] != 0) {
      atomicAdd(&stats[kStatRaw + FullFamily(i / 256) * 256 + i % 256], raw_counts[i]);
    }
  }
}

// ---- host ------------------------------------------------------------------

// ComputeSimpleNaturalCoefficientOrder (ac_group.cpp:381-487): the zig-zag
// over the square of the longer side's frequencies, alternating direction per
// anti-diagonal, keeping the rows that fall on the shorter axis' frequencies;
// the covered-block positions come first (the LLF prefix).
size_t NaturalOrder(int strategy, uint32_t* order) {
  const int cw = CoveredW(strategy), ch = CoveredH(strategy);
  const size_t wide = static_cast<size_t>(std::max(cw, ch));
  const size_t tall = static_cast<size_t>(std::min(cw, ch));
  const size_t ratio = wide / tall;
  const size_t step_mask = ratio - 1;
  int step_shift = 0;
  while ((size_t{1} << step_shift) < ratio) ++step_shift;
  const size_t side = wide * 8;
  const size_t total = side * tall * 8;
  size_t tail = wide * tall;
  auto place = [&](size_t x, size_t y, bool head) -> bool {
    if (y & step_mask) return true;
    y >>= step_shift;
    size_t index;
    if (head && x < wide && y < tall) {
      index = y * wide + x;
    } else {
      index = tail++;
    }
    if (index >= total) return false;
    order[index] = static_cast<uint32_t>(y * side + x);
    return true;
  };
  for (size_t d = 0; d < side; ++d) {
    for (size_t p = 0; p <= d; ++p) {
      const bool odd = d & 1;
      if (!place(odd ? d - p : p, odd ? p : d - p, true)) return 0;
    }
  }
  for (size_t r = side - 1; r > 0; --r) {
    const size_t d = r - 1;
    for (size_t p = 0; p <= d; ++p) {
      const size_t a = side - 1 - (d - p), b = side - 1 - p;
      const bool odd = d & 1;
      if (!place(odd ? b : a, odd ? a : b, false)) return 0;
    }
  }
  return tail == total ? total : 0;
}

// The effort-7 sample decisions (UseCoefficientOrderSample, src/codestream/
// coefficient_order.cpp:115-126, drawn as :626-650 does): one xorshift128+
// draw per block, AC groups in raster order and blocks in raster order
// within each group; a block's transform is counted when the high half of
// the draw's sum is at most half the 32-bit range.  The reference draws per
// transform of a DCT8-only frame, where every block is one transform, which
// is the only case that reads the decisions.
std::vector<uint8_t> SampleDecisions(const AqLayout& l) {
  std::vector<uint8_t> decisions(static_cast<size_t>(l.bw) * l.bh);
  uint64_t first = 0x94D049BB133111EBull, second = 0xBF58476D1CE4E5B9ull;
  for (int g = 0; g < l.groups; ++g) {
    const int x0 = (g % l.groups_w) * kGroupBlocks, y0 = (g / l.groups_w) * kGroupBlocks;
    const int x1 = std::min(x0 + kGroupBlocks, l.bw), y1 = std::min(y0 + kGroupBlocks, l.bh);
    for (int y = y0; y < y1; ++y) {
      for (int x = x0; x < x1; ++x) {
        const uint64_t sum = first + second;
        uint64_t mixed = first ^ (first << 23);
        mixed ^= second ^ (mixed >> 18) ^ (second >> 5);
        first = second;
        second = mixed;
        decisions[static_cast<size_t>(y) * l.bw + x] =
          static_cast<uint8_t>((sum >> 32) <= (0xffffffffull >> 1));
      }
    }
  }
  return decisions;
}

// The natural order of every stage's strategy, built once per process
// (empty on failure).
const std::vector<uint32_t>& StageNaturalOrders() {
  static const std::vector<uint32_t> table = [] {
    std::vector<uint32_t> t(kOrderStride);
    for (int s = 0; s < kStrategyStages; ++s) {
      const int strategy = StageStrategy(s);
      const size_t n = static_cast<size_t>(CoveredW(strategy)) * CoveredH(strategy) * 64;
      if (NaturalOrder(strategy, t.data() + StageOffset(s)) != n) return std::vector<uint32_t>();
    }
    return t;
  }();
  return table;
}

bool ValidOrder(int strategy, const uint32_t* order) {
  const int cw = CoveredW(strategy), ch = CoveredH(strategy);
  const int n = cw * ch * 64;
  const int llf_w = std::max(cw, ch), llf_h = std::min(cw, ch), side = llf_w * 8;
  std::vector<uint8_t> seen(n, 0);
  for (int i = 0; i < n; ++i) {
    if (order[i] >= static_cast<uint32_t>(n) || seen[order[i]]) return false;
    seen[order[i]] = 1;
  }
  for (int y = 0; y < llf_h; ++y) {
    for (int x = 0; x < llf_w; ++x) {
      if (order[y * llf_w + x] != static_cast<uint32_t>(y * side + x)) return false;
    }
  }
  return true;
}

}  // namespace
}  // namespace gjxl_amd

using gjxl_amd::Check;
using gjxl_amd::Fail;

struct GJXLAmdTokenizer {
  GJXLAmdDevice* device = nullptr;
  uint32_t width = 0, height = 0;
  gjxl_amd::AqLayout layout{};
  gjxl_amd::ContextModel model{};
  void* buffers[GJXL_AMD_TOKENS_BUFFER_COUNT] = {};
  size_t sizes[GJXL_AMD_TOKENS_BUFFER_COUNT] = {};     // at the current extent
  size_t capacity[GJXL_AMD_TOKENS_BUFFER_COUNT] = {};  // as allocated
  // pinned staging of the order table and its inverse; staged_copy marks
  // the end of the last copy out of it
  uint32_t* staged_orders = nullptr;
  uint16_t* staged_inverse = nullptr;
  hipEvent_t staged_copy = nullptr;
  bool staged_pending = false;
};

namespace {

template <typename T>
T* Buf(GJXLAmdTokenizer* t, int id) {
  return static_cast<T*>(t->buffers[id]);
}

GJXLResult ValidateTokenizer(GJXLAmdTokenizer* t) {
  if (t == nullptr) return Fail(GJXL_ERROR_INVALID_ARGUMENT, "tokenizer is null");
  return gjxl_amd::ValidateDevice(t->device);
}

// Uploads the orders of every stage and channel with their inverses through
// the pinned staging buffers, without waiting: the copies are ordered before
// every later launch on the stream, and the next upload waits for them
// before it overwrites the staging.
GJXLResult UploadOrders(GJXLAmdTokenizer* t, const std::vector<uint32_t>& table) {
  using namespace gjxl_amd;
  if (t->staged_pending) {
    GJXL_AMD_TRY(Check(hipEventSynchronize(t->staged_copy), "hipEventSynchronize"));
    t->staged_pending = false;
  }
  std::memcpy(t->staged_orders, table.data(), 4 * table.size());
  for (int s = 0; s < kStrategyStages; ++s) {
    const int n = CoveredW(StageStrategy(s)) * CoveredH(StageStrategy(s)) * 64;
    for (int c = 0; c < 3; ++c) {
      const size_t base = 3 * static_cast<size_t>(StageOffset(s)) + c * n;
      for (int i = 0; i < n; ++i) {
        t->staged_inverse[base + table[base + i]] = static_cast<uint16_t>(i);
      }
    }
  }
  hipStream_t st = t->device->stream;
  GJXL_AMD_TRY(Check(hipMemcpyAsync(t->buffers[GJXL_AMD_TOKENS_ORDERS], t->staged_orders,
                                    4 * table.size(), hipMemcpyHostToDevice, st),
                     "hipMemcpyAsync(orders)"));
  GJXL_AMD_TRY(Check(hipMemcpyAsync(t->buffers[GJXL_AMD_TOKENS_INVERSE_ORDERS],
                                    t->staged_inverse, 2 * table.size(),
                                    hipMemcpyHostToDevice, st),
                     "hipMemcpyAsync(inverse orders)"));
  GJXL_AMD_TRY(Check(hipEventRecord(t->staged_copy, st), "hipEventRecord"));
  t->staged_pending = true;
  return GJXL_OK;
}

// The extent fields of a tokenizer and the byte size of each buffer at that
// extent (every size grows with the width and with the height); false when
// the frame exceeds the 32-bit token offsets.
bool TokenizerGeometry(GJXLAmdTokenizer* t, uint32_t width, uint32_t height, size_t* z) {
  using namespace gjxl_amd;
  const AqLayout layout = MakeAqLayout(PadTo8(width), PadTo8(height));
  const size_t blocks = static_cast<size_t>(layout.bw) * layout.bh;
  const size_t groups = layout.groups;
  // every block holds 64 coefficients per channel and every transform adds
  // one count token: 3 * (64 * blocks + anchors) with anchors <= blocks
  const size_t capacity = 3 * 65 * blocks;
  if (capacity > 0xffffffffull) return false;  // token offsets and counts are 32-bit
  t->width = width;
  t->height = height;
  t->layout = layout;
  z[GJXL_AMD_TOKENS_VALUES] = 4 * capacity;
  z[GJXL_AMD_TOKENS_CONTEXTS] = 2 * capacity;
  z[GJXL_AMD_TOKENS_GROUP_OFFSETS] = 4 * (groups + 1);
  z[GJXL_AMD_TOKENS_HISTOGRAM] =
    4 * static_cast<size_t>(kMaxBlockContexts) * (kNonzeroBuckets + kZeroDensityContexts) * kSymbols;
  z[GJXL_AMD_TOKENS_NONZERO_MAP] = 3 * blocks;
  z[GJXL_AMD_TOKENS_TASKS] = 4 * groups * kGroupTasks;
  z[GJXL_AMD_TOKENS_TASK_OFFSETS] = 4 * groups * kGroupTasks;
  z[GJXL_AMD_TOKENS_GROUP_TOKENS] = 4 * groups;
  z[GJXL_AMD_TOKENS_OFFSETS] = 4 * blocks;
  z[GJXL_AMD_TOKENS_ANCHOR_LIST] = 4 * groups * kGroupAnchors;
  z[GJXL_AMD_TOKENS_ANCHOR_COUNT] = 4 * groups;
  z[GJXL_AMD_TOKENS_ORDERS] = 4 * 3 * kOrderStride;
  z[GJXL_AMD_TOKENS_INVERSE_ORDERS] = 2 * 3 * kOrderStride;
  z[GJXL_AMD_TOKENS_STATUS] = 4 * 4;
  z[GJXL_AMD_TOKENS_STATISTICS] = 4 * kStatWords;
  z[GJXL_AMD_TOKENS_SAMPLE_DECISIONS] = blocks;
  return true;
}

GJXLResult UploadSampleDecisions(GJXLAmdTokenizer* t) {
  const std::vector<uint8_t> decisions = gjxl_amd::SampleDecisions(t->layout);
  GJXL_AMD_TRY(Check(hipMemcpyAsync(t->buffers[GJXL_AMD_TOKENS_SAMPLE_DECISIONS],
                                    decisions.data(), decisions.size(),
                                    hipMemcpyHostToDevice, t->device->stream),
                     "hipMemcpyAsync(sample decisions)"));
  // the pageable source is gone when this returns
  return Check(hipStreamSynchronize(t->device->stream), "hipStreamSynchronize");
}

}  // namespace

namespace gjxl_amd {

GJXLResult TokenizeEnqueue(GJXLAmdTokenizer* t, const int32_t* quantized,
                           const uint8_t* cells, const int32_t* raw_quant,
                           bool populations) {
  GJXL_AMD_TRY(ValidateTokenizer(t));
  if (quantized == nullptr || cells == nullptr || raw_quant == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "tokenization arguments are null");
  }
  GJXL_AMD_TRY(Check(hipSetDevice(t->device->ordinal), "hipSetDevice"));
  hipStream_t s = t->device->stream;
  const AqLayout& l = t->layout;
  uint32_t* status = Buf<uint32_t>(t, GJXL_AMD_TOKENS_STATUS);
  uint32_t* offsets = Buf<uint32_t>(t, GJXL_AMD_TOKENS_OFFSETS);
  uint32_t* anchors = Buf<uint32_t>(t, GJXL_AMD_TOKENS_ANCHOR_LIST);
  uint32_t* counts = Buf<uint32_t>(t, GJXL_AMD_TOKENS_ANCHOR_COUNT);
  uint32_t* tasks = Buf<uint32_t>(t, GJXL_AMD_TOKENS_TASKS);
  uint32_t* task_offsets = Buf<uint32_t>(t, GJXL_AMD_TOKENS_TASK_OFFSETS);
  uint32_t* group_tokens = Buf<uint32_t>(t, GJXL_AMD_TOKENS_GROUP_TOKENS);
  uint32_t* group_offsets = Buf<uint32_t>(t, GJXL_AMD_TOKENS_GROUP_OFFSETS);
  uint8_t* map = Buf<uint8_t>(t, GJXL_AMD_TOKENS_NONZERO_MAP);
  uint32_t* histogram = populations ? Buf<uint32_t>(t, GJXL_AMD_TOKENS_HISTOGRAM) : nullptr;
  hipLaunchKernelGGL(FillKernel, dim3(1), dim3(64), 0, s, status, size_t{4}, 0u);
  GJXL_AMD_TRY(CheckLaunch("FillKernel"));
  // status[1] bit 0: the grid does not tile its groups
  GJXL_AMD_TRY(LaunchAnchorLayout(s, cells, l.bw * 8, l.bh * 8, offsets, anchors, counts,
                                  status + 1, nullptr, 0, nullptr));
  const dim3 task_grid(kGroupTasks / kTaskWaves, l.groups);
  hipLaunchKernelGGL(TokenCountKernel, task_grid, dim3(64 * kTaskWaves), 0, s, quantized,
                     l.store, cells, l.bw, l.bh, offsets, anchors, counts,
                     Buf<uint16_t>(t, GJXL_AMD_TOKENS_INVERSE_ORDERS), raw_quant, status,
                     tasks, map);
  GJXL_AMD_TRY(CheckLaunch("TokenCountKernel"));
  hipLaunchKernelGGL(TokenGroupScanKernel, dim3(l.groups), dim3(kGroupAnchors), 0, s, counts,
                     tasks, task_offsets, group_tokens);
  GJXL_AMD_TRY(CheckLaunch("TokenGroupScanKernel"));
  hipLaunchKernelGGL(TokenFrameScanKernel, dim3(1), dim3(1024), 0, s, group_tokens, l.groups,
                     group_offsets, status);
  GJXL_AMD_TRY(CheckLaunch("TokenFrameScanKernel"));
  if (histogram != nullptr) {
    const size_t words = static_cast<size_t>(t->model.num_contexts) *
                         (kNonzeroBuckets + kZeroDensityContexts) * kSymbols;
    hipLaunchKernelGGL(FillKernel, dim3(256), dim3(256), 0, s, histogram, words, 0u);
    GJXL_AMD_TRY(CheckLaunch("FillKernel"));
  }
  hipLaunchKernelGGL(TokenEmitKernel, task_grid, dim3(64 * kTaskWaves), 0, s, quantized,
                     l.store, cells, l.bw, l.bh, offsets, anchors, counts,
                     Buf<uint32_t>(t, GJXL_AMD_TOKENS_ORDERS), raw_quant, tasks, task_offsets,
                     group_offsets, map, t->model, Buf<uint32_t>(t, GJXL_AMD_TOKENS_VALUES),
                     Buf<uint16_t>(t, GJXL_AMD_TOKENS_CONTEXTS), histogram);
  GJXL_AMD_TRY(CheckLaunch("TokenEmitKernel"));
  return GJXL_OK;
}

GJXLResult TokenizeResult(const uint32_t status[4], uint64_t* token_count) {
  if (status[1] & kGridError) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "strategy grid does not tile its AC groups");
  }
  if (status[1] & kRawQuantError) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "AC-group raw quantization is out of range");
  }
  *token_count = status[0];
  return GJXL_OK;
}

}  // namespace gjxl_amd

GJXLResult gjxl_amd_natural_order(int32_t strategy, uint32_t* order, size_t capacity,
                                  size_t* count) {
  using namespace gjxl_amd;
  const bool searched = strategy == kDct8 || strategy == kDct16x16 || strategy == kDct32x32 ||
                        strategy == kDct16x8 || strategy == kDct8x16 ||
                        strategy == kDct32x16 || strategy == kDct16x32;
  if (!searched) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                "natural coefficient order does not support this strategy");
  }
  const size_t n = static_cast<size_t>(CoveredW(strategy)) * CoveredH(strategy) * 64;
  if (order == nullptr || count == nullptr || capacity < n) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "natural order output is too small");
  }
  if (NaturalOrder(strategy, order) != n) {
    return Fail(GJXL_ERROR_INTERNAL, "natural coefficient order is incomplete");
  }
  *count = n;
  return GJXL_OK;
}

GJXLResult gjxl_amd_tokenizer_create(GJXLAmdDevice* device, uint32_t width, uint32_t height,
                                     GJXLAmdTokenizer** out) {
  using namespace gjxl_amd;
  GJXL_AMD_TRY(ValidateDevice(device));
  if (out == nullptr || *out != nullptr || width == 0 || height == 0 ||
      width > (1u << 16) || height > (1u << 16)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "tokenizer geometry is invalid");
  }
  auto* t = new (std::nothrow) GJXLAmdTokenizer;
  if (t == nullptr) return Fail(GJXL_ERROR_OUT_OF_MEMORY, "tokenizer allocation failed");
  t->device = device;
  if (!TokenizerGeometry(t, width, height, t->sizes)) {
    delete t;
    return Fail(GJXL_ERROR_UNSUPPORTED,
                "frame too large for 32-bit token offsets (at most 22,025,473 blocks)");
  }
  size_t* z = t->sizes;
  for (int i = 0; i < GJXL_AMD_TOKENS_BUFFER_COUNT; ++i) {
    t->capacity[i] = z[i];
    const GJXLResult r = gjxl_amd::DeviceAlloc(device, z[i], &t->buffers[i]);
    if (r != GJXL_OK) {
      gjxl_amd_tokenizer_destroy(t);
      return r;
    }
  }

  GJXLResult r = Check(hipHostMalloc(reinterpret_cast<void**>(&t->staged_orders),
                                     4 * 3 * kOrderStride, hipHostMallocDefault),
                       "hipHostMalloc(order staging)");
  if (r == GJXL_OK) {
    r = Check(hipHostMalloc(reinterpret_cast<void**>(&t->staged_inverse),
                            2 * 3 * kOrderStride, hipHostMallocDefault),
              "hipHostMalloc(order staging)");
  }
  if (r == GJXL_OK) r = Check(hipEventCreate(&t->staged_copy), "hipEventCreate");
  if (r == GJXL_OK) r = gjxl_amd_tokenizer_set_orders(t, 0, nullptr);
  if (r == GJXL_OK) r = UploadSampleDecisions(t);
  if (r == GJXL_OK) {
    // DefaultSimpleBlockContextMap: src/codestream/simple_ac_context.h:29-33
    uint8_t map[3 * kOrderFamilies];
    for (int i = 0; i < 3 * kOrderFamilies; ++i) {
      const int row = i / kOrderFamilies, family = i % kOrderFamilies;
      map[i] = static_cast<uint8_t>((row == 0 ? 0 : 2) + (family < 4 ? 0 : 1));
    }
    r = gjxl_amd_tokenizer_set_context_map(t, nullptr, 0, map, 4);
  }
  if (r != GJXL_OK) {
    gjxl_amd_tokenizer_destroy(t);
    return r;
  }
  *out = t;
  return GJXL_OK;
}

void gjxl_amd_tokenizer_destroy(GJXLAmdTokenizer* t) {
  if (t == nullptr) return;
  if (t->device != nullptr) {
    (void)hipSetDevice(t->device->ordinal);
    (void)hipStreamSynchronize(t->device->stream);
  }
  for (void*& b : t->buffers) {
    if (t->device != nullptr) gjxl_amd::DeviceFree(t->device, b);
  }

  if (t->staged_copy != nullptr) (void)hipEventDestroy(t->staged_copy);
  if (t->staged_orders != nullptr) (void)hipHostFree(t->staged_orders);
  if (t->staged_inverse != nullptr) (void)hipHostFree(t->staged_inverse);
  delete t;
}

GJXLResult gjxl_amd_tokenizer_set_orders(GJXLAmdTokenizer* t, uint16_t used_order_mask,
                                         const uint32_t* const* orders) {
  using namespace gjxl_amd;
  GJXL_AMD_TRY(ValidateTokenizer(t));
  if (used_order_mask != 0 && orders == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "coefficient orders are null");
  }
  if (used_order_mask >> kOrderFamilies) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "coefficient-order mask names an absent family");
  }
  const std::vector<uint32_t>& naturals = StageNaturalOrders();
  if (naturals.empty()) return Fail(GJXL_ERROR_INTERNAL, "natural coefficient order is incomplete");
  std::vector<uint32_t> table(3 * kOrderStride);
  for (int s = 0; s < kStrategyStages; ++s) {
    const int strategy = StageStrategy(s);
    const int family = OrderFamily(strategy);
    const size_t n = static_cast<size_t>(CoveredW(strategy)) * CoveredH(strategy) * 64;
    for (int c = 0; c < 3; ++c) {
      uint32_t* dst = table.data() + 3 * static_cast<size_t>(StageOffset(s)) + c * n;
      if (used_order_mask & (1u << family)) {
        const uint32_t* src = orders[3 * family + c];
        if (src == nullptr || !ValidOrder(strategy, src)) {
          return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                      "coefficient order is not a permutation with the LLF prefix");
        }
        std::memcpy(dst, src, 4 * n);
      } else {
        std::memcpy(dst, naturals.data() + StageOffset(s), 4 * n);
      }
    }
  }
  return UploadOrders(t, table);
}

GJXLResult gjxl_amd_tokenizer_set_context_map(GJXLAmdTokenizer* t, const uint32_t* thresholds,
                                              uint32_t threshold_count, const uint8_t* map,
                                              uint32_t num_contexts) {
  using namespace gjxl_amd;
  GJXL_AMD_TRY(ValidateTokenizer(t));
  if (threshold_count > kMaxThresholds || num_contexts == 0 ||
      num_contexts > kMaxBlockContexts || map == nullptr ||
      (threshold_count != 0 && thresholds == nullptr)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "block-context map dimensions are invalid");
  }
  for (uint32_t i = 0; i < threshold_count; ++i) {
    if (thresholds[i] == 0 || thresholds[i] > 255 || (i && thresholds[i - 1] >= thresholds[i])) {
      return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                  "block-context quantization thresholds are invalid");
    }
  }
  const uint32_t entries = 3 * kOrderFamilies * (threshold_count + 1);
  bool seen[kMaxBlockContexts] = {};
  uint32_t labels = 0;
  for (uint32_t i = 0; i < entries; ++i) {
    if (map[i] >= num_contexts) {
      return Fail(GJXL_ERROR_INVALID_ARGUMENT, "block-context map references an absent context");
    }
    seen[map[i]] = true;
    labels = std::max<uint32_t>(labels, map[i] + 1u);
  }
  if (labels != num_contexts) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "block-context map labels are invalid");
  }
  for (uint32_t c = 0; c < num_contexts; ++c) {
    if (!seen[c]) {
      return Fail(GJXL_ERROR_INVALID_ARGUMENT, "block-context map labels are not canonical");
    }
  }
  ContextModel m{};
  std::memcpy(m.map, map, entries);
  for (uint32_t i = 0; i < threshold_count; ++i) m.thresholds[i] = thresholds[i];
  m.threshold_count = threshold_count;
  m.num_contexts = num_contexts;
  t->model = m;
  return GJXL_OK;
}

GJXLResult gjxl_amd_tokenize(GJXLAmdTokenizer* t, const int32_t* quantized,
                             const uint8_t* cells, const int32_t* raw_quant,
                             int32_t populations, uint64_t* token_count) {
  using namespace gjxl_amd;
  if (token_count == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "tokenization arguments are null");
  }
  GJXL_AMD_TRY(TokenizeEnqueue(t, quantized, cells, raw_quant, populations != 0));
  hipStream_t s = t->device->stream;
  uint32_t host[4];
  GJXL_AMD_TRY(Check(hipMemcpyAsync(host, Buf<uint32_t>(t, GJXL_AMD_TOKENS_STATUS), sizeof(host),
                                    hipMemcpyDeviceToHost, s),
                     "hipMemcpyAsync(token status)"));
  GJXL_AMD_TRY(Check(hipStreamSynchronize(s), "hipStreamSynchronize"));
  return TokenizeResult(host, token_count);
}

GJXLResult gjxl_amd_tokenizer_buffer(GJXLAmdTokenizer* t, int32_t id, void** pointer,
                                     size_t* bytes) {
  GJXL_AMD_TRY(ValidateTokenizer(t));
  if (id < 0 || id >= GJXL_AMD_TOKENS_BUFFER_COUNT || pointer == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "tokenizer buffer id is invalid");
  }
  *pointer = t->buffers[id];
  if (bytes != nullptr) *bytes = t->sizes[id];
  return GJXL_OK;
}

namespace gjxl_amd {

GJXLResult FrameStatisticsEnqueue(GJXLAmdTokenizer* t, const int32_t* quantized,
                                  const uint8_t* cells, const int32_t* raw_quant) {
  GJXL_AMD_TRY(ValidateTokenizer(t));
  if (quantized == nullptr || cells == nullptr || raw_quant == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "frame statistics arguments are null");
  }
  hipStream_t s = t->device->stream;
  const AqLayout& l = t->layout;
  uint32_t* stats = Buf<uint32_t>(t, GJXL_AMD_TOKENS_STATISTICS);
  uint32_t* offsets = Buf<uint32_t>(t, GJXL_AMD_TOKENS_OFFSETS);
  uint32_t* anchors = Buf<uint32_t>(t, GJXL_AMD_TOKENS_ANCHOR_LIST);
  uint32_t* counts = Buf<uint32_t>(t, GJXL_AMD_TOKENS_ANCHOR_COUNT);
  hipLaunchKernelGGL(FillKernel, dim3(CeilDiv(kStatWords, 256)), dim3(256), 0, s, stats,
                     size_t{kStatWords}, 0u);
  GJXL_AMD_TRY(CheckLaunch("FillKernel"));
  // error word bit 0: the grid does not tile its groups (no transform of
  // such a group is counted)
  GJXL_AMD_TRY(LaunchAnchorLayout(s, cells, l.bw * 8, l.bh * 8, offsets, anchors, counts,
                                  stats + kStatError, nullptr, 0, nullptr));
  hipLaunchKernelGGL(FrameStatisticsKernel, dim3(l.groups, kStatChunks), dim3(kStatThreads), 0, s,
                     quantized, l.store, cells, l.bw, l.bh, l.groups_w, offsets, anchors,
                     counts, raw_quant, Buf<uint8_t>(t, GJXL_AMD_TOKENS_SAMPLE_DECISIONS),
                     stats);
  return CheckLaunch("FrameStatisticsKernel");
}

GJXLResult FrameStatisticsResult(GJXLAmdTokenizer* t, const uint32_t* words,
                                 GJXLAmdFrameStatistics* out) {
  if (words[kStatError] & kGridError) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "strategy grid does not tile its AC groups");
  }
  if (words[kStatError] & kRawQuantError) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Block-context raw quantization is out of range");
  }
  out->block_width = static_cast<uint32_t>(t->layout.bw);
  out->block_height = static_cast<uint32_t>(t->layout.bh);
  out->present_mask = static_cast<uint16_t>(words[kStatPresent]);
  std::memcpy(out->zero_counts, words, sizeof(out->zero_counts));
  std::memcpy(out->sampled_dct8_zero_counts, words + kStatSampled,
              sizeof(out->sampled_dct8_zero_counts));
  std::memcpy(out->raw_quant_counts, words + kStatRaw, sizeof(out->raw_quant_counts));
  return GJXL_OK;
}

size_t FrameStatisticsWords() { return kStatWords; }

GJXLResult ResizeTokenizer(GJXLAmdTokenizer* t, uint32_t width, uint32_t height) {
  if (width == 0 || height == 0 || width > (1u << 16) || height > (1u << 16)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "tokenizer geometry is invalid");
  }
  GJXLAmdTokenizer probe;
  size_t need[GJXL_AMD_TOKENS_BUFFER_COUNT] = {};
  bool fits = TokenizerGeometry(&probe, width, height, need);
  for (int i = 0; fits && i < GJXL_AMD_TOKENS_BUFFER_COUNT; ++i) fits = need[i] <= t->capacity[i];

  if (!fits) return Fail(GJXL_ERROR_INVALID_ARGUMENT, "extent exceeds the tokenizer's capacity");
  (void)TokenizerGeometry(t, width, height, t->sizes);
  return UploadSampleDecisions(t);
}

}  // namespace gjxl_amd

GJXLResult gjxl_amd_tokenizer_frame_statistics(GJXLAmdTokenizer* t, const int32_t* quantized,
                                               const uint8_t* cells, const int32_t* raw_quant,
                                               GJXLAmdFrameStatistics* statistics) {
  using namespace gjxl_amd;
  if (statistics == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "frame statistics output is null");
  }
  GJXL_AMD_TRY(FrameStatisticsEnqueue(t, quantized, cells, raw_quant));
  std::vector<uint32_t> words(kStatWords);
  hipStream_t s = t->device->stream;
  GJXL_AMD_TRY(Check(hipMemcpyAsync(words.data(), t->buffers[GJXL_AMD_TOKENS_STATISTICS],
                                    4 * words.size(), hipMemcpyDeviceToHost, s),
                     "hipMemcpyAsync(frame statistics)"));
  GJXL_AMD_TRY(Check(hipStreamSynchronize(s), "hipStreamSynchronize"));
  GJXLAmdFrameStatistics result;
  GJXL_AMD_TRY(FrameStatisticsResult(t, words.data(), &result));
  *statistics = result;
  return GJXL_OK;
}
