// Host side of the CPU-chosen tokenization inputs: coefficient orders and the
// adaptive block-context map from the frame statistics that
// gjxl_amd_tokenizer_frame_statistics gathers on the device, and the tokens
// that code the orders.  No kernels here.
//
// Orders: ComputeCoefficientOrdersWithCounts (src/codestream/
// coefficient_order.cpp:736-811).  For each present family and channel, the
// positions after the LLF prefix are ranked by uint64(zeros * n^-1/2 + 0.1f)
// (binary32 product and sum), ascending, with the natural scan order kept
// on equal ranks; the family's order is used when any channel differs from
// the natural one.  At effort 7 a DCT8-only frame ranks by the sampled
// counts.  Frames under 5 blocks in both axes keep the natural orders.
//
// Map: ComputeSimpleBlockContextMapForEncoder (src/codestream/
// block_context_map.cpp:435-476).  Below 1024 blocks the default map; from
// 8192 blocks one raw-quant threshold at the median block; cells (family x
// segment) merged smallest-into-next-smallest down to clamp(blocks / 2048,
// 2, 9) luma contexts (BuildAdaptiveMap, :126-200), labels in order of first
// use, chroma rows next + min(luma label, clamp(blocks / 3072, 1, 5) - 1).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "tables.h"

using gjxl_amd::Fail;

namespace {

constexpr int kFamilies = GJXL_AMD_ORDER_FAMILIES;
constexpr int kCells = kFamilies * (GJXL_AMD_MAX_QF_THRESHOLDS + 1);

// The searched families, their representative strategy (RepresentativeInfo,
// coefficient_order.cpp:128-167) and their counters' offset in a channel.
struct SearchedFamily {
  int family, strategy, offset, size, llf;
};
constexpr SearchedFamily kSearched[5] = {
  {0, gjxl_amd::kDct8, 0, 64, 1},
  {2, gjxl_amd::kDct16x16, 192, 256, 4},
  {3, gjxl_amd::kDct32x32, 960, 1024, 16},
  {4, gjxl_amd::kDct16x8, 64, 128, 2},
  {6, gjxl_amd::kDct32x16, 448, 512, 8},
};
constexpr uint16_t kSearchedMask = (1u << 0) | (1u << 2) | (1u << 3) | (1u << 4) | (1u << 6);

// src/codestream/simple_ac_context.h:29-33
constexpr uint8_t kDefaultMap[3 * kFamilies] = {
  0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1,
  2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3, 3,
  2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3, 3,
};

// The natural orders of the searched families, built once per process.
struct NaturalOrders {
  uint32_t order[5][1024];
  GJXLResult status = GJXL_OK;
  NaturalOrders() {
    for (int i = 0; i < 5 && status == GJXL_OK; ++i) {
      size_t n = 0;
      status = gjxl_amd_natural_order(kSearched[i].strategy, order[i], 1024, &n);
      if (status == GJXL_OK && n != static_cast<size_t>(kSearched[i].size)) {
        status = Fail(GJXL_ERROR_INTERNAL, "natural order size disagrees with its family");
      }
    }
  }
};

const NaturalOrders& Naturals() {
  static const NaturalOrders orders;
  return orders;
}

// Sorts items[0, n) by key[item], stably (least significant 11-bit digit
// first), through tmp.  Keys are below 2^33.
void StableSortByKey(uint32_t* items, size_t n, const uint64_t* key, uint32_t* tmp) {
  uint64_t top = 0;
  for (size_t i = 0; i < n; ++i) top = std::max(top, key[items[i]]);
  uint32_t* from = items;
  uint32_t* to = tmp;
  for (int shift = 0; shift < 64 && (shift == 0 || (top >> shift) != 0); shift += 11) {
    uint32_t start[2048] = {};
    for (size_t i = 0; i < n; ++i) ++start[(key[from[i]] >> shift) & 2047];
    uint32_t sum = 0;
    for (uint32_t& b : start) {
      const uint32_t c = b;
      b = sum;
      sum += c;
    }
    for (size_t i = 0; i < n; ++i) to[start[(key[from[i]] >> shift) & 2047]++] = from[i];
    std::swap(from, to);
  }
  if (from != items) std::copy(from, from + n, items);
}

GJXLResult ComputeOrders(const GJXLAmdFrameStatistics& s, bool sampled_behaviour,
                         GJXLAmdEntropyInputs* out) {
  out->used_order_mask = 0;
  if (s.block_width < 5 && s.block_height < 5) return GJXL_OK;
  const NaturalOrders& naturals = Naturals();
  GJXL_AMD_TRY(naturals.status);
  const bool sample = sampled_behaviour && s.present_mask == 1;
  uint64_t key[1024];
  uint32_t tmp[1024];
  for (int fi = 0; fi < 5; ++fi) {
    const SearchedFamily& f = kSearched[fi];
    if (!(s.present_mask & (1u << f.family))) continue;
    const size_t n = static_cast<size_t>(f.size);
    const uint32_t* natural = naturals.order[fi];
    const float inverse_sqrt_size = 1.0f / std::sqrt(static_cast<float>(n));
    bool custom = false;
    for (int c = 0; c < 3; ++c) {
      const uint32_t* zeros = sample ? s.sampled_dct8_zero_counts[c]
                                     : s.zero_counts[c] + f.offset;
      for (size_t k = 0; k < n; ++k) {
        const float scaled = static_cast<float>(zeros[k]) * inverse_sqrt_size;
        key[k] = static_cast<uint64_t>(scaled + 0.1f);
      }
      // ascending keys, equal keys in natural order: a stable sort of the
      // natural sequence after the LLF prefix
      uint32_t* order = out->orders[f.family][c];
      std::copy(natural, natural + n, order);
      StableSortByKey(order + f.llf, n - f.llf, key, tmp);
      custom |= !std::equal(order, order + n, natural);
    }
    if (custom) out->used_order_mask |= static_cast<uint16_t>(1u << f.family);
  }
  return GJXL_OK;
}

void ComputeMap(const GJXLAmdFrameStatistics& s, GJXLAmdEntropyInputs* out) {
  const uint64_t blocks = static_cast<uint64_t>(s.block_width) * s.block_height;
  out->qf_threshold_count = 0;
  std::memset(out->qf_thresholds, 0, sizeof(out->qf_thresholds));
  std::memset(out->context_map, 0, sizeof(out->context_map));
  if (blocks < 1024) {
    std::memcpy(out->context_map, kDefaultMap, sizeof(kDefaultMap));
    out->num_contexts = 4;
    return;
  }
  // MedianQuantThreshold (:105-124)
  if (blocks >= 8192) {
    uint64_t cumulative = 0;
    for (int q = 0; q < 256; ++q) {
      for (int f = 0; f < kFamilies; ++f) cumulative += s.raw_quant_counts[f][q];
      if (cumulative > blocks / 2) {
        if (q != 0) {
          out->qf_thresholds[0] = static_cast<uint32_t>(q);
          out->qf_threshold_count = 1;
        }
        break;
      }
    }
  }
  const int segments = static_cast<int>(out->qf_threshold_count) + 1;
  const int cells = kFamilies * segments;
  uint64_t weight[kCells] = {};
  for (int f = 0; f < kFamilies; ++f) {
    for (int q = 0; q < 256; ++q) {
      int segment = 0;
      while (segment < segments - 1 &&
             static_cast<uint32_t>(q + 1) > out->qf_thresholds[segment]) {
        ++segment;
      }
      weight[f * segments + segment] += s.raw_quant_counts[f][q];
    }
  }
  // Clusters ordered by weight, heaviest first, equal weights keeping the
  // previous round's order; the lightest joins the next lightest.
  uint8_t parent[kCells];
  std::vector<uint8_t> live(cells);
  for (int i = 0; i < cells; ++i) parent[i] = live[i] = static_cast<uint8_t>(i);
  const size_t luma = static_cast<size_t>(std::clamp<uint64_t>(blocks / 2048, 2, 9));
  while (live.size() > luma) {
    std::stable_sort(live.begin(), live.end(),
                     [&](uint8_t a, uint8_t b) { return weight[a] > weight[b]; });
    const uint8_t into = live[live.size() - 2], from = live.back();
    weight[into] += weight[from];
    weight[from] = 0;
    parent[from] = into;
    live.pop_back();
  }
  uint8_t label_of_root[kCells];
  std::memset(label_of_root, 0xff, sizeof(label_of_root));
  uint8_t labels = 0;
  uint8_t luma_label[kCells];
  for (int i = 0; i < cells; ++i) {
    int root = i;
    while (parent[root] != root) root = parent[root];
    if (label_of_root[root] == 0xff) label_of_root[root] = labels++;
    luma_label[i] = label_of_root[root];
  }
  const int chroma = static_cast<int>(std::clamp<uint64_t>(blocks / 3072, 1, 5));
  uint8_t top = 0;
  for (int i = 0; i < 3 * cells; ++i) {
    const uint8_t l = luma_label[i % cells];
    out->context_map[i] = i < cells ? l
                          : static_cast<uint8_t>(labels + std::min<int>(l, chroma - 1));
    top = std::max(top, out->context_map[i]);
  }
  out->num_contexts = top + 1u;
}

GJXLResult ValidateStatistics(const GJXLAmdFrameStatistics& s) {
  if (s.block_width == 0 || s.block_height == 0) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "frame statistics have no blocks");
  }
  if (s.present_mask & ~kSearchedMask) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                "Coefficient-order population names a family outside the search");
  }
  const uint64_t blocks = static_cast<uint64_t>(s.block_width) * s.block_height;
  uint64_t counted = 0;
  for (int f = 0; f < kFamilies; ++f) {
    for (int q = 0; q < 256; ++q) counted += s.raw_quant_counts[f][q];
  }
  if (counted != blocks) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Block-context counts do not cover the frame");
  }
  for (const SearchedFamily& f : kSearched) {
    const bool present = s.present_mask & (1u << f.family);
    const uint64_t bound = present ? blocks / (f.size / 64) : 0;
    for (int c = 0; c < 3; ++c) {
      for (int k = 0; k < f.size; ++k) {
        if (s.zero_counts[c][f.offset + k] > bound) {
          return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                      "Coefficient-order population exceeds its bounds");
        }
      }
    }
  }
  return GJXL_OK;
}

// EncodeHybridUintValidated with the all-zero configuration
// (src/codestream/entropy_internal.h:31-55), capped at the 8 permutation
// contexts (CoefficientOrderContext, coefficient_order.cpp:403-415).
uint32_t PermutationContext(uint32_t value) {
  if (value == 0) return 0;
  const uint32_t symbol = 1u + (31u - static_cast<uint32_t>(__builtin_clz(value)));
  return std::min<uint32_t>(symbol, 7u);
}

}  // namespace

extern "C" {

GJXLResult gjxl_amd_entropy_inputs_from_statistics(const GJXLAmdFrameStatistics* statistics,
                                                   int32_t effort,
                                                   GJXLAmdEntropyInputs* out) {
  if (statistics == nullptr || out == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "entropy input arguments are null");
  }
  if (effort < 1 || effort > 10) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "effort must be in [1, 10]");
  }
  GJXL_AMD_TRY(ValidateStatistics(*statistics));
  // nothing below fails once the statistics and the natural orders are
  // good, so out is written in place
  GJXL_AMD_TRY(Naturals().status);
  std::memset(out, 0, sizeof(*out));
  GJXL_AMD_TRY(ComputeOrders(*statistics, effort == 7, out));
  ComputeMap(*statistics, out);
  return GJXL_OK;
}

GJXLResult gjxl_amd_tokenizer_set_entropy_inputs(GJXLAmdTokenizer* tokenizer,
                                                 const GJXLAmdEntropyInputs* inputs) {
  if (inputs == nullptr) return Fail(GJXL_ERROR_INVALID_ARGUMENT, "entropy inputs are null");
  const uint32_t* orders[3 * kFamilies] = {};
  for (int f = 0; f < kFamilies; ++f) {
    for (int c = 0; c < 3; ++c) orders[3 * f + c] = inputs->orders[f][c];
  }
  GJXL_AMD_TRY(gjxl_amd_tokenizer_set_orders(tokenizer, inputs->used_order_mask, orders));
  return gjxl_amd_tokenizer_set_context_map(tokenizer, inputs->qf_thresholds,
                                            inputs->qf_threshold_count, inputs->context_map,
                                            inputs->num_contexts);
}

GJXLResult gjxl_amd_coefficient_order_tokens(const GJXLAmdEntropyInputs* inputs,
                                             uint32_t* contexts, uint32_t* values,
                                             size_t capacity, size_t* count) {
  if (inputs == nullptr || count == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "order token arguments are null");
  }
  if (inputs->used_order_mask & ~kSearchedMask) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Coefficient-order mask uses an unsupported family");
  }
  std::vector<uint32_t> ctx, val;
  std::vector<uint32_t> natural(1024), position(1024), lehmer(1024), tree(1025);
  for (int family = 0; family < kFamilies; ++family) {
    if (!(inputs->used_order_mask & (1u << family))) continue;
    const SearchedFamily* f = nullptr;
    for (const SearchedFamily& s : kSearched) {
      if (s.family == family) f = &s;
    }
    size_t n = 0;
    GJXL_AMD_TRY(gjxl_amd_natural_order(f->strategy, natural.data(), natural.size(), &n));
    for (size_t i = 0; i < n; ++i) position[natural[i]] = static_cast<uint32_t>(i);
    for (int c = 0; c < 3; ++c) {
      const uint32_t* order = inputs->orders[family][c];
      // Lehmer code of the order's natural positions: each position less the
      // earlier positions below it (a Fenwick tree over positions)
      std::fill(tree.begin(), tree.begin() + n + 1, 0u);
      for (size_t i = 0; i < n; ++i) {
        if (order[i] >= n) {
          return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Coefficient permutation is invalid");
        }
        const uint32_t p = position[order[i]];
        uint32_t below = 0;
        for (size_t j = p + 1; j > 0; j &= j - 1) below += tree[j];
        if (below > p) {
          return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Coefficient permutation repeats a value");
        }
        lehmer[i] = p - below;
        for (size_t j = p + 1; j <= n; j += j & (~j + 1)) ++tree[j];
      }
      const size_t skip = static_cast<size_t>(f->llf);
      size_t end = n;
      while (end > skip && lehmer[end - 1] == 0) --end;
      ctx.push_back(PermutationContext(static_cast<uint32_t>(n)));
      val.push_back(static_cast<uint32_t>(end - skip));
      uint32_t last = 0;
      for (size_t i = skip; i < end; ++i) {
        ctx.push_back(PermutationContext(last));
        val.push_back(lehmer[i]);
        last = lehmer[i];
      }
    }
  }
  if (ctx.size() > capacity || (!ctx.empty() && (contexts == nullptr || values == nullptr))) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "order token output is too small");
  }
  std::copy(ctx.begin(), ctx.end(), contexts);
  std::copy(val.begin(), val.end(), values);
  *count = ctx.size();
  return GJXL_OK;
}

}  // extern "C"
