/*
 * f1/f2: the CPU-chosen inputs of AC tokenization at effort 7 -- coefficient
 * orders and the adaptive block-context map -- and the frame statistics they
 * are computed from.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates src/codestream/coefficient_order.cpp:
 *   UseCoefficientOrderSample (:115-126), the effort-7 DCT8-only sample
 *   decisions (:626-650), CountGroupZeros (:223-316), the order selection of
 *   ComputeCoefficientOrdersWithCounts (:736-811), ComputeLehmerCode
 *   (:371-401), CoefficientOrderContext (:403-415), TokenizePermutation
 *   (:417-465) and TokenizeSimpleCoefficientOrders (:876-919);
 * and src/codestream/block_context_map.cpp:
 *   CountRawQuantAndOrders (:67-103), MedianQuantThreshold (:105-124),
 *   BuildAdaptiveMap (:126-200) and ComputeSimpleBlockContextMapForEncoder
 *   (:435-476).
 *
 * Pinned: the scalar restatement the reference's own test carries for the
 * orders (tests/coefficient_order_test.cpp:32-118: coefficient-major zero
 * counts, keys uint64(zeros * n^-1/2 + 0.1f), stable sort after the LLF
 * prefix, the xorshift128+ sample with seeds 0x94D049BB133111EB /
 * 0xBF58476D1CE4E5B9), the small-frame cutoff and constant-population
 * cases of that test (:160-186); the reference's default map
 * (tests/block_context_map_test.cpp:15-39); and, for the adaptive map,
 * invariants checked in tests/test_oracle_entropy_inputs.py (canonical
 * labels, cluster count, chroma labels) with brute-force recomputation.
 *
 * Layouts:
 *   statistics zero counts: per channel (X, Y, B) 1984 counters, the order
 *   families DCT8 (64), 16x8 (128), 16x16 (256), 32x16 (512), 32x32 (1024)
 *   at offsets 0, 64, 192, 448, 960; sampled DCT8 counts 3 x 64; raw-quant
 *   counts 13 x 256 (family, raw quant - 1).
 *   orders: the tokenizer's 7-stage table (see tokenize.c), 3 x 2624.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* src/codestream/simple_ac_context.h:20-23 (as tokenize.c) */
static const uint8_t kFamilyOfStrategy[27] = {
  0, 1, 1, 1, 2, 3, 4, 4, 5, 5, 6, 6, 1, 1,
  1, 1, 1, 1, 7, 8, 8, 9, 10, 10, 11, 12, 12,
};
/* offset of a family's counters in one channel's 1984; -1: not searched */
static int FamilyOffset(int family) {
  switch (family) {
    case 0: return 0;
    case 4: return 64;
    case 2: return 192;
    case 6: return 448;
    case 3: return 960;
    default: return -1;
  }
}
/* representative strategy of a family (coefficient_order.cpp:128-167) */
static int FamilyStrategy(int family) {
  switch (family) {
    case 0: return 0;   /* DCT8 */
    case 4: return 6;   /* DCT16x8 */
    case 2: return 4;   /* DCT16x16 */
    case 6: return 10;  /* DCT32x16 */
    case 3: return 5;   /* DCT32x32 */
    default: return -1;
  }
}

void oracle_order_sample_decisions(size_t bw, size_t bh, uint8_t* decisions) {
  uint64_t s0 = 0x94D049BB133111EBull, s1 = 0xBF58476D1CE4E5B9ull;
  const size_t gw = (bw + 31) / 32, gh = (bh + 31) / 32;
  for (size_t gy = 0; gy < gh; ++gy) {
    for (size_t gx = 0; gx < gw; ++gx) {
      const size_t y1 = (gy + 1) * 32 < bh ? (gy + 1) * 32 : bh;
      const size_t x1 = (gx + 1) * 32 < bw ? (gx + 1) * 32 : bw;
      for (size_t y = gy * 32; y < y1; ++y) {
        for (size_t x = gx * 32; x < x1; ++x) {
          /* xorshift128+: the sum of the state, then the shift step */
          const uint64_t a = s0, b = s1;
          const uint64_t bits = a + b;
          uint64_t t = a ^ (a << 23);
          s0 = b;
          s1 = t ^ b ^ (t >> 18) ^ (b >> 5);
          decisions[y * bw + x] =
            (uint8_t)((bits >> 32) <= (0xFFFFFFFFull / 2));
        }
      }
    }
  }
}

int oracle_frame_statistics(const int32_t* quantized, const uint8_t* cells,
                            size_t bw, size_t bh, const int32_t* raw,
                            const uint8_t* decisions, uint32_t* zero_counts,
                            uint32_t* sampled, uint16_t* present_mask,
                            uint32_t* raw_counts) {
  const size_t gw = (bw + 31) / 32, gh = (bh + 31) / 32;
  const size_t store = gw * gh * 65536;
  memset(zero_counts, 0, 3 * 1984 * sizeof(uint32_t));
  memset(sampled, 0, 3 * 64 * sizeof(uint32_t));
  memset(raw_counts, 0, 13 * 256 * sizeof(uint32_t));
  uint16_t present = 0;
  for (size_t by = 0; by < bh; ++by) {
    for (size_t bx = 0; bx < bw; ++bx) {
      const uint8_t cell = cells[by * bw + bx];
      const int32_t q = raw[by * bw + bx];
      if (q < 1 || q > 256) return -1;
      ++raw_counts[kFamilyOfStrategy[cell >> 1] * 256 + (size_t)(q - 1)];
    }
  }
  for (size_t gy = 0; gy < gh; ++gy) {
    for (size_t gx = 0; gx < gw; ++gx) {
      const size_t base = (gy * gw + gx) * 65536;
      size_t offset = 0;
      const size_t y1 = (gy + 1) * 32 < bh ? (gy + 1) * 32 : bh;
      const size_t x1 = (gx + 1) * 32 < bw ? (gx + 1) * 32 : bw;
      for (size_t y = gy * 32; y < y1; ++y) {
        for (size_t x = gx * 32; x < x1; ++x) {
          const uint8_t cell = cells[y * bw + x];
          if (!(cell & 1)) continue;
          const int strategy = cell >> 1;
          int cw, ch;
          if (!oracle_strategy_covered(strategy, &cw, &ch)) return -2;
          const size_t n = (size_t)cw * ch * 64;
          const int family = kFamilyOfStrategy[strategy];
          const int at = FamilyOffset(family);
          if (at < 0) return -3;
          present |= (uint16_t)(1u << family);
          for (int c = 0; c < 3; ++c) {
            const int32_t* coefficients = quantized + (size_t)c * store + base + offset;
            uint32_t* counts = zero_counts + (size_t)c * 1984 + at;
            for (size_t k = 0; k < n; ++k) counts[k] += coefficients[k] == 0;
            if (family == 0 && decisions != NULL && decisions[y * bw + x]) {
              for (size_t k = 0; k < 64; ++k) sampled[c * 64 + k] += coefficients[k] == 0;
            }
          }
          offset += n;
        }
      }
    }
  }
  *present_mask = present;
  return 0;
}

/* One channel of one family: the natural order with the non-LLF part
 * stably sorted by key (insertion sort keeps the natural order on ties). */
static void SortOrder(const uint32_t* natural, size_t n, size_t llf,
                      const uint32_t* counts, uint32_t* out) {
  const float inverse_sqrt = 1.0f / sqrtf((float)n);
  uint64_t* keys = (uint64_t*)malloc(n * sizeof(uint64_t));
  for (size_t i = 0; i < n; ++i) {
    const float scaled = (float)counts[natural[i]] * inverse_sqrt;
    keys[i] = (uint64_t)(scaled + 0.1f);
    out[i] = natural[i];
  }
  for (size_t i = llf + 1; i < n; ++i) {
    const uint64_t key = keys[i];
    const uint32_t value = out[i];
    size_t j = i;
    while (j > llf && keys[j - 1] > key) {
      keys[j] = keys[j - 1];
      out[j] = out[j - 1];
      --j;
    }
    keys[j] = key;
    out[j] = value;
  }
  free(keys);
}

int oracle_coefficient_orders(const uint32_t* zero_counts, const uint32_t* sampled,
                              uint16_t present_mask, size_t bw, size_t bh,
                              int effort7_sampled, uint16_t* used_mask,
                              uint32_t* orders) {
  oracle_natural_order_table(orders);
  *used_mask = 0;
  if (bw < 5 && bh < 5) return 0;
  const int sample = effort7_sampled && present_mask == 1;
  static const int kFamilies[5] = {0, 4, 2, 6, 3};
  for (int f = 0; f < 5; ++f) {
    const int family = kFamilies[f];
    if (!(present_mask & (1u << family))) continue;
    const int strategy = FamilyStrategy(family);
    uint32_t natural[1024];
    const size_t n = oracle_natural_order(strategy, natural);
    int cw, ch;
    oracle_strategy_covered(strategy, &cw, &ch);
    const size_t llf = (size_t)cw * ch;
    uint32_t custom[3][1024];
    int nondefault = 0;
    for (int c = 0; c < 3; ++c) {
      const uint32_t* counts = (family == 0 && sample)
        ? sampled + c * 64 : zero_counts + (size_t)c * 1984 + FamilyOffset(family);
      SortOrder(natural, n, llf, counts, custom[c]);
      nondefault |= memcmp(custom[c], natural, n * sizeof(uint32_t)) != 0;
    }
    if (!nondefault) continue;
    *used_mask |= (uint16_t)(1u << family);
    /* every strategy stage of the family takes the family's order */
    for (int stage = 0; stage < 7; ++stage) {
      const int s = oracle_stage_strategy(stage);
      if (kFamilyOfStrategy[s] != family) continue;
      for (int c = 0; c < 3; ++c) {
        memcpy(orders + 3 * oracle_order_table_offset(stage) + (size_t)c * n,
               custom[c], n * sizeof(uint32_t));
      }
    }
  }
  return 0;
}

/* src/codestream/simple_ac_context.h:29-33 (kSimpleBlockContextMap, the
 * default map below 1024 blocks) */
static const uint8_t kDefaultMap[39] = {
  0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1,
  2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3, 3,
  2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3, 3,
};

int oracle_block_context_map(const uint32_t* raw_counts, size_t block_count,
                             uint32_t* thresholds, size_t* nthr, uint8_t* map,
                             uint32_t* num_contexts) {
  if (block_count < 1024) {
    memcpy(map, kDefaultMap, sizeof(kDefaultMap));
    *nthr = 0;
    *num_contexts = 4;
    return 0;
  }
  /* MedianQuantThreshold over the per-raw-quant totals */
  *nthr = 0;
  if (block_count >= 8192) {
    size_t cumulative = 0;
    for (size_t qf = 0; qf < 256; ++qf) {
      for (int f = 0; f < 13; ++f) cumulative += raw_counts[f * 256 + qf];
      if (cumulative > block_count / 2) {
        if (qf != 0) {
          thresholds[0] = (uint32_t)qf;
          *nthr = 1;
        }
        break;
      }
    }
  }
  const size_t segments = *nthr + 1;
  const size_t cells = 13 * segments;
  uint64_t cell_counts[26] = {0};
  for (size_t f = 0; f < 13; ++f) {
    for (size_t qf = 0; qf < 256; ++qf) {
      size_t segment = 0;
      while (segment < *nthr && (uint32_t)(qf + 1) > thresholds[segment]) ++segment;
      cell_counts[f * segments + segment] += raw_counts[f * 256 + qf];
    }
  }
  /* merge the least populated cluster into the next least, keeping the
   * previous round's order on equal counts, down to the luma target */
  uint8_t remap[26], clusters[26];
  size_t live = cells;
  for (size_t i = 0; i < cells; ++i) remap[i] = clusters[i] = (uint8_t)i;
  size_t luma = block_count / 2048;
  luma = luma < 2 ? 2 : (luma > 9 ? 9 : luma);
  while (live > luma) {
    /* stable insertion sort by count, descending */
    for (size_t i = 1; i < live; ++i) {
      const uint8_t v = clusters[i];
      size_t j = i;
      while (j > 0 && cell_counts[clusters[j - 1]] < cell_counts[v]) {
        clusters[j] = clusters[j - 1];
        --j;
      }
      clusters[j] = v;
    }
    const uint8_t destination = clusters[live - 2], source = clusters[live - 1];
    cell_counts[destination] += cell_counts[source];
    cell_counts[source] = 0;
    remap[source] = destination;
    --live;
  }
  for (size_t i = 0; i < cells; ++i) {
    while (remap[remap[i]] != remap[i]) remap[i] = remap[remap[i]];
  }
  uint8_t labels[26];
  memset(labels, 0xFF, sizeof(labels));
  uint8_t next = 0;
  for (size_t i = 0; i < cells; ++i) {
    if (labels[remap[i]] == 0xFF) labels[remap[i]] = next++;
    remap[i] = labels[remap[i]];
  }
  size_t chroma = block_count / 3072;
  chroma = chroma < 1 ? 1 : (chroma > 5 ? 5 : chroma);
  uint32_t max_label = 0;
  for (size_t i = 0; i < 3 * cells; ++i) {
    const uint8_t luma_label = remap[i % cells];
    map[i] = i < cells ? luma_label
                       : (uint8_t)(next + (luma_label < chroma - 1 ? luma_label : chroma - 1));
    if (map[i] > max_label) max_label = map[i];
  }
  *num_contexts = max_label + 1;
  return 0;
}

/* EncodeHybridUintValidated with the all-zero configuration
 * (src/codestream/entropy_internal.h:31-55), capped at 7 contexts. */
static uint32_t OrderContext(uint32_t value) {
  uint32_t symbol = value;
  if (value >= 1) {
    uint32_t exponent = 31;
    while (!(value >> exponent)) --exponent;
    symbol = 1 + exponent;
  }
  return symbol < 7 ? symbol : 7;
}

int oracle_order_tokens(uint16_t used_mask, const uint32_t* orders,
                        uint32_t* contexts, uint32_t* values, size_t capacity,
                        size_t* count) {
  static const int kFamilies[13] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12};
  size_t emitted = 0;
  for (int fi = 0; fi < 13; ++fi) {
    const int family = kFamilies[fi];
    if (!(used_mask & (1u << family))) continue;
    const int strategy = FamilyStrategy(family);
    if (strategy < 0) return -1;
    const int stage = oracle_strategy_stage(strategy);
    uint32_t natural[1024], lut[1024], permutation[1024], lehmer[1024];
    const size_t n = oracle_natural_order(strategy, natural);
    int cw, ch;
    oracle_strategy_covered(strategy, &cw, &ch);
    const size_t skip = (size_t)cw * ch;
    for (size_t i = 0; i < n; ++i) lut[natural[i]] = (uint32_t)i;
    for (int c = 0; c < 3; ++c) {
      const uint32_t* order = orders + 3 * oracle_order_table_offset(stage) + (size_t)c * n;
      for (size_t i = 0; i < n; ++i) permutation[i] = lut[order[i]];
      /* Lehmer code: each value minus the smaller values already used */
      for (size_t i = 0; i < n; ++i) {
        uint32_t smaller = 0;
        for (size_t j = 0; j < i; ++j) smaller += permutation[j] < permutation[i];
        lehmer[i] = permutation[i] - smaller;
      }
      size_t end = n;
      while (end > skip && lehmer[end - 1] == 0) --end;
      if (emitted + 1 + (end - skip) > capacity) return -2;
      contexts[emitted] = OrderContext((uint32_t)n);
      values[emitted++] = (uint32_t)(end - skip);
      uint32_t last = 0;
      for (size_t i = skip; i < end; ++i) {
        contexts[emitted] = OrderContext(last);
        values[emitted++] = lehmer[i];
        last = lehmer[i];
      }
    }
  }
  *count = emitted;
  return 0;
}
