/*
 * a19: AC tokenization of the quantized frame.  TEST INFRASTRUCTURE ONLY
 * (see oracle.h).
 *
 * Restates src/codestream/ac_group.cpp: the natural coefficient orders
 * (ComputeSimpleNaturalCoefficientOrder, :381-487) and the direct
 * per-group tokenizer (TokenizeSimpleAcGroupDirectValidated, :740-932) with
 * its helpers (NonzeroBucket :82-86, ZeroDensityContext :88-100,
 * PredictNonzeros :102-113, BlockContextValidated :115-131,
 * FinalNonzeroContext / FinalCoefficientContext :133-145,
 * CountNonzerosExceptLlf :268-288, AppendDirectAcToken's population symbol
 * :688-720), and PackSigned (src/codestream/entropy.h).
 *
 * Pinned: the reference's own known answers (tests/ac_group_test.cpp:36-260:
 * FNV-1a hashes of the seven natural orders and of the token streams of a
 * poisoned single-transform group per strategy, first-token block contexts)
 * plus brute-force invariants (orders are permutations with the LLF prefix;
 * the token stream decodes back to the coefficients).
 *
 * Frame layout (shared with the product): 3 planes of groups * 65536 int32
 * coefficients; group g = (by / 32) * groups_w + bx / 32; a group's transforms
 * in raster order of their anchors, each W*H coefficients in the reference's
 * storage layout (longer axis contiguous).
 *
 * Order table layout: the seven strategies in the order DCT8, 16x8, 8x16,
 * 16x16, 32x16, 16x32, 32x32 (sizes 64, 128, 128, 256, 512, 512, 1024);
 * strategy stage s, channel c (0 = X, 1 = Y, 2 = B) starts at
 * 3 * oracle_order_table_offset(s) + c * N_s.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* src/codestream/ac_group.cpp:40-53 */
static const uint16_t kFrequencyContext[64] = {
  0xBAD, 0,  1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14,
  15,    15, 16, 16, 17, 17, 18, 18, 19, 19, 20, 20, 21, 21, 22, 22,
  23,    23, 23, 23, 24, 24, 24, 24, 25, 25, 25, 25, 26, 26, 26, 26,
  27,    27, 27, 27, 28, 28, 28, 28, 29, 29, 29, 29, 30, 30, 30, 30,
};
static const uint16_t kNonzeroContext[64] = {
  0xBAD, 0,   31,  62,  62,  93,  93,  93,  93,  123, 123, 123, 123,
  152,   152, 152, 152, 152, 152, 152, 152, 180, 180, 180, 180, 180,
  180,   180, 180, 180, 180, 180, 180, 206, 206, 206, 206, 206, 206,
  206,   206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206,
  206,   206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206,
};
/* src/codestream/simple_ac_context.h:20-23: strategy code -> order family */
static const uint8_t kStrategyOrder[27] = {
  0, 1, 1, 1, 2, 3, 4, 4, 5, 5, 6, 6, 1, 1,
  1, 1, 1, 1, 7, 8, 8, 9, 10, 10, 11, 12, 12,
};
enum { kOrderFamilies = 13, kNonzeroBuckets = 37, kZeroDensityContexts = 458 };

static const int kStageStrategy[7] = {0, 6, 7, 4, 10, 11, 5};

int oracle_stage_strategy(int stage) {
  return stage >= 0 && stage < 7 ? kStageStrategy[stage] : -1;
}

int oracle_strategy_stage(int strategy) {
  for (int s = 0; s < 7; ++s) {
    if (kStageStrategy[s] == strategy) return s;
  }
  return -1;
}

size_t oracle_order_table_offset(int stage) {
  size_t offset = 0;
  for (int s = 0; s < stage && s < 7; ++s) {
    int cw, ch;
    oracle_strategy_covered(kStageStrategy[s], &cw, &ch);
    offset += (size_t)cw * ch * 64;
  }
  return offset;
}

/* ac_group.cpp:381-487.  Zig-zag over the coefficient_width^2 square of
 * frequencies: along each anti-diagonal, alternating direction, keeping the
 * rows that are multiples of the anisotropy (the transform is stored with its
 * longer axis contiguous); the covered-block positions form the LLF prefix. */
size_t oracle_natural_order(int strategy, uint32_t* order) {
  int cw, ch;
  if (!oracle_strategy_covered(strategy, &cw, &ch)) return 0;
  const size_t blocks = (size_t)(cw > ch ? cw : ch);
  const size_t rows = (size_t)(cw > ch ? ch : cw);
  const size_t aniso = blocks / rows;
  size_t shift = 0;
  while (((size_t)1 << shift) < aniso) ++shift;
  const size_t mask = aniso - 1;
  const size_t width = blocks * 8;
  const size_t count = width * rows * 8;
  size_t next = blocks * rows;
  for (size_t d = 0; d < width; ++d) {
    for (size_t p = 0; p <= d; ++p) {
      size_t x = p, y = d - p;
      if (d & 1) { const size_t t = x; x = y; y = t; }
      if (y & mask) continue;
      y >>= shift;
      const size_t index = x < blocks && y < rows ? y * blocks + x : next++;
      if (index >= count) return 0;
      order[index] = (uint32_t)(y * width + x);
    }
  }
  for (size_t r = width - 1; r > 0; --r) {
    const size_t d = r - 1;
    for (size_t p = 0; p <= d; ++p) {
      size_t x = width - 1 - (d - p), y = width - 1 - p;
      if (d & 1) { const size_t t = x; x = y; y = t; }
      if (y & mask) continue;
      y >>= shift;
      if (next >= count) return 0;
      order[next++] = (uint32_t)(y * width + x);
    }
  }
  return next == count ? count : 0;
}

void oracle_natural_order_table(uint32_t* table) {
  for (int s = 0; s < 7; ++s) {
    const size_t base = 3 * oracle_order_table_offset(s);
    const size_t n = oracle_natural_order(kStageStrategy[s], table + base);
    memcpy(table + base + n, table + base, n * sizeof(uint32_t));
    memcpy(table + base + 2 * n, table + base, n * sizeof(uint32_t));
  }
}

/* entropy.h: zig-zag of a signed value. */
uint32_t oracle_pack_signed(int32_t v) {
  return ((uint32_t)v << 1) ^ (((uint32_t)~v >> 31) - 1u);
}

static uint32_t NonzeroBucket(uint32_t p) {
  return p < 8 ? p : p >= 64 ? 36 : 4 + p / 2;
}

/* AppendDirectAcToken (:688-720): HybridUint(4, 2, 0) symbol of a value. */
uint32_t oracle_token_symbol(uint32_t value) {
  if (value < 16) return value;
  uint32_t e = 31;
  while (!(value >> e)) --e;
  return 16 + ((e - 4) << 2) + ((value - (1u << e)) >> (e - 2));
}

/* ValidateSimpleBlockContextMap (block_context_map.cpp:281-320). */
int oracle_validate_context_map(const uint32_t* thresholds, size_t nthr,
                                const uint8_t* map, uint32_t num_contexts) {
  if (nthr > 15 || num_contexts == 0 || num_contexts > 16) return -1;
  for (size_t i = 0; i < nthr; ++i) {
    if (thresholds[i] == 0 || thresholds[i] > 255) return -1;
    if (i && thresholds[i - 1] >= thresholds[i]) return -1;
  }
  const size_t entries = 3 * kOrderFamilies * (nthr + 1);
  int seen[16] = {0};
  uint32_t max = 0;
  for (size_t i = 0; i < entries; ++i) {
    if (map[i] >= num_contexts) return -1;
    seen[map[i]] = 1;
    if (map[i] > max) max = map[i];
  }
  if (max + 1 != num_contexts) return -1;
  for (uint32_t c = 0; c < num_contexts; ++c) {
    if (!seen[c]) return -1;
  }
  return 0;
}

/* One AC group.  Returns 0, or -2 raw quant outside [1, 256], -3 a strategy
 * grid that does not tile the group, -4 capacity exceeded, -5 an order whose
 * scan misses a nonzero (not an LLF-prefixed permutation). */
int oracle_tokenize_group(const int32_t* quantized, size_t store,
                          size_t group_base, const uint8_t* cells, size_t bw,
                          size_t bh, size_t gx0, size_t gy0,
                          const int32_t* raw, const uint32_t* orders,
                          const uint32_t* thresholds, size_t nthr,
                          const uint8_t* context_map, uint32_t num_contexts,
                          uint32_t* values, uint16_t* contexts,
                          size_t capacity, size_t* count,
                          uint32_t* histogram) {
  const size_t gw = bw - gx0 < 32 ? bw - gx0 : 32;
  const size_t gh = bh - gy0 < 32 ? bh - gy0 : 32;
  uint8_t maps[3][32 * 32];
  uint8_t covered_cells[32 * 32];
  memset(maps, 0, sizeof(maps));
  memset(covered_cells, 0, sizeof(covered_cells));
  size_t n = 0, source = 0;
  static const int kChannelOrder[3] = {1, 0, 2};
  for (size_t ly = 0; ly < gh; ++ly) {
    for (size_t lx = 0; lx < gw; ++lx) {
      const uint8_t cell = cells[(gy0 + ly) * bw + gx0 + lx];
      if (!(cell & 1)) {
        if (!covered_cells[ly * 32 + lx]) return -3;
        continue;
      }
      const int strategy = cell >> 1;
      const int stage = oracle_strategy_stage(strategy);
      int cw, ch;
      if (stage < 0 || !oracle_strategy_covered(strategy, &cw, &ch)) return -3;
      if (lx + cw > gw || ly + ch > gh) return -3;
      for (int dy = 0; dy < ch; ++dy) {
        for (int dx = 0; dx < cw; ++dx) {
          if (covered_cells[(ly + dy) * 32 + lx + dx]) return -3;
          covered_cells[(ly + dy) * 32 + lx + dx] = 1;
        }
      }
      const size_t covered = (size_t)cw * ch;
      const size_t N = covered * 64;
      size_t log2c = 0;
      while (((size_t)1 << log2c) < covered) ++log2c;
      const size_t family = kStrategyOrder[strategy];
      const int32_t rq = raw[(gy0 + ly) * bw + gx0 + lx];
      if (rq < 1 || rq > 256) return -2;
      size_t segment = 0;
      while (segment < nthr && (uint32_t)rq > thresholds[segment]) ++segment;
      const size_t coef_w = (size_t)(cw > ch ? cw : ch) * 8;
      const size_t llf_w = (size_t)(cw > ch ? cw : ch);
      const size_t llf_h = (size_t)(cw > ch ? ch : cw);
      for (int k = 0; k < 3; ++k) {
        const int c = kChannelOrder[k];
        const int32_t* coef = quantized + (size_t)c * store + group_base + source;
        const uint32_t* order =
          orders + 3 * oracle_order_table_offset(stage) + (size_t)c * N;
        /* CountNonzerosExceptLlf (:268-288) */
        int32_t nonzeros = 0;
        for (size_t i = 0; i < N; ++i) nonzeros += coef[i] != 0;
        for (size_t y = 0; y < llf_h; ++y) {
          for (size_t x = 0; x < llf_w; ++x) nonzeros -= coef[y * coef_w + x] != 0;
        }
        const uint8_t scaled =
          (uint8_t)((nonzeros + (int32_t)covered - 1) / (int32_t)covered);
        uint8_t* map = maps[c];
        for (int dy = 0; dy < ch; ++dy) {
          for (int dx = 0; dx < cw; ++dx) map[(ly + dy) * gw + lx + dx] = scaled;
        }
        /* PredictNonzeros (:102-113) on the group's block extent */
        uint32_t prediction;
        if (lx == 0) {
          prediction = ly == 0 ? 32 : map[(ly - 1) * gw];
        } else if (ly == 0) {
          prediction = map[lx - 1];
        } else {
          prediction = (map[(ly - 1) * gw + lx] + map[ly * gw + lx - 1] + 1) / 2;
        }
        const size_t channel_row = c < 2 ? (size_t)(c ^ 1) : 2;
        const uint32_t block_context =
          context_map[(channel_row * kOrderFamilies + family) * (nthr + 1) + segment];
        if (n >= capacity) return -4;
        values[n] = (uint32_t)nonzeros;
        contexts[n] = (uint16_t)(NonzeroBucket(prediction) * num_contexts + block_context);
        if (histogram) ++histogram[(size_t)contexts[n] * 128 + oracle_token_symbol(values[n])];
        ++n;
        int32_t remaining = nonzeros;
        uint32_t previous = nonzeros > (int32_t)(N / 16) ? 0 : 1;
        for (size_t scan = covered; scan < N && remaining != 0; ++scan) {
          const int32_t v = coef[order[scan]];
          const size_t norm_nz = ((size_t)remaining + covered - 1) >> log2c;
          const uint32_t local =
            (kNonzeroContext[norm_nz] + kFrequencyContext[scan >> log2c]) * 2 + previous;
          if (n >= capacity) return -4;
          values[n] = oracle_pack_signed(v);
          contexts[n] = (uint16_t)(num_contexts * kNonzeroBuckets +
                                   kZeroDensityContexts * block_context + local);
          if (histogram) {
            ++histogram[(size_t)contexts[n] * 128 + oracle_token_symbol(values[n])];
          }
          ++n;
          previous = v != 0;
          remaining -= (int32_t)previous;
        }
        if (remaining != 0) return -5;
      }
      source += N;
    }
  }
  *count = n;
  return 0;
}

int oracle_tokenize_frame(const int32_t* quantized, const uint8_t* cells,
                          size_t bw, size_t bh, const int32_t* raw,
                          const uint32_t* orders, const uint32_t* thresholds,
                          size_t nthr, const uint8_t* context_map,
                          uint32_t num_contexts, uint32_t* values,
                          uint16_t* contexts, size_t capacity,
                          uint32_t* group_offsets, uint32_t* histogram) {
  if (oracle_validate_context_map(thresholds, nthr, context_map, num_contexts)) {
    return -1;
  }
  const size_t groups_w = (bw + 31) / 32, groups_h = (bh + 31) / 32;
  const size_t store = groups_w * groups_h * 65536;
  size_t total = 0;
  for (size_t g = 0; g < groups_w * groups_h; ++g) {
    group_offsets[g] = (uint32_t)total;
    size_t n = 0;
    const int rc = oracle_tokenize_group(
      quantized, store, g * 65536, cells, bw, bh, (g % groups_w) * 32,
      (g / groups_w) * 32, raw, orders, thresholds, nthr, context_map,
      num_contexts, values + total, contexts + total, capacity - total, &n,
      histogram);
    if (rc) return rc;
    total += n;
  }
  group_offsets[groups_w * groups_h] = (uint32_t)total;
  return 0;
}
