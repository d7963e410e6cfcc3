/*
 * CPU oracle for the gjxl perceptual-analysis path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference encoder's CPU algorithms; it exists so that tests, the smoke check
 * and the CPU-baseline leg of the benchmark can check the HIP product path.
 * Nothing in the product (gjxl_amd/) may link, load or call it.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the gjxl source tree).  All arithmetic is IEEE binary32/binary64 in source
 * order, compiled with -ffp-contract=off; fmaf/fma are used exactly where the
 * reference calls std::fma.
 *
 * Pinning status of each function is stated beside its declaration.
 */
#ifndef GJXL_ORACLE_H_
#define GJXL_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- a1: packed sRGB8 -> linear planes ------------------------------- */
/* src/c_api/image_conversion.cpp:18-31 (table), :123-145 (conversion).
 * Pinned: closed form (sRGB EOTF) + reference round-trip test values. */
void oracle_srgb_to_linear_table(float table[256]);
void oracle_srgb8_to_linear(const uint8_t* pixels, size_t row_stride_bytes,
                            uint32_t bytes_per_pixel, uint32_t width,
                            uint32_t height, float* linear /* 3*w*h */);

/* ---- a2: linear RGB -> padded XYB ------------------------------------ */
/* src/codec/color_transform.cpp:54-79 (cube root), :233-362 (transform),
 * :395-431 (padding).  Pinned: gray => X=0,Y=B; black => origin; round trip
 * through oracle_opsin_to_linear (color_transform_test.cpp:101-140). */
float oracle_fast_cube_root_and_add(float value, float add);
void oracle_linear_to_padded_opsin(const float* linear, uint32_t width,
                                   uint32_t height, float intensity_target,
                                   uint32_t padded_width,
                                   uint32_t padded_height, float* opsin);
/* src/codec/color_transform.cpp:433-471 */
void oracle_opsin_to_linear(const float* opsin, size_t count,
                            float intensity_target, float* linear);

/* ---- a4: Gaborish inverse ------------------------------------------- */
/* src/core/image_ops.h:20-47 */
size_t oracle_mirror(long long coordinate, size_t size);
/* src/codec/convolution.cpp:38-119; weights = d0, d1, d2, d4, d8, d5.
 * Pinned: bit-exact against the reference's compiled convolution.cpp. */
void oracle_convolve_symmetric5(const float* in, size_t w, size_t h,
                                const float weights[6], float* out);
/* src/codec/gaborish.cpp:22-47 */
void oracle_gaborish_inverse_weights(float multiplier, float weights[6]);
/* src/codec/gaborish.cpp:49-99.  Pinned: bit-exact against the reference's
 * compiled gaborish.cpp, plus tests/gaborish_test.cpp pinned points. */
void oracle_gaborish_inverse(const float* in, size_t w, size_t h,
                             const float multipliers[3], float* out);

/* ---- a5: initial quant field --------------------------------------- */
/* src/codec/adaptive_quantization.cpp:583-738 (non-uniform branch).
 * Pinned: tests/adaptive_quantization_test.cpp:184-250 libjxl goldens
 * (block maps 3e-6, pixel mask 3e-4), closed-form properties.  The unit
 * itself is behind the reference's toolchain gate (not compiled here). */
int oracle_initial_quant_field(const float* opsin, size_t w, size_t h,
                               float target, float rescale, float* quant_field,
                               float* strategy_mask, float* pixel_mask);
/* src/codec/adaptive_quantization.cpp:556-580 */
float oracle_initial_quant_dc(float target);

/* ---- a6-a10: quantization matrices, DCT, CfL, AC strategy ----------- */
/* Strategies searched: 0 DCT8, 4 DCT16x16, 5 DCT32x32, 6 DCT16x8,
 * 7 DCT8x16, 10 DCT32x16, 11 DCT16x32 (src/core/ac_strategy.h:25-148). */
int oracle_strategy_covered(int strategy, int* blocks_w, int* blocks_h);
/* src/codec/quantization_tables_generated.h via
 * scripts/generate_quantization_tables.py; returns the coefficient count.
 * Pinned: bit-exact against the reference's compiled quantization.cpp. */
int oracle_quant_matrix(int strategy, int channel, float* dequant,
                        float* inverse_dequant);
/* src/codec/dct.cpp:135-233.  Pinned: bit-exact against compiled dct.cpp
 * and the FP64 textbook DCT. */
int oracle_forward_dct(int strategy, const float* pixels, float* coefficients);
int oracle_inverse_dct(int strategy, const float* coefficients, float* pixels);
/* src/codec/chroma_from_luma.cpp:112-165.  Pinned: compiled reference. */
int8_t oracle_find_best_multiplier(const float* luma, const float* chroma,
                                   size_t n, float base, int fast,
                                   uint32_t iterations);
/* src/codec/chroma_from_luma.cpp:434-529.  Pinned: compiled reference. */
int oracle_initial_cfl(const float* opsin, size_t w, size_t h, int8_t* y_to_x,
                       int8_t* y_to_b);
/* src/codec/ac_strategy.cpp:823-994.  Pinned: compiled reference. */
int oracle_ac_strategy_cost(int strategy, size_t bx, size_t by,
                            const float* opsin, size_t w, size_t h,
                            const float* quant_field, const float* pixel_mask,
                            float target, float entropy_multiplier,
                            float cfl_x, float cfl_b, float* cost);
/* Candidate cost table (7 stages x blocks, stage order of
 * src/codec/ac_strategy_search_policy.h:20-28) and the table-driven
 * search (src/codec/ac_strategy.cpp:591-773, 996-1099).  Pinned: grids
 * equal to the compiled reference's FindAcStrategyGrid. */
int oracle_stage_index(int strategy);
int oracle_candidate_costs(const float* opsin, size_t w, size_t h,
                           const float* quant_field, const float* pixel_mask,
                           const int8_t* y_to_x, const int8_t* y_to_b,
                           float target, float* costs);
void oracle_search_from_costs(const float* costs, size_t bw, size_t bh,
                              float target, uint8_t* cells);

/* ---- a11-a15: one AQ evaluation without Butteraugli (oracle/aq.c) ----- */
/* src/codestream/workflow.cpp:1318-1484.  stats: x_edge, b_edge, exposed
 * blue over the unpadded image (stride/plane in floats). */
void oracle_matrix_scale_stats(const float* opsin, size_t w, size_t h,
                               size_t stride, size_t plane, float stats[3]);
void oracle_select_matrix_scales(const float stats[3], float target,
                                 int* x_scale, int* b_scale);
float oracle_matrix_multiplier(int scale);
/* src/codec/adaptive_quantization.cpp:740-835, :1178-1242 */
void oracle_adjust_quant_field(const uint8_t* cells, size_t bw, size_t bh,
                               float target, const float* in, float* out);
int oracle_policy_setup(const float* adjusted, size_t n, float target,
                        float* lower, float* upper, float* quant_dc);
/* src/codec/quantization.cpp:536-639.  q: scale, inverse global scale,
 * global scale, quant dc, dc steps (3), inverse dc steps (3). */
int oracle_quantizer_from_field(float quant_dc, const float* field, size_t n,
                                int32_t* raw, float q[10]);
/* src/codec/chroma_from_luma.cpp:543-670 (fast) over per-anchor
 * coefficients stored at offsets[] (oracle_anchor_offsets). */
void oracle_final_cfl_fast(const float* const coefs[3], const uint8_t* cells,
                           size_t bw, size_t bh, const size_t* offsets,
                           const int32_t* raw, float scale, int8_t* y_to_x,
                           int8_t* y_to_b);
/* src/codec/dc_conversion.cpp */
void oracle_llf_to_dc(int cw, int ch, const float* coef, float* dc);
void oracle_dc_to_llf(int cw, int ch, const float* dc, float* coef);
/* src/codec/quantization.cpp:808-835 */
void oracle_dequantize_ac(float inverse_global_scale, int32_t raw, float mult,
                          int channel, int n, const float* dequant,
                          const int32_t* quantized, float* out);
/* src/codec/reconstruction.cpp:565-700 for one transform; dc: 3 x 16. */
void oracle_code_transform(int strategy, const float q[10], int32_t raw,
                           float x_mult, float b_mult, float cfl_x,
                           float cfl_b, float* coef, int32_t* quantized,
                           float* dc, int32_t* decided_raw);
/* src/codec/vardct/dc_quantization.cpp (prediction-aware, weighted). */
void oracle_quantize_dc(const float* dc, size_t bw, size_t bh,
                        const float dc_steps[3], const float inv_dc_steps[3],
                        int extra_precision, int32_t* quantized,
                        float* reconstructed);
/* src/codec/dc_smoothing.cpp */
void oracle_smooth_dc(const float* dc, size_t bw, size_t bh,
                      const float dc_steps[3], float* out);
/* Dense per-group coefficient layout (groups of 256x256 px, 65536 slots
 * each, anchors of a group in raster order). */
void oracle_anchor_offsets(const uint8_t* cells, size_t bw, size_t bh,
                           size_t* offsets);
void oracle_anchor_coefficients(const float* opsin, size_t w, size_t h,
                                const uint8_t* cells, const size_t* offsets,
                                float* coefs);
/* src/codec/reconstruction.cpp:386-760 over the whole frame. */
void oracle_code_coefficients(const float* coefs, const uint8_t* cells,
                              size_t bw, size_t bh, const size_t* offsets,
                              const float q[10], int32_t* raw,
                              const int8_t* y_to_x, const int8_t* y_to_b,
                              float x_mult, float b_mult, int32_t* quantized,
                              int32_t* dc_q, float* dc_rec);
/* src/codec/epf.cpp:207-290 (neutral sharpness) */
void oracle_epf_inverse_sigma(const uint8_t* cells, size_t bw, size_t bh,
                              const int32_t* raw, float scale, float* out);
/* src/codec/reconstruction.cpp:790-936 with adaptive DC smoothing. */
void oracle_reconstruct(const int32_t* quantized, const uint8_t* cells,
                        size_t w, size_t h, const size_t* offsets,
                        const float q[10], const int32_t* raw,
                        const int8_t* y_to_x, const int8_t* y_to_b,
                        float x_mult, float b_mult, const float* dc_rec,
                        float* out);
/* src/codec/gaborish.cpp:101-190, src/codec/loop_filter.cpp:14-65,
 * src/codec/epf.cpp:61-205 on an unpadded image. */
void oracle_gaborish(const float* in, size_t w, size_t h, float* out);
void oracle_loop_filters(const float* in, size_t w, size_t h,
                         const float* inverse_sigma, float* out);

/* ---- a16-a18: Butteraugli, 16-norm reduction, AQ update (butteraugli.c) */
/* src/codec/butteraugli_internal.cpp:565-620: kernel taps (returns size). */
int oracle_butteraugli_kernel(float sigma, float* kernel /* <= 33 */);
/* GaussianBlur (:576-620).  Pinned: bit-exact against the compiled
 * reference; libjxl blur goldens (tests/butteraugli_goldens_generated.h). */
int oracle_butteraugli_blur(const float* in, size_t w, size_t h, float sigma,
                            float* out);
/* OpsinDynamicsImage (:622-700).  Pinned as above. */
int oracle_butteraugli_opsin(const float* rgb, size_t w, size_t h,
                             float intensity, float* xyb);
/* SeparateFrequencies (:702-846): 10 planes LF(3) MF(3) HF(2) UHF(2). */
int oracle_butteraugli_frequencies(const float* xyb, size_t w, size_t h,
                                   float* psycho);
/* FuzzyErosion (src/codec/butteraugli_distance_internal.cpp:579-630). */
void oracle_butteraugli_erosion(const float* in, size_t w, size_t h, float* out);
/* ComputeDifferenceStages (:657-870): 9 stage planes. */
int oracle_butteraugli_differences(const float* ps0, const float* ps1, size_t w,
                                   size_t h, float hf_asymmetry,
                                   float x_multiplier, float* stages);
/* Subsample2x (:380-407): 3 planes ((w+1)/2) x ((h+1)/2). */
void oracle_butteraugli_subsample(const float* in, size_t w, size_t h,
                                  float* out);
/* ComputeButteraugliDistanceNative (:872-1003).  Pinned: compiled reference
 * and libjxl full-map goldens. */
int oracle_butteraugli_distance(const float* ref, const float* dist, size_t w,
                                size_t h, float hf_asymmetry,
                                float x_multiplier, float intensity, float* map,
                                double* score);
/* src/codec/butteraugli.cpp:116-222.  Pinned: compiled reference. */
void oracle_reduce_distance_map(const float* map, size_t w, size_t h,
                                const uint8_t* cells, size_t bw, size_t bh,
                                float* out);
/* src/codec/adaptive_quantization.cpp:1686-1733, one update after
 * evaluation `iteration`.  Pinned: closed-form cases (the unit is behind the
 * reference's toolchain gate). */
void oracle_aq_update(float* field, const float* adjusted_initial,
                      const float* block_distance, size_t n, float target,
                      int iteration, float lower, float upper, float scale,
                      float inverse_global_scale);

/* ---- a19: AC tokenization (tokenize.c) ------------------------------ */
/* src/codestream/ac_group.cpp:381-487.  Returns the size, 0 if unsupported.
 * Pinned: FNV-1a hashes and samples of tests/ac_group_test.cpp:56-73. */
size_t oracle_natural_order(int strategy, uint32_t* order /* <= 1024 */);
/* Strategy stages DCT8, 16x8, 8x16, 16x16, 32x16, 16x32, 32x32. */
int oracle_stage_strategy(int stage);
int oracle_strategy_stage(int strategy);
size_t oracle_order_table_offset(int stage);
/* 3 * 2624 entries: natural orders for every stage and channel. */
void oracle_natural_order_table(uint32_t* table);
uint32_t oracle_pack_signed(int32_t v);
uint32_t oracle_token_symbol(uint32_t value);
/* src/codestream/block_context_map.cpp:281-320 */
int oracle_validate_context_map(const uint32_t* thresholds, size_t nthr,
                                const uint8_t* map, uint32_t num_contexts);
/* TokenizeSimpleAcGroupDirectValidated (ac_group.cpp:740-932) on one group.
 * Pinned: token-stream hashes and block contexts of
 * tests/ac_group_test.cpp:130-260; decode round trip. */
int oracle_tokenize_group(const int32_t* quantized, size_t store,
                          size_t group_base, const uint8_t* cells, size_t bw,
                          size_t bh, size_t gx0, size_t gy0,
                          const int32_t* raw, const uint32_t* orders,
                          const uint32_t* thresholds, size_t nthr,
                          const uint8_t* context_map, uint32_t num_contexts,
                          uint32_t* values, uint16_t* contexts,
                          size_t capacity, size_t* count,
                          uint32_t* histogram /* nullable, contexts*128 */);
/* Every group of the frame in group order; group_offsets: groups + 1. */
int oracle_tokenize_frame(const int32_t* quantized, const uint8_t* cells,
                          size_t bw, size_t bh, const int32_t* raw,
                          const uint32_t* orders, const uint32_t* thresholds,
                          size_t nthr, const uint8_t* context_map,
                          uint32_t num_contexts, uint32_t* values,
                          uint16_t* contexts, size_t capacity,
                          uint32_t* group_offsets, uint32_t* histogram);

/* ---- f1/f2: effort-7 tokenization inputs (entropy_inputs.c) ---------- */
/* UseCoefficientOrderSample draws (coefficient_order.cpp:115-126,626-650),
 * one per block, groups in raster order.  Pinned: the restatement in
 * tests/coefficient_order_test.cpp:47-77 (seeds, threshold). */
void oracle_order_sample_decisions(size_t bw, size_t bh, uint8_t* decisions);
/* CountGroupZeros (coefficient_order.cpp:223-316) and
 * CountRawQuantAndOrders (block_context_map.cpp:67-103); layouts in
 * entropy_inputs.c.  Pinned: brute-force recount in the tests. */
int oracle_frame_statistics(const int32_t* quantized, const uint8_t* cells,
                            size_t bw, size_t bh, const int32_t* raw,
                            const uint8_t* decisions /* nullable */,
                            uint32_t* zero_counts /* 3*1984 */,
                            uint32_t* sampled /* 3*64 */, uint16_t* present_mask,
                            uint32_t* raw_counts /* 13*256 */);
/* ComputeCoefficientOrdersWithCounts (coefficient_order.cpp:736-811) into
 * the 7-stage order table.  Pinned: tests/coefficient_order_test.cpp:32-118
 * (stable sort by uint64(zeros / sqrt(n) + 0.1f) after the LLF prefix). */
int oracle_coefficient_orders(const uint32_t* zero_counts, const uint32_t* sampled,
                              uint16_t present_mask, size_t bw, size_t bh,
                              int effort7_sampled, uint16_t* used_mask,
                              uint32_t* orders /* 3*2624 */);
/* ComputeSimpleBlockContextMapForEncoder (block_context_map.cpp:435-476).
 * Pinned: default map (tests/block_context_map_test.cpp:15-39) and the
 * invariants of tests/test_oracle_entropy_inputs.py; parity unpinned beyond
 * them for the merge order (no known answer in the reference's tests). */
int oracle_block_context_map(const uint32_t* raw_counts, size_t block_count,
                             uint32_t* thresholds, size_t* nthr, uint8_t* map,
                             uint32_t* num_contexts);
/* TokenizeSimpleCoefficientOrders (coefficient_order.cpp:876-919).
 * Pinned: decode round trip of the Lehmer code in the tests. */
int oracle_order_tokens(uint16_t used_mask, const uint32_t* orders,
                        uint32_t* contexts, uint32_t* values, size_t capacity,
                        size_t* count);

/* ---- fast math: src/util/fast_math.h -------------------------------- */
float oracle_fast_log2(float value);
float oracle_fast_pow2(float value);
float oracle_fast_pow(float base, float exponent);

#ifdef __cplusplus
}
#endif

#endif /* GJXL_ORACLE_H_ */
