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

/* ---- fast math: src/util/fast_math.h -------------------------------- */
float oracle_fast_log2(float value);
float oracle_fast_pow2(float value);
float oracle_fast_pow(float base, float exponent);

#ifdef __cplusplus
}
#endif

#endif /* GJXL_ORACLE_H_ */
