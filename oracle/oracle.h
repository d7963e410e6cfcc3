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

/* ---- fast math: src/util/fast_math.h -------------------------------- */
float oracle_fast_log2(float value);
float oracle_fast_pow2(float value);
float oracle_fast_pow(float base, float exponent);

#ifdef __cplusplus
}
#endif

#endif /* GJXL_ORACLE_H_ */
