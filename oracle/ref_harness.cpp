// Harness over the reference encoder's own CPU sources, compiled where they
// lie in the gjxl source tree (see oracle/Makefile).  TEST INFRASTRUCTURE
// ONLY: it validates the plain-C oracle and serves as the reference CPU
// baseline.  Only reference translation units that build with this image's
// toolchain are linked; the rest (color_transform, adaptive_quantization,
// reconstruction, ac_group) stop at the reference's own toolchain gate and are
// pinned by other means.
//
// Each entry point is extern "C" with plain pointers and sizes, planar float
// images with stride == width, and returns the reference Status code (0 = ok).

#include <array>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <span>

#include "c_api/image_conversion.h"
#include "codec/ac_strategy.h"
#include "codec/butteraugli.h"
#include "codec/chroma_from_luma.h"
#include "codec/dct.h"
#include "codec/gaborish.h"
#include "codec/quantization.h"
#include "core/ac_strategy.h"
#include "core/image.h"

namespace {

gjxl::ConstImage3FView ConstImage(const float* data, size_t width,
                                  size_t height) {
  const size_t plane = width * height;
  gjxl::ConstImage3FView view;
  for (size_t c = 0; c < 3; ++c) {
    view.plane[c] = {data + c * plane, {width, height}, width};
  }
  return view;
}

gjxl::Image3FView MutableImage(float* data, size_t width, size_t height) {
  const size_t plane = width * height;
  gjxl::Image3FView view;
  for (size_t c = 0; c < 3; ++c) {
    view.plane[c] = {data + c * plane, {width, height}, width};
  }
  return view;
}

int Code(const gjxl::Status& status) {
  return static_cast<int>(status.code());
}

}  // namespace

extern "C" {

// src/c_api/image_conversion.cpp:123-145
int ref_srgb8_to_linear(const uint8_t* pixels, size_t pixels_size,
                        size_t row_stride_bytes, uint32_t bytes_per_pixel,
                        uint32_t width, uint32_t height, float* linear) {
  gjxl::c_api_internal::PackedSrgbImageView image;
  image.pixels = pixels;
  image.pixels_size = pixels_size;
  image.width = width;
  image.height = height;
  image.row_stride_bytes = row_stride_bytes;
  image.format = bytes_per_pixel == 3
    ? gjxl::c_api_internal::PackedPixelFormat::kRgb8Srgb
    : gjxl::c_api_internal::PackedPixelFormat::kRgba8Srgb;
  return Code(gjxl::c_api_internal::ConvertValidatedPackedSrgbInto(
    image, MutableImage(linear, width, height)));
}

// src/codec/dct.cpp:135-182
int ref_forward_dct(int strategy, const float* pixels, float* coefficients) {
  const auto type = static_cast<gjxl::AcStrategyType>(strategy);
  const auto* info = gjxl::GetAcStrategyInfo(type);
  if (info == nullptr) return -1;
  const size_t n = info->coefficient_count();
  return Code(gjxl::ForwardDctCpu(type, std::span<const float>(pixels, n),
                                  std::span<float>(coefficients, n)));
}

// src/codec/dct.cpp:184-233
int ref_inverse_dct(int strategy, const float* coefficients, float* pixels) {
  const auto type = static_cast<gjxl::AcStrategyType>(strategy);
  const auto* info = gjxl::GetAcStrategyInfo(type);
  if (info == nullptr) return -1;
  const size_t n = info->coefficient_count();
  return Code(gjxl::InverseDctCpu(type, std::span<const float>(coefficients, n),
                                  std::span<float>(pixels, n)));
}

// src/codec/gaborish.cpp:49-99
int ref_gaborish_inverse(const float* input, size_t width, size_t height,
                         const float multipliers[3], float* output) {
  return Code(gjxl::ApplyGaborishInverse(
    ConstImage(input, width, height),
    {multipliers[0], multipliers[1], multipliers[2]},
    MutableImage(output, width, height)));
}

// src/codec/quantization.cpp:641-712 (one channel of one strategy)
int ref_quant_matrix(int strategy, int channel, float* dequant,
                     float* inverse_dequant, size_t* count) {
  gjxl::QuantizationMatrixView view;
  const gjxl::Status status = gjxl::GetDefaultQuantizationMatrix(
    static_cast<gjxl::AcStrategyType>(strategy),
    static_cast<gjxl::XybChannel>(channel), &view);
  if (!status.ok()) return Code(status);
  std::memcpy(dequant, view.dequant.data(), view.dequant.size() * sizeof(float));
  std::memcpy(inverse_dequant, view.inverse_dequant.data(),
              view.inverse_dequant.size() * sizeof(float));
  *count = view.dequant.size();
  return 0;
}

// src/codec/chroma_from_luma.cpp:434-531
int ref_initial_cfl(const float* opsin, size_t width, size_t height,
                    int8_t* y_to_x, int8_t* y_to_b) {
  gjxl::ColorCorrelationMap map;
  const gjxl::Status status = gjxl::ComputeInitialColorCorrelationMap(
    ConstImage(opsin, width, height), &map);
  if (!status.ok()) return Code(status);
  const auto ext = map.tile_extent();
  for (size_t y = 0; y < ext.height; ++y) {
    std::memcpy(y_to_x + y * ext.width, map.y_to_x_map().Row(y), ext.width);
    std::memcpy(y_to_b + y * ext.width, map.y_to_b_map().Row(y), ext.width);
  }
  return 0;
}

// src/codec/ac_strategy.cpp:823-994
int ref_ac_strategy_cost(int strategy, size_t block_x, size_t block_y,
                         const float* opsin, size_t width, size_t height,
                         const float* quant_field, const float* pixel_mask,
                         float butteraugli_target, float entropy_multiplier,
                         float cfl_x, float cfl_b, float* cost) {
  const gjxl::Extent2D blocks{width / 8, height / 8};
  gjxl::AcStrategyCostOptions options;
  options.butteraugli_target = butteraugli_target;
  options.entropy_multiplier = entropy_multiplier;
  options.cfl_factors = {cfl_x, 0.0f, cfl_b};
  return Code(gjxl::EstimateAcStrategyCost(
    static_cast<gjxl::AcStrategyType>(strategy), block_x, block_y,
    ConstImage(opsin, width, height),
    {quant_field, blocks, blocks.width},
    {pixel_mask, {width, height}, width}, options, cost));
}

// src/codec/ac_strategy.cpp:1101-1114, with the initial CfL map computed by
// src/codec/chroma_from_luma.cpp:434 from the same opsin image.  Writes one
// encoded cell per block: strategy << 1 | anchor.
int ref_find_ac_strategy_grid(const float* opsin, size_t width, size_t height,
                              const float* quant_field,
                              const float* pixel_mask,
                              float butteraugli_target, uint8_t* cells) {
  const gjxl::Extent2D blocks{width / 8, height / 8};
  gjxl::ColorCorrelationMap map;
  gjxl::Status status = gjxl::ComputeInitialColorCorrelationMap(
    ConstImage(opsin, width, height), &map);
  if (!status.ok()) return Code(status);
  gjxl::AcStrategyGrid grid;
  status = gjxl::FindAcStrategyGrid(
    ConstImage(opsin, width, height), {quant_field, blocks, blocks.width},
    {pixel_mask, {width, height}, width}, map,
    {.butteraugli_target = butteraugli_target, .dense_dct32_search = false},
    &grid);
  if (!status.ok()) return Code(status);
  for (size_t y = 0; y < blocks.height; ++y) {
    for (size_t x = 0; x < blocks.width; ++x) {
      gjxl::AcStrategyCell cell;
      status = grid.Get(x, y, &cell);
      if (!status.ok()) return Code(status);
      cells[y * blocks.width + x] = static_cast<uint8_t>(
        (static_cast<uint32_t>(cell.strategy) << 1) | (cell.is_anchor ? 1 : 0));
    }
  }
  return 0;
}

}  // extern "C"
