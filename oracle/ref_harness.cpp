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
#include <vector>

#include "c_api/image_conversion.h"
#include "codec/ac_strategy.h"
#include "codec/butteraugli.h"
#include "codec/butteraugli_distance_internal.h"
#include "codec/butteraugli_internal.h"
#include "codec/chroma_from_luma.h"
#include "codec/dc_conversion.h"
#include "codec/dc_quantization.h"
#include "codec/dc_smoothing.h"
#include "codec/dct.h"
#include "codec/epf.h"
#include "codec/loop_filter.h"
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

namespace {
gjxl::Quantizer MakeQuantizer(uint32_t global_scale, uint32_t quant_dc) {
  gjxl::Quantizer q;
  (void)gjxl::Quantizer::Create({global_scale, quant_dc}, &q);
  return q;
}

gjxl::Status GridFromCells(const uint8_t* cells, size_t bw, size_t bh,
                           gjxl::AcStrategyGrid* grid) {
  return gjxl::AcStrategyGrid::CreateFromEncodedCells(
    {bw, bh}, std::span<const uint8_t>(cells, bw * bh), grid);
}
}  // namespace

// src/codec/quantization.cpp:536-639; params[0] = global scale, [1] = quant dc
int ref_quantizer_from_field(float quant_dc, const float* field, size_t bw,
                             size_t bh, int32_t* raw, uint32_t params[2]) {
  gjxl::Quantizer q;
  const gjxl::Status status = gjxl::CreateQuantizerFromField(
    quant_dc, {field, {bw, bh}, bw}, {raw, {bw, bh}, bw}, &q);
  if (!status.ok()) return Code(status);
  params[0] = q.params().global_scale;
  params[1] = q.params().quant_dc;
  return 0;
}

// src/codec/chroma_from_luma.cpp:543-670 (fast or Newton final CfL)
int ref_final_cfl(const float* opsin, size_t width, size_t height,
                  const uint8_t* cells, const int32_t* raw,
                  uint32_t global_scale, uint32_t quant_dc, int fast,
                  int8_t* y_to_x, int8_t* y_to_b) {
  const size_t bw = width / 8, bh = height / 8;
  gjxl::AcStrategyGrid grid;
  gjxl::Status status = GridFromCells(cells, bw, bh, &grid);
  if (!status.ok()) return Code(status);
  gjxl::ColorCorrelationMap map;
  status = gjxl::ComputeFinalColorCorrelationMap(
    ConstImage(opsin, width, height), grid, {raw, {bw, bh}, bw},
    MakeQuantizer(global_scale, quant_dc), fast != 0, &map, 20);
  if (!status.ok()) return Code(status);
  const auto ext = map.tile_extent();
  for (size_t y = 0; y < ext.height; ++y) {
    std::memcpy(y_to_x + y * ext.width, map.y_to_x_map().Row(y), ext.width);
    std::memcpy(y_to_b + y * ext.width, map.y_to_b_map().Row(y), ext.width);
  }
  return 0;
}

// src/codec/quantization.cpp:741-806 then the Y/X/B quantisation of one
// transform as ComputeQuantizedCoefficients performs it
// (src/codec/reconstruction.cpp:565-700, without the DC parts).
int ref_code_transform(int strategy, uint32_t global_scale, uint32_t quant_dc,
                       int32_t raw_quant, float x_mult, float b_mult,
                       float cfl_x, float cfl_b, float* coefficients /*3*n*/,
                       int32_t* quantized /*3*n*/, int32_t* decided_raw) {
  const auto type = static_cast<gjxl::AcStrategyType>(strategy);
  const size_t n = gjxl::GetAcStrategyInfo(type)->coefficient_count();
  const gjxl::Quantizer q = MakeQuantizer(global_scale, quant_dc);
  gjxl::AdjustedAcQuantization decision;
  gjxl::Status status = gjxl::SelectAdjustedAcQuantization(
    type, q, raw_quant, {x_mult, 1.0f, b_mult},
    {std::span<const float>(coefficients, n),
     std::span<const float>(coefficients + n, n),
     std::span<const float>(coefficients + 2 * n, n)}, &decision);
  if (!status.ok()) return Code(status);
  *decided_raw = decision.raw_quant;
  status = gjxl::QuantizeAdjustedYAcBlock(
    type, q, decision, std::span<const float>(coefficients + n, n),
    std::span<int32_t>(quantized + n, n));
  if (!status.ok()) return Code(status);
  status = gjxl::DequantizeAcBlock(
    type, q, decision.raw_quant, {.channel = gjxl::XybChannel::kY},
    std::span<const int32_t>(quantized + n, n),
    std::span<float>(coefficients + n, n));
  if (!status.ok()) return Code(status);
  for (size_t i = 0; i < n; ++i) {
    coefficients[i] -= cfl_x * coefficients[n + i];
    coefficients[2 * n + i] -= cfl_b * coefficients[n + i];
  }
  for (int c : {0, 2}) {
    status = gjxl::QuantizeAcBlock(
      type, q, decision.raw_quant,
      {.channel = static_cast<gjxl::XybChannel>(c),
       .matrix_multiplier = c == 0 ? x_mult : b_mult},
      std::span<const float>(coefficients + c * n, n),
      std::span<int32_t>(quantized + c * n, n));
    if (!status.ok()) return Code(status);
  }
  return 0;
}

// src/codec/quantization.cpp:808-835
int ref_dequantize(int strategy, uint32_t global_scale, uint32_t quant_dc,
                   int32_t raw_quant, int channel, float mult,
                   const int32_t* quantized, float* out) {
  const auto type = static_cast<gjxl::AcStrategyType>(strategy);
  const size_t n = gjxl::GetAcStrategyInfo(type)->coefficient_count();
  return Code(gjxl::DequantizeAcBlock(
    type, MakeQuantizer(global_scale, quant_dc), raw_quant,
    {.channel = static_cast<gjxl::XybChannel>(channel), .matrix_multiplier = mult},
    std::span<const int32_t>(quantized, n), std::span<float>(out, n)));
}

// src/codec/dc_conversion.cpp
int ref_llf_to_dc(int strategy, const float* coefficients, float* dc) {
  const auto type = static_cast<gjxl::AcStrategyType>(strategy);
  const auto* info = gjxl::GetAcStrategyInfo(type);
  return Code(gjxl::ConvertLowFrequenciesToDc(
    type, std::span<const float>(coefficients, info->coefficient_count()),
    {dc, info->covered_blocks, info->covered_blocks.width}));
}

int ref_dc_to_llf(int strategy, const float* dc, float* coefficients) {
  const auto type = static_cast<gjxl::AcStrategyType>(strategy);
  const auto* info = gjxl::GetAcStrategyInfo(type);
  return Code(gjxl::ConvertDcToLowFrequencies(
    type, {dc, info->covered_blocks, info->covered_blocks.width},
    std::span<float>(coefficients, info->coefficient_count())));
}

// src/codec/vardct/dc_quantization.cpp (prediction-aware, weighted, +1 bit)
int ref_quantize_dc(const float* dc, size_t bw, size_t bh,
                    uint32_t global_scale, uint32_t quant_dc,
                    int32_t* quantized, float* reconstructed) {
  const size_t n = bw * bh;
  gjxl::ConstImage3FView in;
  gjxl::Image3I32View q;
  gjxl::Image3FView r;
  for (size_t c = 0; c < 3; ++c) {
    in.plane[c] = {dc + c * n, {bw, bh}, bw};
    q.plane[c] = {quantized + c * n, {bw, bh}, bw};
    r.plane[c] = {reconstructed + c * n, {bw, bh}, bw};
  }
  return Code(gjxl::QuantizeDcCoefficients(
    in, MakeQuantizer(global_scale, quant_dc), {q, r},
    {gjxl::DcQuantizationMode::kPredictionAware,
     gjxl::VarDctDcPrediction::kWeighted, 1}));
}

// src/codec/dc_smoothing.cpp
int ref_smooth_dc(const float* dc, size_t bw, size_t bh, uint32_t global_scale,
                  uint32_t quant_dc, float* out) {
  return Code(gjxl::SmoothDcCoefficients(
    ConstImage(dc, bw, bh), MakeQuantizer(global_scale, quant_dc),
    MutableImage(out, bw, bh)));
}

// src/codec/epf.cpp:207-290 with neutral sharpness
int ref_epf_inverse_sigma(const uint8_t* cells, size_t bw, size_t bh,
                          const int32_t* raw, uint32_t global_scale,
                          uint32_t quant_dc, float* out) {
  gjxl::AcStrategyGrid grid;
  gjxl::Status status = GridFromCells(cells, bw, bh, &grid);
  if (!status.ok()) return Code(status);
  std::vector<uint8_t> sharp(bw * bh, 4);
  return Code(gjxl::ComputeEpfInverseSigma(
    grid, {raw, {bw, bh}, bw}, MakeQuantizer(global_scale, quant_dc),
    {sharp.data(), {bw, bh}, bw}, {}, {out, {bw, bh}, bw}));
}

// src/codec/loop_filter.cpp:14-65 with the default options
int ref_loop_filters(const float* in, size_t width, size_t height,
                     const float* inverse_sigma, float* out) {
  const size_t bw = (width + 7) / 8, bh = (height + 7) / 8;
  return Code(gjxl::ApplyLoopFilters(
    ConstImage(in, width, height), {inverse_sigma, {bw, bh}, bw}, {},
    MutableImage(out, width, height)));
}

namespace {
gjxl::ButteraugliOptions BOptions(float hf_asymmetry, float x_multiplier,
                                  float intensity) {
  return {.hf_asymmetry = hf_asymmetry, .x_multiplier = x_multiplier,
          .intensity_target = intensity};
}
namespace bi = gjxl::butteraugli_internal;
}  // namespace

// src/codec/butteraugli.cpp:98-115
int ref_butteraugli_distance(const float* reference, const float* distorted,
                             size_t width, size_t height, float hf_asymmetry,
                             float x_multiplier, float intensity, float* map,
                             double* score) {
  return Code(gjxl::ComputeButteraugliDistance(
    ConstImage(reference, width, height), ConstImage(distorted, width, height),
    BOptions(hf_asymmetry, x_multiplier, intensity),
    {map, {width, height}, width}, score));
}

// src/codec/butteraugli.cpp:31-71: one Prepare, then `count` Compares of the
// distorted images (count x 3 planes), maps written one after another.
int ref_butteraugli_prepared(const float* reference, const float* distorted,
                             size_t count, size_t width, size_t height,
                             float hf_asymmetry, float x_multiplier,
                             float intensity, float* maps, double* scores) {
  gjxl::PreparedButteraugliReference prepared;
  gjxl::Status status = prepared.Prepare(
    ConstImage(reference, width, height),
    BOptions(hf_asymmetry, x_multiplier, intensity));
  if (!status.ok()) return Code(status);
  const size_t n = width * height;
  for (size_t i = 0; i < count; ++i) {
    status = prepared.Compare(ConstImage(distorted + i * 3 * n, width, height),
                              {maps + i * n, {width, height}, width},
                              scores + i);
    if (!status.ok()) return Code(status);
  }
  return 0;
}

// src/codec/butteraugli_internal.cpp:576-620
int ref_butteraugli_blur(const float* in, size_t width, size_t height,
                         float sigma, float* out) {
  bi::BlurScratch scratch;
  return Code(bi::GaussianBlur({in, {width, height}, width}, sigma, &scratch,
                               {out, {width, height}, width}));
}

// :622-700
int ref_butteraugli_opsin(const float* rgb, size_t width, size_t height,
                          float intensity, float* xyb) {
  bi::OpsinScratch scratch;
  return Code(bi::OpsinDynamicsImage(ConstImage(rgb, width, height), intensity,
                                     &scratch, MutableImage(xyb, width, height)));
}

namespace {
void CopyPsycho(const bi::OwnedPsychoImage& p, size_t n, float* out) {
  const gjxl::ConstPlaneF32View planes[10] = {
    p.LowFrequencyView().plane[0], p.LowFrequencyView().plane[1],
    p.LowFrequencyView().plane[2], p.MediumFrequencyView().plane[0],
    p.MediumFrequencyView().plane[1], p.MediumFrequencyView().plane[2],
    p.HighFrequencyView(0), p.HighFrequencyView(1),
    p.UltraHighFrequencyView(0), p.UltraHighFrequencyView(1)};
  for (size_t k = 0; k < 10; ++k) std::memcpy(out + k * n, planes[k].data, n * 4);
}
}  // namespace

// :702-846; psycho: LF(3), MF(3), HF(2), UHF(2)
int ref_butteraugli_frequencies(const float* xyb, size_t width, size_t height,
                                float* psycho) {
  bi::FrequencyScratch scratch;
  bi::OwnedPsychoImage out;
  gjxl::Status status =
    bi::SeparateFrequencies(ConstImage(xyb, width, height), &scratch, &out);
  if (!status.ok()) return Code(status);
  CopyPsycho(out, width * height, psycho);
  return 0;
}

// src/codec/butteraugli_distance_internal.cpp:657-870 for two linear images:
// both pass through opsin dynamics and frequency separation first.
// stages: 9 planes (DifferenceStage order).
int ref_butteraugli_stages(const float* reference, const float* distorted,
                           size_t width, size_t height, float hf_asymmetry,
                           float x_multiplier, float intensity, float* stages) {
  const size_t n = width * height;
  std::vector<float> x0(3 * n), x1(3 * n);
  bi::OpsinScratch opsin;
  gjxl::Status status = bi::OpsinDynamicsImage(
    ConstImage(reference, width, height), intensity, &opsin,
    MutableImage(x0.data(), width, height));
  if (!status.ok()) return Code(status);
  status = bi::OpsinDynamicsImage(ConstImage(distorted, width, height),
                                  intensity, &opsin,
                                  MutableImage(x1.data(), width, height));
  if (!status.ok()) return Code(status);
  bi::FrequencyScratch frequency;
  bi::OwnedPsychoImage p0, p1;
  status = bi::SeparateFrequencies(ConstImage(x0.data(), width, height),
                                   &frequency, &p0);
  if (!status.ok()) return Code(status);
  status = bi::SeparateFrequencies(ConstImage(x1.data(), width, height),
                                   &frequency, &p1);
  if (!status.ok()) return Code(status);
  bi::DifferenceScratch scratch;
  bi::OwnedDifferenceStages out;
  status = bi::ComputeDifferenceStages(
    p0, p1, {.hf_asymmetry = hf_asymmetry, .x_multiplier = x_multiplier,
             .intensity_target = intensity}, &scratch, &out);
  if (!status.ok()) return Code(status);
  for (size_t k = 0; k < bi::kDifferenceStageCount; ++k) {
    std::memcpy(stages + k * n,
                out.StageView(static_cast<bi::DifferenceStage>(k)).data, n * 4);
  }
  return 0;
}

// src/codec/butteraugli.cpp:116-222
int ref_reduce_distance_map(const float* map, size_t width, size_t height,
                            const uint8_t* cells, size_t bw, size_t bh,
                            float* out) {
  gjxl::AcStrategyGrid grid;
  gjxl::Status status = GridFromCells(cells, bw, bh, &grid);
  if (!status.ok()) return Code(status);
  return Code(gjxl::ReduceButteraugliDistanceMap(
    {map, {width, height}, width}, grid, {out, {bw, bh}, bw}));
}

}  // extern "C"
