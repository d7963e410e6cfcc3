// Device Butteraugli: a prepared reference image (its psycho-visual planes
// and its masking planes at two scales) compared against distorted images,
// all in HBM.  The arithmetic of every kernel follows the reference's native
// scalar Butteraugli (src/codec/butteraugli_internal.cpp,
// src/codec/butteraugli_distance_internal.cpp) output by output, so that the
// maps are bit-identical with the CPU.
#pragma once

#include "common.h"

namespace gjxl_amd {

// Planes of one scale.  The psycho image is 10 planes: LF X,Y,B; MF X,Y,B;
// HF X,Y; UHF X,Y.  The 9 stage planes are the Malta maps (MF Y, MF X, HF Y,
// HF X, UHF Y, UHF X), the reference's mask, the masked AC Y and the final
// composition, in the reference's DifferenceStage order.
struct ButteraugliScale {
  int width = 0, height = 0;
  size_t n = 0;
  float* input = nullptr;        // 3 planes: linear RGB at this scale
  float* xyb = nullptr;          // 3 planes
  float* ref_psycho = nullptr;   // 10 planes
  float* dist_psycho = nullptr;  // 10 planes
  float* ref_blurred = nullptr;  // blurred masking activity of the reference
  float* mask_y = nullptr;       // MaskY(mask) per pixel
  float* mask_dc = nullptr;      // MaskDcY(mask) per pixel
  float* stages = nullptr;       // 9 planes
  float* tmp[3] = {nullptr, nullptr, nullptr};
};

enum ButteraugliSigma { kSigmaOpsin, kSigmaLow, kSigmaHigh, kSigmaUltra, kSigmaMask, kSigmaCount };

struct ButteraugliParams {
  float hf_asymmetry = 1.0f;
  float x_multiplier = 1.0f;
  float intensity_target = 80.0f;
};

}  // namespace gjxl_amd

struct GJXLAmdButteraugli {
  GJXLAmdDevice* device = nullptr;
  gjxl_amd::ButteraugliParams params;
  int width = 0, height = 0;   // requested extent
  bool expanded = false;       // extent below 8 in either axis
  bool has_subscale = false;
  int xborder = 0, yborder = 0;
  bool prepared = false;
  gjxl_amd::ButteraugliScale scale[2];
  float* taps = nullptr;       // kSigmaCount x 33 unscaled kernel taps
  float* map = nullptr;        // requested extent
  uint32_t* score_bits = nullptr;  // max of the map, as float bits
  void* allocation = nullptr;  // one block for everything above
  size_t capacity_floats = 0;  // its size
};

namespace gjxl_amd {

// Enqueues a comparison of `distorted` (3 planes, requested extent, device)
// against the prepared reference: the map goes to b->map and `map_out` when
// it is not null, the score to b->score_bits.  keep_stages also leaves the
// intermediate difference maps (the STAGES buffer) for inspection.
GJXLResult ButteraugliCompare(GJXLAmdButteraugli* b, const float* distorted,
                              float* map_out, bool keep_stages = false);

// Extent fields for (width, height); returns the floats the allocation
// needs at that extent (it grows with the width and with the height).
size_t ButteraugliGeometry(GJXLAmdButteraugli* b, uint32_t width, uint32_t height);
// Points the planes of the current extent into the allocation.
void CarveButteraugli(GJXLAmdButteraugli* b);
// Re-targets the object to an extent whose planes fit its allocation.
GJXLResult ResizeButteraugli(GJXLAmdButteraugli* b, uint32_t width, uint32_t height);

// a17: per-transform 1.2 * (mean d^16)^(1/16) broadcast over the covered
// blocks.  map: width x height (unpadded); cells: (pw/8) x (ph/8).
GJXLResult LaunchReduceDistanceMap(hipStream_t s, const float* map, int width,
                                   int height, const uint8_t* cells, int bw,
                                   int bh, float* out);

}  // namespace gjxl_amd
