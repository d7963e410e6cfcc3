// Read-only tables of the path, generated once per device handle:
//   * libjxl's default AC quantization matrices for the five DCT families the
//     search uses (dequant and inverse-dequant, three channels each);
//   * the binary64 DCT-II / DCT-III bases for lengths 8, 16 and 32.
// Both are computed on the host exactly as the reference computes them and
// uploaded to one device buffer per handle.
#pragma once

#include <cstddef>
#include <cstdint>

namespace gjxl_amd {

// Strategy numbering of the JPEG XL format (src/core/ac_strategy.h:25-52).
enum Strategy : int {
  kDct8 = 0,
  kDct16x16 = 4,
  kDct32x32 = 5,
  kDct16x8 = 6,
  kDct8x16 = 7,
  kDct32x16 = 10,
  kDct16x32 = 11,
};

// Quantization-table families (rows = smaller dimension, cols = larger).
constexpr int kFamilyCount = 5;
constexpr int kFamilyRows[kFamilyCount] = {8, 16, 32, 8, 16};
constexpr int kFamilyCols[kFamilyCount] = {8, 16, 32, 16, 32};
constexpr int kFamilySize[kFamilyCount] = {64, 256, 1024, 128, 512};
constexpr int kFamilyOffset[kFamilyCount] = {0, 64, 320, 1344, 1472};
constexpr int kFamilyTotal = 1984;  // sum of sizes
// dequant[f][c] at (kFamilyOffset[f] * 3 + c * kFamilySize[f]); the inverse
// tables follow all dequant tables.
constexpr int kQuantTableFloats = 2 * 3 * kFamilyTotal;

__host__ __device__ constexpr int FamilyOf(int strategy) {
  return strategy == kDct8 ? 0
       : strategy == kDct16x16 ? 1
       : strategy == kDct32x32 ? 2
       : (strategy == kDct16x8 || strategy == kDct8x16) ? 3
       : 4;
}

__host__ __device__ constexpr int DequantOffset(int strategy, int channel) {
  return kFamilyOffset[FamilyOf(strategy)] * 3 +
         channel * kFamilySize[FamilyOf(strategy)];
}

__host__ __device__ constexpr int InverseOffset(int strategy, int channel) {
  return 3 * kFamilyTotal + DequantOffset(strategy, channel);
}

// DCT bases: forward 8, 16, 32 then inverse 8, 16, 32; basis[f * n + s].
__host__ __device__ constexpr int BasisOffset(int n, bool forward) {
  const int base = forward ? 0 : 1344;
  return base + (n == 8 ? 0 : n == 16 ? 64 : 320);
}
// Transposed forward bases (basis[s * n + f]) follow: 8, 16, 32.
__host__ __device__ constexpr int BasisOffsetT(int n) {
  return 2688 + (n == 8 ? 0 : n == 16 ? 64 : 320);
}
constexpr int kBasisDoubles = 4032;

// Covered 8x8 blocks of the searched strategies.
__host__ __device__ constexpr int CoveredW(int strategy) {
  return strategy == kDct8 ? 1 : strategy == kDct16x16 ? 2
       : strategy == kDct32x32 ? 4 : strategy == kDct16x8 ? 1
       : strategy == kDct8x16 ? 2 : strategy == kDct32x16 ? 2 : 4;
}
__host__ __device__ constexpr int CoveredH(int strategy) {
  return strategy == kDct8 ? 1 : strategy == kDct16x16 ? 2
       : strategy == kDct32x32 ? 4 : strategy == kDct16x8 ? 2
       : strategy == kDct8x16 ? 1 : strategy == kDct32x16 ? 4 : 2;
}

// Host: fills the tables (defined in tables.hip).
void BuildQuantTables(float* out /* kQuantTableFloats */);
void BuildDctBases(double* out /* kBasisDoubles */);

}  // namespace gjxl_amd
