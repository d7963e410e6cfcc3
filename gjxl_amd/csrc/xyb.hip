// a1 + a2: packed sRGB8 -> linear RGB planes -> padded XYB planes.
//
// One thread per padded output pixel; coordinates beyond the source are
// clamped to the last row/column, which reproduces the reference's edge
// replication of transformed values (src/codec/color_transform.cpp:343-351,
// :413-421) because the transform is a pure function of the source pixel.
// HBM-bound: 3 (or 4) bytes in, 12 B linear + 12 B XYB out per pixel.
#include <cmath>

#include "common.h"

namespace gjxl_amd {
namespace {

constexpr float kOpsinBias = 0.0037930732552754493f;

struct OpsinMatrix {
  float m[3][3];
};

// src/codec/color_transform.cpp:35-40.  The kernel multiplies by the
// intensity scale first, as the reference does (scale * matrix entry).
constexpr OpsinMatrix kOpsinMatrix = {{
  {0.30f, 0.622f, 0.078f},
  {0.23f, 0.692f, 0.078f},
  {0.24342268924547819f, 0.20476744424496821f, 0.55180986650955360f},
}};

struct SrgbTable {
  float value[256];
};

// src/c_api/image_conversion.cpp:18-31: the sRGB EOTF in binary64, rounded
// once to binary32 (host libm, evaluated once per process).
const SrgbTable& HostSrgbTable() {
  static const SrgbTable table = [] {
    SrgbTable t{};
    for (int i = 0; i < 256; ++i) {
      const double encoded = static_cast<double>(i) / 255.0;
      const double linear = encoded <= 0.04045
        ? encoded / 12.92 : std::pow((encoded + 0.055) / 1.055, 2.4);
      t.value[i] = static_cast<float>(linear);
    }
    return t;
  }();
  return table;
}

__device__ __forceinline__ void LinearToOpsin(float r, float g, float b,
                                              float scale, float bias_cbrt,
                                              float* x, float* y, float* z) {
  float gamma[3];
#pragma unroll
  for (int row = 0; row < 3; ++row) {
    float v = __builtin_fmaf(scale * kOpsinMatrix.m[row][2], b, kOpsinBias);
    v = __builtin_fmaf(scale * kOpsinMatrix.m[row][1], g, v);
    v = __builtin_fmaf(scale * kOpsinMatrix.m[row][0], r, v);
    gamma[row] = FastCubeRootAndAdd(v > 0.0f ? v : 0.0f, -bias_cbrt);
  }
  *x = 0.5f * (gamma[0] - gamma[1]);
  *y = 0.5f * (gamma[0] + gamma[1]);
  *z = gamma[2];
}

constexpr int kTileX = 64;
constexpr int kTileY = 4;

__global__ __launch_bounds__(kTileX * kTileY) void SrgbToLinearOpsinKernel(
  const uint8_t* __restrict__ srgb, size_t row_stride, uint32_t bpp,
  uint32_t width, uint32_t height, uint32_t pw, uint32_t ph, float scale,
  SrgbTable table, float* __restrict__ linear, float* __restrict__ opsin) {
  __shared__ float lut[256];
  const int tid = threadIdx.y * kTileX + threadIdx.x;
  lut[tid] = table.value[tid];
  __syncthreads();
  const uint32_t x = blockIdx.x * kTileX + threadIdx.x;
  const uint32_t y = blockIdx.y * kTileY + threadIdx.y;
  if (x >= pw || y >= ph) return;
  const uint32_t sx = x < width ? x : width - 1;
  const uint32_t sy = y < height ? y : height - 1;
  const uint8_t* p = srgb + static_cast<size_t>(sy) * row_stride +
                     static_cast<size_t>(sx) * bpp;
  const float r = lut[p[0]], g = lut[p[1]], b = lut[p[2]];
  if (x < width && y < height) {
    const size_t plane = static_cast<size_t>(width) * height;
    const size_t i = static_cast<size_t>(y) * width + x;
    linear[i] = r;
    linear[plane + i] = g;
    linear[2 * plane + i] = b;
  }
  const float bias_cbrt = FastCubeRootAndAdd(kOpsinBias, 0.0f);
  float ox, oy, ob;
  LinearToOpsin(r, g, b, scale, bias_cbrt, &ox, &oy, &ob);
  const size_t oplane = static_cast<size_t>(pw) * ph;
  const size_t o = static_cast<size_t>(y) * pw + x;
  opsin[o] = ox;
  opsin[oplane + o] = oy;
  opsin[2 * oplane + o] = ob;
}

__global__ __launch_bounds__(kTileX * kTileY) void LinearToOpsinKernel(
  const float* __restrict__ linear, uint32_t width, uint32_t height,
  uint32_t pw, uint32_t ph, float scale, float* __restrict__ opsin) {
  const uint32_t x = blockIdx.x * kTileX + threadIdx.x;
  const uint32_t y = blockIdx.y * kTileY + threadIdx.y;
  if (x >= pw || y >= ph) return;
  const uint32_t sx = x < width ? x : width - 1;
  const uint32_t sy = y < height ? y : height - 1;
  const size_t plane = static_cast<size_t>(width) * height;
  const size_t i = static_cast<size_t>(sy) * width + sx;
  const float bias_cbrt = FastCubeRootAndAdd(kOpsinBias, 0.0f);
  float ox, oy, ob;
  LinearToOpsin(linear[i], linear[plane + i], linear[2 * plane + i], scale,
                bias_cbrt, &ox, &oy, &ob);
  const size_t oplane = static_cast<size_t>(pw) * ph;
  const size_t o = static_cast<size_t>(y) * pw + x;
  opsin[o] = ox;
  opsin[oplane + o] = oy;
  opsin[2 * oplane + o] = ob;
}

GJXLResult ValidateIntensity(float intensity_target) {
  if (!std::isfinite(intensity_target) || intensity_target <= 0.0f) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                "Intensity target must be finite and positive");
  }
  return GJXL_OK;
}

}  // namespace

GJXLResult LaunchSrgbToLinearOpsin(GJXLAmdDevice* device, const uint8_t* srgb,
                                   size_t row_stride, uint32_t bpp,
                                   uint32_t width, uint32_t height,
                                   float intensity_target, float* linear,
                                   float* opsin) {
  const uint32_t pw = PadTo8(width), ph = PadTo8(height);
  const dim3 grid(CeilDiv(pw, kTileX), CeilDiv(ph, kTileY));
  hipLaunchKernelGGL(SrgbToLinearOpsinKernel, grid, dim3(kTileX, kTileY), 0,
                     device->stream, srgb, row_stride, bpp, width, height, pw,
                     ph, intensity_target / 255.0f, HostSrgbTable(), linear,
                     opsin);
  return CheckLaunch("SrgbToLinearOpsinKernel");
}

}  // namespace gjxl_amd

using gjxl_amd::Fail;

extern "C" {

GJXLResult gjxl_amd_srgb8_to_linear_and_opsin(
  GJXLAmdDevice* device, const uint8_t* srgb, size_t row_stride_bytes,
  uint32_t bytes_per_pixel, uint32_t width, uint32_t height,
  float intensity_target, float* linear, float* opsin) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  GJXL_AMD_TRY(gjxl_amd::ValidateIntensity(intensity_target));
  if (srgb == nullptr || linear == nullptr || opsin == nullptr ||
      width == 0 || height == 0 ||
      (bytes_per_pixel != 3 && bytes_per_pixel != 4) ||
      row_stride_bytes < static_cast<size_t>(width) * bytes_per_pixel ||
      width > (1u << 30) || height > (1u << 30)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                "sRGB input geometry or buffers are invalid");
  }
  return gjxl_amd::LaunchSrgbToLinearOpsin(device, srgb, row_stride_bytes,
                                           bytes_per_pixel, width, height,
                                           intensity_target, linear, opsin);
}

GJXLResult gjxl_amd_linear_rgb_to_padded_opsin(
  GJXLAmdDevice* device, const float* linear, uint32_t width,
  uint32_t height, float intensity_target, float* opsin) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  GJXL_AMD_TRY(gjxl_amd::ValidateIntensity(intensity_target));
  if (linear == nullptr || opsin == nullptr || width == 0 || height == 0 ||
      width > (1u << 30) || height > (1u << 30)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                "Linear RGB geometry or buffers are invalid");
  }
  const uint32_t pw = gjxl_amd::PadTo8(width), ph = gjxl_amd::PadTo8(height);
  const dim3 grid(gjxl_amd::CeilDiv(pw, gjxl_amd::kTileX),
                  gjxl_amd::CeilDiv(ph, gjxl_amd::kTileY));
  hipLaunchKernelGGL(gjxl_amd::LinearToOpsinKernel, grid,
                     dim3(gjxl_amd::kTileX, gjxl_amd::kTileY), 0,
                     device->stream, linear, width, height, pw, ph,
                     intensity_target / 255.0f, opsin);
  return gjxl_amd::CheckLaunch("LinearToOpsinKernel");
}

}  // extern "C"
