// a4 Gaborish inverse and a5 initial quant field (with its masks).
//
// Symmetric5Kernel: the reference's mirrored 5x5 symmetric convolution
// (src/codec/convolution.cpp:38-119) with its exact summation order, staged
// through an LDS tile of 64x16 outputs plus a 2-pixel halo.  Used for the
// Gaborish inverse (src/codec/gaborish.cpp:49-99) and for blurring the pixel
// mask (src/codec/adaptive_quantization.cpp:263-292).
//
// The initial quant field (src/codec/adaptive_quantization.cpp:583-738):
//   PixelMaskKernel    per-pixel gamma-weighted Laplacian -> raw pixel mask
//                      (fdlibm log1pf) and the 4x4 pre-erosion means;
//   BlockFieldKernel   fuzzy erosion of the pre-erosion map, strategy mask
//                      and the per-block gamma / high-frequency / blue
//                      modulations -> quant field;
//   Symmetric5Kernel   blur of the raw pixel mask.
#include <cmath>

#include "common.h"
#include "libm_port.h"

namespace gjxl_amd {
namespace {

struct Symmetric5Weights {
  float d0, d1, d2, d4, d8, d5;
};

struct Symmetric5WeightSet {
  Symmetric5Weights plane[3];
};

// src/core/image_ops.h:20-47 for any integer coordinate.
__device__ __forceinline__ int Mirror(int c, int size) {
  if (c >= 0 && c < size) return c;
  const long long period = 2LL * size;
  long long phase;
  if (c >= 0) {
    phase = c % period;
  } else {
    const long long mag = -(static_cast<long long>(c) + 1) + 1;
    const long long rem = mag % period;
    phase = rem == 0 ? 0 : period - rem;
  }
  return static_cast<int>(phase < size ? phase : period - 1 - phase);
}

constexpr int kConvX = 64;             // outputs per tile row
constexpr int kConvY = 16;             // output rows per tile
constexpr int kConvRowsPerThread = 4;  // 64 x 4 threads
constexpr int kHaloW = kConvX + 4;
constexpr int kHaloH = kConvY + 4;

// One tile of one plane; blockIdx.z selects the plane and its weights.
__global__ __launch_bounds__(256) void Symmetric5Kernel(
  const float* __restrict__ input, float* __restrict__ output, int width,
  int height, Symmetric5WeightSet weights) {
  __shared__ float tile[kHaloH][kHaloW + 1];
  const size_t plane_offset =
    static_cast<size_t>(blockIdx.z) * width * height;
  const float* in = input + plane_offset;
  const int x0 = blockIdx.x * kConvX;
  const int y0 = blockIdx.y * kConvY;
  const int tid = threadIdx.y * kConvX + threadIdx.x;
  for (int i = tid; i < kHaloH * kHaloW; i += 256) {
    const int ty = i / kHaloW, tx = i % kHaloW;
    const int gx = Mirror(x0 + tx - 2, width);
    const int gy = Mirror(y0 + ty - 2, height);
    tile[ty][tx] = in[static_cast<size_t>(gy) * width + gx];
  }
  __syncthreads();
  const Symmetric5Weights w = weights.plane[blockIdx.z];
  const int x = x0 + threadIdx.x;
  const int lx = threadIdx.x + 2;
#pragma unroll
  for (int r = 0; r < kConvRowsPerThread; ++r) {
    const int ly = threadIdx.y * kConvRowsPerThread + r;
    const int y = y0 + ly;
    // WeightedRow (src/codec/convolution.cpp:16-35) on tile row ty.
    auto row_sum = [&](int ty, float center, float near, float far) {
      const float* t = tile[ty];
      const float f = far * (t[lx - 2] + t[lx + 2]);
      const float n = near * (t[lx - 1] + t[lx + 1]);
      const float c = center * t[lx];
      return f + (n + c);
    };
    const int ty = ly + 2;
    float sum0 = row_sum(ty, w.d0, w.d1, w.d2);
    sum0 += row_sum(ty - 2, w.d2, w.d5, w.d8);
    float sum1 = row_sum(ty + 2, w.d2, w.d5, w.d8);
    sum0 += row_sum(ty - 1, w.d1, w.d4, w.d5);
    sum1 += row_sum(ty + 1, w.d1, w.d4, w.d5);
    if (x < width && y < height) {
      output[plane_offset + static_cast<size_t>(y) * width + x] = sum0 + sum1;
    }
  }
}

GJXLResult LaunchSymmetric5(hipStream_t stream, const float* in, float* out,
                            int width, int height, int planes,
                            const Symmetric5WeightSet& weights) {
  const dim3 grid(CeilDiv(width, kConvX), CeilDiv(height, kConvY), planes);
  hipLaunchKernelGGL(Symmetric5Kernel, grid,
                     dim3(kConvX, 256 / kConvX), 0, stream, in, out, width,
                     height, weights);
  return CheckLaunch("Symmetric5Kernel");
}

// src/codec/gaborish.cpp:22-47 (host; the float partial sum is promoted to
// binary64 exactly where the reference's expression promotes it).
Symmetric5Weights GaborishInverseWeights(float multiplier) {
  const float k0 = -0.09495815671340026f, k1 = -0.041031725066768575f,
              k2 = 0.013710004822696948f, k3 = 0.006510206083837737f,
              k4 = -0.0014789063378272242f;
  const float partial = k0 + k1 + k2 + k4;
  double sum = 1.0 + static_cast<double>(multiplier) * 4.0 *
                       (static_cast<double>(partial) + 2.0 * static_cast<double>(k3));
  if (sum < 1.0e-5) sum = 1.0e-5;
  const float normalize = static_cast<float>(1.0 / sum);
  const float nm = multiplier * normalize;
  return {normalize, nm * k0, nm * k2, nm * k1, nm * k4, nm * k3};
}

// src/codec/adaptive_quantization.cpp:263-292
Symmetric5Weights PixelMaskBlurWeights() {
  const float f0 = 0.364911248f, f1 = 0.05f, f2 = 0.1688888021f,
              f3 = 0.221069183f, f4 = 0.306563504f;
  const float partial = f0 + f1 + f2 + f4;
  double ws = 1.0 + 4.0 * (static_cast<double>(partial) + 2.0 * static_cast<double>(f3));
  if (ws < 1.0e-5) ws = 1.0e-5;
  const float n = static_cast<float>(1.0 / ws);
  return {n, n * f0, n * f2, n * f1, n * f4, n * f3};
}

// ---- initial quant field ---------------------------------------------------

__device__ __forceinline__ float StdMax(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float StdMin(float a, float b) { return b < a ? b : a; }

constexpr float kInverseLog2E = 0.6931471805599453f;

// src/codec/adaptive_quantization.cpp:117-143
template <bool kInvert>
__device__ __forceinline__ float GammaDerivativeRatio(float value) {
  constexpr float kEpsilon = 1.0e-2f;
  constexpr float kSgMul = 226.77216153508914f;
  constexpr float kSgMul2 = 1.0f / 73.377132366608819f;
  constexpr float kSgReturnMul = kSgMul2 * 18.6580932135f * kInverseLog2E;
  constexpr float kSgOffset = 7.7825991679894591f;
  value = StdMax(value, 0.0f);
  const float sq = value * value;
  const float num = __builtin_fmaf(kSgReturnMul * 3.0f * kSgMul, sq, kEpsilon);
  const float den = __builtin_fmaf(kInverseLog2E * kSgMul * value, sq,
                                   kSgOffset * kInverseLog2E + kEpsilon);
  if constexpr (kInvert) return num / den;
  return den / num;
}

// :164-170
__device__ __forceinline__ float MaskingSqrt(float value) {
  constexpr float kLogOffset = 27.505837037000106f;
  constexpr float kMul = 211.66567973503678f;
  const float inner = sqrtf(kMul * 1.0e8f);
  return 0.25f * sqrtf(__builtin_fmaf(value, inner, kLogOffset));
}

// :145-162
__device__ __forceinline__ float ComputeMask(float value) {
  constexpr float kBase = -0.7647f, kMul4 = 9.4708735624378946f,
                  kMul2 = 17.35036561631863f, kOffset2 = 302.59587815579727f,
                  kMul3 = 6.7943250517376494f,
                  kOffset3 = 3.7179635626140772f,
                  kMul0 = 0.80061762862741759f;
  constexpr float kOffset4 = 0.25f * kOffset3;
  const float v1 = StdMax(value * kMul0, 1.0e-3f);
  const float v2 = 1.0f / (v1 + kOffset2);
  const float v3 = 1.0f / __builtin_fmaf(v1, v1, kOffset3);
  const float v4 = 1.0f / __builtin_fmaf(v1, v1, kOffset4);
  return kBase + __builtin_fmaf(kMul4, v4, __builtin_fmaf(kMul2, v2, kMul3 * v3));
}

constexpr int kMaskThreads = 256;

// One thread per pixel column of one 4-row group: raw pixel mask for the 4
// pixels, the column's summed block differences in row order, then one
// pre-erosion mean per 4 columns (:622-664).
__global__ __launch_bounds__(kMaskThreads) void PixelMaskKernel(
  const float* __restrict__ y_plane, int width, int height,
  float* __restrict__ raw_mask, float* __restrict__ pre_erosion) {
  __shared__ float column_sum[kMaskThreads];
  const int x = blockIdx.x * kMaskThreads + threadIdx.x;
  const int group_y = blockIdx.y;
  if (x < width) {
    const int left = x == 0 ? x : x - 1;
    const int right = x + 1 < width ? x + 1 : x;
    float sum = 0.0f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int y = group_y * 4 + r;
      const int top = y == 0 ? y : y - 1;
      const int bottom = y + 1 < height ? y + 1 : y;
      const float* row = y_plane + static_cast<size_t>(y) * width;
      const float center = row[x];
      const float base = 0.25f * (y_plane[static_cast<size_t>(bottom) * width + x] +
                                  y_plane[static_cast<size_t>(top) * width + x] +
                                  row[left] + row[right]);
      const float gamma = GammaDerivativeRatio<false>(center + 0.019f);
      float pixel_difference = fabsf(gamma * (center - base));
      pixel_difference = FdlibmLog1pf(pixel_difference);
      raw_mask[static_cast<size_t>(y) * width + x] =
        1.0f / (pixel_difference + 0.01f);
      float block_difference = gamma * (center - base);
      block_difference *= block_difference;
      block_difference = StdMin(block_difference, 0.2f);
      block_difference = MaskingSqrt(block_difference);
      sum = r == 0 ? block_difference : sum + block_difference;
    }
    column_sum[threadIdx.x] = sum;
  }
  __syncthreads();
  if (x < width && (x & 3) == 0) {
    const int t = threadIdx.x;
    pre_erosion[static_cast<size_t>(group_y) * (width / 4) + x / 4] =
      (column_sum[t] + column_sum[t + 1] + column_sum[t + 2] +
       column_sum[t + 3]) * 0.25f;
  }
}

struct ErosionWeights {
  float w[4];
};

// :172-217
__device__ __forceinline__ void StoreMin4(float v, float* m) {
  if (v >= m[3]) return;
  if (v < m[0]) { m[3] = m[2]; m[2] = m[1]; m[1] = m[0]; m[0] = v; }
  else if (v < m[1]) { m[3] = m[2]; m[2] = m[1]; m[1] = v; }
  else if (v < m[2]) { m[3] = m[2]; m[2] = v; }
  else { m[3] = v; }
}

__device__ __forceinline__ void Sort4(float* v) {
#define GJXL_SWAP(a, b) if (v[a] > v[b]) { const float t = v[a]; v[a] = v[b]; v[b] = t; }
  GJXL_SWAP(0, 1) GJXL_SWAP(0, 2) GJXL_SWAP(0, 3)
  GJXL_SWAP(1, 2) GJXL_SWAP(1, 3) GJXL_SWAP(2, 3)
#undef GJXL_SWAP
}

// One eroded value of the pre-erosion map (w4 x h4) at (x, y) (:256-280).
__device__ __forceinline__ float ErodedValue(const float* __restrict__ src,
                                             int sw, int sh, int x, int y,
                                             const ErosionWeights& ew) {
  const int top = y == 0 ? y : y - 1;
  const int bottom = y + 1 < sh ? y + 1 : y;
  const int left = x == 0 ? x : x - 1;
  const int right = x + 1 < sw ? x + 1 : x;
  auto at = [&](int xx, int yy) { return src[static_cast<size_t>(yy) * sw + xx]; };
  float m[4] = {at(x, y), at(left, y), at(right, y), at(left, top)};
  Sort4(m);
  StoreMin4(at(x, top), m);
  StoreMin4(at(right, top), m);
  StoreMin4(at(left, bottom), m);
  StoreMin4(at(x, bottom), m);
  StoreMin4(at(right, bottom), m);
  return ew.w[0] * m[0] + ew.w[1] * m[1] + ew.w[2] * m[2] + ew.w[3] * m[3];
}

struct ModulationParams {
  float multiplier;
  float addend;
};

// One thread per 8x8 block: fuzzy erosion (2x2 eroded cells accumulated in
// the reference's row-major order), strategy mask, then PerBlockModulations
// (:332-467) with the lane-ordered sums of the reference.
__global__ __launch_bounds__(256) void BlockFieldKernel(
  const float* __restrict__ opsin, int width, int height,
  const float* __restrict__ pre_erosion, ErosionWeights ew,
  ModulationParams mp, float* __restrict__ quant_field,
  float* __restrict__ strategy_mask) {
  const int bw = width / 8, bh = height / 8;
  const int bx = blockIdx.x * blockDim.x + threadIdx.x;
  const int by = blockIdx.y;
  if (bx >= bw || by >= bh) return;
  const int sw = width / 4, sh = height / 4;
  float q = ErodedValue(pre_erosion, sw, sh, 2 * bx, 2 * by, ew);
  q += ErodedValue(pre_erosion, sw, sh, 2 * bx + 1, 2 * by, ew);
  q += ErodedValue(pre_erosion, sw, sh, 2 * bx, 2 * by + 1, ew);
  q += ErodedValue(pre_erosion, sw, sh, 2 * bx + 1, 2 * by + 1, ew);
  const size_t bi = static_cast<size_t>(by) * bw + bx;
  strategy_mask[bi] = 1.0f / (q + 0.001f);

  const size_t plane = static_cast<size_t>(width) * height;
  const float* X = opsin;
  const float* Y = opsin + plane;
  const float* B = opsin + 2 * plane;
  const int px = bx * 8, py = by * 8;
  const float mask_value = ComputeMask(q);

  // GammaModulation (:332-362)
  float lane[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int dy = 0; dy < 8; ++dy) {
    const size_t r = static_cast<size_t>(py + dy) * width + px;
#pragma unroll
    for (int dx = 0; dx < 8; ++dx) {
      const float in_y = Y[r + dx] + 0.16f;
      const float in_x = X[r + dx];
      lane[dx & 3] += GammaDerivativeRatio<true>(in_y - in_x);
      lane[dx & 3] += GammaDerivativeRatio<true>(in_y + in_x);
    }
  }
  const float overall = ((lane[0] + lane[1]) + (lane[2] + lane[3])) *
                        (0.5f / 64.0f);
  float gamma_value = __builtin_nanf("");
  if (isfinite(overall) && overall > 0.0f) {
    gamma_value = __builtin_fmaf(0.1005613337192697f, FastLog2(overall),
                                 mask_value);
  }

  // HighFrequencyModulation (:364-392)
  constexpr float kHfLimit = 0.0206f;
  lane[0] = lane[1] = lane[2] = lane[3] = 0.0f;
  for (int dy = 0; dy < 8; ++dy) {
    const float* row = Y + static_cast<size_t>(py + dy) * width + px;
    const float* next = dy + 1 < 8 ? row + width : row;
#pragma unroll
    for (int dx = 0; dx < 8; ++dx) {
      if (dx + 1 < 8) lane[dx & 3] += StdMin(kHfLimit, fabsf(row[dx] - row[dx + 1]));
      lane[dx & 3] += StdMin(kHfLimit, fabsf(row[dx] - next[dx]));
    }
  }
  const float hf_sum = (lane[0] + lane[1]) + (lane[2] + lane[3]);
  const float hf_value = gamma_value + (hf_sum * -0.38f + 0.42f);

  // BlueModulation (:394-424)
  constexpr float kBlueLimit = 0.010474084867598155f;
  constexpr float kBlueOffset = 0.0031994768654636393f;
  lane[0] = lane[1] = lane[2] = lane[3] = 0.0f;
  for (int dy = 0; dy < 8; ++dy) {
    const size_t r = static_cast<size_t>(py + dy) * width + px;
#pragma unroll
    for (int dx = 0; dx < 8; ++dx) {
      const float ey = Y[r + dx] + kBlueOffset + fabsf(X[r + dx]);
      if (B[r + dx] > ey) lane[dx & 3] += StdMin(B[r + dx] - ey, kBlueLimit);
    }
  }
  float blue_sum = (lane[0] + lane[1]) + (lane[2] + lane[3]);
  if (blue_sum >= 32.0f * kBlueLimit) blue_sum = 64.0f * kBlueLimit - blue_sum;
  blue_sum = StdMin(blue_sum, 15.463398341612438f * kBlueLimit);
  const float blue_value = gamma_value + blue_sum * 0.90590804735610064f;

  const float exponent = StdMin(hf_value, blue_value);
  const float e2 = exponent * 1.442695041f;
  const float p = (e2 >= -126.0f && e2 <= 127.0f) ? FastPow2(e2)
                                                  : __builtin_nanf("");
  quant_field[bi] = p * mp.multiplier + mp.addend;
}

}  // namespace

// Host-side dispatch used by the C entry points and the pipeline driver.
GJXLResult LaunchGaborishInverse(hipStream_t stream, const float* opsin,
                                 int width, int height,
                                 const float multipliers[3], float* out) {
  Symmetric5WeightSet set;
  for (int c = 0; c < 3; ++c) set.plane[c] = GaborishInverseWeights(multipliers[c]);
  return LaunchSymmetric5(stream, opsin, out, width, height, 3, set);
}

size_t InitialQuantScratchFloats(int width, int height) {
  return static_cast<size_t>(width) * height +
         static_cast<size_t>(width / 4) * (height / 4);
}

// Workspace of gjxl_amd_initial_quant_field_no_alloc: the raw pixel mask and
// the pre-erosion map of one padded image of at most
// GJXL_AMD_NO_ALLOC_MAX_PIXELS pixels (17/16 floats per pixel, 17.8 MB).
// Each device holds its own copy for the life of the process.  In host code
// the name is not a device address; kernels receive the address that
// LookUpInitialQuantStaticScratch obtained.
constexpr size_t kInitialQuantStaticFloats =
  static_cast<size_t>(GJXL_AMD_NO_ALLOC_MAX_PIXELS) +
  static_cast<size_t>(GJXL_AMD_NO_ALLOC_MAX_PIXELS) / 16;
static __device__ __align__(256) float initial_quant_scratch[kInitialQuantStaticFloats];

GJXLResult LookUpInitialQuantStaticScratch(float** out) {
  void* address = nullptr;
  GJXL_AMD_TRY(Check(hipGetSymbolAddress(&address, HIP_SYMBOL(initial_quant_scratch)),
                     "hipGetSymbolAddress(initial quant scratch)"));
  if (address == nullptr) {
    return Fail(GJXL_ERROR_BACKEND, "initial quant scratch has no device address");
  }
  *out = static_cast<float*>(address);
  return GJXL_OK;
}

GJXLResult LaunchInitialQuantField(hipStream_t stream, const float* opsin,
                                   int width, int height, float target,
                                   float rescale, float* quant_field,
                                   float* strategy_mask, float* pixel_mask,
                                   float* scratch) {
  float* raw_mask = scratch;
  float* pre = scratch + static_cast<size_t>(width) * height;
  const float* y_plane = opsin + static_cast<size_t>(width) * height;
  hipLaunchKernelGGL(PixelMaskKernel,
                     dim3(CeilDiv(width, kMaskThreads), height / 4),
                     dim3(kMaskThreads), 0, stream, y_plane, width, height,
                     raw_mask, pre);
  GJXL_AMD_TRY(CheckLaunch("PixelMaskKernel"));

  // FuzzyErosion weights (:228-254), evaluated as the reference does.
  const float mul_base[4] = {0.125f, 0.1f, 0.09f, 0.06f};
  const float mul_add[4] = {0.0f, -0.1f, -0.09f, -0.06f};
  const float total = 0.29959705784054957f;
  float mix = 0.0f;
  if (target < 2.0f) mix = (2.0f - target) * 0.5f;
  ErosionWeights ew;
  float weight_sum = 0.0f;
  for (int i = 0; i < 4; ++i) {
    ew.w[i] = mul_base[i] + mix * mul_add[i];
    weight_sum += ew.w[i];
  }
  for (int i = 0; i < 4; ++i) ew.w[i] *= total / weight_sum;

  // PerBlockModulations scale (:426-440)
  const float scale = 0.765f / target * rescale;
  const float base_level = 0.48f * scale;
  float dampen = 1.0f;
  if (target >= 2.0f) {
    dampen = 1.0f - (target - 2.0f) / 12.0f;
    dampen = dampen < 0.0f ? 0.0f : dampen;
  }
  const ModulationParams mp{scale * dampen, (1.0f - dampen) * base_level};
  const int bw = width / 8, bh = height / 8;
  hipLaunchKernelGGL(BlockFieldKernel, dim3(CeilDiv(bw, 64), bh), dim3(64), 0,
                     stream, opsin, width, height, pre, ew, mp, quant_field,
                     strategy_mask);
  GJXL_AMD_TRY(CheckLaunch("BlockFieldKernel"));

  Symmetric5WeightSet set;
  set.plane[0] = set.plane[1] = set.plane[2] = PixelMaskBlurWeights();
  return LaunchSymmetric5(stream, raw_mask, pixel_mask, width, height, 1, set);
}

}  // namespace gjxl_amd

using gjxl_amd::Fail;

namespace {
GJXLResult ValidatePadded(uint32_t width, uint32_t height) {
  if (width == 0 || height == 0 || width % 8 != 0 || height % 8 != 0 ||
      width > (1u << 28) || height > (1u << 28)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                "Padded image dimensions must be non-zero multiples of 8");
  }
  return GJXL_OK;
}
}  // namespace

extern "C" {

GJXLResult gjxl_amd_gaborish_inverse(GJXLAmdDevice* device, const float* opsin,
                                     uint32_t width, uint32_t height,
                                     const float multipliers[3], float* out) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  GJXL_AMD_TRY(ValidatePadded(width, height));
  if (opsin == nullptr || out == nullptr || opsin == out ||
      multipliers == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                "Gaborish buffers must be non-null and distinct");
  }
  for (int c = 0; c < 3; ++c) {
    if (!std::isfinite(multipliers[c])) {
      return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Gaborish multipliers must be finite");
    }
  }
  return gjxl_amd::LaunchGaborishInverse(device->stream, opsin, width, height,
                                         multipliers, out);
}

size_t gjxl_amd_initial_quant_field_scratch_bytes(uint32_t width,
                                                  uint32_t height) {
  return sizeof(float) * gjxl_amd::InitialQuantScratchFloats(width, height);
}

GJXLResult gjxl_amd_initial_quant_field(GJXLAmdDevice* device,
                                        const float* opsin, uint32_t width,
                                        uint32_t height, float target,
                                        float rescale, float* quant_field,
                                        float* strategy_mask,
                                        float* pixel_mask, float* scratch) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  GJXL_AMD_TRY(ValidatePadded(width, height));
  if (opsin == nullptr || quant_field == nullptr || strategy_mask == nullptr ||
      pixel_mask == nullptr || scratch == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Initial quant buffers must be non-null");
  }
  if (!std::isfinite(target) || target <= 0.0f || !std::isfinite(rescale) ||
      rescale <= 0.0f) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                "Initial quantization options must be finite and positive");
  }
  return gjxl_amd::LaunchInitialQuantField(device->stream, opsin, width, height,
                                           target, rescale, quant_field,
                                           strategy_mask, pixel_mask, scratch);
}

GJXLResult gjxl_amd_initial_quant_field_no_alloc(GJXLAmdDevice* device,
                                                 const float* opsin, uint32_t width,
                                                 uint32_t height, float target,
                                                 float rescale, float* quant_field,
                                                 float* strategy_mask,
                                                 float* pixel_mask) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  GJXL_AMD_TRY(ValidatePadded(width, height));
  if (static_cast<uint64_t>(width) * height > GJXL_AMD_NO_ALLOC_MAX_PIXELS) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                "Initial quant field without allocation takes at most "
                "GJXL_AMD_NO_ALLOC_MAX_PIXELS (2048 * 2048) padded pixels");
  }
  if (device->initial_quant_static_scratch == nullptr) {
    return Fail(GJXL_ERROR_BACKEND, "Device handle has no static scratch address");
  }
  if (opsin == nullptr || quant_field == nullptr || strategy_mask == nullptr ||
      pixel_mask == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Initial quant buffers must be non-null");
  }
  if (!std::isfinite(target) || target <= 0.0f || !std::isfinite(rescale) ||
      rescale <= 0.0f) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                "Initial quantization options must be finite and positive");
  }
  // Every workspace word the kernels read is written earlier in the same
  // call (raw mask and pre-erosion map by PixelMaskKernel), so nothing is
  // cleared.
  return gjxl_amd::LaunchInitialQuantField(device->stream, opsin, width, height,
                                           target, rescale, quant_field,
                                           strategy_mask, pixel_mask,
                                           device->initial_quant_static_scratch);
}

}  // extern "C"
