// a14 loop filters (Gaborish 3x3, EPF passes 1 and 2) and a15 XYB -> linear
// RGB on the unpadded reconstruction, as EvaluateQuantization applies them
// (src/codec/adaptive_quantization.cpp:972-1000): the padded reconstruction
// is cropped to the source extent and every filter mirrors at that extent.
#include <cmath>

#include "common.h"

namespace gjxl_amd {
namespace {

__device__ __forceinline__ int MirrorInt(int c, int size) {
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

struct GaborishWeights {
  float center, axis, diagonal;
};

// ApplyGaborish (src/codec/gaborish.cpp:101-190).  in: stride in_stride,
// plane in_plane; out: unpadded w x h planes.
__global__ __launch_bounds__(256) void GaborishKernel(
  const float* __restrict__ in, int in_stride, size_t in_plane, int w, int h,
  GaborishWeights g, float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  const int c = blockIdx.z;
  if (x >= w) return;
  const float* p = in + c * in_plane;
  auto S = [&](int dx, int dy) {
    return p[static_cast<size_t>(MirrorInt(y + dy, h)) * in_stride + MirrorInt(x + dx, w)];
  };
  const float axes = (S(-1, 0) + S(1, 0)) + (S(0, -1) + S(0, 1));
  const float diagonals = (S(-1, -1) + S(1, -1)) + (S(-1, 1) + S(1, 1));
  out[static_cast<size_t>(c) * w * h + static_cast<size_t>(y) * w + x] =
    g.center * S(0, 0) + g.axis * axes + g.diagonal * diagonals;
}

// ApplyEpfPass (src/codec/epf.cpp:61-205), pass 1 (plus-patch SAD) or pass 2
// (pixel SAD), four cardinal neighbours, default options.  A work group
// filters a 64x16 tile from an LDS copy of its three channels with a
// mirrored 2-pixel halo (every sample either pass reads is within 2 pixels).
constexpr int kEpfTileW = 64, kEpfTileH = 16, kEpfHalo = 2;
constexpr int kEpfLw = kEpfTileW + 2 * kEpfHalo, kEpfLh = kEpfTileH + 2 * kEpfHalo;

template <int kPass>
__global__ __launch_bounds__(256) void EpfKernel(
  const float* __restrict__ in, int w, int h,
  const float* __restrict__ inverse_sigma, float* __restrict__ out) {
  __shared__ float tile[3][kEpfLh][kEpfLw];
  const int x0 = blockIdx.x * kEpfTileW, y0 = blockIdx.y * kEpfTileH;
  const size_t plane = static_cast<size_t>(w) * h;
  for (int e = threadIdx.x; e < 3 * kEpfLh * kEpfLw; e += blockDim.x) {
    const int c = e / (kEpfLh * kEpfLw), r = (e / kEpfLw) % kEpfLh, col = e % kEpfLw;
    const int sy = MirrorInt(y0 + r - kEpfHalo, h), sx = MirrorInt(x0 + col - kEpfHalo, w);
    tile[c][r][col] = in[c * plane + static_cast<size_t>(sy) * w + sx];
  }
  __syncthreads();
  const int bw = (w + 7) / 8;
  constexpr float kMinInverseSigma = -1.1715728752538099024f / 0.3f;
  constexpr float kSigmaScale = 1.65f * (kPass == 2 ? 6.5f : 1.0f);
  constexpr int kCard[4][2] = {{0, -1}, {-1, 0}, {1, 0}, {0, 1}};
  constexpr int kPlus[5][2] = {{0, 0}, {0, -1}, {-1, 0}, {0, 1}, {1, 0}};
  constexpr float kScale[3] = {40.0f, 5.0f, 3.5f};
  const int tx = threadIdx.x % kEpfTileW;
  for (int ty = threadIdx.x / kEpfTileW; ty < kEpfTileH; ty += blockDim.x / kEpfTileW) {
    const int x = x0 + tx, y = y0 + ty;
    if (x >= w || y >= h) continue;
    const size_t i = static_cast<size_t>(y) * w + x;
    // tile sample at offset (dx, dy) from (x, y)
    auto S = [&](int c, int dx, int dy) {
      return tile[c][ty + kEpfHalo + dy][tx + kEpfHalo + dx];
    };
    const float bis = inverse_sigma[static_cast<size_t>(y / 8) * bw + x / 8];
    if (bis < kMinInverseSigma) {
      for (int c = 0; c < 3; ++c) out[c * plane + i] = S(c, 0, 0);
      continue;
    }
    const bool border = x % 8 == 0 || x % 8 == 7 || y % 8 == 0 || y % 8 == 7;
    const float sis = bis * kSigmaScale * (border ? 2.0f / 3.0f : 1.0f);
    float sum[3] = {S(0, 0, 0), S(1, 0, 0), S(2, 0, 0)};
    float weight_sum = 1.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int dx = kCard[k][0], dy = kCard[k][1];
      float sad = 0.0f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if constexpr (kPass == 1) {
          float cs = 0.0f;
#pragma unroll
          for (int o = 0; o < 5; ++o) {
            cs += fabsf(S(c, kPlus[o][0], kPlus[o][1]) -
                        S(c, dx + kPlus[o][0], dy + kPlus[o][1]));
          }
          sad = __builtin_fmaf(cs, kScale[c], sad);
        } else {
          sad = __builtin_fmaf(fabsf(S(c, 0, 0) - S(c, dx, dy)), kScale[c], sad);
        }
      }
      float weight = __builtin_fmaf(sad, sis, 1.0f);
      weight = 0.0f < weight ? weight : 0.0f;
      weight_sum += weight;
#pragma unroll
      for (int c = 0; c < 3; ++c) sum[c] = __builtin_fmaf(weight, S(c, dx, dy), sum[c]);
    }
    for (int c = 0; c < 3; ++c) out[c * plane + i] = sum[c] / weight_sum;
  }
}

struct InverseOpsin {
  float m[3][3];
  float scale;       // 255 / intensity
  float bias_cbrt;   // host cbrtf(bias)
};

// OpsinToLinearRgb (src/codec/color_transform.cpp:433-471).
__global__ __launch_bounds__(256) void OpsinToLinearKernel(
  const float* __restrict__ in, size_t count, InverseOpsin p,
  float* __restrict__ out) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  constexpr float kOpsinBias = 0.0037930732552754493f;
  const float x = in[i], y = in[count + i], b = in[2 * count + i];
  const float gamma[3] = {y + x + p.bias_cbrt, y - x + p.bias_cbrt, b + p.bias_cbrt};
  float mixed[3];
  for (int c = 0; c < 3; ++c) mixed[c] = gamma[c] * gamma[c] * gamma[c] - kOpsinBias;
  for (int row = 0; row < 3; ++row) {
    float v = p.scale * p.m[row][0] * mixed[0];
    v = __builtin_fmaf(p.scale * p.m[row][1], mixed[1], v);
    v = __builtin_fmaf(p.scale * p.m[row][2], mixed[2], v);
    out[row * count + i] = v;
  }
}

}  // namespace

// Gaborish from the padded reconstruction (stride pw, plane pw*ph) into
// scratch_a, EPF pass 1 into scratch_b, pass 2 into filtered, then XYB ->
// linear into linear.  All outputs unpadded w x h planes.
GJXLResult LaunchLoopFiltersAndColor(hipStream_t s, const float* reconstructed,
                                     int pw, int ph, int w, int h,
                                     const float* inverse_sigma,
                                     float* scratch_a, float* scratch_b,
                                     float* filtered, float* linear,
                                     float intensity_target) {
  const float w1 = 1.1f * 0.104699568f, w2 = 1.1f * 0.055680538f;
  const float divisor = 1.0f + 4.0f * (w1 + w2);
  const GaborishWeights g{1.0f / divisor, w1 / divisor, w2 / divisor};
  const dim3 grid(CeilDiv(w, 256), h, 3);
  hipLaunchKernelGGL(GaborishKernel, grid, dim3(256), 0, s, reconstructed, pw,
                     static_cast<size_t>(pw) * ph, w, h, g, scratch_a);
  GJXL_AMD_TRY(CheckLaunch("GaborishKernel"));
  const dim3 grid2(CeilDiv(w, kEpfTileW), CeilDiv(h, kEpfTileH));
  hipLaunchKernelGGL(EpfKernel<1>, grid2, dim3(256), 0, s, scratch_a, w, h,
                     inverse_sigma, scratch_b);
  GJXL_AMD_TRY(CheckLaunch("EpfKernel<1>"));
  hipLaunchKernelGGL(EpfKernel<2>, grid2, dim3(256), 0, s, scratch_b, w, h,
                     inverse_sigma, filtered);
  GJXL_AMD_TRY(CheckLaunch("EpfKernel<2>"));
  // The host libm's cbrtf, as the reference evaluates it; the volatile keeps
  // the compiler from folding it with a different method.
  volatile float bias = 0.0037930732552754493f;
  InverseOpsin p{{{11.031566901960783f, -9.866943921568629f, -0.16462299647058826f},
                  {-3.254147380392157f, 4.418770392156863f, -0.16462299647058826f},
                  {-3.6588512862745097f, 2.7129230470588235f, 1.9459282392156863f}},
                 255.0f / intensity_target,
                 cbrtf(bias)};
  const size_t count = static_cast<size_t>(w) * h;
  hipLaunchKernelGGL(OpsinToLinearKernel, dim3(CeilDiv(static_cast<uint32_t>(count), 256)),
                     dim3(256), 0, s, filtered, count, p, linear);
  return CheckLaunch("OpsinToLinearKernel");
}

}  // namespace gjxl_amd
