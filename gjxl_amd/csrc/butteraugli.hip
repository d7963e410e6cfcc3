// a16-a17: Butteraugli on the device and the per-transform 16-norm
// reduction of its map.
//
//   OpsinDynamicsKernel   5x5 mirrored blur + opsin dynamics (Gamma via the
//                         unfused FastLog2), butteraugli_internal.cpp:622-700
//   BlurRows/ColsKernel   separable Gaussian of radius 3, 6, 7 or 16 with the
//                         reference's interior summation order and its
//                         renormalised border columns (:245-620)
//   LowSplit/HighSplit/UltraSplit
//                         SeparateFrequencies (:702-846), as epilogues of the
//                         blurs' column passes
//   MaskSource/MaskKernel masking activity, blur, fuzzy erosion, MaskY/MaskDcY
//   DifferencesKernel     MaltaDiffMap of the six stages (butteraugli_
//                         distance_internal.cpp:482-577, zero outside the
//                         image), their accumulation (UHF, HF, MF), the L2
//                         terms, masked AC Y and the final composition
//                         (:657-870), per tile
//   Subsample/Expand/ComposeKernel
//                         the two-scale map and its maximum (:872-1003)
//   ReduceKernel          ReduceButteraugliDistanceMap (butteraugli.cpp:
//                         116-222)
//
// Every output is computed by one thread in the reference's own operation
// order (binary32, no contraction; binary64 where the reference promotes), so
// the maps equal the CPU's bit for bit.  The reference image's planes and
// masking terms are computed once per prepare and stay resident.
#include <cmath>
#include <cstring>
#include <new>

#include "butteraugli.h"
#include "libm_port.h"
#include "tables.h"

namespace gjxl_amd {
namespace {

constexpr int kThreads = 256;
constexpr int kTapStride = 64;  // 33 unscaled taps, then 17 scaled taps
constexpr float kSigmas[kSigmaCount] = {1.2f, 7.15593339443f, 3.22489901262f,
                                        1.56416327805f, 2.7f};

__device__ __forceinline__ int BMirror(int c, int size) {
  while (c < 0 || c >= size) c = c < 0 ? -c - 1 : 2 * size - 1 - c;
  return c;
}
// BMirror for -size <= c < 2 * size (one reflection), by selects.
__device__ __forceinline__ int BMirrorOnce(int c, int size) {
  c = c < 0 ? -c - 1 : c;
  return c >= size ? 2 * size - 1 - c : c;
}
__device__ __forceinline__ float BMax(float a, float b) { return a < b ? b : a; }

// ---- Gaussian blurs --------------------------------------------------------

// An interior output of ConvolutionWithTranspose from the 2R + 1 samples
// s[0 .. 2R] around it (s[R] is the centre): mirrored pairs times the scaled
// taps, accumulated into four interleaved partial sums, in the reference's
// order.
template <int R>
__device__ __forceinline__ float ConvolveInterior(const float* s, const float* sk) {
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const float pair = (s[i] + s[2 * R - i]) * sk[i];
    acc[i % 4] = i < 4 ? pair : acc[i % 4] + pair;
  }
  float r = s[R] * sk[R];
#pragma unroll
  for (int i = 0; i < 4 && i < R; ++i) r = r + acc[i];
  return r;
}

// A border output of ConvolutionWithTranspose (fewer than R samples on one
// side): the taps that fall inside the line, renormalised; sample(sx) gives
// the line's sample sx.  Every tap is visited in ascending order and the
// ones outside the line are skipped by selects, so the sums are the
// reference's; samples are read at clamped positions, which stay within
// pos - R .. pos + R.
template <int R, typename Sample>
__device__ __forceinline__ float ConvolveBorder(Sample sample, int pos, int len,
                                                const float* __restrict__ k) {
  float weight = 0.0f;
#pragma unroll
  for (int i = 0; i <= 2 * R; ++i) {
    const int sx = pos - R + i;
    const float next = weight + k[i];
    weight = sx >= 0 && sx < len ? next : weight;
  }
  const float scale = 1.0f / weight;
  float v[2 * R + 1];
#pragma unroll
  for (int i = 0; i <= 2 * R; ++i) {
    const int sx = pos - R + i;
    v[i] = sample(sx < 0 ? 0 : sx < len ? sx : len - 1);
  }
  float sum = 0.0f;
#pragma unroll
  for (int i = 0; i <= 2 * R; ++i) {
    const int sx = pos - R + i;
    const float next = sum + v[i] * k[i];
    sum = sx >= 0 && sx < len ? next : sum;
  }
  return sum * scale;
}

// One output of ConvolutionWithTranspose: position `pos` of a line of `len`
// samples spaced `stride` floats apart.
template <int R>
__device__ __forceinline__ float ConvolveLine(const float* __restrict__ line,
                                              int stride, int pos, int len,
                                              const float* __restrict__ taps) {
  const float* k = taps;
  const float* sk = taps + 33;
  const int first = len < R ? len : R;
  const int last = len > R ? len - R : 0;
  if (pos >= first && pos < last) {
    const float* s = line + static_cast<ptrdiff_t>(pos - R) * stride;
    float window[2 * R + 1];
#pragma unroll
    for (int i = 0; i <= 2 * R; ++i) window[i] = s[i * stride];
    return ConvolveInterior<R>(window, sk);
  }
  return ConvolveBorder<R>(
    [&](int sx) { return line[static_cast<ptrdiff_t>(sx) * stride]; }, pos, len, k);
}

// Rows: a work group blurs kRowRows rows of kRowSpan pixels, one wave per
// row, from an LDS copy of the rows and their halo.  Each lane produces four
// adjacent outputs from one register window that it reads with 16-byte LDS
// reads (lanes read consecutive 16-byte words: conflict-free).
constexpr int kRowSpan = 256;
constexpr int kRowLanes = kRowSpan / 4;
constexpr int kRowRows = kThreads / kRowLanes;

template <int R>
__global__ __launch_bounds__(kThreads) void BlurRowsKernel(
  const float* __restrict__ in, int w, int h, const float* __restrict__ taps,
  float* __restrict__ out) {
  // samples x - R .. x + 3 + R of a lane's outputs x .. x + 3, whole float4s
  constexpr int kWindow = (2 * R + 4 + 3) / 4 * 4;
  constexpr int kUsed = kRowSpan - 4 + kWindow;
  // whole loads per lane, so that every load issues before the first store
  constexpr int kLoads = (kUsed + kRowLanes - 1) / kRowLanes;
  constexpr int kStride = kLoads * kRowLanes;
  __shared__ __align__(16) float tile[kRowRows][kStride];
  const TileIndex tile_index = XcdTile();
  const int x0 = tile_index.x * kRowSpan;
  const int row = threadIdx.x / kRowLanes, lane = threadIdx.x % kRowLanes;
  const int y = tile_index.y * kRowRows + row;
  // tile column c holds sample x0 - R + c of the row
  const float* src = in + static_cast<size_t>(y < h ? y : h - 1) * w;
#pragma unroll
  for (int k = 0; k < kLoads; ++k) {
    const int c = lane + k * kRowLanes;
    const int gx = x0 - R + c;
    // clamped reads and a select instead of a branch per load
    const float v = src[gx < 0 ? 0 : gx < w ? gx : w - 1];
    tile[row][c] = gx >= 0 && gx < w && y < h ? v : 0.0f;
  }
  __syncthreads();
  if (y >= h) return;
  const float* line = tile[row];
  const float* sk = taps + 33;
  const int first = w < R ? w : R;
  const int last = w > R ? w - R : 0;
  // interior outputs: four adjacent ones per lane
  const int x = x0 + 4 * lane;
  float* dst = out + static_cast<size_t>(y) * w;
  if (x >= first && x + 3 < last) {
    float window[kWindow];
    const float4* s4 = reinterpret_cast<const float4*>(&line[4 * lane]);
#pragma unroll
    for (int m = 0; m < kWindow / 4; ++m) {
      const float4 v = s4[m];
      window[4 * m] = v.x;
      window[4 * m + 1] = v.y;
      window[4 * m + 2] = v.z;
      window[4 * m + 3] = v.w;
    }
    float r[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) r[j] = ConvolveInterior<R>(window + j, sk);
    if ((w & 3) == 0) {
      // rows start on 16-byte boundaries
      *reinterpret_cast<float4*>(dst + x) = make_float4(r[0], r[1], r[2], r[3]);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) dst[x + j] = r[j];
    }
  } else {
    // a lane that straddles an end of the interior
#pragma unroll 1
    for (int j = 0; j < 4; ++j) {
      const int xj = x + j;
      if (xj >= first && xj < last && xj < w) {
        dst[xj] = ConvolveInterior<R>(line + 4 * lane + j, sk);
      }
    }
  }
  // border outputs (x < first or x >= max(first, last)), one per lane: at
  // most R at each end of the row
  const int left = min(max(first - x0, 0), kRowSpan);
  const int right_begin = max(max(last, first), x0);
  const int right = max(min(w, x0 + kRowSpan) - right_begin, 0);
  if (lane < left + right) {
    const int xb = lane < left ? x0 + lane : right_begin + (lane - left);
    dst[xb] = ConvolveBorder<R>([&](int sx) { return line[sx - x0 + R]; }, xb, w, taps);
  }
}

// Columns: each thread produces kColOutputs outputs down one column from
// one register window (coalesced across the wave).  The epilogue takes the
// blurred value v of pixel i: a store, or the frequency split that consumes
// the blur, so that the blurred plane is never written.  `Load(i, src, in)`
// reads the kLoads values the split of pixel i needs (src: the blur's
// input plane) and `Store(i, v, in)`
// writes its results; a thread reads its window and every output's split
// inputs before it stores anything (its stores may alias its later reads
// as far as the compiler knows, which would otherwise make each output
// wait for its own reads).
constexpr int kColOutputs = 8;

template <int R, typename Epilogue>
__global__ __launch_bounds__(kThreads) void BlurColsKernel(
  const float* __restrict__ in, int w, int h, const float* __restrict__ taps,
  Epilogue epilogue) {
  constexpr int kOutputs = kColOutputs;
  // dispatch order (no XCD remap): horizontally adjacent groups read the
  // same rows, and a band's halo rows are re-read from L2 either way
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y0 = blockIdx.y * kOutputs;
  if (x >= w) return;
  const int first = h < R ? h : R;
  const int last = h > R ? h - R : 0;
  if (y0 >= first && y0 + kOutputs - 1 < last) {
    float window[kOutputs + 2 * R];
#pragma unroll
    for (int j = 0; j < kOutputs + 2 * R; ++j) {
      window[j] = in[static_cast<size_t>(y0 - R + j) * w + x];
    }
    float split[kOutputs][Epilogue::kLoads];
#pragma unroll
    for (int p = 0; p < kOutputs; ++p) epilogue.Load(static_cast<size_t>(y0 + p) * w + x, in, split[p]);
    const float* sk = taps + 33;
#pragma unroll
    for (int p = 0; p < kOutputs; ++p) {
      epilogue.Store(static_cast<size_t>(y0 + p) * w + x, ConvolveInterior<R>(window + p, sk),
                     split[p]);
    }
    return;
  }
  for (int p = 0; p < kOutputs; ++p) {
    const int y = y0 + p;
    if (y >= h) break;
    const size_t i = static_cast<size_t>(y) * w + x;
    float split[Epilogue::kLoads];
    epilogue.Load(i, in, split);
    epilogue.Store(i, ConvolveLine<R>(in + x, w, y, h, taps), split);
  }
}

// ---- opsin dynamics ---------------------------------------------------------

__device__ __forceinline__ float ButteraugliFastLog2(float v) {
  const uint32_t bits = __float_as_uint(v);
  const int32_t e = static_cast<int32_t>(bits - 0x3f2aaaabu) >> 23;
  const float x = __uint_as_float(bits - (static_cast<uint32_t>(e) << 23)) - 1.0f;
  float num = 0.74245873327820566f * x + 1.4287160470083755f;
  num = num * x + -1.8503833400518310e-06f;
  float den = 0.17409343003366853f * x + 1.0096718572241148f;
  den = den * x + 0.99032814277590719f;
  return num / den + static_cast<float>(e);
}

__device__ __forceinline__ float ButteraugliGamma(float v) {
  constexpr float kRetMul = 19.245013259874995f * 0.6931471805599453f;
  v = BMax(v, 0.0f);
  return kRetMul * ButteraugliFastLog2(v + 9.9710635769299145f) + -23.16046239805755f;
}

__device__ __forceinline__ void Absorbance(float r, float g, float b, bool clamp,
                                           float out[3]) {
  out[0] = 0.29956550340058319f * r +
           (0.63373087833825936f * g + (0.077705617820981968f * b + 1.7557483643287353f));
  out[1] = 0.22158691104574774f * r +
           (0.69391388044116142f * g + (0.0987313588422f * b + 1.7557483643287353f));
  out[2] = 0.02f * r + (0.02f * g + (0.20480129041026129f * b + 12.226454707163354f));
  if (clamp) {
    out[0] = BMax(out[0], 1.7557483643287353f);
    out[1] = BMax(out[1], 1.7557483643287353f);
    out[2] = BMax(out[2], 12.226454707163354f);
  }
}

// One work group per 64x16 tile: the three planes' tile with a mirrored
// 2-pixel halo in LDS, then per pixel the 5-tap horizontal sums of the five
// rows it needs and their vertical combination, in the reference's order.
constexpr int kOpsinTileW = 64, kOpsinTileH = 16;
constexpr int kOpsinRows = kOpsinTileH / (kThreads / kOpsinTileW);  // per thread

// kOnce: w and h are at least 2, so every coordinate of the 2-pixel halo
// needs one reflection.
template <bool kOnce>
__global__ __launch_bounds__(kThreads) void OpsinDynamicsKernel(
  const float* __restrict__ rgb, int w, int h, float intensity,
  const float* __restrict__ taps, float* __restrict__ xyb) {
  constexpr int TW = kOpsinTileW + 4, TH = kOpsinTileH + 4;
  __shared__ float tile[3][TH][TW];
  const TileIndex tile_index = XcdTile();
  const int x0 = tile_index.x * kOpsinTileW, y0 = tile_index.y * kOpsinTileH;
  const size_t n = static_cast<size_t>(w) * h;
  auto mirror = [](int c, int size) { return kOnce ? BMirrorOnce(c, size) : BMirror(c, size); };
  // The tile's 3 x TH rows of TW = 64 + 4 columns: each wave reads kWaveRows
  // whole rows, one lane per column of the first 64, with the row address
  // uniform (scalar arithmetic); the last 4 columns of every row are one
  // read per thread.  Every read issues before the first store.
  static_assert(TW == 64 + 4 && (3 * TH) % (kThreads / 64) == 0 && 3 * TH * 4 <= kThreads,
                "tile rows split over the waves");
  constexpr int kWaveRows = 3 * TH / (kThreads / 64);
  constexpr int kStripCount = 3 * TH * 4;
  const int lane = threadIdx.x % 64;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / 64));
  const int sx = mirror(x0 - 2 + lane, w);
  float rows[kWaveRows];
#pragma unroll
  for (int i = 0; i < kWaveRows; ++i) {
    const int row = wave * kWaveRows + i;
    const int c = row / TH, r = row % TH;
    rows[i] = rgb[c * n + static_cast<size_t>(mirror(y0 - 2 + r, h)) * w + sx];
  }
  float strip;
  {
    const int e = min(static_cast<int>(threadIdx.x), kStripCount - 1);
    const int row = e / 4, c = row / TH, r = row % TH;
    strip = rgb[c * n + static_cast<size_t>(mirror(y0 - 2 + r, h)) * w +
                mirror(x0 - 2 + 64 + e % 4, w)];
  }
#pragma unroll
  for (int i = 0; i < kWaveRows; ++i) {
    (&tile[0][0][0])[(wave * kWaveRows + i) * TW + lane] = rows[i];
  }
  if (threadIdx.x < kStripCount) {
    (&tile[0][0][0])[(threadIdx.x / 4) * TW + 64 + threadIdx.x % 4] = strip;
  }
  __syncthreads();
  const float w0 = taps[33 + 2], w1 = taps[33 + 1], w2 = taps[33 + 0];
  const int lx = threadIdx.x % kOpsinTileW;
  const int ly0 = (threadIdx.x / kOpsinTileW) * kOpsinRows;
  const int x = x0 + lx;
  if (x >= w) return;
  // horizontal sums of tile rows ly0 .. ly0 + kOpsinRows + 3, per channel
  float hz[3][kOpsinRows + 4];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
#pragma unroll
    for (int r = 0; r < kOpsinRows + 4; ++r) {
      const float* row = &tile[c][ly0 + r][lx + 2];
      float v = row[0] * w0;
      v = (row[-1] + row[1]) * w1 + v;
      v = (row[-2] + row[2]) * w2 + v;
      hz[c][r] = v;
    }
  }
#pragma unroll
  for (int j = 0; j < kOpsinRows; ++j) {
    const int y = y0 + ly0 + j;
    if (y >= h) break;
    float blurred[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float r = hz[c][j + 2] * w0;
      r = (hz[c][j + 1] + hz[c][j + 3]) * w1 + r;
      r = (hz[c][j] + hz[c][j + 4]) * w2 + r;
      blurred[c] = r;
    }
    const size_t i = static_cast<size_t>(y) * w + x;
    float pre[3], cur[3], sens[3];
    Absorbance(blurred[0] * intensity, blurred[1] * intensity, blurred[2] * intensity,
               true, pre);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      pre[c] = BMax(pre[c], 1.0e-4f);
      sens[c] = BMax(ButteraugliGamma(pre[c]) / pre[c], 1.0e-4f);
    }
    const int ly = ly0 + j + 2, tx = lx + 2;
    Absorbance(tile[0][ly][tx] * intensity, tile[1][ly][tx] * intensity,
               tile[2][ly][tx] * intensity, false, cur);
    cur[0] = BMax(cur[0] * sens[0], 1.7557483643287353f);
    cur[1] = BMax(cur[1] * sens[1], 1.7557483643287353f);
    cur[2] = BMax(cur[2] * sens[2], 12.226454707163354f);
    xyb[i] = cur[0] - cur[1];
    xyb[n + i] = cur[0] + cur[1];
    xyb[2 * n + i] = cur[2];
  }
}

// ---- frequency separation ---------------------------------------------------

__device__ __forceinline__ float MaximumClamp(float v, float maximum) {
  constexpr float k = 0.724216145665f;
  if (v >= maximum) return (v - maximum) * k + maximum;
  if (v < -maximum) return (v + maximum) * k + -maximum;
  return v;
}
__device__ __forceinline__ float RemoveRange(float v, float width) {
  if (v > width) return v - width;
  if (v < -width) return v + width;
  return 0.0f;
}
__device__ __forceinline__ float AmplifyRange(float v, float width) {
  if (v > width) return v + width;
  if (v < -width) return v - width;
  return v + v;
}

// Column-pass epilogues of the separable blurs (SeparateFrequencies).

// Each split reads kLoads values per pixel with unconditional reads: a
// value a channel does not use is read from the blur's own input at the
// same pixel, which the thread's window has just read (no branch around a
// read, no extra memory traffic).
struct StoreBlur {
  static constexpr int kLoads = 1;
  float* out;
  __device__ __forceinline__ void Load(size_t, const float*, float (&in)[kLoads]) const {
    in[0] = 0.0f;
  }
  __device__ __forceinline__ void Store(size_t i, float v, const float (&)[kLoads]) const {
    out[i] = v;
  }
};

// The LF blur v of channel c: MF = XYB - LF, then the LF plane scaled.  The
// channels go Y, X, B: Y's LF stays raw until B's pass, which needs it and
// then scales it.
struct LowSplit {
  static constexpr int kLoads = 2;
  int c;
  const float* xyb;
  float* psycho;
  size_t n;
  __device__ __forceinline__ void Load(size_t i, const float* blur_in, float (&in)[kLoads]) const {
    in[0] = xyb[c * n + i];
    in[1] = (c == 2 ? psycho + n : blur_in)[i];  // raw LF Y for channel 2
  }
  __device__ __forceinline__ void Store(size_t i, float v, const float (&in)[kLoads]) const {
    float* lf = psycho;
    float* mf = psycho + 3 * n;
    mf[c * n + i] = in[0] - v;
    if (c == 0) {
      lf[i] = v * 33.832837186260f;
    } else if (c == 1) {
      lf[n + i] = v;
    } else {
      const float ly = in[1];
      lf[2 * n + i] = (-0.362267051518f * ly + v) * 49.87984651440f;
      lf[n + i] = ly * 14.458268100570f;
    }
  }
};

// The 3.22 blur v of MF channel c: HF and MF of channel c; channel 1 also
// applies SuppressXByY to HF X.
struct HighSplit {
  static constexpr int kLoads = 2;
  int c;
  float* psycho;
  size_t n;
  __device__ __forceinline__ void Load(size_t i, const float* blur_in, float (&in)[kLoads]) const {
    in[0] = (c == 2 ? blur_in : psycho + (3 + c) * n)[i];  // MF of channels 0, 1
    in[1] = (c == 1 ? psycho + 6 * n : blur_in)[i];        // HF X for channel 1
  }
  __device__ __forceinline__ void Store(size_t i, float v, const float (&in)[kLoads]) const {
    float* mf = psycho + 3 * n;
    float* hf = psycho + 6 * n;
    if (c == 2) {
      mf[2 * n + i] = v;
      return;
    }
    const float high = in[0] - v;
    hf[c * n + i] = high;
    mf[c * n + i] = c == 0 ? RemoveRange(v, 0.29f) : AmplifyRange(v, 0.1f);
    if (c == 1) {
      constexpr float kS = 46.0f, kMin = 0.653020556257f;
      const float scaler = (kS / (high * high + kS)) * (1.0f - kMin) + kMin;
      hf[i] = scaler * in[1];
    }
  }
};

// The 1.56 blur v of HF channel c: UHF and HF of channel c.
struct UltraSplit {
  static constexpr int kLoads = 1;
  int c;
  float* psycho;
  size_t n;
  __device__ __forceinline__ void Load(size_t i, const float*, float (&in)[kLoads]) const {
    in[0] = psycho[(6 + c) * n + i];
  }
  __device__ __forceinline__ void Store(size_t i, float v, const float (&in)[kLoads]) const {
    float* hf = psycho + (6 + c) * n;
    float* uhf = psycho + (8 + c) * n;
    const float original = in[0];
    float hv = v;
    if (c == 0) {
      uhf[i] = RemoveRange(original - hv, 0.04f);
      hf[i] = RemoveRange(hv, 1.5f);
    } else {
      hv = MaximumClamp(hv, 28.4691806922f);
      float uv = original - hv;
      uv = MaximumClamp(uv, 5.19175294647f);
      uhf[i] = uv * 2.69313763794f;
      hf[i] = AmplifyRange(hv * 2.155f, 0.132f);
    }
  }
};

// ---- difference stages -------------------------------------------------------

struct MaltaScale {
  double gt;   // norm2_0_gt_1, a binary32 value
  double lt;   // norm2_0_lt_1, a binary32 value
  float norm;  // (float)norm
};

// The scaled difference MaltaDiffMap feeds its lines (butteraugli_distance_
// internal.cpp:512-555).
//
// The two binary32 quotients gt / den and lt / den come from one binary64
// reciprocal: r = 1 / den to within 2^-52 relative (hardware estimate and two
// Newton steps), then the binary64 product rounded to binary32.  A quotient
// of two binary32 values is never a binary32 rounding midpoint and lies at
// least 2^-49 (relative) from every midpoint, so a binary64 value within
// 2^-51 of it rounds to the same binary32 value: the correctly rounded
// quotient, as binary32 division gives.
//
// The four corrections s -/+ ds2 * (...) are one s +/- ds2 * (u - k) with
// u = |v1| mirrored by the sign of v0 and k = too_small or too_big picked by
// selects: negation is exact and rounding is symmetric, so -(x + y) is
// (-x) + (-y) and x - y is x + (-y), and the result is the reference's.
__device__ __forceinline__ float MaltaScaledDiff(float v0, float v1, const MaltaScale& m) {
  const float absolute = 0.5f * (fabsf(v0) + fabsf(v1));
  const float diff = v0 - v1;
  const double den = m.norm + absolute;
  double r = __builtin_amdgcn_rcp(den);
  r = fma(fma(-den, r, 1.0), r, r);
  r = fma(fma(-den, r, 1.0), r, r);
  const float scaler = static_cast<float>(m.gt * r);
  const float s = scaler * diff;
  const double ds2 = static_cast<float>(m.lt * r);
  const double magnitude = fabs(static_cast<double>(v0));
  const double too_small = 0.55 * magnitude, too_big = 1.05 * magnitude;
  // v0 < 0: dv1 > -too_small: s - ds2 * (dv1 + too_small)
  //         dv1 < -too_big:   s + ds2 * (-dv1 - too_big)
  // v0 >= 0: dv1 < too_small: s + ds2 * (too_small - dv1)
  //          dv1 > too_big:   s - ds2 * (dv1 - too_big)
  const bool negative = v0 < 0.0f;
  const double u = negative ? -v1 : v1;
  const bool near = u < too_small;
  const bool far = u > too_big;
  const double k = near ? too_small : too_big;
  const double w = ds2 * (u - k);
  const double ds = s;
  const float corrected = static_cast<float>(negative ? ds + w : ds - w);
  return near || far ? corrected : s;
}

// Keeps one Malta window's responses computed before the next window's LDS
// reads: an empty asm statement that takes the four values and orders
// memory, so that the compiler neither sinks their arithmetic below it nor
// hoists the next reads above it (with all windows live at once the kernel
// would need over 400 registers).
__device__ __forceinline__ void Settle(const float (&v)[4]) {
  asm volatile("" : : "v"(v[0]), "v"(v[1]), "v"(v[2]), "v"(v[3]) : "memory");
}

__device__ __forceinline__ float Sum4(float a, float b, float c, float d) {
  return (a + b) + (c + d);
}

// The 16 Malta lines (MaltaUnitLf / MaltaUnitFull).  Offsets as (dx, dy).
struct MaltaTap {
  signed char dx, dy;
};
constexpr MaltaTap kLfLines[16][5] = {
  {{-4, 0}, {-2, 0}, {0, 0}, {2, 0}, {4, 0}},
  {{0, -4}, {0, -2}, {0, 0}, {0, 2}, {0, 4}},
  {{-3, -3}, {-2, -2}, {0, 0}, {2, 2}, {3, 3}},
  {{3, -3}, {2, -2}, {0, 0}, {-2, 2}, {-3, 3}},
  {{1, -4}, {1, -2}, {0, 0}, {-1, 2}, {-1, 4}},
  {{-1, -4}, {-1, -2}, {0, 0}, {1, 2}, {1, 4}},
  {{-4, -1}, {-2, -1}, {0, 0}, {2, 1}, {4, 1}},
  {{-4, 1}, {-2, 1}, {0, 0}, {2, -1}, {4, -1}},
  {{-2, -3}, {-1, -2}, {0, 0}, {1, 2}, {2, 3}},
  {{2, -3}, {1, -2}, {0, 0}, {-1, 2}, {-2, 3}},
  {{-3, -2}, {-2, -1}, {0, 0}, {2, 1}, {3, 2}},
  {{3, -2}, {2, -1}, {0, 0}, {-2, 1}, {-3, 2}},
  {{-4, 2}, {-2, 1}, {0, 0}, {2, -1}, {4, -2}},
  {{-4, -2}, {-2, -1}, {0, 0}, {2, 1}, {4, 2}},
  {{-2, -4}, {-1, -2}, {0, 0}, {1, 2}, {2, 4}},
  {{2, -4}, {1, -2}, {0, 0}, {-1, 2}, {-2, 4}},
};
constexpr int kFullTaps[16] = {9, 9, 7, 7, 9, 9, 9, 9, 7, 7, 7, 7, 9, 9, 9, 9};
constexpr MaltaTap kFullLines[16][9] = {
  {{-4, 0}, {-3, 0}, {-2, 0}, {-1, 0}, {0, 0}, {1, 0}, {2, 0}, {3, 0}, {4, 0}},
  {{0, -4}, {0, -3}, {0, -2}, {0, -1}, {0, 0}, {0, 1}, {0, 2}, {0, 3}, {0, 4}},
  {{-3, -3}, {-2, -2}, {-1, -1}, {0, 0}, {1, 1}, {2, 2}, {3, 3}},
  {{3, -3}, {2, -2}, {1, -1}, {0, 0}, {-1, 1}, {-2, 2}, {-3, 3}},
  {{1, -4}, {1, -3}, {1, -2}, {0, -1}, {0, 0}, {0, 1}, {-1, 2}, {-1, 3}, {-1, 4}},
  {{-1, -4}, {-1, -3}, {-1, -2}, {0, -1}, {0, 0}, {0, 1}, {1, 2}, {1, 3}, {1, 4}},
  {{-4, -1}, {-3, -1}, {-2, -1}, {-1, 0}, {0, 0}, {1, 0}, {2, 1}, {3, 1}, {4, 1}},
  {{-4, 1}, {-3, 1}, {-2, 1}, {-1, 0}, {0, 0}, {1, 0}, {2, -1}, {3, -1}, {4, -1}},
  {{-2, -3}, {-1, -2}, {-1, -1}, {0, 0}, {1, 1}, {1, 2}, {2, 3}},
  {{2, -3}, {1, -2}, {1, -1}, {0, 0}, {-1, 1}, {-1, 2}, {-2, 3}},
  {{-3, -2}, {-2, -1}, {-1, -1}, {0, 0}, {1, 1}, {2, 1}, {3, 2}},
  {{3, -2}, {2, -1}, {1, -1}, {0, 0}, {-1, 1}, {-2, 1}, {-3, 2}},
  {{-4, 1}, {-3, 1}, {-2, 1}, {-1, 0}, {0, 0}, {1, 0}, {2, -1}, {3, -1}, {4, -1}},
  {{-4, -1}, {-3, -1}, {-2, -1}, {-1, 0}, {0, 0}, {1, 0}, {2, 1}, {3, 1}, {4, 1}},
  {{-1, -4}, {-1, -3}, {-1, -2}, {0, -1}, {0, 0}, {0, 1}, {1, 2}, {1, 3}, {1, 4}},
  {{1, -4}, {1, -3}, {1, -2}, {0, -1}, {0, 0}, {0, 1}, {-1, 2}, {-1, 3}, {-1, 4}},
};

// Tiles of the difference stages: 64x32 outputs, a 4-pixel halo.
constexpr int kMaltaTileW = 64, kMaltaTileH = 32, kMaltaHalo = 4;
constexpr int kMaltaLw = kMaltaTileW + 2 * kMaltaHalo, kMaltaLh = kMaltaTileH + 2 * kMaltaHalo;
constexpr int kMaltaQuads = kMaltaTileW / 4;
constexpr int kMaltaWinW = 4 + 2 * kMaltaHalo, kMaltaWinH = 1 + 2 * kMaltaHalo;
constexpr int kMaltaRowsPerPass = kThreads / kMaltaQuads;
static_assert(kMaltaTileH == 2 * kMaltaRowsPerPass, "two rows of four outputs per thread");
static_assert(kMaltaLw % 4 == 0, "16-byte aligned LDS rows");

struct L2Weights {
  float asym_gt[2];  // weight_0_gt_1 * 0.8 for HF X, Y
  float asym_lt[2];  // weight_0_lt_1 * 0.8
  bool asym_on[2];
  float mf[3];
  float lf[3];
};

// The 16 Malta lines of one pixel (MaltaUnitLf / MaltaUnitFull) from a
// register window P(dx, dy), summed in the reference's order, 0 + response.
template <bool kLowFrequency, typename Window>
__device__ __forceinline__ float MaltaResponse(Window P) {
  float result = 0.0f;
#pragma unroll
  for (int l = 0; l < 16; ++l) {
    float s;
    if constexpr (kLowFrequency) {
      const MaltaTap* t = kLfLines[l];
      s = Sum4(P(t[0].dx, t[0].dy), P(t[1].dx, t[1].dy), P(t[2].dx, t[2].dy),
               P(t[3].dx, t[3].dy) + P(t[4].dx, t[4].dy));
    } else {
      const MaltaTap* t = kFullLines[l];
      float v[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) v[k] = k < kFullTaps[l] ? P(t[k].dx, t[k].dy) : 0.0f;
      s = kFullTaps[l] == 9
            ? (Sum4(v[0], v[1], v[2], v[3]) + Sum4(v[4], v[5], v[6], v[7])) + v[8]
            : Sum4(v[0], v[1], v[2], Sum4(v[3], v[4], v[5], v[6]));
    }
    result = l == 0 ? s * s : s * s + result;
  }
  return 0.0f + result;
}

// The psycho plane of Malta stage k (MF Y, MF X, HF Y, HF X, UHF Y, UHF X)
// and its line set (low-frequency lines for the first four).
constexpr int kMaltaPlanes[6] = {4, 3, 7, 6, 9, 8};
struct MaltaScales {
  MaltaScale m[6];
};

// The difference stages of one scale, fused per 64x32 tile:
// * MaltaDiffMap of the six stages, three at a time: the scaled differences
//   of the tile and its 4-pixel halo for each of three stages go to LDS
//   (zero outside the image, as the reference's ReadWithZeroBorder), the
//   reads of one stage issued before the previous stage's differences; each
//   thread then evaluates four adjacent outputs of a row, for two rows, from
//   9 x 12 register windows of 16-byte LDS reads, and stores the stage;
// * the Malta accumulation in the reference's order (UHF, HF, MF per
//   channel): the first three stages are UHF Y, HF Y and UHF X, and their
//   partial sums ((0 + UHF Y) + HF Y, 0 + UHF X) go to planes 0 and 1 of
//   `stages`, which the same thread reads back after the last three (HF X,
//   MF X, MF Y);
// * then the asymmetric HF and symmetric MF L2 terms, the DC differences
//   from LF, the Y blurred-activity term, masking and the final map
//   (MaskPsychoImage / CombineChannelsToDiffmap): stage 8.
// kKeepStages stores the six Malta maps (stages 0-5) and the masked Y term
// (stage 7) for inspection instead, and forms the partial sums from the
// first three maps read back; without it only the partial sums and the
// final map are written.
// Three stages of LDS (34.6 KB) let four work groups share a CU.
template <bool kKeepStages>
__global__ __launch_bounds__(kThreads) void DifferencesKernel(
  const float* __restrict__ ps0, const float* __restrict__ ps1, int w, int h,
  MaltaScales ms, L2Weights lw, const float* __restrict__ ref_blurred,
  const float* __restrict__ dist_blurred, const float* __restrict__ mask_y,
  const float* __restrict__ mask_dc, float x_multiplier, float* __restrict__ stages) {
  __shared__ __align__(16) float tile[3][kMaltaLh][kMaltaLw];
  const TileIndex tile_index = XcdTile();
  const int x0 = tile_index.x * kMaltaTileW, y0 = tile_index.y * kMaltaTileH;
  const size_t n = static_cast<size_t>(w) * h;
  const int tid = threadIdx.x;
  // A stage's tile is kMaltaLh rows of kMaltaLw = 64 + 8 columns.  Each wave
  // forms kFormRows whole rows, one lane per column of the first 64, with
  // the row's address and its inside test uniform (scalar arithmetic); the
  // last 8 columns of every row are kFormStrip elements per thread.  Reads
  // are clamped to the image (32-bit byte offsets: a plane is under 4 GiB)
  // and elements outside it are zero.
  static_assert(kMaltaLw == 64 + 8 && kMaltaLh % (kThreads / 64) == 0, "rows split over the waves");
  constexpr int kFormRows = kMaltaLh / (kThreads / 64);
  constexpr int kStripCount = kMaltaLh * 8;
  constexpr int kFormStrip = (kStripCount + kThreads - 1) / kThreads;
  constexpr int kFormSlots = kFormRows + kFormStrip;
  const int wave = __builtin_amdgcn_readfirstlane(tid / 64);
  // The index arithmetic of a formation is recomputed each time from an
  // opaque copy of the thread index: shared with the first formation, its
  // results would stay live through the first Malta phase and cost the
  // registers of a fourth wave per SIMD.
  struct FormIndex {
    uint32_t offset[kFormSlots];  // byte offset within a plane
    bool inside[kFormSlots];
    int lds[kFormStrip];          // tile index of the strip elements
  };
  auto index = [&](FormIndex& f) {
    int t = tid;
    asm volatile("" : "+v"(t));
    const int lane = t % 64;
    const int sx = x0 + lane - kMaltaHalo;
    const uint32_t sx4 = 4u * static_cast<uint32_t>(min(max(sx, 0), w - 1));
    const bool col_in = sx >= 0 && sx < w;
#pragma unroll
    for (int j = 0; j < kFormRows; ++j) {
      const int sy = y0 + wave * kFormRows + j - kMaltaHalo;
      f.offset[j] = 4u * static_cast<uint32_t>(min(max(sy, 0), h - 1) * w) + sx4;
      f.inside[j] = col_in && sy >= 0 && sy < h;
    }
#pragma unroll
    for (int j = 0; j < kFormStrip; ++j) {
      const int e = min(t + j * kThreads, kStripCount - 1);
      const int r = e / 8, col = 64 + e % 8;
      const int sy = y0 + r - kMaltaHalo, sxs = x0 + col - kMaltaHalo;
      f.offset[kFormRows + j] =
        4u * static_cast<uint32_t>(min(max(sy, 0), h - 1) * w + min(max(sxs, 0), w - 1));
      f.inside[kFormRows + j] = sxs >= 0 && sy >= 0 && sxs < w && sy < h;
      f.lds[j] = t + j * kThreads < kStripCount ? r * kMaltaLw + col : -1;
    }
  };
  auto load = [&](int k, const FormIndex& f, float (&v0)[kFormSlots], float (&v1)[kFormSlots]) {
    const char* p0 = reinterpret_cast<const char*>(ps0 + kMaltaPlanes[k] * n);
    const char* p1 = reinterpret_cast<const char*>(ps1 + kMaltaPlanes[k] * n);
#pragma unroll
    for (int j = 0; j < kFormSlots; ++j) {
      v0[j] = *reinterpret_cast<const float*>(p0 + f.offset[j]);
      v1[j] = *reinterpret_cast<const float*>(p1 + f.offset[j]);
    }
  };
  auto store = [&](int k, int slot, const FormIndex& f, const float (&v0)[kFormSlots],
                   const float (&v1)[kFormSlots]) {
    float* stage = &tile[slot][0][0];
    const int lane = tid % 64;
#pragma unroll
    for (int j = 0; j < kFormRows; ++j) {
      stage[(wave * kFormRows + j) * kMaltaLw + lane] =
        f.inside[j] ? MaltaScaledDiff(v0[j], v1[j], ms.m[k]) : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < kFormStrip; ++j) {
      if (f.lds[j] >= 0) {
        stage[f.lds[j]] = f.inside[kFormRows + j]
                            ? MaltaScaledDiff(v0[kFormRows + j], v1[kFormRows + j], ms.m[k])
                            : 0.0f;
      }
    }
  };
  // stages k0, k1, k2 into LDS slots 0, 1, 2
  auto form = [&](int k0, int k1, int k2) {
    FormIndex f;
    index(f);
    float a0[kFormSlots], a1[kFormSlots], b0[kFormSlots], b1[kFormSlots];
    load(k0, f, a0, a1);
    load(k1, f, b0, b1);
    store(k0, 0, f, a0, a1);
    load(k2, f, a0, a1);
    store(k1, 1, f, b0, b1);
    store(k2, 2, f, a0, a1);
  };
  const int q = tid % kMaltaQuads;
  const int x = x0 + 4 * q;
  // a thread past the right edge still takes part in forming the stages
  const bool active = x < w;
  // four consecutive pixels of a plane: one 16-byte access when rows start
  // on 16-byte boundaries (then x + 3 < w), else one access per pixel
  // (32-bit byte offsets from the plane's base: a plane is under 4 GiB)
  const bool vector = (w & 3) == 0;
  auto load4 = [&](const float* plane, size_t i, float (&out)[4]) {
    const char* at = reinterpret_cast<const char*>(plane) + 4u * static_cast<uint32_t>(i);
    if (vector) {
      const float4 v = *reinterpret_cast<const float4*>(at);
      out[0] = v.x;
      out[1] = v.y;
      out[2] = v.z;
      out[3] = v.w;
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        out[j] = x + j < w ? reinterpret_cast<const float*>(at)[j] : 0.0f;
      }
    }
  };
  auto store4 = [&](float* plane, size_t i, const float (&v)[4]) {
    char* at = reinterpret_cast<char*>(plane) + 4u * static_cast<uint32_t>(i);
    if (vector) {
      *reinterpret_cast<float4*>(at) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (x + j < w) reinterpret_cast<float*>(at)[j] = v[j];
      }
    }
  };
  // Malta of stage k (in LDS slot `slot`) for the four outputs of row ty
  auto malta = [&](auto stage, int slot, int ty, float (&out)[4]) {
    constexpr int k = decltype(stage)::value;
    // win[r][c] = tile[slot][ty + r][4q + c]: rows y - 4 .. y + 4, columns
    // x - 4 .. x + 7
    float win[kMaltaWinH][kMaltaWinW];
#pragma unroll
    for (int r = 0; r < kMaltaWinH; ++r) {
      const float4* s4 = reinterpret_cast<const float4*>(&tile[slot][ty + r][4 * q]);
#pragma unroll
      for (int c4 = 0; c4 < kMaltaWinW / 4; ++c4) {
        const float4 v = s4[c4];
        win[r][4 * c4] = v.x;
        win[r][4 * c4 + 1] = v.y;
        win[r][4 * c4 + 2] = v.z;
        win[r][4 * c4 + 3] = v.w;
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      auto P = [&](int dx, int dy) -> float {
        return win[kMaltaHalo + dy][j + kMaltaHalo + dx];
      };
      out[j] = MaltaResponse<(k < 4)>(P);
    }
  };
  form(4, 2, 5);
  __syncthreads();
  if (active) {
#pragma unroll 1
    for (int pass = 0; pass < 2; ++pass) {
      const int ty = tid / kMaltaQuads + pass * kMaltaRowsPerPass;
      const int y = y0 + ty;
      if (y >= h) break;
      const size_t base = static_cast<size_t>(y) * w + x;
      float sv[3][4];
      // one Malta window live at a time: the scheduling barriers stand where
      // the stores of the inspected stages keep the windows apart
      malta(std::integral_constant<int, 4>{}, 0, ty, sv[0]);
      if constexpr (kKeepStages) store4(stages + 4 * n, base, sv[0]);
      Settle(sv[0]);
      malta(std::integral_constant<int, 2>{}, 1, ty, sv[1]);
      if constexpr (kKeepStages) store4(stages + 2 * n, base, sv[1]);
      Settle(sv[1]);
      malta(std::integral_constant<int, 5>{}, 2, ty, sv[2]);
      if constexpr (kKeepStages) {
        store4(stages + 5 * n, base, sv[2]);
      } else {
        float acc_y[4], acc_x[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          acc_y[j] = 0.0f;
          acc_y[j] += sv[0][j];
          acc_y[j] += sv[1][j];
          acc_x[j] = 0.0f;
          acc_x[j] += sv[2][j];
        }
        store4(stages, base, acc_y);
        store4(stages + n, base, acc_x);
      }
    }
  }
  __syncthreads();
  form(3, 1, 0);
  __syncthreads();
  if (!active) return;
#pragma unroll 1
  for (int pass = 0; pass < 2; ++pass) {
    const int ty = tid / kMaltaQuads + pass * kMaltaRowsPerPass;
    const int y = y0 + ty;
    if (y >= h) return;
    const size_t base = static_cast<size_t>(y) * w + x;
    float sv[6][4];
    malta(std::integral_constant<int, 3>{}, 0, ty, sv[3]);
    if constexpr (kKeepStages) store4(stages + 3 * n, base, sv[3]);
    Settle(sv[3]);
    malta(std::integral_constant<int, 1>{}, 1, ty, sv[1]);
    if constexpr (kKeepStages) store4(stages + 1 * n, base, sv[1]);
    Settle(sv[1]);
    malta(std::integral_constant<int, 0>{}, 2, ty, sv[0]);
    if constexpr (kKeepStages) store4(stages, base, sv[0]);
    Settle(sv[0]);
    // the first half's partial sums as this thread stored them (or its
    // stages 4, 2, 5)
    float acc_y[4], acc_x[4];
    if constexpr (kKeepStages) {
      load4(stages + 4 * n, base, sv[4]);
      load4(stages + 2 * n, base, sv[2]);
      load4(stages + 5 * n, base, sv[5]);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        acc_y[j] = 0.0f;
        acc_y[j] += sv[4][j];
        acc_y[j] += sv[2][j];
        acc_x[j] = 0.0f;
        acc_x[j] += sv[5][j];
      }
    } else {
      load4(stages, base, acc_y);
      load4(stages + n, base, acc_x);
    }
    // the per-pixel inputs of the L2 terms and the final map
    float p0[8][4], p1[8][4];  // planes LF 0-2, MF 3-5, HF 6-7
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      load4(ps0 + c * n, base, p0[c]);
      load4(ps1 + c * n, base, p1[c]);
    }
    float rb[4], db[4], my[4], mdc[4];
    load4(ref_blurred, base, rb);
    load4(dist_blurred, base, db);
    load4(mask_y, base, my);
    load4(mask_dc, base, mdc);
    float masked_y[4], final_map[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float a[3], dcv[3];
      // Y: (UHF + HF) + MF; X: (UHF + HF) + MF, from 0 as the reference sums
      a[1] = acc_y[j];
      a[1] += sv[0][j];
      a[0] = acc_x[j];
      a[0] += sv[3][j];
      a[0] += sv[1][j];
      a[2] = 0.0f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (c < 2 && lw.asym_on[c]) {
          const float v0 = p0[6 + c][j], v1 = p1[6 + c][j];
          const float d = v0 - v1;
          const float total = (d * d) * lw.asym_gt[c] + a[c];
          const float mag = fabsf(v0);
          const float small = 0.4f * mag;
          // v0 < 0:  v1 > -small: v1 + small;  else v1 < -mag: -v1 - mag
          // v0 >= 0: v1 < small:  small - v1;  else v1 > mag:  v1 - mag
          // (else 0), as selects on the same comparisons
          const float below = v0 < 0.0f ? (v1 > -small ? v1 + small : (v1 < -mag ? -v1 - mag : 0.0f))
                                        : 0.0f;
          const float above = v1 < small ? small - v1 : (v1 > mag ? v1 - mag : 0.0f);
          const float sec = v0 < 0.0f ? below : above;
          a[c] = lw.asym_lt[c] * (sec * sec) + total;
        }
        if (lw.mf[c] != 0.0f) {
          const float d = p0[3 + c][j] - p1[3 + c][j];
          a[c] = (d * d) * lw.mf[c] + a[c];
        }
        const float e = p0[c][j] - p1[c][j];
        dcv[c] = lw.lf[c] == 0.0f ? 0.0f : e * e * lw.lf[c];
      }
      const float d = rb[j] - db[j];
      const float ay = a[1] + 10.0f * d * d;
      masked_y[j] = ay;
      const float dc0 = dcv[0] * x_multiplier;
      const float ac0 = a[0] * x_multiplier;
      const float masked_dc = dc0 * mdc[j] + dcv[1] * mdc[j] + dcv[2] * mdc[j];
      const float masked_ac = ac0 * my[j] + ay * my[j] + a[2] * my[j];
      final_map[j] = sqrtf(masked_dc + masked_ac);
    }
    if constexpr (kKeepStages) store4(stages + 7 * n, base, masked_y);
    store4(stages + 8 * n, base, final_map);
  }
}

// CombineChannelsForMasking + DiffPrecompute.
__global__ __launch_bounds__(kThreads) void MaskSourceKernel(
  const float* __restrict__ ps, size_t n, float* __restrict__ out) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* hf = ps + 6 * n;
  const float* uhf = ps + 8 * n;
  constexpr float kMul = 6.19424080439f;
  const float bias = kMul * 12.61050594197f;
  const float sqrt_bias = sqrtf(bias);
  const float xd = (uhf[i] + hf[i]) * 2.5f;
  const float yd = uhf[n + i] * 0.4f + hf[n + i] * 0.4f;
  const float activity = sqrtf(xd * xd + yd * yd);
  out[i] = sqrtf(kMul * fabsf(activity) + bias) - sqrt_bias;
}

// StoreMin3 of the reference as selects: the same comparisons decide the
// same moves, also when m0 <= m1 <= m2 does not hold.
__device__ __forceinline__ void StoreMin3(float v, float& m0, float& m1, float& m2) {
  const bool below2 = v < m2, below0 = v < m0, below1 = v < m1;
  // if (v < m2) { if (v < m0) {m2 = m1; m1 = m0; m0 = v;}
  //               else if (v < m1) {m2 = m1; m1 = v;} else {m2 = v;} }
  const float n2 = below0 || below1 ? m1 : v;
  const float n1 = below0 ? m0 : (below1 ? v : m1);
  const float n0 = below0 ? v : m0;
  m2 = below2 ? n2 : m2;
  m1 = below2 ? n1 : m1;
  m0 = below2 ? n0 : m0;
}

// FuzzyErosion of the reference's blurred activity, and MaskY / MaskDcY of
// the result.
__global__ __launch_bounds__(kThreads) void MaskKernel(
  const float* __restrict__ in, int w, int h, float* __restrict__ mask,
  float* __restrict__ mask_y, float* __restrict__ mask_dc) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w) return;
  constexpr int s = 3;
  // every neighbour is read (at clamped positions) before the selection,
  // which then takes the ones inside the image in the reference's order
  const int xl = max(x - s, 0), xr = min(x + s, w - 1);
  const int yu = max(y - s, 0), yd = min(y + s, h - 1);
  auto at = [&](int xx, int yy) { return in[static_cast<size_t>(yy) * w + xx]; };
  const float c0 = at(x, y);
  const float l0 = at(xl, y), lu = at(xl, yu), ld = at(xl, yd);
  const float r0 = at(xr, y), ru = at(xr, yu), rd = at(xr, yd);
  const float u0 = at(x, yu), d0 = at(x, yd);
  const bool left = x >= s, right = x + s < w, up = y >= s, down = y + s < h;
  // a neighbour outside the image is +inf, which no StoreMin3 comparison
  // lets in (v < m2 is false for every m2), as the reference skips it
  constexpr float kSkip = __builtin_inff();
  float m0 = c0;
  float m1 = 2.0f * m0, m2 = m1;
  StoreMin3(left ? l0 : kSkip, m0, m1, m2);
  StoreMin3(left && up ? lu : kSkip, m0, m1, m2);
  StoreMin3(left && down ? ld : kSkip, m0, m1, m2);
  StoreMin3(right ? r0 : kSkip, m0, m1, m2);
  StoreMin3(right && up ? ru : kSkip, m0, m1, m2);
  StoreMin3(right && down ? rd : kSkip, m0, m1, m2);
  StoreMin3(up ? u0 : kSkip, m0, m1, m2);
  StoreMin3(down ? d0 : kSkip, m0, m1, m2);
  const float m = 0.45f * m0 + 0.3f * m1 + 0.25f * m2;
  const size_t i = static_cast<size_t>(y) * w + x;
  mask[i] = m;
  constexpr double kGlobal = 1.0 / (17.83 * 0.79079917404);
  const double vy = kGlobal * (1.0 + 2.5485944793 / (0.451936922203 * m + 0.829591754942));
  const double vd = kGlobal * (1.0 + 0.505054525019 / (3.87449418804 * m + 0.20025578522));
  mask_y[i] = static_cast<float>(vy * vy);
  mask_dc[i] = static_cast<float>(vd * vd);
}

// ---- scales ------------------------------------------------------------------

__global__ __launch_bounds__(kThreads) void SubsampleKernel(
  const float* __restrict__ in, int w, int h, float* __restrict__ out) {
  const int sw = (w + 1) / 2, sh = (h + 1) / 2;
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  const int c = blockIdx.z;
  if (x >= sw) return;
  const float* p = in + static_cast<size_t>(c) * w * h;
  const int x0 = 2 * x, y0 = 2 * y;
  // the four reads are unconditional (clamped into the image); the sum takes
  // the in-image ones in the reference's order
  const bool right = x0 + 1 < w, below = y0 + 1 < h;
  const size_t r0 = static_cast<size_t>(y0) * w, r1 = static_cast<size_t>(below ? y0 + 1 : y0) * w;
  const int x1 = right ? x0 + 1 : x0;
  float a = p[r0 + x0], b = p[r0 + x1], d = p[r1 + x0], e = p[r1 + x1];
  // (the empty asm keeps the compiler from moving a read under its select)
  asm volatile("" : "+v"(a), "+v"(b), "+v"(d), "+v"(e));
  float v = 0.0f + 0.25f * a;
  v = right ? v + 0.25f * b : v;
  const float lower = v + 0.25f * d;
  v = below ? (right ? lower + 0.25f * e : lower) : v;
  if ((w & 1) && x == sw - 1) v *= 2.0f;
  if ((h & 1) && y == sh - 1) v *= 2.0f;
  out[static_cast<size_t>(c) * sw * sh + static_cast<size_t>(y) * sw + x] = v;
}

// ExpandImage for extents below 8: edge-replicated into ew x eh.
__global__ __launch_bounds__(kThreads) void ExpandKernel(
  const float* __restrict__ in, int w, int h, int xb, int yb, int ew, int eh,
  float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  const int c = blockIdx.z;
  if (x >= ew) return;
  int sy = y > yb ? y - yb : 0;
  sy = sy < h - 1 ? sy : h - 1;
  int sx = x > xb ? x - xb : 0;
  sx = sx < w - 1 ? sx : w - 1;
  out[static_cast<size_t>(c) * ew * eh + static_cast<size_t>(y) * ew + x] =
    in[static_cast<size_t>(c) * w * h + static_cast<size_t>(sy) * w + sx];
}

// The requested-extent map: 0.85 main + 0.5 sub (two scales), the main map
// cropped from the expanded extent, or the main map alone; and its maximum
// (non-negative floats compare as their bit patterns), one atomic per group.
__global__ __launch_bounds__(kThreads) void ComposeKernel(
  const float* __restrict__ main_map, int mw, const float* __restrict__ sub_map,
  int sw, int w, int h, int xb, int yb, float* __restrict__ map,
  float* __restrict__ map_out, uint32_t* __restrict__ score_bits) {
  uint32_t best[1] = {0u};
  // rows strided over the groups, columns over the threads (no division)
  for (int y = blockIdx.x; y < h; y += gridDim.x) {
    const float* main_row = main_map + static_cast<size_t>(y + yb) * mw + xb;
    const float* sub_row = sub_map != nullptr ? sub_map + static_cast<size_t>(y / 2) * sw : nullptr;
    const size_t row = static_cast<size_t>(y) * w;
    for (int x = threadIdx.x; x < w; x += blockDim.x) {
      const float m = main_row[x];
      const float v = sub_row != nullptr ? m * 0.85f + 0.5f * sub_row[x / 2] : m;
      map[row + x] = v;
      if (map_out != nullptr) map_out[row + x] = v;
      const uint32_t bits = __float_as_uint(v);
      best[0] = bits > best[0] ? bits : best[0];
    }
  }
  BlockReduceU32(best, U32Max());
  if (threadIdx.x == 0) atomicMax(score_bits, best[0]);
}

__global__ void FillWordKernel(uint32_t* p, uint32_t value) {
  if (threadIdx.x == 0) *p = value;
}

// ---- a17 reduction ------------------------------------------------------------

// std::pow(x, 1.0f / 16.0f) as the reference's libm evaluates it.
__device__ __forceinline__ float Root16(float x) { return GlibcPowf(x, 1.0f / 16.0f); }

__global__ __launch_bounds__(kThreads) void ReduceKernel(
  const float* __restrict__ map, int w, int h, const uint8_t* __restrict__ cells,
  int bw, int bh, float* __restrict__ out) {
  const int bx = blockIdx.x * blockDim.x + threadIdx.x, by = blockIdx.y;
  if (bx >= bw) return;
  const uint8_t c = cells[static_cast<size_t>(by) * bw + bx];
  if (!(c & 1)) return;
  const int cw = CoveredW(c >> 1), ch = CoveredH(c >> 1);
  // a footprint past the grid is not a valid transform: nothing is written
  if (bx + cw > bw || by + ch > bh) return;
  const int x0 = bx * 8, y0 = by * 8;
  const int x1 = min(w, x0 + cw * 8), y1 = min(h, y0 + ch * 8);
  const int xw = x1 - x0;
  const int count = xw * max(y1 - y0, 0);
  if (count <= 0) return;
  // rows of up to 32 pixels, each read whole (16-byte reads when rows start
  // on 16-byte boundaries) one row ahead of its ordered adds (raster order)
  constexpr int kQuads = 8;
  const bool vector = (w & 3) == 0;
  auto load_row = [&](int y, float4 (&v)[kQuads]) {
    const float* row = map + static_cast<size_t>(min(y, y1 - 1)) * w + x0;
#pragma unroll
    for (int k = 0; k < kQuads; ++k) {
      if (4 * k < xw) {
        if (vector) {
          v[k] = *reinterpret_cast<const float4*>(row + 4 * k);
        } else {
          v[k].x = row[4 * k];
          v[k].y = 4 * k + 1 < xw ? row[4 * k + 1] : 0.0f;
          v[k].z = 4 * k + 2 < xw ? row[4 * k + 2] : 0.0f;
          v[k].w = 4 * k + 3 < xw ? row[4 * k + 3] : 0.0f;
        }
      } else {
        v[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      }
    }
  };
  float norm = 0.0f;
  auto add = [&](float v, bool live) {
    v *= v;
    v *= v;
    v *= v;
    v *= v;
    const float next = norm + v;
    norm = live ? next : norm;
  };
  float4 cur[kQuads], nxt[kQuads];
  load_row(y0, cur);
  for (int y = y0; y < y1; ++y) {
    load_row(y + 1, nxt);
#pragma unroll
    for (int k = 0; k < kQuads; ++k) {
      add(cur[k].x, 4 * k < xw);
      add(cur[k].y, 4 * k + 1 < xw);
      add(cur[k].z, 4 * k + 2 < xw);
      add(cur[k].w, 4 * k + 3 < xw);
    }
#pragma unroll
    for (int k = 0; k < kQuads; ++k) cur[k] = nxt[k];
  }
  const float d = 1.2f * Root16(norm / static_cast<float>(count));
  for (int dy = 0; dy < ch; ++dy)
    for (int dx = 0; dx < cw; ++dx) out[static_cast<size_t>(by + dy) * bw + bx + dx] = d;
}

// ---- host side -------------------------------------------------------------------

void HostTaps(float sigma, float* out) {
  float k[33] = {};
  const float scaled = 2.25f * std::fabs(sigma);
  int radius = static_cast<int>(scaled);
  radius = radius < 1 ? 1 : radius;
  const int size = 2 * radius + 1;
  // the host libm's exp, as the reference evaluates it; volatile keeps the
  // compiler from folding it
  volatile double exponent_scale = -1.0 / (2.0 * sigma * sigma);
  for (int i = -radius; i <= radius; ++i) {
    k[i + radius] = static_cast<float>(std::exp(exponent_scale * i * i));
  }
  float full = 0.0f;
  for (int i = 0; i < size; ++i) full += k[i];
  const float scale = 1.0f / full;
  for (int i = 0; i < 33; ++i) out[i] = k[i];
  for (int i = 0; i < 17; ++i) out[33 + i] = i <= radius ? k[i] * scale : 0.0f;
}

inline dim3 Grid2(int w, int h, int z = 1) {
  return dim3(CeilDiv(static_cast<uint32_t>(w), kThreads), h, z);
}
inline dim3 Grid1(size_t n) {
  return dim3(static_cast<uint32_t>((n + kThreads - 1) / kThreads));
}

const float* Taps(const GJXLAmdButteraugli* b, ButteraugliSigma s) {
  return b->taps + s * kTapStride;
}

template <int R, typename Epilogue>
GJXLResult Blur(hipStream_t st, const float* in, int w, int h, const float* taps,
                float* tmp, Epilogue epilogue) {
  hipLaunchKernelGGL(BlurRowsKernel<R>,
                     dim3(CeilDiv(static_cast<uint32_t>(w), kRowSpan),
                          CeilDiv(static_cast<uint32_t>(h), kRowRows)),
                     dim3(kThreads), 0, st, in, w, h, taps, tmp);
  hipLaunchKernelGGL((BlurColsKernel<R, Epilogue>),
                     dim3(CeilDiv(static_cast<uint32_t>(w), kThreads),
                          CeilDiv(static_cast<uint32_t>(h), kColOutputs)),
                     dim3(kThreads), 0, st, tmp, w, h, taps, epilogue);
  return CheckLaunch("BlurKernel");
}

// Opsin dynamics and frequency separation of `in` into `psycho`.
GJXLResult Psycho(GJXLAmdButteraugli* b, ButteraugliScale& s, const float* in,
                  float* psycho) {
  hipStream_t st = b->device->stream;
  const int w = s.width, h = s.height;
  const size_t n = s.n;
  const dim3 opsin_grid(CeilDiv(static_cast<uint32_t>(w), kOpsinTileW),
                        CeilDiv(static_cast<uint32_t>(h), kOpsinTileH));
  if (w >= 2 && h >= 2) {
    hipLaunchKernelGGL(OpsinDynamicsKernel<true>, opsin_grid, dim3(kThreads), 0, st, in, w, h,
                       b->params.intensity_target, Taps(b, kSigmaOpsin), s.xyb);
  } else {
    hipLaunchKernelGGL(OpsinDynamicsKernel<false>, opsin_grid, dim3(kThreads), 0, st, in, w,
                       h, b->params.intensity_target, Taps(b, kSigmaOpsin), s.xyb);
  }
  GJXL_AMD_TRY(CheckLaunch("OpsinDynamicsKernel"));
  // LF of Y, X, B with the MF = XYB - LF split (B's pass scales Y's LF)
  const int lf_order[3] = {1, 0, 2};
  for (int c : lf_order) {
    GJXL_AMD_TRY(Blur<16>(st, s.xyb + c * n, w, h, Taps(b, kSigmaLow), s.tmp[0],
                          LowSplit{c, s.xyb, psycho, n}));
  }
  for (int c = 0; c < 3; ++c) {
    GJXL_AMD_TRY(Blur<7>(st, psycho + (3 + c) * n, w, h, Taps(b, kSigmaHigh), s.tmp[0],
                         HighSplit{c, psycho, n}));
  }
  for (int c = 0; c < 2; ++c) {
    GJXL_AMD_TRY(Blur<3>(st, psycho + (6 + c) * n, w, h, Taps(b, kSigmaUltra), s.tmp[0],
                         UltraSplit{c, psycho, n}));
  }
  return GJXL_OK;
}

// Blurred masking activity of a psycho image into `out`.
GJXLResult MaskActivity(GJXLAmdButteraugli* b, ButteraugliScale& s, const float* psycho,
                        float* out) {
  hipStream_t st = b->device->stream;
  hipLaunchKernelGGL(MaskSourceKernel, Grid1(s.n), dim3(kThreads), 0, st, psycho, s.n,
                     s.tmp[0]);
  GJXL_AMD_TRY(CheckLaunch("MaskSourceKernel"));
  return Blur<6>(st, s.tmp[0], s.width, s.height, Taps(b, kSigmaMask), s.tmp[1],
                 StoreBlur{out});
}

constexpr double kMaltaWeights[6] = {37.0819870399, 8246.75321353, 18.7237414387,
                                     6923.99476109, 1.10039032555, 173.5};
constexpr double kMaltaNorms[6] = {130262059.556, 1009002.70582, 4498534.45232,
                                   8051.15833247, 71.7800275169, 5.0};
constexpr float kL2[9] = {400.0f, 1.50815703118f, 0.0f, 2150.0f, 10.6195433239f,
                          16.2176043152f, 29.2353797994f, 0.844626970982f,
                          0.703646627719f};

GJXLResult Differences(GJXLAmdButteraugli* b, ButteraugliScale& s, bool keep_stages) {
  hipStream_t st = b->device->stream;
  const int w = s.width, h = s.height;
  const size_t n = s.n;
  const float asym = b->params.hf_asymmetry;
  const float sqrt_asym = std::sqrt(asym);
  const double up[6] = {kMaltaWeights[0], kMaltaWeights[1], kMaltaWeights[2] * sqrt_asym,
                        kMaltaWeights[3] * sqrt_asym, kMaltaWeights[4] * asym,
                        kMaltaWeights[5] * asym};
  const double down[6] = {kMaltaWeights[0], kMaltaWeights[1], kMaltaWeights[2] / sqrt_asym,
                          kMaltaWeights[3] / sqrt_asym, kMaltaWeights[4] / asym,
                          kMaltaWeights[5] / asym};
  MaltaScales ms;
  for (int k = 0; k < 6; ++k) {
    const bool lf = k < 4;
    const double multiplier = lf ? 0.611612573796 : 0.39905817637;
    const double pre_gt = multiplier * std::sqrt(0.5f * up[k]) / (3.75 * 2.0 + 1.0);
    const double pre_lt = multiplier * std::sqrt(0.33f * down[k]) / (3.75 * 2.0 + 1.0);
    ms.m[k] = MaltaScale{static_cast<float>(pre_gt * kMaltaNorms[k]),
                         static_cast<float>(pre_lt * kMaltaNorms[k]),
                         static_cast<float>(kMaltaNorms[k])};  // gt, lt: binary32 values
  }
  L2Weights lw{};
  for (int c = 0; c < 2; ++c) {
    const float gt = kL2[c] * asym, lt = kL2[c] / asym;
    lw.asym_on[c] = !(gt == 0.0f && lt == 0.0f);
    lw.asym_gt[c] = gt * 0.8f;
    lw.asym_lt[c] = lt * 0.8f;
  }
  for (int c = 0; c < 3; ++c) {
    lw.mf[c] = kL2[3 + c];
    lw.lf[c] = kL2[6 + c];
  }
  // the difference kernel addresses a plane with 32-bit byte offsets
  if (static_cast<uint64_t>(w) * h >= (uint64_t{1} << 30)) {
    return Fail(GJXL_ERROR_UNSUPPORTED, "Butteraugli differences: plane of 2^30 or more pixels");
  }
  GJXL_AMD_TRY(MaskActivity(b, s, s.dist_psycho, s.tmp[2]));
  const dim3 tiles(CeilDiv(static_cast<uint32_t>(w), kMaltaTileW),
                   CeilDiv(static_cast<uint32_t>(h), kMaltaTileH));
  const bool probed = ProbeBegin(b->device, GJXL_AMD_KERNEL_BUTTERAUGLI_DIFFERENCES);
  if (keep_stages) {
    hipLaunchKernelGGL(DifferencesKernel<true>, tiles, dim3(kThreads), 0, st, s.ref_psycho,
                       s.dist_psycho, w, h, ms, lw, s.ref_blurred, s.tmp[2], s.mask_y,
                       s.mask_dc, b->params.x_multiplier, s.stages);
  } else {
    hipLaunchKernelGGL(DifferencesKernel<false>, tiles, dim3(kThreads), 0, st, s.ref_psycho,
                       s.dist_psycho, w, h, ms, lw, s.ref_blurred, s.tmp[2], s.mask_y,
                       s.mask_dc, b->params.x_multiplier, s.stages);
  }
  ProbeEnd(b->device, probed);
  return CheckLaunch("DifferencesKernel");
}

constexpr int kScalePlanes = 3 + 3 + 10 + 10 + 3 + 9 + 3;

}  // namespace

GJXLResult ButteraugliCompare(GJXLAmdButteraugli* b, const float* distorted,
                              float* map_out, bool keep_stages) {
  hipStream_t st = b->device->stream;
  ButteraugliScale& m = b->scale[0];
  const float* main_in = distorted;
  if (b->expanded) {
    hipLaunchKernelGGL(ExpandKernel, Grid2(m.width, m.height, 3), dim3(kThreads), 0, st,
                       distorted, b->width, b->height, b->xborder, b->yborder, m.width,
                       m.height, m.input);
    GJXL_AMD_TRY(CheckLaunch("ExpandKernel"));
    main_in = m.input;
  }
  GJXL_AMD_TRY(Psycho(b, m, main_in, m.dist_psycho));
  GJXL_AMD_TRY(Differences(b, m, keep_stages));
  const float* sub_map = nullptr;
  int sw = 0;
  if (b->has_subscale) {
    ButteraugliScale& s = b->scale[1];
    hipLaunchKernelGGL(SubsampleKernel, Grid2(s.width, s.height, 3), dim3(kThreads), 0, st,
                       distorted, b->width, b->height, s.input);
    GJXL_AMD_TRY(CheckLaunch("SubsampleKernel"));
    GJXL_AMD_TRY(Psycho(b, s, s.input, s.dist_psycho));
    GJXL_AMD_TRY(Differences(b, s, keep_stages));
    sub_map = s.stages + 8 * s.n;
    sw = s.width;
  }
  hipLaunchKernelGGL(FillWordKernel, dim3(1), dim3(64), 0, st, b->score_bits, 0u);
  hipLaunchKernelGGL(ComposeKernel,
                     dim3(ReductionGroups(static_cast<size_t>(b->width) * b->height)),
                     dim3(kThreads), 0, st,
                     m.stages + 8 * m.n, m.width, sub_map, sw, b->width, b->height,
                     b->xborder, b->yborder, b->map, map_out, b->score_bits);
  return CheckLaunch("ComposeKernel");
}

GJXLResult LaunchReduceDistanceMap(hipStream_t s, const float* map, int width,
                                   int height, const uint8_t* cells, int bw,
                                   int bh, float* out) {
  hipLaunchKernelGGL(ReduceKernel, Grid2(bw, bh), dim3(kThreads), 0, s, map, width,
                     height, cells, bw, bh, out);
  return CheckLaunch("ReduceKernel");
}

size_t ButteraugliGeometry(GJXLAmdButteraugli* b, uint32_t width, uint32_t height) {
  b->width = static_cast<int>(width);
  b->height = static_cast<int>(height);
  b->expanded = width < 8 || height < 8;
  b->has_subscale = !b->expanded && width >= 15 && height >= 15;
  b->xborder = width < 8 ? (8 - static_cast<int>(width)) / 2 : 0;
  b->yborder = height < 8 ? (8 - static_cast<int>(height)) / 2 : 0;
  const int mw = width < 8 ? 8 : static_cast<int>(width);
  const int mh = height < 8 ? 8 : static_cast<int>(height);
  b->scale[0].width = mw;
  b->scale[0].height = mh;
  b->scale[1].width = (static_cast<int>(width) + 1) / 2;
  b->scale[1].height = (static_cast<int>(height) + 1) / 2;
  size_t floats = kSigmaCount * 64 + static_cast<size_t>(width) * height + 64;
  const int scales = b->has_subscale ? 2 : 1;
  for (int k = 0; k < scales; ++k) {
    auto& s = b->scale[k];
    s.n = static_cast<size_t>(s.width) * s.height;
    floats += kScalePlanes * s.n;
  }
  return floats;
}

void CarveButteraugli(GJXLAmdButteraugli* b) {
  float* p = static_cast<float*>(b->allocation);
  b->taps = p;
  p += kSigmaCount * 64;
  b->score_bits = reinterpret_cast<uint32_t*>(p);
  p += 64;
  b->map = p;
  p += static_cast<size_t>(b->width) * b->height;
  const int scales = b->has_subscale ? 2 : 1;
  for (int k = 0; k < scales; ++k) {
    auto& s = b->scale[k];
    const size_t n = s.n;
    s.input = p; p += 3 * n;
    s.xyb = p; p += 3 * n;
    s.ref_psycho = p; p += 10 * n;
    s.dist_psycho = p; p += 10 * n;
    s.ref_blurred = p; p += n;
    s.mask_y = p; p += n;
    s.mask_dc = p; p += n;
    s.stages = p; p += 9 * n;
    for (auto& t : s.tmp) { t = p; p += n; }
  }
}

GJXLResult ResizeButteraugli(GJXLAmdButteraugli* b, uint32_t width, uint32_t height) {
  if (width == 0 || height == 0 || width > 65535 || height > 65535) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Butteraugli geometry is invalid");
  }
  GJXLAmdButteraugli probe;
  if (ButteraugliGeometry(&probe, width, height) > b->capacity_floats) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "extent exceeds the Butteraugli capacity");
  }
  // the taps stay where they are: they lead the allocation at every extent
  (void)ButteraugliGeometry(b, width, height);
  CarveButteraugli(b);
  b->prepared = false;
  return GJXL_OK;
}

}  // namespace gjxl_amd

using gjxl_amd::Check;
using gjxl_amd::Fail;

extern "C" {

GJXLResult gjxl_amd_butteraugli_create(GJXLAmdDevice* device, uint32_t width,
                                       uint32_t height, float hf_asymmetry,
                                       float x_multiplier, float intensity_target,
                                       GJXLAmdButteraugli** out) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  if (out == nullptr || *out != nullptr || width == 0 || height == 0 ||
      width > 65535 || height > 65535 || !std::isfinite(hf_asymmetry) ||
      hf_asymmetry <= 0.0f || !std::isfinite(x_multiplier) || x_multiplier <= 0.0f ||
      !std::isfinite(intensity_target) || intensity_target <= 0.0f) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Butteraugli geometry or options are invalid");
  }
  auto* b = new (std::nothrow) GJXLAmdButteraugli;
  if (b == nullptr) return Fail(GJXL_ERROR_OUT_OF_MEMORY, "Butteraugli allocation failed");
  b->device = device;
  b->params = {hf_asymmetry, x_multiplier, intensity_target};
  const size_t floats = gjxl_amd::ButteraugliGeometry(b, width, height);
  GJXLResult r = Check(hipSetDevice(device->ordinal), "hipSetDevice");
  if (r == GJXL_OK) r = gjxl_amd::DeviceAlloc(device, floats * sizeof(float), &b->allocation);
  if (r != GJXL_OK) {
    delete b;
    return r;
  }
  b->capacity_floats = floats;
  gjxl_amd::CarveButteraugli(b);
  float host[gjxl_amd::kSigmaCount * 64] = {};
  for (int k = 0; k < gjxl_amd::kSigmaCount; ++k) {
    gjxl_amd::HostTaps(gjxl_amd::kSigmas[k], host + k * 64);
  }
  r = Check(hipMemcpyAsync(b->taps, host, sizeof(host), hipMemcpyHostToDevice, device->stream),
            "hipMemcpyAsync(Butteraugli taps)");
  if (r == GJXL_OK) r = Check(hipStreamSynchronize(device->stream), "hipStreamSynchronize");
  if (r != GJXL_OK) {
    gjxl_amd_butteraugli_destroy(b);
    return r;
  }
  *out = b;
  return GJXL_OK;
}

void gjxl_amd_butteraugli_destroy(GJXLAmdButteraugli* b) {
  if (b == nullptr) return;
  if (b->device != nullptr) {
    (void)hipSetDevice(b->device->ordinal);
    (void)hipStreamSynchronize(b->device->stream);
  }
  if (b->device != nullptr) gjxl_amd::DeviceFree(b->device, b->allocation);
  delete b;
}

GJXLResult gjxl_amd_butteraugli_prepare(GJXLAmdButteraugli* b,
                                        const float* reference_linear) {
  if (b == nullptr || reference_linear == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Butteraugli prepare arguments are null");
  }
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(b->device));
  using namespace gjxl_amd;
  b->prepared = false;
  hipStream_t st = b->device->stream;
  for (int k = 0; k < (b->has_subscale ? 2 : 1); ++k) {
    ButteraugliScale& s = b->scale[k];
    const float* in = reference_linear;
    if (k == 0 && b->expanded) {
      hipLaunchKernelGGL(ExpandKernel, Grid2(s.width, s.height, 3), dim3(kThreads), 0, st,
                         reference_linear, b->width, b->height, b->xborder, b->yborder,
                         s.width, s.height, s.input);
      GJXL_AMD_TRY(CheckLaunch("ExpandKernel"));
      in = s.input;
    } else if (k == 1) {
      hipLaunchKernelGGL(SubsampleKernel, Grid2(s.width, s.height, 3), dim3(kThreads), 0,
                         st, reference_linear, b->width, b->height, s.input);
      GJXL_AMD_TRY(CheckLaunch("SubsampleKernel"));
      in = s.input;
    }
    GJXL_AMD_TRY(Psycho(b, s, in, s.ref_psycho));
    GJXL_AMD_TRY(MaskActivity(b, s, s.ref_psycho, s.ref_blurred));
    hipLaunchKernelGGL(MaskKernel, Grid2(s.width, s.height), dim3(kThreads), 0, st,
                       s.ref_blurred, s.width, s.height, s.stages + 6 * s.n, s.mask_y,
                       s.mask_dc);
    GJXL_AMD_TRY(CheckLaunch("MaskKernel"));
  }
  b->prepared = true;
  return GJXL_OK;
}

GJXLResult gjxl_amd_butteraugli_compare(GJXLAmdButteraugli* b,
                                        const float* distorted_linear,
                                        float* distance_map) {
  if (b == nullptr || distorted_linear == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Butteraugli compare arguments are null");
  }
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(b->device));
  if (!b->prepared) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Butteraugli reference has not been prepared");
  }
  return gjxl_amd::ButteraugliCompare(b, distorted_linear, distance_map, true);
}

GJXLResult gjxl_amd_butteraugli_score(GJXLAmdButteraugli* b, double* score) {
  if (b == nullptr || score == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Butteraugli score arguments are null");
  }
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(b->device));
  uint32_t bits = 0;
  GJXL_AMD_TRY(Check(hipMemcpyAsync(&bits, b->score_bits, 4, hipMemcpyDeviceToHost,
                                    b->device->stream), "hipMemcpyAsync(score)"));
  GJXL_AMD_TRY(Check(hipStreamSynchronize(b->device->stream), "hipStreamSynchronize"));
  float v;
  std::memcpy(&v, &bits, 4);
  *score = v;
  return GJXL_OK;
}

GJXLResult gjxl_amd_butteraugli_buffer(GJXLAmdButteraugli* b, int32_t scale,
                                       int32_t id, void** pointer, size_t* bytes) {
  if (b == nullptr || pointer == nullptr || scale < 0 || scale > 1 ||
      (scale == 1 && !b->has_subscale)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Butteraugli buffer request is invalid");
  }
  const gjxl_amd::ButteraugliScale& s = b->scale[scale];
  const size_t plane = s.n * sizeof(float);
  void* p = nullptr;
  size_t z = 0;
  switch (id) {
    case GJXL_AMD_BUTTERAUGLI_MAP: p = b->map; z = sizeof(float) * b->width * b->height; break;
    case GJXL_AMD_BUTTERAUGLI_REFERENCE_PSYCHO: p = s.ref_psycho; z = 10 * plane; break;
    case GJXL_AMD_BUTTERAUGLI_DISTORTED_PSYCHO: p = s.dist_psycho; z = 10 * plane; break;
    case GJXL_AMD_BUTTERAUGLI_STAGES: p = s.stages; z = 9 * plane; break;
    case GJXL_AMD_BUTTERAUGLI_XYB: p = s.xyb; z = 3 * plane; break;
    case GJXL_AMD_BUTTERAUGLI_SCORE: p = b->score_bits; z = 4; break;
    default: return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Butteraugli buffer id is invalid");
  }
  *pointer = p;
  if (bytes) *bytes = z;
  return GJXL_OK;
}

GJXLResult gjxl_amd_reduce_distance_map(GJXLAmdDevice* device, const float* map,
                                        uint32_t width, uint32_t height,
                                        const uint8_t* cells, float* block_distance) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  if (map == nullptr || cells == nullptr || block_distance == nullptr || width == 0 ||
      height == 0) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Distance-map reduction arguments are invalid");
  }
  const int bw = static_cast<int>(gjxl_amd::CeilDiv(width, 8));
  const int bh = static_cast<int>(gjxl_amd::CeilDiv(height, 8));
  return gjxl_amd::LaunchReduceDistanceMap(device->stream, map, width, height, cells, bw,
                                           bh, block_distance);
}

}  // extern "C"
