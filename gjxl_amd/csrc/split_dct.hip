// Split-precision forward DCT on the matrix cores: the chosen precision form
// of the DCT (not the plain path, whose DCTs are binary64 chains so that
// every coefficient equals the reference's).
//
// A W x H transform (W, H in {8, 16, 32, 64}) is two matrix products,
//   T = X * Fw^T   (rows: T[y][u] = sum_x X[y][x] Fw[u][x])
//   C = Fh * T     (columns: C[v][u] = sum_y Fh[v][y] T[y][u]),
// with F the forward DCT-II basis of src/codec/dct.cpp:135-182 (sqrt(2)/n,
// 1/sqrt(2) on the first row).  Each operand is split into binary16 halves,
// a = hi + lo with hi = fp16(a) and lo = fp16(a - hi), and each product is
// lo*hi + hi*lo + hi*hi on v_mfma_f32_16x16x16_f16 with binary32
// accumulation (the lo*lo term, 2^-22 of the product, is dropped).  Before
// splitting, a transform's pixels are scaled by a power of two that brings
// their maximum to [2^11, 2^12), and the basis by n (both exact), so that
// the lo halves stay normal binary16 numbers; the coefficients are scaled
// back exactly.  Dimensions of 8 are padded to 16 with zeros (exact).
//
// One wave per transform, its three channels in turn.  Tiles: the MFMA
// operand layouts of gfx950's 16x16x16 f16 instruction -- lane l holds
// A[l % 16][4 (l / 16) + i], B[4 (l / 16) + i][l % 16] and the result
// D[4 (l / 16) + i][l % 16], i = 0..3 -- so that a stage-1 result tile is
// already the stage-2 B operand of the same rows: no data exchange between
// the two products.
#include <cmath>

#include "aq_common.h"
#include "common.h"

namespace gjxl_amd {
namespace {

typedef _Float16 Half4 __attribute__((ext_vector_type(4)));
typedef float Float4 __attribute__((ext_vector_type(4)));

constexpr int kSplitWaves = 4;

__host__ __device__ constexpr int Padded(int n) { return n < 16 ? 16 : n; }

// offsets of the n-point tables (Padded(n)^2 entries each) within one of
// the hi and lo halves of the split basis
__host__ __device__ constexpr int SplitBasisOffset(int n) {
  return n == 8 ? 0 : n == 16 ? 256 : n == 32 ? 512 : 1536;
}
constexpr int kSplitBasisHalf = 1536 + 64 * 64;

// The transform's pixels: rows of the image (three planes) or of a block
// batch (one plane).
struct FrameSource {
  const float* image;
  int width;
  size_t plane;
  __device__ const float* Row(int bx, int by, int c, int y) const {
    return image + c * plane + static_cast<size_t>(by * 8 + y) * width + bx * 8;
  }
};

__device__ __forceinline__ void Split(Float4 v, Half4* hi, Half4* lo) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const _Float16 h = static_cast<_Float16>(v[i]);
    (*hi)[i] = h;
    (*lo)[i] = static_cast<_Float16>(v[i] - static_cast<float>(h));
  }
}

__device__ __forceinline__ Float4 SplitProduct(Half4 ahi, Half4 alo, Half4 bhi, Half4 blo,
                                               Float4 acc) {
  acc = __builtin_amdgcn_mfma_f32_16x16x16f16(alo, bhi, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x16f16(ahi, blo, acc, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x16f16(ahi, bhi, acc, 0, 0, 0);
}

__device__ __forceinline__ float WaveMax(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
  return v;
}

// One channel of one transform.  `row(y)` is the first pixel of row y;
// `store(i, v)` writes coefficient i (storage order) of value v.
template <int W, int H, typename Row, typename Store>
__device__ __forceinline__ void SplitTransform(const _Float16* __restrict__ basis, Row row,
                                               Store store) {
  constexpr int PW = Padded(W), PH = Padded(H);
  constexpr int TW = PW / 16, TH = PH / 16;
  const int lane = threadIdx.x & 63;
  const int r16 = lane & 15, q4 = (lane >> 4) * 4;
  const _Float16* fw_hi = basis + SplitBasisOffset(W);
  const _Float16* fw_lo = fw_hi + kSplitBasisHalf;
  const _Float16* fh_hi = basis + SplitBasisOffset(H);
  const _Float16* fh_lo = fh_hi + kSplitBasisHalf;
  // X tiles: a[mt][kt] = X[16 mt + r16][16 kt + q4 + i]
  Float4 a[TH][TW];
  float peak = 0.0f;
#pragma unroll
  for (int mt = 0; mt < TH; ++mt) {
    const int y = 16 * mt + r16;
#pragma unroll
    for (int kt = 0; kt < TW; ++kt) {
      const int x = 16 * kt + q4;
      Float4 v = {0.0f, 0.0f, 0.0f, 0.0f};
      if (y < H && x < W) {
        const float4 p = *reinterpret_cast<const float4*>(row(y) + x);
        v = Float4{p.x, p.y, p.z, p.w};
      }
      a[mt][kt] = v;
      peak = fmaxf(peak, fmaxf(fmaxf(fabsf(v[0]), fabsf(v[1])), fmaxf(fabsf(v[2]), fabsf(v[3]))));
    }
  }
  peak = WaveMax(peak);
  int exponent = 0;
  if (peak > 0.0f) {
    (void)frexpf(peak, &exponent);  // peak = m 2^exponent, m in [0.5, 1)
    exponent = 12 - exponent;       // peak 2^exponent in [2^11, 2^12)
  }
  // stage 1: t[mt][nt] = (X 2^e)(W Fw^T), tile rows y, columns u
  Float4 t[TH][TW];
#pragma unroll
  for (int mt = 0; mt < TH; ++mt) {
    Half4 ahi[TW], alo[TW];
#pragma unroll
    for (int kt = 0; kt < TW; ++kt) {
      Float4 s = a[mt][kt];
#pragma unroll
      for (int i = 0; i < 4; ++i) s[i] = ldexpf(s[i], exponent);
      Split(s, &ahi[kt], &alo[kt]);
    }
#pragma unroll
    for (int nt = 0; nt < TW; ++nt) {
      Float4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kt = 0; kt < TW; ++kt) {
        // B[x][u] = Fw[u][x]: four consecutive x of row u
        const int at = (16 * nt + r16) * PW + 16 * kt + q4;
        const Half4 bhi = *reinterpret_cast<const Half4*>(fw_hi + at);
        const Half4 blo = *reinterpret_cast<const Half4*>(fw_lo + at);
        acc = SplitProduct(ahi[kt], alo[kt], bhi, blo, acc);
      }
      t[mt][nt] = acc;
    }
  }
  // stage 2: c[mt][nt] = (H Fh)(T / W), rows v, columns u; t[kt][nt] is the
  // B operand (rows y = 16 kt + q4 + i, column u = 16 nt + r16) as it stands
  Half4 thi[TH][TW], tlo[TH][TW];
#pragma unroll
  for (int kt = 0; kt < TH; ++kt) {
#pragma unroll
    for (int nt = 0; nt < TW; ++nt) {
      Float4 s = t[kt][nt];
#pragma unroll
      for (int i = 0; i < 4; ++i) s[i] = s[i] * (1.0f / static_cast<float>(W));
      Split(s, &thi[kt][nt], &tlo[kt][nt]);
    }
  }
  const float unscale = ldexpf(1.0f / static_cast<float>(H), -exponent);
#pragma unroll
  for (int mt = 0; mt < TH; ++mt) {
    Half4 fhi[TH], flo[TH];
#pragma unroll
    for (int kt = 0; kt < TH; ++kt) {
      // A[v][y] = Fh[v][y]: four consecutive y of row v
      const int at = (16 * mt + r16) * PH + 16 * kt + q4;
      fhi[kt] = *reinterpret_cast<const Half4*>(fh_hi + at);
      flo[kt] = *reinterpret_cast<const Half4*>(fh_lo + at);
    }
#pragma unroll
    for (int nt = 0; nt < TW; ++nt) {
      Float4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kt = 0; kt < TH; ++kt) {
        acc = SplitProduct(fhi[kt], flo[kt], thi[kt][nt], tlo[kt][nt], acc);
      }
      const int u = 16 * nt + r16;
      const int v0 = 16 * mt + q4;
      if (u < W && v0 < H) {
#pragma unroll
        for (int i = 0; i < 4; ++i) store(CoefIndex(W, H, v0 + i, u), acc[i] * unscale);
      }
    }
  }
}

// Every transform of one strategy of a frame: one wave each, the three
// channels in turn, coefficients into the per-group store.
template <int kStrategy>
__global__ __launch_bounds__(64 * kSplitWaves) void SplitAnchorDctKernel(
  const float* __restrict__ opsin, int width, int height,
  const uint32_t* __restrict__ entries, uint32_t count,
  const uint32_t* __restrict__ offsets, const _Float16* __restrict__ basis,
  size_t store, float* __restrict__ coefs) {
  constexpr int W = CoveredW(kStrategy) * 8, H = CoveredH(kStrategy) * 8;
  const uint32_t t = blockIdx.x * kSplitWaves + (threadIdx.x >> 6);
  if (t >= count) return;  // whole waves
  const uint32_t packed = entries[t];
  const int bx = packed & 0xffff, by = packed >> 16;
  const uint32_t offset = offsets[static_cast<size_t>(by) * (width / 8) + bx];
  const FrameSource src{opsin, width, static_cast<size_t>(width) * height};
  for (int c = 0; c < 3; ++c) {
    float* out = coefs + c * store + offset;
    SplitTransform<W, H>(basis, [&](int y) { return src.Row(bx, by, c, y); },
                         [&](int i, float v) { out[i] = v; });
  }
}

// A batch of W x H blocks (row-major, contiguous) into coefficient blocks of
// W * H values in storage order.
template <int W, int H>
__global__ __launch_bounds__(64 * kSplitWaves) void SplitBlockDctKernel(
  const float* __restrict__ blocks, size_t count, const _Float16* __restrict__ basis,
  float* __restrict__ coefs) {
  const size_t t = static_cast<size_t>(blockIdx.x) * kSplitWaves + (threadIdx.x >> 6);
  if (t >= count) return;  // whole waves
  const float* in = blocks + t * (W * H);
  float* out = coefs + t * (W * H);
  SplitTransform<W, H>(basis, [&](int y) { return in + y * W; },
                       [&](int i, float v) { out[i] = v; });
}

template <int W, int H>
GJXLResult LaunchBlocks(GJXLAmdDevice* d, const float* blocks, size_t count, float* coefs) {
  const dim3 grid(static_cast<uint32_t>((count + kSplitWaves - 1) / kSplitWaves));
  hipLaunchKernelGGL((SplitBlockDctKernel<W, H>), grid, dim3(64 * kSplitWaves), 0, d->stream,
                     blocks, count, d->split_bases, coefs);
  return CheckLaunch("SplitBlockDctKernel");
}

template <int W>
GJXLResult DispatchHeight(GJXLAmdDevice* d, int h, const float* blocks, size_t count,
                          float* coefs) {
  switch (h) {
    case 8: return LaunchBlocks<W, 8>(d, blocks, count, coefs);
    case 16: return LaunchBlocks<W, 16>(d, blocks, count, coefs);
    case 32: return LaunchBlocks<W, 32>(d, blocks, count, coefs);
    default: return LaunchBlocks<W, 64>(d, blocks, count, coefs);
  }
}

}  // namespace

void BuildSplitBases(_Float16* out) {
  const double sqrt_two = 1.4142135623730951;
  const double pi = 3.141592653589793;
  for (int i = 0; i < 2 * kSplitBasisHalf; ++i) out[i] = static_cast<_Float16>(0.0f);
  for (int n : {8, 16, 32, 64}) {
    const int p = Padded(n);
    for (int f = 0; f < n; ++f) {
      const double alpha = f == 0 ? 1.0 / sqrt_two : 1.0;
      for (int s = 0; s < n; ++s) {
        const double angle = (static_cast<double>(s) + 0.5) * static_cast<double>(f) * pi /
                             static_cast<double>(n);
        // the basis times n (a power of two): sqrt(2) alpha cos(angle)
        const float scaled = static_cast<float>(sqrt_two / static_cast<double>(n) * alpha *
                                                std::cos(angle) * static_cast<double>(n));
        const _Float16 hi = static_cast<_Float16>(scaled);
        const _Float16 lo = static_cast<_Float16>(scaled - static_cast<float>(hi));
        out[SplitBasisOffset(n) + f * p + s] = hi;
        out[kSplitBasisHalf + SplitBasisOffset(n) + f * p + s] = lo;
      }
    }
  }
}

size_t SplitBasisHalves() { return 2 * kSplitBasisHalf; }

GJXLResult LaunchSplitAnchorDct(GJXLAmdDevice* d, const float* opsin, int width, int height,
                                const uint32_t* offsets, const StrategyLists& lists,
                                size_t store, float* coefs) {
  for (int stage = 0; stage < kStrategyStages; ++stage) {
    const uint32_t count = lists.counts[stage];
    if (count == 0) continue;
    const uint32_t* entries = lists.entries + stage * lists.capacity;
    const dim3 grid((count + kSplitWaves - 1) / kSplitWaves);
    const dim3 block(64 * kSplitWaves);
    switch (stage) {
      case 0:
        hipLaunchKernelGGL(SplitAnchorDctKernel<kDct8>, grid, block, 0, d->stream, opsin, width,
                           height, entries, count, offsets, d->split_bases, store, coefs);
        break;
      case 1:
        hipLaunchKernelGGL(SplitAnchorDctKernel<kDct16x8>, grid, block, 0, d->stream, opsin,
                           width, height, entries, count, offsets, d->split_bases, store, coefs);
        break;
      case 2:
        hipLaunchKernelGGL(SplitAnchorDctKernel<kDct8x16>, grid, block, 0, d->stream, opsin,
                           width, height, entries, count, offsets, d->split_bases, store, coefs);
        break;
      case 3:
        hipLaunchKernelGGL(SplitAnchorDctKernel<kDct16x16>, grid, block, 0, d->stream, opsin,
                           width, height, entries, count, offsets, d->split_bases, store, coefs);
        break;
      case 4:
        hipLaunchKernelGGL(SplitAnchorDctKernel<kDct32x16>, grid, block, 0, d->stream, opsin,
                           width, height, entries, count, offsets, d->split_bases, store, coefs);
        break;
      case 5:
        hipLaunchKernelGGL(SplitAnchorDctKernel<kDct16x32>, grid, block, 0, d->stream, opsin,
                           width, height, entries, count, offsets, d->split_bases, store, coefs);
        break;
      default:
        hipLaunchKernelGGL(SplitAnchorDctKernel<kDct32x32>, grid, block, 0, d->stream, opsin,
                           width, height, entries, count, offsets, d->split_bases, store, coefs);
        break;
    }
    GJXL_AMD_TRY(CheckLaunch("SplitAnchorDctKernel"));
  }
  return GJXL_OK;
}

}  // namespace gjxl_amd

extern "C" {

GJXLResult gjxl_amd_split_dct_blocks(GJXLAmdDevice* device, uint32_t width, uint32_t height,
                                     const float* blocks, size_t count, float* coefficients) {
  using namespace gjxl_amd;
  GJXL_AMD_TRY(ValidateDevice(device));
  auto valid = [](uint32_t n) { return n == 8 || n == 16 || n == 32 || n == 64; };
  if (!valid(width) || !valid(height)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "split DCT sides must be 8, 16, 32 or 64");
  }
  if (count == 0) return GJXL_OK;
  if (blocks == nullptr || coefficients == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "split DCT buffers are null");
  }
  if (count > (size_t{1} << 30)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "split DCT batch is too large");
  }
  switch (width) {
    case 8: return DispatchHeight<8>(device, static_cast<int>(height), blocks, count, coefficients);
    case 16:
      return DispatchHeight<16>(device, static_cast<int>(height), blocks, count, coefficients);
    case 32:
      return DispatchHeight<32>(device, static_cast<int>(height), blocks, count, coefficients);
    default:
      return DispatchHeight<64>(device, static_cast<int>(height), blocks, count, coefficients);
  }
}

}  // extern "C"
