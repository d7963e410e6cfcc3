// f2: sparse read-back packing of quantized AC coefficients, in the format
// of the reference's LaunchCudaSparseAc (src/gpu/cuda/cuda_sparse_ac_kernels
// .h:11-17, .cu:12-46) and SparseCoefficientSpan (src/codec/
// sparse_coefficients.h:90-107): per 64 logical values one mask of the
// nonzero positions and the absolute offset of that word's first payload;
// the payloads of a word keep coefficient order, the words of one work
// group are contiguous, and work groups take their payload ranges from one
// counter (so ranges may be in any order).  On 64-lane waves one ballot is
// one mask word.
#include "common.h"

namespace gjxl_amd {
namespace {

constexpr int kSparseThreads = 256;
constexpr int kSparseWaves = kSparseThreads / 64;

__global__ void ClearWordKernel(uint32_t* __restrict__ word) {
  if (threadIdx.x == 0) *word = 0;
}

template <typename T>
__global__ __launch_bounds__(kSparseThreads) void SparseAcKernel(
  const T* __restrict__ source, uint32_t count, uint64_t* __restrict__ masks,
  uint32_t* __restrict__ offsets, T* __restrict__ values, uint32_t* __restrict__ total) {
  __shared__ uint32_t prefix[kSparseWaves], base;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t index = static_cast<uint64_t>(blockIdx.x) * kSparseThreads + threadIdx.x;
  const T value = index < count ? source[index] : T{0};
  const uint64_t mask = __ballot(value != 0);
  if (lane == 0) prefix[wave] = static_cast<uint32_t>(__popcll(mask));
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t sum = 0;
    for (int w = 0; w < kSparseWaves; ++w) {
      const uint32_t n = prefix[w];
      prefix[w] = sum;
      sum += n;
    }
    base = sum != 0 ? atomicAdd(total, sum) : 0;
  }
  __syncthreads();
  const uint64_t word = static_cast<uint64_t>(blockIdx.x) * kSparseWaves + wave;
  if (lane == 0 && word < (static_cast<uint64_t>(count) + 63) / 64) {
    masks[word] = mask;
    offsets[word] = base + prefix[wave];
  }
  if (value != 0) {
    const uint64_t below = lane == 0 ? 0 : mask & (~uint64_t{0} >> (64 - lane));
    values[base + prefix[wave] + __popcll(below)] = value;
  }
}

}  // namespace
}  // namespace gjxl_amd

extern "C" {

GJXLResult gjxl_amd_sparse_ac(GJXLAmdDevice* device, const void* source, uint32_t count,
                              uint32_t width, uint64_t* masks, uint32_t* offsets, void* values,
                              uint32_t* total) {
  using namespace gjxl_amd;
  GJXL_AMD_TRY(ValidateDevice(device));
  if (width != 1 && width != 2 && width != 4) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "sparse AC width must be 1, 2 or 4 bytes");
  }
  if (source == nullptr || masks == nullptr || offsets == nullptr || values == nullptr ||
      total == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "sparse AC buffers must be non-null");
  }
  hipStream_t st = device->stream;
  hipLaunchKernelGGL(ClearWordKernel, dim3(1), dim3(64), 0, st, total);
  GJXL_AMD_TRY(CheckLaunch("ClearWordKernel"));
  if (count == 0) return GJXL_OK;
  const dim3 grid(CeilDiv(count, kSparseThreads)), block(kSparseThreads);
  if (width == 1) {
    hipLaunchKernelGGL(SparseAcKernel<int8_t>, grid, block, 0, st,
                       static_cast<const int8_t*>(source), count, masks, offsets,
                       static_cast<int8_t*>(values), total);
  } else if (width == 2) {
    hipLaunchKernelGGL(SparseAcKernel<int16_t>, grid, block, 0, st,
                       static_cast<const int16_t*>(source), count, masks, offsets,
                       static_cast<int16_t*>(values), total);
  } else {
    hipLaunchKernelGGL(SparseAcKernel<int32_t>, grid, block, 0, st,
                       static_cast<const int32_t*>(source), count, masks, offsets,
                       static_cast<int32_t*>(values), total);
  }
  return CheckLaunch("SparseAcKernel");
}

}  // extern "C"
