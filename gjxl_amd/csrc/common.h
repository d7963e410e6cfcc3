// Shared host/device plumbing for the gjxl_amd HIP library (gfx950).
//
// Every translation unit is compiled with -ffp-contract=off: a*b+c is two
// roundings unless the code says __builtin_fmaf, matching the reference's
// CPU arithmetic (which has no FMA contraction on baseline x86-64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/gjxl_amd.h"

namespace gjxl_amd {

// ---- errors ------------------------------------------------------------

// Records this thread's diagnostic and returns `code`.
GJXLResult Fail(GJXLResult code, const std::string& message);
// Maps a HIP error to GJXL_ERROR_BACKEND with context; GJXL_OK on success.
GJXLResult Check(hipError_t error, const char* what);
// Checks the launch that was just enqueued.
GJXLResult CheckLaunch(const char* kernel);

#define GJXL_AMD_TRY(expr)                   \
  do {                                       \
    const GJXLResult gjxl_amd_result_ = (expr); \
    if (gjxl_amd_result_ != GJXL_OK) return gjxl_amd_result_; \
  } while (0)

// ---- device handle ------------------------------------------------------

}  // namespace gjxl_amd

struct GJXLAmdDevice {
  int ordinal = 0;
  hipStream_t stream = nullptr;
  hipEvent_t events[GJXL_AMD_EVENT_SLOTS] = {};
  // Read-only tables uploaded at creation (see tables.h).
  float* quant_tables = nullptr;
  double* dct_bases = nullptr;
  // binary16 hi and lo halves of the split-precision DCT bases (split_dct.hip)
  _Float16* split_bases = nullptr;
  // Device address of the initial-quant-field static workspace on this
  // device (frontend.hip), looked up at creation.
  float* initial_quant_static_scratch = nullptr;
  // Launch timing of one kernel family (gjxl_amd_device_probe).
  int32_t probe_kernel = GJXL_AMD_KERNEL_NONE;
  int32_t probe_count = 0;
  hipEvent_t probe_begin[GJXL_AMD_PROBE_SLOTS] = {};
  hipEvent_t probe_end[GJXL_AMD_PROBE_SLOTS] = {};
};

namespace gjxl_amd {

GJXLResult ValidateDevice(const GJXLAmdDevice* device);
// Device memory of the objects of a device handle (paths, sessions,
// Butteraugli, tokenizers), allocated in the handle's stream order from the
// device's memory pool.  The pool keeps freed memory for later allocations
// instead of returning it to the driver, so that creating and destroying
// objects of changing extents (the batch driver) neither calls the driver's
// allocator nor waits for the whole device.  Memory is usable by work
// enqueued on the handle's stream after the call, and by the host or other
// streams after a wait on that stream.
GJXLResult DeviceAlloc(GJXLAmdDevice* device, size_t bytes, void** out);
// Releases in stream order (after the work enqueued before it).
void DeviceFree(GJXLAmdDevice* device, void* pointer);
// Device address of frontend.hip's static initial-quant workspace on the
// current device.
GJXLResult LookUpInitialQuantStaticScratch(float** out);
// Host: the split-precision DCT bases (SplitBasisHalves() binary16 values).
void BuildSplitBases(_Float16* out);
size_t SplitBasisHalves();

// Begin and end of one launch of `kernel` on the device stream, recorded
// when that family is being probed.  Usage:
//   const bool probed = ProbeBegin(d, GJXL_AMD_KERNEL_X);
//   hipLaunchKernelGGL(...);
//   ProbeEnd(d, probed);
inline bool ProbeBegin(GJXLAmdDevice* d, int32_t kernel) {
  if (d->probe_kernel != kernel || d->probe_count >= GJXL_AMD_PROBE_SLOTS) return false;
  return hipEventRecord(d->probe_begin[d->probe_count], d->stream) == hipSuccess;
}
inline void ProbeEnd(GJXLAmdDevice* d, bool probed) {
  if (probed && hipEventRecord(d->probe_end[d->probe_count], d->stream) == hipSuccess) {
    ++d->probe_count;
  }
}

inline uint32_t CeilDiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }
inline uint32_t PadTo8(uint32_t v) { return CeilDiv(v, 8) * 8; }

// ---- device math --------------------------------------------------------

__device__ __forceinline__ uint32_t FloatBits(float v) {
  return __float_as_uint(v);
}
__device__ __forceinline__ float BitsFloat(uint32_t u) {
  return __uint_as_float(u);
}

// libjxl CubeRootAndAdd: exponent-bit seed for 1/cbrt, three Newton steps and
// one refinement, then r^2 * v + add (src/codec/color_transform.cpp:54-79).
__device__ __forceinline__ float FastCubeRootAndAdd(float value, float add) {
  const float one_third = 1.0f / 3.0f;
  const float four_thirds = 4.0f / 3.0f;
  const uint32_t bits = FloatBits(value);
  const uint32_t seed =
    bits == 0 ? 0u : 0x54800000u - (bits >> 23) * 0x002AAAAAu;
  float r = BitsFloat(seed);
  const float divided = one_third * value;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float sq = r * r;
    r = __builtin_fmaf(-divided, sq * sq, four_thirds * r);
  }
  float sq = r * r;
  r = __builtin_fmaf(one_third, __builtin_fmaf(-value, sq * sq, r), r);
  sq = r * r;
  return __builtin_fmaf(sq, value, add);
}

// libjxl FastLog2f rational polynomial (src/util/fast_math.h:14-38); the
// caller guarantees a finite positive input.
__device__ __forceinline__ float FastLog2(float value) {
  const uint32_t bits = FloatBits(value);
  const int32_t e = static_cast<int32_t>(bits - 0x3f2aaaabu) >> 23;
  const float x = BitsFloat(bits - (static_cast<uint32_t>(e) << 23)) - 1.0f;
  float num = __builtin_fmaf(0.74245873327820566f, x, 1.4287160470083755f);
  num = __builtin_fmaf(num, x, -1.8503833400518310e-06f);
  float den = __builtin_fmaf(0.17409343003366853f, x, 1.0096718572241148f);
  den = __builtin_fmaf(den, x, 0.99032814277590719f);
  return num / den + static_cast<float>(e);
}

// libjxl FastPow2f rational polynomial (src/util/fast_math.h:40-70); the
// caller guarantees an input in [-126, 127].
__device__ __forceinline__ float FastPow2(float value) {
  const float fl = floorf(value);
  const int32_t e = static_cast<int32_t>(fl) + 127;
  const float ev = BitsFloat(static_cast<uint32_t>(e) << 23);
  const float f = value - fl;
  float num = f + 1.01749063e+01f;
  num = __builtin_fmaf(num, f, 4.88687798e+01f);
  num = __builtin_fmaf(num, f, 9.85506591e+01f);
  num *= ev;
  float den = __builtin_fmaf(f, 2.10242958e-01f, -2.22328856e-02f);
  den = __builtin_fmaf(den, f, -1.94414990e+01f);
  den = __builtin_fmaf(den, f, 9.85506633e+01f);
  return num / den;
}

__device__ __forceinline__ float FastPow(float base, float exponent) {
  return FastPow2(FastLog2(base) * exponent);
}

// Work-group reduction of K words with `op` (max or min of unsigned bit
// patterns); the result is valid in thread 0.  blockDim.x must be a multiple
// of 64 and at most 1024; every thread of the group must call it.
template <int K, typename Op>
__device__ inline void BlockReduceU32(uint32_t (&v)[K], Op op) {
  __shared__ uint32_t partial[16][K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    for (int off = 32; off > 0; off >>= 1) v[k] = op(v[k], __shfl_xor(v[k], off));
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) partial[wave][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < static_cast<int>(blockDim.x >> 6); ++w) {
#pragma unroll
      for (int k = 0; k < K; ++k) v[k] = op(v[k], partial[w][k]);
    }
  }
}

// A work-group barrier that orders LDS accesses only: global loads and stores
// issued before it stay in flight (a thread's own global traffic needs no
// barrier, and waiting for it at every step of a wavefront would expose its
// full latency).
__device__ __forceinline__ void LdsBarrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Exclusive prefix sum of one word per thread over the work group (thread
// order), with the group total in *total.  blockDim.x must be a multiple of
// 64 and at most 1024; every thread of the group must call it.
__device__ inline uint32_t BlockExclusiveScanU32(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wave_sums[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t inclusive = v;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t other = __shfl_up(inclusive, off);
    if (lane >= off) inclusive += other;
  }
  if (lane == 63) wave_sums[wave] = inclusive;
  __syncthreads();
  uint32_t before = 0, sum = 0;
  const int waves = static_cast<int>(blockDim.x >> 6);
  for (int w = 0; w < waves; ++w) {
    const uint32_t s = wave_sums[w];
    if (w < wave) before += s;
    sum += s;
  }
  __syncthreads();  // wave_sums may be reused by the next call
  *total = sum;
  return before + inclusive - v;
}

// XCD-aware tile order for a 2-D grid.  Work groups are dispatched to the 8
// XCDs round-robin (group i to XCD i % 8), and each XCD has its own L2; this
// maps the groups of one XCD to a contiguous run of tiles (row-major), so
// that neighbouring tiles, which share halo cache lines, are processed by
// the same L2.  The last (count % 8) groups keep their own index.
struct TileIndex {
  int x, y;
};
__device__ __forceinline__ TileIndex XcdTile() {
  constexpr unsigned kXcds = 8;
  const unsigned count = gridDim.x * gridDim.y;
  const unsigned i = blockIdx.y * gridDim.x + blockIdx.x;
  const unsigned per = count / kXcds;
  const unsigned t = i < per * kXcds ? (i % kXcds) * per + i / kXcds : i;
  return {static_cast<int>(t % gridDim.x), static_cast<int>(t / gridDim.x)};
}

struct U32Max {
  __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a < b ? b : a; }
};
struct U32Min {
  __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return b < a ? b : a; }
};

// Work groups for a grid-stride reduction over n items.
inline uint32_t ReductionGroups(size_t n, uint32_t threads = 256) {
  const size_t groups = (n + threads - 1) / threads;
  return static_cast<uint32_t>(groups < 1024 ? (groups < 1 ? 1 : groups) : 1024);
}

}  // namespace gjxl_amd
