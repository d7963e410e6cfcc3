// Device handle, memory, copies, events and thread-local diagnostics.
// Replaces the allocation/copy/submission part of GpuBackend
// (src/gpu/backend.h:24-105) for the perceptual-analysis path.
#include <cstdio>
#include <new>
#include <string>
#include <vector>

#include "common.h"
#include "tables.h"

namespace gjxl_amd {
namespace {
thread_local std::string g_last_error;
}  // namespace

GJXLResult Fail(GJXLResult code, const std::string& message) {
  g_last_error = message;
  return code;
}

GJXLResult Check(hipError_t error, const char* what) {
  if (error == hipSuccess) return GJXL_OK;
  const GJXLResult code = error == hipErrorOutOfMemory
    ? GJXL_ERROR_OUT_OF_MEMORY : GJXL_ERROR_BACKEND;
  return Fail(code, std::string(what) + ": " + hipGetErrorString(error));
}

GJXLResult CheckLaunch(const char* kernel) {
  return Check(hipGetLastError(), kernel);
}

GJXLResult ValidateDevice(const GJXLAmdDevice* device) {
  if (device == nullptr || device->stream == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Device handle is null");
  }
  return Check(hipSetDevice(device->ordinal), "hipSetDevice");
}

GJXLResult DeviceAlloc(GJXLAmdDevice* device, size_t bytes, void** out) {
  *out = nullptr;
  return Check(hipMallocAsync(out, bytes, device->stream), "hipMallocAsync");
}

void DeviceFree(GJXLAmdDevice* device, void* pointer) {
  if (pointer != nullptr) (void)hipFreeAsync(pointer, device->stream);
}

}  // namespace gjxl_amd

using gjxl_amd::Check;
using gjxl_amd::Fail;

extern "C" {

GJXLResult gjxl_amd_device_create(int ordinal, GJXLAmdDevice** out) {
  if (out == nullptr || *out != nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                "Device output must be non-null and empty");
  }
  int count = 0;
  GJXL_AMD_TRY(Check(hipGetDeviceCount(&count), "hipGetDeviceCount"));
  if (ordinal < 0 || ordinal >= count) {
    return Fail(GJXL_ERROR_UNAVAILABLE, "HIP device ordinal is unavailable");
  }
  GJXL_AMD_TRY(Check(hipSetDevice(ordinal), "hipSetDevice"));
  auto* device = new (std::nothrow) GJXLAmdDevice;
  if (device == nullptr) {
    return Fail(GJXL_ERROR_OUT_OF_MEMORY, "Unable to allocate device handle");
  }
  device->ordinal = ordinal;
  GJXLResult result = Check(
    hipStreamCreateWithFlags(&device->stream, hipStreamNonBlocking),
    "hipStreamCreateWithFlags");
  for (int i = 0; result == GJXL_OK && i < GJXL_AMD_EVENT_SLOTS; ++i) {
    result = Check(hipEventCreate(&device->events[i]), "hipEventCreate");
  }
  if (result == GJXL_OK) {
    // the device's pool keeps what objects free for the next allocation
    hipMemPool_t pool = nullptr;
    result = Check(hipDeviceGetDefaultMemPool(&pool, ordinal), "hipDeviceGetDefaultMemPool");
    uint64_t keep = ~uint64_t{0};
    if (result == GJXL_OK) {
      result = Check(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep),
                     "hipMemPoolSetAttribute");
    }
  }
  if (result == GJXL_OK) {
    // with this ordinal current: each device has its own copy of the array
    result = gjxl_amd::LookUpInitialQuantStaticScratch(&device->initial_quant_static_scratch);
  }
  if (result == GJXL_OK) {
    result = Check(hipMalloc(reinterpret_cast<void**>(&device->quant_tables),
                             sizeof(float) * gjxl_amd::kQuantTableFloats),
                   "hipMalloc(quant tables)");
  }
  if (result == GJXL_OK) {
    result = Check(hipMalloc(reinterpret_cast<void**>(&device->dct_bases),
                             sizeof(double) * gjxl_amd::kBasisDoubles),
                   "hipMalloc(DCT bases)");
  }
  if (result == GJXL_OK) {
    result = Check(hipMalloc(reinterpret_cast<void**>(&device->split_bases),
                             sizeof(_Float16) * gjxl_amd::SplitBasisHalves()),
                   "hipMalloc(split DCT bases)");
  }
  if (result == GJXL_OK) {
    static std::vector<_Float16> split(gjxl_amd::SplitBasisHalves());
    static const bool split_built = [] {
      gjxl_amd::BuildSplitBases(split.data());
      return true;
    }();
    (void)split_built;
    result = Check(hipMemcpy(device->split_bases, split.data(), sizeof(_Float16) * split.size(),
                             hipMemcpyHostToDevice), "hipMemcpy(split DCT bases)");
  }
  if (result == GJXL_OK) {
    static float tables[gjxl_amd::kQuantTableFloats];
    static double bases[gjxl_amd::kBasisDoubles];
    static const bool built = [] {
      gjxl_amd::BuildQuantTables(tables);
      gjxl_amd::BuildDctBases(bases);
      return true;
    }();
    (void)built;
    result = Check(hipMemcpy(device->quant_tables, tables, sizeof(tables),
                             hipMemcpyHostToDevice), "hipMemcpy(quant tables)");
    if (result == GJXL_OK) {
      result = Check(hipMemcpy(device->dct_bases, bases, sizeof(bases),
                               hipMemcpyHostToDevice), "hipMemcpy(DCT bases)");
    }
  }
  if (result != GJXL_OK) {
    gjxl_amd_device_destroy(device);
    return result;
  }
  *out = device;
  return GJXL_OK;
}

void gjxl_amd_device_destroy(GJXLAmdDevice* device) {
  if (device == nullptr) return;
  (void)hipSetDevice(device->ordinal);
  if (device->stream != nullptr) (void)hipStreamSynchronize(device->stream);
  for (hipEvent_t& event : device->events) {
    if (event != nullptr) (void)hipEventDestroy(event);
  }
  for (int i = 0; i < GJXL_AMD_PROBE_SLOTS; ++i) {
    if (device->probe_begin[i] != nullptr) (void)hipEventDestroy(device->probe_begin[i]);
    if (device->probe_end[i] != nullptr) (void)hipEventDestroy(device->probe_end[i]);
  }
  if (device->quant_tables != nullptr) (void)hipFree(device->quant_tables);
  if (device->dct_bases != nullptr) (void)hipFree(device->dct_bases);
  if (device->split_bases != nullptr) (void)hipFree(device->split_bases);
  if (device->stream != nullptr) (void)hipStreamDestroy(device->stream);
  delete device;
}

GJXLResult gjxl_amd_alloc(GJXLAmdDevice* device, size_t bytes, void** out) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  if (out == nullptr || bytes == 0) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Allocation size or output is invalid");
  }
  *out = nullptr;
  return Check(hipMalloc(out, bytes), "hipMalloc");
}

GJXLResult gjxl_amd_free(GJXLAmdDevice* device, void* pointer) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  if (pointer == nullptr) return GJXL_OK;
  GJXL_AMD_TRY(Check(hipStreamSynchronize(device->stream), "hipStreamSynchronize"));
  return Check(hipFree(pointer), "hipFree");
}

GJXLResult gjxl_amd_upload(GJXLAmdDevice* device, void* dst, const void* src,
                           size_t bytes) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  if (bytes == 0) return GJXL_OK;
  if (dst == nullptr || src == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Upload pointer is null");
  }
  GJXL_AMD_TRY(Check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice,
                                    device->stream), "hipMemcpyAsync(H2D)"));
  return Check(hipStreamSynchronize(device->stream), "hipStreamSynchronize");
}

GJXLResult gjxl_amd_download(GJXLAmdDevice* device, void* dst, const void* src,
                             size_t bytes) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  if (bytes == 0) return GJXL_OK;
  if (dst == nullptr || src == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Download pointer is null");
  }
  GJXL_AMD_TRY(Check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost,
                                    device->stream), "hipMemcpyAsync(D2H)"));
  return Check(hipStreamSynchronize(device->stream), "hipStreamSynchronize");
}

GJXLResult gjxl_amd_copy(GJXLAmdDevice* device, void* dst, const void* src,
                         size_t bytes) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  if (bytes == 0) return GJXL_OK;
  if (dst == nullptr || src == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Copy pointer is null");
  }
  return Check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, device->stream),
               "hipMemcpyAsync(D2D)");
}

GJXLResult gjxl_amd_synchronize(GJXLAmdDevice* device) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  return Check(hipStreamSynchronize(device->stream), "hipStreamSynchronize");
}

const char* gjxl_amd_last_error(void) {
  return gjxl_amd::g_last_error.c_str();
}

GJXLResult gjxl_amd_event_record(GJXLAmdDevice* device, int slot) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  if (slot < 0 || slot >= GJXL_AMD_EVENT_SLOTS) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Event slot is out of range");
  }
  return Check(hipEventRecord(device->events[slot], device->stream),
               "hipEventRecord");
}

GJXLResult gjxl_amd_event_elapsed(GJXLAmdDevice* device, int begin_slot,
                                  int end_slot, float* milliseconds) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  if (begin_slot < 0 || begin_slot >= GJXL_AMD_EVENT_SLOTS || end_slot < 0 ||
      end_slot >= GJXL_AMD_EVENT_SLOTS || milliseconds == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Event slot is out of range");
  }
  GJXL_AMD_TRY(Check(hipEventSynchronize(device->events[end_slot]),
                     "hipEventSynchronize"));
  return Check(hipEventElapsedTime(milliseconds, device->events[begin_slot],
                                   device->events[end_slot]),
               "hipEventElapsedTime");
}

GJXLResult gjxl_amd_device_probe(GJXLAmdDevice* device, int32_t kernel) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  if (kernel < GJXL_AMD_KERNEL_NONE || kernel >= GJXL_AMD_KERNEL_COUNT) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Probe kernel id is out of range");
  }
  // events are created on first use (timing enabled: elapsed times are read)
  for (int i = 0; kernel != GJXL_AMD_KERNEL_NONE && i < GJXL_AMD_PROBE_SLOTS; ++i) {
    if (device->probe_begin[i] == nullptr) {
      GJXL_AMD_TRY(Check(hipEventCreate(&device->probe_begin[i]), "hipEventCreate(probe)"));
    }
    if (device->probe_end[i] == nullptr) {
      GJXL_AMD_TRY(Check(hipEventCreate(&device->probe_end[i]), "hipEventCreate(probe)"));
    }
  }
  device->probe_kernel = kernel;
  device->probe_count = 0;
  return GJXL_OK;
}

GJXLResult gjxl_amd_device_probe_times(GJXLAmdDevice* device, float* milliseconds,
                                       int32_t capacity, int32_t* count) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  if (count == nullptr || capacity < 0 || (capacity > 0 && milliseconds == nullptr)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Probe time outputs are invalid");
  }
  const int n = device->probe_count;
  if (n > 0) {
    GJXL_AMD_TRY(Check(hipEventSynchronize(device->probe_end[n - 1]), "hipEventSynchronize"));
  }
  for (int i = 0; i < n && i < capacity; ++i) {
    GJXL_AMD_TRY(Check(hipEventElapsedTime(&milliseconds[i], device->probe_begin[i],
                                           device->probe_end[i]), "hipEventElapsedTime"));
  }
  *count = n < capacity ? n : capacity;
  device->probe_count = 0;
  return GJXL_OK;
}

}  // extern "C"
