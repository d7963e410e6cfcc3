// C ABI of the adaptive-quantization evaluation: a session owns the device
// storage of one image's evaluation (the counterpart of the reference's
// PreparedAqEvaluation, src/gpu/ops/aq_evaluation.h:219-420).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "aq_common.h"
#include "butteraugli.h"
#include "common.h"


struct GJXLAmdAqSession {
  GJXLAmdDevice* device = nullptr;
  uint32_t width = 0, height = 0, pw = 0, ph = 0;
  gjxl_amd::AqLayout layout{};
  // borrowed inputs
  const float* opsin = nullptr;
  const float* pre = nullptr;
  const uint8_t* cells = nullptr;
  float target = 1.0f;
  // policy (host)
  float lower = 0.0f, upper = 0.0f, quant_dc = 0.0f;
  int x_scale = 2, b_scale = 2;
  float x_mult = 1.0f, b_mult = 1.0f;
  bool prepared = false;
  uint32_t strategy_counts[gjxl_amd::kStrategyStages] = {};
  // owned device storage, indexed by GJXL_AMD_AQ_* ids
  void* buffers[GJXL_AMD_AQ_BUFFER_COUNT] = {};
  size_t sizes[GJXL_AMD_AQ_BUFFER_COUNT] = {};     // at the current extent
  size_t capacity[GJXL_AMD_AQ_BUFFER_COUNT] = {};  // as allocated
};

using gjxl_amd::Check;
using gjxl_amd::Fail;

namespace {

template <typename T>
T* Buf(GJXLAmdAqSession* s, int id) {
  return static_cast<T*>(s->buffers[id]);
}

gjxl_amd::StrategyLists Lists(GJXLAmdAqSession* s) {
  gjxl_amd::StrategyLists l{};
  l.entries = Buf<uint32_t>(s, GJXL_AMD_AQ_STRATEGY_LISTS);
  l.capacity = static_cast<size_t>(s->layout.bw) * s->layout.bh;
  for (int i = 0; i < gjxl_amd::kStrategyStages; ++i) l.counts[i] = s->strategy_counts[i];
  return l;
}

GJXLResult ValidateSession(GJXLAmdAqSession* s) {
  if (s == nullptr) return Fail(GJXL_ERROR_INVALID_ARGUMENT, "AQ session is null");
  return gjxl_amd::ValidateDevice(s->device);
}

constexpr int kMaxIterations = 62;  // scores hold iterations + 1 words
constexpr int kErrorWord = 7;       // in the stats buffer

__global__ void CopyWordKernel(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst) {
  if (threadIdx.x == 0) *dst = *src;
}

__global__ void ZeroWordKernel(uint32_t* __restrict__ dst) {
  if (threadIdx.x == 0) *dst = 0;
}

__global__ void ZeroWordsKernel(uint32_t* __restrict__ p, size_t n) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    p[i] = 0;
  }
}

__global__ void CopyFloatsKernel(const float* __restrict__ src, size_t n,
                                 float* __restrict__ dst) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// One update of RunAdaptiveQuantizationPolicyImpl
// (src/codec/adaptive_quantization.cpp:1686-1733) per block.  x^0.2 is the
// binary64 power rounded to binary32, as std::pow(float, double) is.
__global__ __launch_bounds__(256) void AqUpdateKernel(
  float* __restrict__ field, const float* __restrict__ initial,
  const float* __restrict__ distance, size_t n, float target, int iteration,
  float lower, float upper, const gjxl_amd::QuantizerState* __restrict__ q,
  uint32_t* __restrict__ error) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float f = field[i];
  if (iteration == 1) {
    const float clamp = 0.4f * f + 0.6f * initial[i];
    if (f < clamp) f = clamp < lower ? lower : (upper < clamp ? upper : clamp);
  }
  const float difference = distance[i] / target;
  if (!isfinite(difference) || difference < 0.0f) {
    atomicOr(error, 1u);
    return;
  }
  if (difference <= 1.0f) {
    if (iteration < 2) f *= static_cast<float>(pow(static_cast<double>(difference), 0.2));
  } else {
    const float old = f;
    f *= difference;
    const float igs = q->inverse_global_scale;
    // lround: nearest, halves away from zero
    if (roundf(old * igs) == roundf(f * igs)) f = old + q->scale;
  }
  field[i] = f < lower ? lower : (upper < f ? upper : f);
}

// src/codestream/workflow.cpp:1435-1484 SelectQuantizationMatrixScales
void SelectMatrixScales(const float stats[3], float target, int* xs, int* bs) {
  int x = 3, b = 2;
  for (float threshold : {2.5f, 5.5f, 9.5f}) {
    if (target > threshold) ++x;
  }
  int xa = 0;
  if (stats[0] >= 0.026f) xa = 3;
  else if (stats[0] >= 0.022f) xa = 2;
  else if (stats[0] >= 0.015f) xa = 1;
  if (x < 2 + xa) x = 2 + xa;
  const int eb = stats[2] >= 0.13f ? 1 : 0;
  int ba = 0;
  if (stats[1] > 0.38f) ba = 2 + eb;
  else if (stats[1] > 0.33f) ba = 1 + eb;
  else if (stats[1] > 0.28f) ba = eb;
  *xs = x;
  *bs = b + ba;
}

}  // namespace

namespace {

// The extent fields of a session and the byte size of each of its buffers
// at that extent (every size grows with the width and with the height).
void SessionGeometry(GJXLAmdAqSession* s, uint32_t width, uint32_t height, size_t* z) {
  s->width = width;
  s->height = height;
  s->pw = gjxl_amd::PadTo8(width);
  s->ph = gjxl_amd::PadTo8(height);
  s->layout = gjxl_amd::MakeAqLayout(s->pw, s->ph);
  const size_t n = static_cast<size_t>(s->layout.bw) * s->layout.bh;
  const size_t tiles = static_cast<size_t>(gjxl_amd::CeilDiv(s->layout.bw, 8)) *
                       gjxl_amd::CeilDiv(s->layout.bh, 8);
  const size_t store = s->layout.store;
  const size_t groups = s->layout.groups;
  const size_t pixels = static_cast<size_t>(width) * height;
  const size_t padded = static_cast<size_t>(s->pw) * s->ph;
  z[GJXL_AMD_AQ_OFFSETS] = 4 * n;
  z[GJXL_AMD_AQ_ANCHOR_LIST] = 4 * groups * gjxl_amd::kGroupAnchors;
  z[GJXL_AMD_AQ_ANCHOR_COUNT] = 4 * groups;
  z[GJXL_AMD_AQ_COEFFICIENTS] = 4 * 3 * store;
  z[GJXL_AMD_AQ_ADJUSTED_FIELD] = 4 * n;
  z[GJXL_AMD_AQ_STATS] = 4 * 16;
  z[GJXL_AMD_AQ_SELECT] = 4 * (4 + 256);
  z[GJXL_AMD_AQ_QUANTIZER] = sizeof(gjxl_amd::QuantizerState);
  z[GJXL_AMD_AQ_RAW_QUANT] = 4 * n;
  z[GJXL_AMD_AQ_Y_TO_X] = tiles;
  z[GJXL_AMD_AQ_Y_TO_B] = tiles;
  z[GJXL_AMD_AQ_QUANTIZED] = 4 * 3 * store;
  z[GJXL_AMD_AQ_DC] = 4 * 3 * n;
  z[GJXL_AMD_AQ_DC_QUANTIZED] = 4 * 3 * n;
  z[GJXL_AMD_AQ_DC_RECONSTRUCTED] = 4 * 3 * n;
  z[GJXL_AMD_AQ_DC_SMOOTHED] = 4 * 3 * n;
  z[GJXL_AMD_AQ_INVERSE_SIGMA] = 4 * n;
  z[GJXL_AMD_AQ_RECONSTRUCTED] = 4 * 3 * padded;
  z[GJXL_AMD_AQ_FILTERED] = 4 * 3 * pixels;
  z[GJXL_AMD_AQ_LINEAR] = 4 * 3 * pixels;
  z[GJXL_AMD_AQ_FIELD] = 4 * n;
  z[GJXL_AMD_AQ_BLOCK_DISTANCE] = 4 * n;
  z[GJXL_AMD_AQ_SCORES] = 4 * 64;
  z[GJXL_AMD_AQ_STRATEGY_LISTS] = 4 * gjxl_amd::kStrategyStages * n;
}

// Zeroes the coefficient stores (their unused group tails must read zero)
// and the statistics words on the device stream, then waits.
GJXLResult ClearStores(GJXLAmdAqSession* s) {
  GJXLResult r = GJXL_OK;
  for (int id : {GJXL_AMD_AQ_COEFFICIENTS, GJXL_AMD_AQ_QUANTIZED, GJXL_AMD_AQ_STATS}) {
    if (r != GJXL_OK) break;
    hipLaunchKernelGGL(ZeroWordsKernel, dim3(1024), dim3(256), 0, s->device->stream,
                       static_cast<uint32_t*>(s->buffers[id]), s->sizes[id] / 4);
    r = gjxl_amd::CheckLaunch("ZeroWordsKernel");
  }
  if (r == GJXL_OK) r = Check(hipStreamSynchronize(s->device->stream), "hipStreamSynchronize");
  return r;
}

}  // namespace

namespace gjxl_amd {

GJXLResult ResizeAqSession(GJXLAmdAqSession* s, uint32_t width, uint32_t height) {
  if (width == 0 || height == 0 || width > (1u << 16) || height > (1u << 16)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "AQ session geometry is invalid");
  }
  GJXLAmdAqSession probe;
  size_t need[GJXL_AMD_AQ_BUFFER_COUNT] = {};
  SessionGeometry(&probe, width, height, need);
  for (int i = 0; i < GJXL_AMD_AQ_BUFFER_COUNT; ++i) {
    if (need[i] > s->capacity[i]) {
      return Fail(GJXL_ERROR_INVALID_ARGUMENT, "extent exceeds the AQ session's capacity");
    }
  }
  SessionGeometry(s, width, height, s->sizes);
  s->prepared = false;
  s->opsin = s->pre = nullptr;
  s->cells = nullptr;
  return ClearStores(s);
}

// The quantizing part of an evaluation: quantizer, final CfL, coefficient
// coding and DC (what tokenization reads); the session is validated and
// prepared.
GJXLResult AqQuantize(GJXLAmdAqSession* s, const float* quant_field) {
  GJXLAmdDevice* d = s->device;
  hipStream_t st = d->stream;
  const size_t n = static_cast<size_t>(s->layout.bw) * s->layout.bh;
  uint32_t* select = Buf<uint32_t>(s, GJXL_AMD_AQ_SELECT);
  auto* q = Buf<QuantizerState>(s, GJXL_AMD_AQ_QUANTIZER);
  GJXL_AMD_TRY(LaunchQuantizer(st, quant_field, n, s->quant_dc, select, select + 4, q,
                               Buf<int32_t>(s, GJXL_AMD_AQ_RAW_QUANT)));
  GJXL_AMD_TRY(LaunchFinalCfl(d, Buf<float>(s, GJXL_AMD_AQ_COEFFICIENTS), s->pw, s->ph,
                              s->cells, Buf<uint32_t>(s, GJXL_AMD_AQ_OFFSETS),
                              Buf<int32_t>(s, GJXL_AMD_AQ_RAW_QUANT), q,
                              Buf<int8_t>(s, GJXL_AMD_AQ_Y_TO_X),
                              Buf<int8_t>(s, GJXL_AMD_AQ_Y_TO_B)));
  GJXL_AMD_TRY(LaunchCodeCoefficients(
    d, Buf<float>(s, GJXL_AMD_AQ_COEFFICIENTS), s->pw, s->ph, s->cells,
    Buf<uint32_t>(s, GJXL_AMD_AQ_OFFSETS), Lists(s), Buf<int32_t>(s, GJXL_AMD_AQ_RAW_QUANT),
    Buf<int8_t>(s, GJXL_AMD_AQ_Y_TO_X), Buf<int8_t>(s, GJXL_AMD_AQ_Y_TO_B), q,
    s->x_mult, s->b_mult, Buf<int32_t>(s, GJXL_AMD_AQ_QUANTIZED),
    Buf<float>(s, GJXL_AMD_AQ_DC), Buf<int32_t>(s, GJXL_AMD_AQ_DC_QUANTIZED),
    Buf<float>(s, GJXL_AMD_AQ_DC_RECONSTRUCTED)));
  return GJXL_OK;
}

}  // namespace gjxl_amd

extern "C" {

GJXLResult gjxl_amd_aq_session_create(GJXLAmdDevice* device, uint32_t width,
                                      uint32_t height,
                                      GJXLAmdAqSession** out) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  if (out == nullptr || *out != nullptr || width == 0 || height == 0 ||
      width > (1u << 16) || height > (1u << 16)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "AQ session geometry is invalid");
  }
  auto* s = new (std::nothrow) GJXLAmdAqSession;
  if (s == nullptr) return Fail(GJXL_ERROR_OUT_OF_MEMORY, "AQ session allocation failed");
  s->device = device;
  SessionGeometry(s, width, height, s->sizes);
  size_t* z = s->sizes;
  for (int i = 0; i < GJXL_AMD_AQ_BUFFER_COUNT; ++i) {
    s->capacity[i] = z[i];
    const GJXLResult r = gjxl_amd::DeviceAlloc(device, z[i], &s->buffers[i]);
    if (r != GJXL_OK) {
      gjxl_amd_aq_session_destroy(s);
      return r;
    }
  }
  // the coefficient stores hold zeros in unused group tails (cleared on the
  // device stream, which every later stage of the session is ordered after)
  const GJXLResult r = ClearStores(s);
  if (r != GJXL_OK) {
    gjxl_amd_aq_session_destroy(s);
    return r;
  }
  *out = s;
  return GJXL_OK;
}

void gjxl_amd_aq_session_destroy(GJXLAmdAqSession* s) {
  if (s == nullptr) return;
  if (s->device != nullptr) {
    (void)hipSetDevice(s->device->ordinal);
    (void)hipStreamSynchronize(s->device->stream);
  }
  for (void*& b : s->buffers) {
    if (s->device != nullptr) gjxl_amd::DeviceFree(s->device, b);
  }
  delete s;
}

GJXLResult gjxl_amd_aq_prepare(GJXLAmdAqSession* s, const float* opsin,
                               const float* preprocessed, const uint8_t* cells,
                               const float* initial_quant_field,
                               float butteraugli_target) {
  GJXL_AMD_TRY(ValidateSession(s));
  if (opsin == nullptr || preprocessed == nullptr || cells == nullptr ||
      initial_quant_field == nullptr || !std::isfinite(butteraugli_target) ||
      butteraugli_target <= 0.0f) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "AQ preparation inputs are invalid");
  }
  GJXLAmdDevice* d = s->device;
  hipStream_t st = d->stream;
  s->prepared = false;
  s->opsin = opsin;
  s->pre = preprocessed;
  s->cells = cells;
  s->target = butteraugli_target;
  using namespace gjxl_amd;
  // the invalid-distance flag of the updates that follow this preparation
  hipLaunchKernelGGL(ZeroWordKernel, dim3(1), dim3(64), 0, st,
                     Buf<uint32_t>(s, GJXL_AMD_AQ_STATS) + kErrorWord);
  GJXL_AMD_TRY(CheckLaunch("ZeroWordKernel"));
  GJXL_AMD_TRY(LaunchMatrixStats(st, opsin, s->width, s->height, s->pw, s->ph,
                                 Buf<uint32_t>(s, GJXL_AMD_AQ_STATS)));
  const size_t blocks = static_cast<size_t>(s->layout.bw) * s->layout.bh;
  GJXL_AMD_TRY(LaunchAnchorLayout(st, cells, s->pw, s->ph,
                                  Buf<uint32_t>(s, GJXL_AMD_AQ_OFFSETS),
                                  Buf<uint32_t>(s, GJXL_AMD_AQ_ANCHOR_LIST),
                                  Buf<uint32_t>(s, GJXL_AMD_AQ_ANCHOR_COUNT),
                                  Buf<uint32_t>(s, GJXL_AMD_AQ_STATS) + 6,
                                  Buf<uint32_t>(s, GJXL_AMD_AQ_STRATEGY_LISTS), blocks,
                                  Buf<uint32_t>(s, GJXL_AMD_AQ_STATS) + 8));
  GJXL_AMD_TRY(LaunchAdjustField(st, cells, s->pw, s->ph, butteraugli_target,
                                 initial_quant_field,
                                 Buf<float>(s, GJXL_AMD_AQ_ADJUSTED_FIELD)));
  const size_t n = static_cast<size_t>(s->layout.bw) * s->layout.bh;
  uint32_t* stats = Buf<uint32_t>(s, GJXL_AMD_AQ_STATS);
  GJXL_AMD_TRY(LaunchRange(st, Buf<float>(s, GJXL_AMD_AQ_ADJUSTED_FIELD), n, stats + 4));
  uint32_t host[16];
  GJXL_AMD_TRY(Check(hipMemcpyAsync(host, stats, sizeof(host), hipMemcpyDeviceToHost, st),
                     "hipMemcpyAsync(stats)"));
  GJXL_AMD_TRY(Check(hipStreamSynchronize(st), "hipStreamSynchronize"));
  if (host[6] != 0) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                "Strategy cells do not form a valid grid of searched transforms");
  }
  for (int i = 0; i < kStrategyStages; ++i) s->strategy_counts[i] = host[8 + i];
  GJXL_AMD_TRY(LaunchAnchorDct(d, preprocessed, s->pw, s->ph,
                               Buf<uint32_t>(s, GJXL_AMD_AQ_OFFSETS), Lists(s),
                               Buf<float>(s, GJXL_AMD_AQ_COEFFICIENTS)));
  float fstats[3];
  for (int i = 0; i < 3; ++i) std::memcpy(&fstats[i], &host[i], 4);
  SelectMatrixScales(fstats, butteraugli_target, &s->x_scale, &s->b_scale);
  // src/codec/codestream.cpp:92-96
  s->x_mult = static_cast<float>(std::pow(1.25, static_cast<double>(s->x_scale - 2)));
  s->b_mult = static_cast<float>(std::pow(1.25, static_cast<double>(s->b_scale - 2)));
  // PrepareButteraugliPolicyFromRange (adaptive_quantization.cpp:1197-1242)
  float mn, mx;
  std::memcpy(&mn, &host[4], 4);
  std::memcpy(&mx, &host[5], 4);
  const float ratio = mx / mn;
  const float max_dev = std::sqrt(250.0f / ratio);
  const float asym = max_dev < 2.0f ? max_dev : 2.0f;
  s->lower = mn / (asym * max_dev);
  s->upper = mx * (max_dev / asym);
  // ComputeInitialQuantDc (adaptive_quantization.cpp:556-580)
  const float nonlinear = 0.3f * std::pow(butteraugli_target / 0.3f, 0.83f);
  const float tdc = std::max(0.5f * butteraugli_target,
                             std::min(butteraugli_target, nonlinear));
  s->quant_dc = std::min(1.095924047623553f / tdc, 50.0f);
  if (!(s->upper / s->lower < 253.0f)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Initial quant field cannot form libjxl AQ bounds");
  }
  s->prepared = true;
  return GJXL_OK;
}

GJXLResult gjxl_amd_aq_policy(GJXLAmdAqSession* s, float* lower, float* upper,
                              float* quant_dc, int32_t* x_scale,
                              int32_t* b_scale) {
  GJXL_AMD_TRY(ValidateSession(s));
  if (!s->prepared) return Fail(GJXL_ERROR_INVALID_ARGUMENT, "AQ session is not prepared");
  if (lower) *lower = s->lower;
  if (upper) *upper = s->upper;
  if (quant_dc) *quant_dc = s->quant_dc;
  if (x_scale) *x_scale = s->x_scale;
  if (b_scale) *b_scale = s->b_scale;
  return GJXL_OK;
}

GJXLResult gjxl_amd_aq_split_forward_dct(GJXLAmdAqSession* s) {
  GJXL_AMD_TRY(ValidateSession(s));
  if (!s->prepared || s->pre == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "split forward DCT needs a prepared session");
  }
  return gjxl_amd::LaunchSplitAnchorDct(s->device, s->pre, s->pw, s->ph,
                                        Buf<uint32_t>(s, GJXL_AMD_AQ_OFFSETS), Lists(s),
                                        s->layout.store, Buf<float>(s, GJXL_AMD_AQ_COEFFICIENTS));
}

GJXLResult gjxl_amd_aq_evaluate(GJXLAmdAqSession* s, const float* quant_field) {
  GJXL_AMD_TRY(ValidateSession(s));
  if (!s->prepared || quant_field == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "AQ evaluation needs a prepared session and a field");
  }
  using namespace gjxl_amd;
  GJXLAmdDevice* d = s->device;
  hipStream_t st = d->stream;
  auto* q = Buf<QuantizerState>(s, GJXL_AMD_AQ_QUANTIZER);
  GJXL_AMD_TRY(AqQuantize(s, quant_field));
  GJXL_AMD_TRY(LaunchReconstruct(
    d, Buf<int32_t>(s, GJXL_AMD_AQ_QUANTIZED), s->pw, s->ph, s->cells,
    Buf<uint32_t>(s, GJXL_AMD_AQ_OFFSETS), Lists(s), Buf<int32_t>(s, GJXL_AMD_AQ_RAW_QUANT),
    Buf<int8_t>(s, GJXL_AMD_AQ_Y_TO_X), Buf<int8_t>(s, GJXL_AMD_AQ_Y_TO_B), q, s->x_mult,
    s->b_mult, Buf<float>(s, GJXL_AMD_AQ_DC_RECONSTRUCTED),
    Buf<float>(s, GJXL_AMD_AQ_DC_SMOOTHED), Buf<float>(s, GJXL_AMD_AQ_INVERSE_SIGMA),
    Buf<float>(s, GJXL_AMD_AQ_RECONSTRUCTED)));
  return LaunchLoopFiltersAndColor(
    st, Buf<float>(s, GJXL_AMD_AQ_RECONSTRUCTED), s->pw, s->ph, s->width, s->height,
    Buf<float>(s, GJXL_AMD_AQ_INVERSE_SIGMA), Buf<float>(s, GJXL_AMD_AQ_FILTERED),
    Buf<float>(s, GJXL_AMD_AQ_LINEAR), 255.0f);
}

GJXLResult gjxl_amd_aq_update(GJXLAmdAqSession* s, float* quant_field,
                              const float* block_distance, int32_t iteration,
                              float butteraugli_target) {
  GJXL_AMD_TRY(ValidateSession(s));
  if (!s->prepared || quant_field == nullptr || block_distance == nullptr ||
      iteration < 0 || !std::isfinite(butteraugli_target) || butteraugli_target <= 0.0f) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "AQ update arguments are invalid");
  }
  const size_t n = static_cast<size_t>(s->layout.bw) * s->layout.bh;
  hipLaunchKernelGGL(AqUpdateKernel, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256),
                     0, s->device->stream, quant_field,
                     Buf<float>(s, GJXL_AMD_AQ_ADJUSTED_FIELD), block_distance, n,
                     butteraugli_target, static_cast<int>(iteration), s->lower, s->upper,
                     Buf<gjxl_amd::QuantizerState>(s, GJXL_AMD_AQ_QUANTIZER),
                     Buf<uint32_t>(s, GJXL_AMD_AQ_STATS) + kErrorWord);
  return gjxl_amd::CheckLaunch("AqUpdateKernel");
}

GJXLResult gjxl_amd_aq_run(GJXLAmdAqSession* s, GJXLAmdButteraugli* b,
                           int32_t iterations, float butteraugli_target,
                           double* scores) {
  GJXL_AMD_TRY(ValidateSession(s));
  if (b == nullptr || !b->prepared || b->device != s->device ||
      b->width != static_cast<int>(s->width) || b->height != static_cast<int>(s->height) ||
      !s->prepared || iterations < 0 || iterations > kMaxIterations) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                "AQ run needs a prepared session and a prepared Butteraugli of its extent");
  }
  hipStream_t st = s->device->stream;
  const size_t n = static_cast<size_t>(s->layout.bw) * s->layout.bh;
  float* field = Buf<float>(s, GJXL_AMD_AQ_FIELD);
  float* distance = Buf<float>(s, GJXL_AMD_AQ_BLOCK_DISTANCE);
  uint32_t* score_bits = Buf<uint32_t>(s, GJXL_AMD_AQ_SCORES);
  uint32_t* error = Buf<uint32_t>(s, GJXL_AMD_AQ_STATS) + kErrorWord;
  hipLaunchKernelGGL(ZeroWordKernel, dim3(1), dim3(64), 0, st, error);
  hipLaunchKernelGGL(CopyFloatsKernel, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256),
                     0, st, Buf<float>(s, GJXL_AMD_AQ_ADJUSTED_FIELD), n, field);
  GJXL_AMD_TRY(gjxl_amd::CheckLaunch("CopyFloatsKernel"));
  for (int it = 0; it <= iterations; ++it) {
    GJXL_AMD_TRY(gjxl_amd_aq_evaluate(s, field));
    GJXL_AMD_TRY(gjxl_amd::ButteraugliCompare(b, Buf<float>(s, GJXL_AMD_AQ_LINEAR), nullptr));
    hipLaunchKernelGGL(CopyWordKernel, dim3(1), dim3(64), 0, st, b->score_bits,
                       score_bits + it);
    GJXL_AMD_TRY(gjxl_amd::LaunchReduceDistanceMap(st, b->map, s->width, s->height, s->cells,
                                                   s->layout.bw, s->layout.bh, distance));
    if (it == iterations) break;
    GJXL_AMD_TRY(gjxl_amd_aq_update(s, field, distance, it, butteraugli_target));
  }
  if (scores == nullptr) return GJXL_OK;
  uint32_t host[kMaxIterations + 2] = {};
  GJXL_AMD_TRY(Check(hipMemcpyAsync(host, score_bits, 4 * (iterations + 1),
                                    hipMemcpyDeviceToHost, st), "hipMemcpyAsync(scores)"));
  uint32_t err = 0;
  GJXL_AMD_TRY(Check(hipMemcpyAsync(&err, error, 4, hipMemcpyDeviceToHost, st),
                     "hipMemcpyAsync(error)"));
  GJXL_AMD_TRY(Check(hipStreamSynchronize(st), "hipStreamSynchronize"));
  if (err != 0) {
    return Fail(GJXL_ERROR_INTERNAL, "Adaptive quantization produced an invalid block distance");
  }
  for (int it = 0; it <= iterations; ++it) {
    float v;
    std::memcpy(&v, &host[it], 4);
    scores[it] = v;
  }
  return GJXL_OK;
}

GJXLResult gjxl_amd_aq_check(GJXLAmdAqSession* s) {
  GJXL_AMD_TRY(ValidateSession(s));
  if (!s->prepared) return Fail(GJXL_ERROR_INVALID_ARGUMENT, "AQ session is not prepared");
  hipStream_t st = s->device->stream;
  uint32_t err = 0;
  GJXL_AMD_TRY(Check(hipMemcpyAsync(&err, Buf<uint32_t>(s, GJXL_AMD_AQ_STATS) + kErrorWord, 4,
                                    hipMemcpyDeviceToHost, st),
                     "hipMemcpyAsync(error)"));
  GJXL_AMD_TRY(Check(hipStreamSynchronize(st), "hipStreamSynchronize"));
  if (err != 0) {
    return Fail(GJXL_ERROR_INTERNAL, "Adaptive quantization produced an invalid block distance");
  }
  return GJXL_OK;
}

GJXLResult gjxl_amd_aq_buffer(GJXLAmdAqSession* s, int32_t id, void** pointer,
                              size_t* bytes) {
  GJXL_AMD_TRY(ValidateSession(s));
  if (id < 0 || id >= GJXL_AMD_AQ_BUFFER_COUNT || pointer == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "AQ buffer id is invalid");
  }
  *pointer = s->buffers[id];
  if (bytes) *bytes = s->sizes[id];
  return GJXL_OK;
}

}  // extern "C"
