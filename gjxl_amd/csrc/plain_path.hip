// The plain path of one image extent as one C++ host sequence on the device
// handle's stream: upload of the packed sRGB8 pixels, XYB, Gaborish inverse,
// initial quant field, initial CfL, candidate costs, strategy search, AQ
// preparation, iterations + 1 Butteraugli-driven evaluations with field
// updates, AC tokenization of the last evaluation and the read-back of its
// token stream.  At effort 7 the coefficient orders and the block-context map
// of the tokens are chosen on the host from the final frame's statistics, as
// the reference's serializer chooses them (src/codestream/vardct/
// encoder.cpp:1985-2010).
//
// The order is RunPreparedQuantizationPipelineWithProviders
// (src/codec/quantization_pipeline.cpp:428-586) with every provider on the
// GPU, followed by the AcTokenizationProvider hand-off
// (src/codestream/ac_tokenization_provider_internal.h:18-47), as PrepareWorkflow
// (src/codestream/workflow.cpp:636-768) runs it for gjxl_encode
// (src/c_api/gjxl.cpp:470-637) at distance d, effort 7.  The Butteraugli
// reference side runs on a second stream from the linear image on, beside
// the stages up to the AQ loop.  The host waits three times:
// once for the quantization-matrix statistics that choose the matrix scales
// (src/codestream/workflow.cpp:1318-1484), once for the frame statistics and
// once for the token count.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <new>

#include "aq_common.h"
#include "butteraugli.h"
#include "common.h"
#include "proposal.h"
#include "tokenize.h"

namespace {

constexpr int kMaxMarks = 48;
constexpr int kMaxEvaluations = 8;

// The path's own device buffers
enum PathBuffer {
  kSrgb, kLinear, kOpsin, kPreprocessed, kQuantField, kStrategyMask, kPixelMask,
  kIqfScratch, kYToX, kYToB, kCosts, kCells, kScoreWords, kFeatures, kAllow, kPathBuffers
};

}  // namespace

struct GJXLAmdPlainPath {
  GJXLAmdDevice* device = nullptr;
  uint32_t width = 0, height = 0, pw = 0, ph = 0, bw = 0, bh = 0, groups = 0;
  GJXLAmdPlainPathOptions options{};
  int32_t iterations = 2;
  // device storage: own[] as allocated (capacity[] bytes) and typed views
  void* own[kPathBuffers] = {};
  size_t capacity[kPathBuffers] = {};
  uint8_t* srgb = nullptr;
  float* linear = nullptr;
  float* opsin = nullptr;
  float* preprocessed = nullptr;
  float* quant_field = nullptr;
  float* strategy_mask = nullptr;
  float* pixel_mask = nullptr;
  float* iqf_scratch = nullptr;
  int8_t* y_to_x = nullptr;
  int8_t* y_to_b = nullptr;
  float* costs = nullptr;
  uint8_t* cells = nullptr;
  uint32_t* score_words = nullptr;
  // the learned run's block features and allowed stages
  uint16_t* features = nullptr;
  uint8_t* allow = nullptr;
  GJXLAmdProposal* proposal = nullptr;
  GJXLAmdProposalOptions proposal_options{};
  int32_t evaluations_run = 0;
  float proposal_norm = 0.0f;
  const float* proposal_head = nullptr;  // the learned run's head, for the back half
  bool front_done = false;
  GJXLAmdAqSession* session = nullptr;
  GJXLAmdButteraugli* butteraugli = nullptr;
  GJXLAmdTokenizer* tokenizer = nullptr;
  // a second handle (its own stream) for the Butteraugli reference side,
  // which needs only the linear image: it overlaps the stages before the
  // AQ loop; linear_ready and reference_ready order the two streams
  GJXLAmdDevice* side = nullptr;
  hipEvent_t linear_ready = nullptr;
  hipEvent_t reference_ready = nullptr;
  // pinned host words: token status (4), AQ error (1), scores (kMaxEvaluations)
  uint32_t* host_words = nullptr;
  // pinned host copy of the frame statistics; the entropy inputs chosen
  // from them (effort 7) for the last run
  uint32_t* host_statistics = nullptr;
  GJXLAmdFrameStatistics* statistics = nullptr;
  GJXLAmdEntropyInputs* entropy_inputs = nullptr;
  // stage marks of the last run
  hipEvent_t marks[kMaxMarks] = {};
  int32_t mark_stage[kMaxMarks] = {};
  int mark_count = 0;
  uint64_t token_count = 0;
  float scores[kMaxEvaluations] = {};
  bool ran = false;
};

using gjxl_amd::Check;
using gjxl_amd::Fail;

namespace {

GJXLResult Validate(const GJXLAmdPlainPath* p) {
  if (p == nullptr) return Fail(GJXL_ERROR_INVALID_ARGUMENT, "plain path is null");
  return gjxl_amd::ValidateDevice(p->device);
}

GJXLResult Mark(GJXLAmdPlainPath* p, int32_t stage) {
  if (p->mark_count >= kMaxMarks) return GJXL_OK;
  p->mark_stage[p->mark_count] = stage;
  return Check(hipEventRecord(p->marks[p->mark_count++], p->device->stream), "hipEventRecord");
}

// The extent fields of a path and the byte size of each of its own buffers
// at that extent (every size grows with the width and with the height).
void PathGeometry(GJXLAmdPlainPath* p, uint32_t width, uint32_t height, size_t* z) {
  p->width = width;
  p->height = height;
  p->pw = gjxl_amd::PadTo8(width);
  p->ph = gjxl_amd::PadTo8(height);
  p->bw = p->pw / 8;
  p->bh = p->ph / 8;
  p->groups = gjxl_amd::CeilDiv(p->bw, 32) * gjxl_amd::CeilDiv(p->bh, 32);
  const size_t pixels = static_cast<size_t>(width) * height;
  const size_t padded = static_cast<size_t>(p->pw) * p->ph;
  const size_t blocks = static_cast<size_t>(p->bw) * p->bh;
  const size_t tiles = static_cast<size_t>(gjxl_amd::CeilDiv(p->bw, 8)) *
                       gjxl_amd::CeilDiv(p->bh, 8);
  z[kSrgb] = 4 * pixels;  // room for RGBA rows without padding
  z[kLinear] = 4 * 3 * pixels;
  z[kOpsin] = 4 * 3 * padded;
  z[kPreprocessed] = 4 * 3 * padded;
  z[kQuantField] = 4 * blocks;
  z[kStrategyMask] = 4 * blocks;
  z[kPixelMask] = 4 * padded;
  z[kIqfScratch] = gjxl_amd_initial_quant_field_scratch_bytes(p->pw, p->ph);
  z[kYToX] = tiles;
  z[kYToB] = tiles;
  z[kCosts] = 4 * 7 * blocks;
  z[kCells] = blocks;
  // scores, the judged norm, then (8-byte aligned) the norm's partial sums
  z[kScoreWords] = 4 * (kMaxEvaluations + 2) + 8 * gjxl_amd::kProposalNormPartials;
  z[kFeatures] = 2 * GJXL_AMD_PROPOSAL_FEATURES * blocks;
  z[kAllow] = blocks;
}

void PointViews(GJXLAmdPlainPath* p) {
  p->srgb = static_cast<uint8_t*>(p->own[kSrgb]);
  p->linear = static_cast<float*>(p->own[kLinear]);
  p->opsin = static_cast<float*>(p->own[kOpsin]);
  p->preprocessed = static_cast<float*>(p->own[kPreprocessed]);
  p->quant_field = static_cast<float*>(p->own[kQuantField]);
  p->strategy_mask = static_cast<float*>(p->own[kStrategyMask]);
  p->pixel_mask = static_cast<float*>(p->own[kPixelMask]);
  p->iqf_scratch = static_cast<float*>(p->own[kIqfScratch]);
  p->y_to_x = static_cast<int8_t*>(p->own[kYToX]);
  p->y_to_b = static_cast<int8_t*>(p->own[kYToB]);
  p->costs = static_cast<float*>(p->own[kCosts]);
  p->cells = static_cast<uint8_t*>(p->own[kCells]);
  p->score_words = static_cast<uint32_t*>(p->own[kScoreWords]);
  p->features = static_cast<uint16_t*>(p->own[kFeatures]);
  p->allow = static_cast<uint8_t*>(p->own[kAllow]);
}

template <typename T>
T* AqBuffer(GJXLAmdAqSession* s, int id) {
  void* p = nullptr;
  size_t bytes = 0;
  (void)gjxl_amd_aq_buffer(s, id, &p, &bytes);
  return static_cast<T*>(p);
}

template <typename T>
T* TokenBuffer(GJXLAmdTokenizer* t, int id) {
  void* p = nullptr;
  size_t bytes = 0;
  (void)gjxl_amd_tokenizer_buffer(t, id, &p, &bytes);
  return static_cast<T*>(p);
}

__global__ void CopyScoreKernel(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst) {
  if (threadIdx.x == 0) *dst = *src;
}

GJXLResult ValidateOptions(const GJXLAmdPlainPathOptions* o) {
  if (o == nullptr || o->struct_size < sizeof(GJXLAmdPlainPathOptions)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "plain path options are missing or too small");
  }
  if (!std::isfinite(o->distance) || o->distance <= 0.0f || o->distance > 25.0f) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "distance must be in (0, 25]");
  }
  if (o->effort != 7) {
    return Fail(GJXL_ERROR_UNSUPPORTED,
                "the device plain path implements effort 7 (two AQ updates)");
  }
  if (!std::isfinite(o->intensity_target) || o->intensity_target <= 0.0f) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "intensity target must be positive");
  }
  return GJXL_OK;
}

// The run after the upload (device-resident packed pixels in p->srgb or the
// caller's device buffer).
// The front half: XYB, the Butteraugli reference side (second stream),
// Gaborish inverse, initial quant field, the learned proposal if attached,
// initial CfL; enqueued, no host wait.  What the candidate costs read is
// ready on the device stream at its end.
GJXLResult RunFront(GJXLAmdPlainPath* p, const uint8_t* srgb, size_t row_stride,
                    uint32_t bpp) {
  GJXLAmdDevice* d = p->device;
  const float t = p->options.distance;
  GJXL_AMD_TRY(gjxl_amd_srgb8_to_linear_and_opsin(d, srgb, row_stride, bpp, p->width,
                                                  p->height, p->options.intensity_target,
                                                  p->linear, p->opsin));
  GJXL_AMD_TRY(Mark(p, GJXL_AMD_STAGE_SRGB_TO_LINEAR_AND_OPSIN));
  // the Butteraugli reference side on the side stream, from the linear image
  // (not needed by an unjudged learned run of one pass, which compares
  // nothing)
  const GJXLAmdProposalOptions& po = p->proposal_options;
  const bool compares =
    p->proposal == nullptr || po.final_compare != 0 || po.evaluations > 1;
  if (compares) {
    GJXL_AMD_TRY(Check(hipEventRecord(p->linear_ready, d->stream), "hipEventRecord"));
    GJXL_AMD_TRY(Check(hipStreamWaitEvent(p->side->stream, p->linear_ready, 0),
                       "hipStreamWaitEvent"));
    p->butteraugli->device = p->side;
    const GJXLResult prepared = gjxl_amd_butteraugli_prepare(p->butteraugli, p->linear);
    p->butteraugli->device = d;
    GJXL_AMD_TRY(prepared);
    GJXL_AMD_TRY(Check(hipEventRecord(p->reference_ready, p->side->stream), "hipEventRecord"));
  }
  const float ones[3] = {1.0f, 1.0f, 1.0f};
  GJXL_AMD_TRY(gjxl_amd_gaborish_inverse(d, p->opsin, p->pw, p->ph, ones, p->preprocessed));
  GJXL_AMD_TRY(Mark(p, GJXL_AMD_STAGE_GABORISH_INVERSE));
  GJXL_AMD_TRY(gjxl_amd_initial_quant_field(d, p->opsin, p->pw, p->ph, t, 1.0f, p->quant_field,
                                            p->strategy_mask, p->pixel_mask, p->iqf_scratch));
  GJXL_AMD_TRY(Mark(p, GJXL_AMD_STAGE_INITIAL_QUANT_FIELD));
  const float* head = nullptr;
  if (p->proposal != nullptr) {
    // the learned form: features, the network's forward and the allowed
    // stages of every block, on this path's streams whatever handle the
    // network was created on; with no Butteraugli reference side to run,
    // on the side stream beside the CfL
    GJXLAmdDevice* pd = compares ? d : p->side;
    if (!compares) {
      GJXL_AMD_TRY(Check(hipEventRecord(p->linear_ready, d->stream), "hipEventRecord"));
      GJXL_AMD_TRY(Check(hipStreamWaitEvent(pd->stream, p->linear_ready, 0),
                         "hipStreamWaitEvent"));
    }
    GJXL_AMD_TRY(gjxl_amd::LaunchProposalFeatures(pd, p->preprocessed, p->pw, p->ph,
                                                  p->quant_field, p->features));
    GJXL_AMD_TRY(gjxl_amd::ProposalForward(p->proposal, pd->stream, p->features, 1, p->bh,
                                           p->bw, &head));
    GJXL_AMD_TRY(gjxl_amd::LaunchProposalDecide(pd, head, p->bw, p->bh, po.threshold,
                                                static_cast<int>(po.top_k), po.p_min, p->allow));
    if (!compares) {
      GJXL_AMD_TRY(Check(hipEventRecord(p->reference_ready, pd->stream), "hipEventRecord"));
    }
    GJXL_AMD_TRY(Mark(p, GJXL_AMD_STAGE_PROPOSAL));
  }
  GJXL_AMD_TRY(gjxl_amd_initial_cfl(d, p->preprocessed, p->pw, p->ph, p->y_to_x, p->y_to_b));
  GJXL_AMD_TRY(Mark(p, GJXL_AMD_STAGE_INITIAL_CFL));
  if (!compares) {
    // the decisions, which the candidate costs read
    GJXL_AMD_TRY(Check(hipStreamWaitEvent(d->stream, p->reference_ready, 0),
                       "hipStreamWaitEvent"));
  }
  // what is left of the reference side once the device stream gets here:
  // the side stream shares the GPU with the latency-bound stages before the
  // candidate costs, which the costs themselves would not leave room for
  if (compares) {
    GJXL_AMD_TRY(Check(hipStreamWaitEvent(d->stream, p->reference_ready, 0),
                       "hipStreamWaitEvent"));
    GJXL_AMD_TRY(Mark(p, GJXL_AMD_STAGE_BUTTERAUGLI_PREPARE));
  }
  p->proposal_head = head;
  return GJXL_OK;
}

// The back half: the candidate costs (unless a caller has written them
// into p->costs on this stream, costs_ready), the strategy search, AQ
// preparation, the Butteraugli loop, the entropy inputs and the tokens.
GJXLResult RunBack(GJXLAmdPlainPath* p, bool costs_ready) {
  GJXLAmdDevice* d = p->device;
  const float t = p->options.distance;
  const GJXLAmdProposalOptions& po = p->proposal_options;
  const float* head = p->proposal_head;
  if (!costs_ready) {
    GJXL_AMD_TRY(gjxl_amd_ac_strategy_costs_pruned(
      d, p->preprocessed, p->pw, p->ph, p->quant_field, p->pixel_mask, p->y_to_x, p->y_to_b, t,
      p->proposal != nullptr ? p->allow : nullptr, p->costs));
  }
  GJXL_AMD_TRY(Mark(p, GJXL_AMD_STAGE_AC_STRATEGY_COSTS));
  GJXL_AMD_TRY(gjxl_amd_ac_strategy_search(d, p->costs, p->pw, p->ph, t, p->cells));
  GJXL_AMD_TRY(Mark(p, GJXL_AMD_STAGE_AC_STRATEGY_SEARCH));
  GJXL_AMD_TRY(gjxl_amd_aq_prepare(p->session, p->opsin, p->preprocessed, p->cells,
                                   p->quant_field, t));
  GJXL_AMD_TRY(Mark(p, GJXL_AMD_STAGE_AQ_PREPARE));
  // RunAdaptiveQuantizationPolicyImpl (src/codec/adaptive_quantization.cpp:
  // 1600-1738): evaluate, compare, reduce, update
  hipStream_t st = d->stream;
  float* field = AqBuffer<float>(p->session, GJXL_AMD_AQ_FIELD);
  float* distance = AqBuffer<float>(p->session, GJXL_AMD_AQ_BLOCK_DISTANCE);
  const float* decoded = AqBuffer<float>(p->session, GJXL_AMD_AQ_LINEAR);
  const size_t n = static_cast<size_t>(p->bw) * p->bh;
  const float* adjusted = AqBuffer<float>(p->session, GJXL_AMD_AQ_ADJUSTED_FIELD);
  // updates planned before the first evaluation: the plain path's, or the
  // learned run's (whose judge may add more, up to the plain count)
  int planned = p->iterations;
  if (head != nullptr) {
    float lower = 0.0f, upper = 0.0f;
    GJXL_AMD_TRY(gjxl_amd_aq_policy(p->session, &lower, &upper, nullptr, nullptr, nullptr));
    GJXL_AMD_TRY(gjxl_amd::LaunchProposalWarmStart(d, head, adjusted, n, po.field_limit,
                                                   po.field_offset, lower, upper, field));
    planned = po.evaluations - 1;
  } else {
    GJXL_AMD_TRY(gjxl_amd_copy(d, field, adjusted, 4 * n));
  }
  // the unjudged learned run: its planned updates, then a last pass that
  // only quantizes (what the tokens read), without a comparison
  const bool unjudged = head != nullptr && po.final_compare == 0;
  if (unjudged) p->proposal_norm = std::nanf("");
  int it = 0;
  for (;; ++it) {
    if (unjudged && it == planned) {
      GJXL_AMD_TRY(gjxl_amd::AqQuantize(p->session, field));
      GJXL_AMD_TRY(Mark(p, GJXL_AMD_STAGE_AQ_EVALUATE));
      break;
    }
    GJXL_AMD_TRY(gjxl_amd_aq_evaluate(p->session, field));
    GJXL_AMD_TRY(Mark(p, GJXL_AMD_STAGE_AQ_EVALUATE));
    GJXL_AMD_TRY(gjxl_amd::ButteraugliCompare(p->butteraugli, decoded, nullptr));
    GJXL_AMD_TRY(Mark(p, GJXL_AMD_STAGE_BUTTERAUGLI_COMPARE));
    hipLaunchKernelGGL(CopyScoreKernel, dim3(1), dim3(64), 0, st, p->butteraugli->score_bits,
                       p->score_words + it);
    GJXL_AMD_TRY(gjxl_amd::CheckLaunch("CopyScoreKernel"));
    GJXL_AMD_TRY(gjxl_amd::LaunchReduceDistanceMap(st, p->butteraugli->map, p->width, p->height,
                                                   p->cells, p->bw, p->bh, distance));
    bool more = it < planned;
    if (!more && head != nullptr) {
      // the judge: the 3-norm of the block distances; the learned run goes
      // on with the plain updates while it is above judge_ratio * distance
      float* norm = reinterpret_cast<float*>(p->score_words + kMaxEvaluations);
      double* partial = reinterpret_cast<double*>(p->score_words + kMaxEvaluations + 2);
      GJXL_AMD_TRY(gjxl_amd::LaunchProposalNorm3(d, distance, n, partial, norm));
      GJXL_AMD_TRY(Check(hipMemcpyAsync(p->host_words + 24, norm, 4, hipMemcpyDeviceToHost, st),
                         "hipMemcpyAsync(norm)"));
      GJXL_AMD_TRY(Check(hipStreamSynchronize(st), "hipStreamSynchronize"));
      std::memcpy(&p->proposal_norm, p->host_words + 24, 4);
      more = it < p->iterations && !(p->proposal_norm <= po.judge_ratio * t);
    }
    if (more) GJXL_AMD_TRY(gjxl_amd_aq_update(p->session, field, distance, it, t));
    GJXL_AMD_TRY(Mark(p, GJXL_AMD_STAGE_AQ_REDUCE_UPDATE));
    if (!more) break;
  }
  // evaluations with a comparison (the unjudged last pass has none)
  p->evaluations_run = unjudged ? it : it + 1;
  // The CPU-chosen tokenization inputs of effort 7 (coefficient orders and
  // the adaptive block-context map) from the final frame's statistics, with
  // the AQ error flag and the scores in the same wait
  const int32_t* quantized = AqBuffer<int32_t>(p->session, GJXL_AMD_AQ_QUANTIZED);
  const int32_t* raw_quant = AqBuffer<int32_t>(p->session, GJXL_AMD_AQ_RAW_QUANT);
  GJXL_AMD_TRY(gjxl_amd::FrameStatisticsEnqueue(p->tokenizer, quantized, p->cells, raw_quant));
  uint32_t* host = p->host_words;
  uint32_t* stats = AqBuffer<uint32_t>(p->session, GJXL_AMD_AQ_STATS);
  GJXL_AMD_TRY(Check(hipMemcpyAsync(p->host_statistics,
                                    TokenBuffer<uint32_t>(p->tokenizer,
                                                          GJXL_AMD_TOKENS_STATISTICS),
                                    4 * gjxl_amd::FrameStatisticsWords(),
                                    hipMemcpyDeviceToHost, st),
                     "hipMemcpyAsync(frame statistics)"));
  GJXL_AMD_TRY(Check(hipMemcpyAsync(host + 4, stats + 7, 4, hipMemcpyDeviceToHost, st),
                     "hipMemcpyAsync(AQ error)"));
  if (p->evaluations_run > 0) {
    GJXL_AMD_TRY(Check(hipMemcpyAsync(host + 8, p->score_words, 4 * p->evaluations_run,
                                      hipMemcpyDeviceToHost, st), "hipMemcpyAsync(scores)"));
  }
  GJXL_AMD_TRY(Check(hipStreamSynchronize(st), "hipStreamSynchronize"));
  if (host[4] != 0) {
    return Fail(GJXL_ERROR_INTERNAL, "Adaptive quantization produced an invalid block distance");
  }
  GJXL_AMD_TRY(gjxl_amd::FrameStatisticsResult(p->tokenizer, p->host_statistics,
                                               p->statistics));
  GJXL_AMD_TRY(gjxl_amd_entropy_inputs_from_statistics(p->statistics, p->options.effort,
                                                       p->entropy_inputs));
  GJXL_AMD_TRY(gjxl_amd_tokenizer_set_entropy_inputs(p->tokenizer, p->entropy_inputs));
  GJXL_AMD_TRY(Mark(p, GJXL_AMD_STAGE_ENTROPY_INPUTS));
  // AC tokens of the last evaluation, then their status words
  GJXL_AMD_TRY(gjxl_amd::TokenizeEnqueue(p->tokenizer, quantized, p->cells, raw_quant, false));
  GJXL_AMD_TRY(Mark(p, GJXL_AMD_STAGE_AC_TOKENIZE));
  GJXL_AMD_TRY(Check(hipMemcpyAsync(host, TokenBuffer<uint32_t>(p->tokenizer,
                                                               GJXL_AMD_TOKENS_STATUS),
                                    16, hipMemcpyDeviceToHost, st), "hipMemcpyAsync(status)"));
  GJXL_AMD_TRY(Check(hipStreamSynchronize(st), "hipStreamSynchronize"));
  GJXL_AMD_TRY(gjxl_amd::TokenizeResult(host, &p->token_count));
  for (int e = 0; e < p->evaluations_run; ++e) std::memcpy(&p->scores[e], &host[8 + e], 4);
  return GJXL_OK;
}

GJXLResult RunResident(GJXLAmdPlainPath* p, const uint8_t* srgb, size_t row_stride,
                       uint32_t bpp) {
  GJXL_AMD_TRY(RunFront(p, srgb, row_stride, bpp));
  return RunBack(p, false);
}

GJXLResult ReadTokens(GJXLAmdPlainPath* p, const GJXLAmdTokenOutput* out) {
  if (out == nullptr) return GJXL_OK;
  if (out->values == nullptr && out->contexts == nullptr && out->group_offsets == nullptr) {
    return GJXL_OK;
  }
  if ((out->values != nullptr || out->contexts != nullptr) && out->capacity < p->token_count) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "token output capacity is below the token count");
  }
  if (out->group_offsets != nullptr && out->group_capacity < p->groups + 1u) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "group offset output holds fewer than groups + 1");
  }
  hipStream_t st = p->device->stream;
  const size_t count = static_cast<size_t>(p->token_count);
  if (out->values != nullptr && count > 0) {
    GJXL_AMD_TRY(Check(hipMemcpyAsync(out->values,
                                      TokenBuffer<uint32_t>(p->tokenizer, GJXL_AMD_TOKENS_VALUES),
                                      4 * count, hipMemcpyDeviceToHost, st),
                       "hipMemcpyAsync(token values)"));
  }
  if (out->contexts != nullptr && count > 0) {
    GJXL_AMD_TRY(Check(hipMemcpyAsync(out->contexts,
                                      TokenBuffer<uint16_t>(p->tokenizer,
                                                            GJXL_AMD_TOKENS_CONTEXTS),
                                      2 * count, hipMemcpyDeviceToHost, st),
                       "hipMemcpyAsync(token contexts)"));
  }
  if (out->group_offsets != nullptr) {
    GJXL_AMD_TRY(Check(hipMemcpyAsync(out->group_offsets,
                                      TokenBuffer<uint32_t>(p->tokenizer,
                                                            GJXL_AMD_TOKENS_GROUP_OFFSETS),
                                      4 * (p->groups + 1), hipMemcpyDeviceToHost, st),
                       "hipMemcpyAsync(group offsets)"));
  }
  GJXL_AMD_TRY(Mark(p, GJXL_AMD_STAGE_TOKEN_READBACK));
  return Check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

GJXLResult FillFrame(GJXLAmdPlainPath* p, GJXLAmdPlainFrame* f) {
  if (f == nullptr) return GJXL_OK;
  GJXLAmdPlainFrame r{};
  r.block_width = p->bw;
  r.block_height = p->bh;
  r.groups = p->groups;
  r.cells = p->cells;
  r.quant_field = AqBuffer<float>(p->session, GJXL_AMD_AQ_FIELD);
  r.block_distance = AqBuffer<float>(p->session, GJXL_AMD_AQ_BLOCK_DISTANCE);
  r.raw_quant = AqBuffer<int32_t>(p->session, GJXL_AMD_AQ_RAW_QUANT);
  r.quantized = AqBuffer<int32_t>(p->session, GJXL_AMD_AQ_QUANTIZED);
  r.dc_quantized = AqBuffer<int32_t>(p->session, GJXL_AMD_AQ_DC_QUANTIZED);
  r.y_to_x = AqBuffer<int8_t>(p->session, GJXL_AMD_AQ_Y_TO_X);
  r.y_to_b = AqBuffer<int8_t>(p->session, GJXL_AMD_AQ_Y_TO_B);
  r.token_values = TokenBuffer<uint32_t>(p->tokenizer, GJXL_AMD_TOKENS_VALUES);
  r.token_contexts = TokenBuffer<uint16_t>(p->tokenizer, GJXL_AMD_TOKENS_CONTEXTS);
  r.token_group_offsets = TokenBuffer<uint32_t>(p->tokenizer, GJXL_AMD_TOKENS_GROUP_OFFSETS);
  r.token_count = p->token_count;
  float lo = 0, hi = 0, qdc = 0;
  GJXL_AMD_TRY(gjxl_amd_aq_policy(p->session, &lo, &hi, &qdc, &r.x_qm_scale, &r.b_qm_scale));
  r.quant_dc = qdc;
  r.evaluations = p->evaluations_run;
  r.proposal_allow = p->proposal != nullptr ? p->allow : nullptr;
  r.proposal_norm = p->proposal != nullptr ? p->proposal_norm : 0.0f;
  for (int i = 0; i < r.evaluations; ++i) r.scores[i] = p->scores[i];
  r.entropy_inputs = p->entropy_inputs;
  *f = r;
  return GJXL_OK;
}

}  // namespace

extern "C" {

GJXLResult gjxl_amd_plain_path_options_init(GJXLAmdPlainPathOptions* options,
                                            size_t struct_size) {
  if (options == nullptr || struct_size < sizeof(GJXLAmdPlainPathOptions)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "plain path options are missing or too small");
  }
  std::memset(options, 0, sizeof(GJXLAmdPlainPathOptions));
  options->struct_size = sizeof(GJXLAmdPlainPathOptions);
  options->distance = 1.0f;  // gjxl_encoder_options_init (include/gjxl/gjxl.h:207-211)
  options->effort = 7;
  options->intensity_target = 255.0f;
  return GJXL_OK;
}

GJXLResult gjxl_amd_plain_path_create(GJXLAmdDevice* device, uint32_t width, uint32_t height,
                                      const GJXLAmdPlainPathOptions* options,
                                      GJXLAmdPlainPath** out) {
  GJXL_AMD_TRY(gjxl_amd::ValidateDevice(device));
  if (out == nullptr || *out != nullptr || width == 0 || height == 0 ||
      width > (1u << 16) || height > (1u << 16)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "plain path geometry is invalid");
  }
  GJXLAmdPlainPathOptions defaults{};
  if (options == nullptr) {
    GJXL_AMD_TRY(gjxl_amd_plain_path_options_init(&defaults, sizeof(defaults)));
    options = &defaults;
  }
  GJXL_AMD_TRY(ValidateOptions(options));
  auto* p = new (std::nothrow) GJXLAmdPlainPath;
  if (p == nullptr) return Fail(GJXL_ERROR_OUT_OF_MEMORY, "plain path allocation failed");
  p->device = device;
  p->options = *options;
  p->iterations = 2;
  PathGeometry(p, width, height, p->capacity);
  GJXLResult r = GJXL_OK;
  for (int i = 0; r == GJXL_OK && i < kPathBuffers; ++i) {
    r = gjxl_amd::DeviceAlloc(device, p->capacity[i], &p->own[i]);
  }
  PointViews(p);
  if (r == GJXL_OK) {
    r = Check(hipHostMalloc(reinterpret_cast<void**>(&p->host_words), 4 * 32,
                            hipHostMallocDefault), "hipHostMalloc(plain path)");
  }
  if (r == GJXL_OK) {
    r = Check(hipHostMalloc(reinterpret_cast<void**>(&p->host_statistics),
                            4 * gjxl_amd::FrameStatisticsWords(), hipHostMallocDefault),
              "hipHostMalloc(frame statistics)");
  }
  if (r == GJXL_OK) {
    p->statistics = new (std::nothrow) GJXLAmdFrameStatistics;
    p->entropy_inputs = new (std::nothrow) GJXLAmdEntropyInputs;
    if (p->statistics == nullptr || p->entropy_inputs == nullptr) {
      r = Fail(GJXL_ERROR_OUT_OF_MEMORY, "plain path allocation failed");
    }
  }
  for (int i = 0; r == GJXL_OK && i < kMaxMarks; ++i) {
    r = Check(hipEventCreate(&p->marks[i]), "hipEventCreate");
  }
  if (r == GJXL_OK) r = gjxl_amd_aq_session_create(device, width, height, &p->session);
  if (r == GJXL_OK) {
    r = gjxl_amd_butteraugli_create(device, width, height, 1.0f, 1.0f, 80.0f, &p->butteraugli);
  }
  if (r == GJXL_OK) r = gjxl_amd_device_create(device->ordinal, &p->side);
  if (r == GJXL_OK) {
    r = Check(hipEventCreateWithFlags(&p->linear_ready, hipEventDisableTiming),
              "hipEventCreateWithFlags");
  }
  if (r == GJXL_OK) {
    r = Check(hipEventCreateWithFlags(&p->reference_ready, hipEventDisableTiming),
              "hipEventCreateWithFlags");
  }
  if (r == GJXL_OK) r = gjxl_amd_tokenizer_create(device, width, height, &p->tokenizer);
  if (r != GJXL_OK) {
    gjxl_amd_plain_path_destroy(p);
    return r;
  }
  *out = p;
  return GJXL_OK;
}

void gjxl_amd_plain_path_destroy(GJXLAmdPlainPath* p) {
  if (p == nullptr) return;
  if (p->device != nullptr) {
    (void)hipSetDevice(p->device->ordinal);
    (void)hipStreamSynchronize(p->device->stream);
  }
  if (p->side != nullptr) (void)hipStreamSynchronize(p->side->stream);
  gjxl_amd_tokenizer_destroy(p->tokenizer);
  gjxl_amd_butteraugli_destroy(p->butteraugli);
  gjxl_amd_aq_session_destroy(p->session);
  for (hipEvent_t e : p->marks) {
    if (e != nullptr) (void)hipEventDestroy(e);
  }
  if (p->linear_ready != nullptr) (void)hipEventDestroy(p->linear_ready);
  if (p->reference_ready != nullptr) (void)hipEventDestroy(p->reference_ready);
  gjxl_amd_device_destroy(p->side);
  if (p->host_words != nullptr) (void)hipHostFree(p->host_words);
  if (p->host_statistics != nullptr) (void)hipHostFree(p->host_statistics);
  delete p->statistics;
  delete p->entropy_inputs;
  for (void* b : p->own) {
    gjxl_amd::DeviceFree(p->device, b);
  }
  delete p;
}

GJXLResult gjxl_amd_plain_path_reconfigure(GJXLAmdPlainPath* p, uint32_t width,
                                           uint32_t height,
                                           const GJXLAmdPlainPathOptions* options) {
  GJXL_AMD_TRY(Validate(p));
  if (width == 0 || height == 0 || width > (1u << 16) || height > (1u << 16)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "plain path geometry is invalid");
  }
  if (options != nullptr) GJXL_AMD_TRY(ValidateOptions(options));
  GJXLAmdPlainPath probe;
  size_t need[kPathBuffers] = {};
  PathGeometry(&probe, width, height, need);
  for (int i = 0; i < kPathBuffers; ++i) {
    if (need[i] > p->capacity[i]) {
      return Fail(GJXL_ERROR_INVALID_ARGUMENT, "extent exceeds the plain path's capacity");
    }
  }
  GJXL_AMD_TRY(Check(hipSetDevice(p->device->ordinal), "hipSetDevice"));
  // the last run's device views go stale from here on
  p->ran = false;
  p->token_count = 0;
  // the parts check their own capacities; a refusal leaves every part at
  // the old extent
  GJXL_AMD_TRY(gjxl_amd::ResizeAqSession(p->session, width, height));
  GJXLResult r = gjxl_amd::ResizeButteraugli(p->butteraugli, width, height);
  if (r == GJXL_OK) r = gjxl_amd::ResizeTokenizer(p->tokenizer, width, height);
  if (r != GJXL_OK) {
    (void)gjxl_amd::ResizeAqSession(p->session, p->width, p->height);
    (void)gjxl_amd::ResizeButteraugli(p->butteraugli, p->width, p->height);
    (void)gjxl_amd::ResizeTokenizer(p->tokenizer, p->width, p->height);
    return r;
  }
  PathGeometry(p, width, height, need);
  if (options != nullptr) p->options = *options;
  p->front_done = false;
  p->ran = false;
  return GJXL_OK;
}

}  // extern "C"

namespace {

// Validates the image, starts the stage marks and uploads host pixels into
// the path's device copy; *srgb, *stride: the resident pixels.
GJXLResult BeginRun(GJXLAmdPlainPath* p, const GJXLAmdImage* image, const uint8_t** srgb_out,
                    size_t* stride_out) {
  GJXL_AMD_TRY(Validate(p));
  if (image == nullptr || image->pixels == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "plain path image is null");
  }
  if (image->width != p->width || image->height != p->height) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "image extent differs from the plain path's");
  }
  const uint32_t bpp = image->bytes_per_pixel;
  if ((bpp != 3 && bpp != 4) ||
      image->row_stride_bytes < static_cast<size_t>(image->width) * bpp) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "pixel layout must be packed RGB8 or RGBA8");
  }
  const size_t row = static_cast<size_t>(image->width) * bpp;
  const size_t need = (static_cast<size_t>(image->height) - 1) * image->row_stride_bytes + row;
  if (image->pixels_size < need) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "pixel buffer is smaller than its rows");
  }
  p->ran = false;
  p->front_done = false;
  p->mark_count = 0;
  GJXL_AMD_TRY(Mark(p, GJXL_AMD_STAGE_BEGIN));
  const uint8_t* srgb = static_cast<const uint8_t*>(image->pixels);
  size_t stride = image->row_stride_bytes;
  if (!image->device_resident) {
    // rows packed into the device copy (pinned host memory gives one DMA)
    hipStream_t st = p->device->stream;
    if (stride == row) {
      GJXL_AMD_TRY(Check(hipMemcpyAsync(p->srgb, srgb, row * image->height,
                                        hipMemcpyHostToDevice, st), "hipMemcpyAsync(sRGB)"));
    } else {
      GJXL_AMD_TRY(Check(hipMemcpy2DAsync(p->srgb, row, srgb, stride, row, image->height,
                                          hipMemcpyHostToDevice, st), "hipMemcpy2DAsync(sRGB)"));
    }
    srgb = p->srgb;
    stride = row;
  }
  GJXL_AMD_TRY(Mark(p, GJXL_AMD_STAGE_UPLOAD));
  *srgb_out = srgb;
  *stride_out = stride;
  return GJXL_OK;
}

}  // namespace

extern "C" {

GJXLResult gjxl_amd_plain_path_run(GJXLAmdPlainPath* p, const GJXLAmdImage* image,
                                   const GJXLAmdTokenOutput* tokens,
                                   GJXLAmdPlainFrame* frame) {
  const uint8_t* srgb = nullptr;
  size_t stride = 0;
  GJXL_AMD_TRY(BeginRun(p, image, &srgb, &stride));
  GJXL_AMD_TRY(RunResident(p, srgb, stride, image->bytes_per_pixel));
  GJXL_AMD_TRY(ReadTokens(p, tokens));
  p->ran = true;
  return FillFrame(p, frame);
}

GJXLResult gjxl_amd_plain_path_run_front(GJXLAmdPlainPath* p, const GJXLAmdImage* image) {
  const uint8_t* srgb = nullptr;
  size_t stride = 0;
  GJXL_AMD_TRY(BeginRun(p, image, &srgb, &stride));
  GJXL_AMD_TRY(RunFront(p, srgb, stride, image->bytes_per_pixel));
  p->front_done = true;
  return GJXL_OK;
}

GJXLResult gjxl_amd_plain_path_cost_frame(GJXLAmdPlainPath* p, GJXLAmdCostFrame* frame) {
  GJXL_AMD_TRY(Validate(p));
  if (frame == nullptr) return Fail(GJXL_ERROR_INVALID_ARGUMENT, "cost frame is null");
  GJXLAmdCostFrame f{};
  f.opsin = p->preprocessed;
  f.quant_field = p->quant_field;
  f.pixel_mask = p->pixel_mask;
  f.y_to_x = p->y_to_x;
  f.y_to_b = p->y_to_b;
  f.allow = p->proposal != nullptr ? p->allow : nullptr;
  f.costs = p->costs;
  f.width = p->pw;
  f.height = p->ph;
  f.butteraugli_target = p->options.distance;
  *frame = f;
  return GJXL_OK;
}

GJXLResult gjxl_amd_plain_path_run_back(GJXLAmdPlainPath* p, int32_t costs_ready,
                                        const GJXLAmdTokenOutput* tokens,
                                        GJXLAmdPlainFrame* frame) {
  GJXL_AMD_TRY(Validate(p));
  if (!p->front_done) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "the front half of this run has not been enqueued");
  }
  p->front_done = false;
  GJXL_AMD_TRY(RunBack(p, costs_ready != 0));
  GJXL_AMD_TRY(ReadTokens(p, tokens));
  p->ran = true;
  return FillFrame(p, frame);
}

GJXLResult gjxl_amd_plain_path_set_proposal(GJXLAmdPlainPath* p, GJXLAmdProposal* net,
                                            const GJXLAmdProposalOptions* options) {
  GJXL_AMD_TRY(Validate(p));
  if (net == nullptr) {
    p->proposal = nullptr;
    return GJXL_OK;
  }
  GJXLAmdProposalOptions o{};
  GJXL_AMD_TRY(gjxl_amd_proposal_options_init(&o, sizeof(o)));
  if (options != nullptr) {
    // a caller built before final_compare passes the shorter struct: its
    // fields over the defaults
    constexpr size_t kFirstSize = offsetof(GJXLAmdProposalOptions, final_compare);
    if (options->struct_size < kFirstSize) {
      return Fail(GJXL_ERROR_INVALID_ARGUMENT, "proposal options are too small");
    }
    std::memcpy(&o, options, std::min(options->struct_size, sizeof(o)));
    o.struct_size = sizeof(o);
  }
  if (!(o.threshold >= 0.0f && o.threshold <= 1.0f) || o.top_k > 7 ||
      !std::isfinite(o.p_min) || o.evaluations < 1 || o.evaluations > p->iterations + 1 ||
      !(o.field_limit >= 0.0f && o.field_limit <= 16.0f) || !std::isfinite(o.field_offset) ||
      !(o.judge_ratio > 0.0f) ||
      !std::isfinite(o.judge_ratio) || (o.final_compare != 0 && o.final_compare != 1)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "proposal options out of range");
  }
  void* head = nullptr;
  size_t bytes = 0;
  GJXL_AMD_TRY(gjxl_amd_proposal_buffer(net, GJXL_AMD_PROPOSAL_HEAD, &head, &bytes));
  const size_t positions = bytes / (4 * GJXL_AMD_PROPOSAL_CHANNELS);
  if (positions < p->capacity[kAllow]) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                "proposal network holds fewer positions than the path's blocks");
  }
  p->proposal = net;
  p->proposal_options = o;
  return GJXL_OK;
}

GJXLResult gjxl_amd_plain_path_proposal_sample(GJXLAmdPlainPath* p, uint16_t* features,
                                               uint8_t* labels, float* targets) {
  GJXL_AMD_TRY(Validate(p));
  if (!p->ran) return Fail(GJXL_ERROR_INVALID_ARGUMENT, "the plain path has not completed a run");
  if (features == nullptr || labels == nullptr || targets == nullptr) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "proposal sample outputs must be non-null");
  }
  GJXL_AMD_TRY(gjxl_amd::LaunchProposalFeatures(p->device, p->preprocessed, p->pw, p->ph,
                                                p->quant_field, features));
  return gjxl_amd::LaunchProposalSample(
    p->device, p->cells, AqBuffer<float>(p->session, GJXL_AMD_AQ_FIELD),
    AqBuffer<float>(p->session, GJXL_AMD_AQ_ADJUSTED_FIELD),
    static_cast<size_t>(p->bw) * p->bh, labels, targets);
}

GJXLResult gjxl_amd_plain_path_stage_times(GJXLAmdPlainPath* p, int32_t* stages,
                                           float* milliseconds, int32_t capacity,
                                           int32_t* count) {
  GJXL_AMD_TRY(Validate(p));
  if (count == nullptr || (capacity > 0 && (stages == nullptr || milliseconds == nullptr))) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "stage time outputs are null");
  }
  if (!p->ran) return Fail(GJXL_ERROR_INVALID_ARGUMENT, "the plain path has not completed a run");
  const int n = p->mark_count - 1;
  *count = n;
  for (int i = 0; i < n && i < capacity; ++i) {
    GJXL_AMD_TRY(Check(hipEventElapsedTime(&milliseconds[i], p->marks[i], p->marks[i + 1]),
                       "hipEventElapsedTime"));
    stages[i] = p->mark_stage[i + 1];
  }
  return GJXL_OK;
}

GJXLResult gjxl_amd_host_alloc(size_t bytes, void** out) {
  if (out == nullptr || *out != nullptr || bytes == 0) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "host allocation output must be empty");
  }
  return Check(hipHostMalloc(out, bytes, hipHostMallocDefault), "hipHostMalloc");
}

void gjxl_amd_host_free(void* pointer) {
  if (pointer != nullptr) (void)hipHostFree(pointer);
}

}  // extern "C"
