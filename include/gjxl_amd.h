/*
 * gjxl_amd: MI355X (gfx950) device layer for gjxl's perceptual-analysis path.
 *
 * A C ABI with plain pointers and sizes.  Device images are planar binary32:
 * plane c of an image with row stride S and height H starts at
 * base + c * S * H.  Padded ("coding") images have width and height rounded
 * up to multiples of 8; block fields have one value per 8x8 block.
 *
 * Every call returns a GJXLResult (codes as in gjxl.h:34-45).  On failure
 * gjxl_amd_last_error() returns this thread's diagnostic, valid until the next
 * failing call on the thread (gjxl.h:247-248 contract).  Stage functions
 * enqueue on the device handle's stream and return without waiting;
 * gjxl_amd_synchronize() waits and reports asynchronous kernel errors.
 *
 * Each stage entry point names the reference interface it replaces
 * (paths relative to the gjxl source tree).
 */
#ifndef GJXL_AMD_H_
#define GJXL_AMD_H_

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__) || defined(__clang__)
#define GJXL_AMD_API __attribute__((visibility("default")))
#else
#define GJXL_AMD_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#ifndef GJXL_GJXL_H_
typedef int32_t GJXLResult;
enum {
  GJXL_OK = 0,
  GJXL_ERROR_INVALID_ARGUMENT = 1,
  GJXL_ERROR_UNSUPPORTED = 2,
  GJXL_ERROR_UNAVAILABLE = 3,
  GJXL_ERROR_OUT_OF_MEMORY = 4,
  GJXL_ERROR_BACKEND = 5,
  GJXL_ERROR_INTERNAL = 6,
  GJXL_ERROR_RESOURCE_PLAN_EXCEEDED = 7,
};
#endif

/* ---------------------------------------------------------------------- */
/* Device handle: one HIP device and one stream.                           */
/* Replaces GpuBackend (src/gpu/backend.h:24-105) for this path: allocation, */
/* copies and submission; CudaBackend binds one device_ordinal the same way */
/* (src/gpu/cuda/cuda_backend.cpp:1043-1059).                                */
/* ---------------------------------------------------------------------- */
typedef struct GJXLAmdDevice GJXLAmdDevice;

GJXL_AMD_API GJXLResult gjxl_amd_device_create(int ordinal,
                                               GJXLAmdDevice** out);
GJXL_AMD_API void gjxl_amd_device_destroy(GJXLAmdDevice* device);
GJXL_AMD_API GJXLResult gjxl_amd_alloc(GJXLAmdDevice* device, size_t bytes,
                                       void** out);
GJXL_AMD_API GJXLResult gjxl_amd_free(GJXLAmdDevice* device, void* pointer);
/* Synchronous copies on the device stream. */
GJXL_AMD_API GJXLResult gjxl_amd_upload(GJXLAmdDevice* device, void* dst,
                                        const void* src, size_t bytes);
GJXL_AMD_API GJXLResult gjxl_amd_download(GJXLAmdDevice* device, void* dst,
                                          const void* src, size_t bytes);
/* Device-to-device copy, enqueued on the device stream. */
GJXL_AMD_API GJXLResult gjxl_amd_copy(GJXLAmdDevice* device, void* dst,
                                      const void* src, size_t bytes);
GJXL_AMD_API GJXLResult gjxl_amd_synchronize(GJXLAmdDevice* device);
GJXL_AMD_API const char* gjxl_amd_last_error(void);

/* Timing on the device stream (HIP events), for measurement.  Slots are
 * 0..GJXL_AMD_EVENT_SLOTS-1; elapsed time is in milliseconds. */
enum { GJXL_AMD_EVENT_SLOTS = 64 };
GJXL_AMD_API GJXLResult gjxl_amd_event_record(GJXLAmdDevice* device,
                                              int slot);
GJXL_AMD_API GJXLResult gjxl_amd_event_elapsed(GJXLAmdDevice* device,
                                               int begin_slot, int end_slot,
                                               float* milliseconds);

/* Per-launch timing of one kernel family inside any sequence on this device
 * handle (the plain path included), for roofline figures: while a kernel id
 * is selected, each launch of that family records a begin and an end event
 * on the device stream (up to GJXL_AMD_PROBE_SLOTS launches between reads).
 * No counterpart in the reference, which times stages on the host.
 * gjxl_amd_device_probe selects (0 = none) and clears; probe_times waits for
 * the recorded launches, writes their durations in launch order and
 * clears. */
enum {
  GJXL_AMD_KERNEL_NONE = 0,
  GJXL_AMD_KERNEL_CANDIDATE_COST = 1,           /* one launch per strategy */
  GJXL_AMD_KERNEL_DC_QUANT = 2,                 /* the DC wavefront */
  GJXL_AMD_KERNEL_BUTTERAUGLI_DIFFERENCES = 3,  /* one launch per scale */
  GJXL_AMD_KERNEL_COUNT = 4
};
enum { GJXL_AMD_PROBE_SLOTS = 64 };
GJXL_AMD_API GJXLResult gjxl_amd_device_probe(GJXLAmdDevice* device,
                                              int32_t kernel);
GJXL_AMD_API GJXLResult gjxl_amd_device_probe_times(GJXLAmdDevice* device,
                                                    float* milliseconds,
                                                    int32_t capacity,
                                                    int32_t* count);

/* ---------------------------------------------------------------------- */
/* a1 + a2: packed sRGB8 -> linear RGB planes and padded XYB planes.        */
/* Replaces PrepareResidentInput (src/gpu/cuda/cuda_input_preparation.cpp:  */
/* 299): ConvertValidatedPackedSrgbInto (src/c_api/image_conversion.cpp:    */
/* 123-145) fused with LinearRgbToPaddedOpsin (src/codec/color_transform.cpp*/
/* :395-431).  srgb: device copy of the packed pixels (3 or 4 bytes/pixel,   */
/* alpha ignored after host validation).  linear: 3*width*height floats.    */
/* opsin: 3*pw*ph floats, pw = ceil8(width), ph = ceil8(height).            */
/* ---------------------------------------------------------------------- */
GJXL_AMD_API GJXLResult gjxl_amd_srgb8_to_linear_and_opsin(
  GJXLAmdDevice* device, const uint8_t* srgb, size_t row_stride_bytes,
  uint32_t bytes_per_pixel, uint32_t width, uint32_t height,
  float intensity_target, float* linear, float* opsin);

/* a2 alone: LinearRgbToPaddedOpsin (src/codec/color_transform_internal.h:17,
 * src/codec/color_transform.cpp:395-431).  linear: 3*width*height floats. */
GJXL_AMD_API GJXLResult gjxl_amd_linear_rgb_to_padded_opsin(
  GJXLAmdDevice* device, const float* linear, uint32_t width,
  uint32_t height, float intensity_target, float* opsin);

/* ---------------------------------------------------------------------- */
/* a4: Gaborish inverse on a padded XYB image (3 planes of w*h floats).     */
/* Replaces ApplyGaborishInverse (src/codec/gaborish.cpp:49-99) behind      */
/* GaborishInverseProvider (src/codec/quantization_pipeline_internal.h:19-  */
/* 73).  w, h: non-zero multiples of 8.  in and out must not alias.         */
/* ---------------------------------------------------------------------- */
GJXL_AMD_API GJXLResult gjxl_amd_gaborish_inverse(
  GJXLAmdDevice* device, const float* opsin, uint32_t width, uint32_t height,
  const float multipliers[3], float* out);

/* ---------------------------------------------------------------------- */
/* a5: initial quant field, strategy mask and blurred pixel mask from the   */
/* pre-Gaborish padded XYB image.  Replaces ComputeInitialQuantField        */
/* (src/codec/adaptive_quantization.cpp:583-738) as driven by               */
/* PreparedAqEvaluation::ComputeInitialQuantization                          */
/* (src/gpu/ops/aq_evaluation.h:219-420).  quant_field, strategy_mask:      */
/* (w/8)*(h/8) floats; pixel_mask: w*h floats; scratch: device memory of    */
/* gjxl_amd_initial_quant_field_scratch_bytes(w, h) bytes.                  */
/* ---------------------------------------------------------------------- */
GJXL_AMD_API size_t gjxl_amd_initial_quant_field_scratch_bytes(
  uint32_t width, uint32_t height);
GJXL_AMD_API GJXLResult gjxl_amd_initial_quant_field(
  GJXLAmdDevice* device, const float* opsin, uint32_t width, uint32_t height,
  float butteraugli_target, float rescale, float* quant_field,
  float* strategy_mask, float* pixel_mask, float* scratch);

/* No counterpart in the reference.  gjxl_amd_initial_quant_field for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): the same kernels, with
 * the workspace in a static device array of the library instead of the
 * scratch argument.  It calls no hipMalloc, hipFree or allocating library
 * routine.  Limit: w * h <= GJXL_AMD_NO_ALLOC_MAX_PIXELS padded pixels;
 * larger extents return GJXL_ERROR_INVALID_ARGUMENT.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (two streams, two
 * device handles, two host threads) race and must be serialized by the
 * caller. */
#define GJXL_AMD_NO_ALLOC_MAX_PIXELS (2048u * 2048u)
GJXL_AMD_API GJXLResult gjxl_amd_initial_quant_field_no_alloc(
  GJXLAmdDevice* device, const float* opsin, uint32_t width, uint32_t height,
  float butteraugli_target, float rescale, float* quant_field,
  float* strategy_mask, float* pixel_mask);

/* ---------------------------------------------------------------------- */
/* a6: initial chroma-from-luma map from the Gaborish-inverted XYB image.   */
/* Replaces ComputeInitialColorCorrelationMap (src/codec/chroma_from_luma.  */
/* cpp:434-529; DCT8 coefficients, 20 Newton steps per 64x64 tile).         */
/* y_to_x, y_to_b: ceil(w/64)*ceil(h/64) int8 each.                         */
/* ---------------------------------------------------------------------- */
GJXL_AMD_API GJXLResult gjxl_amd_initial_cfl(
  GJXLAmdDevice* device, const float* opsin, uint32_t width, uint32_t height,
  int8_t* y_to_x, int8_t* y_to_b);

/* ---------------------------------------------------------------------- */
/* a7-a9: AC-strategy candidate costs.  Replaces                             */
/* GpuAcStrategyEvaluation::EvaluateAcStrategyCandidateBatches               */
/* (src/gpu/ops/ac_strategy.h:118-164), each cost equal to                   */
/* EstimateAcStrategyCost (src/codec/ac_strategy.cpp:823-994).  costs: 7     */
/* planes of (w/8)*(h/8) floats in the stage order of kCandidateStages       */
/* (src/codec/ac_strategy_search_policy.h:20-28: DCT8, 16x8, 8x16, 16x16,    */
/* 32x16, 16x32, 32x32); every anchor whose footprint lies inside its 64x64  */
/* tile is evaluated (the 32-pixel families on even tile positions); other   */
/* entries are NaN.                                                          */
/* ---------------------------------------------------------------------- */
GJXL_AMD_API GJXLResult gjxl_amd_ac_strategy_costs(
  GJXLAmdDevice* device, const float* opsin, uint32_t width, uint32_t height,
  const float* quant_field, const float* pixel_mask, const int8_t* y_to_x,
  const int8_t* y_to_b, float butteraugli_target, float* costs);
/* The learned form's pruned costs (no counterpart in the reference): as
 * gjxl_amd_ac_strategy_costs, but a candidate whose stage's bit is clear in
 * allow at its anchor ((w/8)*(h/8) bytes, bit s = stage s of the order
 * above, from gjxl_amd_proposal_decide) costs +infinity, which the greedy
 * merge never selects; allowed candidates cost what
 * gjxl_amd_ac_strategy_costs gives them, bit for bit.  Only the allowed
 * ones are evaluated: the larger transforms' allowed anchors are listed on
 * the device and evaluated several to a work group.  Scratch for the lists
 * (about 4 bytes per candidate position) comes from the device's pool in
 * stream order.  A null allow evaluates every candidate. */
GJXL_AMD_API GJXLResult gjxl_amd_ac_strategy_costs_pruned(
  GJXLAmdDevice* device, const float* opsin, uint32_t width, uint32_t height,
  const float* quant_field, const float* pixel_mask, const int8_t* y_to_x,
  const int8_t* y_to_b, float butteraugli_target, const uint8_t* allow,
  float* costs);

/* ---------------------------------------------------------------------- */
/* a10: greedy merge of the candidate costs into a strategy grid.  Replaces  */
/* FindAcStrategyGridFromCandidateCosts (src/codec/ac_strategy_search_      */
/* internal.h:24, src/codec/ac_strategy.cpp:591-773) on the device, one      */
/* thread per 64x64 tile.  cells: (w/8)*(h/8) bytes, strategy << 1 | anchor  */
/* (AcStrategyGrid::CreateFromEncodedCells, src/core/ac_strategy.h:196).     */
/* ---------------------------------------------------------------------- */
GJXL_AMD_API GJXLResult gjxl_amd_ac_strategy_search(
  GJXLAmdDevice* device, const float* costs, uint32_t width, uint32_t height,
  float butteraugli_target, uint8_t* cells);

/* ---------------------------------------------------------------------- */
/* a3, a11-a15: adaptive-quantization evaluation session.  Replaces         */
/* PreparedAqEvaluation (src/gpu/ops/aq_evaluation.h:219-420) and the        */
/* GpuQuantizationEvaluationSession behind it (src/gpu/ops/aq_evaluation.h:  */
/* 112-160): the session owns the device storage of one image's evaluation.  */
/* width, height: the source extent; the coding extent is ceil8 of each.     */
/* ---------------------------------------------------------------------- */
typedef struct GJXLAmdAqSession GJXLAmdAqSession;

/* Buffer ids for gjxl_amd_aq_buffer.  Block fields are (pw/8)*(ph/8);
 * coefficient stores hold 3 planes of groups*65536 values, one 65536-slot
 * region per 256x256 AC group with its transforms in raster order of their
 * anchors (the reference's per-group coefficient layout, src/codec/
 * reconstruction.cpp:386-760). */
enum {
  GJXL_AMD_AQ_OFFSETS = 0,          /* u32 per block: store offset or ~0 */
  GJXL_AMD_AQ_ANCHOR_LIST,          /* u32 groups*1024: by << 16 | bx */
  GJXL_AMD_AQ_ANCHOR_COUNT,         /* u32 per group */
  GJXL_AMD_AQ_COEFFICIENTS,         /* f32 forward DCT of every transform */
  GJXL_AMD_AQ_ADJUSTED_FIELD,       /* f32 AdjustQuantField output */
  GJXL_AMD_AQ_STATS,                /* u32[16]: matrix stats 0-2, range 4-5,
                                       grid error 6, strategy counts 8-14 */
  GJXL_AMD_AQ_SELECT,               /* u32[4 + 256]: median / MAD selection */
  GJXL_AMD_AQ_QUANTIZER,            /* quantizer state (10 words) */
  GJXL_AMD_AQ_RAW_QUANT,            /* i32 per block: adjusted raw quant */
  GJXL_AMD_AQ_Y_TO_X,               /* i8 per 64x64 tile: final CfL */
  GJXL_AMD_AQ_Y_TO_B,
  GJXL_AMD_AQ_QUANTIZED,            /* i32 quantized AC, store layout */
  GJXL_AMD_AQ_DC,                   /* f32 3 block planes: DC before coding */
  GJXL_AMD_AQ_DC_QUANTIZED,         /* i32 3 block planes */
  GJXL_AMD_AQ_DC_RECONSTRUCTED,     /* f32 3 block planes */
  GJXL_AMD_AQ_DC_SMOOTHED,          /* f32 3 block planes */
  GJXL_AMD_AQ_INVERSE_SIGMA,        /* f32 per block: EPF inverse sigma */
  GJXL_AMD_AQ_RECONSTRUCTED,        /* f32 3 planes pw*ph: XYB */
  GJXL_AMD_AQ_FILTERED,             /* f32 3 planes w*h: after EPF pass 2 */
  GJXL_AMD_AQ_LINEAR,               /* f32 3 planes w*h: linear RGB */
  GJXL_AMD_AQ_FIELD,                /* f32 per block: gjxl_amd_aq_run field */
  GJXL_AMD_AQ_BLOCK_DISTANCE,       /* f32 per block: last 16-norms */
  GJXL_AMD_AQ_SCORES,               /* u32[64]: score bits per evaluation */
  GJXL_AMD_AQ_STRATEGY_LISTS,       /* u32 7 x blocks: anchors per strategy */
  GJXL_AMD_AQ_BUFFER_COUNT
};

GJXL_AMD_API GJXLResult gjxl_amd_aq_session_create(GJXLAmdDevice* device,
                                                   uint32_t width,
                                                   uint32_t height,
                                                   GJXLAmdAqSession** out);
GJXL_AMD_API void gjxl_amd_aq_session_destroy(GJXLAmdAqSession* session);

/* Per-encode preparation (the inputs stay borrowed until the session is
 * destroyed or prepared again): quantization-matrix scale statistics and
 * selection (src/codestream/workflow.cpp:1318-1484), AdjustQuantField
 * (src/codec/adaptive_quantization.cpp:740-835), the Butteraugli policy
 * bounds (:1178-1242) and initial DC quant (:556-580), the AC-group layout
 * of the strategy grid and the forward DCT of every transform.
 * opsin: padded pre-Gaborish XYB; preprocessed: padded Gaborish-inverted
 * XYB; cells: strategy cells as gjxl_amd_ac_strategy_search writes them;
 * initial_quant_field: block field before adjustment.  Waits for the stream
 * (the matrix scales are chosen on the host). */
GJXL_AMD_API GJXLResult gjxl_amd_aq_prepare(
  GJXLAmdAqSession* session, const float* opsin, const float* preprocessed,
  const uint8_t* cells, const float* initial_quant_field,
  float butteraugli_target);

/* The host-side policy of the last preparation. */
GJXL_AMD_API GJXLResult gjxl_amd_aq_policy(GJXLAmdAqSession* session,
                                           float* lower_bound,
                                           float* upper_bound,
                                           float* quant_dc, int32_t* x_scale,
                                           int32_t* b_scale);

/* One evaluation of a quant field (device, (pw/8)*(ph/8) floats), up to the
 * linear-RGB decode: CreateQuantizerFromField (src/codec/quantization.cpp:
 * 536-639), ComputeFinalColorCorrelationMap (src/codec/chroma_from_luma.cpp:
 * 543-670), ComputeQuantizedCoefficients with prediction-aware DC
 * (src/codec/reconstruction.cpp:386-760, src/codec/vardct/dc_quantization.
 * cpp:121-240), ReconstructQuantizedCoefficients with DC smoothing
 * (src/codec/reconstruction.cpp:790-936), ApplyLoopFilters (src/codec/
 * loop_filter.cpp:14-65) and OpsinToLinearRgb (src/codec/color_transform.
 * cpp:433-471): EvaluateQuantization (src/codec/adaptive_quantization.cpp:
 * 940-1010) without the distance.  Enqueued on the device stream. */
GJXL_AMD_API GJXLResult gjxl_amd_aq_evaluate(GJXLAmdAqSession* session,
                                             const float* quant_field);

/* ---------------------------------------------------------------------- */
/* Chosen precision form (not the plain path): the forward DCT as batched   */
/* matrix products on MFMA in binary16 with split hi + lo operands and      */
/* binary32 accumulation (hi*hi + hi*lo + lo*hi), in place of the binary64 */
/* bases of ForwardDctCpu (src/codec/dct.cpp:135-182).  No counterpart in   */
/* the reference, whose GPU DCTs are binary64 (src/gpu/cuda/                */
/* cuda_exact_arithmetic_test.cpp:161-186).                                 */
/* ---------------------------------------------------------------------- */
/* count blocks of width x height floats (row-major, contiguous; sides 8,
 * 16, 32 or 64) into count blocks of coefficients in the storage order of
 * src/core/ac_strategy.h:95-108 (the longer side contiguous).  Enqueued. */
GJXL_AMD_API GJXLResult gjxl_amd_split_dct_blocks(GJXLAmdDevice* device,
                                                  uint32_t width, uint32_t height,
                                                  const float* blocks, size_t count,
                                                  float* coefficients);
/* Rewrites GJXL_AMD_AQ_COEFFICIENTS of a prepared session with the split
 * form of every transform's forward DCT (from the Gaborish-inverted image
 * the session was prepared with); later evaluations code those
 * coefficients.  Enqueued. */
GJXL_AMD_API GJXLResult gjxl_amd_aq_split_forward_dct(GJXLAmdAqSession* session);

/* Device pointer and size in bytes of one session buffer. */
GJXL_AMD_API GJXLResult gjxl_amd_aq_buffer(GJXLAmdAqSession* session,
                                           int32_t id, void** pointer,
                                           size_t* bytes);

/* ---------------------------------------------------------------------- */
/* a16: Butteraugli with a prepared reference.  Replaces                    */
/* PreparedButteraugliReference::{Prepare, Compare} (src/codec/butteraugli.h:*/
/* 27-54, src/codec/butteraugli.cpp:31-71; the native scalar pipeline of     */
/* src/codec/butteraugli_distance_internal.cpp:872-1202) and                 */
/* DeviceButteraugliOperation (src/gpu/ops/butteraugli.h:110).  Images are   */
/* 3 planes of width*height linear RGB floats on the device.                 */
/* ---------------------------------------------------------------------- */
typedef struct GJXLAmdButteraugli GJXLAmdButteraugli;

enum {
  GJXL_AMD_BUTTERAUGLI_MAP = 0,          /* f32 width*height: last map */
  GJXL_AMD_BUTTERAUGLI_REFERENCE_PSYCHO, /* 10 planes: LF(3) MF(3) HF(2) UHF(2) */
  GJXL_AMD_BUTTERAUGLI_DISTORTED_PSYCHO,
  GJXL_AMD_BUTTERAUGLI_STAGES,           /* 9 planes, DifferenceStage order */
  GJXL_AMD_BUTTERAUGLI_XYB,              /* 3 planes: last opsin image */
  GJXL_AMD_BUTTERAUGLI_SCORE             /* u32: last score's float bits */
};

GJXL_AMD_API GJXLResult gjxl_amd_butteraugli_create(
  GJXLAmdDevice* device, uint32_t width, uint32_t height, float hf_asymmetry,
  float x_multiplier, float intensity_target, GJXLAmdButteraugli** out);
GJXL_AMD_API void gjxl_amd_butteraugli_destroy(GJXLAmdButteraugli* b);
GJXL_AMD_API GJXLResult gjxl_amd_butteraugli_prepare(
  GJXLAmdButteraugli* b, const float* reference_linear);
/* Enqueues one comparison; distance_map (device, width*height) may be null
 * (the map stays in the object's own buffer).  Leaves every difference
 * stage in the STAGES buffer; the AQ loop's comparisons (gjxl_amd_aq_run,
 * the plain path) write only its final map (stage 8) and stage 6. */
GJXL_AMD_API GJXLResult gjxl_amd_butteraugli_compare(
  GJXLAmdButteraugli* b, const float* distorted_linear, float* distance_map);
/* Waits for the stream and returns the last comparison's score (the map's
 * maximum). */
GJXL_AMD_API GJXLResult gjxl_amd_butteraugli_score(GJXLAmdButteraugli* b,
                                                   double* score);
/* A device buffer of scale 0 (full) or 1 (half resolution, present from 15
 * pixels in both axes): see GJXL_AMD_BUTTERAUGLI_*. */
GJXL_AMD_API GJXLResult gjxl_amd_butteraugli_buffer(GJXLAmdButteraugli* b,
                                                    int32_t scale, int32_t id,
                                                    void** pointer,
                                                    size_t* bytes);

/* ---------------------------------------------------------------------- */
/* a17: per-transform 16-norm of a distance map.  Replaces                   */
/* ReduceButteraugliDistanceMap (src/codec/butteraugli.cpp:116-222).  map:   */
/* width*height floats; cells and block_distance: ceil(w/8)*ceil(h/8).       */
/* ---------------------------------------------------------------------- */
GJXL_AMD_API GJXLResult gjxl_amd_reduce_distance_map(
  GJXLAmdDevice* device, const float* map, uint32_t width, uint32_t height,
  const uint8_t* cells, float* block_distance);

/* ---------------------------------------------------------------------- */
/* a18: the Butteraugli AQ loop.  Replaces RunAdaptiveQuantizationPolicyImpl */
/* (src/codec/adaptive_quantization.cpp:1600-1738) behind                    */
/* PreparedAqEvaluation::EvaluateResidentButteraugliPolicy (src/gpu/ops/     */
/* aq_evaluation.h:219-420).                                                 */
/* gjxl_amd_aq_update: one field update after evaluation `iteration` of a    */
/* prepared session, with the session's last quantizer, adjusted initial     */
/* field and bounds (device field and block_distance, (pw/8)*(ph/8)).        */
/* gjxl_amd_aq_run: iterations + 1 evaluations from the adjusted initial     */
/* field, each compared with Butteraugli (b prepared with the source linear  */
/* RGB) and reduced, with updates between them.  The final field, block      */
/* distances and the last evaluation's frame stay in the session             */
/* (GJXL_AMD_AQ_FIELD, GJXL_AMD_AQ_BLOCK_DISTANCE, ...); scores (host,       */
/* iterations + 1 doubles) may be null, otherwise the call waits for them.   */
/* ---------------------------------------------------------------------- */
GJXL_AMD_API GJXLResult gjxl_amd_aq_update(GJXLAmdAqSession* session,
                                           float* quant_field,
                                           const float* block_distance,
                                           int32_t iteration,
                                           float butteraugli_target);
GJXL_AMD_API GJXLResult gjxl_amd_aq_run(GJXLAmdAqSession* session,
                                        GJXLAmdButteraugli* butteraugli,
                                        int32_t iterations,
                                        float butteraugli_target,
                                        double* scores);
/* Waits for the stream and fails with GJXL_ERROR_INTERNAL if an update since
 * the last preparation met a non-finite or negative block distance (the
 * reference's Status::Internal, src/codec/adaptive_quantization.cpp:1686-
 * 1695); such blocks keep their field value. */
GJXL_AMD_API GJXLResult gjxl_amd_aq_check(GJXLAmdAqSession* session);

/* ---------------------------------------------------------------------- */
/* a19: AC tokenization of a quantized frame.  Replaces                      */
/* AcTokenizationProvider::{Begin, Finish} (src/codestream/                  */
/* ac_tokenization_provider_internal.h:18-31) over the per-group tokenizer   */
/* TokenizeSimpleAcGroupDirectValidated (src/codestream/ac_group.cpp:        */
/* 740-932): per transform and channel (Y, X, B) the non-LLF nonzero count   */
/* with its predicted-nonzero context, then the zig-zag signed coefficients  */
/* along the coefficient order up to the last nonzero with zero-density      */
/* contexts.  The stream of every AC group is the reference's; the groups    */
/* are concatenated in group order (GJXL_AMD_TOKENS_GROUP_OFFSETS).  Tokens  */
/* are SoA: u32 values and u16 contexts (src/codestream/entropy.h:33-39).    */
/* A tokenizer starts with the natural orders and the default block-context  */
/* map (src/codestream/block_context_map.cpp:245-252).                       */
/* ---------------------------------------------------------------------- */
typedef struct GJXLAmdTokenizer GJXLAmdTokenizer;

enum {
  GJXL_AMD_TOKENS_VALUES = 0,       /* u32 per token */
  GJXL_AMD_TOKENS_CONTEXTS,         /* u16 per token */
  GJXL_AMD_TOKENS_GROUP_OFFSETS,    /* u32 groups + 1: first token per group,
                                       then the total */
  GJXL_AMD_TOKENS_HISTOGRAM,        /* u32 (num_contexts * 495) x 128:
                                       tokens per context and symbol */
  GJXL_AMD_TOKENS_NONZERO_MAP,      /* u8 3 block planes: scaled nonzeros */
  GJXL_AMD_TOKENS_TASKS,            /* u32 groups x 3072: nonzeros | tokens << 16
                                       per (anchor, channel) in stream order */
  GJXL_AMD_TOKENS_TASK_OFFSETS,     /* u32 groups x 3072: first token in group */
  GJXL_AMD_TOKENS_GROUP_TOKENS,     /* u32 per group */
  GJXL_AMD_TOKENS_OFFSETS,          /* u32 per block: anchor layout */
  GJXL_AMD_TOKENS_ANCHOR_LIST,      /* u32 groups*1024: by << 16 | bx */
  GJXL_AMD_TOKENS_ANCHOR_COUNT,     /* u32 per group */
  GJXL_AMD_TOKENS_ORDERS,           /* u32 3 x 2624: scan -> coefficient */
  GJXL_AMD_TOKENS_INVERSE_ORDERS,   /* u16 3 x 2624: coefficient -> scan */
  GJXL_AMD_TOKENS_STATUS,           /* u32[4]: total tokens, error bits */
  GJXL_AMD_TOKENS_STATISTICS,       /* u32: frame statistics (GJXLAmdFrame-
                                       Statistics order) + present mask + error */
  GJXL_AMD_TOKENS_SAMPLE_DECISIONS, /* u8 per block: effort-7 DCT8 sample */
  GJXL_AMD_TOKENS_BUFFER_COUNT
};

/* width, height: the source extent (blocks of its 8-padded extent). */
GJXL_AMD_API GJXLResult gjxl_amd_tokenizer_create(GJXLAmdDevice* device,
                                                  uint32_t width,
                                                  uint32_t height,
                                                  GJXLAmdTokenizer** out);
GJXL_AMD_API void gjxl_amd_tokenizer_destroy(GJXLAmdTokenizer* tokenizer);

/* ComputeSimpleNaturalCoefficientOrder (src/codestream/ac_group.cpp:381-487)
 * on the host: scan position -> coefficient index of one of the seven
 * searched strategies. */
GJXL_AMD_API GJXLResult gjxl_amd_natural_order(int32_t strategy,
                                               uint32_t* order,
                                               size_t capacity, size_t* count);

/* The coefficient orders of SimpleCoefficientOrders (src/codestream/
 * coefficient_order.h:32-41): for each order family f set in
 * used_order_mask, orders[3 * f + c] (host, channel c = X, Y, B) replaces the
 * natural order; it must be a permutation with the LLF prefix, as
 * ValidateSimpleCoefficientOrders requires.  orders may be null when the
 * mask is 0. */
GJXL_AMD_API GJXLResult gjxl_amd_tokenizer_set_orders(
  GJXLAmdTokenizer* tokenizer, uint16_t used_order_mask,
  const uint32_t* const* orders);

/* A SimpleBlockContextMap (src/codestream/block_context_map.h): at most 15
 * increasing quant-field thresholds in [1, 255], 3 * 13 * (thresholds + 1)
 * labels, num_contexts in [1, 16] with every label used
 * (ValidateSimpleBlockContextMap, src/codestream/block_context_map.cpp:
 * 281-320). */
GJXL_AMD_API GJXLResult gjxl_amd_tokenizer_set_context_map(
  GJXLAmdTokenizer* tokenizer, const uint32_t* qf_thresholds,
  uint32_t threshold_count, const uint8_t* context_map,
  uint32_t num_contexts);

/* Tokenizes a frame on the device: quantized AC in the per-group store
 * layout (3 planes of groups * 65536 i32, as GJXL_AMD_AQ_QUANTIZED), the
 * strategy cells and the raw quant field (i32 per block).  populations != 0
 * also fills GJXL_AMD_TOKENS_HISTOGRAM (the sum of the groups' fixed-ANS
 * populations, AppendDirectAcToken, ac_group.cpp:688-720).  Waits for the
 * stream; *token_count receives the total. */
GJXL_AMD_API GJXLResult gjxl_amd_tokenize(GJXLAmdTokenizer* tokenizer,
                                          const int32_t* quantized,
                                          const uint8_t* cells,
                                          const int32_t* raw_quant,
                                          int32_t populations,
                                          uint64_t* token_count);

GJXL_AMD_API GJXLResult gjxl_amd_tokenizer_buffer(GJXLAmdTokenizer* tokenizer,
                                                  int32_t id, void** pointer,
                                                  size_t* bytes);

/* ---------------------------------------------------------------------- */
/* f1/f2: the CPU-chosen inputs of AC tokenization, from frame statistics   */
/* gathered on the device.  The reference counts, on the CPU, the zero      */
/* coefficients of every order family, channel and position                  */
/* (CountGroupZeros, src/codestream/coefficient_order.cpp:223-316; at        */
/* effort 7 over a xorshift128+ sample of the transforms when the frame is   */
/* DCT8 only, :626-650) and the blocks per order family and raw quant       */
/* (CountRawQuantAndOrders, src/codestream/block_context_map.cpp:67-103),   */
/* then sorts the orders (:736-811) and clusters the block contexts         */
/* (BuildAdaptiveMap, :126-200, via ComputeSimpleBlockContextMapForEncoder, */
/* :435-476).  Here the counting is one kernel over the quantized store and */
/* the sorting and clustering run on the host from 38 KB of counts.         */
/* ---------------------------------------------------------------------- */
enum {
  GJXL_AMD_ORDER_FAMILIES = 13,     /* kSimpleCoefficientOrderCount */
  GJXL_AMD_ZERO_COUNTS = 1984,      /* counters per channel, see below */
  GJXL_AMD_MAX_QF_THRESHOLDS = 15,
  GJXL_AMD_MAX_BLOCK_CONTEXTS = 16
};

/* zero_counts[c] (c = X, Y, B): the families of the searched strategies at
 * offsets DCT8 0 (64 positions), DCT16x8 and 8x16 64 (128), DCT16x16 192
 * (256), DCT32x16 and 16x32 448 (512), DCT32x32 960 (1024); position k is
 * the coefficient index in the transform's storage layout.
 * sampled_dct8_zero_counts: the DCT8 counts over the sampled transforms
 * only (meaningful when DCT8 is the only family present).
 * raw_quant_counts[f][q - 1]: blocks of order family f with raw quant q. */
typedef struct {
  uint32_t block_width, block_height;
  uint16_t present_mask;            /* order families with anchors */
  uint32_t zero_counts[3][GJXL_AMD_ZERO_COUNTS];
  uint32_t sampled_dct8_zero_counts[3][64];
  uint32_t raw_quant_counts[GJXL_AMD_ORDER_FAMILIES][256];
} GJXLAmdFrameStatistics;

/* SimpleCoefficientOrders (src/codestream/coefficient_order.h:32-41) and
 * SimpleBlockContextMap (src/codestream/block_context_map.h): for each
 * family f in used_order_mask, orders[f][c] holds its first n entries
 * (scan position -> coefficient index, n = the family's size); the map is
 * 3 * 13 * (qf_threshold_count + 1) labels. */
typedef struct {
  uint16_t used_order_mask;
  uint32_t orders[GJXL_AMD_ORDER_FAMILIES][3][1024];
  uint32_t qf_threshold_count;
  uint32_t qf_thresholds[GJXL_AMD_MAX_QF_THRESHOLDS];
  uint32_t num_contexts;
  uint8_t context_map[3 * GJXL_AMD_ORDER_FAMILIES * (GJXL_AMD_MAX_QF_THRESHOLDS + 1)];
} GJXLAmdEntropyInputs;

/* The statistics of a quantized frame (inputs as gjxl_amd_tokenize).
 * Waits for the stream.  Raw quant outside [1, 256] is
 * GJXL_ERROR_INVALID_ARGUMENT, as CountRawQuantAndOrders has it. */
GJXL_AMD_API GJXLResult gjxl_amd_tokenizer_frame_statistics(
  GJXLAmdTokenizer* tokenizer, const int32_t* quantized, const uint8_t* cells,
  const int32_t* raw_quant, GJXLAmdFrameStatistics* statistics);

/* Host only.  The orders of ComputeSimpleCoefficientOrdersForEncoder with
 * the behaviour of the effort (effort 7: kEffort7Dct8Sampled, otherwise
 * kFull; src/codestream/workflow.cpp:2066-2072) and the map of
 * ComputeSimpleBlockContextMapForEncoder.  out is written only on success. */
GJXL_AMD_API GJXLResult gjxl_amd_entropy_inputs_from_statistics(
  const GJXLAmdFrameStatistics* statistics, int32_t effort,
  GJXLAmdEntropyInputs* out);

/* gjxl_amd_tokenizer_set_orders and gjxl_amd_tokenizer_set_context_map
 * with the inputs above. */
GJXL_AMD_API GJXLResult gjxl_amd_tokenizer_set_entropy_inputs(
  GJXLAmdTokenizer* tokenizer, const GJXLAmdEntropyInputs* inputs);

/* Host only.  TokenizeSimpleCoefficientOrders (src/codestream/
 * coefficient_order.cpp:876-919): per used family and channel, the Lehmer
 * code of the order relative to the natural order, as {context, value}
 * tokens.  *count receives the number of tokens; capacity is checked
 * first (at most 3 * 1984 + 15 tokens). */
GJXL_AMD_API GJXLResult gjxl_amd_coefficient_order_tokens(
  const GJXLAmdEntropyInputs* inputs, uint32_t* contexts, uint32_t* values,
  size_t capacity, size_t* count);

/* ---------------------------------------------------------------------- */
/* f2: sparse read-back of quantized AC coefficients.  Replaces             */
/* LaunchCudaSparseAc (src/gpu/cuda/cuda_sparse_ac_kernels.h:11-17) in the  */
/* format SparseCoefficientSpan reads (src/codec/sparse_coefficients.h:     */
/* 90-107): for count values of `width` bytes (1, 2 or 4, signed) at       */
/* source, masks[ceil(count / 64)] hold the nonzero positions of each 64   */
/* values, offsets[] the absolute payload offset of each word's first      */
/* nonzero, values[] the nonzero payloads (coefficient order within a      */
/* word; ranges of different work groups in any order), *total their       */
/* count.  values needs room for count payloads.  Unlike the reference,    */
/* total is cleared here (by a kernel).  Enqueued; source and values must  */
/* not overlap.                                                             */
/* ---------------------------------------------------------------------- */
GJXL_AMD_API GJXLResult gjxl_amd_sparse_ac(GJXLAmdDevice* device, const void* source,
                                           uint32_t count, uint32_t width, uint64_t* masks,
                                           uint32_t* offsets, void* values, uint32_t* total);

/* ---------------------------------------------------------------------- */
/* f3: rate control.  Replaces SearchTargetSize (src/codestream/rate_control */
/* .cpp:246-337, src/codestream/rate_control_internal.h:35-86), which       */
/* gjxl's workflow drives for a target file size (src/codestream/workflow.  */
/* cpp:1683-1835): a bounded search over the Butteraugli target (distance)  */
/* by complete encodes, the lower and upper targets first, then midpoints   */
/* of the widest remaining interval.  The caller's evaluator encodes at one */
/* target and reports the size; the search keeps the best candidate by the  */
/* reference's selection rules.  Host code; nothing is enqueued.            */
/* ---------------------------------------------------------------------- */
enum {
  GJXL_AMD_RATE_LARGEST_AT_OR_BELOW = 0,  /* TargetSizeSelectionPolicy */
  GJXL_AMD_RATE_CLOSEST_ABSOLUTE = 1
};

typedef struct {
  size_t struct_size;
  uint64_t target_bytes;     /* > 0 */
  uint64_t tolerance_bytes;
  uint32_t maximum_attempts; /* 1-64, default 12 */
  int32_t selection;         /* default largest at or below */
  float minimum_target;      /* default 0.01 */
  float maximum_target;      /* default 10 */
} GJXLAmdRateControlOptions;

/* One encode's outcome.  final_score breaks size ties (lower wins) only
 * when score_evaluated is non-zero and resident is zero: the score of a
 * resident GPU loop is a diagnostic in the reference. */
typedef struct {
  uint64_t encoded_bytes;    /* > 0 */
  int32_t score_evaluated;
  int32_t resident;
  double final_score;        /* finite when evaluated */
} GJXLAmdRateCandidate;

/* Encodes at butteraugli_target into *out.  A failure is local to that
 * candidate, except GJXL_ERROR_RESOURCE_PLAN_EXCEEDED, which ends the
 * search with that code. */
typedef GJXLResult (*GJXLAmdRateEvaluator)(void* user, float butteraugli_target,
                                           GJXLAmdRateCandidate* out);

typedef struct {
  float selected_target;
  GJXLAmdRateCandidate selected;
  uint32_t attempts, failed_attempts;
  int32_t target_size_met, search_exhausted;
} GJXLAmdRateControlResult;

GJXL_AMD_API GJXLResult gjxl_amd_rate_control_options_init(GJXLAmdRateControlOptions* options,
                                                           size_t struct_size);
/* On failure *result is left unchanged. */
GJXL_AMD_API GJXLResult gjxl_amd_rate_control_search(const GJXLAmdRateControlOptions* options,
                                                     GJXLAmdRateEvaluator evaluator,
                                                     void* user,
                                                     GJXLAmdRateControlResult* result);

/* ---------------------------------------------------------------------- */
/* Chosen batching form (not the plain path; no counterpart in the          */
/* reference, whose batching is host-level: src/codestream/batch_workflow.h */
/* :47-109): the candidate costs of a batch of frames of uneven extents as  */
/* persistent work groups pulling (frame, 64x64 tile, chunk of anchors)     */
/* items of one strategy from a work queue in device memory, one captured   */
/* hipGraph per queue (plan, NaN fill, seven strategies) replayed for every */
/* batch, and an autotuner that times persistent grids per strategy on a    */
/* batch and keeps the fastest.  Each frame's costs equal                   */
/* gjxl_amd_ac_strategy_costs_pruned's for the same inputs.                 */
/* ---------------------------------------------------------------------- */
typedef struct GJXLAmdCostQueue GJXLAmdCostQueue;

/* One frame of a batch: the arguments of gjxl_amd_ac_strategy_costs_pruned
 * (device pointers; allow may be null). */
typedef struct {
  const float* opsin;
  const float* quant_field;
  const float* pixel_mask;
  const int8_t* y_to_x;
  const int8_t* y_to_b;
  const uint8_t* allow;
  float* costs;
  uint32_t width, height;  /* padded, multiples of 8, at most 65536 */
  float butteraugli_target;
} GJXLAmdCostFrame;

/* A queue for batches of up to max_frames (at most 256) frames on the
 * device's stream; its graph is captured here, with persistent grids of
 * two groups per CU for every strategy until tuned. */
GJXL_AMD_API GJXLResult gjxl_amd_cost_queue_create(GJXLAmdDevice* device,
                                                   uint32_t max_frames,
                                                   GJXLAmdCostQueue** out);
GJXL_AMD_API void gjxl_amd_cost_queue_destroy(GJXLAmdCostQueue* queue);
/* Uploads the batch's table (waiting for the previous upload to finish
 * reading the host copy) and launches the graph.  Enqueued: the buffers
 * must stay valid until the stream reaches the end of the graph. */
GJXL_AMD_API GJXLResult gjxl_amd_cost_queue_run(GJXLAmdCostQueue* queue,
                                                const GJXLAmdCostFrame* frames,
                                                uint32_t count);
/* Times each strategy's queue on this batch (repetitions runs per
 * variant, after one warm-up) with persistent grids of one to four groups
 * per CU and with one group per item, keeps the fastest grid per strategy
 * and captures the graph again.  Synchronous; writes the frames' costs. */
GJXL_AMD_API GJXLResult gjxl_amd_cost_queue_tune(GJXLAmdCostQueue* queue,
                                                 const GJXLAmdCostFrame* frames,
                                                 uint32_t count, uint32_t repetitions);
/* The grid of each strategy (7, stage order) and its tuned time in ms
 * (0 before tuning; milliseconds may be null). */
GJXL_AMD_API GJXLResult gjxl_amd_cost_queue_grids(GJXLAmdCostQueue* queue,
                                                  uint32_t* grids, float* milliseconds);

/* ---------------------------------------------------------------------- */
/* The whole plain path as one call.  Replaces, for one image, the stage     */
/* sequence that gjxl_encode (src/c_api/gjxl.cpp:470-637) runs through       */
/* PrepareWorkflow (src/codestream/workflow.cpp:636-768) and                 */
/* RunPreparedQuantizationPipelineWithProviders (src/codec/quantization_     */
/* pipeline.cpp:428-586) with the GPU providers installed, up to the         */
/* AcTokenizationProvider hand-off (src/codestream/ac_tokenization_provider_ */
/* internal.h:18-47): upload, XYB, Gaborish inverse, initial quant field,    */
/* initial CfL, AC strategy search, AQ preparation, three Butteraugli-driven */
/* evaluations with two field updates (effort 7), AC tokenization of the     */
/* last evaluation and the read-back of its tokens.  One stream; the host    */
/* waits for the matrix-scale statistics and for the token count.  The       */
/* entropy coder and bitstream writer that consume the frame stay on the     */
/* CPU (src/codestream/vardct/encoder.cpp).                                  */
/* ---------------------------------------------------------------------- */
typedef struct GJXLAmdPlainPath GJXLAmdPlainPath;

/* Subset of GJXLEncoderOptions (include/gjxl/gjxl.h:129-160) that the path
 * reads: distance and effort (7 only; other efforts return
 * GJXL_ERROR_UNSUPPORTED), plus the input's intensity target.  The DC
 * options are those of effort 7: weighted prediction, prediction-aware
 * quantization, adaptive smoothing. */
typedef struct {
  size_t struct_size;
  float distance;          /* default 1.0 */
  int32_t effort;          /* default 7 */
  float intensity_target;  /* default 255 */
} GJXLAmdPlainPathOptions;

/* A packed RGB8 or RGBA8 sRGB image (GJXLImageView, include/gjxl/gjxl.h:
 * 176-184), on the host (pinned memory from gjxl_amd_host_alloc gives one
 * DMA copy) or, when device_resident is non-zero, in device memory. */
typedef struct {
  uint32_t width, height;
  uint32_t bytes_per_pixel;  /* 3 or 4 (alpha ignored) */
  int32_t device_resident;
  const void* pixels;
  size_t pixels_size;
  size_t row_stride_bytes;
} GJXLAmdImage;

/* Host buffers for the token stream.  Null pointers skip that read-back;
 * capacity (tokens) must cover the frame's count, group_capacity its
 * groups + 1 offsets. */
typedef struct {
  uint32_t* values;
  uint16_t* contexts;
  uint32_t* group_offsets;
  size_t capacity;
  size_t group_capacity;
} GJXLAmdTokenOutput;

/* The frame of the last run: device pointers into the path's storage, valid
 * until the next run or destroy.  Layouts as the session and tokenizer
 * buffers above (GJXL_AMD_AQ_*, GJXL_AMD_TOKENS_*). */
typedef struct {
  uint32_t block_width, block_height, groups;
  const uint8_t* cells;             /* strategy << 1 | anchor per block */
  const float* quant_field;         /* the last evaluation's field */
  const float* block_distance;      /* its 16-norms */
  const int32_t* raw_quant;         /* adjusted raw quant per block */
  const int32_t* quantized;         /* 3 planes of groups * 65536 */
  const int32_t* dc_quantized;      /* 3 block planes */
  const int8_t* y_to_x;             /* final CfL per 64x64 tile */
  const int8_t* y_to_b;
  const uint32_t* token_values;
  const uint16_t* token_contexts;
  const uint32_t* token_group_offsets;
  uint64_t token_count;
  int32_t x_qm_scale, b_qm_scale;
  float quant_dc;
  int32_t evaluations;
  float scores[8];                  /* Butteraugli score per evaluation */
  /* host: the coefficient orders and block-context map of the tokens */
  const GJXLAmdEntropyInputs* entropy_inputs;
  /* the learned run's allowed stages per block (null in a plain run) and
   * the 3-norm of its last evaluation's block distances (0 in a plain run) */
  const uint8_t* proposal_allow;
  float proposal_norm;
} GJXLAmdPlainFrame;

/* Stage ids of gjxl_amd_plain_path_stage_times. */
enum {
  GJXL_AMD_STAGE_BEGIN = 0,
  GJXL_AMD_STAGE_UPLOAD,
  GJXL_AMD_STAGE_SRGB_TO_LINEAR_AND_OPSIN,
  GJXL_AMD_STAGE_GABORISH_INVERSE,
  GJXL_AMD_STAGE_INITIAL_QUANT_FIELD,
  GJXL_AMD_STAGE_INITIAL_CFL,
  GJXL_AMD_STAGE_AC_STRATEGY_COSTS,
  GJXL_AMD_STAGE_AC_STRATEGY_SEARCH,
  GJXL_AMD_STAGE_AQ_PREPARE,
  GJXL_AMD_STAGE_BUTTERAUGLI_PREPARE,
  GJXL_AMD_STAGE_AQ_EVALUATE,
  GJXL_AMD_STAGE_BUTTERAUGLI_COMPARE,
  GJXL_AMD_STAGE_AQ_REDUCE_UPDATE,
  GJXL_AMD_STAGE_AC_TOKENIZE,
  GJXL_AMD_STAGE_ENTROPY_INPUTS,
  GJXL_AMD_STAGE_TOKEN_READBACK,
  GJXL_AMD_STAGE_PROPOSAL, /* the learned run's features, forward and decisions */
  GJXL_AMD_STAGE_COUNT
};

GJXL_AMD_API GJXLResult gjxl_amd_plain_path_options_init(
  GJXLAmdPlainPathOptions* options, size_t struct_size);
/* options may be null (defaults).  Allocates every device buffer of the
 * extent once; runs reuse them. */
GJXL_AMD_API GJXLResult gjxl_amd_plain_path_create(
  GJXLAmdDevice* device, uint32_t width, uint32_t height,
  const GJXLAmdPlainPathOptions* options, GJXLAmdPlainPath** out);
GJXL_AMD_API void gjxl_amd_plain_path_destroy(GJXLAmdPlainPath* path);
/* Re-targets the path to another extent (and, when options is not null,
 * other options) inside the buffers allocated at creation: every extent
 * whose width and height do not exceed the creation extent's fits.  No
 * device allocation; the views of the last run's frame go stale.  A refused
 * extent (GJXL_ERROR_INVALID_ARGUMENT) leaves the path as it was.  The
 * reference builds its per-image state in each encode call
 * (src/c_api/gjxl.cpp:470-637); this is the reuse its batch encoder
 * (src/codestream/batch_workflow.h:47-109) leaves to the allocator. */
GJXL_AMD_API GJXLResult gjxl_amd_plain_path_reconfigure(
  GJXLAmdPlainPath* path, uint32_t width, uint32_t height,
  const GJXLAmdPlainPathOptions* options);
/* Runs the path on one image of the path's extent and returns when the
 * tokens are on the host (tokens may be null: they stay on the device).
 * frame may be null. */
GJXL_AMD_API GJXLResult gjxl_amd_plain_path_run(GJXLAmdPlainPath* path,
                                                const GJXLAmdImage* image,
                                                const GJXLAmdTokenOutput* tokens,
                                                GJXLAmdPlainFrame* frame);
/* The run in two halves, for callers that batch the candidate costs of
 * several paths (gjxl_amd_cost_queue_run): run_front uploads and enqueues
 * every stage before the candidate costs (no host wait); cost_frame
 * describes the path's cost inputs and planes; run_back, after the costs
 * are on the path's device stream (costs_ready non-zero: written there by
 * the caller, ordered before this call by an event the stream waits for;
 * zero: computed here), runs the rest as gjxl_amd_plain_path_run does. */
GJXL_AMD_API GJXLResult gjxl_amd_plain_path_run_front(GJXLAmdPlainPath* path,
                                                      const GJXLAmdImage* image);
GJXL_AMD_API GJXLResult gjxl_amd_plain_path_cost_frame(GJXLAmdPlainPath* path,
                                                       GJXLAmdCostFrame* frame);
GJXL_AMD_API GJXLResult gjxl_amd_plain_path_run_back(GJXLAmdPlainPath* path,
                                                     int32_t costs_ready,
                                                     const GJXLAmdTokenOutput* tokens,
                                                     GJXLAmdPlainFrame* frame);
/* Device time of each stage of the last completed run, from HIP events on
 * the device stream: stages[i] (GJXL_AMD_STAGE_*) took milliseconds[i];
 * *count receives the number of entries (evaluation stages repeat). */
GJXL_AMD_API GJXLResult gjxl_amd_plain_path_stage_times(
  GJXLAmdPlainPath* path, int32_t* stages, float* milliseconds,
  int32_t capacity, int32_t* count);

/* Pinned host memory (hipHostMalloc) for inputs and token outputs. */
GJXL_AMD_API GJXLResult gjxl_amd_host_alloc(size_t bytes, void** out);
GJXL_AMD_API void gjxl_amd_host_free(void* pointer);

/* ---------------------------------------------------------------------- */
/* f4: batch driver.  Replaces VarDctBatchEncoder (src/codestream/          */
/* batch_workflow.h:47-109) for the plain path: a persistent,               */
/* bounded-concurrency driver for independent images of any extent.  Each   */
/* of max_in_flight workers is one host thread with its own device handle   */
/* (its own stream) on one GPU and keeps its plain path between images of   */
/* the same extent and options; images of different workers run on the     */
/* GPU at once.  One encode call runs every request and keeps request       */
/* order in the results; an image's failure is reported in its result and   */
/* does not stop the others.  Results are published when every image is    */
/* done; invalid driver or output arguments leave the results unchanged.    */
/* Concurrent encode calls on one driver are serialized.  shutdown closes   */
/* the driver: it drains the active call and joins the workers; later      */
/* calls return GJXL_ERROR_UNAVAILABLE.                                     */
/* ---------------------------------------------------------------------- */
typedef struct GJXLAmdBatch GJXLAmdBatch;

typedef struct {
  GJXLAmdImage image;               /* host pixels (device_resident == 0) */
  GJXLAmdPlainPathOptions options;  /* struct_size 0: the defaults */
} GJXLAmdBatchRequest;

/* One image's outcome (VarDctBatchEncodingResult, batch_workflow.h:40-45).
 * On success the token stream, its group offsets and the entropy inputs are
 * library-owned host memory until gjxl_amd_batch_results_free.  Timing is
 * wall time on the calling side (VarDctBatchSchedulingTiming, :22-38):
 * queue = encode call to the worker taking the image, service = worker
 * start to result, ready = queue + service. */
typedef struct {
  GJXLResult status;
  char message[256];                /* the failure's diagnostic */
  uint32_t width, height, groups;
  uint64_t token_count;
  uint32_t* token_values;
  uint16_t* token_contexts;
  uint32_t* token_group_offsets;    /* groups + 1 */
  GJXLAmdEntropyInputs* entropy_inputs;
  int32_t evaluations;
  float scores[8];
  uint64_t queue_nanoseconds, service_nanoseconds, ready_nanoseconds;
} GJXLAmdBatchResult;

GJXL_AMD_API GJXLResult gjxl_amd_batch_create(int device_ordinal,
                                              uint32_t max_in_flight,
                                              GJXLAmdBatch** out);
GJXL_AMD_API uint32_t gjxl_amd_batch_max_in_flight(const GJXLAmdBatch* batch);
/* results: count entries, zeroed (or freed) by the caller beforehand. */
GJXL_AMD_API GJXLResult gjxl_amd_batch_encode(GJXLAmdBatch* batch,
                                              const GJXLAmdBatchRequest* requests,
                                              size_t count,
                                              GJXLAmdBatchResult* results);
/* Frees the host memory of count results and zeroes them. */
GJXL_AMD_API void gjxl_amd_batch_results_free(GJXLAmdBatchResult* results,
                                              size_t count);
/* The chosen batching form for this driver (no counterpart in the
 * reference): with enable non-zero, later encode calls take their images
 * largest first; every worker runs its image's front half
 * (gjxl_amd_plain_path_run_front) and submits its cost inputs, and whenever
 * no worker is still in a front half, one launch of a tile-queue graph
 * (gjxl_amd_cost_queue_*, on a stream of its own) computes the candidate
 * costs of every submitted image; each worker then runs its back half.
 * Results, their order and their values are those of the default mode.
 * tune_repetitions non-zero tunes the queue's grids at the next launch.
 * Not while an encode call is in progress. */
GJXL_AMD_API GJXLResult gjxl_amd_batch_set_tile_queue(GJXLAmdBatch* batch, int32_t enable,
                                                      uint32_t tune_repetitions);
GJXL_AMD_API void gjxl_amd_batch_shutdown(GJXLAmdBatch* batch);
/* Shuts down, then frees the driver; no call may be in progress. */
GJXL_AMD_API void gjxl_amd_batch_destroy(GJXLAmdBatch* batch);


/* ---------------------------------------------------------------------- */
/* Chosen learned form (not the plain path; no counterpart in the           */
/* reference): a proposal network that predicts, per 8x8 block, logits over */
/* the format's 27 strategy codes and the log ratio of the Butteraugli      */
/* loop's final quant field to the adjusted initial field, trained on       */
/* labels of the plain search (src/codec/ac_strategy.cpp:1101               */
/* FindAcStrategyGrid; the loop of src/codec/adaptive_quantization.cpp:     */
/* 1600-1738).  Its confident tiles prune the candidate costs              */
/* (gjxl_amd_ac_strategy_costs_pruned) and its field warm-starts the loop.  */
/*                                                                          */
/* Network: six 3x3 convolutions with zero padding over the block grid,     */
/* 16 input features, 32 channels, ReLU after the first five; weights and   */
/* activations bfloat16, products accumulated in binary32 on MFMA, biases   */
/* binary32.  Output channels 0-26 are the logits, 27 the field ratio.      */
/* Tensors are channel-last: features [n][h][w][16] and activations         */
/* [n][h][w][32] of bfloat16 bits (uint16), the head [n][h][w][32] binary32. */
/* Parameters: layer by layer, weights [9][32][inputs] (tap = 3 ky + kx     */
/* reads row y + ky - 1, column x + kx - 1) then 32 biases, binary32.       */
/* ---------------------------------------------------------------------- */
typedef struct GJXLAmdProposal GJXLAmdProposal;

enum {
  GJXL_AMD_PROPOSAL_FEATURES = 16,
  GJXL_AMD_PROPOSAL_CHANNELS = 32,
  GJXL_AMD_PROPOSAL_LAYERS = 6,
  GJXL_AMD_PROPOSAL_CLASSES = 27,
  GJXL_AMD_PROPOSAL_FIELD_CHANNEL = 27,
  GJXL_AMD_PROPOSAL_IGNORE = 255 /* label of a block outside the loss */
};

/* Buffers of gjxl_amd_proposal_buffer. */
enum {
  GJXL_AMD_PROPOSAL_HEAD = 0,       /* [positions][32] binary32, last forward */
  GJXL_AMD_PROPOSAL_PARAMETERS = 1, /* binary32 master parameters */
  GJXL_AMD_PROPOSAL_GRADIENTS = 2,  /* binary32, of the last backward (summed loss) */
  GJXL_AMD_PROPOSAL_ACTIVATIONS = 3 /* five [positions][32] bfloat16 planes */
};

/* Number of binary32 parameters (50880). */
GJXL_AMD_API size_t gjxl_amd_proposal_parameter_count(void);
/* A network on `device` for up to max_positions block positions per
 * forward (n * h * w); parameters zero until set. */
GJXL_AMD_API GJXLResult gjxl_amd_proposal_create(GJXLAmdDevice* device,
                                                 size_t max_positions,
                                                 GJXLAmdProposal** out);
GJXL_AMD_API void gjxl_amd_proposal_destroy(GJXLAmdProposal* net);
/* Host parameters in, out (count = gjxl_amd_proposal_parameter_count());
 * setting them resets the optimizer state.  Synchronous. */
GJXL_AMD_API GJXLResult gjxl_amd_proposal_set_parameters(GJXLAmdProposal* net,
                                                         const float* host,
                                                         size_t count);
GJXL_AMD_API GJXLResult gjxl_amd_proposal_get_parameters(GJXLAmdProposal* net,
                                                         float* host, size_t count);
GJXL_AMD_API GJXLResult gjxl_amd_proposal_buffer(GJXLAmdProposal* net, int32_t id,
                                                 void** pointer, size_t* bytes);
/* Features of every 8x8 block of a padded XYB image (3 planes of
 * width x height binary32, multiples of 8) and its initial quant field
 * ((w/8)*(h/8)): features [h/8][w/8][16] bfloat16 -- the block mean of
 * each channel (X times 16), log1p of the orthonormal-DCT8 energy of each
 * channel in the bands u + v = 1-2, 3-5, 6-14 times 1e4 (X) or 1e2 (Y,
 * B), the log of the field, three zeros.  Enqueued. */
GJXL_AMD_API GJXLResult gjxl_amd_proposal_features(GJXLAmdDevice* device,
                                                   const float* opsin, uint32_t width,
                                                   uint32_t height,
                                                   const float* quant_field,
                                                   uint16_t* features);
/* Forward over n frames of h x w blocks (device features); the head and
 * the activations stay in the network for the backward, which reads the
 * features again: they must stay valid until it has run.  Enqueued. */
GJXL_AMD_API GJXLResult gjxl_amd_proposal_forward(GJXLAmdProposal* net,
                                                  const uint16_t* features,
                                                  uint32_t n, uint32_t h, uint32_t w);
/* Loss and gradients of the last forward: labels (strategy code or
 * GJXL_AMD_PROPOSAL_IGNORE) and field targets per position (device).  The
 * loss is the sum over positions of the softmax cross-entropy plus
 * qf_weight times the squared error of channel 27.  sums (host, may be
 * null) receives the cross-entropy sum, the squared-error sum and the
 * count of positions whose largest logit is the label (waits for the
 * device then). */
GJXL_AMD_API GJXLResult gjxl_amd_proposal_backward(GJXLAmdProposal* net,
                                                   const uint8_t* labels,
                                                   const float* targets,
                                                   float qf_weight, double* sums);
/* One Adam step on the gradients of the last backward times grad_scale
 * (step counts from 1 after set_parameters).  Enqueued. */
GJXL_AMD_API GJXLResult gjxl_amd_proposal_adam(GJXLAmdProposal* net, float learning_rate,
                                               float beta1, float beta2, float epsilon,
                                               float grad_scale);
/* Allowed stages per block from a head of bh x bw positions: per block a
 * softmax over the seven candidate strategies' logits; stage s is allowed
 * if it is DCT8, among the top_k probabilities (ties to the lower stage) or
 * at least p_min.  A block whose largest probability is below threshold
 * allows every stage (the plain search).  Enqueued. */
GJXL_AMD_API GJXLResult gjxl_amd_proposal_decide(GJXLAmdDevice* device, const float* head,
                                                 uint32_t block_width, uint32_t block_height,
                                                 float threshold, uint32_t top_k,
                                                 float p_min, uint8_t* allow);
/* The loop's starting field: adjusted * exp(clamp(head channel 27 +
 * offset, +-limit)) clamped to [lower, upper].  Enqueued. */
GJXL_AMD_API GJXLResult gjxl_amd_proposal_warm_start(GJXLAmdDevice* device,
                                                     const float* head,
                                                     const float* adjusted, size_t count,
                                                     float limit, float offset, float lower,
                                                     float upper, float* field);

/* The learned run of a plain path.  With a network attached, a run adds
 * the features and a forward after the initial quant field, prunes the
 * candidate costs by the decisions, starts the Butteraugli loop from the
 * warm-started field and runs `evaluations` evaluations (1-3).  The judge
 * is the 3-norm of the last evaluation's block distances, (mean of d^3)^
 * (1/3): while it exceeds judge_ratio * distance and fewer than three
 * evaluations ran, the run goes on with the plain loop's updates (plain
 * runs of the synthetic workloads end at 0.60-0.64 times the distance
 * between distances 0.9 and 1.2).  The frame reports the evaluations run,
 * the allowed stages and the judged norm.  With final_compare zero the run
 * is unjudged: evaluations - 1 evaluations with their updates, then a last
 * pass that quantizes the updated field (the coefficients, DC and raw
 * quantization the tokens read) without reconstruction, comparison or
 * judge; the tokens equal those of a run that compares that pass too.  The
 * frame then reports the comparisons run as its evaluations (none for
 * evaluations 1, which also skips the Butteraugli reference side), a NaN
 * judged norm, and the block distances and reconstruction of its last
 * comparison (with none, those of an earlier run).  set_proposal also takes a struct that ends before
 * final_compare (struct_size), with final_compare at its default. */
typedef struct {
  size_t struct_size;
  float threshold;       /* block confidence, default 0.3 */
  uint32_t top_k;        /* default 2 */
  float p_min;           /* default 0.1 */
  int32_t evaluations;   /* default 2 */
  float field_limit;     /* default 2 */
  float field_offset;    /* log offset of the predicted field, default 0 */
  float judge_ratio;     /* default 0.7 */
  int32_t final_compare; /* 1 (default) or 0: the unjudged last pass above */
} GJXLAmdProposalOptions;

GJXL_AMD_API GJXLResult gjxl_amd_proposal_options_init(GJXLAmdProposalOptions* options,
                                                       size_t struct_size);
/* Attaches (net non-null; it must outlive the attachment and hold
 * max_positions >= the path's block count) or detaches (null) the
 * network.  Not while a run is in progress.  A run uses the network on
 * the path's own streams, whatever handle created it: work enqueued on the
 * network's handle (gjxl_amd_proposal_adam and the like) must be complete
 * before the run (gjxl_amd_proposal_set_parameters waits for its own). */
GJXL_AMD_API GJXLResult gjxl_amd_plain_path_set_proposal(GJXLAmdPlainPath* path,
                                                         GJXLAmdProposal* net,
                                                         const GJXLAmdProposalOptions* options);
/* Training data from the last run of a plain path (device outputs of
 * (w/8)*(h/8) entries): the features of its Gaborish-inverted image and
 * initial field, labels = the strategy code of each block's transform,
 * targets = log(final field / adjusted initial field).  Enqueued on the
 * path's device stream. */
GJXL_AMD_API GJXLResult gjxl_amd_plain_path_proposal_sample(GJXLAmdPlainPath* path,
                                                            uint16_t* features,
                                                            uint8_t* labels,
                                                            float* targets);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif /* GJXL_AMD_H_ */
