// Batch driver: a persistent pool of workers, each a host thread with its own
// device handle (stream) on one GPU, running the plain path on independent
// images.  The semantics follow VarDctBatchEncoder (src/codestream/
// batch_workflow.h:47-109): request order is kept, a failed image is reported
// in its own result, results are published only when the whole call is done,
// encode calls on one driver are serialized, and shutdown drains the active
// call before joining the workers.  Images of different workers overlap on
// the GPU (one image's DC wavefront uses eight work groups; the others fill
// the remaining CUs).  A worker keeps one plain path, and so its device
// buffers, across images: it creates the path at the bounding extent of the
// call's requests and re-targets it to each image's extent and options
// inside those buffers.
//
// With the tile queue on (gjxl_amd_batch_set_tile_queue, the chosen batching
// form), the workers take the call's images largest first and submit each
// image's cost inputs after its front half; whenever no worker is still in a
// front half, the submitted frames' candidate costs go through one launch of
// the queue's captured graph on the queue's own stream, and each worker runs
// its back half once its costs are on its stream (an event).  A worker waits
// only for the others' front halves, never for their back halves.  No
// kernels here.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "common.h"

using gjxl_amd::Fail;

namespace {

using Clock = std::chrono::steady_clock;

uint64_t Nanoseconds(Clock::duration d) {
  return static_cast<uint64_t>(std::chrono::duration_cast<std::chrono::nanoseconds>(d).count());
}

struct Call {
  const GJXLAmdBatchRequest* requests = nullptr;
  size_t count = 0;
  std::vector<GJXLAmdBatchResult>* staged = nullptr;
  Clock::time_point arrival;
  uint32_t bound_width = 0, bound_height = 0;  // over the requests
  std::atomic<size_t> next{0};
  bool grouped = false;
  std::vector<size_t> order;  // grouped: request indices, largest image first
  size_t done = 0;  // guarded by the driver's mutex
};

struct Worker {
  size_t index = 0;
  GJXLAmdDevice* device = nullptr;
  GJXLAmdPlainPath* path = nullptr;
  uint32_t capacity_width = 0, capacity_height = 0;  // the path's creation extent
  uint32_t width = 0, height = 0;                    // its current extent
  GJXLAmdPlainPathOptions options{};
  std::thread thread;
};

void FreeResult(GJXLAmdBatchResult* r) {
  std::free(r->token_values);
  std::free(r->token_contexts);
  std::free(r->token_group_offsets);
  delete r->entropy_inputs;
  std::memset(r, 0, sizeof(*r));
}

void SetFailure(GJXLAmdBatchResult* r, GJXLResult code) {
  r->status = code;
  const char* message = gjxl_amd_last_error();
  std::strncpy(r->message, message != nullptr ? message : "", sizeof(r->message) - 1);
  r->message[sizeof(r->message) - 1] = '\0';
}

bool SameOptions(const GJXLAmdPlainPathOptions& a, const GJXLAmdPlainPathOptions& b) {
  return a.distance == b.distance && a.effort == b.effort &&
         a.intensity_target == b.intensity_target;
}

void DropPath(Worker* w) {
  gjxl_amd_plain_path_destroy(w->path);
  w->path = nullptr;
  w->capacity_width = w->capacity_height = w->width = w->height = 0;
}

// Points the worker's path at the image's extent and options: a reconfigure
// inside the current buffers when they fit, else a new path at the call's
// bounding extent (or at the image's own, when the bound is refused).
GJXLResult PreparePath(Worker* w, const Call& call, uint32_t width, uint32_t height,
                       const GJXLAmdPlainPathOptions& options) {
  if (w->path != nullptr && width <= w->capacity_width && height <= w->capacity_height) {
    if (width == w->width && height == w->height && SameOptions(w->options, options)) {
      return GJXL_OK;
    }
    const GJXLResult r = gjxl_amd_plain_path_reconfigure(w->path, width, height, &options);
    if (r == GJXL_OK) {
      w->width = width;
      w->height = height;
      w->options = options;
    } else if (r != GJXL_ERROR_INVALID_ARGUMENT && r != GJXL_ERROR_UNSUPPORTED) {
      DropPath(w);  // refused arguments leave the path as it was; anything else may not
    }
    return r;
  }
  DropPath(w);
  uint32_t cw = std::max(width, call.bound_width);
  uint32_t ch = std::max(height, call.bound_height);
  GJXLResult r = gjxl_amd_plain_path_create(w->device, cw, ch, &options, &w->path);
  if (r == GJXL_OK && (cw != width || ch != height)) {
    r = gjxl_amd_plain_path_reconfigure(w->path, width, height, &options);
    if (r != GJXL_OK) DropPath(w);
  }
  if (r != GJXL_OK && (cw != width || ch != height)) {
    w->path = nullptr;
    cw = width;
    ch = height;
    r = gjxl_amd_plain_path_create(w->device, cw, ch, &options, &w->path);
  }
  if (r != GJXL_OK) {
    w->path = nullptr;
    return r;
  }
  w->capacity_width = cw;
  w->capacity_height = ch;
  w->width = width;
  w->height = height;
  w->options = options;
  return GJXL_OK;
}

// The options and path of one request; a refusal lands in *r (which starts
// zeroed).
GJXLResult PrepareImage(Worker* w, const Call& call, const GJXLAmdBatchRequest& request,
                        GJXLAmdBatchResult* r) {
  GJXLAmdPlainPathOptions options{};
  if (request.options.struct_size == 0) {
    (void)gjxl_amd_plain_path_options_init(&options, sizeof(options));
  } else if (request.options.struct_size < sizeof(GJXLAmdPlainPathOptions)) {
    const GJXLResult e = Fail(GJXL_ERROR_INVALID_ARGUMENT, "batch request options are too small");
    SetFailure(r, e);
    return e;
  } else {
    options = request.options;
  }
  const GJXLAmdImage& image = request.image;
  if (image.device_resident) {
    const GJXLResult e = Fail(GJXL_ERROR_INVALID_ARGUMENT, "batch requests take host pixels");
    SetFailure(r, e);
    return e;
  }
  r->width = image.width;
  r->height = image.height;
  const GJXLResult prepared = PreparePath(w, call, image.width, image.height, options);
  if (prepared != GJXL_OK) SetFailure(r, prepared);
  return prepared;
}

// The outputs of a finished run (status) into *r.
void CollectImage(Worker* w, GJXLResult status, const GJXLAmdPlainFrame& frame,
                  GJXLAmdBatchResult* r) {
  if (status != GJXL_OK) {
    // a failed run may leave the path mid-sequence: start the next image afresh
    DropPath(w);
    return SetFailure(r, status);
  }
  const size_t count = static_cast<size_t>(frame.token_count);
  r->groups = frame.groups;
  r->token_count = frame.token_count;
  r->token_values = static_cast<uint32_t*>(std::malloc(std::max<size_t>(4 * count, 4)));
  r->token_contexts = static_cast<uint16_t*>(std::malloc(std::max<size_t>(2 * count, 2)));
  r->token_group_offsets = static_cast<uint32_t*>(std::malloc(4 * (frame.groups + 1u)));
  r->entropy_inputs = new (std::nothrow) GJXLAmdEntropyInputs;
  if (r->token_values == nullptr || r->token_contexts == nullptr ||
      r->token_group_offsets == nullptr || r->entropy_inputs == nullptr) {
    FreeResult(r);
    return SetFailure(r, Fail(GJXL_ERROR_OUT_OF_MEMORY, "batch result allocation failed"));
  }
  status = gjxl_amd_download(w->device, r->token_values, frame.token_values, 4 * count);
  if (status == GJXL_OK) {
    status = gjxl_amd_download(w->device, r->token_contexts, frame.token_contexts, 2 * count);
  }
  if (status == GJXL_OK) {
    status = gjxl_amd_download(w->device, r->token_group_offsets, frame.token_group_offsets,
                               4 * (frame.groups + 1u));
  }
  if (status != GJXL_OK) {
    FreeResult(r);
    return SetFailure(r, status);
  }
  std::memcpy(r->entropy_inputs, frame.entropy_inputs, sizeof(GJXLAmdEntropyInputs));
  r->evaluations = frame.evaluations;
  for (int i = 0; i < 8; ++i) r->scores[i] = frame.scores[i];
  r->status = GJXL_OK;
}

// One image on one worker into *r (which starts zeroed).
void RunImage(Worker* w, const Call& call, const GJXLAmdBatchRequest& request,
              GJXLAmdBatchResult* r) {
  if (PrepareImage(w, call, request, r) != GJXL_OK) return;
  GJXLAmdPlainFrame frame{};
  const GJXLResult status = gjxl_amd_plain_path_run(w->path, &request.image, nullptr, &frame);
  CollectImage(w, status, frame, r);
}

}  // namespace

struct GJXLAmdBatch {
  int ordinal = 0;
  std::vector<std::unique_ptr<Worker>> workers;
  std::mutex call_mutex;  // serializes encode calls and shutdown's drain
  std::mutex mutex;       // guards what follows
  std::condition_variable work, finished;
  Call* call = nullptr;
  uint64_t generation = 0;
  bool closed = false, stopping = false, joined = false;
  // the tile queue on its own device handle (stream) and the hand-over:
  // per worker its frame, the events of its front half and of its costs,
  // and whether its submitted frame has been launched (and ran)
  GJXLAmdDevice* queue_device = nullptr;
  GJXLAmdCostQueue* queue = nullptr;
  uint32_t tune_repetitions = 0;  // tune at the next launch when non-zero
  std::vector<GJXLAmdCostFrame> frames;
  std::vector<hipEvent_t> front_done, costs_done;
  std::vector<char> launched, queue_ok;
  std::mutex group_mutex;  // guards the hand-over state and the queue
  std::condition_variable group_cv;
  std::vector<size_t> pending;  // submitted workers, not yet launched
  size_t in_front = 0;          // workers between taking an image and submitting
  uint64_t launches = 0;
};

namespace {

// Under group_mutex: the pending frames' candidate costs through one launch
// of the queue's graph, after their front halves (events); each launched
// worker's costs event is recorded after it.
void LaunchPending(GJXLAmdBatch* b) {
  hipStream_t st = b->queue_device->stream;
  std::vector<GJXLAmdCostFrame> frames;
  bool ok = true;
  for (size_t k : b->pending) {
    ok = ok && hipStreamWaitEvent(st, b->front_done[k], 0) == hipSuccess;
    frames.push_back(b->frames[k]);
  }
  const uint32_t n = static_cast<uint32_t>(frames.size());
  if (ok && b->tune_repetitions != 0) {
    ok = gjxl_amd_cost_queue_tune(b->queue, frames.data(), n, b->tune_repetitions) == GJXL_OK;
    b->tune_repetitions = 0;
  }
  ok = ok && gjxl_amd_cost_queue_run(b->queue, frames.data(), n) == GJXL_OK;
  for (size_t k : b->pending) {
    b->queue_ok[k] = ok && hipEventRecord(b->costs_done[k], st) == hipSuccess;
    b->launched[k] = 1;
  }
  b->pending.clear();
  ++b->launches;
  b->group_cv.notify_all();
}

// A call in tile-queue mode: images taken in call->order (largest first).
void RunQueued(GJXLAmdBatch* b, Worker* w, Call* call) {
  const size_t j = w->index;
  for (;;) {
    size_t slot;
    {
      std::lock_guard<std::mutex> lock(b->group_mutex);
      slot = call->next.fetch_add(1);
      if (slot >= call->count) break;
      ++b->in_front;
    }
    const size_t i = call->order[slot];
    GJXLAmdBatchResult* r = &(*call->staged)[i];
    const Clock::time_point start = Clock::now();
    bool front = false;
    if (PrepareImage(w, *call, call->requests[i], r) == GJXL_OK) {
      const GJXLResult status = gjxl_amd_plain_path_run_front(w->path, &call->requests[i].image);
      if (status == GJXL_OK) {
        front = gjxl_amd_plain_path_cost_frame(w->path, &b->frames[j]) == GJXL_OK &&
                hipEventRecord(b->front_done[j], w->device->stream) == hipSuccess;
      } else {
        DropPath(w);
        SetFailure(r, status);
      }
    }
    bool ready = false;
    {
      std::unique_lock<std::mutex> lock(b->group_mutex);
      --b->in_front;
      if (front) {
        b->launched[j] = 0;
        b->pending.push_back(j);
      }
      // the last worker out of its front half launches what is pending
      if (!b->pending.empty() && b->in_front == 0) LaunchPending(b);
      if (front) {
        b->group_cv.wait(lock, [&] { return b->launched[j] != 0; });
        ready = b->queue_ok[j] != 0;
      }
    }
    if (front) {
      ready = ready && hipStreamWaitEvent(w->device->stream, b->costs_done[j], 0) == hipSuccess;
      GJXLAmdPlainFrame frame{};
      const GJXLResult status = gjxl_amd_plain_path_run_back(w->path, ready ? 1 : 0, nullptr,
                                                             &frame);
      CollectImage(w, status, frame, r);
    }
    const Clock::time_point end = Clock::now();
    r->queue_nanoseconds = Nanoseconds(start - call->arrival);
    r->service_nanoseconds = Nanoseconds(end - start);
    r->ready_nanoseconds = r->queue_nanoseconds + r->service_nanoseconds;
    std::lock_guard<std::mutex> lock(b->mutex);
    if (++call->done == call->count) b->finished.notify_all();
  }
}

void WorkerLoop(GJXLAmdBatch* b, Worker* w) {
  uint64_t seen = 0;
  for (;;) {
    Call* call = nullptr;
    {
      std::unique_lock<std::mutex> lock(b->mutex);
      b->work.wait(lock, [&] { return b->stopping || (b->call != nullptr && b->generation != seen); });
      if (b->call == nullptr || b->generation == seen) return;  // stopping, nothing new
      call = b->call;
      seen = b->generation;
    }
    if (call->grouped) {
      RunQueued(b, w, call);
      continue;
    }
    for (;;) {
      const size_t i = call->next.fetch_add(1);
      if (i >= call->count) break;
      GJXLAmdBatchResult& r = (*call->staged)[i];
      const Clock::time_point start = Clock::now();
      RunImage(w, *call, call->requests[i], &r);
      const Clock::time_point end = Clock::now();
      r.queue_nanoseconds = Nanoseconds(start - call->arrival);
      r.service_nanoseconds = Nanoseconds(end - start);
      r.ready_nanoseconds = r.queue_nanoseconds + r.service_nanoseconds;
      std::lock_guard<std::mutex> lock(b->mutex);
      if (++call->done == call->count) b->finished.notify_all();
    }
  }
}

void DropQueue(GJXLAmdBatch* b) {
  gjxl_amd_cost_queue_destroy(b->queue);
  b->queue = nullptr;
  for (auto* events : {&b->front_done, &b->costs_done}) {
    for (hipEvent_t& e : *events) {
      if (e != nullptr) (void)hipEventDestroy(e);
    }
    events->clear();
  }
  gjxl_amd_device_destroy(b->queue_device);
  b->queue_device = nullptr;
}

void JoinAndRelease(GJXLAmdBatch* b) {
  for (auto& w : b->workers) {
    if (w->thread.joinable()) w->thread.join();
  }
  DropQueue(b);
  for (auto& w : b->workers) {
    DropPath(w.get());
    gjxl_amd_device_destroy(w->device);
    w->device = nullptr;
  }
}

}  // namespace

extern "C" {

GJXLResult gjxl_amd_batch_create(int device_ordinal, uint32_t max_in_flight,
                                 GJXLAmdBatch** out) {
  if (out == nullptr || *out != nullptr || max_in_flight == 0 || max_in_flight > 16) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT,
                "batch output must be empty and max_in_flight in [1, 16]");
  }
  auto* b = new (std::nothrow) GJXLAmdBatch;
  if (b == nullptr) return Fail(GJXL_ERROR_OUT_OF_MEMORY, "batch allocation failed");
  b->ordinal = device_ordinal;
  GJXLResult r = GJXL_OK;
  try {
    for (uint32_t i = 0; r == GJXL_OK && i < max_in_flight; ++i) {
      auto w = std::make_unique<Worker>();
      w->index = i;
      r = gjxl_amd_device_create(device_ordinal, &w->device);
      if (r == GJXL_OK) b->workers.push_back(std::move(w));
    }
    if (r == GJXL_OK) {
      for (auto& w : b->workers) w->thread = std::thread(WorkerLoop, b, w.get());
    }
  } catch (...) {
    r = Fail(GJXL_ERROR_OUT_OF_MEMORY, "batch worker start failed");
  }
  if (r != GJXL_OK) {
    {
      std::lock_guard<std::mutex> lock(b->mutex);
      b->closed = b->stopping = true;
    }
    b->work.notify_all();
    JoinAndRelease(b);
    delete b;
    return r;
  }
  *out = b;
  return GJXL_OK;
}

uint32_t gjxl_amd_batch_max_in_flight(const GJXLAmdBatch* batch) {
  return batch == nullptr ? 0u : static_cast<uint32_t>(batch->workers.size());
}

GJXLResult gjxl_amd_batch_encode(GJXLAmdBatch* b, const GJXLAmdBatchRequest* requests,
                                 size_t count, GJXLAmdBatchResult* results) {
  if (b == nullptr || (count != 0 && (requests == nullptr || results == nullptr))) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "batch arguments are null");
  }
  const Clock::time_point arrival = Clock::now();
  std::lock_guard<std::mutex> serial(b->call_mutex);
  {
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->closed) return Fail(GJXL_ERROR_UNAVAILABLE, "batch driver is shut down");
  }
  if (count == 0) return GJXL_OK;
  std::vector<GJXLAmdBatchResult> staged;
  try {
    staged.resize(count);
  } catch (...) {
    return Fail(GJXL_ERROR_OUT_OF_MEMORY, "batch result staging failed");
  }
  std::memset(staged.data(), 0, count * sizeof(GJXLAmdBatchResult));
  Call call;
  call.requests = requests;
  call.count = count;
  call.staged = &staged;
  call.arrival = arrival;
  if (b->queue != nullptr) {
    // largest first (results keep request order)
    try {
      call.order.resize(count);
    } catch (...) {
      return Fail(GJXL_ERROR_OUT_OF_MEMORY, "batch order staging failed");
    }
    for (size_t i = 0; i < count; ++i) call.order[i] = i;
    auto pixels = [&](size_t i) {
      return static_cast<uint64_t>(requests[i].image.width) * requests[i].image.height;
    };
    std::stable_sort(call.order.begin(), call.order.end(),
                     [&](size_t a, size_t c) { return pixels(a) > pixels(c); });
    call.grouped = true;
  }
  for (size_t i = 0; i < count; ++i) {
    const GJXLAmdImage& image = requests[i].image;
    if (image.width == 0 || image.height == 0 || image.width > (1u << 16) ||
        image.height > (1u << 16)) {
      continue;  // refused in its own result
    }
    call.bound_width = std::max(call.bound_width, image.width);
    call.bound_height = std::max(call.bound_height, image.height);
  }
  {
    std::lock_guard<std::mutex> lock(b->mutex);
    b->call = &call;
    ++b->generation;
  }
  b->work.notify_all();
  {
    std::unique_lock<std::mutex> lock(b->mutex);
    b->finished.wait(lock, [&] { return call.done == call.count; });
    b->call = nullptr;
  }
  std::memcpy(results, staged.data(), count * sizeof(GJXLAmdBatchResult));
  return GJXL_OK;
}

GJXLResult gjxl_amd_batch_set_tile_queue(GJXLAmdBatch* b, int32_t enable,
                                         uint32_t tune_repetitions) {
  if (b == nullptr) return Fail(GJXL_ERROR_INVALID_ARGUMENT, "batch is null");
  if (tune_repetitions > 100) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "tune repetitions must be at most 100");
  }
  std::lock_guard<std::mutex> serial(b->call_mutex);
  {
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->closed) return Fail(GJXL_ERROR_UNAVAILABLE, "batch driver is shut down");
  }
  DropQueue(b);
  if (!enable) return GJXL_OK;
  const size_t n = b->workers.size();
  GJXLResult r = gjxl_amd_device_create(b->ordinal, &b->queue_device);
  if (r == GJXL_OK) {
    r = gjxl_amd_cost_queue_create(b->queue_device, static_cast<uint32_t>(n), &b->queue);
  }
  try {
    b->frames.assign(n, GJXLAmdCostFrame{});
    b->front_done.assign(n, nullptr);
    b->costs_done.assign(n, nullptr);
    b->launched.assign(n, 0);
    b->queue_ok.assign(n, 0);
    b->pending.clear();
    b->pending.reserve(n);
  } catch (...) {
    if (r == GJXL_OK) r = Fail(GJXL_ERROR_OUT_OF_MEMORY, "tile queue staging failed");
  }
  for (size_t i = 0; r == GJXL_OK && i < n; ++i) {
    r = gjxl_amd::Check(hipEventCreateWithFlags(&b->front_done[i], hipEventDisableTiming),
                        "hipEventCreate");
    if (r == GJXL_OK) {
      r = gjxl_amd::Check(hipEventCreateWithFlags(&b->costs_done[i], hipEventDisableTiming),
                          "hipEventCreate");
    }
  }
  if (r != GJXL_OK) {
    DropQueue(b);
    return r;
  }
  b->tune_repetitions = tune_repetitions;
  return GJXL_OK;
}

void gjxl_amd_batch_results_free(GJXLAmdBatchResult* results, size_t count) {
  if (results == nullptr) return;
  for (size_t i = 0; i < count; ++i) FreeResult(&results[i]);
}

void gjxl_amd_batch_shutdown(GJXLAmdBatch* b) {
  if (b == nullptr) return;
  {
    std::lock_guard<std::mutex> lock(b->mutex);
    b->closed = true;
  }
  // the active call (if any) finishes first; calls queued behind it see
  // `closed` and return
  std::lock_guard<std::mutex> serial(b->call_mutex);
  {
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->joined) return;
    b->stopping = true;
  }
  b->work.notify_all();
  JoinAndRelease(b);
  std::lock_guard<std::mutex> lock(b->mutex);
  b->joined = true;
}

void gjxl_amd_batch_destroy(GJXLAmdBatch* b) {
  if (b == nullptr) return;
  gjxl_amd_batch_shutdown(b);
  delete b;
}

}  // extern "C"
