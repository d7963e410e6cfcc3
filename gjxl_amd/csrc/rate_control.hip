// f3: the target-size search (rate control) of the reference, host code.
//
// src/codestream/rate_control.cpp:246-320 SearchTargetSizeImpl: the encoder
// at the lower Butteraugli target, then at the upper one unless the first
// met the tolerance, then at the midpoint of the widest remaining interval
// (ties to the lower interval; binary32 arithmetic) until the size is met,
// the attempts run out or no interval splits.  The caller's evaluator runs
// one complete encode per target (here the device path and whatever sizes
// its output); candidates compare as BetterCandidate (:56-103).  Evaluator
// failures are candidate-local except GJXL_ERROR_RESOURCE_PLAN_EXCEEDED
// (:147-150); an invalid successful candidate is fatal (:157-168); a failed
// search leaves the result unchanged.  No kernels here.
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "common.h"

using gjxl_amd::Fail;

namespace {

constexpr float kMinimumTarget = 0.01f;   // kMinimumTargetSizeButteraugliTarget
constexpr float kMaximumTarget = 10.0f;   // kMaximumTargetSizeButteraugliTarget
constexpr uint32_t kMaximumAttempts = 64; // kMaximumTargetSizeEncodeAttempts

struct Interval {
  float lower, upper;
};

struct Best {
  bool valid = false;
  float target = 0.0f;
  GJXLAmdRateCandidate candidate{};
};

uint64_t AbsoluteError(uint64_t bytes, uint64_t target) {
  return bytes < target ? target - bytes : bytes - target;
}

bool Within(uint64_t bytes, const GJXLAmdRateControlOptions& o) {
  if (o.selection == GJXL_AMD_RATE_CLOSEST_ABSOLUTE) {
    return AbsoluteError(bytes, o.target_bytes) <= o.tolerance_bytes;
  }
  return bytes <= o.target_bytes && o.target_bytes - bytes <= o.tolerance_bytes;
}

// FinalScore: a resident GPU run's score, or an unevaluated one, breaks no tie
double FinalScore(const GJXLAmdRateCandidate& c) {
  return c.resident != 0 || c.score_evaluated == 0 ? std::numeric_limits<double>::infinity()
                                                   : c.final_score;
}

bool Better(const GJXLAmdRateControlOptions& o, float target, const GJXLAmdRateCandidate& c,
            const Best& best) {
  if (!best.valid) return true;
  const uint64_t t = o.target_bytes;
  const bool c_under = c.encoded_bytes <= t;
  const bool b_under = best.candidate.encoded_bytes <= t;
  if (o.selection == GJXL_AMD_RATE_CLOSEST_ABSOLUTE) {
    const uint64_t ce = AbsoluteError(c.encoded_bytes, t);
    const uint64_t be = AbsoluteError(best.candidate.encoded_bytes, t);
    if (ce != be) return ce < be;
    if (c_under != b_under) return c_under;
  } else {
    if (c_under != b_under) return c_under;
    if (c.encoded_bytes != best.candidate.encoded_bytes) {
      return c_under ? c.encoded_bytes > best.candidate.encoded_bytes
                     : c.encoded_bytes < best.candidate.encoded_bytes;
    }
  }
  const double cs = FinalScore(c), bs = FinalScore(best.candidate);
  if (cs != bs) return cs < bs;
  return target < best.target;
}

// the widest interval, ties to the lower one (WidestInterval)
size_t Widest(const std::vector<Interval>& v) {
  size_t best = 0;
  for (size_t i = 1; i < v.size(); ++i) {
    const float width = v[i].upper - v[i].lower;
    const float best_width = v[best].upper - v[best].lower;
    if (width > best_width || (width == best_width && v[i].lower < v[best].lower)) best = i;
  }
  return best;
}

struct Search {
  const GJXLAmdRateControlOptions& o;
  GJXLAmdRateEvaluator evaluator;
  void* user;
  Best best;
  uint32_t attempts = 0, failed = 0;
  GJXLResult first_failure = GJXL_OK;
  std::string first_message;

  // GJXL_OK unless the search must stop with that code
  GJXLResult Evaluate(float target) {
    ++attempts;
    GJXLAmdRateCandidate c{};
    const GJXLResult status = evaluator(user, target, &c);
    if (status != GJXL_OK) {
      if (status == GJXL_ERROR_RESOURCE_PLAN_EXCEEDED) return status;
      ++failed;
      if (first_failure == GJXL_OK) {
        first_failure = status;
        const char* m = gjxl_amd_last_error();
        first_message = m != nullptr ? m : "";
      }
      return GJXL_OK;
    }
    if (c.encoded_bytes == 0 || (c.score_evaluated != 0 && !std::isfinite(c.final_score))) {
      return Fail(GJXL_ERROR_INTERNAL, "Target-size evaluator returned an invalid candidate");
    }
    if (Better(o, target, c, best)) {
      best.valid = true;
      best.target = target;
      best.candidate = c;
    }
    return GJXL_OK;
  }

  bool Pending() const { return !best.valid || !Within(best.candidate.encoded_bytes, o); }
};

}  // namespace

extern "C" {

GJXLResult gjxl_amd_rate_control_options_init(GJXLAmdRateControlOptions* o, size_t struct_size) {
  if (o == nullptr || struct_size < sizeof(GJXLAmdRateControlOptions)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "rate control options are missing or too small");
  }
  std::memset(o, 0, sizeof(*o));
  o->struct_size = sizeof(*o);
  o->selection = GJXL_AMD_RATE_LARGEST_AT_OR_BELOW;
  o->maximum_attempts = 12;
  o->minimum_target = kMinimumTarget;
  o->maximum_target = kMaximumTarget;
  return GJXL_OK;
}

GJXLResult gjxl_amd_rate_control_search(const GJXLAmdRateControlOptions* options,
                                        GJXLAmdRateEvaluator evaluator, void* user,
                                        GJXLAmdRateControlResult* result) {
  if (options == nullptr || options->struct_size < sizeof(GJXLAmdRateControlOptions)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "rate control options are missing or too small");
  }
  const GJXLAmdRateControlOptions& o = *options;
  if ((o.selection != GJXL_AMD_RATE_LARGEST_AT_OR_BELOW &&
       o.selection != GJXL_AMD_RATE_CLOSEST_ABSOLUTE)) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Target-size selection policy is invalid");
  }
  if (result == nullptr || evaluator == nullptr || o.target_bytes == 0 ||
      o.maximum_attempts == 0 || o.maximum_attempts > kMaximumAttempts ||
      !std::isfinite(o.minimum_target) || !std::isfinite(o.maximum_target) ||
      o.minimum_target <= 0.0f || o.maximum_target <= o.minimum_target) {
    return Fail(GJXL_ERROR_INVALID_ARGUMENT, "Target-size search options are invalid");
  }
  Search s{o, evaluator, user};
  const float lower = o.minimum_target, upper = o.maximum_target;
  GJXL_AMD_TRY(s.Evaluate(lower));
  if (s.attempts < o.maximum_attempts && s.Pending()) GJXL_AMD_TRY(s.Evaluate(upper));
  std::vector<Interval> intervals;
  intervals.reserve(2 * kMaximumAttempts);
  intervals.push_back({lower, upper});
  while (s.attempts < o.maximum_attempts && s.Pending() && !intervals.empty()) {
    const size_t index = Widest(intervals);
    const Interval interval = intervals[index];
    intervals.erase(intervals.begin() + static_cast<std::ptrdiff_t>(index));
    const float midpoint = interval.lower + 0.5f * (interval.upper - interval.lower);
    if (midpoint == interval.lower || midpoint == interval.upper) continue;
    GJXL_AMD_TRY(s.Evaluate(midpoint));
    intervals.push_back({interval.lower, midpoint});
    intervals.push_back({midpoint, interval.upper});
  }
  if (!s.best.valid) {
    return s.first_failure == GJXL_OK
      ? Fail(GJXL_ERROR_INTERNAL, "Target-size search produced no valid candidate")
      : Fail(s.first_failure, s.first_message);
  }
  GJXLAmdRateControlResult r{};
  r.selected_target = s.best.target;
  r.selected = s.best.candidate;
  r.attempts = s.attempts;
  r.failed_attempts = s.failed;
  r.target_size_met = Within(s.best.candidate.encoded_bytes, o);
  r.search_exhausted = !r.target_size_met && (s.attempts == o.maximum_attempts || intervals.empty());
  *result = r;
  return GJXL_OK;
}

}  // extern "C"
