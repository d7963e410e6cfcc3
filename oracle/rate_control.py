"""Python restatement of the reference's target-size search (rate control).
TEST INFRASTRUCTURE ONLY: tests/ use it as the checker of
gjxl_amd_rate_control_search; the product never imports it.

It follows src/codestream/rate_control.cpp:246-320 SearchTargetSizeImpl and
its helpers (:24-200): evaluate the lower Butteraugli target, then the upper
one unless the first met the tolerance, then midpoints of the widest
remaining interval (ties to the lower interval) until the size is met, the
attempts run out or no interval can be split.  Candidates compare by
`better_candidate` (:56-103); evaluator failures are candidate-local except a
resource-plan overrun (:147-150); an invalid successful candidate is fatal
(:157-168); a failed search leaves the result unchanged.

Pinned by every case of the reference's own tests/rate_control_search_test.cpp
(tests/test_oracle_rate_control.py).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

MINIMUM_TARGET = 0.01   # kMinimumTargetSizeButteraugliTarget (float)
MAXIMUM_TARGET = 10.0   # kMaximumTargetSizeButteraugliTarget
MAXIMUM_ATTEMPTS = 64   # kMaximumTargetSizeEncodeAttempts
LARGEST_AT_OR_BELOW, CLOSEST_ABSOLUTE = 0, 1

OK, INVALID_ARGUMENT, INTERNAL, RESOURCE_PLAN_EXCEEDED, FAILED = 0, 1, 6, 7, 99


def f32(x: float) -> float:
    import numpy as np
    return float(np.float32(x))


@dataclass
class Candidate:
    encoded_bytes: int
    score_evaluated: bool = True
    resident: bool = False
    final_score: float = 0.0
    target: float = 0.0

    def score(self) -> float:
        # FinalScore: resident GPU scores and unevaluated ones are no tie-break
        return math.inf if self.resident or not self.score_evaluated else self.final_score


@dataclass
class Result:
    best: Candidate | None = None
    attempts: int = 0
    failed_attempts: int = 0
    target_size_met: bool = False
    search_exhausted: bool = False
    evaluated: list = field(default_factory=list)  # targets in evaluation order


def within(size: int, o: dict) -> bool:
    if o["selection"] == CLOSEST_ABSOLUTE:
        return abs(size - o["target_bytes"]) <= o["tolerance_bytes"]
    return size <= o["target_bytes"] and o["target_bytes"] - size <= o["tolerance_bytes"]


def better(o: dict, c: Candidate, best: Candidate | None) -> bool:
    if best is None:
        return True
    t = o["target_bytes"]
    cu, bu = c.encoded_bytes <= t, best.encoded_bytes <= t
    if o["selection"] == CLOSEST_ABSOLUTE:
        ce, be = abs(c.encoded_bytes - t), abs(best.encoded_bytes - t)
        if ce != be:
            return ce < be
        if cu != bu:
            return cu
    else:
        if cu != bu:
            return cu
        if c.encoded_bytes != best.encoded_bytes:
            return c.encoded_bytes > best.encoded_bytes if cu else c.encoded_bytes < best.encoded_bytes
    cs, bs = c.score(), best.score()
    if cs != bs:
        return cs < bs
    return c.target < best.target


def options(target_bytes, maximum_attempts, tolerance_bytes=0, selection=LARGEST_AT_OR_BELOW,
            minimum=MINIMUM_TARGET, maximum=MAXIMUM_TARGET) -> dict:
    return dict(target_bytes=target_bytes, tolerance_bytes=tolerance_bytes,
                maximum_attempts=maximum_attempts, selection=selection,
                minimum=f32(minimum), maximum=f32(maximum))


def search(o: dict, evaluator) -> tuple[int, Result | None]:
    """evaluator(target) -> (status, Candidate | None).  Returns (status,
    result); result is None when the search fails."""
    if (o["selection"] not in (LARGEST_AT_OR_BELOW, CLOSEST_ABSOLUTE) or evaluator is None or
            o["target_bytes"] == 0 or o["maximum_attempts"] == 0 or
            o["maximum_attempts"] > MAXIMUM_ATTEMPTS or not math.isfinite(o["minimum"]) or
            not math.isfinite(o["maximum"]) or o["minimum"] <= 0 or
            o["maximum"] <= o["minimum"]):
        return INVALID_ARGUMENT, None
    r = Result()
    first_failure = [OK]

    def evaluate(target: float) -> int:
        r.attempts += 1
        r.evaluated.append(target)
        status, c = evaluator(target)
        if status != OK:
            if status == RESOURCE_PLAN_EXCEEDED:
                return status
            r.failed_attempts += 1
            if first_failure[0] == OK:
                first_failure[0] = status
            return OK
        if (c is None or c.encoded_bytes <= 0 or
                (c.score_evaluated and not math.isfinite(c.final_score))):
            return INTERNAL
        c.target = target
        if better(o, c, r.best):
            r.best = c
        return OK

    lower, upper = o["minimum"], o["maximum"]
    s = evaluate(lower)
    if s != OK:
        return s, None
    if r.attempts < o["maximum_attempts"] and (
            r.best is None or not within(r.best.encoded_bytes, o)):
        s = evaluate(upper)
        if s != OK:
            return s, None
    intervals = [(lower, upper)]
    while (r.attempts < o["maximum_attempts"] and
           (r.best is None or not within(r.best.encoded_bytes, o)) and intervals):
        # the widest interval, ties to the lower one (WidestInterval)
        index = 0
        for i in range(1, len(intervals)):
            w = f32(intervals[i][1] - intervals[i][0])
            bw = f32(intervals[index][1] - intervals[index][0])
            if w > bw or (w == bw and intervals[i][0] < intervals[index][0]):
                index = i
        lo, hi = intervals.pop(index)
        mid = f32(lo + f32(0.5 * f32(hi - lo)))
        if mid == lo or mid == hi:
            continue
        s = evaluate(mid)
        if s != OK:
            return s, None
        intervals.append((lo, mid))
        intervals.append((mid, hi))
    if r.best is None:
        return (INTERNAL if first_failure[0] == OK else first_failure[0]), None
    r.target_size_met = within(r.best.encoded_bytes, o)
    r.search_exhausted = not r.target_size_met and (
        r.attempts == o["maximum_attempts"] or not intervals)
    return OK, r
