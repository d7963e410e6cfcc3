"""CPU: the target-size search's oracle (oracle/rate_control.py) against every
case of the reference's tests/rate_control_search_test.cpp (the expected
sizes, attempt counts, flags and selected targets are that file's)."""
import math

from oracle import rate_control as R


def evaluator(size_for_target, score_for_target=lambda t: t, evaluated=True, resident=False):
    def run(target):
        return R.OK, R.Candidate(size_for_target(target), evaluated, resident,
                                 score_for_target(target))
    return run


def monotonic(target):  # MonotonicSize: 2000 - min(1000, int(target * 100)) (float)
    return 2000 - min(1000, int(R.f32(R.f32(target) * R.f32(100.0))))


def test_bracketed_search():
    s, r = R.search(R.options(1500, 12), evaluator(monotonic))
    assert s == R.OK and r.best.encoded_bytes == 1500 and r.attempts == 3 and r.target_size_met


def test_infeasible_targets():
    s, r = R.search(R.options(2500, 12), evaluator(monotonic))
    assert s == R.OK and r.best.encoded_bytes == 1999 and r.attempts == 12
    assert not r.target_size_met and r.search_exhausted and r.best.target == R.f32(R.MINIMUM_TARGET)
    s, r = R.search(R.options(900, 12), evaluator(monotonic))
    assert s == R.OK and r.best.encoded_bytes == 1000 and r.attempts == 12
    assert not r.target_size_met and r.search_exhausted and r.best.target == R.f32(R.MAXIMUM_TARGET)


def test_attempt_limit_and_plateau():
    s, r = R.search(R.options(1500, 2), evaluator(monotonic))
    assert s == R.OK and r.best.encoded_bytes == 1000 and r.attempts == 2
    assert not r.target_size_met and r.search_exhausted
    s, r = R.search(R.options(900, 12), evaluator(lambda t: 1000))
    assert s == R.OK and r.best.encoded_bytes == 1000 and r.attempts == 12
    assert not r.target_size_met and r.search_exhausted and r.best.target == R.f32(R.MINIMUM_TARGET)


def test_non_monotonic_selection():
    def size(t):
        if t > 7.0:
            return 1000
        if t > 4.0:
            return 1400
        if t > 2.0:
            return 1300
        if t > 1.0:
            return 1500
        return 2000
    s, r = R.search(R.options(1500, 8), evaluator(size))
    assert s == R.OK and r.best.encoded_bytes == 1500 and r.attempts == 7
    assert r.target_size_met and not r.search_exhausted


def test_failures_are_atomic_or_candidate_local():
    s, r = R.search(R.options(1500, 12), lambda t: (R.FAILED, None))
    assert s == R.FAILED and r is None
    s, r = R.search(R.options(1500, 12), lambda t: (R.OK, R.Candidate(0)))
    assert s == R.INTERNAL and r is None
    s, r = R.search(R.options(1500, 12), lambda t: (R.RESOURCE_PLAN_EXCEEDED, None))
    assert s == R.RESOURCE_PLAN_EXCEEDED and r is None
    failures = []

    def intermittent(t):
        if t == R.f32(R.MAXIMUM_TARGET):
            failures.append(t)
            return R.FAILED, None
        return R.OK, R.Candidate(monotonic(t), True, False, t)
    s, r = R.search(R.options(1500, 8), intermittent)
    assert s == R.OK and r.best.encoded_bytes == 1500 and r.attempts == 3
    assert r.failed_attempts == 1 and len(failures) == 1 and r.target_size_met


def test_selection_policies():
    ev = evaluator(lambda t: 1600 if t < 1.0 else 1300)
    s, under = R.search(R.options(1500, 2, selection=R.LARGEST_AT_OR_BELOW), ev)
    s2, closest = R.search(R.options(1500, 2, selection=R.CLOSEST_ABSOLUTE), ev)
    assert s == s2 == R.OK
    assert under.best.encoded_bytes == 1300 and closest.best.encoded_bytes == 1600
    assert not under.target_size_met and not closest.target_size_met
    assert under.search_exhausted and closest.search_exhausted
    s, sym = R.search(R.options(1500, 8, tolerance_bytes=100, selection=R.CLOSEST_ABSOLUTE), ev)
    assert s == R.OK and sym.best.encoded_bytes == 1600 and sym.attempts == 1
    assert sym.target_size_met and not sym.search_exhausted
    s, tie = R.search(R.options(1500, 2, selection=R.CLOSEST_ABSOLUTE),
                      evaluator(lambda t: 1600 if t < 1.0 else 1400))
    assert s == R.OK and tie.best.encoded_bytes == 1400


def test_scoreless_and_diagnostic_scores():
    s, r = R.search(R.options(1000, 4), evaluator(lambda t: 1000, evaluated=False))
    assert s == R.OK and r.best.encoded_bytes == 1000 and r.attempts == 1
    assert r.target_size_met and not r.search_exhausted
    score = lambda t: 100.0 if t == R.f32(R.MINIMUM_TARGET) else 0.0  # noqa: E731
    for kw in (dict(evaluated=False), dict(resident=True)):
        s, r = R.search(R.options(1000, 2), evaluator(lambda t: 1100, score, **kw))
        assert s == R.OK and r.best.encoded_bytes == 1100 and r.attempts == 2
        assert not r.target_size_met and r.search_exhausted
        assert r.best.target == R.f32(R.MINIMUM_TARGET)


def test_invalid_options():
    ev = evaluator(monotonic)
    for o in (R.options(0, 12), R.options(1000, 0), R.options(1000, R.MAXIMUM_ATTEMPTS + 1),
              R.options(1000, 12, minimum=1.0, maximum=1.0),
              R.options(1000, 12, minimum=math.nan), R.options(1000, 12, selection=99)):
        assert R.search(o, ev) == (R.INVALID_ARGUMENT, None)
    assert R.search(R.options(1000, 12), None) == (R.INVALID_ARGUMENT, None)
