"""CPU: the library's target-size search (gjxl_amd_rate_control_search, host
code) against the oracle (oracle/rate_control.py, itself pinned to the
reference's rate_control_search_test.cpp): the same evaluated targets in the
same order, the same selection, counts and flags, on the reference's cases
and on seeded random size functions."""
import numpy as np
import pytest

import gjxl_amd
from gjxl_amd import rate_control as rc
from oracle import rate_control as R


def _both(target_bytes, size_for_target, score=lambda t: t, evaluated=True, resident=False,
          fail=lambda t: 0, **kw):
    lib_tried, ora_tried = [], []

    def lib_eval(t):
        lib_tried.append(t)
        code = fail(t)
        return code, size_for_target(t), evaluated, resident, score(t)

    def ora_eval(t):
        ora_tried.append(t)
        code = fail(t)
        if code:
            return (R.RESOURCE_PLAN_EXCEEDED if code == 7 else R.FAILED), None
        return R.OK, R.Candidate(size_for_target(t), evaluated, resident, score(t))

    o = R.options(target_bytes, kw.get("maximum_attempts", 12),
                  kw.get("tolerance_bytes", 0), kw.get("selection", 0))
    status, want = R.search(o, ora_eval)
    try:
        got = rc.search(rc.options(target_bytes, **kw), lib_eval)
        code = 0
    except gjxl_amd.GjxlError as e:
        got, code = None, e.code
    return status, want, code, got, ora_tried, lib_tried


def _same(status, want, code, got, ora_tried, lib_tried):
    assert lib_tried == ora_tried
    if status != R.OK:
        assert got is None and code == {R.FAILED: 5, R.INTERNAL: 6,
                                        R.RESOURCE_PLAN_EXCEEDED: 7}[status]
        return
    assert code == 0
    assert got.selected.encoded_bytes == want.best.encoded_bytes
    assert got.selected_target == want.best.target
    assert (got.attempts, got.failed_attempts) == (want.attempts, want.failed_attempts)
    assert bool(got.target_size_met) == want.target_size_met
    assert bool(got.search_exhausted) == want.search_exhausted


def monotonic(t):
    return 2000 - min(1000, int(R.f32(R.f32(t) * R.f32(100.0))))


@pytest.mark.parametrize("target,attempts", [(1500, 12), (2500, 12), (900, 12), (1500, 2)])
def test_reference_cases(target, attempts):
    _same(*_both(target, monotonic, maximum_attempts=attempts))


def test_plateau_selection_and_failures():
    _same(*_both(900, lambda t: 1000))
    _same(*_both(1500, lambda t: 1600 if t < 1.0 else 1300, maximum_attempts=2,
                 selection=rc.CLOSEST_ABSOLUTE))
    _same(*_both(1500, lambda t: 1600 if t < 1.0 else 1300, maximum_attempts=8,
                 tolerance_bytes=100, selection=rc.CLOSEST_ABSOLUTE))
    _same(*_both(1500, monotonic, maximum_attempts=8,
                 fail=lambda t: 5 if t == R.f32(R.MAXIMUM_TARGET) else 0))
    _same(*_both(1500, monotonic, fail=lambda t: 5))
    _same(*_both(1500, monotonic, fail=lambda t: 7))
    _same(*_both(1000, lambda t: 1100, score=lambda t: 100.0 if t == R.f32(0.01) else 0.0,
                 resident=True, maximum_attempts=2))


@pytest.mark.parametrize("seed", range(12))
def test_random_size_functions(seed):
    rng = np.random.default_rng(seed)
    steps = np.sort(rng.uniform(0.01, 10.0, 6))
    sizes = np.sort(rng.integers(500, 5000, 7))[::-1]
    if seed % 3 == 0:
        sizes = rng.integers(500, 5000, 7)  # non-monotonic
    size = lambda t: int(sizes[int(np.searchsorted(steps, t))])  # noqa: E731
    target = int(rng.integers(400, 5200))
    _same(*_both(target, size, maximum_attempts=int(rng.integers(1, 20)),
                 tolerance_bytes=int(rng.integers(0, 200)), selection=int(seed % 2)))


def test_invalid_options_leave_the_result():
    r = rc.Result()
    r.attempts = 77
    for kw in (dict(maximum_attempts=0), dict(maximum_attempts=65),
               dict(minimum_target=1.0, maximum_target=1.0), dict(selection=9)):
        with pytest.raises(gjxl_amd.GjxlError, match="INVALID_ARGUMENT"):
            rc.search(rc.options(1000, **kw), lambda t: (0, 1, 1, 0, 0.0), r)
        assert r.attempts == 77
    with pytest.raises(gjxl_amd.GjxlError, match="INVALID_ARGUMENT"):
        rc.search(rc.options(0), lambda t: (0, 1, 1, 0, 0.0), r)
