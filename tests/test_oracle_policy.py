"""Pins the AQ policy pieces of the oracle whose reference units are behind the
reference's toolchain gate (src/codestream/workflow.cpp,
src/codec/adaptive_quantization.cpp): matrix-scale statistics and selection,
AdjustQuantField and the Butteraugli policy bounds.  Known answers come from
the reference's own tests (tests/codestream_workflow_test.cpp:89-263,
tests/adaptive_quantization_test.cpp:502-557); the bounds from their closed
form.
"""
import ctypes

import numpy as np
import pytest

from oracle import binding as b


def _stats(planes, w, h):
    b._declare_aq()
    opsin = np.ascontiguousarray(np.array(planes, np.float32).reshape(3, h, w))
    out = np.empty(3, np.float32)
    b.oracle().oracle_matrix_scale_stats(b.ptr(opsin), w, h, w, w * h, b.ptr(out))
    return out


def _select(stats, target):
    b._declare_aq()
    s = np.array(stats, np.float32)
    xs, bs = ctypes.c_int(), ctypes.c_int()
    b.oracle().oracle_select_matrix_scales(b.ptr(s), target, ctypes.byref(xs),
                                           ctypes.byref(bs))
    return xs.value, bs.value


def test_matrix_scale_stats_known_answer():
    # tests/codestream_workflow_test.cpp:89-104
    st = _stats([[0.0, -0.1, 0.05, 0.2], [0.0, 0.1, 0.2, 0.1], [0.0, 0.3, 0.2, 0.5]], 2, 2)
    assert abs(st[0] - 0.3) <= 1e-6 and abs(st[1] - 0.4) <= 1e-6
    assert abs(st[2] - 0.19) <= 1e-6
    # a single pixel has no neighbours: all statistics stay zero (:150-165)
    assert np.all(_stats([[0.2], [0.1], [0.4]], 1, 1) == 0)


def _up(v):
    return float(np.nextafter(np.float32(v), np.float32(np.inf)))


def _down(v):
    return float(np.nextafter(np.float32(v), np.float32(-np.inf)))


@pytest.mark.parametrize("stats,target,expected", [
    # tests/codestream_workflow_test.cpp:218-263
    ((0, 0, 0), 2.5, (3, 2)), ((0, 0, 0), _up(2.5), (4, 2)),
    ((0, 0, 0), 5.5, (4, 2)), ((0, 0, 0), _up(5.5), (5, 2)),
    ((0, 0, 0), 9.5, (5, 2)), ((0, 0, 0), _up(9.5), (6, 2)),
    ((_down(0.022), 0, 0), 1.0, (3, 2)), ((0.015, 0, 0), 1.0, (3, 2)),
    ((0.022, 0, 0), 1.0, (4, 2)), ((_down(0.026), 0, 0), 1.0, (4, 2)),
    ((0.026, 0, 0), 1.0, (5, 2)), ((0, 0.33, 0), 1.0, (3, 2)),
    ((0, _up(0.33), 0), 1.0, (3, 3)), ((0, 0.38, 0), 1.0, (3, 3)),
    ((0, _up(0.38), 0), 1.0, (3, 4)), ((0, 0.28, 0.13), 1.0, (3, 2)),
    ((0, _up(0.28), 0.13), 1.0, (3, 3)), ((0, 0.34, _down(0.13)), 1.0, (3, 3)),
    ((0, 0.34, 0.13), 1.0, (3, 4)), ((0, _up(0.38), 0.13), 1.0, (3, 5)),
])
def test_matrix_scale_selection_thresholds(stats, target, expected):
    assert _select(stats, target) == expected


def test_adjust_quant_field_pinned_libjxl():
    # tests/adaptive_quantization_test.cpp:502-557: one DCT32x32 over 4x4
    # blocks, field 0.48 + 0.01 (x + y)
    b._declare_aq()
    cells = np.full((4, 4), 5 << 1, np.uint8)
    cells[0, 0] |= 1
    field = np.array([[0.48 + 0.01 * (x + y) for x in range(4)] for y in range(4)],
                     np.float32)
    out = np.empty_like(field)
    b.oracle().oracle_adjust_quant_field(b.ptr(cells), 4, 4, 3.0, b.ptr(field), b.ptr(out))
    assert np.all(np.abs(out - np.float32(0.515324116)) <= 2e-7)
    b.oracle().oracle_adjust_quant_field(b.ptr(cells), 4, 4, 1.2, b.ptr(field), b.ptr(out))
    assert np.all(np.abs(out - np.float32(0.54)) <= 2e-7)
    # DCT8 cells keep their own value
    dct8 = np.ones((4, 4), np.uint8)
    b.oracle().oracle_adjust_quant_field(b.ptr(dct8), 4, 4, 3.0, b.ptr(field), b.ptr(out))
    assert np.array_equal(out, field)


def test_policy_bounds_closed_form():
    """PrepareButteraugliPolicyFromRange (adaptive_quantization.cpp:1197-1242):
    ratio r = max/min, d = sqrt(250 / r), a = min(2, d); bounds
    min / (a d) and max (d / a); quant dc from ComputeInitialQuantDc."""
    b._declare_aq()
    field = np.array([0.5, 1.0, 2.0, 0.75], np.float32)
    lo, hi, qdc = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    rc = b.oracle().oracle_policy_setup(b.ptr(field), 4, 1.0, ctypes.byref(lo),
                                        ctypes.byref(hi), ctypes.byref(qdc))
    assert rc == 0
    f = np.float32
    d = np.sqrt(f(250.0) / (f(2.0) / f(0.5)), dtype=np.float32)
    a = min(f(2.0), d)
    assert lo.value == f(f(0.5) / (a * d))
    assert hi.value == f(f(2.0) * (d / a))
    assert qdc.value == b.oracle().oracle_initial_quant_dc(1.0)
    # the bounds always span a ratio of r d^2 = 250, whatever the range
    for lo_v, hi_v in ((0.5, 2.0), (0.001, 1.0), (1.0, 1.0), (0.2, 40.0)):
        f2 = np.array([lo_v, hi_v], np.float32)
        assert b.oracle().oracle_policy_setup(b.ptr(f2), 2, 1.0, ctypes.byref(lo),
                                              ctypes.byref(hi), ctypes.byref(qdc)) == 0
        assert abs(hi.value / lo.value - 250.0) < 1e-2
        if hi_v / lo_v <= 250.0:  # a range within 250 lies inside its bounds
            assert lo.value <= lo_v and hi.value >= hi_v
