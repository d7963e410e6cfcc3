"""CPU: pin the oracle's sRGB table, fast math and XYB transform.

The reference's colour transform (src/codec/color_transform.cpp) sits behind
the reference's GCC-13.3-only toolchain gate, so the XYB restatement is pinned
by the reference test's own closed-form ScalarOpsin within its 1e-6 bound
(tests/color_transform_test.cpp:33-46,340-364), by its invariants
(:101-160) and by closed forms; the sRGB table is compared bit for bit with
the reference's compiled image_conversion.cpp.
"""
import math

import numpy as np
import pytest

from oracle import binding


def test_srgb_table_bit_exact_with_reference(ref):
    values = np.arange(256, dtype=np.uint8)
    srgb = np.stack([values, values[::-1], np.roll(values, 7)], -1)[None]
    srgb = np.ascontiguousarray(srgb)  # 1 x 256 x 3
    expected = np.empty((3, 1, 256), np.float32)
    rc = ref.ref_srgb8_to_linear(binding.ptr(srgb), srgb.nbytes, 256 * 3, 3,
                                 256, 1, binding.ptr(expected))
    assert rc == 0
    got = binding.srgb8_to_linear(srgb)
    assert np.array_equal(got.view(np.uint32), expected.view(np.uint32))


def test_srgb_table_closed_form():
    table = np.empty(256, np.float32)
    binding.oracle().oracle_srgb_to_linear_table(binding.ptr(table))
    for i in (0, 1, 10, 11, 128, 254, 255):
        e = i / 255.0
        lin = e / 12.92 if e <= 0.04045 else ((e + 0.055) / 1.055) ** 2.4
        assert table[i] == np.float32(lin)
    assert table[0] == 0.0 and table[255] == 1.0
    assert np.all(np.diff(table) > 0)


def test_fast_log2_and_pow2_accuracy():
    lib = binding.oracle()
    xs = np.geomspace(1e-6, 1e6, 2001).astype(np.float32)
    err = max(abs(lib.oracle_fast_log2(float(x)) - math.log2(float(x)))
              for x in xs)
    assert err < 1e-4  # libjxl documents ~ 1e-5 absolute
    ys = np.linspace(-20, 20, 2001).astype(np.float32)
    rel = max(abs(lib.oracle_fast_pow2(float(y)) / 2.0 ** float(y) - 1.0)
              for y in ys)
    assert rel < 1e-5
    assert math.isnan(lib.oracle_fast_log2(0.0))
    assert math.isnan(lib.oracle_fast_pow2(200.0))
    assert abs(lib.oracle_fast_pow(9.0, 0.5) - 3.0) < 1e-4


def test_fast_cube_root_within_6_ulp():
    lib = binding.oracle()
    xs = np.geomspace(1e-9, 1e3, 4001).astype(np.float32)
    for x in xs:
        got = np.float32(lib.oracle_fast_cube_root_and_add(float(x), 0.0))
        want = np.float32(np.cbrt(np.float64(x)))
        ulp = np.spacing(want)
        assert abs(float(got) - float(want)) <= 6 * float(ulp), x
    assert lib.oracle_fast_cube_root_and_add(0.0, 0.25) == 0.25


def _fill_linear(h, w):
    y, x = np.mgrid[0:h, 0:w]
    img = np.empty((3, h, w), np.float32)
    img[0] = x.astype(np.float32) / np.float32(w - 1)
    img[1] = y.astype(np.float32) / np.float32(h - 1)
    img[2] = ((3 * x + 5 * y) % 17).astype(np.float32) / np.float32(16)
    return img


def test_xyb_round_trip_and_invariants():
    # tests/color_transform_test.cpp:101-160 bounds
    h, w = 13, 17
    lin = _fill_linear(h, w)
    opsin = binding.linear_to_padded_opsin(lin)
    assert opsin.shape == (3, 16, 24)
    back = binding.opsin_to_linear(np.ascontiguousarray(opsin[:, :h, :w]))
    assert np.max(np.abs(back - lin)) <= 2e-5
    black = binding.linear_to_padded_opsin(np.zeros((3, 8, 8), np.float32))
    assert abs(black[0, 0, 0]) <= 1e-8 and abs(black[1, 0, 0]) <= 1e-7
    assert abs(black[2, 0, 0]) <= 1e-7
    y, x = np.mgrid[0:h, 0:w]
    gray = np.repeat(((x + y + 1) / 17.0).astype(np.float32)[None], 3, 0)
    g = binding.linear_to_padded_opsin(gray)[:, :h, :w]
    assert np.max(np.abs(g[0])) <= 1e-6
    assert np.max(np.abs(g[1] - g[2])) <= 3e-5


def test_xyb_padding_replicates_edges():
    lin = np.random.default_rng(3).uniform(0, 1, (3, 13, 17)).astype(np.float32)
    opsin = binding.linear_to_padded_opsin(lin)
    for c in range(3):
        assert np.array_equal(opsin[c, :13, 17:], np.repeat(opsin[c, :13, 16:17], 7, 1))
        assert np.array_equal(opsin[c, 13:, :], np.repeat(opsin[c, 12:13, :], 3, 0))


def test_xyb_detects_swapped_matrix_rows():
    # A plausible mistake (X computed as gamma1 - gamma0) flips the sign of X
    # for a saturated red pixel; the oracle must give X > 0 for red.
    red = np.zeros((3, 8, 8), np.float32)
    red[0] = 1.0
    assert binding.linear_to_padded_opsin(red)[0, 0, 0] > 0.01


# tests/color_transform_test.cpp:24-30
_OPSIN_BIAS = np.float32(0.0037930732552754493)
_OPSIN_MATRIX = np.array([[0.30, 0.622, 0.078], [0.23, 0.692, 0.078],
                          [0.24342268924547819, 0.20476744424496821,
                           0.55180986650955360]], np.float32)


def _fma32(a, b, c):
    # binary32 operands: the binary64 product is exact
    return (a.astype(np.float64) * np.float64(b) + c.astype(np.float64)).astype(np.float32)


def _scalar_opsin(rgb: np.ndarray) -> np.ndarray:
    """ScalarOpsin (tests/color_transform_test.cpp:33-46) per pixel: fma
    order blue, green, red onto the bias, cube root, X = (g0 - g1) / 2,
    Y = (g0 + g1) / 2, B = g2."""
    bias_root = np.float32(np.cbrt(np.float64(_OPSIN_BIAS)))
    gamma = []
    for row in range(3):
        mixed = _fma32(rgb[2], _OPSIN_MATRIX[row][2], np.full_like(rgb[2], _OPSIN_BIAS))
        mixed = _fma32(rgb[1], _OPSIN_MATRIX[row][1], mixed)
        mixed = _fma32(rgb[0], _OPSIN_MATRIX[row][0], mixed)
        root = np.cbrt(np.maximum(mixed, 0).astype(np.float64)).astype(np.float32)
        gamma.append((root - bias_root).astype(np.float32))
    return np.stack([np.float32(0.5) * (gamma[0] - gamma[1]),
                     np.float32(0.5) * (gamma[0] + gamma[1]), gamma[2]])


def _ramp_input(w=256, h=256):
    """The reference test's 256x256 input (tests/color_transform_test.cpp:
    310-318)."""
    y, x = np.mgrid[0:h, 0:w]
    return np.stack([((3 * x + 5 * y) % 257) / np.float32(256.0),
                     ((7 * x + 2 * y) % 251) / np.float32(250.0),
                     ((11 * x + 13 * y) % 263) / np.float32(262.0)]).astype(np.float32)


@pytest.mark.parametrize("case", ["ramp", "uniform", "synthetic"])
def test_opsin_matches_scalar_opsin_within_1e6(case):
    if case == "ramp":
        rgb = _ramp_input()
    elif case == "uniform":
        rgb = np.random.default_rng(5).uniform(0, 1, (3, 61, 77)).astype(np.float32)
    else:
        from gjxl_amd import synthetic
        rgb = binding.srgb8_to_linear(synthetic.generate(96, 64, 3, "textured"))
    _, h, w = rgb.shape
    got = binding.linear_to_padded_opsin(rgb)[:, :h, :w]
    want = _scalar_opsin(rgb)
    err = float(np.max(np.abs(got - want)))
    assert err <= 1e-6, err
    # a transposed matrix (columns for rows) is far outside the bound
    swapped = _OPSIN_MATRIX.copy()
    try:
        _OPSIN_MATRIX[:] = swapped.T
        assert float(np.max(np.abs(got - _scalar_opsin(rgb)))) > 1e-3
    finally:
        _OPSIN_MATRIX[:] = swapped
