"""CPU: pin the oracle's Gaborish inverse and initial quant field.

Gaborish inverse: bit-exact against the reference's compiled gaborish.cpp /
convolution.cpp, and the libjxl points pinned by tests/gaborish_test.cpp.
Initial quant field: the reference unit sits behind its toolchain gate, so it
is pinned by the libjxl goldens in tests/adaptive_quantization_test.cpp and
by properties a dropped term or swapped index would break.
"""
import ctypes
import ctypes.util
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import binding

GOLDEN = json.loads((Path(__file__).parent / "golden" /
                     "reference_goldens.json").read_text())
_libm = ctypes.CDLL(ctypes.util.find_library("m"))
for _name in ("sinf", "cosf"):
    getattr(_libm, _name).argtypes = [ctypes.c_float]
    getattr(_libm, _name).restype = ctypes.c_float
F = np.float32


def sinf(v):
    return F(_libm.sinf(F(v)))


def cosf(v):
    return F(_libm.cosf(F(v)))


def test_mirror_matches_reference_rule():
    lib = binding.oracle()
    assert [lib.oracle_mirror(c, 5) for c in range(-3, 9)] == \
        [2, 1, 0, 0, 1, 2, 3, 4, 4, 3, 2, 1]
    assert [lib.oracle_mirror(c, 1) for c in (-2, -1, 0, 1, 2)] == [0, 0, 0, 0, 0]


@pytest.mark.parametrize("w,h,seed", [(24, 16, 1), (64, 40, 2), (8, 8, 3),
                                      (5, 3, 4)])
def test_gaborish_inverse_bit_exact_with_reference(ref, w, h, seed):
    img = np.random.default_rng(seed).normal(0, 0.3, (3, h, w)).astype(F)
    mult = np.array([1.0, 0.7, 1.15], F)
    want = np.empty_like(img)
    assert ref.ref_gaborish_inverse(binding.ptr(img), w, h, binding.ptr(mult),
                                    binding.ptr(want)) == 0
    got = binding.gaborish_inverse(img, mult)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_gaborish_inverse_pinned_libjxl_points():
    g = GOLDEN["gaborish_inverse"]
    w, h = g["width"], g["height"]
    img = np.empty((3, h, w), F)
    for c in range(3):
        for y in range(h):
            for x in range(w):
                img[c, y, x] = F(F(0.2) * F(c + 1)) + F(F(0.013) * F(x * x + 3 * y)) \
                    + sinf(F(0.17) * F((c + 2) * x + y))
    out = binding.gaborish_inverse(img, g["multipliers"])
    for c in range(3):
        for (x, y), want in zip(g["points"], g["goldens"][c]):
            assert abs(float(out[c, y, x]) - want) <= g["tolerance"]


def test_gaborish_weights_normalise_to_one():
    w = np.empty(6, F)
    binding.oracle().oracle_gaborish_inverse_weights(1.0, binding.ptr(w))
    d0, d1, d2, d4, d8, d5 = map(float, w)
    # 1 centre, 4 each of d1, d2, d4, d8 and 8 of d5 sum to one
    assert abs(d0 + 4 * (d1 + d2 + d4 + d8) + 8 * d5 - 1.0) < 1e-6
    # a constant image is preserved
    flat = np.full((3, 16, 16), 0.37, F)
    assert np.max(np.abs(binding.gaborish_inverse(flat) - 0.37)) < 1e-6


def _aq_input():
    g = GOLDEN["initial_quant_field"]
    w, h = g["width"], g["height"]
    img = np.empty((3, h, w), F)
    for y in range(h):
        for x in range(w):
            img[0, y, x] = F(F(0.025) * sinf(F(x + 1) * F(0.37))) - \
                F(F(0.012) * cosf(F(y + 2) * F(0.23)))
            img[1, y, x] = F(F(F(F(0.12) + F(F(0.035) * sinf(F(x + 2) * F(0.19)))) +
                             F(F(0.022) * cosf(F(y + 1) * F(0.31)))) +
                             F(F(0.002) * F((x * 7 + y * 11) % 9)))
            img[2, y, x] = F(F(F(0.145) + F(F(0.04) * sinf(F(x + y + 1) * F(0.17)))) -
                             F(F(0.018) * cosf(F(y + 3) * F(0.29))))
    return img


def test_initial_quant_field_pinned_libjxl_goldens():
    g = GOLDEN["initial_quant_field"]
    out = binding.initial_quant_field(_aq_input(), g["target"], g["rescale"])
    assert np.max(np.abs(out["quant_field"].ravel() - g["quant_field"])) <= \
        g["block_tolerance"]
    assert np.max(np.abs(out["strategy_mask"].ravel() - g["strategy_mask"])) <= \
        g["block_tolerance"]
    for (x, y), want in zip(g["pixel_points"], g["pixel_mask"]):
        assert abs(float(out["pixel_mask"][y, x]) - want) <= g["pixel_tolerance"]


def test_initial_quant_dc_pinned():
    g = GOLDEN["initial_quant_dc"]
    lib = binding.oracle()
    for t, want in zip(g["targets"], g["quant_dc"]):
        assert abs(lib.oracle_initial_quant_dc(t) - want) <= g["tolerance"]


def test_initial_quant_field_properties():
    # flat image: no Laplacian energy -> pixel mask is 1/0.01 everywhere and
    # every block gets the same field value
    flat = np.zeros((3, 32, 32), F)
    flat[1] = 0.3
    out = binding.initial_quant_field(flat)
    assert np.allclose(out["pixel_mask"], 100.0, rtol=1e-6)
    assert np.all(out["quant_field"] == out["quant_field"][0, 0])
    # more texture -> lower (finer) quant field: the field must react to the
    # Laplacian, which a dropped term would hide
    rng = np.random.default_rng(0)
    busy = flat.copy()
    busy[1] += rng.normal(0, 0.02, (32, 32)).astype(F)
    q_busy = binding.initial_quant_field(busy)["quant_field"]
    assert np.all(np.isfinite(q_busy)) and np.all(q_busy > 0)
    assert not np.allclose(q_busy, out["quant_field"])
