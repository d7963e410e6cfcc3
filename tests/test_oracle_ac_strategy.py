"""CPU: pin the oracle's quantization matrices, DCT, initial CfL, candidate
cost and greedy merge against the reference's own compiled sources
(quantization.cpp, dct.cpp, chroma_from_luma.cpp, ac_strategy.cpp) and
against closed forms that a transposed or dropped term would break."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from gjxl_amd import synthetic
from oracle import binding

F = np.float32
STRATEGIES = (0, 4, 5, 6, 7, 10, 11)


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("strategy", STRATEGIES)
def test_quant_matrices_bit_exact_with_reference(ref, strategy):
    import ctypes
    for channel in range(3):
        dq = np.empty(1024, F)
        inv = np.empty(1024, F)
        n = ctypes.c_size_t()
        assert ref.ref_quant_matrix(strategy, channel, binding.ptr(dq),
                                    binding.ptr(inv), ctypes.byref(n)) == 0
        odq, oinv = binding.quant_matrix(strategy, channel)
        assert len(odq) == n.value
        assert np.array_equal(_bits(odq), _bits(dq[:n.value]))
        assert np.array_equal(_bits(oinv), _bits(inv[:n.value]))


def test_quant_matrix_shape_properties():
    dq, inv = binding.quant_matrix(0, 1)
    assert inv[0] == 0.0 and np.all(inv[1:] > 0)
    # weights grow with frequency along both axes
    w = 1.0 / dq.reshape(8, 8)
    assert np.all(np.diff(w[0]) <= 0) or np.all(np.diff(w[0]) >= 0)
    # the DC weight is the first band parameter of each channel
    assert abs(float(w[0, 0]) - 560.0) < 1e-3
    assert abs(float(1.0 / binding.quant_matrix(0, 0)[0][0]) - 3150.0) < 1e-2


@pytest.mark.parametrize("strategy", STRATEGIES)
def test_dct_bit_exact_with_reference_and_textbook(ref, strategy):
    cw, ch = binding.COVERED[strategy]
    W, H = cw * 8, ch * 8
    rng = np.random.default_rng(strategy)
    pix = rng.normal(0, 1, W * H).astype(F)
    want = np.empty(W * H, F)
    assert ref.ref_forward_dct(strategy, binding.ptr(pix), binding.ptr(want)) == 0
    got = np.empty(W * H, F)
    binding.oracle().oracle_forward_dct(strategy, binding.ptr(pix), binding.ptr(got))
    assert np.array_equal(_bits(got), _bits(want))
    back_ref = np.empty(W * H, F)
    back = np.empty(W * H, F)
    ref.ref_inverse_dct(strategy, binding.ptr(want), binding.ptr(back_ref))
    binding.oracle().oracle_inverse_dct(strategy, binding.ptr(got), binding.ptr(back))
    assert np.array_equal(_bits(back), _bits(back_ref))
    assert np.max(np.abs(back - pix)) < 1e-5  # orthonormal round trip
    # textbook orthonormal DCT-II with the libjxl scaling (1/N per axis)
    img = pix.astype(np.float64).reshape(H, W)

    def mat(n):
        k = np.arange(n)[:, None]
        s = np.arange(n)[None, :]
        m = np.sqrt(2.0) / n * np.cos((s + 0.5) * k * np.pi / n)
        m[0] /= np.sqrt(2.0)
        return m
    natural = mat(H) @ img @ mat(W).T  # [v][u]
    stored = natural if H < W else natural.T
    assert np.max(np.abs(stored.ravel() - got)) < 1e-5


def _opsin(w, h, seed):
    srgb = np.ascontiguousarray(synthetic.generate(max(w, 16), max(h, 16), seed)[:h, :w])
    ops = binding.linear_to_padded_opsin(binding.srgb8_to_linear(srgb))
    return ops, binding.gaborish_inverse(ops)


@pytest.mark.parametrize("w,h,seed", [(128, 128, 1), (200, 72, 2)])
def test_initial_cfl_equals_reference(ref, w, h, seed):
    _, pre = _opsin(w, h, seed)
    _, ph, pw = pre.shape
    tw, th = (pw + 63) // 64, (ph + 63) // 64
    rx = np.empty((th, tw), np.int8)
    rb = np.empty((th, tw), np.int8)
    assert ref.ref_initial_cfl(binding.ptr(pre), pw, ph, binding.ptr(rx),
                               binding.ptr(rb)) == 0
    ox, ob = binding.initial_cfl(pre)
    assert np.array_equal(ox, rx) and np.array_equal(ob, rb)


def test_find_best_multiplier_recovers_planted_factor():
    # chroma = k/84 * luma exactly -> the LS minimiser is x = k (then the
    # towards-zero quantiser subtracts 2.6 and rounds)
    rng = np.random.default_rng(0)
    luma = rng.normal(0, 30, 4096).astype(F)
    for k in (-40.0, 17.0):
        chroma = (luma * F(k / 84.0)).astype(F)
        fast = binding.oracle().oracle_find_best_multiplier(
            binding.ptr(luma), binding.ptr(chroma), 4096, 0.0, 1, 20)
        assert fast == int(round(k - math.copysign(2.6, k)))


def test_costs_bit_exact_with_reference(ref):
    ops, pre = _opsin(128, 96, 3)
    _, h, w = pre.shape
    aq = binding.initial_quant_field(ops)
    qf, mask = aq["quant_field"], aq["pixel_mask"]
    rng = np.random.default_rng(5)
    import ctypes
    for strategy in STRATEGIES:
        cw, ch = binding.COVERED[strategy]
        for _ in range(6):
            bx = int(rng.integers(0, w // 8 - cw + 1))
            by = int(rng.integers(0, h // 8 - ch + 1))
            cx, cb = float(rng.integers(-30, 30)) / 84, 1 + float(rng.integers(-30, 30)) / 84
            want = ctypes.c_float()
            got = ctypes.c_float()
            assert ref.ref_ac_strategy_cost(strategy, bx, by, binding.ptr(pre), w, h,
                                            binding.ptr(qf), binding.ptr(mask), 1.0,
                                            1.34, cx, cb, ctypes.byref(want)) == 0
            assert binding.oracle().oracle_ac_strategy_cost(
                strategy, bx, by, binding.ptr(pre), w, h, binding.ptr(qf),
                binding.ptr(mask), 1.0, 1.34, cx, cb, ctypes.byref(got)) == 0
            assert _bits(got.value) == _bits(want.value)


@pytest.mark.parametrize("w,h,seed", [(256, 256, 1), (200, 136, 4), (72, 40, 7)])
def test_strategy_grid_equals_reference(ref, w, h, seed):
    ops, pre = _opsin(w, h, seed)
    _, ph, pw = pre.shape
    aq = binding.initial_quant_field(ops)
    cfl_x, cfl_b = binding.initial_cfl(pre)
    costs = binding.candidate_costs(pre, aq["quant_field"], aq["pixel_mask"],
                                    cfl_x, cfl_b)
    cells = binding.search_from_costs(costs)
    want = np.empty((ph // 8, pw // 8), np.uint8)
    assert ref.ref_find_ac_strategy_grid(
        binding.ptr(pre), pw, ph, binding.ptr(aq["quant_field"]),
        binding.ptr(aq["pixel_mask"]), 1.0, binding.ptr(want)) == 0
    assert np.array_equal(cells, want)
    if w >= 256:
        # the search picks several transform families on the larger image
        assert len(np.unique(cells >> 1)) >= 2
