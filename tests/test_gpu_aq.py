"""GPU parity: one adaptive-quantization evaluation (quantizer, final CfL,
coefficient coding with prediction-aware DC, reconstruction with DC
smoothing, Gaborish + EPF, XYB -> linear) and its per-encode preparation.

Every output is bit-exact with the oracle on the same inputs: the strategy
grid, fields and XYB planes are taken from the oracle's frontend (small
sizes) or from the device frontend (full size), so that this test isolates
the evaluation.
"""
import numpy as np
import pytest

import gjxl_amd
from gjxl_amd import synthetic
from oracle import binding

pytestmark = pytest.mark.gpu


def _device_inputs(device, fe):
    return (device.array(fe["opsin"]), device.array(fe["pre"]),
            device.array(np.ascontiguousarray(fe["cells"])),
            device.array(fe["quant_field"]))


def _session_outputs(s: gjxl_amd.AqSession) -> dict:
    bh, bw, ph, pw = s.bh, s.bw, s.ph, s.pw
    h, w = s.height, s.width
    store = binding.store_size(bw, bh)
    return {
        "raw": s.buffer("raw_quant", (bh, bw), np.int32).numpy(),
        "y_to_x": s.buffer("y_to_x", ((bh + 7) // 8, (bw + 7) // 8), np.int8).numpy(),
        "y_to_b": s.buffer("y_to_b", ((bh + 7) // 8, (bw + 7) // 8), np.int8).numpy(),
        "quantized": s.buffer("quantized", (3, store), np.int32).numpy(),
        "dc_q": s.buffer("dc_quantized", (3, bh, bw), np.int32).numpy(),
        "dc_rec": s.buffer("dc_reconstructed", (3, bh, bw)).numpy(),
        "inv_sigma": s.buffer("inverse_sigma", (bh, bw)).numpy(),
        "reconstructed": s.buffer("reconstructed", (3, ph, pw)).numpy(),
        "filtered": s.buffer("filtered", (3, h, w)).numpy(),
        "linear": s.buffer("linear", (3, h, w)).numpy(),
    }


def _same(name, got, want):
    if got.dtype.kind == "f":
        ok = np.array_equal(got.view(np.int32), want.astype(got.dtype).view(np.int32))
    else:
        ok = np.array_equal(got, want.astype(got.dtype))
    if not ok:
        diff = np.argwhere(got != want)
        first = tuple(diff[0]) if len(diff) else None
        raise AssertionError(
            f"{name}: {len(diff)} entries differ, first at {first}: "
            f"{got[first] if first else ''} vs {want[first] if first else ''}")


def _check_prepare(s, setup):
    pol = s.policy()
    assert pol["x_scale"] == setup["x_scale"] and pol["b_scale"] == setup["b_scale"]
    for key in ("lower", "upper", "quant_dc"):
        assert np.float32(pol[key]) == np.float32(setup[key]), key
    offsets = s.buffer("offsets", (s.bh, s.bw), np.uint32).numpy()
    want = setup["offsets"].astype(np.uint64)
    want = np.where(want == np.uint64(2**64 - 1), np.uint64(0xFFFFFFFF), want)
    _same("offsets", offsets, want.astype(np.uint32))
    _same("adjusted_field", s.buffer("adjusted_field", (s.bh, s.bw)).numpy(),
          setup["adjusted"])
    store = binding.store_size(s.bw, s.bh)
    _same("coefficients", s.buffer("coefficients", (3, store)).numpy(), setup["coefs"])


def _check_evaluation(s, fe, setup, field, w, h):
    want = binding.aq_evaluate(fe, setup, field, w, h)
    got = _session_outputs(s)
    q = s.buffer("quantizer", (10,), np.float32).numpy()
    qi = q.view(np.uint32)
    assert q[0] == np.float32(want["q"][0]) and q[1] == np.float32(want["q"][1])
    assert qi[2] == int(want["q"][2]) and qi[3] == int(want["q"][3])
    _same("dc_steps", q[4:7], want["q"][4:7])
    _same("inverse_dc_steps", q[7:10], want["q"][7:10])
    for key in ("y_to_x", "y_to_b", "raw", "quantized", "dc_q", "dc_rec",
                "inv_sigma", "reconstructed", "filtered", "linear"):
        _same(key, got[key], want[key])


# 3x5, 2x9 and 11x3 mirror the loop filters' tiles more than once per side
@pytest.mark.parametrize("w,h,seed", [(512, 512, 1), (200, 136, 4), (72, 40, 7),
                                      (8, 8, 3), (136, 264, 9), (2104, 48, 5),
                                      (40, 2072, 6), (3, 5, 2), (2, 9, 8), (11, 3, 10)])
def test_aq_evaluation_parity(device, w, h, seed):
    print(f"aq parity {w}x{h} seed {seed}", flush=True)
    srgb = np.ascontiguousarray(
        synthetic.generate(max(w, 16), max(h, 16), seed, "textured")[:h, :w])
    fe = binding.frontend(srgb)
    setup = binding.aq_setup(fe, w, h)
    d_ops, d_pre, d_cells, d_field = _device_inputs(device, fe)
    s = gjxl_amd.AqSession(device, w, h)
    s.prepare(d_ops, d_pre, d_cells, d_field, 1.0)
    device.synchronize()
    _check_prepare(s, setup)
    # the adjusted field, then a coarser one (a second evaluation reuses the
    # session and its coefficient store)
    adjusted = s.buffer("adjusted_field", (s.bh, s.bw))
    s.evaluate(adjusted)
    device.synchronize()
    _check_evaluation(s, fe, setup, setup["adjusted"], w, h)
    coarse = (setup["adjusted"] * np.float32(0.55)).astype(np.float32)
    s.evaluate(device.array(coarse))
    device.synchronize()
    _check_evaluation(s, fe, setup, coarse, w, h)
    s.close()


def test_aq_evaluation_full_size_4032x3024(device):
    w, h = 4032, 3024
    print("aq parity 4032x3024", flush=True)
    srgb = synthetic.generate(**synthetic.WORKLOADS["4032x3024"])
    d_srgb = device.array(srgb)
    _, d_ops = device.srgb8_to_linear_and_opsin(d_srgb, w, h)
    d_pre = device.gaborish_inverse(d_ops)
    aq = device.initial_quant_field(d_ops)
    cfl = device.initial_cfl(d_pre)
    costs = device.ac_strategy_costs(d_pre, aq["quant_field"], aq["pixel_mask"],
                                     cfl["y_to_x"], cfl["y_to_b"])
    d_cells = device.ac_strategy_search(costs)
    fe = {"opsin": d_ops.numpy(), "pre": d_pre.numpy(), "cells": d_cells.numpy(),
          "quant_field": aq["quant_field"].numpy()}
    setup = binding.aq_setup(fe, w, h)
    s = gjxl_amd.AqSession(device, w, h)
    s.prepare(d_ops, d_pre, d_cells, aq["quant_field"], 1.0)
    device.synchronize()
    _check_prepare(s, setup)
    s.evaluate(s.buffer("adjusted_field", (s.bh, s.bw)))
    device.synchronize()
    _check_evaluation(s, fe, setup, setup["adjusted"], w, h)
    s.close()


def test_aq_dc_wide_groups(device):
    """The DC wavefront evaluates each DC group in 32-bit arithmetic and
    reruns a group in 64-bit arithmetic once a quantized value reaches 2^18
    (or a prediction error 2^22): here the left DC group's XYB is scaled by
    400 (Y DC values near 4e5), the right group keeps ordinary values."""
    w, h = 2104, 48
    print(f"aq dc wide groups {w}x{h}", flush=True)
    srgb = np.ascontiguousarray(synthetic.generate(w, h, 5, "textured")[:h, :w])
    fe = binding.frontend(srgb)
    pre = fe["pre"].copy()
    pre[:, :, :1024] *= np.float32(400)
    fe = dict(fe, pre=pre)
    setup = binding.aq_setup(fe, w, h)
    d_ops, d_pre, d_cells, d_field = _device_inputs(device, fe)
    s = gjxl_amd.AqSession(device, w, h)
    s.prepare(d_ops, d_pre, d_cells, d_field, 1.0)
    device.synchronize()
    _check_prepare(s, setup)
    s.evaluate(s.buffer("adjusted_field", (s.bh, s.bw)))
    device.synchronize()
    want = binding.aq_evaluate(fe, setup, setup["adjusted"], w, h)
    assert np.abs(want["dc_q"][:, :, :256]).max() >= 2**18
    assert np.abs(want["dc_q"][:, :, 256:]).max() < 2**18
    _check_evaluation(s, fe, setup, setup["adjusted"], w, h)
    s.close()
