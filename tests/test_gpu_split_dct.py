"""GPU: the split-precision forward DCT on MFMA (the chosen precision form).

* Blocks of every side pair in {8, 16, 32, 64} against a binary64 DCT
  (numpy, the basis of src/codec/dct.cpp:135-182: sqrt(2)/n, 1/sqrt(2) on
  the first row), in the storage order of src/core/ac_strategy.h:95-108.
  The error bound is stated per size: the split products keep about 22 bits
  of each operand and accumulate in binary32.  A wrong operand layout or a
  dropped term fails it by orders of magnitude.
* The seven searched strategies against the oracle's DCT
  (oracle_forward_dct, bit-exact with the reference's compiled dct.cpp).
* A frame: the split coefficients of every transform of the 512x512
  workload against the plain path's, and the share of quantized
  coefficients that change when the evaluation codes them (measured, with
  a loose bound).
"""
import ctypes

import numpy as np
import pytest

import gjxl_amd
from gjxl_amd import synthetic
from oracle import binding

pytestmark = pytest.mark.gpu

SIDES = (8, 16, 32, 64)


def basis(n):
    f = np.arange(n)[:, None]
    s = np.arange(n)[None, :]
    b = np.sqrt(2.0) / n * np.cos((s + 0.5) * f * np.pi / n)
    b[0] /= np.sqrt(2.0)
    return b


def dct64(blocks):
    """Binary64 2-D DCT of (count, H, W) blocks, storage order."""
    _, h, w = blocks.shape
    c = np.einsum("vy,kyx,ux->kvu", basis(h), blocks.astype(np.float64), basis(w))
    if h < w:
        return c.reshape(len(blocks), -1)
    return np.transpose(c, (0, 2, 1)).reshape(len(blocks), -1)


def split_dct(device, blocks):
    count, h, w = blocks.shape
    d_in = device.array(np.ascontiguousarray(blocks, np.float32))
    d_out = device.empty((count, h * w))
    gjxl_amd._check(gjxl_amd.library().gjxl_amd_split_dct_blocks(
        device.handle, w, h, d_in.ptr, count, d_out.ptr))
    return d_out.numpy()


@pytest.mark.parametrize("w", SIDES)
@pytest.mark.parametrize("h", SIDES)
def test_blocks_against_binary64(device, w, h):
    rng = np.random.default_rng(w * 100 + h)
    count = 261  # a partial last work group
    # XYB-like magnitudes: smooth ramps plus texture, some blocks tiny
    ramp = np.linspace(0, 1, w)[None, None, :] + np.linspace(0, 1, h)[None, :, None]
    scale = rng.choice([1.0, 0.05, 1e-3], size=(count, 1, 1))
    blocks = (scale * (0.3 * ramp + rng.normal(0, 0.2, size=(count, h, w)))).astype(np.float32)
    blocks[0] = 0.0  # all zero: exact
    got = split_dct(device, blocks)
    want = dct64(blocks)
    peak = np.abs(blocks).reshape(count, -1).max(axis=1)[:, None]
    err = np.abs(got - want)
    rel = (err / np.maximum(peak, 1e-30)).max()
    print(f"split DCT {w}x{h}: max abs {err.max():.3g}, max / block peak {rel:.3g}", flush=True)
    assert np.all(got[0] == 0.0)
    assert rel < 4e-6


@pytest.mark.parametrize("strategy", [0, 4, 5, 6, 7, 10, 11])
def test_strategies_against_the_oracle_dct(device, strategy):
    cw, ch = {0: (1, 1), 4: (2, 2), 5: (4, 4), 6: (1, 2), 7: (2, 1), 10: (2, 4),
              11: (4, 2)}[strategy]
    w, h = 8 * cw, 8 * ch
    rng = np.random.default_rng(strategy)
    blocks = rng.uniform(-0.2, 0.8, size=(64, h, w)).astype(np.float32)
    got = split_dct(device, blocks)
    lib = binding.oracle()
    want = np.zeros_like(got)
    for k in range(64):
        px = np.ascontiguousarray(blocks[k])
        out = np.zeros(w * h, np.float32)
        assert lib.oracle_forward_dct(strategy, binding.ptr(px), binding.ptr(out)) == 0
        want[k] = out
    assert np.abs(got - want).max() < 4e-6 * 0.8


def test_frame_coefficients_and_quantized_mismatch(device):
    spec = synthetic.WORKLOADS["512x512"]
    w, h = spec["width"], spec["height"]
    srgb = synthetic.generate(**spec)
    path = gjxl_amd.pipeline.PlainPath(device, w, h)
    path.run(device.array(srgb))
    s = path.session
    store = ((s.bw + 31) // 32) * ((s.bh + 31) // 32) * 65536
    coefs = s.buffer("coefficients", (3, store))
    plain = coefs.numpy()
    quantized_plain = path.quantized.numpy()
    field = path.field.numpy()
    gjxl_amd._check(gjxl_amd.library().gjxl_amd_aq_split_forward_dct(s.handle))
    device.synchronize()
    split = coefs.numpy()
    used = plain != 0
    err = np.abs(split - plain)
    print(f"frame: max abs {err.max():.3g} over {used.sum()} coefficients", flush=True)
    assert err.max() < 1e-5
    # code the split coefficients with the final field of the plain loop
    s.evaluate(path.field)
    device.synchronize()
    quantized_split = path.quantized.numpy()
    assert np.array_equal(path.field.numpy(), field)
    changed = np.count_nonzero(quantized_split != quantized_plain)
    rate = changed / quantized_plain.size
    print(f"frame: quantized mismatch {changed} of {quantized_plain.size} ({rate:.2e})",
          flush=True)
    assert rate < 1e-3
    path.close()
