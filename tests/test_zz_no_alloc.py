"""GPU: gjxl_amd_initial_quant_field_no_alloc (static device workspace, no
allocation) gives the outputs of gjxl_amd_initial_quant_field (caller's
workspace) bit for bit, on two calls in a row, the second on a smaller
image, and rejects extents above its limit."""
import ctypes

import numpy as np
import pytest

import gjxl_amd
from gjxl_amd import synthetic

pytestmark = pytest.mark.gpu

MAX_PIXELS = 2048 * 2048


def _opsin(device, w, h, seed):
    srgb = np.ascontiguousarray(synthetic.generate(w, h, seed=seed))
    d_srgb = device.array(srgb)
    _, opsin = device.srgb8_to_linear_and_opsin(d_srgb, w, h)
    return opsin


def _no_alloc(device, opsin, target, rescale):
    _, h, w = opsin.shape
    out = {"quant_field": device.empty((h // 8, w // 8)),
           "strategy_mask": device.empty((h // 8, w // 8)),
           "pixel_mask": device.empty((h, w))}
    lib = gjxl_amd.library()
    gjxl_amd._check(lib.gjxl_amd_initial_quant_field_no_alloc(
        device.handle, opsin.ptr, w, h, ctypes.c_float(target), ctypes.c_float(rescale),
        out["quant_field"].ptr, out["strategy_mask"].ptr, out["pixel_mask"].ptr))
    return out


@pytest.fixture(scope="module")
def dev():
    d = gjxl_amd.Device(0)
    yield d
    d.close()


def test_no_alloc_matches_allocating_entry_point_twice(dev):
    # (w, h, seed, target): the second call is smaller, so stale words of
    # the first call's workspace would show up if any were read
    for w, h, seed, target in [(1024, 768, 11, 1.0), (264, 136, 12, 2.5)]:
        opsin = _opsin(dev, w, h, seed)
        got = _no_alloc(dev, opsin, target, 1.0)
        want = dev.initial_quant_field(opsin, target, 1.0)
        dev.synchronize()
        for name in ("quant_field", "strategy_mask", "pixel_mask"):
            a = got[name].numpy().view(np.uint32)
            b = want[name].numpy().view(np.uint32)
            assert np.array_equal(a, b), (w, h, name)


def test_no_alloc_rejects_extent_above_limit(dev):
    w, h = 2048, 2056  # padded, one block row above the limit
    assert w * h > MAX_PIXELS
    lib = gjxl_amd.library()
    # the size check comes before any buffer is touched: null outputs are
    # never reached, and the opsin pointer is not read
    opsin = dev.empty((3, 8, 8))
    q = dev.empty((1, 1))
    rc = lib.gjxl_amd_initial_quant_field_no_alloc(
        dev.handle, opsin.ptr, w, h, ctypes.c_float(1.0), ctypes.c_float(1.0),
        q.ptr, q.ptr, q.ptr)
    assert rc != 0
    assert b"GJXL_AMD_NO_ALLOC_MAX_PIXELS" in lib.gjxl_amd_last_error()
