"""GPU parity: Butteraugli (prepared reference, repeated comparisons), its
per-transform 16-norm reduction and the complete AQ loop.

Distance maps, intermediate planes, block distances, scores and the AQ
loop's final fields are bit-exact with the oracle.  The 16-norm takes
x^(1/16) with the C library's powf (src/codec/butteraugli.cpp:188-190); the
device evaluates glibc's powf algorithm (gjxl_amd/csrc/libm_port.h), which
the CPU tests check against the host libm.
"""
import numpy as np
import pytest

import gjxl_amd
from gjxl_amd import synthetic
from oracle import binding

pytestmark = pytest.mark.gpu


def _pair(w, h, seed, noise=0.03):
    rng = np.random.default_rng(seed)
    r = rng.uniform(0.0, 1.0, (3, h, w)).astype(np.float32)
    d = np.clip(r + rng.normal(0, noise, (3, h, w)), 0, 1).astype(np.float32)
    return r, d


def _bits_equal(a, c):
    return np.array_equal(np.ascontiguousarray(a, np.float32).view(np.int32),
                          np.ascontiguousarray(c, np.float32).view(np.int32))


@pytest.mark.parametrize("w,h", [(64, 48), (37, 26), (15, 15), (14, 40), (8, 8),
                                 (5, 3), (1, 1), (7, 9), (256, 200)])
def test_butteraugli_parity(device, w, h):
    print(f"butteraugli {w}x{h}", flush=True)
    r, _ = _pair(w, h, w * 7 + h)
    b = gjxl_amd.Butteraugli(device, w, h)
    b.prepare(device.array(r))
    for noise in (0.01, 0.06):
        _, d = _pair(w, h, w * 7 + h, noise)
        out = device.empty((h, w))
        b.compare(device.array(d), out)
        score = b.score()
        want, want_score = binding.butteraugli_distance(r, d)
        got = out.numpy()
        assert _bits_equal(got, want), f"{np.abs(got - want).max()}"
        assert _bits_equal(b.buffer("map").numpy(), want)
        assert score == want_score
    b.close()


def test_butteraugli_stages(device):
    w, h = 48, 40
    r, d = _pair(w, h, 3)
    b = gjxl_amd.Butteraugli(device, w, h)
    b.prepare(device.array(r))
    b.compare(device.array(d))
    device.synchronize()
    x0 = binding.butteraugli_opsin(r)
    p0 = binding.butteraugli_frequencies(x0)
    assert _bits_equal(b.buffer("reference_psycho").numpy(), p0)
    x1 = binding.butteraugli_opsin(d)
    assert _bits_equal(b.buffer("xyb").numpy(), x1)
    assert _bits_equal(b.buffer("distorted_psycho").numpy(),
                       binding.butteraugli_frequencies(x1))
    assert _bits_equal(b.buffer("stages").numpy(), binding.butteraugli_stages(r, d))
    # the half-resolution scale
    rs = np.empty((3, (h + 1) // 2, (w + 1) // 2), np.float32)
    ds = np.empty_like(rs)
    binding._declare_butteraugli()
    binding.oracle().oracle_butteraugli_subsample(binding.ptr(r), w, h, binding.ptr(rs))
    binding.oracle().oracle_butteraugli_subsample(binding.ptr(d), w, h, binding.ptr(ds))
    assert _bits_equal(b.buffer("stages", 1).numpy(), binding.butteraugli_stages(rs, ds))
    b.close()


def test_reduce_distance_map(device):
    w, h = 200, 136
    srgb = np.ascontiguousarray(synthetic.generate(w, h, 4, "textured"))
    fe = binding.frontend(srgb)
    rng = np.random.default_rng(1)
    dmap = rng.uniform(0.0, 3.0, (h, w)).astype(np.float32)
    got = gjxl_amd.reduce_distance_map(device, device.array(dmap),
                                       device.array(fe["cells"])).numpy()
    want = binding.reduce_distance_map(dmap, fe["cells"])
    assert _bits_equal(got, want)


@pytest.mark.parametrize("w,h,seed", [(64, 64, 1), (200, 136, 4), (8, 8, 3), (72, 40, 7),
                                      (512, 512, 1)])
def test_aq_loop_parity(device, w, h, seed):
    print(f"aq loop {w}x{h} seed {seed}", flush=True)
    srgb = np.ascontiguousarray(
        synthetic.generate(max(w, 16), max(h, 16), seed, "textured")[:h, :w])
    fe = binding.frontend(srgb)
    setup = binding.aq_setup(fe, w, h)
    want = binding.aq_loop(fe, setup, fe["linear"], w, h)
    s = gjxl_amd.AqSession(device, w, h)
    s.prepare(device.array(fe["opsin"]), device.array(fe["pre"]),
              device.array(np.ascontiguousarray(fe["cells"])),
              device.array(fe["quant_field"]), 1.0)
    b = gjxl_amd.Butteraugli(device, w, h)
    b.prepare(device.array(fe["linear"]))
    scores = s.run(b, 2, 1.0)
    field = s.buffer("field", (s.bh, s.bw)).numpy()
    dist = s.buffer("block_distance", (s.bh, s.bw)).numpy()
    dmap = b.buffer("map").numpy()
    linear = s.buffer("linear", (3, h, w)).numpy()
    assert _bits_equal(field, want["field"])
    assert _bits_equal(dist, want["block_distance"])
    assert scores == want["scores"]
    assert _bits_equal(dmap, want["distance_map"])
    assert _bits_equal(linear, want["evaluation"]["linear"])
    store = binding.store_size(s.bw, s.bh)
    q = s.buffer("quantized", (3, store), np.int32).numpy()
    assert np.array_equal(q, want["evaluation"]["quantized"])
    s.close()
    b.close()
