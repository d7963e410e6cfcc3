"""GPU: the tile queue for batches of candidate costs (gjxl_amd_cost_queue_*).

Frames of uneven extents (one with pruning bits) go through one captured
graph of persistent tile-queue kernels; every frame's seven cost planes
equal gjxl_amd_ac_strategy_costs_pruned's bit for bit (NaN entries
included), on three replays with other batches and eager work between them,
and again after the autotuner has re-captured the graph with its grids.
The pruned per-frame costs themselves agree with the full evaluation at
every allowed candidate.
"""
import ctypes

import numpy as np
import pytest

import gjxl_amd
from gjxl_amd import synthetic

pytestmark = pytest.mark.gpu

EXTENTS = [(256, 200, 3), (520, 344, 4), (64, 64, 5), (1000, 776, 6), (136, 72, 7),
           (8, 8, 8)]


def _frame(device, w, h, seed, allow_seed=None):
    img = np.ascontiguousarray(synthetic.generate(w, h, seed=seed, recipe="textured"))
    _, opsin = device.srgb8_to_linear_and_opsin(device.array(img), w, h)
    pre = device.gaborish_inverse(opsin)
    iqf = device.initial_quant_field(opsin, 1.0)
    cfl = device.initial_cfl(pre)
    ph, pw = pre.shape[1:]
    f = {"opsin": pre, "quant_field": iqf["quant_field"], "pixel_mask": iqf["pixel_mask"],
         "y_to_x": cfl["y_to_x"], "y_to_b": cfl["y_to_b"],
         "costs": device.empty((7, ph // 8, pw // 8)), "target": 1.0}
    if allow_seed is not None:
        rng = np.random.default_rng(allow_seed)
        bits = rng.integers(0, 128, (ph // 8, pw // 8)).astype(np.uint8) | 1
        f["allow"] = device.array(bits)
    return f


def _eager(device, f):
    out = device.empty(f["costs"].shape)
    _, ph, pw = f["opsin"].shape
    allow = f.get("allow")
    gjxl_amd._check(gjxl_amd.library().gjxl_amd_ac_strategy_costs_pruned(
        device.handle, f["opsin"].ptr, pw, ph, f["quant_field"].ptr, f["pixel_mask"].ptr,
        f["y_to_x"].ptr, f["y_to_b"].ptr, f["target"],
        allow.ptr if allow is not None else None, out.ptr))
    return out.numpy().view(np.uint32)


def _same(device, frames, want):
    for i, f in enumerate(frames):
        got = f["costs"].numpy().view(np.uint32)
        assert np.array_equal(got, want[i]), i


def test_cost_queue_matches_per_frame_costs(device):
    frames = [_frame(device, w, h, s) for w, h, s in EXTENTS]
    frames.append(_frame(device, 392, 264, 9, allow_seed=1))
    frames[1]["target"] = 2.0
    want = [_eager(device, f) for f in frames]
    queue = gjxl_amd.CostQueue(device, 16)
    print(f"cost queue: {len(frames)} frames", flush=True)
    queue.run(frames)
    _same(device, frames, want)
    # another batch (a subset, reordered), eager work, then the first again
    sub = [frames[3], frames[0], frames[6]]
    for f in sub:
        f["costs"].copy_from(np.zeros(f["costs"].shape, np.float32))
    queue.run(sub)
    _same(device, sub, [want[3], want[0], want[6]])
    assert np.array_equal(_eager(device, frames[2]), want[2])
    queue.run(frames)
    _same(device, frames, want)
    grids, ms = queue.grids()
    assert all(g > 0 for g in grids) and ms == [0.0] * 7
    queue.tune(frames, repetitions=2)
    grids, ms = queue.grids()
    print(f"tuned grids {grids}, ms {[round(m, 4) for m in ms]}", flush=True)
    assert all(g > 0 for g in grids) and all(m > 0 for m in ms)
    for f in frames:
        f["costs"].copy_from(np.zeros(f["costs"].shape, np.float32))
    queue.run(frames)
    _same(device, frames, want)
    queue.close()


@pytest.mark.parametrize("w,h,seed,allow_seed", [(392, 264, 9, 1), (1000, 776, 6, 2),
                                                  (136, 72, 7, 3), (2048, 1536, 11, 4)])
def test_pruned_costs_follow_allow_bits(device, w, h, seed, allow_seed):
    """Pruned costs: an allowed candidate's cost equals the full evaluation's
    bit for bit, a fitting candidate whose stage bit is clear costs +inf, a
    position that does not fit stays NaN; all bits set gives the full planes."""
    f = _frame(device, w, h, seed, allow_seed=allow_seed)
    if allow_seed % 2 == 0:  # DCT8 pruned too (the proposal never does this)
        rng = np.random.default_rng(allow_seed + 100)
        f["allow"] = device.array(rng.integers(0, 128, f["allow"].shape).astype(np.uint8))
    print(f"pruned costs {w}x{h} seed {seed} allow {allow_seed}", flush=True)
    pruned = _eager(device, f)
    bits = f["allow"].numpy()
    f_all = dict(f)
    del f_all["allow"]
    full = _eager(device, f_all)
    f_all["allow"] = device.array(np.full(bits.shape, 0x7F, np.uint8))
    assert np.array_equal(_eager(device, f_all), full)
    fp = full.view(np.float32)
    inf = np.array(np.inf, np.float32).view(np.uint32)
    for s in range(7):
        allowed = ((bits >> s) & 1).astype(bool)
        fits = ~np.isnan(fp[s])
        assert np.array_equal(pruned[s][allowed], full[s][allowed]), s
        assert (pruned[s][fits & ~allowed] == inf).all(), s
        assert np.isnan(pruned[s].view(np.float32)[~fits]).all(), s


def test_cost_queue_refuses_bad_batches(device):
    queue = gjxl_amd.CostQueue(device, 2)
    f = _frame(device, 64, 64, 1)
    with pytest.raises(gjxl_amd.GjxlError, match="INVALID_ARGUMENT"):
        queue.run([f, f, f])
    bad = dict(f)
    bad["target"] = -1.0
    with pytest.raises(gjxl_amd.GjxlError, match="INVALID_ARGUMENT"):
        queue.run([bad])
    table = (gjxl_amd.CostFrame * 1)()
    with pytest.raises(gjxl_amd.GjxlError, match="INVALID_ARGUMENT"):
        gjxl_amd._check(gjxl_amd.library().gjxl_amd_cost_queue_run(queue.handle, table, 1))
    with pytest.raises(gjxl_amd.GjxlError, match="INVALID_ARGUMENT"):
        gjxl_amd.CostQueue(device, 257)
    queue.close()
