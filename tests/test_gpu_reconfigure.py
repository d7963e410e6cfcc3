"""GPU: re-targeting one plain path to other extents and options
(gjxl_amd_plain_path_reconfigure) inside the buffers of its creation extent.

A path created at 640x520 runs a 256x200 image, then the 512x512 workload,
then a 130x66 crop, then the 512x512 workload again at distance 2: every
result equals a path created for that extent and those options alone (and
the 512x512 runs equal the oracle's digests), so nothing a larger or earlier
image left in the buffers leaks into a later run.  An extent wider or taller
than the creation extent is refused and leaves the path as it was.
"""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

import gjxl_amd
from gjxl_amd import synthetic

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
from make_plain_path_hashes import digest  # noqa: E402

pytestmark = pytest.mark.gpu

GOLDEN = json.loads(
    (Path(__file__).resolve().parent / "golden" / "plain_path_hashes.json").read_text())


def _frame(path):
    t = path.tokens()
    return {"tokens": t, "scores": path.scores(),
            "cells": path.view("cells").numpy(),
            "quant_field": path.view("quant_field").numpy(),
            "quantized": path.view("quantized").numpy(),
            "dc_quantized": path.view("dc_quantized").numpy(),
            "orders": path.entropy_inputs().order_table()}


def _same(a, b):
    for k in ("values", "contexts", "group_offsets"):
        assert np.array_equal(a["tokens"][k], b["tokens"][k]), k
    assert a["scores"] == b["scores"]
    for k in ("cells", "quant_field", "quantized", "dc_quantized", "orders"):
        assert np.array_equal(a[k], b[k]), k


def _lone(device, img, distance=1.0):
    h, w = img.shape[:2]
    lone = gjxl_amd.NativePlainPath(device, w, h, distance=distance)
    lone.run(img)
    out = _frame(lone)
    lone.close()
    return out


def test_reconfigure_matches_fresh_paths(device):
    big = synthetic.generate(640, 520, seed=11)
    spec = synthetic.WORKLOADS["512x512"]
    workload = synthetic.generate(**spec)
    cases = [(np.ascontiguousarray(synthetic.generate(256, 200, seed=3)), 1.0),
             (workload, 1.0),
             (np.ascontiguousarray(synthetic.generate(136, 72, seed=5)[:66, :130]), 1.0),
             (workload, 2.0)]
    path = gjxl_amd.NativePlainPath(device, 640, 520)
    print("reconfigure: 640x520 first", flush=True)
    path.run(big)
    for img, distance in cases:
        h, w = img.shape[:2]
        print(f"reconfigure: {w}x{h} at distance {distance}", flush=True)
        path.reconfigure(w, h, distance=distance)
        count = path.run(img)
        got = _frame(path)
        _same(got, _lone(device, img, distance))
        if (w, h) == (512, 512) and distance == 1.0:
            want = GOLDEN["512x512"]
            assert count == want["token_count"]
            assert digest(got["tokens"]["values"]) == want["token_values"]
            assert digest(got["quantized"]) == want["quantized"]
    path.close()


def test_reconfigure_refuses_a_larger_extent(device):
    img = np.ascontiguousarray(synthetic.generate(256, 200, seed=3))
    path = gjxl_amd.NativePlainPath(device, 256, 200)
    path.run(img)
    first = path.tokens()
    for w, h in ((264, 200), (256, 208), (0, 10)):
        with pytest.raises(gjxl_amd.GjxlError, match="INVALID_ARGUMENT"):
            path.reconfigure(w, h)
    with pytest.raises(gjxl_amd.GjxlError, match="INVALID_ARGUMENT"):
        path.reconfigure(128, 100, distance=30.0)
    assert (path.width, path.height) == (256, 200)
    path.run(img)
    again = path.tokens()
    for k in first:
        assert np.array_equal(first[k], again[k]), k
    path.close()
