"""Several device handles on one GPU, each driven from its own host thread.

Each handle has its own stream, buffers and sessions; the library keeps no
shared mutable state.  Three concurrent passes of the whole plain path over
the same image must each give the results of a pass run alone (scores, AC
token count, strategy cells and quantized coefficients, bit for bit).
"""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import gjxl_amd
from gjxl_amd import pipeline, synthetic

pytestmark = pytest.mark.gpu


def _results(path: pipeline.PlainPath) -> dict:
    return {
        "scores": path.scores(),
        "tokens": path.token_count,
        "cells": path.cells.numpy().copy(),
        "quantized": path.quantized.numpy().copy(),
    }


def test_concurrent_images_match_single():
    w, h = 512, 384
    srgb = synthetic.generate(w, h, seed=5, recipe="textured")
    gjxl_amd.library()
    # one image alone
    dev = gjxl_amd.Device(0)
    single = pipeline.PlainPath(dev, w, h)
    single.run(dev.array(srgb))
    dev.synchronize()
    expected = _results(single)
    single.close()
    dev.close()
    # three at once, one host thread and device handle each
    devs = [gjxl_amd.Device(0) for _ in range(3)]
    paths = [pipeline.PlainPath(d, w, h) for d in devs]
    srcs = [d.array(srgb) for d in devs]

    def run(i):
        for _ in range(2):
            paths[i].run(srcs[i])
        devs[i].synchronize()
        return _results(paths[i])

    try:
        with ThreadPoolExecutor(max_workers=3) as pool:
            got = list(pool.map(run, range(3)))
    finally:
        for p in paths:
            p.close()
        for d in devs:
            d.close()
    for g in got:
        assert g["scores"] == expected["scores"]
        assert g["tokens"] == expected["tokens"]
        np.testing.assert_array_equal(g["cells"], expected["cells"])
        np.testing.assert_array_equal(g["quantized"], expected["quantized"])
