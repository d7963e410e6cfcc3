"""GPU parity of the whole plain path on both seeded workloads.

The device path (gjxl_amd.pipeline.PlainPath: XYB, Gaborish inverse, initial
quant field, initial CfL, candidate costs, strategy search, AQ preparation,
three Butteraugli-driven evaluations with two field updates, AC tokens) runs
on 512x512 ("basic", seed 1) and 4032x3024 ("textured", seed 2).  Every
stage output is compared, by SHA-256 digest, with the oracle's output on the
same image: tests/golden/plain_path_hashes.json, written in the CPU
container by tests/golden/make_plain_path_hashes.py.  Each evaluation's
field, block distances and score are compared, so a divergence names the
first stage and evaluation where it appears.
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


@pytest.mark.parametrize("name", ["512x512", "4032x3024"])
def test_plain_path_matches_oracle_digests(device, name):
    want = GOLDEN[name]
    spec = synthetic.WORKLOADS[name]
    w, h = spec["width"], spec["height"]
    print(f"plain path {name}", flush=True)
    srgb = synthetic.generate(**spec)
    assert digest(srgb) == want["srgb"], "the seeded generator changed"
    path = gjxl_amd.pipeline.PlainPath(device, w, h)
    fields, distances = [], []

    def mark(stage):
        if stage == "butteraugli_compare":
            fields.append(digest(path.field.numpy()))
        elif stage == "aq_reduce_update":
            distances.append(digest(path.block_distance.numpy()))

    path.run(device.array(srgb), mark=mark)
    device.synchronize()
    s = path.session
    tiles = ((s.bh + 7) // 8, (s.bw + 7) // 8)
    store = ((s.bw + 31) // 32) * ((s.bh + 31) // 32) * 65536
    tokens = path.tokenizer.tokens(path.token_count)
    pol = s.policy()
    got = {
        "linear": digest(path.linear.numpy()),
        "opsin": digest(path.opsin.numpy()),
        "preprocessed": digest(path.preprocessed.numpy()),
        "initial_quant_field": digest(path.aq["quant_field"].numpy()),
        "pixel_mask": digest(path.aq["pixel_mask"].numpy()),
        "initial_y_to_x": digest(path.cfl["y_to_x"].numpy()),
        "initial_y_to_b": digest(path.cfl["y_to_b"].numpy()),
        "costs": digest(path.costs.numpy()),
        "cells": digest(path.cells.numpy()),
        "adjusted_field": digest(path.adjusted.numpy()),
        "x_scale": pol["x_scale"], "b_scale": pol["b_scale"],
        "lower": float(np.float32(pol["lower"])),
        "upper": float(np.float32(pol["upper"])),
        "quant_dc": float(np.float32(pol["quant_dc"])),
        "fields": fields,
        "block_distances": distances,
        "score_bits": [int(np.float32(v).view(np.uint32)) for v in path.scores()],
        "final_field": digest(path.field.numpy()),
        "raw_quant": digest(path.raw_quant.numpy()),
        "y_to_x": digest(s.buffer("y_to_x", tiles, np.int8).numpy()),
        "y_to_b": digest(s.buffer("y_to_b", tiles, np.int8).numpy()),
        "quantized": digest(s.buffer("quantized", (3, store), np.int32).numpy()),
        "dc_quantized": digest(s.buffer("dc_quantized", (3, s.bh, s.bw), np.int32).numpy()),
        "reconstructed_linear": digest(path.reconstructed_linear.numpy()),
        "distance_map": digest(path.distance_map.numpy()),
        "token_count": int(path.token_count),
        "token_values": digest(tokens["values"]),
        "token_contexts": digest(tokens["contexts"]),
        "token_group_offsets": digest(tokens["group_offsets"]),
    }
    path.close()
    # stage order: the first differing entry names where the paths part
    order = ["linear", "opsin", "preprocessed", "initial_quant_field", "pixel_mask",
             "initial_y_to_x", "initial_y_to_b", "costs", "cells", "adjusted_field",
             "x_scale", "b_scale", "lower", "upper", "quant_dc", "fields",
             "block_distances", "score_bits", "final_field", "raw_quant", "y_to_x",
             "y_to_b", "quantized", "dc_quantized", "reconstructed_linear",
             "distance_map", "token_count", "token_values", "token_contexts",
             "token_group_offsets"]
    differ = [k for k in order if got[k] != want[k]]
    assert not differ, f"{name}: differs from the oracle at {differ}"
