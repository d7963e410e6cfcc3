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
    got.update(entropy_digests(path.statistics.arrays(), path.entropy_inputs))
    path.close()
    # stage order: the first differing entry names where the paths part
    order = ["linear", "opsin", "preprocessed", "initial_quant_field", "pixel_mask",
             "initial_y_to_x", "initial_y_to_b", "costs", "cells", "adjusted_field",
             "x_scale", "b_scale", "lower", "upper", "quant_dc", "fields",
             "block_distances", "score_bits", "final_field", "raw_quant", "y_to_x",
             "y_to_b", "quantized", "dc_quantized", "reconstructed_linear",
             "distance_map", *ENTROPY_KEYS, "token_count", "token_values",
             "token_contexts", "token_group_offsets"]
    differ = [k for k in order if got[k] != want[k]]
    assert not differ, f"{name}: differs from the oracle at {differ}"


ENTROPY_KEYS = ("zero_counts", "sampled_zero_counts", "raw_quant_counts", "present_mask",
                "used_order_mask", "orders", "qf_thresholds", "context_map", "num_contexts",
                "order_token_count", "order_tokens")
FRAME_KEYS = ("cells", "final_field", "raw_quant", "y_to_x", "y_to_b", "quantized",
              "dc_quantized", "used_order_mask", "orders", "qf_thresholds", "context_map",
              "num_contexts", "order_token_count", "order_tokens", "token_count",
              "token_values", "token_contexts", "token_group_offsets", "score_bits")


def entropy_digests(stats, inputs):
    """Digests of the frame statistics (None: not available) and of the
    effort-7 orders, block-context map and order tokens, as the oracle's
    table has them."""
    thresholds, cmap, nctx = inputs.context_model()
    order_ctx, order_val = gjxl_amd.order_tokens(inputs)
    out = {"used_order_mask": int(inputs.used_order_mask),
           "orders": digest(inputs.order_table()),
           "qf_thresholds": [int(v) for v in thresholds],
           "context_map": digest(cmap), "num_contexts": nctx,
           "order_token_count": int(order_val.size),
           "order_tokens": digest(np.stack([order_ctx, order_val]))}
    if stats is not None:
        out.update({"zero_counts": digest(stats["zero_counts"]),
                    "sampled_zero_counts": digest(stats["sampled"]),
                    "raw_quant_counts": digest(stats["raw_counts"]),
                    "present_mask": stats["present_mask"]})
    return out


@pytest.mark.parametrize("name", ["512x512", "4032x3024"])
def test_native_plain_path_matches_oracle_digests(device, name):
    """The C++ sequence (gjxl_amd_plain_path_run: upload from pinned host
    memory, every stage, token read-back) gives the oracle's frame and
    tokens, on two runs in a row."""
    want = GOLDEN[name]
    spec = synthetic.WORKLOADS[name]
    w, h = spec["width"], spec["height"]
    print(f"native plain path {name}", flush=True)
    srgb = synthetic.generate(**spec)
    path = gjxl_amd.NativePlainPath(device, w, h)
    for run in range(2):
        count = path.run(srgb)
        tokens = path.tokens()
        got = {
            "cells": digest(path.view("cells").numpy()),
            "final_field": digest(path.view("quant_field").numpy()),
            "raw_quant": digest(path.view("raw_quant").numpy()),
            "y_to_x": digest(path.view("y_to_x").numpy()),
            "y_to_b": digest(path.view("y_to_b").numpy()),
            "quantized": digest(path.view("quantized").numpy()),
            "dc_quantized": digest(path.view("dc_quantized").numpy()),
            "token_count": count,
            "token_values": digest(tokens["values"]),
            "token_contexts": digest(tokens["contexts"]),
            "token_group_offsets": digest(tokens["group_offsets"]),
            "score_bits": [int(np.float32(v).view(np.uint32)) for v in path.scores()],
        }
        got.update(entropy_digests(None, path.entropy_inputs()))
        differ = [k for k in FRAME_KEYS if got[k] != want[k]]
        assert not differ, f"{name} run {run}: differs from the oracle at {differ}"
    stages = [s for s, _ in path.stage_times()]
    assert stages[0] == "upload" and stages[-1] == "token_readback"
    assert "entropy_inputs" in stages
    assert stages.count("aq_evaluate") == 3
    path.close()


def test_cpp_caller_matches_oracle_digests(tmp_path):
    """A C++ program built with g++ against include/gjxl_amd.h alone
    (tests/plain_path_caller.cpp) drives the path and writes its outputs."""
    import subprocess

    from gjxl_amd import build

    name = "512x512"
    want = GOLDEN[name]
    spec = synthetic.WORKLOADS[name]
    w, h = spec["width"], spec["height"]
    srgb = synthetic.generate(**spec)
    raw = tmp_path / "input.rgb"
    raw.write_bytes(np.ascontiguousarray(srgb).tobytes())
    prefix = tmp_path / "out"
    assert build.CALLER.exists(), "run gjxl_amd.build.build_caller() first"
    print("c++ caller 512x512", flush=True)
    proc = subprocess.run([str(build.CALLER), str(w), str(h), str(raw), str(prefix), "2"],
                          capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stderr
    report = json.loads(proc.stdout.strip().splitlines()[-1])
    bw, bh = w // 8, h // 8

    def load(ext, dtype):
        return np.fromfile(f"{prefix}.{ext}", dtype=dtype)

    got = {
        "cells": digest(load("cells", np.uint8).reshape(bh, bw)),
        "final_field": digest(load("field", np.float32).reshape(bh, bw)),
        "raw_quant": digest(load("raw_quant", np.int32).reshape(bh, bw)),
        "y_to_x": digest(load("y_to_x", np.int8).reshape((bh + 7) // 8, (bw + 7) // 8)),
        "y_to_b": digest(load("y_to_b", np.int8).reshape((bh + 7) // 8, (bw + 7) // 8)),
        "quantized": digest(load("quantized", np.int32).reshape(3, -1)),
        "dc_quantized": digest(load("dc_quantized", np.int32).reshape(3, bh, bw)),
        "token_count": report["token_count"],
        "token_values": digest(load("values", np.uint32)),
        "token_contexts": digest(load("contexts", np.uint16)),
        "token_group_offsets": digest(load("group_offsets", np.uint32)),
        "score_bits": [int(np.float32(v).view(np.uint32)) for v in report["scores"]],
    }
    differ = [k for k in FRAME_KEYS if k in got and got[k] != want[k]]
    assert not differ, f"C++ caller differs from the oracle at {differ}"
