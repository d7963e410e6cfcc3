"""Writes tests/golden/plain_path_hashes.json: digests of every stage output
of the oracle's plain path (XYB, Gaborish inverse, initial quant field,
initial CfL, candidate costs, strategy grid, AQ setup, the three
Butteraugli-driven evaluations of d1.0 e7, the frame statistics, the
effort-7 coefficient orders and block-context map, and the AC tokens of the
last evaluation with them) on the two seeded workloads, 512x512 ("basic") and 4032x3024
("textured").

The oracle is the plain-C restatement in oracle/ (each function pinned
against the reference's compiled CPU code or its known answers by the CPU
tests).  The 4032x3024 run takes a few minutes of one CPU core.  The GPU test
tests/test_gpu_plain_path.py runs the device path on the same inputs and
compares digests.

Usage: python tests/golden/make_plain_path_hashes.py [512x512] [4032x3024]
"""
from __future__ import annotations

import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from gjxl_amd import synthetic  # noqa: E402
from oracle import binding  # noqa: E402

OUT = Path(__file__).resolve().parent / "plain_path_hashes.json"


def digest(array: np.ndarray) -> str:
    """SHA-256 (first 16 hex digits) of the array's bytes in C order; every
    NaN of a binary32 array is first made the quiet NaN 0x7fc00000."""
    a = np.ascontiguousarray(array)
    if a.dtype == np.float32:
        a = np.where(np.isnan(a), np.float32(np.nan), a).astype(np.float32)
        a = np.where(np.isnan(a), np.uint32(0x7FC00000), a.view(np.uint32)).astype(np.uint32)
    return hashlib.sha256(a.tobytes()).hexdigest()[:16]


def plain_path(name: str) -> dict:
    spec = synthetic.WORKLOADS[name]
    w, h = spec["width"], spec["height"]
    srgb = synthetic.generate(**spec)
    t0 = time.time()
    fe = binding.frontend(srgb)
    setup = binding.aq_setup(fe, w, h)
    loop = binding.aq_loop(fe, setup, fe["linear"], w, h)
    ev = loop["evaluation"]
    tokens = binding.effort7_tokens(ev["quantized"], fe["cells"], ev["raw"])
    order_ctx, order_val = binding.order_tokens(tokens["used_order_mask"], tokens["orders"])
    stats = tokens["statistics"]
    seconds = time.time() - t0
    out = {
        "width": w, "height": h,
        "oracle_seconds": round(seconds, 1),
        "srgb": digest(srgb),
        "linear": digest(fe["linear"]),
        "opsin": digest(fe["opsin"]),
        "preprocessed": digest(fe["pre"]),
        "initial_quant_field": digest(fe["quant_field"]),
        "pixel_mask": digest(fe["pixel_mask"]),
        "initial_y_to_x": digest(fe["init_y_to_x"]),
        "initial_y_to_b": digest(fe["init_y_to_b"]),
        "costs": digest(fe["costs"]),
        "cells": digest(np.ascontiguousarray(fe["cells"])),
        "adjusted_field": digest(setup["adjusted"]),
        "x_scale": setup["x_scale"], "b_scale": setup["b_scale"],
        "lower": float(np.float32(setup["lower"])),
        "upper": float(np.float32(setup["upper"])),
        "quant_dc": float(np.float32(setup["quant_dc"])),
        "fields": [digest(f) for f in loop["fields"]],
        "block_distances": [digest(d) for d in loop["block_distances"]],
        "score_bits": [int(np.float32(s).view(np.uint32)) for s in loop["scores"]],
        "final_field": digest(loop["field"]),
        "raw_quant": digest(ev["raw"]),
        "y_to_x": digest(ev["y_to_x"]),
        "y_to_b": digest(ev["y_to_b"]),
        "quantized": digest(ev["quantized"]),
        "dc_quantized": digest(ev["dc_q"]),
        "reconstructed_linear": digest(ev["linear"]),
        "distance_map": digest(loop["distance_map"]),
        "token_count": int(tokens["group_offsets"][-1]),
        "token_values": digest(tokens["values"]),
        "token_contexts": digest(tokens["contexts"]),
        "token_group_offsets": digest(tokens["group_offsets"]),
        "zero_counts": digest(stats["zero_counts"]),
        "sampled_zero_counts": digest(stats["sampled"]),
        "raw_quant_counts": digest(stats["raw_counts"]),
        "present_mask": stats["present_mask"],
        "used_order_mask": tokens["used_order_mask"],
        "orders": digest(tokens["orders"]),
        "qf_thresholds": [int(v) for v in tokens["thresholds"]],
        "context_map": digest(tokens["context_map"]),
        "num_contexts": tokens["num_contexts"],
        "order_token_count": int(order_val.size),
        "order_tokens": digest(np.stack([order_ctx, order_val])),
        "strategy_counts": np.bincount(
            (np.ascontiguousarray(fe["cells"])[fe["cells"] & 1 == 1] >> 1).astype(np.int64),
            minlength=16).tolist(),
    }
    return out


def main(names: list[str]) -> None:
    table = json.loads(OUT.read_text()) if OUT.exists() else {}
    for name in names or list(synthetic.WORKLOADS):
        print(f"oracle plain path {name} ...", flush=True)
        table[name] = plain_path(name)
        print(f"  {table[name]['oracle_seconds']} s, scores "
              f"{[np.uint32(b).view(np.float32) for b in table[name]['score_bits']]}",
              flush=True)
        OUT.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    main(sys.argv[1:])
