"""Benchmark of the plain perceptual-analysis path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 4032x3024]
                    [--images 3]

A step is one pass of the path over each of `--images` copies of one
synthetic 8-bit sRGB image whose packed pixels are already resident in HBM
when the timed region starts.  Each image has its own device handle (its own
HIP stream, default priority), buffers and host thread, so the images run
concurrently: while one runs its DC wavefront (eight work groups), the other
fills the remaining CUs.  `value` is the images' pixels per second of the
timed region; the run fails if the images' results differ.
Prints one JSON line (rank 0).  Per-stage kernel time is measured with HIP
events on one image's stream in single-image steps after the timed region;
`roofline` describes the dominant stage of those steps.
`cpu_baseline` times the oracle (the CPU restatement) on the host on a bounded
crop of the same image.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0
FP64_PEAK_TFLOPS = 78.6
CALIBRATION_BYTES = 256 * 1024 * 1024
# HBM bytes per AQ evaluation and per Butteraugli compare, from the
# FETCH_SIZE / WRITE_SIZE counter runs summarised in this file (see
# --pmc-summary)
TRAFFIC_PROFILE = ROOT / "profiles" / "hbm_traffic_4032x3024.md"


def stage_bytes(name: str, w: int, h: int, tokens: int = 0) -> int:
    """Algorithmic HBM bytes of one launch: declared inputs read once and
    declared outputs written once."""
    n = w * h
    pn = ((w + 7) // 8 * 8) * ((h + 7) // 8 * 8)
    if name == "srgb_to_linear_and_opsin":
        return 3 * n + 12 * n + 12 * pn
    if name == "gaborish_inverse":
        return 12 * pn + 12 * pn
    if name == "initial_quant_field":
        # opsin in; pixel mask + quant field + strategy mask out
        return 12 * pn + 4 * pn + 2 * 4 * (pn // 64)
    if name == "initial_cfl":
        return 12 * pn + 2 * (pn // 4096)
    if name == "ac_strategy_costs":
        # XYB + pixel mask + quant field in, 7 cost planes out
        return 12 * pn + 4 * pn + 4 * (pn // 64) + 7 * 4 * (pn // 64)
    if name == "ac_strategy_search":
        return 7 * 4 * (pn // 64) + pn // 64
    if name == "aq_prepare":
        # opsin (matrix stats) and preprocessed XYB in, coefficients out
        return 12 * pn + 12 * pn + 12 * pn
    if name == "butteraugli_prepare":
        # reference linear RGB in; 10 psycho planes + 3 mask planes out at
        # full scale, a quarter of that at half scale
        return int((12 * n + 52 * n) * 1.25)
    if name == "aq_evaluate":
        # coefficients -> quantized (12 + 12), reconstructed XYB out (12),
        # loop filters and inverse XYB: XYB in, linear RGB out (12 + 12)
        return 36 * pn + 24 * n
    if name == "butteraugli_compare":
        # distorted RGB in (12), resident reference planes read once
        # (10 psycho + 3 mask: 52), map out (4); the half-resolution scale
        # adds a quarter
        return int((12 + 52 + 4) * n * 1.25)
    if name == "aq_reduce_update":
        return 4 * n + 4 * 4 * (pn // 64)
    if name == "ac_tokenize":
        # quantized AC (3 x 64 int32 per block), cells and raw quant in;
        # tokens (u32 value + u16 context) out
        return 3 * 4 * pn + (1 + 4) * (pn // 64) + 6 * tokens
    raise KeyError(name)


STAGE_SHAPES = ((1, 1, 1), (1, 2, 1), (2, 1, 1), (2, 2, 1), (2, 4, 2),
                (4, 2, 2), (4, 4, 2))  # covered w, h, anchor step


def candidate_flops(w: int, h: int) -> int:
    """binary64 flops of all candidate DCTs (forward + inverse, 3 channels):
    12 * W * H * (W + H) per candidate of W x H pixels."""
    bw, bh = (w + 7) // 8, (h + 7) // 8
    total = 0
    for cw, ch, step in STAGE_SHAPES:
        count = 0
        for by in range(bh):
            ly = by % 8
            tbh = min(8, bh - (by - ly))
            if ly + ch > tbh or (step == 2 and ly & 1):
                continue
            for lx in range(8):
                nx = 0
                for tx in range(0, bw, 8):
                    tbw = min(8, bw - tx)
                    if lx < tbw and lx + cw <= tbw and not (step == 2 and lx & 1):
                        nx += 1
                count += nx
        W, H = cw * 8, ch * 8
        total += count * 12 * W * H * (W + H)
    return total


def calibration_copy() -> None:
    """Three device copies of CALIBRATION_BYTES, so that counter runs can
    scale FETCH_SIZE / WRITE_SIZE by what they read for a copy of known
    size."""
    import torch
    a = torch.ones(CALIBRATION_BYTES // 4, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"calibration copy: {CALIBRATION_BYTES} bytes read and written, 3 times",
          flush=True)


def _pmc_dispatches(path: str, counter: str) -> list[tuple[int, str, float]]:
    import csv
    rows = []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    return rows


def _stage_spans(names: list[str]) -> dict[str, list[tuple[int, int]]]:
    """Dispatch index ranges (inclusive) of the AQ evaluations and the
    Butteraugli compares in one counter run: an evaluation runs from a
    FillWordsKernel (its first dispatch) to the next LoopFilterKernel; a
    compare from the dispatch after that to the next ComposeKernel."""
    spans = {"aq_evaluate": [], "butteraugli_compare": []}
    last_fill = None
    i = 0
    while i < len(names):
        if "FillWordsKernel" in names[i]:
            last_fill = i
        if "LoopFilterKernel" in names[i]:
            if last_fill is not None:
                spans["aq_evaluate"].append((last_fill, i))
            j = i + 1
            while j < len(names) and "ComposeKernel" not in names[j]:
                j += 1
            if j < len(names):
                spans["butteraugli_compare"].append((i + 1, j))
            last_fill = None
        i += 1
    return spans


def pmc_summary(fetch_csv: str, write_csv: str) -> dict:
    """HBM bytes per AQ evaluation and per Butteraugli compare from two
    rocprofv3 --pmc runs of `bench.py --calibration-copy --steps 1 --warmup 1
    --no-cpu-baseline` (FETCH_SIZE, then WRITE_SIZE; both in KiB per
    dispatch).  Each counter is scaled by what it reads for the calibration
    copy (CALIBRATION_BYTES read and written)."""
    out = {}
    for counter, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        rows = _pmc_dispatches(path, counter)
        copies = [v for _, name, v in rows if "copy" in name.lower() and v > 1e5]
        factor = CALIBRATION_BYTES / (float(np.median(copies)) * 1024.0)
        names = [name for _, name, _ in rows]
        values = [v for _, _, v in rows]
        for stage, spans in _stage_spans(names).items():
            per = [sum(values[a:b + 1]) for a, b in spans]
            entry = out.setdefault(stage, {"launches": len(per)})
            entry[counter] = round(float(np.mean(per)) * 1024.0 * factor)
        out.setdefault("calibration_factor", {})[counter] = round(factor, 4)
    for stage in ("aq_evaluate", "butteraugli_compare"):
        out[stage]["traffic_bytes_per_launch"] = (out[stage]["FETCH_SIZE"] +
                                                  out[stage]["WRITE_SIZE"])
    return out


def committed_traffic(stage: str) -> int | None:
    """The per-launch HBM traffic of `stage` recorded in TRAFFIC_PROFILE."""
    import re
    try:
        text = TRAFFIC_PROFILE.read_text()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch %s: (\d+)$" % stage, text, re.M)
    return int(m.group(1)) if m else None


def run_step(dev, path):
    """Enqueues one pass; records event slot i+1 after stage i."""
    names = []

    def mark(name):
        names.append(name)
        dev.record(len(names))

    dev.record(0)
    path.run(path.srgb_device, mark=mark)
    return names


def cpu_baseline(srgb: np.ndarray, budget_s: float = 15.0) -> dict:
    """Times the oracle's whole path on the host on a crop of the image."""
    from oracle import binding

    h, w = srgb.shape[:2]
    ch, cw = min(h, 512), min(w, 512)
    crop = np.ascontiguousarray(srgb[:ch, :cw])
    done = 0
    t0 = time.perf_counter()
    while True:
        fe = binding.frontend(crop)
        setup = binding.aq_setup(fe, cw, ch)
        loop = binding.aq_loop(fe, setup, fe["linear"], cw, ch)
        ev = loop["evaluation"]
        binding.tokenize_frame(ev["quantized"], fe["cells"], ev["raw"])
        done += 1
        if time.perf_counter() - t0 > budget_s or done >= 50:
            break
    elapsed = time.perf_counter() - t0
    return {
        "value": round(done * cw * ch / 1e6 / elapsed, 4),
        "unit": "MP/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{done} passes over a {cw}x{ch} crop of the same image, "
                  "same stages as the GPU step (frontend, AC search, AQ "
                  "setup, 3 evaluations with Butteraugli, AC tokens), single "
                  "thread",
    }


def main() -> int:
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=20)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--workload", default="4032x3024")
    parser.add_argument("--images", type=int, default=3,
                        help="images per step, each on its own stream and host "
                             "thread (the DC wavefront of one leaves most CUs "
                             "to the others)")
    parser.add_argument("--breakdown-steps", type=int, default=5,
                        help="single-image steps after the timed region for "
                             "the stage times and the roofline")
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--calibration-copy", action="store_true",
                        help="run a device copy of known size first (counter runs)")
    parser.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                        help="print the HBM traffic per AQ evaluation and per "
                             "Butteraugli compare from two counter runs and exit "
                             "(no GPU needed)")
    args = parser.parse_args()

    if args.pmc_summary:
        print(json.dumps(pmc_summary(*args.pmc_summary)))
        return 0

    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU "
              "path (images and tiles are independent; see DESIGN.md for how "
              "the path would shard).", file=sys.stderr)
        return 2

    import gjxl_amd
    from gjxl_amd import synthetic

    if args.calibration_copy:
        calibration_copy()
    spec = synthetic.WORKLOADS[args.workload]
    w, h = spec["width"], spec["height"]
    srgb = synthetic.generate(**spec)
    images = max(1, args.images)
    gjxl_amd.library()
    # one device handle (its own stream), path and input per concurrent image
    devs = [gjxl_amd.Device(0) for _ in range(images)]
    paths = [gjxl_amd.pipeline.PlainPath(d, w, h) for d in devs]
    for d, p in zip(devs, paths):
        p.srgb_device = d.array(srgb)

    def one_image(i):
        paths[i].run(paths[i].srgb_device)
        devs[i].synchronize()

    with ThreadPoolExecutor(max_workers=images) as pool:
        def step():
            list(pool.map(one_image, range(images)))

        for _ in range(args.warmup):
            step()
        for d in devs:
            d.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        for d in devs:
            d.synchronize()
        elapsed = time.perf_counter() - t0

    # stage times and the roofline from single-image steps on an otherwise
    # idle device (not part of the timed region)
    dev, path = devs[0], paths[0]
    calls: dict[str, list[float]] = {}
    for _ in range(max(1, args.breakdown_steps)):
        names = run_step(dev, path)
        dev.synchronize()
        for i, name in enumerate(names):
            calls.setdefault(name, []).append(dev.elapsed_ms(i, i + 1))
    breakdown_steps = max(1, args.breakdown_steps)

    ms_per_step = elapsed * 1e3 / args.steps
    mp = w * h / 1e6
    per_step = {k: float(np.sum(v)) / breakdown_steps for k, v in calls.items()}
    per_call = {k: float(np.mean(v)) for k, v in calls.items()}
    dominant = max(per_step, key=per_step.get)
    if dominant == "ac_strategy_costs":
        flops = candidate_flops(w, h)
        achieved = flops / (per_call[dominant] * 1e-3) / 1e12
        roofline = {"kernel": "CandidateCostKernel (7 strategy launches)",
                    "bound": "fp64", "achieved": round(achieved, 3),
                    "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                    "traffic": None}
    else:
        nbytes = stage_bytes(dominant, w, h, path.token_count)
        achieved = nbytes / (per_call[dominant] * 1e-3) / 1e9
        traffic = None
        if args.workload == "4032x3024":
            traffic = committed_traffic(dominant)
        roofline = {"kernel": dominant, "bound": "hbm",
                    "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                    "ms_per_launch": round(per_call[dominant], 4),
                    "algorithmic_bytes": nbytes, "traffic": traffic}
    result = {
        "metric": "MP/s at d1.0 e7, plain GPU path (XYB, AC strategy search, "
                  "Butteraugli AQ loop, AC tokenization), 1 GPU",
        "value": round(images * mp / (ms_per_step * 1e-3), 3),
        "unit": "MP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{w}x{h} 8-bit sRGB, seeded synthetic "
                        f"({spec['recipe']}), distance 1.0 effort 7",
            "images_per_step": images,
            "streams": images,
            "stages": list(per_step),
            "stage_ms_one_image": {k: round(v, 5) for k, v in per_step.items()},
            "ms_one_image": round(sum(per_step.values()), 4),
        },
        "roofline": roofline,
    }
    scores = [p.scores() for p in paths]
    if any(sc != scores[0] for sc in scores) or any(
            p.token_count != path.token_count for p in paths):
        print("concurrent images disagree", file=sys.stderr)
        return 1
    result["config"]["scores"] = [round(v, 6) for v in scores[0]]
    result["config"]["ac_tokens"] = path.token_count
    for p in paths:
        p.close()
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(srgb)
    for d in devs:
        d.close()
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
