"""Benchmark of the plain perceptual-analysis path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 4032x3024]

A step is one call of gjxl_amd_plain_path_run (the C++ host sequence) on one
synthetic 8-bit sRGB image at distance 1.0, effort 7: the packed pixels are
resident in HBM when the timed region starts, and the step ends when the
token values, contexts and group offsets are in pinned host memory.  `value`
is pixels per second of the timed region (one image per step, one stream).
Secondary figures in `config`: the same step with the upload from pinned
host memory, and several images at once on their own device handles and
host threads (`concurrent_images`).  Every run's tokens and scores must
equal the first run's.

`roofline`: the kernel family with the most device time per image is timed
per launch with HIP events on the device stream over the whole timed region
(gjxl_amd_device_probe); `achieved` is its algorithmic bytes (or binary64
flops) per launch over the mean launch time.  `cpu_baseline` times the
oracle (the CPU restatement) on the host on a bounded crop of the same
image.
"""
from __future__ import annotations

import argparse
import hashlib
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


# kernel families whose HBM bytes per launch the counter summary reports
TRAFFIC_KERNELS = ("CandidateCostKernel", "DcQuantKernel", "DifferencesKernel")


def pmc_summary(fetch_csv: str, write_csv: str) -> dict:
    """HBM bytes per AQ evaluation, per Butteraugli compare and per launch of
    each TRAFFIC_KERNELS family from two rocprofv3 --pmc runs of `bench.py
    --calibration-copy --steps 1 --warmup 1 --no-cpu-baseline` (FETCH_SIZE,
    then WRITE_SIZE; both in KiB per dispatch).  Each counter is scaled by
    what it reads for the calibration copy (CALIBRATION_BYTES read and
    written)."""
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
        for family in TRAFFIC_KERNELS:
            per = [v for name, v in zip(names, values) if family in name]
            entry = out.setdefault(family, {"launches": len(per)})
            entry[counter] = round(float(np.mean(per)) * 1024.0 * factor) if per else 0
        out.setdefault("calibration_factor", {})[counter] = round(factor, 4)
    for stage in ("aq_evaluate", "butteraugli_compare") + TRAFFIC_KERNELS:
        out[stage]["traffic_bytes_per_launch"] = (out[stage]["FETCH_SIZE"] +
                                                  out[stage]["WRITE_SIZE"])
    return out


def committed_traffic(stage: str) -> int | None:
    """The per-launch HBM traffic of `stage` (a stage or a kernel family)
    recorded in TRAFFIC_PROFILE."""
    import re
    try:
        text = TRAFFIC_PROFILE.read_text()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch %s: (\d+)$" % stage, text, re.M)
    return int(m.group(1)) if m else None


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




# Speed of the oracle against the reference's own CPU encoder, both timed in
# the build container (8 x Xeon @ 2.70 GHz, one busy core each) on 512x512
# synthetic images at d1.0 e7: the reference encoder (`gjxl_encode --backend
# cpu`, whole encode including entropy coding) took 2.91 s per image, the
# oracle's plain path 4.45 s (median of 3 after one warm-up, seed-1 "basic"
# image).  The ratio carries the GPU host's oracle figure over to an estimate
# of the reference encoder on the same host.
REFERENCE_CPU_SECONDS_512 = 2.91
ORACLE_CPU_SECONDS_512 = 4.45

# Algorithmic bytes (or binary64 flops) of one launch of each probed kernel
# family at image extent w x h.
def probe_units(kernel: str, w: int, h: int) -> tuple[str, float]:
    pn = ((w + 7) // 8 * 8) * ((h + 7) // 8 * 8)
    blocks = pn // 64
    if kernel == "DcQuantKernel":
        # three DC planes in (float), three quantized planes out (int32)
        return "bytes", 24.0 * blocks
    if kernel == "DifferencesKernel":
        # per scale (full and half resolution, one launch each; the mean
        # launch covers 1.25 / 2 of the full-resolution pixels): 10 + 10
        # psycho planes, both blurred activities, MaskY and MaskDcY in, the
        # two final planes out
        return "bytes", 104.0 * w * h * 1.25 / 2
    if kernel == "CandidateCostKernel":
        # binary64 DCT flops of all candidates over the 7 launches
        return "flops", candidate_flops(w, h) / 7.0
    raise KeyError(kernel)


def _digest(path) -> str:
    import hashlib
    n = int(path.frame.token_count)
    h = hashlib.sha256()
    h.update(path.values.array[:n].tobytes())
    h.update(path.contexts.array[:n].tobytes())
    h.update(path.group_offsets.array.tobytes())
    h.update(np.asarray(path.scores(), np.float32).tobytes())
    return h.hexdigest()[:16]


def _concurrent(gjxl_amd, srgb, w, h, images: int, steps: int) -> dict:
    """`images` device handles (one stream each) and host threads running the
    path on copies of the image at once; returns the aggregate rate and the
    digests of every handle's last run."""
    devs = [gjxl_amd.Device(0) for _ in range(images)]
    paths = [gjxl_amd.NativePlainPath(d, w, h) for d in devs]
    inputs = [d.array(srgb) for d in devs]

    def one(i):
        paths[i].run(resident=inputs[i])

    with ThreadPoolExecutor(max_workers=images) as pool:
        list(pool.map(one, range(images)))
        t0 = time.perf_counter()
        for _ in range(steps):
            list(pool.map(one, range(images)))
        elapsed = time.perf_counter() - t0
    digests = [_digest(p) for p in paths]
    for p in paths:
        p.close()
    for d in devs:
        d.close()
    return {"images_per_step": images, "streams": images, "steps": steps,
            "ms_per_step": round(elapsed * 1e3 / steps, 4),
            "value": round(images * steps * w * h / 1e6 / elapsed, 3),
            "unit": "MP/s", "digests": digests}


def batch_workload(count: int, seed: int = 4, largest: int = 4096) -> list[np.ndarray]:
    """`count` images whose width and height are drawn uniformly from 256 to
    `largest` in steps of 8, cropped at random offsets from one seeded
    4096x4096 synthetic image ("textured", seed 2)."""
    from gjxl_amd import synthetic
    rng = np.random.default_rng(seed)
    source = synthetic.generate(4096, 4096, seed=2, recipe="textured")
    out = []
    for _ in range(count):
        w, h = (int(v) for v in rng.integers(32, largest // 8 + 1, size=2) * 8)
        x0, y0 = int(rng.integers(0, 4096 - w + 1)), int(rng.integers(0, 4096 - h + 1))
        out.append(np.ascontiguousarray(source[y0:y0 + h, x0:x0 + w]))
    return out


def _batch(gjxl_amd, images: list[np.ndarray], lanes: int,
           calls: int = 2, tile_queue: bool = False) -> tuple[list[float], list[int]]:
    """Seconds of each of `calls` encode calls over the same images on one
    batch driver with `lanes` workers (the first call creates each worker's
    plain path at the bounding extent, later calls reuse it), and every
    image's token count.  tile_queue: group mode, tuned on the first call's
    first group."""
    batch = gjxl_amd.Batch(0, lanes)
    if tile_queue:
        batch.set_tile_queue(True, tune_repetitions=2)
    seconds, counts = [], None
    for _ in range(calls):
        t0 = time.perf_counter()
        results = batch.encode(images)
        seconds.append(time.perf_counter() - t0)
        got = []
        for i in range(len(images)):
            if results[i].status != 0:
                raise RuntimeError(f"batch image {i}: {results[i].message.decode()}")
            got.append(int(results[i].token_count))
        batch.free_results(results, len(images))
        if counts is not None and got != counts:
            raise RuntimeError("batch calls disagree")
        counts = got
    batch.close()
    return seconds, counts


MFMA_FP16_PEAK_TFLOPS = 2500.0


def cost_queue_figures(gjxl_amd, dev, images: list[np.ndarray], reps: int = 3) -> dict:
    """The candidate costs of a batch of uneven frames: per-frame launches
    (gjxl_amd_ac_strategy_costs, one frame after another on one stream)
    against one launch of the tuned tile-queue graph; device time from HIP
    events, the costs compared bit for bit."""
    lib = gjxl_amd.library()
    frames, pixels = [], 0
    for img in images:
        h, w = img.shape[:2]
        srgb = dev.array(img)
        linear, opsin = dev.srgb8_to_linear_and_opsin(srgb, w, h)
        pre = dev.gaborish_inverse(opsin)
        iqf = dev.initial_quant_field(opsin, 1.0)
        cfl = dev.initial_cfl(pre)
        ph, pw = pre.shape[1:]
        frames.append({"opsin": pre, "quant_field": iqf["quant_field"],
                       "pixel_mask": iqf["pixel_mask"], "y_to_x": cfl["y_to_x"],
                       "y_to_b": cfl["y_to_b"], "costs": dev.empty((7, ph // 8, pw // 8)),
                       "target": 1.0})
        iqf["strategy_mask"].free()
        for a in (srgb, linear, opsin):
            a.free()
        pixels += w * h
    dev.synchronize()

    def eager():
        for f in frames:
            _, ph, pw = f["opsin"].shape
            gjxl_amd._check(lib.gjxl_amd_ac_strategy_costs(
                dev.handle, f["opsin"].ptr, pw, ph, f["quant_field"].ptr, f["pixel_mask"].ptr,
                f["y_to_x"].ptr, f["y_to_b"].ptr, 1.0, f["costs"].ptr))

    def timed(fn):
        fn()
        dev.synchronize()
        dev.record(0)
        for _ in range(reps):
            fn()
        dev.record(1)
        dev.synchronize()
        return dev.elapsed_ms(0, 1) / reps

    eager_ms = timed(eager)
    want = [hashlib.sha256(f["costs"].numpy().tobytes()).hexdigest() for f in frames]
    queue = gjxl_amd.CostQueue(dev, len(frames))
    queue.tune(frames, repetitions=2)
    queue_ms = timed(lambda: queue.run(frames))
    same = all(hashlib.sha256(f["costs"].numpy().tobytes()).hexdigest() == want[i]
               for i, f in enumerate(frames))
    grids, tuned = queue.grids()
    queue.close()
    for f in frames:
        for a in f.values():
            if hasattr(a, "free"):
                a.free()
    return {"frames": len(frames), "megapixels": round(pixels / 1e6, 2),
            "per_frame_launches_ms": round(eager_ms, 3),
            "tile_queue_graph_ms": round(queue_ms, 3),
            "speedup": round(eager_ms / queue_ms, 3), "bit_exact": same,
            "tuned_grids": grids, "tuned_stage_ms": [round(m, 4) for m in tuned]}


def _dct_basis(n: int) -> np.ndarray:
    f = np.arange(n)[:, None]
    b = np.sqrt(2.0) / n * np.cos((np.arange(n)[None, :] + 0.5) * f * np.pi / n)
    b[0] /= np.sqrt(2.0)
    return b


def split_dct_errors(gjxl_amd, dev, count: int = 128) -> dict:
    """Maximum error of the split-precision DCT against a binary64 DCT, per
    side pair, on seeded blocks of XYB-like magnitudes (relative to each
    block's largest pixel)."""
    lib = gjxl_amd.library()
    out = {}
    for w in (8, 16, 32, 64):
        for h in (8, 16, 32, 64):
            rng = np.random.default_rng(w * 100 + h)
            ramp = np.linspace(0, 1, w)[None, None, :] + np.linspace(0, 1, h)[None, :, None]
            scale = rng.choice([1.0, 0.05, 1e-3], size=(count, 1, 1))
            blocks = (scale * (0.3 * ramp + rng.normal(0, 0.2, size=(count, h, w)))
                      ).astype(np.float32)
            d_in = dev.array(blocks)
            d_out = dev.empty((count, w * h))
            gjxl_amd._check(lib.gjxl_amd_split_dct_blocks(dev.handle, w, h, d_in.ptr, count,
                                                          d_out.ptr))
            got = d_out.numpy()
            c = np.einsum("vy,kyx,ux->kvu", _dct_basis(h), blocks.astype(np.float64),
                          _dct_basis(w))
            want = c.reshape(count, -1) if h < w else np.transpose(c, (0, 2, 1)).reshape(count, -1)
            peak = np.abs(blocks).reshape(count, -1).max(axis=1)[:, None]
            out[f"{w}x{h}"] = float((np.abs(got - want) / peak).max())
    return out


def split_dct_figures(gjxl_amd, dev, srgb: np.ndarray, distances) -> dict:
    """The split-precision forward DCT of every transform of the frame (the
    chosen precision form): its time per frame from HIP events on the device
    stream, HBM bytes and MFMA flops against their peaks, its error against
    the plain path's binary64-basis coefficients, and the share of quantized
    coefficients that change when the final evaluation codes them, per
    distance."""
    h, w = srgb.shape[:2]
    lib = gjxl_amd.library()
    out = {"max_error_per_block_peak": split_dct_errors(gjxl_amd, dev), "per_distance": []}
    for d in distances:
        path = gjxl_amd.pipeline.PlainPath(dev, w, h, target=d)
        path.run(dev.array(srgb))
        s = path.session
        store = ((s.bw + 31) // 32) * ((s.bh + 31) // 32) * 65536
        coefs = s.buffer("coefficients", (3, store))
        plain = coefs.numpy()
        q_plain = path.quantized.numpy()
        reps = 10
        dev.synchronize()
        dev.record(0)
        for _ in range(reps):
            gjxl_amd._check(lib.gjxl_amd_aq_split_forward_dct(s.handle))
        dev.record(1)
        ms = dev.elapsed_ms(0, 1) / reps
        split = coefs.numpy()
        s.evaluate(path.field)
        q_split = path.quantized.numpy()
        changed = int(np.count_nonzero(q_split != q_plain))
        row = {"distance": d, "ms": round(ms, 4),
               "max_abs_error_vs_plain": float(np.abs(split - plain).max()),
               "quantized_mismatch": changed,
               "quantized_mismatch_rate": changed / q_plain.size}
        if d == 1.0:
            cells = path.cells.numpy()
            anchors = cells[(cells & 1) == 1] >> 1
            sides = {0: (8, 8), 4: (16, 16), 5: (32, 32), 6: (8, 16), 7: (16, 8),
                     10: (16, 32), 11: (32, 16)}
            flops = 0
            for st, n in zip(*np.unique(anchors, return_counts=True)):
                tw, th = sides[int(st)]
                pw_, ph_ = max(tw, 16), max(th, 16)
                # 3 channels x 3 split products x 2 flops per multiply-add
                flops += int(n) * 3 * 3 * 2 * (ph_ * pw_ * pw_ + ph_ * ph_ * pw_)
            nbytes = 2 * 3 * 4 * s.pw * s.ph  # pixels in, coefficients out
            out.update({"frame": f"{w}x{h}", "mfma_flops": flops, "algorithmic_bytes": nbytes,
                        "achieved_gbps": round(nbytes / (ms * 1e-3) / 1e9, 1),
                        "hbm_frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
                        "achieved_mfma_tflops": round(flops / (ms * 1e-3) / 1e12, 2),
                        "mfma_frac": round(flops / (ms * 1e-3) / 1e12 / MFMA_FP16_PEAK_TFLOPS,
                                           4)})
        out["per_distance"].append(row)
        path.close()
    return out


def _token_bits(tokens: dict) -> float:
    """Entropy of the token stream given its contexts (bits): the size of an
    ideal coder with static per-context statistics, a proxy for the ANS
    stage, which this library does not run."""
    v = tokens["values"].astype(np.uint64)
    c = tokens["contexts"].astype(np.uint64)
    _, cnt = np.unique((c << np.uint64(32)) | v, return_counts=True)
    _, ccnt = np.unique(c, return_counts=True)
    return float(-(cnt * np.log2(cnt)).sum() + (ccnt * np.log2(ccnt)).sum())


def _rate_quality(path) -> tuple[float, float]:
    """(3-norm of the last evaluation's block distances, token bits)."""
    d = path.view("block_distance").numpy().astype(np.float64)
    return float(np.mean(d ** 3) ** (1.0 / 3.0)), _token_bits(path.tokens())


def learned_figures(gjxl_amd, dev, srgb: np.ndarray, steps: int) -> dict:
    """The learned proposal (chosen form) against the plain path on the
    bench image: MP/s with the input resident in HBM and the tokens read
    back, the rate excess at equal quality against the plain path's own
    rate-quality curve (distances 0.9-1.2; quality = 3-norm of the block
    distances, rate = token entropy), with the trained network and with none
    (zero parameters), grid agreement and the share of blocks left to the
    plain search; for the trained network also the unjudged runs (last pass
    quantized without a comparison), whose quality is that of the same run
    comparing its last pass (same tokens, checked)."""
    from gjxl_amd import proposal as prop

    h, w = srgb.shape[:2]
    mp = w * h / 1e6
    path = gjxl_amd.NativePlainPath(dev, w, h)
    resident = dev.array(srgb)

    def timed():
        path.run(resident=resident)
        dev.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            path.run(resident=resident)
        return (time.perf_counter() - t0) / steps

    curve = []
    for dist in (0.9, 1.0, 1.1, 1.2):
        path.reconfigure(w, h, distance=dist)
        path.run(resident=resident)
        curve.append(_rate_quality(path))
    path.reconfigure(w, h, distance=1.0)
    plain_s = timed()
    plain_cells = path.view("cells").numpy()
    qs = np.array([c[0] for c in curve])
    lb = np.log(np.array([c[1] for c in curve]))

    def excess(q, b):
        return round(float(b / np.exp(np.interp(q, qs, lb)) - 1.0), 5)

    out = {"parameters": "gjxl_amd/proposal_parameters.npz",
           "training": {k: v for k, v in prop.load_metadata().items() if k != "history"},
           "plain": {"value": round(mp / plain_s, 3), "ms_per_image": round(plain_s * 1e3, 3),
                     "quality_3norm": round(curve[1][0], 5), "bits": round(curve[1][1])},
           "runs": []}
    trained = prop.load_parameters()
    for name, params in (("trained", trained), ("none", np.zeros_like(trained))):
        net = prop.ProposalNet(dev, path.bw * path.bh, params)
        runs = [{"evaluations": 1, "judge_ratio": 100.0},
                {"evaluations": 2, "judge_ratio": 100.0}, {}]
        if name == "trained":
            runs += [{"evaluations": 1, "judge_ratio": 100.0, "final_compare": 0},
                     {"evaluations": 2, "judge_ratio": 100.0, "final_compare": 0}]
        for opts in runs:
            path.set_proposal(net, **opts)
            seconds = timed()
            evaluations = int(path.frame.evaluations)
            stages = {}
            for stage, ms in path.stage_times():
                stages[stage] = stages.get(stage, 0.0) + ms
            if opts.get("final_compare") == 0:
                # the unjudged last pass compares nothing: its quality is that
                # of the same run comparing the last pass, whose tokens (and
                # so field and quantization) are the same
                tokens = path.tokens()
                path.set_proposal(net, **dict(opts, final_compare=1))
                path.run(resident=resident)
                same = all(np.array_equal(tokens[k], path.tokens()[k]) for k in tokens)
                if not same:
                    raise RuntimeError("the unjudged last pass changed the tokens")
            q, b = _rate_quality(path)
            cells = path.view("cells").numpy()
            allow = path.view("proposal_allow").numpy()
            out["runs"].append({
                "network": name, "options": opts or "defaults",
                "evaluations": evaluations,
                "value": round(mp / seconds, 3), "ms_per_image": round(seconds * 1e3, 3),
                "speedup_vs_plain": round(plain_s / seconds, 3),
                "quality_3norm": round(q, 5), "rate_excess_at_equal_quality": excess(q, b),
                "grid_agreement": round(float(((cells >> 1) == (plain_cells >> 1)).mean()), 4),
                "plain_search_blocks": round(float((allow == 0x7F).mean()), 4),
                "stage_ms": {k: round(v, 3) for k, v in stages.items()
                             if k in ("proposal", "ac_strategy_costs", "aq_evaluate",
                                      "butteraugli_compare")}})
        path.set_proposal(None)
        net.close()
    path.close()
    resident.free()
    return out


def main() -> int:
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=500)
    parser.add_argument("--warmup", type=int, default=10)
    parser.add_argument("--workload", default="4032x3024")
    parser.add_argument("--probe", default="auto",
                        help="kernel family timed per launch in the timed region "
                             "(auto: the one with the most device time per image)")
    parser.add_argument("--pcie-steps", type=int, default=50,
                        help="steps with the upload from pinned host memory "
                             "(secondary figure)")
    parser.add_argument("--concurrent-images", type=int, default=3)
    parser.add_argument("--concurrent-steps", type=int, default=60)
    parser.add_argument("--batch-images", type=int, default=64,
                        help="images of uneven extents for the batch-driver figure "
                             "(0: skip)")
    parser.add_argument("--batch-lanes", type=int, default=3)
    parser.add_argument("--split-distances", default="0.5,1,2,3",
                        help="distances of the split-precision DCT figures ('' to skip)")
    parser.add_argument("--learned-steps", type=int, default=20,
                        help="timed steps per learned-proposal configuration (0: skip)")
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
        print("bench.py measures one GPU only: the library has no multi-GPU "
              "path (images and tiles are independent, so it would shard by "
              "image, or by tile for one large image).", file=sys.stderr)
        return 2

    import gjxl_amd
    from gjxl_amd import synthetic

    if args.calibration_copy:
        calibration_copy()
    spec = synthetic.WORKLOADS[args.workload]
    w, h = spec["width"], spec["height"]
    mp = w * h / 1e6
    srgb = synthetic.generate(**spec)
    gjxl_amd.library()
    dev = gjxl_amd.Device(0)
    path = gjxl_amd.NativePlainPath(dev, w, h)
    resident = dev.array(srgb)        # packed sRGB8 in HBM
    path.pixels.array[...] = srgb.reshape(-1)  # and in pinned host memory
    print(f"bench: {args.workload}, warmup {args.warmup}", file=sys.stderr, flush=True)
    for _ in range(max(1, args.warmup)):
        path.run(resident=resident)
    reference_digest = _digest(path)

    # the probed kernel family: the one with the most device time per image
    # (one pass per family, before the timed region)
    probe = args.probe
    family_ms = {}
    for name in dev.PROBE_KERNELS:
        dev.probe(name)
        path.run(resident=resident)
        family_ms[name] = float(np.sum(dev.probe_times()))
    if probe == "auto":
        probe = max(family_ms, key=family_ms.get)
    dev.probe(probe)

    # timed region: one image per step, packed pixels resident in HBM, every
    # stage, then the token values, contexts and group offsets read back
    # into pinned host memory
    print(f"bench: timing {args.steps} steps (probe {probe})", file=sys.stderr, flush=True)
    stage_ms: dict[str, list[float]] = {}
    launches: list[float] = []
    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        path.run(resident=resident)
        launches.extend(dev.probe_times())
        for name, ms in path.stage_times():
            stage_ms.setdefault(name, []).append(ms)
    dev.synchronize()
    elapsed = time.perf_counter() - t0
    dev.probe(None)
    if _digest(path) != reference_digest:
        print("timed runs disagree with the warm-up run", file=sys.stderr)
        return 1
    ms_per_step = elapsed * 1e3 / args.steps

    # secondary: the same step with the upload from pinned host memory
    t0 = time.perf_counter()
    for _ in range(args.pcie_steps):
        path.run()
    pcie_s = time.perf_counter() - t0
    if _digest(path) != reference_digest:
        print("runs from host memory disagree", file=sys.stderr)
        return 1
    tokens = int(path.frame.token_count)
    scores = path.scores()
    path.close()

    concurrent = None
    if args.concurrent_images > 1:
        print(f"bench: {args.concurrent_images} concurrent images", file=sys.stderr, flush=True)
        concurrent = _concurrent(gjxl_amd, srgb, w, h, args.concurrent_images,
                                 args.concurrent_steps)
        if any(d != reference_digest for d in concurrent.pop("digests")):
            print("concurrent images disagree", file=sys.stderr)
            return 1

    batch = None
    if args.batch_images > 0:
        print(f"bench: batch of {args.batch_images} uneven images", file=sys.stderr, flush=True)
        images = batch_workload(args.batch_images)
        total_mp = sum(a.shape[0] * a.shape[1] for a in images) / 1e6
        _batch(gjxl_amd, images[:2], args.batch_lanes, calls=1)  # warm-up
        alone_s, alone_counts = _batch(gjxl_amd, images, 1)
        lanes_s, lanes_counts = _batch(gjxl_amd, images, args.batch_lanes)
        queue_s, queue_counts = _batch(gjxl_amd, images, args.batch_lanes, tile_queue=True)
        wide = max(args.batch_lanes, 6)
        wide_lanes_s, wide_lanes_counts = _batch(gjxl_amd, images, wide)
        wide_s, wide_counts = _batch(gjxl_amd, images, wide, tile_queue=True)
        if not (alone_counts == lanes_counts == queue_counts == wide_counts ==
                wide_lanes_counts):
            print("batch lanes disagree", file=sys.stderr)
            return 1

        def figure(seconds):
            return {"seconds": round(seconds[-1], 4),
                    "value": round(total_mp / seconds[-1], 3), "unit": "MP/s",
                    "first_call_seconds": round(seconds[0], 4)}

        batch = {"images": args.batch_images, "megapixels": round(total_mp, 2),
                 "extents": "width, height uniform in [256, 4096] step 8",
                 "note": "second encode call of the same images; the first also "
                         "creates each worker's plain path at the bounding extent",
                 "one_at_a_time": figure(alone_s),
                 f"lanes_{args.batch_lanes}": figure(lanes_s),
                 f"tile_queue_lanes_{args.batch_lanes}": figure(queue_s),
                 f"lanes_{wide}": figure(wide_lanes_s),
                 f"tile_queue_lanes_{wide}": figure(wide_s),
                 "candidate_costs": cost_queue_figures(gjxl_amd, dev, images),
                 "candidate_costs_sides_256_to_1024": cost_queue_figures(
                     gjxl_amd, dev, batch_workload(args.batch_images, seed=5, largest=1024))}

    split = None
    if args.split_distances:
        print("bench: split-precision DCT", file=sys.stderr, flush=True)
        split = split_dct_figures(gjxl_amd, dev, srgb,
                                  [float(v) for v in args.split_distances.split(",")])

    learned = None
    if args.learned_steps > 0:
        print("bench: learned proposal", file=sys.stderr, flush=True)
        learned = learned_figures(gjxl_amd, dev, srgb, args.learned_steps)

    per_step = {k: float(np.sum(v)) / args.steps for k, v in stage_ms.items()}
    kind, units = probe_units(probe, w, h)
    launch_ms = float(np.mean(launches))
    if kind == "flops":
        achieved = units / (launch_ms * 1e-3) / 1e12
        roofline = {"kernel": probe, "bound": "fp64 (VALU, no MFMA)",
                    "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                    "traffic": committed_traffic(probe) if args.workload == "4032x3024"
                    else None}
    else:
        achieved = units / (launch_ms * 1e-3) / 1e9
        roofline = {"kernel": probe, "bound": "hbm", "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBPS, 4),
                    "traffic": committed_traffic(probe) if args.workload == "4032x3024"
                    else None}
    roofline.update({"launches": len(launches), "ms_per_launch": round(launch_ms, 4),
                     "ms_per_image": round(launch_ms * len(launches) / args.steps, 4),
                     "algorithmic_" + kind + "_per_launch": round(units),
                     "family_ms_per_image": {k: round(v, 4) for k, v in family_ms.items()}})
    result = {
        "metric": "MP/s at d1.0 e7, plain GPU path (XYB, AC strategy search, "
                  "Butteraugli AQ loop, AC tokenization), 1 GPU",
        "value": round(mp / (ms_per_step * 1e-3), 3),
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
            "workload": f"{w}x{h} 8-bit sRGB, seeded synthetic ({spec['recipe']}, "
                        f"seed {spec['seed']}), distance 1.0 effort 7, one image per step",
            "span": "packed sRGB8 resident in HBM -> every stage -> token values, "
                    "contexts and group offsets in pinned host memory",
            "stage_ms": {k: round(v, 4) for k, v in per_step.items()},
            "ac_tokens": tokens,
            "scores": [round(v, 6) for v in scores],
            "pcie_inclusive": {
                "span": "upload from pinned host memory included",
                "steps": args.pcie_steps,
                "ms_per_step": round(pcie_s * 1e3 / args.pcie_steps, 4),
                "value": round(mp * args.pcie_steps / pcie_s, 3),
                "unit": "MP/s"} if args.pcie_steps > 0 else None,
            "concurrent_images": concurrent,
            "batch_driver": batch,
            "split_precision_dct": split,
            "learned_proposal": learned,
        },
        "roofline": roofline,
    }
    dev.close()
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(srgb)
        cpu["reference_cpu_estimate"] = {
            "value": round(cpu["value"] * ORACLE_CPU_SECONDS_512 / REFERENCE_CPU_SECONDS_512, 4),
            "unit": "MP/s",
            "how": "oracle rate on this host x (oracle 4.45 s / reference CPU encoder "
                   "2.91 s per 512x512 image at d1 e7, both timed on one build-container "
                   "core)"}
        result["cpu_baseline"] = cpu
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
