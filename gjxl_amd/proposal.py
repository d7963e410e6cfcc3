"""The learned AC-strategy proposal, a chosen form beside the plain path.

A ctypes mirror of gjxl_amd_proposal_* (include/gjxl_amd.h), the collection
of training data from plain runs (gjxl_amd_plain_path_proposal_sample), the
training loop on the library's own forward, backward and Adam kernels, and
the parameters file.

The network predicts per 8x8 block logits over the format's 27 strategy
codes and the log ratio of the Butteraugli loop's final quant field to the
adjusted initial field.  Its labels are what the plain path chose on seeded
synthetic images: the strategy map of the greedy merge
(src/codec/ac_strategy.cpp:1101 FindAcStrategyGrid in the reference) and the
field after the loop (src/codec/adaptive_quantization.cpp:1600-1738).  A
plain path with the network attached (NativePlainPath.set_proposal) prunes
the candidate costs of confident tiles and starts the loop from the
predicted field.

Usage (on a GPU):
  python -m gjxl_amd.proposal train --out PATH [--images N ...]
"""
from __future__ import annotations

import argparse
import concurrent.futures
import ctypes
import json
import multiprocessing
import time
import weakref
from pathlib import Path

import numpy as np

from . import (DeviceArray, NativePlainPath, _check, _F, _P, _SZ, _U32, library,
               synthetic)

FEATURES = 16
CHANNELS = 32
LAYERS = 6
CLASSES = 27
FIELD_CHANNEL = 27
IGNORE = 255
TAPS = 9
# strategy code of each candidate stage (the cost planes' order)
STRATEGIES = (0, 6, 7, 4, 10, 11, 5)
PARAMETERS_FILE = Path(__file__).resolve().with_name("proposal_parameters.npz")
CROP = 32  # blocks per side of a training crop (256 x 256 pixels)

_HEAD, _PARAMETERS, _GRADIENTS, _ACTIVATIONS = 0, 1, 2, 3


class ProposalOptions(ctypes.Structure):
    _fields_ = [("struct_size", _SZ), ("threshold", _F), ("top_k", _U32), ("p_min", _F),
                ("evaluations", ctypes.c_int32), ("field_limit", _F), ("field_offset", _F),
                ("judge_ratio", _F), ("final_compare", ctypes.c_int32)]

    @classmethod
    def defaults(cls) -> "ProposalOptions":
        o = cls()
        _check(library().gjxl_amd_proposal_options_init(ctypes.byref(o), ctypes.sizeof(o)))
        return o


def parameter_count() -> int:
    return int(library().gjxl_amd_proposal_parameter_count())


def layer_inputs(layer: int) -> int:
    return FEATURES if layer == 0 else CHANNELS


def init_parameters(seed: int) -> np.ndarray:
    """He-normal weights and zero biases; the head's weights are scaled by
    0.1 and those of its unused channels are zero."""
    rng = np.random.default_rng(seed)
    parts = []
    for L in range(LAYERS):
        cin = layer_inputs(L)
        std = np.sqrt(2.0 / (TAPS * cin)) * (0.1 if L == LAYERS - 1 else 1.0)
        w = rng.standard_normal((TAPS, CHANNELS, cin)) * std
        if L == LAYERS - 1:
            w[:, CLASSES + 1:, :] = 0.0
        parts += [w.reshape(-1), np.zeros(CHANNELS)]
    return np.concatenate(parts).astype(np.float32)


def load_parameters(path: Path = PARAMETERS_FILE) -> np.ndarray:
    with np.load(path, allow_pickle=False) as z:
        params = np.asarray(z["parameters"], np.float32)
    if params.size != parameter_count():
        raise ValueError(f"{path}: {params.size} parameters, expected {parameter_count()}")
    return params


def load_metadata(path: Path = PARAMETERS_FILE) -> dict:
    with np.load(path, allow_pickle=False) as z:
        return json.loads(str(z["metadata"]))


def save_parameters(params: np.ndarray, path: Path, metadata: dict) -> None:
    np.savez(path, parameters=np.asarray(params, np.float32),
             metadata=np.array(json.dumps(metadata)))


class ProposalNet:
    """One network on a device (gjxl_amd_proposal_create) with room for
    max_positions block positions per forward."""

    def __init__(self, device, max_positions: int, parameters: np.ndarray | None = None):
        self.device = device
        handle = ctypes.c_void_p()
        _check(library().gjxl_amd_proposal_create(device.handle, max_positions,
                                                  ctypes.byref(handle)))
        self.handle = handle.value
        self.max_positions = max_positions
        self._views = weakref.WeakSet()
        device._children.add(self)
        if parameters is not None:
            self.set_parameters(parameters)

    def close(self) -> None:
        for view in list(self._views):
            view.ptr = None
        if self.handle:
            library().gjxl_amd_proposal_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_parameters(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32)
        _check(library().gjxl_amd_proposal_set_parameters(self.handle, p.ctypes.data, p.size))

    def parameters(self) -> np.ndarray:
        out = np.empty(parameter_count(), np.float32)
        _check(library().gjxl_amd_proposal_get_parameters(self.handle, out.ctypes.data,
                                                          out.size))
        return out

    def buffer(self, which: int, shape, dtype) -> DeviceArray:
        ptr = ctypes.c_void_p()
        size = ctypes.c_size_t()
        _check(library().gjxl_amd_proposal_buffer(self.handle, which, ctypes.byref(ptr),
                                                  ctypes.byref(size)))
        view = DeviceArray(self.device, shape, dtype, pointer=ptr.value, owner=self)
        if view.nbytes > size.value:
            raise ValueError("view exceeds the buffer")
        return view

    def head(self, positions: int | None = None) -> DeviceArray:
        return self.buffer(_HEAD, (positions or self.max_positions, CHANNELS), np.float32)

    def gradients(self) -> np.ndarray:
        return self.buffer(_GRADIENTS, (parameter_count(),), np.float32).numpy()

    def activations(self) -> DeviceArray:
        return self.buffer(_ACTIVATIONS, (LAYERS - 1, self.max_positions, CHANNELS), np.uint16)

    def forward(self, features: DeviceArray, n: int, h: int, w: int) -> None:
        """Enqueues the forward; `features` must stay alive until the next
        backward has run (it reads them again)."""
        if features.nbytes < n * h * w * FEATURES * 2:
            raise ValueError("features are smaller than n * h * w positions")
        _check(library().gjxl_amd_proposal_forward(self.handle, features.ptr, n, h, w))
        self._features = features  # kept alive for the backward

    def backward(self, labels: DeviceArray, targets: DeviceArray, qf_weight: float,
                 sums: bool = True):
        """Gradients of the last forward; (cross-entropy sum, squared-error
        sum, correct count) when sums (a device wait)."""
        out = (ctypes.c_double * 3)()
        _check(library().gjxl_amd_proposal_backward(self.handle, labels.ptr, targets.ptr,
                                                    qf_weight, out if sums else None))
        return (out[0], out[1], int(round(out[2]))) if sums else None

    def adam(self, learning_rate: float, beta1: float = 0.9, beta2: float = 0.999,
             epsilon: float = 1e-8, grad_scale: float = 1.0) -> None:
        _check(library().gjxl_amd_proposal_adam(self.handle, learning_rate, beta1, beta2,
                                                epsilon, grad_scale))


def features(device, opsin: DeviceArray, width: int, height: int,
             quant_field: DeviceArray, out: DeviceArray | None = None) -> DeviceArray:
    """Block features of a padded XYB image (width, height multiples of 8)."""
    out = out or device.empty((height // 8, width // 8, FEATURES), np.uint16)
    _check(library().gjxl_amd_proposal_features(device.handle, opsin.ptr, width, height,
                                                quant_field.ptr, out.ptr))
    return out


def decide(device, head: DeviceArray, bw: int, bh: int, threshold: float, top_k: int,
           p_min: float, out: DeviceArray | None = None) -> DeviceArray:
    out = out or device.empty((bh, bw), np.uint8)
    _check(library().gjxl_amd_proposal_decide(device.handle, head.ptr, bw, bh, threshold,
                                              top_k, p_min, out.ptr))
    return out


def warm_start(device, head: DeviceArray, adjusted: DeviceArray, count: int, limit: float,
               lower: float, upper: float, offset: float = 0.0,
               out: DeviceArray | None = None) -> DeviceArray:
    out = out or device.empty((count,), np.float32)
    _check(library().gjxl_amd_proposal_warm_start(device.handle, head.ptr, adjusted.ptr,
                                                  count, limit, offset, lower, upper, out.ptr))
    return out


# ---- training data ------------------------------------------------------------

def _generate(args):
    width, height, seed, recipe = args
    return synthetic.generate(width, height, seed=seed, recipe=recipe)


def collect(device, count: int, width: int, height: int, seed0: int,
            recipe: str = "textured", workers: int = 8, log=None):
    """Runs the plain path on `count` seeded synthetic images (seeds seed0,
    seed0 + 1, ...) and cuts their block grids into CROP x CROP crops:
    features [N][CROP][CROP][16] (bfloat16 bits), labels [N][CROP][CROP]
    (strategy codes) and targets [N][CROP][CROP] (log field ratios)."""
    path = NativePlainPath(device, width, height)
    bw, bh = path.bw, path.bh
    f_dev = device.empty((bh, bw, FEATURES), np.uint16)
    l_dev = device.empty((bh, bw), np.uint8)
    t_dev = device.empty((bh, bw), np.float32)
    feats, labels, targets = [], [], []
    jobs = [(width, height, seed0 + i, recipe) for i in range(count)]
    ctx = multiprocessing.get_context("spawn")
    with concurrent.futures.ProcessPoolExecutor(max_workers=workers, mp_context=ctx) as pool:
        for i, img in enumerate(pool.map(_generate, jobs)):
            path.run(img, read_tokens=False)
            path.proposal_sample(f_dev, l_dev, t_dev)
            f, lab, t = f_dev.numpy(), l_dev.numpy(), t_dev.numpy()
            for y0 in range(0, bh - CROP + 1, CROP):
                for x0 in range(0, bw - CROP + 1, CROP):
                    feats.append(f[y0:y0 + CROP, x0:x0 + CROP])
                    labels.append(lab[y0:y0 + CROP, x0:x0 + CROP])
                    targets.append(t[y0:y0 + CROP, x0:x0 + CROP])
            if log and (i + 1) % 50 == 0:
                log(f"collect: {i + 1}/{count} images, {len(feats)} crops")
    path.close()
    for a in (f_dev, l_dev, t_dev):
        a.free()
    return np.stack(feats), np.stack(labels), np.stack(targets)


def train(device, feats: np.ndarray, labels: np.ndarray, targets: np.ndarray,
          epochs: int, batch: int = 32, learning_rate: float = 2e-3,
          qf_weight: float = 1.0, seed: int = 0, params: np.ndarray | None = None,
          log=None):
    """Adam with a cosine learning-rate decay over shuffled minibatches of
    crops, each step the library's forward, backward and Adam kernels.
    Returns (parameters, per-epoch history)."""
    n = feats.shape[0]
    h, w = feats.shape[1:3]
    positions = batch * h * w
    net = ProposalNet(device, positions, params if params is not None else init_parameters(seed))
    f_dev = device.empty((batch, h, w, FEATURES), np.uint16)
    l_dev = device.empty((batch, h, w), np.uint8)
    t_dev = device.empty((batch, h, w), np.float32)
    rng = np.random.default_rng(seed)
    steps = epochs * (n // batch)
    history = []
    step = 0
    begin = time.time()
    for epoch in range(epochs):
        order = rng.permutation(n)
        ce = se = correct = 0.0
        seen = 0
        for b in range(n // batch):
            idx = order[b * batch:(b + 1) * batch]
            f_dev.copy_from(feats[idx])
            l_dev.copy_from(labels[idx])
            t_dev.copy_from(targets[idx])
            net.forward(f_dev, batch, h, w)
            s = net.backward(l_dev, t_dev, qf_weight, sums=True)
            lr = learning_rate * 0.5 * (1.0 + np.cos(np.pi * step / max(steps, 1)))
            net.adam(lr, grad_scale=1.0 / positions)
            ce, se, correct = ce + s[0], se + s[1], correct + s[2]
            seen += positions
            step += 1
        entry = {"epoch": epoch + 1, "cross_entropy": ce / seen, "field_mse": se / seen,
                 "accuracy": correct / seen, "seconds": round(time.time() - begin, 2)}
        history.append(entry)
        if log:
            log(f"train: {json.dumps(entry)}")
    out = net.parameters()
    net.close()
    for a in (f_dev, l_dev, t_dev):
        a.free()
    return out, history


def evaluate(device, params: np.ndarray, feats: np.ndarray, labels: np.ndarray,
             targets: np.ndarray, batch: int = 32) -> dict:
    """Top-1 and top-2 accuracy over the seven candidate strategies and the
    field error, against the all-zero field prediction."""
    n, h, w = labels.shape
    net = ProposalNet(device, batch * h * w, params)
    f_dev = device.empty((batch, h, w, FEATURES), np.uint16)
    top1 = top2 = total = 0
    se = se0 = 0.0
    for b in range(0, n - batch + 1, batch):
        f_dev.copy_from(feats[b:b + batch])
        net.forward(f_dev, batch, h, w)
        head = net.head(batch * h * w).numpy().reshape(batch, h, w, CHANNELS)
        z = head[..., list(STRATEGIES)]
        lab = labels[b:b + batch]
        stage = np.full(lab.shape, -1)
        for s, code in enumerate(STRATEGIES):
            stage[lab == code] = s
        order = np.argsort(-z, axis=-1, kind="stable")
        live = stage >= 0
        top1 += int(((order[..., 0] == stage) & live).sum())
        top2 += int((((order[..., 0] == stage) | (order[..., 1] == stage)) & live).sum())
        total += int(live.sum())
        t = targets[b:b + batch]
        se += float(((head[..., FIELD_CHANNEL] - t) ** 2).sum())
        se0 += float((t ** 2).sum())
    net.close()
    f_dev.free()
    return {"top1": top1 / total, "top2": top2 / total, "field_mse": se / total,
            "field_mse_zero": se0 / total, "blocks": total}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train", help="collect plain-path labels and train on the GPU")
    t.add_argument("--out", required=True)
    t.add_argument("--images", type=int, default=427)
    t.add_argument("--width", type=int, default=2048)
    t.add_argument("--height", type=int, default=1536)
    t.add_argument("--seed0", type=int, default=1000)
    t.add_argument("--held-out", type=int, default=8)
    t.add_argument("--epochs", type=int, default=20)
    t.add_argument("--batch", type=int, default=32)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--qf-weight", type=float, default=1.0)
    t.add_argument("--workers", type=int, default=14)
    args = ap.parse_args()
    from . import Device

    def log(msg):
        print(msg, flush=True)

    device = Device(0)
    begin = time.time()
    feats, labels, targets = collect(device, args.images, args.width, args.height,
                                     args.seed0, workers=args.workers, log=log)
    log(f"collect: {feats.shape[0]} crops in {time.time() - begin:.1f} s")
    vf, vl, vt = collect(device, args.held_out, args.width, args.height,
                         args.seed0 + args.images, workers=args.workers)
    begin = time.time()
    params, history = train(device, feats, labels, targets, args.epochs, args.batch,
                            args.lr, args.qf_weight, log=log)
    seconds = time.time() - begin
    held = evaluate(device, params, vf, vl, vt)
    untrained = evaluate(device, init_parameters(0), vf, vl, vt)
    counts = {int(c): int((labels == c).sum()) for c in STRATEGIES}
    meta = {"images": args.images, "extent": [args.width, args.height], "seed0": args.seed0,
            "recipe": "textured", "crops": int(feats.shape[0]), "crop_blocks": CROP,
            "epochs": args.epochs, "batch": args.batch, "learning_rate": args.lr,
            "qf_weight": args.qf_weight, "train_seconds": round(seconds, 2),
            "history": history, "held_out": held, "held_out_untrained": untrained,
            "label_counts": counts}
    save_parameters(params, Path(args.out), meta)
    log(json.dumps({"held_out": held, "untrained": untrained, "train_seconds": seconds}))
    device.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
