"""CPU restatement of the learned AC-strategy proposal network.
TEST INFRASTRUCTURE ONLY: tests/, the smoke check and bench.py's checker legs
may use it; the product path (gjxl_amd/csrc/proposal.hip) never does.

The network has no counterpart in the reference: it learns to predict the
plain search's choices (src/codec/ac_strategy.cpp:1101 FindAcStrategyGrid, the
strategy map) and the Butteraugli loop's final quant field
(src/codec/adaptive_quantization.cpp:1600-1738) from per-block features.

What it computes, per 8x8 block of the padded XYB image:

* features (16 channels, bfloat16): the block mean of each channel (X scaled
  by 16), log1p of the orthonormal-DCT8 energy of each channel in three
  bands of u + v (1-2, 3-5, 6-14) times a per-channel gain, log of the
  initial quant field, three zero channels;
* six 3x3 convolutions with zero padding, 32 outputs each; weights and
  activations bfloat16 (round to nearest even), products accumulated in
  binary32 (binary64 here), bias in binary32; ReLU after the first five, the
  sixth's outputs kept in binary32: channels 0-26 are logits over the
  format's 27 strategy codes, channel 27 the log ratio of the final to the
  adjusted initial quant field, 28-31 unused;
* loss: per block, softmax cross-entropy of the logits against the plain
  search's strategy code plus qf_weight times the squared error of channel
  27, summed over blocks; the output gradient is stored as bfloat16;
* backward: the ReLU mask from the stored bfloat16 outputs, input gradients
  through the transposed, tap-flipped weights (rounded to bfloat16 per
  layer), weight and bias gradients in binary32;
* Adam on binary32 master weights.

Pinned by: float64 central differences of the loss against the analytic
gradients with rounding off (tests/test_oracle_proposal.py), a delta input
whose convolution is the weights' taps, softmax closed forms, and Adam's
first step (whose update is -lr * sign(g) up to eps).
"""
from __future__ import annotations

import numpy as np

FEATURES = 16
HIDDEN = 32
OUT = 32
LAYERS = 6
TAPS = 9
CLASSES = 27
QF_CHANNEL = 27
# candidate strategies of the plain search by stage (the order of its cost
# planes): DCT8, DCT16x8, DCT8x16, DCT16x16, DCT32x16, DCT16x32, DCT32x32
STRATEGIES = (0, 6, 7, 4, 10, 11, 5)
MEAN_SCALE = (16.0, 1.0, 1.0)
ENERGY_GAIN = (1.0e4, 1.0e2, 1.0e2)
IGNORE = 255


def layer_inputs(layer: int) -> int:
    return FEATURES if layer == 0 else HIDDEN


def param_count() -> int:
    return sum(TAPS * OUT * layer_inputs(L) + OUT for L in range(LAYERS))


def param_offsets() -> list[int]:
    out, at = [], 0
    for L in range(LAYERS):
        out.append(at)
        at += TAPS * OUT * layer_inputs(L) + OUT
    return out


def unpack(params: np.ndarray) -> list[tuple[np.ndarray, np.ndarray]]:
    """Flat binary32 parameters -> [(W [9][32][Cin], b [32])] per layer."""
    layers, at = [], 0
    for L in range(LAYERS):
        cin = layer_inputs(L)
        w = params[at:at + TAPS * OUT * cin].reshape(TAPS, OUT, cin)
        at += TAPS * OUT * cin
        b = params[at:at + OUT]
        at += OUT
        layers.append((w, b))
    return layers


def pack(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([w.reshape(-1), b]) for w, b in layers]).astype(
        np.float32)


def init_params(seed: int) -> np.ndarray:
    """He-normal weights, zero biases (the network's initial state)."""
    rng = np.random.default_rng(seed)
    layers = []
    for L in range(LAYERS):
        cin = layer_inputs(L)
        std = np.sqrt(2.0 / (TAPS * cin))
        if L == LAYERS - 1:
            std *= 0.1
        w = (rng.standard_normal((TAPS, OUT, cin)) * std).astype(np.float32)
        if L == LAYERS - 1:
            w[:, CLASSES + 1:, :] = 0.0
        layers.append((w, np.zeros(OUT, np.float32)))
    return pack(layers)


def bf16(x) -> np.ndarray:
    """Round binary32 values to bfloat16 (nearest, ties to even), as binary32."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def dct8_basis() -> np.ndarray:
    k = np.arange(8)[:, None]
    n = np.arange(8)[None, :]
    d = np.cos(np.pi * (2 * n + 1) * k / 16.0) * np.sqrt(2.0 / 8.0)
    d[0] *= np.sqrt(0.5)
    return d


def features(opsin: np.ndarray, quant_field: np.ndarray) -> np.ndarray:
    """opsin [3][ph][pw] binary32, quant_field [bh][bw] -> [bh][bw][16]
    (bfloat16 values as binary32)."""
    _, ph, pw = opsin.shape
    bh, bw = ph // 8, pw // 8
    blocks = opsin.astype(np.float64).reshape(3, bh, 8, bw, 8).transpose(0, 1, 3, 2, 4)
    d = dct8_basis()
    coef = np.einsum("vy,cbayx,ux->cbavu", d, blocks, d)
    s = np.arange(8)[:, None] + np.arange(8)[None, :]
    bands = [(s >= 1) & (s <= 2), (s >= 3) & (s <= 5), s >= 6]
    out = np.zeros((bh, bw, FEATURES), np.float64)
    for c in range(3):
        out[..., c] = blocks[c].mean(axis=(2, 3)) * MEAN_SCALE[c]
        for b, m in enumerate(bands):
            e = (coef[c] ** 2 * m).sum(axis=(2, 3))
            out[..., 3 + 3 * c + b] = np.log1p(ENERGY_GAIN[c] * e)
    out[..., 12] = np.log(quant_field.astype(np.float64))
    return bf16(out.astype(np.float32))


def _shifted(x: np.ndarray, dy: int, dx: int) -> np.ndarray:
    """x[n][y + dy][x + dx][c] with zeros outside."""
    out = np.zeros_like(x)
    n, h, w, _ = x.shape
    ys, yd = slice(max(0, dy), h + min(0, dy)), slice(max(0, -dy), h + min(0, -dy))
    xs, xd = slice(max(0, dx), w + min(0, dx)), slice(max(0, -dx), w + min(0, -dx))
    out[:, yd, xd] = x[:, ys, xs]
    return out


def conv(x: np.ndarray, w: np.ndarray, b: np.ndarray | None) -> np.ndarray:
    """x [N][H][W][Cin], w [9][Cout][Cin] -> [N][H][W][Cout] in binary64;
    tap t = 3 * ky + kx reads x[y + ky - 1][x + kx - 1]."""
    x = x.astype(np.float64)
    out = np.zeros(x.shape[:3] + (w.shape[1],), np.float64)
    for t in range(TAPS):
        dy, dx = t // 3 - 1, t % 3 - 1
        out += _shifted(x, dy, dx) @ w[t].astype(np.float64).T
    if b is not None:
        out += b.astype(np.float64)
    return out


def forward(x: np.ndarray, params: np.ndarray, rounding: bool = True):
    """Returns (inputs per layer, outputs per layer); the last output is the
    binary32 head, the others the stored (bfloat16) ReLU activations."""
    rnd = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params)
    ins, outs = [], []
    h = rnd(x)
    for L, (w, b) in enumerate(layers):
        ins.append(h)
        pre = conv(h, rnd(w), b)
        if L < LAYERS - 1:
            h = rnd(np.maximum(pre, 0.0).astype(np.float32 if rounding else np.float64))
            outs.append(h)
        else:
            outs.append(pre.astype(np.float32) if rounding else pre)
    return ins, outs


def loss_and_output_grad(head: np.ndarray, labels: np.ndarray, targets: np.ndarray,
                         qf_weight: float, rounding: bool = True):
    """head [..][32], labels [..] (strategy code or IGNORE), targets [..].
    Returns (cross-entropy sum, squared-error sum, correct count, grad)."""
    z = head.astype(np.float64)
    logits = z[..., :CLASSES]
    live = labels != IGNORE
    lab = np.where(live, labels, 0).astype(np.int64)
    m = logits.max(axis=-1, keepdims=True)
    e = np.exp(logits - m)
    s = e.sum(axis=-1, keepdims=True)
    p = e / s
    picked = np.take_along_axis(logits, lab[..., None], axis=-1)[..., 0]
    ce = (np.log(s[..., 0]) + m[..., 0] - picked) * live
    err = (z[..., QF_CHANNEL] - targets) * live
    grad = np.zeros_like(z)
    grad[..., :CLASSES] = p
    np.put_along_axis(grad[..., :CLASSES], lab[..., None],
                      np.take_along_axis(p, lab[..., None], axis=-1) - 1.0, axis=-1)
    grad[..., :CLASSES] *= live[..., None]
    grad[..., QF_CHANNEL] = 2.0 * qf_weight * err
    correct = ((np.argmax(logits, axis=-1) == lab) & live).sum()
    if rounding:
        grad = bf16(grad.astype(np.float32))
    return float(ce.sum()), float((err ** 2).sum()), int(correct), grad


def backward(ins, outs, params: np.ndarray, grad_head: np.ndarray, rounding: bool = True):
    """Gradients of the summed loss: flat array shaped like params."""
    rnd = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params)
    grads = [None] * LAYERS
    g = np.asarray(grad_head, np.float64)
    for L in range(LAYERS - 1, -1, -1):
        w, _ = layers[L]
        d = g if L == LAYERS - 1 else g * (outs[L] > 0)
        x = ins[L].astype(np.float64)
        dw = np.zeros(w.shape, np.float64)
        for t in range(TAPS):
            dy, dx = t // 3 - 1, t % 3 - 1
            xs = _shifted(x, dy, dx)
            dw[t] = np.einsum("nhwo,nhwi->oi", d, xs)
        db = d.sum(axis=(0, 1, 2))
        grads[L] = (dw, db)
        if L > 0:
            wr = rnd(w).astype(np.float64)
            # transposed, tap-flipped weights: tap t of the input gradient
            # reads d[y - dy][x - dx]
            dx_in = np.zeros(x.shape, np.float64)
            for t in range(TAPS):
                dy, ddx = t // 3 - 1, t % 3 - 1
                dx_in += _shifted(d, -dy, -ddx) @ wr[t]
            g = rnd(dx_in.astype(np.float32)) if rounding else dx_in
    return np.concatenate([np.concatenate([dw.reshape(-1), db]) for dw, db in grads])


def loss_total(x, params, labels, targets, qf_weight, rounding=False) -> float:
    _, outs = forward(x, params, rounding)
    ce, se, _, _ = loss_and_output_grad(outs[-1], labels, targets, qf_weight, rounding)
    return ce + qf_weight * se


def adam(params, m, v, grads, step, lr, beta1=0.9, beta2=0.999, eps=1e-8, scale=1.0):
    """One Adam step on binary32 state (step counts from 1); returns new
    (params, m, v)."""
    g = grads.astype(np.float32) * np.float32(scale)
    m = np.float32(beta1) * m + np.float32(1.0 - beta1) * g
    v = np.float32(beta2) * v + np.float32(1.0 - beta2) * g * g
    bc1 = np.float32(1.0 - beta1 ** step)
    bc2 = np.float32(1.0 - beta2 ** step)
    update = (m / bc1) / (np.sqrt(v / bc2) + np.float32(eps))
    return (params - np.float32(lr) * update).astype(np.float32), m, v


def decide(head: np.ndarray, threshold: float, top_k: int, p_min: float) -> np.ndarray:
    """head [bh][bw][32] -> allowed-stage bits per block [bh][bw] (uint8).

    Per block: softmax over the seven candidate strategies' logits; stage s
    is allowed if s is DCT8, among the top_k probabilities (ties to the lower
    stage) or at least p_min.  A block whose largest probability is below
    threshold allows every stage (the plain search)."""
    bh, bw, _ = head.shape
    z = head[..., list(STRATEGIES)].astype(np.float64)
    z = z - z.max(axis=-1, keepdims=True)
    p = np.exp(z)
    p /= p.sum(axis=-1, keepdims=True)
    # rank of each stage: how many stages beat it (ties to the lower stage)
    rank = np.zeros(p.shape, np.int64)
    for s in range(7):
        for o in range(7):
            if o == s:
                continue
            rank[..., s] += (p[..., o] > p[..., s]) | ((p[..., o] == p[..., s]) & (o < s))
    bits = ((rank < top_k) | (p >= p_min)).astype(np.uint8)
    bits[..., 0] = 1
    allow = np.zeros((bh, bw), np.uint8)
    for s in range(7):
        allow |= bits[..., s] << s
    allow[p.max(axis=-1) < threshold] = 0x7F
    return allow


def warm_start(head: np.ndarray, adjusted: np.ndarray, lower: float, upper: float,
               limit: float = 2.0, offset: float = 0.0) -> np.ndarray:
    """The loop's starting field: adjusted * exp(clamp(head[..., 27] +
    offset, +-limit)), clamped to the policy bounds (binary32)."""
    r = np.clip(head[..., QF_CHANNEL].astype(np.float32) + np.float32(offset), -limit, limit)
    f = adjusted.astype(np.float32) * np.exp(r).astype(np.float32)
    return np.clip(f, np.float32(lower), np.float32(upper)).astype(np.float32)
