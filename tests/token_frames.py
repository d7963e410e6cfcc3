"""Synthetic quantized frames for the tokenization tests: a strategy grid
tiled from 32x32-pixel patterns of the seven searched strategies, quantized
coefficients in the per-group store layout (3 planes of groups * 65536
int32) and a raw quant field."""
import numpy as np

MASK64 = (1 << 64) - 1
COVERED = {0: (1, 1), 4: (2, 2), 5: (4, 4), 6: (1, 2), 7: (2, 1), 10: (2, 4), 11: (4, 2)}


def _pattern(placements):
    """4x4 cells from (strategy, x, y) anchors."""
    cells = np.zeros((4, 4), np.uint8)
    for st, x, y in placements:
        cw, ch = COVERED[st]
        cells[y:y + ch, x:x + cw] = st << 1
        cells[y, x] |= 1
    return cells


PATTERNS = [
    _pattern([(0, x, y) for y in range(4) for x in range(4)]),
    _pattern([(5, 0, 0)]),
    _pattern([(4, 0, 0), (4, 2, 0), (4, 0, 2), (4, 2, 2)]),
    _pattern([(10, 0, 0), (10, 2, 0)]),
    _pattern([(11, 0, 0), (11, 0, 2)]),
    _pattern([(6, x, y) for y in (0, 2) for x in range(4)]),
    _pattern([(7, x, y) for y in range(4) for x in (0, 2)]),
    _pattern([(4, 0, 0), (7, 2, 0), (7, 2, 1), (6, 0, 2), (6, 1, 2), (0, 2, 2),
              (0, 3, 2), (0, 2, 3), (0, 3, 3)]),
]


def strategy_grid(rng, bw, bh, weights=None):
    """Cells of a bw x bh block grid: random patterns on the whole 4x4 tiles,
    DCT8 on the partial tiles at the right and bottom edges."""
    cells = np.ones((bh, bw), np.uint8)
    tw, th = bw // 4, bh // 4
    if tw and th:
        p = np.asarray(weights if weights is not None else np.ones(len(PATTERNS)), float)
        choice = rng.choice(len(PATTERNS), size=(th, tw), p=p / p.sum())
        tiles = np.stack(PATTERNS)[choice]              # th, tw, 4, 4
        cells[:th * 4, :tw * 4] = tiles.transpose(0, 2, 1, 3).reshape(th * 4, tw * 4)
    return cells


def frame(rng, bw, bh, density=0.15, spread=30, weights=None):
    cells = strategy_grid(rng, bw, bh, weights)
    groups = ((bw + 31) // 32) * ((bh + 31) // 32)
    q = np.zeros((3, groups * 65536), np.int32)
    mask = rng.random(q.shape) < density
    q[mask] = rng.integers(-spread, spread + 1, size=int(mask.sum()))
    raw = rng.integers(1, 257, size=(bh, bw)).astype(np.int32)
    return cells, q, raw


def fnv_values(values):
    """FNV-1a 64 over words (tests/ac_group_test.cpp:36-43)."""
    h = 1469598103934665603
    for v in values:
        h = ((h ^ int(v)) * 1099511628211) & MASK64
    return h


def fnv_tokens(contexts, values):
    """FNV-1a 64 over (context, value) pairs (tests/ac_group_test.cpp:45-54)."""
    h = 1469598103934665603
    for c, v in zip(contexts, values):
        h = ((h ^ int(c)) * 1099511628211) & MASK64
        h = ((h ^ int(v)) * 1099511628211) & MASK64
    return h


def poison_llf(strategy, base):
    """One channel of a transform with its LLF entries base, base + 1, ...
    (tests/ac_group_test.cpp:157-167)."""
    cw, ch = COVERED[strategy]
    lw, lh, W = max(cw, ch), min(cw, ch), max(cw, ch) * 8
    v = np.zeros(cw * ch * 64, np.int32)
    for y in range(lh):
        for x in range(lw):
            v[y * W + x] = base + y * lw + x
    return v
