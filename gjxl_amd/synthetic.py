"""Seeded synthetic 8-bit sRGB test images.

Two recipes, both deterministic for a given (width, height, seed):

* ``basic``: smooth 2-D colour gradients + band-limited Gaussian noise
  (sigma 6 in 8-bit units) + 40 sharp-edged filled rectangles + 200
  high-contrast strokes 1-3 px wide (the 512x512 workload).
* ``textured``: ``basic`` plus tiled texture patches at four scales
  (periods 2, 4, 8 and 16 px) and flat regions covering about 10 % of the
  area (the 4032x3024 workload).

Returns an (height, width, 3) uint8 array (packed RGB, row stride 3*width).
"""
from __future__ import annotations

import numpy as np


def _band_limited_noise(rng, height, width, sigma_px=2.0):
    from scipy.ndimage import gaussian_filter

    noise = rng.standard_normal((3, height, width)).astype(np.float32)
    for c in range(3):
        noise[c] = gaussian_filter(noise[c], sigma_px, mode="wrap")
        noise[c] /= max(float(noise[c].std()), 1e-6)
    return noise


def _gradients(rng, height, width):
    y = np.linspace(0.0, 1.0, height, dtype=np.float32)[:, None]
    x = np.linspace(0.0, 1.0, width, dtype=np.float32)[None, :]
    image = np.empty((3, height, width), np.float32)
    for c in range(3):
        a, b, d = rng.uniform(-1.0, 1.0, 3)
        fx, fy = rng.uniform(0.5, 2.5, 2)
        phase = rng.uniform(0, 2 * np.pi)
        image[c] = 128.0 + 60.0 * (a * x + b * y) + 40.0 * d * np.sin(
            2 * np.pi * (fx * x + fy * y) + phase)
    return image


def _rectangles(rng, image, count):
    _, height, width = image.shape
    for _ in range(count):
        w = int(rng.integers(4, max(5, width // 6)))
        h = int(rng.integers(4, max(5, height // 6)))
        x0 = int(rng.integers(0, width - w))
        y0 = int(rng.integers(0, height - h))
        image[:, y0:y0 + h, x0:x0 + w] = rng.uniform(0, 255, 3)[:, None, None]


def _strokes(rng, image, count):
    _, height, width = image.shape
    for _ in range(count):
        thickness = int(rng.integers(1, 4))
        x0, x1 = rng.uniform(0, width, 2)
        y0, y1 = rng.uniform(0, height, 2)
        length = max(abs(x1 - x0), abs(y1 - y0))
        # keep strokes moderate so they do not dominate the image
        scale = min(1.0, (0.25 * max(width, height)) / max(length, 1.0))
        x1 = x0 + (x1 - x0) * scale
        y1 = y0 + (y1 - y0) * scale
        n = int(max(abs(x1 - x0), abs(y1 - y0))) + 1
        t = np.linspace(0.0, 1.0, n)
        xs = np.round(x0 + (x1 - x0) * t).astype(np.int64)
        ys = np.round(y0 + (y1 - y0) * t).astype(np.int64)
        color = np.where(rng.uniform(size=3) < 0.5, 0.0, 255.0)
        for dy in range(thickness):
            for dx in range(thickness):
                yy = np.clip(ys + dy, 0, height - 1)
                xx = np.clip(xs + dx, 0, width - 1)
                image[:, yy, xx] = color[:, None]


def _textures(rng, image, patches_per_scale):
    _, height, width = image.shape
    for period in (2, 4, 8, 16):
        for _ in range(patches_per_scale):
            w = int(rng.integers(period * 4, max(period * 5, width // 8)))
            h = int(rng.integers(period * 4, max(period * 5, height // 8)))
            w, h = min(w, width - 1), min(h, height - 1)
            x0 = int(rng.integers(0, width - w))
            y0 = int(rng.integers(0, height - h))
            yy, xx = np.mgrid[0:h, 0:w]
            kind = int(rng.integers(0, 3))
            if kind == 0:
                pattern = ((xx // period + yy // period) % 2).astype(np.float32)
            elif kind == 1:
                pattern = 0.5 + 0.5 * np.sin(2 * np.pi * xx / period)
            else:
                pattern = 0.5 + 0.5 * np.sin(2 * np.pi * (xx + yy) / period)
            lo, hi = rng.uniform(0, 255, (2, 3))
            image[:, y0:y0 + h, x0:x0 + w] = (
                lo[:, None, None] + (hi - lo)[:, None, None] * pattern[None])


def _flat_regions(rng, image, fraction):
    _, height, width = image.shape
    target = fraction * height * width
    covered = 0
    while covered < target:
        w = int(rng.integers(16, max(17, width // 5)))
        h = int(rng.integers(16, max(17, height // 5)))
        w, h = min(w, width), min(h, height)
        x0 = int(rng.integers(0, width - w + 1))
        y0 = int(rng.integers(0, height - h + 1))
        image[:, y0:y0 + h, x0:x0 + w] = rng.uniform(0, 255, 3)[:, None, None]
        covered += w * h


def generate(width: int, height: int, seed: int = 1,
             recipe: str = "basic") -> np.ndarray:
    rng = np.random.default_rng(seed)
    image = _gradients(rng, height, width)
    image += 6.0 * _band_limited_noise(rng, height, width)
    if recipe == "textured":
        _textures(rng, image, patches_per_scale=max(2, (width * height) // 400_000))
        _flat_regions(rng, image, 0.10)
    elif recipe != "basic":
        raise ValueError(f"unknown recipe {recipe!r}")
    _rectangles(rng, image, 40)
    _strokes(rng, image, 200)
    out = np.clip(np.round(image), 0, 255).astype(np.uint8)
    return np.ascontiguousarray(out.transpose(1, 2, 0))


WORKLOADS = {
    "512x512": dict(width=512, height=512, seed=1, recipe="basic"),
    "4032x3024": dict(width=4032, height=3024, seed=2, recipe="textured"),
}
