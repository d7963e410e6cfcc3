"""Writes tests/golden/butteraugli_goldens.json from the Butteraugli known-answer
vectors that the reference's own tests hold.

Expected outputs: tests/butteraugli_goldens_generated.h of the gjxl source
tree (libjxl e8ff0976 scalar Butteraugli, read here as data: the hex float
bits of four 32x24 full distance maps with their scores, and of 27
intermediate-stage planes at 16x12).

Inputs: the reference's fixture generator (tests/butteraugli_fixtures.cpp:
44-160, FillPattern) restated for the kinds the goldens use, with binary32
arithmetic in source order and the C library's sinf / cosf, so that the
inputs equal the ones the reference's tests feed.  Both are stored as
float bit patterns.

Run in a container that holds the gjxl sources:
    python tests/golden/make_butteraugli_goldens.py [/path/to/gjxl]
"""
from __future__ import annotations

import ctypes
import json
import re
import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent / "butteraugli_goldens.json"
F = np.float32

_libm = ctypes.CDLL("libm.so.6")
for _name in ("sinf", "cosf"):
    getattr(_libm, _name).argtypes = [ctypes.c_float]
    getattr(_libm, _name).restype = ctypes.c_float


def sinf(v) -> F:
    return F(_libm.sinf(F(v)))


def cosf(v) -> F:
    return F(_libm.cosf(F(v)))


def fill(kind: str, w: int, h: int):
    """FillPattern for the flat, texture, contrast, chromatic and
    intermediate kinds (reference = value, distorted = value + error)."""
    ref = np.zeros((3, h, w), F)
    dist = np.zeros((3, h, w), F)
    for y in range(h):
        for x in range(w):
            for c in range(3):
                fx, fy = F(x), F(y)
                error = F(0)
                if kind == "flat":
                    value = F(0.18) + F(0.04) * F(c)
                    if 10 <= x < 22 and 7 <= y < 17:
                        error = F(0.012) * F(c + 1)
                elif kind == "texture":
                    value = (F(0.25) + F(0.11) * sinf(F(0.41) * F((c + 1) * x + 2 * y))) \
                        + F(0.07) * cosf(F(0.27) * (F(3.0) * fx - fy))
                    error = F(0.018) * sinf(F(0.73) * F(5 * x + 3 * y + c))
                elif kind == "contrast":
                    value = F(0.02) if ((x // 4 + y // 4) & 1) == 0 else F(0.92)
                    if x % 4 == 0 or y % 4 == 0:
                        error = F(0.035) if value < F(0.5) else F(-0.035)
                elif kind == "chromatic":
                    if c == 0:
                        value = F(0.05) + (F(0.75) * fx) / F(max(1, w - 1))
                        error = F(0.02) * sinf(F(0.31) * fy)
                    elif c == 1:
                        value = F(0.08) + (F(0.65) * fy) / F(max(1, h - 1))
                    else:
                        value = F(0.72) - (F(0.55) * (fx + fy)) / F(max(1, w + h - 2))
                        error = F(-0.018) * cosf(F(0.29) * fx)
                elif kind == "intermediate":
                    value = F(0.08) + F(0.021) * fx
                    value = value + F(0.013) * fy
                    value = value + F(0.09) * sinf(F(0.43) * F((c + 1) * x + y))
                    value = value + F(0.04) * cosf(F(0.31) * (fx - F(2.0) * fy))
                    value = min(max(value, F(-0.1)), F(1.2))
                    error = F(0.024) * sinf(F(0.67) * F(3 * x + 5 * y + c))
                    error = error + (F(0.013) if (x + 2 * y + c) % 7 == 0 else F(0))
                else:
                    raise ValueError(kind)
                ref[c, y, x] = value
                dist[c, y, x] = F(value + error)
    return ref, dist


def parse_arrays(text: str, name: str) -> list[list[int]]:
    start = text.index(name)
    body = text[text.index("{{", start):]
    depth, end = 0, 0
    for i, ch in enumerate(body):
        depth += ch == "{"
        depth -= ch == "}"
        if depth == 0:
            end = i
            break
    inner = body[1:end]
    groups = re.findall(r"\{\{(.*?)\}\}", inner, re.S)
    return [[int(v, 16) for v in re.findall(r"0x([0-9a-fA-F]+)u", g)] for g in groups]


def bits(a: np.ndarray) -> list[int]:
    return [int(v) for v in np.ascontiguousarray(a, F).view(np.uint32).ravel()]


def main() -> None:
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    header = (root / "tests" / "butteraugli_goldens_generated.h").read_text()
    maps = parse_arrays(header, "kFullMapBits =")
    scores = [int(v, 16) for v in re.findall(
        r"0x([0-9a-fA-F]+)ull", header[header.index("kFullMapScoreBits"):])]
    stages = parse_arrays(header, "kIntermediateStageBits =")
    assert len(maps) == 4 and all(len(m) == 32 * 24 for m in maps)
    assert len(stages) == 27 and all(len(s) == 16 * 12 for s in stages)
    full = []
    for index, kind in enumerate(("flat", "texture", "contrast", "chromatic")):
        ref, dist = fill(kind, 32, 24)
        full.append({"kind": kind, "width": 32, "height": 24,
                     "reference_bits": bits(ref), "distorted_bits": bits(dist),
                     "map_bits": maps[index], "score_bits": scores[index]})
    ref, dist = fill("intermediate", 16, 12)
    out = {
        "_about": (
            "Butteraugli known-answer vectors of the reference's tests "
            "(tests/butteraugli_goldens_generated.h: libjxl e8ff0976 scalar "
            "outputs; inputs from tests/butteraugli_fixtures.cpp FillPattern). "
            "Data only; written by make_butteraugli_goldens.py.  Options: "
            "hf_asymmetry 1, x_multiplier 1, intensity_target 80.  Stage order: "
            "blur sigma 1.2, 7.156, 3.225, 1.564, 2.7 of reference plane 0; "
            "opsin X,Y,B; LF X,Y,B; MF X,Y,B; HF X,Y; UHF X,Y (reference "
            "image); Malta MF Y, MF X, HF Y, HF X, UHF Y, UHF X; mask; masked "
            "AC Y; final composition."),
        "full_maps": full,
        "intermediate": {"width": 16, "height": 12, "reference_bits": bits(ref),
                         "distorted_bits": bits(dist), "stage_bits": stages},
    }
    OUT.write_text(json.dumps(out, separators=(",", ":")))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
