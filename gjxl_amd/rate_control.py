"""f3: the target-size search (gjxl_amd_rate_control_search), as the
reference's SearchTargetSize (src/codestream/rate_control.cpp:246-337).

`search(options, evaluate)` runs the library's search with a Python
evaluator; `search_plain_path(path, image, target_bytes, ...)` evaluates
candidates with the device plain path at each Butteraugli target and sizes
each encode by the entropy of its token stream given the contexts (an ideal
coder with static per-context statistics).  ANS and the bitstream are not
built here, so that size stands in for the file size; the search itself is
the reference's.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import GjxlError, _check, _F, _SZ, _U32, library

LARGEST_AT_OR_BELOW, CLOSEST_ABSOLUTE = 0, 1


class Options(ctypes.Structure):
    _fields_ = [("struct_size", _SZ), ("target_bytes", ctypes.c_uint64),
                ("tolerance_bytes", ctypes.c_uint64), ("maximum_attempts", _U32),
                ("selection", ctypes.c_int32), ("minimum_target", _F), ("maximum_target", _F)]


class Candidate(ctypes.Structure):
    _fields_ = [("encoded_bytes", ctypes.c_uint64), ("score_evaluated", ctypes.c_int32),
                ("resident", ctypes.c_int32), ("final_score", ctypes.c_double)]


class Result(ctypes.Structure):
    _fields_ = [("selected_target", _F), ("selected", Candidate), ("attempts", _U32),
                ("failed_attempts", _U32), ("target_size_met", ctypes.c_int32),
                ("search_exhausted", ctypes.c_int32)]


EVALUATOR = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.c_float,
                             ctypes.POINTER(Candidate))


def options(target_bytes: int, **kw) -> Options:
    o = Options()
    _check(library().gjxl_amd_rate_control_options_init(ctypes.byref(o), ctypes.sizeof(o)))
    o.target_bytes = target_bytes
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def search(opts: Options, evaluate, result: Result | None = None) -> Result:
    """evaluate(target) -> (code, encoded_bytes, score_evaluated, resident,
    final_score); a non-zero code is that candidate's failure.  Raises
    GjxlError when the search fails (result then unchanged)."""
    errors = []

    def call(_user, target, out):
        try:
            code, size, evaluated, resident, score = evaluate(float(target))
        except Exception as e:  # an exception is a candidate-local failure
            errors.append(e)
            return 5
        out[0] = Candidate(size, int(evaluated), int(resident), score)
        return code

    fn = EVALUATOR(call)
    out = result if result is not None else Result()
    _check(library().gjxl_amd_rate_control_search(ctypes.byref(opts), fn, None,
                                                   ctypes.byref(out)))
    return out


def token_bytes(tokens: dict) -> int:
    """Bytes of an ideal coder with static per-context statistics of the
    token values."""
    v = tokens["values"].astype(np.uint64)
    c = tokens["contexts"].astype(np.uint64)
    if v.size == 0:
        return 1
    _, cnt = np.unique((c << np.uint64(32)) | v, return_counts=True)
    _, ccnt = np.unique(c, return_counts=True)
    bits = float(-(cnt * np.log2(cnt)).sum() + (ccnt * np.log2(ccnt)).sum())
    return max(1, int(math.ceil(bits / 8.0)))


def search_plain_path(path, image: np.ndarray, target_bytes: int, **kw) -> tuple[Result, list]:
    """The search with the device plain path as the encoder: each candidate
    re-targets `path` to its Butteraugli target and sizes the tokens; the
    path's Butteraugli loop runs on the device, so its score is a
    diagnostic (resident).  Returns (result, [(target, bytes), ...])."""
    h, w = image.shape[:2]
    tried = []

    def evaluate(target):
        path.reconfigure(w, h, distance=target)
        path.run(image)
        size = token_bytes(path.tokens())
        tried.append((target, size))
        return 0, size, 1, 1, float(path.scores()[-1])

    return search(options(target_bytes, **kw), evaluate), tried
