"""ctypes binding of the CPU checker.  TEST INFRASTRUCTURE ONLY.

``oracle()`` loads liboracle.so (the plain-C restatement in this directory);
``reference()`` loads _ref/libgjxlref.so (the reference encoder's own CPU
sources compiled in place, plus ref_harness.cpp) and returns None when it has
not been built.  Only tests/, the smoke check and the CPU-baseline leg of
bench.py may use this module.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
ORACLE_LIB = ORACLE_DIR / "liboracle.so"
REF_LIB = ORACLE_DIR / "_ref" / "libgjxlref.so"

_P = ctypes.c_void_p
_F = ctypes.c_float
_U32 = ctypes.c_uint32
_SZ = ctypes.c_size_t
_I = ctypes.c_int


def build(reference: bool = True) -> None:
    target = "all" if reference else "oracle"
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR), "-j8", target],
                   check=True)


_oracle = None
_ref = None


def oracle() -> ctypes.CDLL:
    global _oracle
    if _oracle is None:
        if not ORACLE_LIB.exists():
            build(reference=False)
        lib = ctypes.CDLL(str(ORACLE_LIB))
        lib.oracle_srgb_to_linear_table.argtypes = [_P]
        lib.oracle_srgb_to_linear_table.restype = None
        lib.oracle_srgb8_to_linear.argtypes = [_P, _SZ, _U32, _U32, _U32, _P]
        lib.oracle_srgb8_to_linear.restype = None
        lib.oracle_fast_cube_root_and_add.argtypes = [_F, _F]
        lib.oracle_fast_cube_root_and_add.restype = _F
        lib.oracle_linear_to_padded_opsin.argtypes = [
            _P, _U32, _U32, _F, _U32, _U32, _P]
        lib.oracle_linear_to_padded_opsin.restype = None
        lib.oracle_opsin_to_linear.argtypes = [_P, _SZ, _F, _P]
        lib.oracle_opsin_to_linear.restype = None
        for name in ("oracle_fast_log2", "oracle_fast_pow2"):
            getattr(lib, name).argtypes = [_F]
            getattr(lib, name).restype = _F
        lib.oracle_fast_pow.argtypes = [_F, _F]
        lib.oracle_fast_pow.restype = _F
        lib.oracle_mirror.argtypes = [ctypes.c_longlong, _SZ]
        lib.oracle_mirror.restype = _SZ
        lib.oracle_convolve_symmetric5.argtypes = [_P, _SZ, _SZ, _P, _P]
        lib.oracle_convolve_symmetric5.restype = None
        lib.oracle_gaborish_inverse_weights.argtypes = [_F, _P]
        lib.oracle_gaborish_inverse_weights.restype = None
        lib.oracle_gaborish_inverse.argtypes = [_P, _SZ, _SZ, _P, _P]
        lib.oracle_gaborish_inverse.restype = None
        lib.oracle_initial_quant_field.argtypes = [
            _P, _SZ, _SZ, _F, _F, _P, _P, _P]
        lib.oracle_initial_quant_field.restype = _I
        lib.oracle_initial_quant_dc.argtypes = [_F]
        lib.oracle_initial_quant_dc.restype = _F
        lib.oracle_quant_matrix.argtypes = [_I, _I, _P, _P]
        lib.oracle_quant_matrix.restype = _I
        lib.oracle_forward_dct.argtypes = [_I, _P, _P]
        lib.oracle_forward_dct.restype = _I
        lib.oracle_inverse_dct.argtypes = [_I, _P, _P]
        lib.oracle_inverse_dct.restype = _I
        lib.oracle_find_best_multiplier.argtypes = [_P, _P, _SZ, _F, _I, _U32]
        lib.oracle_find_best_multiplier.restype = ctypes.c_int8
        lib.oracle_initial_cfl.argtypes = [_P, _SZ, _SZ, _P, _P]
        lib.oracle_initial_cfl.restype = _I
        lib.oracle_ac_strategy_cost.argtypes = [
            _I, _SZ, _SZ, _P, _SZ, _SZ, _P, _P, _F, _F, _F, _F,
            ctypes.POINTER(_F)]
        lib.oracle_ac_strategy_cost.restype = _I
        lib.oracle_stage_index.argtypes = [_I]
        lib.oracle_stage_index.restype = _I
        lib.oracle_candidate_costs.argtypes = [_P, _SZ, _SZ, _P, _P, _P, _P,
                                               _F, _P]
        lib.oracle_candidate_costs.restype = _I
        lib.oracle_search_from_costs.argtypes = [_P, _SZ, _SZ, _F, _P]
        lib.oracle_search_from_costs.restype = None
        _oracle = lib
    return _oracle


def reference():
    global _ref
    if _ref is None and REF_LIB.exists():
        lib = ctypes.CDLL(str(REF_LIB))
        lib.ref_srgb8_to_linear.argtypes = [_P, _SZ, _SZ, _U32, _U32, _U32, _P]
        lib.ref_forward_dct.argtypes = [_I, _P, _P]
        lib.ref_inverse_dct.argtypes = [_I, _P, _P]
        lib.ref_gaborish_inverse.argtypes = [_P, _SZ, _SZ, _P, _P]
        lib.ref_quant_matrix.argtypes = [_I, _I, _P, _P, ctypes.POINTER(_SZ)]
        lib.ref_initial_cfl.argtypes = [_P, _SZ, _SZ, _P, _P]
        lib.ref_ac_strategy_cost.argtypes = [
            _I, _SZ, _SZ, _P, _SZ, _SZ, _P, _P, _F, _F, _F, _F,
            ctypes.POINTER(_F)]
        lib.ref_find_ac_strategy_grid.argtypes = [_P, _SZ, _SZ, _P, _P, _F, _P]
        _ref = lib
    return _ref


def ptr(array: np.ndarray) -> int:
    assert array.flags["C_CONTIGUOUS"]
    return array.ctypes.data


def pad8(v: int) -> int:
    return (v + 7) // 8 * 8


# ---- convenience wrappers (numpy in, numpy out) -------------------------

def srgb8_to_linear(srgb: np.ndarray) -> np.ndarray:
    h, w, bpp = srgb.shape
    srgb = np.ascontiguousarray(srgb)
    out = np.empty((3, h, w), np.float32)
    oracle().oracle_srgb8_to_linear(ptr(srgb), w * bpp, bpp, w, h, ptr(out))
    return out


def linear_to_padded_opsin(linear: np.ndarray,
                           intensity_target: float = 255.0) -> np.ndarray:
    _, h, w = linear.shape
    linear = np.ascontiguousarray(linear, np.float32)
    out = np.empty((3, pad8(h), pad8(w)), np.float32)
    oracle().oracle_linear_to_padded_opsin(
        ptr(linear), w, h, intensity_target, pad8(w), pad8(h), ptr(out))
    return out


def opsin_to_linear(opsin: np.ndarray,
                    intensity_target: float = 255.0) -> np.ndarray:
    opsin = np.ascontiguousarray(opsin, np.float32)
    count = opsin[0].size
    out = np.empty_like(opsin)
    oracle().oracle_opsin_to_linear(ptr(opsin), count, intensity_target,
                                    ptr(out))
    return out


def gaborish_inverse(opsin: np.ndarray, multipliers=(1.0, 1.0, 1.0)) -> np.ndarray:
    opsin = np.ascontiguousarray(opsin, np.float32)
    _, h, w = opsin.shape
    out = np.empty_like(opsin)
    mult = np.asarray(multipliers, np.float32)
    oracle().oracle_gaborish_inverse(ptr(opsin), w, h, ptr(mult), ptr(out))
    return out


def initial_quant_field(opsin: np.ndarray, target: float = 1.0,
                        rescale: float = 1.0) -> dict:
    opsin = np.ascontiguousarray(opsin, np.float32)
    _, h, w = opsin.shape
    q = np.empty((h // 8, w // 8), np.float32)
    s = np.empty_like(q)
    m = np.empty((h, w), np.float32)
    rc = oracle().oracle_initial_quant_field(ptr(opsin), w, h, target, rescale,
                                             ptr(q), ptr(s), ptr(m))
    if rc != 0:
        raise ValueError(f"oracle initial quant field failed ({rc})")
    return {"quant_field": q, "strategy_mask": s, "pixel_mask": m}


STAGE_STRATEGIES = (0, 6, 7, 4, 10, 11, 5)
STAGE_ENTROPY = (1.0, 1.21, 1.21, 1.34, 1.49, 1.49, 1.48)
COVERED = {0: (1, 1), 4: (2, 2), 5: (4, 4), 6: (1, 2), 7: (2, 1), 10: (2, 4),
           11: (4, 2)}


def quant_matrix(strategy: int, channel: int):
    dq = np.empty(1024, np.float32)
    inv = np.empty(1024, np.float32)
    n = oracle().oracle_quant_matrix(strategy, channel, ptr(dq), ptr(inv))
    return dq[:n].copy(), inv[:n].copy()


def initial_cfl(opsin: np.ndarray):
    opsin = np.ascontiguousarray(opsin, np.float32)
    _, h, w = opsin.shape
    tw, th = (w + 63) // 64, (h + 63) // 64
    x = np.empty((th, tw), np.int8)
    b = np.empty((th, tw), np.int8)
    assert oracle().oracle_initial_cfl(ptr(opsin), w, h, ptr(x), ptr(b)) == 0
    return x, b


def candidate_costs(opsin, quant_field, pixel_mask, y_to_x, y_to_b,
                    target: float = 1.0) -> np.ndarray:
    opsin = np.ascontiguousarray(opsin, np.float32)
    _, h, w = opsin.shape
    costs = np.empty((7, h // 8, w // 8), np.float32)
    args = [np.ascontiguousarray(a) for a in
            (quant_field, pixel_mask, y_to_x, y_to_b)]
    rc = oracle().oracle_candidate_costs(ptr(opsin), w, h, *map(ptr, args),
                                         target, ptr(costs))
    assert rc == 0
    return costs


def search_from_costs(costs: np.ndarray, target: float = 1.0) -> np.ndarray:
    costs = np.ascontiguousarray(costs, np.float32)
    _, bh, bw = costs.shape
    cells = np.empty((bh, bw), np.uint8)
    oracle().oracle_search_from_costs(ptr(costs), bw, bh, target, ptr(cells))
    return cells
