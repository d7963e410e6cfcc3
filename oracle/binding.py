"""ctypes binding of the CPU checker.  TEST INFRASTRUCTURE ONLY.

``oracle()`` loads liboracle.so (the plain-C restatement in this directory);
``reference()`` loads _ref/libgjxlref.so (the reference encoder's own CPU
sources compiled in place, plus ref_harness.cpp) and returns None when it has
not been built; only the CPU tests call it, through the ``ref`` fixture, so
no GPU process maps it.  Only tests/, the smoke check and the CPU-baseline
leg of bench.py may use this module.
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
        _declare_reference_frame(lib)
        _ref = lib
    return _ref


def _declare_reference_frame(ref) -> None:
    """Signatures of ref_harness.cpp's frame, loop-filter and Butteraugli
    entry points."""
    P, F_, SZ, I, U32 = _P, _F, _SZ, _I, _U32
    D = ctypes.POINTER(ctypes.c_double)
    rsig = {
        "ref_quantizer_from_field": [F_, P, SZ, SZ, P, P],
        "ref_final_cfl": [P, SZ, SZ, P, P, U32, U32, I, P, P],
        "ref_code_transform": [I, U32, U32, ctypes.c_int32, F_, F_, F_, F_, P, P, P],
        "ref_dequantize": [I, U32, U32, ctypes.c_int32, I, F_, P, P],
        "ref_llf_to_dc": [I, P, P],
        "ref_dc_to_llf": [I, P, P],
        "ref_quantize_dc": [P, SZ, SZ, U32, U32, P, P],
        "ref_smooth_dc": [P, SZ, SZ, U32, U32, P],
        "ref_epf_inverse_sigma": [P, SZ, SZ, P, U32, U32, P],
        "ref_loop_filters": [P, SZ, SZ, P, P],
        "ref_butteraugli_distance": [P, P, SZ, SZ, F_, F_, F_, P, D],
        "ref_butteraugli_prepared": [P, P, SZ, SZ, SZ, F_, F_, F_, P, P],
        "ref_butteraugli_blur": [P, SZ, SZ, F_, P],
        "ref_butteraugli_opsin": [P, SZ, SZ, F_, P],
        "ref_butteraugli_frequencies": [P, SZ, SZ, P],
        "ref_butteraugli_stages": [P, P, SZ, SZ, F_, F_, F_, P],
        "ref_reduce_distance_map": [P, SZ, SZ, P, SZ, SZ, P],
    }
    for name, args in rsig.items():
        fn = getattr(ref, name)
        fn.argtypes = args
        fn.restype = I


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


# ---- a11-a15: AQ evaluation without Butteraugli ---------------------------

_AQ_DECLARED = False


def _declare_aq():
    global _AQ_DECLARED
    if _AQ_DECLARED:
        return
    lib = oracle()
    P, F_, SZ, I = _P, _F, _SZ, _I
    sig = {
        "oracle_matrix_scale_stats": ([P, SZ, SZ, SZ, SZ, P], None),
        "oracle_select_matrix_scales": ([P, F_, P, P], None),
        "oracle_matrix_multiplier": ([I], F_),
        "oracle_adjust_quant_field": ([P, SZ, SZ, F_, P, P], None),
        "oracle_policy_setup": ([P, SZ, F_, P, P, P], I),
        "oracle_quantizer_from_field": ([F_, P, SZ, P, P], I),
        "oracle_final_cfl_fast": ([P, P, SZ, SZ, P, P, F_, P, P], None),
        "oracle_llf_to_dc": ([I, I, P, P], None),
        "oracle_dc_to_llf": ([I, I, P, P], None),
        "oracle_dequantize_ac": ([F_, ctypes.c_int32, F_, I, I, P, P, P], None),
        "oracle_code_transform": ([I, P, ctypes.c_int32, F_, F_, F_, F_, P, P, P, P], None),
        "oracle_quantize_dc": ([P, SZ, SZ, P, P, I, P, P], None),
        "oracle_smooth_dc": ([P, SZ, SZ, P, P], None),
        "oracle_anchor_offsets": ([P, SZ, SZ, P], None),
        "oracle_anchor_coefficients": ([P, SZ, SZ, P, P, P], None),
        "oracle_code_coefficients": ([P, P, SZ, SZ, P, P, P, P, P, F_, F_, P, P, P], None),
        "oracle_epf_inverse_sigma": ([P, SZ, SZ, P, F_, P], None),
        "oracle_reconstruct": ([P, P, SZ, SZ, P, P, P, P, P, F_, F_, P, P], None),
        "oracle_gaborish": ([P, SZ, SZ, P], None),
        "oracle_loop_filters": ([P, SZ, SZ, P, P], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _AQ_DECLARED = True


def store_size(bw: int, bh: int) -> int:
    return ((bw + 31) // 32) * ((bh + 31) // 32) * 65536


def quantizer_from_field(field: np.ndarray, quant_dc: float):
    _declare_aq()
    field = np.ascontiguousarray(field, np.float32)
    raw = np.empty(field.shape, np.int32)
    q = np.empty(10, np.float32)
    oracle().oracle_quantizer_from_field(quant_dc, ptr(field), field.size,
                                         ptr(raw), ptr(q))
    return raw, q


def frontend(srgb: np.ndarray, target: float = 1.0) -> dict:
    """The oracle's plain frontend: XYB, Gaborish inverse, initial quant
    field, initial CfL, candidate costs and the strategy grid."""
    lin = srgb8_to_linear(srgb)
    ops = linear_to_padded_opsin(lin)
    pre = gaborish_inverse(ops)
    aq = initial_quant_field(ops, target, 1.0)
    cx, cb = initial_cfl(pre)
    costs = candidate_costs(pre, aq["quant_field"], aq["pixel_mask"], cx, cb,
                            target)
    cells = search_from_costs(costs, target)
    return {"linear": lin, "opsin": ops, "pre": pre, "cells": cells,
            "costs": costs, **aq, "init_y_to_x": cx, "init_y_to_b": cb}


def aq_setup(fe: dict, width: int, height: int, target: float = 1.0) -> dict:
    """Per-encode AQ setup: matrix scales, adjusted field, bounds, anchor
    layout and the anchors' forward coefficients."""
    _declare_aq()
    lib = oracle()
    ops = fe["opsin"]
    _, ph, pw = ops.shape
    bh, bw = ph // 8, pw // 8
    stats = np.empty(3, np.float32)
    lib.oracle_matrix_scale_stats(ptr(ops), width, height, pw, pw * ph,
                                  ptr(stats))
    xs, bs = ctypes.c_int(), ctypes.c_int()
    lib.oracle_select_matrix_scales(ptr(stats), target, ctypes.byref(xs),
                                    ctypes.byref(bs))
    cells = np.ascontiguousarray(fe["cells"])
    adjusted = np.empty((bh, bw), np.float32)
    lib.oracle_adjust_quant_field(ptr(cells), bw, bh, target,
                                  ptr(np.ascontiguousarray(fe["quant_field"])),
                                  ptr(adjusted))
    lo, hi, qdc = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    lib.oracle_policy_setup(ptr(adjusted), adjusted.size, target,
                            ctypes.byref(lo), ctypes.byref(hi),
                            ctypes.byref(qdc))
    offsets = np.empty((bh, bw), np.uint64)
    lib.oracle_anchor_offsets(ptr(cells), bw, bh, ptr(offsets))
    coefs = np.zeros((3, store_size(bw, bh)), np.float32)
    lib.oracle_anchor_coefficients(ptr(np.ascontiguousarray(fe["pre"])), pw,
                                   ph, ptr(cells), ptr(offsets), ptr(coefs))
    return {"stats": stats, "x_scale": xs.value, "b_scale": bs.value,
            "x_mult": lib.oracle_matrix_multiplier(xs.value),
            "b_mult": lib.oracle_matrix_multiplier(bs.value),
            "adjusted": adjusted, "lower": lo.value, "upper": hi.value,
            "quant_dc": qdc.value, "offsets": offsets, "coefs": coefs}


def aq_evaluate(fe: dict, setup: dict, field: np.ndarray, width: int,
                height: int) -> dict:
    """One evaluation up to the loop filters (no Butteraugli)."""
    _declare_aq()
    lib = oracle()
    cells = np.ascontiguousarray(fe["cells"])
    bh, bw = cells.shape
    raw, q = quantizer_from_field(field, setup["quant_dc"])
    coefs = setup["coefs"]
    planes = (ctypes.c_void_p * 3)(*(coefs[k].ctypes.data for k in range(3)))
    tiles = ((bh + 7) // 8, (bw + 7) // 8)
    fx = np.empty(tiles, np.int8)
    fb = np.empty(tiles, np.int8)
    lib.oracle_final_cfl_fast(planes, ptr(cells), bw, bh,
                              ptr(setup["offsets"]), ptr(raw), float(q[0]),
                              ptr(fx), ptr(fb))
    field_raw = raw.copy()
    quantized = np.zeros((3, store_size(bw, bh)), np.int32)
    dc_q = np.empty((3, bh, bw), np.int32)
    dc_rec = np.empty((3, bh, bw), np.float32)
    lib.oracle_code_coefficients(ptr(coefs), ptr(cells), bw, bh,
                                 ptr(setup["offsets"]), ptr(q), ptr(raw),
                                 ptr(fx), ptr(fb), setup["x_mult"],
                                 setup["b_mult"], ptr(quantized), ptr(dc_q),
                                 ptr(dc_rec))
    inv_sigma = np.empty((bh, bw), np.float32)
    lib.oracle_epf_inverse_sigma(ptr(cells), bw, bh, ptr(raw), float(q[0]),
                                 ptr(inv_sigma))
    rec = np.empty((3, bh * 8, bw * 8), np.float32)
    lib.oracle_reconstruct(ptr(quantized), ptr(cells), bw * 8, bh * 8,
                           ptr(setup["offsets"]), ptr(q), ptr(raw), ptr(fx),
                           ptr(fb), setup["x_mult"], setup["b_mult"],
                           ptr(dc_rec), ptr(rec))
    cropped = np.ascontiguousarray(rec[:, :height, :width])
    filtered = np.empty_like(cropped)
    lib.oracle_loop_filters(ptr(cropped), width, height, ptr(inv_sigma),
                            ptr(filtered))
    linear = opsin_to_linear(filtered)
    return {"q": q, "field_raw": field_raw, "raw": raw, "y_to_x": fx,
            "y_to_b": fb, "quantized": quantized, "dc_q": dc_q,
            "dc_rec": dc_rec, "inv_sigma": inv_sigma, "reconstructed": rec,
            "filtered": filtered, "linear": linear}


# ---- a16-a18: Butteraugli, 16-norm reduction, AQ update -------------------

_B_DECLARED = False


def _declare_butteraugli():
    global _B_DECLARED
    if _B_DECLARED:
        return
    lib = oracle()
    P, F_, SZ, I = _P, _F, _SZ, _I
    sig = {
        "oracle_butteraugli_kernel": ([F_, P], I),
        "oracle_butteraugli_blur": ([P, SZ, SZ, F_, P], I),
        "oracle_butteraugli_opsin": ([P, SZ, SZ, F_, P], I),
        "oracle_butteraugli_frequencies": ([P, SZ, SZ, P], I),
        "oracle_butteraugli_erosion": ([P, SZ, SZ, P], None),
        "oracle_butteraugli_differences": ([P, P, SZ, SZ, F_, F_, P], I),
        "oracle_butteraugli_subsample": ([P, SZ, SZ, P], None),
        "oracle_butteraugli_distance": ([P, P, SZ, SZ, F_, F_, F_, P,
                                         ctypes.POINTER(ctypes.c_double)], I),
        "oracle_reduce_distance_map": ([P, SZ, SZ, P, SZ, SZ, P], None),
        "oracle_aq_update": ([P, P, P, SZ, F_, I, F_, F_, F_, F_], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _B_DECLARED = True


def _f32(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, np.float32)


def butteraugli_blur(plane: np.ndarray, sigma: float) -> np.ndarray:
    _declare_butteraugli()
    plane = _f32(plane)
    h, w = plane.shape
    out = np.empty_like(plane)
    assert oracle().oracle_butteraugli_blur(ptr(plane), w, h, sigma, ptr(out)) == 0
    return out


def butteraugli_opsin(rgb: np.ndarray, intensity: float = 80.0) -> np.ndarray:
    _declare_butteraugli()
    rgb = _f32(rgb)
    _, h, w = rgb.shape
    out = np.empty_like(rgb)
    assert oracle().oracle_butteraugli_opsin(ptr(rgb), w, h, intensity, ptr(out)) == 0
    return out


def butteraugli_frequencies(xyb: np.ndarray) -> np.ndarray:
    _declare_butteraugli()
    xyb = _f32(xyb)
    _, h, w = xyb.shape
    out = np.empty((10, h, w), np.float32)
    assert oracle().oracle_butteraugli_frequencies(ptr(xyb), w, h, ptr(out)) == 0
    return out


def butteraugli_stages(reference_rgb: np.ndarray, distorted_rgb: np.ndarray,
                       hf_asymmetry: float = 1.0, x_multiplier: float = 1.0,
                       intensity: float = 80.0) -> np.ndarray:
    """The 9 difference-stage planes of one scale (no subscale)."""
    x0 = butteraugli_opsin(reference_rgb, intensity)
    x1 = butteraugli_opsin(distorted_rgb, intensity)
    p0, p1 = butteraugli_frequencies(x0), butteraugli_frequencies(x1)
    _, h, w = x0.shape
    out = np.empty((9, h, w), np.float32)
    assert oracle().oracle_butteraugli_differences(
        ptr(p0), ptr(p1), w, h, hf_asymmetry, x_multiplier, ptr(out)) == 0
    return out


def butteraugli_distance(reference_rgb: np.ndarray, distorted_rgb: np.ndarray,
                         hf_asymmetry: float = 1.0, x_multiplier: float = 1.0,
                         intensity: float = 80.0):
    _declare_butteraugli()
    r, d = _f32(reference_rgb), _f32(distorted_rgb)
    _, h, w = r.shape
    out = np.empty((h, w), np.float32)
    score = ctypes.c_double()
    assert oracle().oracle_butteraugli_distance(
        ptr(r), ptr(d), w, h, hf_asymmetry, x_multiplier, intensity, ptr(out),
        ctypes.byref(score)) == 0
    return out, score.value


def reduce_distance_map(distance_map: np.ndarray, cells: np.ndarray) -> np.ndarray:
    _declare_butteraugli()
    m = _f32(distance_map)
    cells = np.ascontiguousarray(cells, np.uint8)
    h, w = m.shape
    bh, bw = cells.shape
    out = np.empty((bh, bw), np.float32)
    oracle().oracle_reduce_distance_map(ptr(m), w, h, ptr(cells), bw, bh, ptr(out))
    return out


def aq_update(field: np.ndarray, adjusted_initial: np.ndarray,
              block_distance: np.ndarray, iteration: int, lower: float,
              upper: float, q: np.ndarray, target: float = 1.0) -> np.ndarray:
    """One field update after evaluation `iteration`; q is the evaluation's
    quantizer as oracle_quantizer_from_field fills it."""
    _declare_butteraugli()
    out = _f32(field).copy()
    a0, bd = _f32(adjusted_initial), _f32(block_distance)
    oracle().oracle_aq_update(ptr(out), ptr(a0), ptr(bd), out.size, target,
                              iteration, lower, upper, float(q[0]), float(q[1]))
    return out


def aq_loop(fe: dict, setup: dict, linear: np.ndarray, width: int, height: int,
            iterations: int = 2, target: float = 1.0) -> dict:
    """RunAdaptiveQuantizationPolicyImpl (src/codec/adaptive_quantization.cpp:
    1600-1738): iterations + 1 evaluations with Butteraugli against the source
    linear RGB, field updates between them."""
    field = setup["adjusted"].copy()
    scores, fields, distances = [], [], []
    for it in range(iterations + 1):
        ev = aq_evaluate(fe, setup, field, width, height)
        dmap, score = butteraugli_distance(linear, ev["linear"])
        bd = reduce_distance_map(dmap, fe["cells"])
        scores.append(score)
        fields.append(field.copy())
        distances.append(bd)
        if it == iterations:
            break
        field = aq_update(field, setup["adjusted"], bd, it, setup["lower"],
                          setup["upper"], ev["q"], target)
    return {"field": field, "evaluation": ev, "distance_map": dmap,
            "block_distance": bd, "scores": scores, "fields": fields,
            "block_distances": distances}


# ---- a19: AC tokenization -------------------------------------------------

_T_DECLARED = False
# strategy code of each stage of the order table, and the stage sizes
STAGE_STRATEGIES = (0, 6, 7, 4, 10, 11, 5)
ORDER_TABLE_SIZE = 3 * 2624
# src/codestream/simple_ac_context.h:29-33 (default map, no thresholds)
DEFAULT_CONTEXT_MAP = np.array([0] * 4 + [1] * 9 + ([2] * 4 + [3] * 9) * 2, np.uint8)


def _declare_tokenize():
    global _T_DECLARED
    if _T_DECLARED:
        return
    lib = oracle()
    P, SZ, I, U32 = _P, _SZ, _I, _U32
    sig = {
        "oracle_natural_order": ([I, P], SZ),
        "oracle_stage_strategy": ([I], I),
        "oracle_strategy_stage": ([I], I),
        "oracle_order_table_offset": ([I], SZ),
        "oracle_natural_order_table": ([P], None),
        "oracle_pack_signed": ([ctypes.c_int32], U32),
        "oracle_token_symbol": ([U32], U32),
        "oracle_validate_context_map": ([P, SZ, P, U32], I),
        "oracle_tokenize_group": ([P, SZ, SZ, P, SZ, SZ, SZ, SZ, P, P, P, SZ, P,
                                   U32, P, P, SZ, ctypes.POINTER(ctypes.c_size_t),
                                   P], I),
        "oracle_tokenize_frame": ([P, P, SZ, SZ, P, P, P, SZ, P, U32, P, P, SZ, P,
                                   P], I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _T_DECLARED = True


def natural_order(strategy: int) -> np.ndarray:
    _declare_tokenize()
    out = np.zeros(1024, np.uint32)
    n = oracle().oracle_natural_order(strategy, ptr(out))
    if n == 0:
        raise ValueError(f"no natural order for strategy {strategy}")
    return out[:n].copy()


def natural_order_table() -> np.ndarray:
    _declare_tokenize()
    table = np.zeros(ORDER_TABLE_SIZE, np.uint32)
    oracle().oracle_natural_order_table(ptr(table))
    return table


def order_slice(table: np.ndarray, strategy: int, channel: int) -> np.ndarray:
    """The entries of `table` holding the order of (strategy, channel)."""
    _declare_tokenize()
    lib = oracle()
    stage = lib.oracle_strategy_stage(strategy)
    n = natural_order(strategy).size
    base = 3 * lib.oracle_order_table_offset(stage) + channel * n
    return table[base:base + n]


def token_capacity(cells: np.ndarray) -> int:
    anchors = int(np.count_nonzero(cells & 1))
    return 3 * (64 * cells.size + anchors)


def tokenize_frame(quantized: np.ndarray, cells: np.ndarray, raw: np.ndarray,
                   orders: np.ndarray | None = None, thresholds=(),
                   context_map: np.ndarray | None = None,
                   num_contexts: int | None = None,
                   histogram: bool = False) -> dict:
    """Tokens of every AC group of a frame in the per-group coefficient
    layout (3 planes of groups * 65536 int32)."""
    _declare_tokenize()
    bh, bw = cells.shape
    groups = ((bw + 31) // 32) * ((bh + 31) // 32)
    quantized = np.ascontiguousarray(quantized, np.int32)
    assert quantized.size == 3 * groups * 65536
    cells = np.ascontiguousarray(cells, np.uint8)
    raw = np.ascontiguousarray(raw, np.int32)
    orders = natural_order_table() if orders is None else np.ascontiguousarray(orders, np.uint32)
    thr = np.ascontiguousarray(np.array(thresholds, np.uint32).reshape(-1))
    cmap = DEFAULT_CONTEXT_MAP if context_map is None else np.ascontiguousarray(context_map, np.uint8)
    nctx = int(cmap.max()) + 1 if num_contexts is None else num_contexts
    cap = token_capacity(cells)
    values = np.zeros(max(cap, 1), np.uint32)
    contexts = np.zeros(max(cap, 1), np.uint16)
    offsets = np.zeros(groups + 1, np.uint32)
    hist = np.zeros(nctx * 495 * 128, np.uint32) if histogram else None
    rc = oracle().oracle_tokenize_frame(
        ptr(quantized), ptr(cells), bw, bh, ptr(raw), ptr(orders),
        ptr(thr) if thr.size else None, thr.size, ptr(cmap), nctx, ptr(values),
        ptr(contexts), cap, ptr(offsets), ptr(hist) if histogram else None)
    if rc:
        raise ValueError(f"oracle_tokenize_frame failed with {rc}")
    n = int(offsets[-1])
    out = {"values": values[:n].copy(), "contexts": contexts[:n].copy(),
           "group_offsets": offsets}
    if histogram:
        out["histogram"] = hist.reshape(nctx * 495, 128)
    return out


# ---- f1/f2: effort-7 tokenization inputs -----------------------------------

_E_DECLARED = False


def _declare_entropy_inputs():
    global _E_DECLARED
    if _E_DECLARED:
        return
    _declare_tokenize()
    lib = oracle()
    P, SZ, I = _P, _SZ, _I
    sig = {
        "oracle_order_sample_decisions": ([SZ, SZ, P], None),
        "oracle_frame_statistics": ([P, P, SZ, SZ, P, P, P, P, P, P], I),
        "oracle_coefficient_orders": ([P, P, ctypes.c_uint16, SZ, SZ, I, P, P], I),
        "oracle_block_context_map": ([P, SZ, P, ctypes.POINTER(ctypes.c_size_t), P,
                                      ctypes.POINTER(ctypes.c_uint32)], I),
        "oracle_order_tokens": ([ctypes.c_uint16, P, P, P, SZ,
                                 ctypes.POINTER(ctypes.c_size_t)], I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _E_DECLARED = True


def order_sample_decisions(bw: int, bh: int) -> np.ndarray:
    _declare_entropy_inputs()
    out = np.zeros((bh, bw), np.uint8)
    oracle().oracle_order_sample_decisions(bw, bh, ptr(out))
    return out


def frame_statistics(quantized: np.ndarray, cells: np.ndarray, raw: np.ndarray,
                     decisions: np.ndarray | None = None) -> dict:
    _declare_entropy_inputs()
    bh, bw = cells.shape
    quantized = np.ascontiguousarray(quantized, np.int32)
    cells = np.ascontiguousarray(cells, np.uint8)
    raw = np.ascontiguousarray(raw, np.int32)
    if decisions is None:
        decisions = order_sample_decisions(bw, bh)
    zeros = np.zeros((3, 1984), np.uint32)
    sampled = np.zeros((3, 64), np.uint32)
    raw_counts = np.zeros((13, 256), np.uint32)
    present = ctypes.c_uint16()
    rc = oracle().oracle_frame_statistics(
        ptr(quantized), ptr(cells), bw, bh, ptr(raw), ptr(decisions), ptr(zeros),
        ptr(sampled), ctypes.byref(present), ptr(raw_counts))
    if rc:
        raise ValueError(f"oracle_frame_statistics failed with {rc}")
    return {"zero_counts": zeros, "sampled": sampled, "raw_counts": raw_counts,
            "present_mask": int(present.value), "bw": bw, "bh": bh}


def coefficient_orders(stats: dict, effort7_sampled: bool = True) -> tuple[int, np.ndarray]:
    """(used_order_mask, 7-stage order table)."""
    _declare_entropy_inputs()
    orders = np.zeros(ORDER_TABLE_SIZE, np.uint32)
    used = ctypes.c_uint16()
    rc = oracle().oracle_coefficient_orders(
        ptr(stats["zero_counts"]), ptr(stats["sampled"]), stats["present_mask"],
        stats["bw"], stats["bh"], int(effort7_sampled), ctypes.byref(used), ptr(orders))
    if rc:
        raise ValueError(f"oracle_coefficient_orders failed with {rc}")
    return int(used.value), orders


def block_context_map(raw_counts: np.ndarray, blocks: int) -> dict:
    _declare_entropy_inputs()
    raw_counts = np.ascontiguousarray(raw_counts, np.uint32)
    thr = np.zeros(15, np.uint32)
    nthr = ctypes.c_size_t()
    cmap = np.zeros(3 * 13 * 16, np.uint8)
    nctx = ctypes.c_uint32()
    rc = oracle().oracle_block_context_map(ptr(raw_counts), blocks, ptr(thr),
                                           ctypes.byref(nthr), ptr(cmap), ctypes.byref(nctx))
    if rc:
        raise ValueError(f"oracle_block_context_map failed with {rc}")
    n = int(nthr.value)
    return {"thresholds": thr[:n].copy(), "context_map": cmap[:3 * 13 * (n + 1)].copy(),
            "num_contexts": int(nctx.value)}


def order_tokens(used_mask: int, orders: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    _declare_entropy_inputs()
    cap = 3 * 1984 + 15
    ctx = np.zeros(cap, np.uint32)
    val = np.zeros(cap, np.uint32)
    n = ctypes.c_size_t()
    orders = np.ascontiguousarray(orders, np.uint32)
    rc = oracle().oracle_order_tokens(used_mask, ptr(orders), ptr(ctx), ptr(val), cap,
                                      ctypes.byref(n))
    if rc:
        raise ValueError(f"oracle_order_tokens failed with {rc}")
    return ctx[:n.value].copy(), val[:n.value].copy()


def effort7_tokens(quantized: np.ndarray, cells: np.ndarray, raw: np.ndarray) -> dict:
    """The frame's tokens with the coefficient orders and block-context map
    that effort 7 chooses from its statistics (the reference's serializer,
    src/codestream/vardct/encoder.cpp:1985-2010), plus those inputs."""
    stats = frame_statistics(quantized, cells, raw)
    used, orders = coefficient_orders(stats, True)
    bh, bw = cells.shape
    model = block_context_map(stats["raw_counts"], bw * bh)
    tokens = tokenize_frame(quantized, cells, raw, orders=orders,
                            thresholds=model["thresholds"], context_map=model["context_map"],
                            num_contexts=model["num_contexts"])
    tokens.update({"used_order_mask": used, "orders": orders, "statistics": stats, **model})
    return tokens

