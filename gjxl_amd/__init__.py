"""gjxl_amd: MI355X-native perceptual-analysis path of the gjxl encoder.

Python host mirror over the C ABI in include/gjxl_amd.h.  The library
(libgjxl_amd.so, built by gjxl_amd.build) holds every kernel and the C++
host orchestration; this module only binds it with ctypes and moves numpy
arrays to and from device memory.  There is no CPU fallback: if the library
is missing or no HIP device is present, calls raise.
"""
from __future__ import annotations

import ctypes
import os
import weakref
from pathlib import Path

import numpy as np

PACKAGE_DIR = Path(__file__).resolve().parent
LIBRARY_PATH = PACKAGE_DIR / "libgjxl_amd.so"

GJXL_OK = 0
_ERROR_NAMES = {
    1: "INVALID_ARGUMENT", 2: "UNSUPPORTED", 3: "UNAVAILABLE",
    4: "OUT_OF_MEMORY", 5: "BACKEND", 6: "INTERNAL",
    7: "RESOURCE_PLAN_EXCEEDED",
}


class GjxlError(RuntimeError):
    """A non-OK GJXLResult, with the library's thread-local diagnostic."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{_ERROR_NAMES.get(code, code)}: {message}")
        self.code = code


_lib = None


def library() -> ctypes.CDLL:
    """Loads libgjxl_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIBRARY_PATH.exists():
            raise RuntimeError(
                f"{LIBRARY_PATH} is missing: run `python -m gjxl_amd.build`")
        _lib = ctypes.CDLL(str(LIBRARY_PATH))
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_SZ = ctypes.c_size_t
_F = ctypes.c_float
_I = ctypes.c_int

# name -> argument types (all return GJXLResult unless listed in _VOID/_STR)
_SIGNATURES = {
    "gjxl_amd_device_create": [_I, ctypes.POINTER(_P)],
    "gjxl_amd_device_destroy": [_P],
    "gjxl_amd_alloc": [_P, _SZ, ctypes.POINTER(_P)],
    "gjxl_amd_free": [_P, _P],
    "gjxl_amd_upload": [_P, _P, _P, _SZ],
    "gjxl_amd_download": [_P, _P, _P, _SZ],
    "gjxl_amd_copy": [_P, _P, _P, _SZ],
    "gjxl_amd_synchronize": [_P],
    "gjxl_amd_last_error": [],
    "gjxl_amd_event_record": [_P, _I],
    "gjxl_amd_event_elapsed": [_P, _I, _I, ctypes.POINTER(_F)],
    "gjxl_amd_device_probe": [_P, ctypes.c_int32],
    "gjxl_amd_device_probe_times": [_P, _P, ctypes.c_int32,
                                    ctypes.POINTER(ctypes.c_int32)],
    "gjxl_amd_srgb8_to_linear_and_opsin":
        [_P, _P, _SZ, _U32, _U32, _U32, _F, _P, _P],
    "gjxl_amd_linear_rgb_to_padded_opsin": [_P, _P, _U32, _U32, _F, _P],
    "gjxl_amd_gaborish_inverse": [_P, _P, _U32, _U32, _P, _P],
    "gjxl_amd_initial_quant_field_scratch_bytes": [_U32, _U32],
    "gjxl_amd_initial_quant_field":
        [_P, _P, _U32, _U32, _F, _F, _P, _P, _P, _P],
    "gjxl_amd_initial_quant_field_no_alloc":
        [_P, _P, _U32, _U32, _F, _F, _P, _P, _P],
    "gjxl_amd_initial_cfl": [_P, _P, _U32, _U32, _P, _P],
    "gjxl_amd_ac_strategy_costs":
        [_P, _P, _U32, _U32, _P, _P, _P, _P, _F, _P],
    "gjxl_amd_ac_strategy_search": [_P, _P, _U32, _U32, _F, _P],
    "gjxl_amd_ac_strategy_costs_pruned":
        [_P, _P, _U32, _U32, _P, _P, _P, _P, _F, _P, _P],
    "gjxl_amd_sparse_ac": [_P, _P, _U32, _U32, _P, _P, _P, _P],
    "gjxl_amd_rate_control_options_init": [_P, _SZ],
    "gjxl_amd_rate_control_search": [_P, _P, _P, _P],
    "gjxl_amd_cost_queue_create": [_P, _U32, ctypes.POINTER(_P)],
    "gjxl_amd_cost_queue_destroy": [_P],
    "gjxl_amd_cost_queue_run": [_P, _P, _U32],
    "gjxl_amd_cost_queue_tune": [_P, _P, _U32, _U32],
    "gjxl_amd_cost_queue_grids": [_P, _P, _P],
    "gjxl_amd_proposal_parameter_count": [],
    "gjxl_amd_proposal_create": [_P, _SZ, ctypes.POINTER(_P)],
    "gjxl_amd_proposal_destroy": [_P],
    "gjxl_amd_proposal_set_parameters": [_P, _P, _SZ],
    "gjxl_amd_proposal_get_parameters": [_P, _P, _SZ],
    "gjxl_amd_proposal_buffer": [_P, ctypes.c_int32, ctypes.POINTER(_P),
                                 ctypes.POINTER(_SZ)],
    "gjxl_amd_proposal_features": [_P, _P, _U32, _U32, _P, _P],
    "gjxl_amd_proposal_forward": [_P, _P, _U32, _U32, _U32],
    "gjxl_amd_proposal_backward": [_P, _P, _P, _F, _P],
    "gjxl_amd_proposal_adam": [_P, _F, _F, _F, _F, _F],
    "gjxl_amd_proposal_decide": [_P, _P, _U32, _U32, _F, _U32, _F, _P],
    "gjxl_amd_proposal_warm_start": [_P, _P, _P, _SZ, _F, _F, _F, _F, _P],
    "gjxl_amd_proposal_options_init": [_P, _SZ],
    "gjxl_amd_plain_path_set_proposal": [_P, _P, _P],
    "gjxl_amd_plain_path_proposal_sample": [_P, _P, _P, _P],
    "gjxl_amd_aq_session_create": [_P, _U32, _U32, ctypes.POINTER(_P)],
    "gjxl_amd_aq_session_destroy": [_P],
    "gjxl_amd_aq_prepare": [_P, _P, _P, _P, _P, _F],
    "gjxl_amd_aq_policy": [_P, ctypes.POINTER(_F), ctypes.POINTER(_F),
                           ctypes.POINTER(_F), ctypes.POINTER(ctypes.c_int32),
                           ctypes.POINTER(ctypes.c_int32)],
    "gjxl_amd_aq_evaluate": [_P, _P],
    "gjxl_amd_aq_buffer": [_P, ctypes.c_int32, ctypes.POINTER(_P),
                           ctypes.POINTER(_SZ)],
    "gjxl_amd_butteraugli_create": [_P, _U32, _U32, _F, _F, _F, ctypes.POINTER(_P)],
    "gjxl_amd_butteraugli_destroy": [_P],
    "gjxl_amd_butteraugli_prepare": [_P, _P],
    "gjxl_amd_butteraugli_compare": [_P, _P, _P],
    "gjxl_amd_butteraugli_score": [_P, ctypes.POINTER(ctypes.c_double)],
    "gjxl_amd_butteraugli_buffer": [_P, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.POINTER(_P), ctypes.POINTER(_SZ)],
    "gjxl_amd_reduce_distance_map": [_P, _P, _U32, _U32, _P, _P],
    "gjxl_amd_aq_update": [_P, _P, _P, ctypes.c_int32, _F],
    "gjxl_amd_aq_run": [_P, _P, ctypes.c_int32, _F, _P],
    "gjxl_amd_tokenizer_create": [_P, _U32, _U32, ctypes.POINTER(_P)],
    "gjxl_amd_tokenizer_destroy": [_P],
    "gjxl_amd_natural_order": [ctypes.c_int32, _P, _SZ, ctypes.POINTER(_SZ)],
    "gjxl_amd_tokenizer_set_orders": [_P, ctypes.c_uint16, _P],
    "gjxl_amd_tokenizer_set_context_map": [_P, _P, _U32, _P, _U32],
    "gjxl_amd_tokenize": [_P, _P, _P, _P, ctypes.c_int32,
                          ctypes.POINTER(ctypes.c_uint64)],
    "gjxl_amd_tokenizer_buffer": [_P, ctypes.c_int32, ctypes.POINTER(_P),
                                  ctypes.POINTER(_SZ)],
    "gjxl_amd_aq_check": [_P],
    "gjxl_amd_tokenizer_frame_statistics": [_P, _P, _P, _P, _P],
    "gjxl_amd_entropy_inputs_from_statistics": [_P, ctypes.c_int32, _P],
    "gjxl_amd_tokenizer_set_entropy_inputs": [_P, _P],
    "gjxl_amd_coefficient_order_tokens": [_P, _P, _P, _SZ, ctypes.POINTER(_SZ)],
    "gjxl_amd_plain_path_options_init": [_P, _SZ],
    "gjxl_amd_plain_path_create": [_P, _U32, _U32, _P, ctypes.POINTER(_P)],
    "gjxl_amd_plain_path_destroy": [_P],
    "gjxl_amd_plain_path_reconfigure": [_P, _U32, _U32, _P],
    "gjxl_amd_plain_path_run": [_P, _P, _P, _P],
    "gjxl_amd_plain_path_stage_times": [_P, _P, _P, ctypes.c_int32,
                                        ctypes.POINTER(ctypes.c_int32)],
    "gjxl_amd_host_alloc": [_SZ, ctypes.POINTER(_P)],
    "gjxl_amd_host_free": [_P],
    "gjxl_amd_split_dct_blocks": [_P, _U32, _U32, _P, _SZ, _P],
    "gjxl_amd_aq_split_forward_dct": [_P],
    "gjxl_amd_batch_create": [_I, _U32, ctypes.POINTER(_P)],
    "gjxl_amd_batch_max_in_flight": [_P],
    "gjxl_amd_batch_encode": [_P, _P, _SZ, _P],
    "gjxl_amd_batch_set_tile_queue": [_P, ctypes.c_int32, _U32],
    "gjxl_amd_plain_path_run_front": [_P, _P],
    "gjxl_amd_plain_path_cost_frame": [_P, _P],
    "gjxl_amd_plain_path_run_back": [_P, ctypes.c_int32, _P, _P],
    "gjxl_amd_batch_results_free": [_P, _SZ],
    "gjxl_amd_batch_shutdown": [_P],
    "gjxl_amd_batch_destroy": [_P],
}
_VOID = {"gjxl_amd_device_destroy", "gjxl_amd_aq_session_destroy",
         "gjxl_amd_butteraugli_destroy", "gjxl_amd_tokenizer_destroy",
         "gjxl_amd_plain_path_destroy", "gjxl_amd_host_free",
         "gjxl_amd_batch_results_free", "gjxl_amd_batch_shutdown",
         "gjxl_amd_batch_destroy", "gjxl_amd_proposal_destroy",
         "gjxl_amd_cost_queue_destroy"}
_STR = {"gjxl_amd_last_error"}
_SIZE = {"gjxl_amd_initial_quant_field_scratch_bytes", "gjxl_amd_proposal_parameter_count"}


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def _declare(lib: ctypes.CDLL) -> None:
    for name, argtypes in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        if name in _VOID:
            fn.restype = None
        elif name in _STR:
            fn.restype = ctypes.c_char_p
        elif name in _SIZE:
            fn.restype = ctypes.c_size_t
        elif name == "gjxl_amd_batch_max_in_flight":
            fn.restype = ctypes.c_uint32
        else:
            fn.restype = ctypes.c_int32


def _check(code: int) -> None:
    if code != GJXL_OK:
        msg = library().gjxl_amd_last_error()
        raise GjxlError(code, msg.decode() if msg else "")


def _drop_views(owner) -> None:
    """Clears the pointers of the views an owner lent out, so that a view used
    after its owner closed raises instead of reading freed memory."""
    for view in list(getattr(owner, "_views", ())):
        view.ptr = None


def pad8(v: int) -> int:
    return (v + 7) // 8 * 8


class DeviceArray:
    """A device allocation with a numpy shape and dtype."""

    def __init__(self, device: "Device", shape, dtype, pointer: int | None = None,
                 owner=None):
        self.device = device
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        # a borrowed pointer (a session buffer) is never freed here; its owner
        # clears it when the owner closes
        self.owned = pointer is None
        self.ptr = device._alloc(max(self.nbytes, 1)) if pointer is None else pointer
        if self.owned:
            device._arrays.add(self)
        elif owner is not None:
            owner._views.add(self)

    def _live(self) -> int:
        if not self.ptr:
            raise RuntimeError("device array is freed or its owner is closed")
        return self.ptr

    def numpy(self) -> np.ndarray:
        self._live()
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _check(library().gjxl_amd_download(
                self.device.handle, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def copy_from(self, array: np.ndarray) -> "DeviceArray":
        self._live()
        array = np.ascontiguousarray(array, dtype=self.dtype)
        if array.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        if self.nbytes:
            _check(library().gjxl_amd_upload(
                self.device.handle, self.ptr, array.ctypes.data, self.nbytes))
        return self

    def free(self) -> None:
        if self.ptr and self.owned:
            _check(library().gjxl_amd_free(self.device.handle, self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.owned and self.device.handle:
                library().gjxl_amd_free(self.device.handle, self.ptr)
        except Exception:
            pass


class Device:
    """One HIP device and its stream (gjxl_amd_device_create)."""

    def __init__(self, ordinal: int = 0):
        handle = ctypes.c_void_p()
        _check(library().gjxl_amd_device_create(ordinal, ctypes.byref(handle)))
        self.handle = handle.value
        # sessions and comparators bound to this device: closed first
        self._children = weakref.WeakSet()
        # arrays this device allocated: freed when it closes
        self._arrays = weakref.WeakSet()

    def close(self) -> None:
        if self.handle:
            for child in list(self._children):
                child.close()
            for array in list(self._arrays):
                array.free()
            library().gjxl_amd_device_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _alloc(self, nbytes: int) -> int:
        ptr = ctypes.c_void_p()
        _check(library().gjxl_amd_alloc(self.handle, nbytes, ctypes.byref(ptr)))
        return ptr.value

    def empty(self, shape, dtype=np.float32) -> DeviceArray:
        return DeviceArray(self, shape, dtype)

    def array(self, array: np.ndarray, dtype=None) -> DeviceArray:
        array = np.asarray(array)
        d = DeviceArray(self, array.shape, dtype or array.dtype)
        return d.copy_from(array)

    def synchronize(self) -> None:
        _check(library().gjxl_amd_synchronize(self.handle))

    def copy(self, dst: DeviceArray, src: DeviceArray, nbytes: int | None = None,
             dst_offset: int = 0, src_offset: int = 0) -> None:
        """Device-to-device copy on the stream (byte offsets)."""
        nbytes = min(dst.nbytes, src.nbytes) if nbytes is None else nbytes
        _check(library().gjxl_amd_copy(self.handle, dst.ptr + dst_offset,
                                       src.ptr + src_offset, nbytes))

    def record(self, slot: int) -> None:
        _check(library().gjxl_amd_event_record(self.handle, slot))

    def elapsed_ms(self, begin: int, end: int) -> float:
        ms = ctypes.c_float()
        _check(library().gjxl_amd_event_elapsed(
            self.handle, begin, end, ctypes.byref(ms)))
        return ms.value

    # kernel ids of probe() (GJXL_AMD_KERNEL_*)
    PROBE_KERNELS = {"CandidateCostKernel": 1, "DcQuantKernel": 2,
                     "DifferencesKernel": 3}

    def probe(self, kernel: str | None) -> None:
        """Times every launch of `kernel` (a PROBE_KERNELS name, or None to
        stop) on this handle's stream from now on."""
        kid = 0 if kernel is None else self.PROBE_KERNELS[kernel]
        _check(library().gjxl_amd_device_probe(self.handle, kid))

    def probe_times(self) -> list[float]:
        """Milliseconds of each probed launch since the last call."""
        ms = (ctypes.c_float * 64)()
        count = ctypes.c_int32()
        _check(library().gjxl_amd_device_probe_times(self.handle, ms, 64,
                                                     ctypes.byref(count)))
        return [ms[i] for i in range(count.value)]

    # ---- a1 + a2 --------------------------------------------------------
    def srgb8_to_linear_and_opsin(self, srgb: DeviceArray, width: int,
                                  height: int, bytes_per_pixel: int = 3,
                                  row_stride: int | None = None,
                                  intensity_target: float = 255.0,
                                  linear: DeviceArray | None = None,
                                  opsin: DeviceArray | None = None):
        row_stride = row_stride or width * bytes_per_pixel
        linear = linear or self.empty((3, height, width))
        opsin = opsin or self.empty((3, pad8(height), pad8(width)))
        _check(library().gjxl_amd_srgb8_to_linear_and_opsin(
            self.handle, srgb.ptr, row_stride, bytes_per_pixel, width, height,
            intensity_target, linear.ptr, opsin.ptr))
        return linear, opsin

    def linear_rgb_to_padded_opsin(self, linear: DeviceArray, width: int,
                                   height: int, intensity_target: float = 255.0,
                                   opsin: DeviceArray | None = None):
        opsin = opsin or self.empty((3, pad8(height), pad8(width)))
        _check(library().gjxl_amd_linear_rgb_to_padded_opsin(
            self.handle, linear.ptr, width, height, intensity_target, opsin.ptr))
        return opsin


    # ---- a4 ---------------------------------------------------------------
    def gaborish_inverse(self, opsin: DeviceArray, multipliers=(1.0, 1.0, 1.0),
                         out: DeviceArray | None = None) -> DeviceArray:
        _, h, w = opsin.shape
        out = out or self.empty(opsin.shape)
        mult = (ctypes.c_float * 3)(*multipliers)
        _check(library().gjxl_amd_gaborish_inverse(
            self.handle, opsin.ptr, w, h, mult, out.ptr))
        return out

    # ---- a5 ---------------------------------------------------------------
    def initial_quant_field(self, opsin: DeviceArray, target: float = 1.0,
                            rescale: float = 1.0, out: dict | None = None,
                            scratch: DeviceArray | None = None) -> dict:
        _, h, w = opsin.shape
        out = out if out is not None else {}
        if "quant_field" not in out:
            out["quant_field"] = self.empty((h // 8, w // 8))
            out["strategy_mask"] = self.empty((h // 8, w // 8))
            out["pixel_mask"] = self.empty((h, w))
        if scratch is None:
            nbytes = library().gjxl_amd_initial_quant_field_scratch_bytes(w, h)
            scratch = self.empty((nbytes,), np.uint8)
        _check(library().gjxl_amd_initial_quant_field(
            self.handle, opsin.ptr, w, h, target, rescale,
            out["quant_field"].ptr, out["strategy_mask"].ptr,
            out["pixel_mask"].ptr, scratch.ptr))
        return out

    # ---- a6 ---------------------------------------------------------------
    def initial_cfl(self, opsin: DeviceArray, out: dict | None = None) -> dict:
        _, h, w = opsin.shape
        out = out if out is not None else {}
        if "y_to_x" not in out:
            tiles = ((h + 63) // 64, (w + 63) // 64)
            out["y_to_x"] = self.empty(tiles, np.int8)
            out["y_to_b"] = self.empty(tiles, np.int8)
        _check(library().gjxl_amd_initial_cfl(
            self.handle, opsin.ptr, w, h, out["y_to_x"].ptr, out["y_to_b"].ptr))
        return out

    # ---- a7-a9 ------------------------------------------------------------
    def ac_strategy_costs(self, opsin: DeviceArray, quant_field: DeviceArray,
                          pixel_mask: DeviceArray, y_to_x: DeviceArray,
                          y_to_b: DeviceArray, target: float = 1.0,
                          costs: DeviceArray | None = None) -> DeviceArray:
        _, h, w = opsin.shape
        costs = costs or self.empty((7, h // 8, w // 8))
        _check(library().gjxl_amd_ac_strategy_costs(
            self.handle, opsin.ptr, w, h, quant_field.ptr, pixel_mask.ptr,
            y_to_x.ptr, y_to_b.ptr, target, costs.ptr))
        return costs

    # ---- a10 --------------------------------------------------------------
    def ac_strategy_search(self, costs: DeviceArray, target: float = 1.0,
                           cells: DeviceArray | None = None) -> DeviceArray:
        _, bh, bw = costs.shape
        cells = cells or self.empty((bh, bw), np.uint8)
        _check(library().gjxl_amd_ac_strategy_search(
            self.handle, costs.ptr, bw * 8, bh * 8, target, cells.ptr))
        return cells


# Session buffer ids (include/gjxl_amd.h GJXL_AMD_AQ_*), in order.
AQ_BUFFERS = (
    "offsets", "anchor_list", "anchor_count", "coefficients", "adjusted_field",
    "stats", "select", "quantizer", "raw_quant", "y_to_x", "y_to_b",
    "quantized", "dc", "dc_quantized", "dc_reconstructed",
    "dc_smoothed", "inverse_sigma", "reconstructed",
    "filtered", "linear", "field", "block_distance", "scores", "strategy_lists",
)


class AqSession:
    """Adaptive-quantization evaluation session (gjxl_amd_aq_session_*)."""

    def __init__(self, device: Device, width: int, height: int):
        self.device = device
        self.width, self.height = width, height
        self.pw, self.ph = pad8(width), pad8(height)
        self.bw, self.bh = self.pw // 8, self.ph // 8
        handle = ctypes.c_void_p()
        _check(library().gjxl_amd_aq_session_create(
            device.handle, width, height, ctypes.byref(handle)))
        self.handle = handle.value
        self._borrowed = None
        self._views = weakref.WeakSet()
        device._children.add(self)

    def close(self) -> None:
        _drop_views(self)
        if self.handle:
            library().gjxl_amd_aq_session_destroy(self.handle)
            self.handle = None
        self._borrowed = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare(self, opsin: DeviceArray, preprocessed: DeviceArray,
                cells: DeviceArray, initial_quant_field: DeviceArray,
                target: float = 1.0) -> None:
        """The inputs stay borrowed by the session until it is closed or
        prepared again (the evaluations read the strategy cells), so the
        session keeps them alive."""
        _check(library().gjxl_amd_aq_prepare(
            self.handle, opsin.ptr, preprocessed.ptr, cells.ptr,
            initial_quant_field.ptr, target))
        self._borrowed = (opsin, preprocessed, cells, initial_quant_field)

    def policy(self) -> dict:
        lo, hi, qdc = _F(), _F(), _F()
        xs, bs = ctypes.c_int32(), ctypes.c_int32()
        _check(library().gjxl_amd_aq_policy(
            self.handle, ctypes.byref(lo), ctypes.byref(hi), ctypes.byref(qdc),
            ctypes.byref(xs), ctypes.byref(bs)))
        return {"lower": lo.value, "upper": hi.value, "quant_dc": qdc.value,
                "x_scale": xs.value, "b_scale": bs.value}

    def check(self) -> None:
        """Waits and raises if an update met an invalid block distance."""
        _check(library().gjxl_amd_aq_check(self.handle))

    def evaluate(self, quant_field: DeviceArray) -> None:
        _check(library().gjxl_amd_aq_evaluate(self.handle, quant_field.ptr))

    def buffer(self, name: str, shape=None, dtype=np.float32) -> DeviceArray:
        """A borrowed view of one session buffer (whole buffer by default)."""
        ptr, size = ctypes.c_void_p(), ctypes.c_size_t()
        _check(library().gjxl_amd_aq_buffer(
            self.handle, AQ_BUFFERS.index(name), ctypes.byref(ptr),
            ctypes.byref(size)))
        dtype = np.dtype(dtype)
        if shape is None:
            shape = (size.value // dtype.itemsize,)
        view = DeviceArray(self.device, shape, dtype, pointer=ptr.value, owner=self)
        if view.nbytes > size.value:
            raise ValueError(f"{name}: view of {view.nbytes} bytes exceeds {size.value}")
        return view


    def update(self, quant_field: DeviceArray, block_distance: DeviceArray,
               iteration: int, target: float = 1.0) -> None:
        _check(library().gjxl_amd_aq_update(
            self.handle, quant_field.ptr, block_distance.ptr, iteration, target))

    def run(self, butteraugli: "Butteraugli", iterations: int = 2,
            target: float = 1.0, scores: bool = True) -> list[float] | None:
        """The Butteraugli AQ loop.  With scores, waits and returns the score
        of every evaluation; otherwise only enqueues."""
        out = (ctypes.c_double * (iterations + 1))() if scores else None
        _check(library().gjxl_amd_aq_run(self.handle, butteraugli.handle,
                                         iterations, target, out))
        return list(out) if scores else None


BUTTERAUGLI_BUFFERS = ("map", "reference_psycho", "distorted_psycho", "stages", "xyb",
                       "score")


class Butteraugli:
    """Device Butteraugli with a prepared reference (gjxl_amd_butteraugli_*)."""

    def __init__(self, device: Device, width: int, height: int,
                 hf_asymmetry: float = 1.0, x_multiplier: float = 1.0,
                 intensity_target: float = 80.0):
        self.device = device
        self.width, self.height = width, height
        handle = ctypes.c_void_p()
        _check(library().gjxl_amd_butteraugli_create(
            device.handle, width, height, hf_asymmetry, x_multiplier,
            intensity_target, ctypes.byref(handle)))
        self.handle = handle.value
        self._views = weakref.WeakSet()
        device._children.add(self)

    def close(self) -> None:
        _drop_views(self)
        if self.handle:
            library().gjxl_amd_butteraugli_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare(self, reference_linear: DeviceArray) -> None:
        _check(library().gjxl_amd_butteraugli_prepare(self.handle, reference_linear.ptr))

    def compare(self, distorted_linear: DeviceArray,
                distance_map: DeviceArray | None = None) -> None:
        _check(library().gjxl_amd_butteraugli_compare(
            self.handle, distorted_linear.ptr,
            distance_map.ptr if distance_map is not None else None))

    def score(self) -> float:
        v = ctypes.c_double()
        _check(library().gjxl_amd_butteraugli_score(self.handle, ctypes.byref(v)))
        return v.value

    def buffer(self, name: str, scale: int = 0) -> DeviceArray:
        ptr, size = ctypes.c_void_p(), ctypes.c_size_t()
        _check(library().gjxl_amd_butteraugli_buffer(
            self.handle, scale, BUTTERAUGLI_BUFFERS.index(name), ctypes.byref(ptr),
            ctypes.byref(size)))
        if name == "map":
            shape = (self.height, self.width)
        elif name == "score":
            return DeviceArray(self.device, (1,), np.uint32, pointer=ptr.value, owner=self)
        else:
            w = self.width if scale == 0 else (self.width + 1) // 2
            h = self.height if scale == 0 else (self.height + 1) // 2
            if scale == 0 and (self.width < 8 or self.height < 8):
                w, h = max(w, 8), max(h, 8)
            shape = (size.value // (4 * w * h), h, w)
        return DeviceArray(self.device, shape, np.float32, pointer=ptr.value, owner=self)


def reduce_distance_map(device: Device, distance_map: DeviceArray,
                        cells: DeviceArray,
                        out: DeviceArray | None = None) -> DeviceArray:
    h, w = distance_map.shape
    out = out or device.empty(cells.shape)
    _check(library().gjxl_amd_reduce_distance_map(
        device.handle, distance_map.ptr, w, h, cells.ptr, out.ptr))
    return out


TOKEN_BUFFERS = (
    "values", "contexts", "group_offsets", "histogram", "nonzero_map", "tasks",
    "task_offsets", "group_tokens", "offsets", "anchor_list", "anchor_count",
    "orders", "inverse_orders", "status", "statistics", "sample_decisions",
)
ORDER_FAMILIES = 13
# searched order families: (family, representative strategy, counter offset,
# size, LLF size); counters of one channel are GJXLAmdFrameStatistics'
FAMILIES = ((0, 0, 0, 64, 1), (4, 6, 64, 128, 2), (2, 4, 192, 256, 4),
            (6, 10, 448, 512, 8), (3, 5, 960, 1024, 16))


class FrameStatistics(ctypes.Structure):
    """GJXLAmdFrameStatistics."""
    _fields_ = [("block_width", _U32), ("block_height", _U32),
                ("present_mask", ctypes.c_uint16),
                ("zero_counts", (_U32 * 1984) * 3),
                ("sampled_dct8_zero_counts", (_U32 * 64) * 3),
                ("raw_quant_counts", (_U32 * 256) * ORDER_FAMILIES)]

    def arrays(self) -> dict:
        return {"present_mask": int(self.present_mask),
                "zero_counts": np.ctypeslib.as_array(self.zero_counts).copy(),
                "sampled": np.ctypeslib.as_array(self.sampled_dct8_zero_counts).copy(),
                "raw_counts": np.ctypeslib.as_array(self.raw_quant_counts).copy()}


class EntropyInputs(ctypes.Structure):
    """GJXLAmdEntropyInputs: coefficient orders and block-context map."""
    _fields_ = [("used_order_mask", ctypes.c_uint16),
                ("orders", ((_U32 * 1024) * 3) * ORDER_FAMILIES),
                ("qf_threshold_count", _U32), ("qf_thresholds", _U32 * 15),
                ("num_contexts", _U32),
                ("context_map", ctypes.c_uint8 * (3 * ORDER_FAMILIES * 16))]

    def order_table(self) -> np.ndarray:
        """The orders as the tokenizer's 7-stage table (3 x 2624, natural
        where the family is unused)."""
        stage_strategies = (0, 6, 7, 4, 10, 11, 5)
        family_of = {0: 0, 6: 4, 7: 4, 4: 2, 10: 6, 11: 6, 5: 3}
        orders = np.ctypeslib.as_array(self.orders)
        parts = []
        for st in stage_strategies:
            f = family_of[st]
            nat = natural_order(st)
            for c in range(3):
                parts.append(orders[f, c, :nat.size] if self.used_order_mask >> f & 1
                             else nat)
        return np.concatenate(parts).astype(np.uint32)

    def context_model(self) -> tuple[np.ndarray, np.ndarray, int]:
        n = int(self.qf_threshold_count)
        entries = 3 * ORDER_FAMILIES * (n + 1)
        return (np.array(self.qf_thresholds[:n], np.uint32),
                np.array(self.context_map[:entries], np.uint8), int(self.num_contexts))


def entropy_inputs(statistics: FrameStatistics, effort: int = 7) -> EntropyInputs:
    """gjxl_amd_entropy_inputs_from_statistics (host code)."""
    out = EntropyInputs()
    _check(library().gjxl_amd_entropy_inputs_from_statistics(
        ctypes.byref(statistics), effort, ctypes.byref(out)))
    return out


def order_tokens(inputs: EntropyInputs) -> tuple[np.ndarray, np.ndarray]:
    """gjxl_amd_coefficient_order_tokens: (contexts, values)."""
    cap = 3 * 1984 + 15
    ctx = np.zeros(cap, np.uint32)
    val = np.zeros(cap, np.uint32)
    n = ctypes.c_size_t()
    _check(library().gjxl_amd_coefficient_order_tokens(
        ctypes.byref(inputs), ctx.ctypes.data, val.ctypes.data, cap, ctypes.byref(n)))
    return ctx[:n.value].copy(), val[:n.value].copy()


def natural_order(strategy: int) -> np.ndarray:
    """ComputeSimpleNaturalCoefficientOrder of the library (host code)."""
    out = np.zeros(1024, np.uint32)
    n = ctypes.c_size_t()
    _check(library().gjxl_amd_natural_order(
        strategy, out.ctypes.data, out.size, ctypes.byref(n)))
    return out[:n.value].copy()


class Tokenizer:
    """AC tokenization of a quantized frame (gjxl_amd_tokenizer_*)."""

    def __init__(self, device: Device, width: int, height: int):
        self.device = device
        self.width, self.height = width, height
        self.bw, self.bh = pad8(width) // 8, pad8(height) // 8
        self.groups = ((self.bw + 31) // 32) * ((self.bh + 31) // 32)
        self.num_contexts = 4
        handle = ctypes.c_void_p()
        _check(library().gjxl_amd_tokenizer_create(
            device.handle, width, height, ctypes.byref(handle)))
        self.handle = handle.value
        self._views = weakref.WeakSet()
        device._children.add(self)

    def close(self) -> None:
        _drop_views(self)
        if self.handle:
            library().gjxl_amd_tokenizer_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_orders(self, used_order_mask: int = 0, orders: dict | None = None) -> None:
        """orders[(family, channel)] -> uint32 array for every family in the
        mask (channel 0 = X, 1 = Y, 2 = B)."""
        table = (_P * (3 * ORDER_FAMILIES))()
        keep = []
        for (family, channel), order in (orders or {}).items():
            a = np.ascontiguousarray(order, np.uint32)
            keep.append(a)
            table[3 * family + channel] = a.ctypes.data
        _check(library().gjxl_amd_tokenizer_set_orders(
            self.handle, used_order_mask, table if used_order_mask else None))

    def set_context_map(self, thresholds, context_map, num_contexts: int) -> None:
        thr = np.ascontiguousarray(np.array(thresholds, np.uint32).reshape(-1))
        cmap = np.ascontiguousarray(context_map, np.uint8)
        _check(library().gjxl_amd_tokenizer_set_context_map(
            self.handle, thr.ctypes.data if thr.size else None, thr.size,
            cmap.ctypes.data, num_contexts))
        self.num_contexts = num_contexts

    def frame_statistics(self, quantized: DeviceArray, cells: DeviceArray,
                         raw_quant: DeviceArray) -> FrameStatistics:
        out = FrameStatistics()
        _check(library().gjxl_amd_tokenizer_frame_statistics(
            self.handle, quantized.ptr, cells.ptr, raw_quant.ptr, ctypes.byref(out)))
        return out

    def set_entropy_inputs(self, inputs: EntropyInputs) -> None:
        _check(library().gjxl_amd_tokenizer_set_entropy_inputs(self.handle,
                                                               ctypes.byref(inputs)))
        self.num_contexts = int(inputs.num_contexts)

    def tokenize(self, quantized: DeviceArray, cells: DeviceArray,
                 raw_quant: DeviceArray, populations: bool = False) -> int:
        n = ctypes.c_uint64()
        _check(library().gjxl_amd_tokenize(
            self.handle, quantized.ptr, cells.ptr, raw_quant.ptr,
            1 if populations else 0, ctypes.byref(n)))
        return n.value

    def buffer(self, name: str, shape=None, dtype=None) -> DeviceArray:
        ptr, size = ctypes.c_void_p(), ctypes.c_size_t()
        _check(library().gjxl_amd_tokenizer_buffer(
            self.handle, TOKEN_BUFFERS.index(name), ctypes.byref(ptr),
            ctypes.byref(size)))
        if dtype is None:
            dtype = {"contexts": np.uint16, "inverse_orders": np.uint16,
                     "nonzero_map": np.uint8}.get(name, np.uint32)
        dtype = np.dtype(dtype)
        if shape is None:
            shape = (size.value // dtype.itemsize,)
        view = DeviceArray(self.device, shape, dtype, pointer=ptr.value, owner=self)
        if view.nbytes > size.value:
            raise ValueError(f"{name}: view of {view.nbytes} bytes exceeds {size.value}")
        return view

    def tokens(self, count: int) -> dict:
        """The last stream on the host: values, contexts, group offsets."""
        out = {"values": self.buffer("values", (count,)).numpy() if count else
               np.zeros(0, np.uint32),
               "contexts": self.buffer("contexts", (count,)).numpy() if count else
               np.zeros(0, np.uint16),
               "group_offsets": self.buffer("group_offsets", (self.groups + 1,)).numpy()}
        return out

    def histogram(self) -> np.ndarray:
        rows = self.num_contexts * 495
        return self.buffer("histogram", (rows, 128)).numpy()


def gpu_available() -> bool:
    """True when a HIP device can be opened (does not fall back to CPU)."""
    if os.environ.get("GJXL_AMD_NO_GPU"):
        return False
    try:
        count = ctypes.c_int()
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        return hip.hipGetDeviceCount(ctypes.byref(count)) == 0 and count.value > 0
    except OSError:
        return False


from . import pipeline  # noqa: E402,F401  (stage sequencing over this API)


# ---- f2: sparse AC read-back (gjxl_amd_sparse_ac) ----------------------------------

def sparse_ac(device: Device, source: DeviceArray, count: int | None = None):
    """The sparse form of `count` values of `source` (int8, int16 or int32):
    (masks uint64, offsets uint32, values, total) read back to the host."""
    width = source.dtype.itemsize
    count = source.nbytes // width if count is None else count
    words = (count + 63) // 64
    masks = device.empty((max(words, 1),), np.uint64)
    offsets = device.empty((max(words, 1),), np.uint32)
    values = device.empty((max(count, 1),), source.dtype)
    total = device.empty((1,), np.uint32)
    _check(library().gjxl_amd_sparse_ac(device.handle, source.ptr, count, width, masks.ptr,
                                        offsets.ptr, values.ptr, total.ptr))
    n = int(total.numpy()[0])
    out = (masks.numpy()[:words], offsets.numpy()[:words], values.numpy()[:n], n)
    for a in (masks, offsets, values, total):
        a.free()
    return out


# ---- the tile queue for batches of candidate costs (gjxl_amd_cost_queue_*) --------

class CostFrame(ctypes.Structure):
    _fields_ = [("opsin", _P), ("quant_field", _P), ("pixel_mask", _P), ("y_to_x", _P),
                ("y_to_b", _P), ("allow", _P), ("costs", _P), ("width", _U32),
                ("height", _U32), ("butteraugli_target", _F)]


class CostQueue:
    """Candidate costs of batches of frames through one captured graph of
    persistent tile-queue kernels (gjxl_amd_cost_queue_*).  A frame is a dict
    of DeviceArrays: opsin (Gaborish-inverted, 3 x h x w), quant_field,
    pixel_mask, y_to_x, y_to_b, costs (7 x bh x bw), optionally allow, and
    the target."""

    def __init__(self, device: Device, max_frames: int):
        self.device = device
        handle = ctypes.c_void_p()
        _check(library().gjxl_amd_cost_queue_create(device.handle, max_frames,
                                                    ctypes.byref(handle)))
        self.handle = handle.value
        device._children.add(self)

    def close(self) -> None:
        if self.handle:
            library().gjxl_amd_cost_queue_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _table(frames) -> ctypes.Array:
        table = (CostFrame * len(frames))()
        for i, f in enumerate(frames):
            _, h, w = f["opsin"].shape
            allow = f.get("allow")
            table[i] = CostFrame(f["opsin"].ptr, f["quant_field"].ptr, f["pixel_mask"].ptr,
                                 f["y_to_x"].ptr, f["y_to_b"].ptr,
                                 allow.ptr if allow is not None else None, f["costs"].ptr,
                                 w, h, f.get("target", 1.0))
        return table

    def run(self, frames) -> None:
        table = self._table(frames)
        _check(library().gjxl_amd_cost_queue_run(self.handle, table, len(frames)))
        self._last = (table, frames)  # kept alive until the next call

    def tune(self, frames, repetitions: int = 3) -> None:
        table = self._table(frames)
        _check(library().gjxl_amd_cost_queue_tune(self.handle, table, len(frames),
                                                  repetitions))

    def grids(self) -> tuple[list[int], list[float]]:
        g = (ctypes.c_uint32 * 7)()
        ms = (ctypes.c_float * 7)()
        _check(library().gjxl_amd_cost_queue_grids(self.handle, g, ms))
        return list(g), list(ms)


# ---- the whole plain path (gjxl_amd_plain_path_*) ------------------------------

class PlainPathOptions(ctypes.Structure):
    _fields_ = [("struct_size", _SZ), ("distance", _F), ("effort", ctypes.c_int32),
                ("intensity_target", _F)]


class Image(ctypes.Structure):
    _fields_ = [("width", _U32), ("height", _U32), ("bytes_per_pixel", _U32),
                ("device_resident", ctypes.c_int32), ("pixels", _P),
                ("pixels_size", _SZ), ("row_stride_bytes", _SZ)]


class TokenOutput(ctypes.Structure):
    _fields_ = [("values", _P), ("contexts", _P), ("group_offsets", _P),
                ("capacity", _SZ), ("group_capacity", _SZ)]


class PlainFrame(ctypes.Structure):
    _fields_ = [("block_width", _U32), ("block_height", _U32), ("groups", _U32),
                ("cells", _P), ("quant_field", _P), ("block_distance", _P),
                ("raw_quant", _P), ("quantized", _P), ("dc_quantized", _P),
                ("y_to_x", _P), ("y_to_b", _P), ("token_values", _P),
                ("token_contexts", _P), ("token_group_offsets", _P),
                ("token_count", ctypes.c_uint64), ("x_qm_scale", ctypes.c_int32),
                ("b_qm_scale", ctypes.c_int32), ("quant_dc", _F),
                ("evaluations", ctypes.c_int32), ("scores", _F * 8),
                ("entropy_inputs", _P), ("proposal_allow", _P), ("proposal_norm", _F)]


STAGE_NAMES = ("begin", "upload", "srgb_to_linear_and_opsin", "gaborish_inverse",
               "initial_quant_field", "initial_cfl", "ac_strategy_costs",
               "ac_strategy_search", "aq_prepare", "butteraugli_prepare", "aq_evaluate",
               "butteraugli_compare", "aq_reduce_update", "ac_tokenize", "entropy_inputs",
               "token_readback", "proposal")


class HostBuffer:
    """Pinned host memory (gjxl_amd_host_alloc) with a numpy view."""

    def __init__(self, shape, dtype):
        self.dtype = np.dtype(dtype)
        self.shape = tuple(int(v) for v in shape)
        nbytes = max(int(np.prod(self.shape)) * self.dtype.itemsize, 1)
        ptr = ctypes.c_void_p()
        _check(library().gjxl_amd_host_alloc(nbytes, ctypes.byref(ptr)))
        self.ptr = ptr.value
        raw = (ctypes.c_uint8 * nbytes).from_address(self.ptr)
        self.array = np.frombuffer(raw, dtype=np.uint8)[:int(np.prod(self.shape)) *
                                                         self.dtype.itemsize]
        self.array = self.array.view(self.dtype).reshape(self.shape)

    def free(self) -> None:
        if self.ptr:
            self.array = None
            library().gjxl_amd_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class NativePlainPath:
    """The plain path as one C++ call (gjxl_amd_plain_path_*): upload from
    pinned host memory, every stage on the device stream, token read-back
    into pinned host buffers."""

    def __init__(self, device: Device, width: int, height: int,
                 distance: float = 1.0, effort: int = 7,
                 intensity_target: float = 255.0):
        self.device = device
        self.width, self.height = width, height
        opts = PlainPathOptions()
        _check(library().gjxl_amd_plain_path_options_init(ctypes.byref(opts),
                                                          ctypes.sizeof(opts)))
        opts.distance, opts.effort, opts.intensity_target = distance, effort, intensity_target
        handle = ctypes.c_void_p()
        _check(library().gjxl_amd_plain_path_create(device.handle, width, height,
                                                     ctypes.byref(opts),
                                                     ctypes.byref(handle)))
        self.handle = handle.value
        self._views = weakref.WeakSet()
        device._children.add(self)
        self.options = opts
        self.capacity_extent = (width, height)
        self._extent(width, height)
        capacity = 3 * 65 * self.bw * self.bh
        self.pixels = HostBuffer((height * width * 3,), np.uint8)
        self.values = HostBuffer((capacity,), np.uint32)
        self.contexts = HostBuffer((capacity,), np.uint16)
        self.group_offsets = HostBuffer((self.groups + 1,), np.uint32)
        self.frame = PlainFrame()

    def _extent(self, width: int, height: int) -> None:
        self.width, self.height = width, height
        bw, bh = pad8(width) // 8, pad8(height) // 8
        self.bw, self.bh = bw, bh
        self.groups = ((bw + 31) // 32) * ((bh + 31) // 32)

    def reconfigure(self, width: int, height: int, distance: float | None = None,
                    intensity_target: float | None = None) -> None:
        """Re-targets the path to another extent no wider and no taller than
        the creation extent (and optionally another distance or intensity
        target) inside the buffers it holds; device views of the last run go
        stale and are dropped."""
        opts = PlainPathOptions()
        ctypes.memmove(ctypes.byref(opts), ctypes.byref(self.options), ctypes.sizeof(opts))
        if distance is not None:
            opts.distance = distance
        if intensity_target is not None:
            opts.intensity_target = intensity_target
        _check(library().gjxl_amd_plain_path_reconfigure(self.handle, width, height,
                                                          ctypes.byref(opts)))
        _drop_views(self)
        self.options = opts
        self.frame = PlainFrame()
        self._extent(width, height)

    def close(self) -> None:
        _drop_views(self)
        if self.handle:
            library().gjxl_amd_plain_path_destroy(self.handle)
            self.handle = None
        for b in ("pixels", "values", "contexts", "group_offsets"):
            buf = getattr(self, b, None)
            if buf is not None:
                buf.free()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, srgb: np.ndarray | None = None, read_tokens: bool = True,
            resident: DeviceArray | None = None) -> int:
        """One pass over `srgb` (h, w, 3 uint8; copied into the pinned input
        first), over the pinned input as it stands, or over `resident`
        (packed pixels already in device memory: no upload); returns the
        token count.  With read_tokens the stream lands in the pinned
        buffers."""
        n = self.width * self.height * 3
        if resident is not None:
            if resident.nbytes < n:
                raise ValueError("resident input is smaller than the image")
            img = Image(self.width, self.height, 3, 1, resident.ptr,
                        resident.nbytes, self.width * 3)
        else:
            if srgb is not None:
                if srgb.shape != (self.height, self.width, 3):
                    raise ValueError("image extent differs from the plain path's")
                self.pixels.array[:n] = np.ascontiguousarray(srgb, np.uint8).reshape(-1)
            img = Image(self.width, self.height, 3, 0, self.pixels.ptr, n, self.width * 3)
        out = TokenOutput(self.values.ptr, self.contexts.ptr, self.group_offsets.ptr,
                          self.values.shape[0], self.group_offsets.shape[0])
        _check(library().gjxl_amd_plain_path_run(
            self.handle, ctypes.byref(img), ctypes.byref(out) if read_tokens else None,
            ctypes.byref(self.frame)))
        return int(self.frame.token_count)

    def tokens(self) -> dict:
        n = int(self.frame.token_count)
        return {"values": self.values.array[:n].copy(),
                "contexts": self.contexts.array[:n].copy(),
                "group_offsets": self.group_offsets.array[:self.groups + 1].copy()}

    def scores(self) -> list[float]:
        return [float(self.frame.scores[i]) for i in range(self.frame.evaluations)]

    def entropy_inputs(self) -> EntropyInputs:
        """A copy of the last run's coefficient orders and block-context map."""
        if not self.frame.entropy_inputs:
            raise RuntimeError("the plain path has not completed a run")
        out = EntropyInputs()
        ctypes.memmove(ctypes.byref(out), self.frame.entropy_inputs, ctypes.sizeof(out))
        return out

    def view(self, name: str) -> DeviceArray:
        """A device view of one frame buffer of the last run."""
        bw, bh, store = self.bw, self.bh, self.groups * 65536
        tiles = ((bh + 7) // 8, (bw + 7) // 8)
        shapes = {"cells": ((bh, bw), np.uint8), "quant_field": ((bh, bw), np.float32),
                  "block_distance": ((bh, bw), np.float32),
                  "raw_quant": ((bh, bw), np.int32), "quantized": ((3, store), np.int32),
                  "dc_quantized": ((3, bh, bw), np.int32), "y_to_x": (tiles, np.int8),
                  "y_to_b": (tiles, np.int8), "proposal_allow": ((bh, bw), np.uint8)}
        shape, dtype = shapes[name]
        pointer = getattr(self.frame, name)
        if not pointer:
            raise RuntimeError(f"the last run has no {name} buffer")
        return DeviceArray(self.device, shape, dtype, pointer=pointer, owner=self)

    def set_proposal(self, net, **options) -> None:
        """Attaches a proposal network (gjxl_amd.proposal.ProposalNet) for
        learned runs, with GJXLAmdProposalOptions fields as keywords, or
        detaches it (net None)."""
        from .proposal import ProposalOptions

        opts = ProposalOptions.defaults()
        for k, v in options.items():
            if not hasattr(opts, k):
                raise TypeError(f"unknown proposal option {k!r}")
            setattr(opts, k, v)
        _check(library().gjxl_amd_plain_path_set_proposal(
            self.handle, None if net is None else net.handle, ctypes.byref(opts)))
        self.proposal = net

    def proposal_sample(self, features: DeviceArray, labels: DeviceArray,
                        targets: DeviceArray) -> None:
        """Training data of the last run: features, strategy-code labels and
        log(final / adjusted field) targets per block."""
        n = self.bw * self.bh
        if features.nbytes < 32 * n or labels.nbytes < n or targets.nbytes < 4 * n:
            raise ValueError("proposal sample buffers are smaller than the block grid")
        _check(library().gjxl_amd_plain_path_proposal_sample(
            self.handle, features.ptr, labels.ptr, targets.ptr))

    def stage_times(self) -> list[tuple[str, float]]:
        stages = (ctypes.c_int32 * 64)()
        ms = (ctypes.c_float * 64)()
        count = ctypes.c_int32()
        _check(library().gjxl_amd_plain_path_stage_times(self.handle, stages, ms, 64,
                                                         ctypes.byref(count)))
        return [(STAGE_NAMES[stages[i]], float(ms[i])) for i in range(min(count.value, 64))]


# ---- batch driver (gjxl_amd_batch_*) -------------------------------------------

class BatchRequest(ctypes.Structure):
    _fields_ = [("image", Image), ("options", PlainPathOptions)]


class BatchResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("message", ctypes.c_char * 256),
                ("width", _U32), ("height", _U32), ("groups", _U32),
                ("token_count", ctypes.c_uint64), ("token_values", _P),
                ("token_contexts", _P), ("token_group_offsets", _P),
                ("entropy_inputs", _P), ("evaluations", ctypes.c_int32),
                ("scores", _F * 8), ("queue_nanoseconds", ctypes.c_uint64),
                ("service_nanoseconds", ctypes.c_uint64),
                ("ready_nanoseconds", ctypes.c_uint64)]

    def tokens(self) -> dict:
        n = int(self.token_count)

        def grab(ptr, count, ctype, dtype):
            if count == 0:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array((ctype * count).from_address(ptr)).copy()

        return {"values": grab(self.token_values, n, ctypes.c_uint32, np.uint32),
                "contexts": grab(self.token_contexts, n, ctypes.c_uint16, np.uint16),
                "group_offsets": grab(self.token_group_offsets, int(self.groups) + 1,
                                      ctypes.c_uint32, np.uint32)}

    def scores_list(self) -> list[float]:
        return [float(self.scores[i]) for i in range(self.evaluations)]

    def entropy(self) -> EntropyInputs:
        out = EntropyInputs()
        ctypes.memmove(ctypes.byref(out), self.entropy_inputs, ctypes.sizeof(out))
        return out


class Batch:
    """The batch driver: max_in_flight workers, each with its own device
    handle and stream, running the plain path on independent images."""

    def __init__(self, ordinal: int = 0, max_in_flight: int = 3):
        handle = ctypes.c_void_p()
        _check(library().gjxl_amd_batch_create(ordinal, max_in_flight,
                                                ctypes.byref(handle)))
        self.handle = handle.value

    @property
    def max_in_flight(self) -> int:
        return int(library().gjxl_amd_batch_max_in_flight(self.handle))

    def encode(self, images: list[np.ndarray], options: list[dict] | None = None,
               results=None):
        """images: (h, w, 3|4) uint8 arrays; returns the BatchResult array
        (free it with free_results).  Raises on a driver-level failure."""
        count = len(images)
        keep = [np.ascontiguousarray(a, np.uint8) for a in images]
        reqs = (BatchRequest * max(count, 1))()
        for i, a in enumerate(keep):
            h, w, bpp = a.shape
            reqs[i].image = Image(w, h, bpp, 0, a.ctypes.data, a.nbytes, w * bpp)
            o = (options or [{}] * count)[i]
            if o:
                reqs[i].options = PlainPathOptions(ctypes.sizeof(PlainPathOptions),
                                                   o.get("distance", 1.0), o.get("effort", 7),
                                                   o.get("intensity_target", 255.0))
        out = results if results is not None else (BatchResult * max(count, 1))()
        _check(library().gjxl_amd_batch_encode(self.handle, reqs, count, out))
        return out

    def set_tile_queue(self, enable: bool = True, tune_repetitions: int = 0) -> None:
        """Group mode: the candidate costs of each group of max_in_flight
        images through one tile-queue graph (tuned on the next call's first
        group when tune_repetitions > 0)."""
        _check(library().gjxl_amd_batch_set_tile_queue(self.handle, int(enable),
                                                       tune_repetitions))

    @staticmethod
    def free_results(results, count: int) -> None:
        library().gjxl_amd_batch_results_free(results, count)

    def shutdown(self) -> None:
        if self.handle:
            library().gjxl_amd_batch_shutdown(self.handle)

    def close(self) -> None:
        if self.handle:
            library().gjxl_amd_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

