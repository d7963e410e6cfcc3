"""Stage sequencing of the plain path on one device.

Mirrors the order of RunPreparedQuantizationPipelineWithProviders
(src/codec/quantization_pipeline.cpp:428-586) and PrepareWorkflow
(src/codestream/workflow.cpp:636-768) for the stages built so far.  Every
stage is a HIP kernel behind the C ABI; this module only allocates and calls.
"""
from __future__ import annotations

import numpy as np

from . import Device, DeviceArray, pad8  # noqa: F401


def run_frontend(dev: Device, srgb: np.ndarray, intensity_target: float = 255.0,
                 srgb_device: DeviceArray | None = None,
                 out: dict | None = None) -> dict:
    """packed sRGB8 (h, w, 3|4) -> {'linear', 'opsin'} device arrays."""
    h, w, bpp = srgb.shape
    if srgb_device is None:
        srgb_device = dev.array(np.ascontiguousarray(srgb))
    out = out if out is not None else {}
    if "linear" not in out:
        out["linear"] = dev.empty((3, h, w))
        out["opsin"] = dev.empty((3, pad8(h), pad8(w)))
    dev.srgb8_to_linear_and_opsin(srgb_device, w, h, bytes_per_pixel=bpp,
                                  intensity_target=intensity_target,
                                  linear=out["linear"], opsin=out["opsin"])
    return out
