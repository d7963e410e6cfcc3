"""Stage sequencing of the plain path on one device, call by call.

The product's sequence is the C++ one behind gjxl_amd_plain_path_run
(gjxl_amd.NativePlainPath); this module repeats it through the per-stage
entry points so that tests can read every intermediate buffer.

Mirrors the order of RunPreparedQuantizationPipelineWithProviders
(src/codec/quantization_pipeline.cpp:428-586) and PrepareWorkflow
(src/codestream/workflow.cpp:636-768): XYB, Gaborish inverse, initial quant
field, initial CfL, AC strategy search, then the Butteraugli AQ loop
(src/codec/adaptive_quantization.cpp:1600-1738).  Every stage is a HIP
kernel behind the C ABI; this module only allocates and calls.
"""
from __future__ import annotations

from typing import Callable

import numpy as np

from . import (AqSession, Butteraugli, Device, DeviceArray, Tokenizer, entropy_inputs,
               library, pad8, reduce_distance_map)


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


class PlainPath:
    """Device buffers for one image extent and the plain path over them.

    run() takes the packed sRGB8 pixels already on the device and leaves the
    strategy cells, the final quant field, the block distances and the last
    evaluation's frame in the session (d1.0 e7: two AQ updates, three
    evaluations), and that frame's AC tokens in the tokenizer, with the
    coefficient orders and block-context map that effort 7 chooses from the
    frame's statistics."""

    def __init__(self, dev: Device, width: int, height: int,
                 target: float = 1.0, iterations: int = 2):
        self.dev = dev
        self.width, self.height = width, height
        self.target, self.iterations = target, iterations
        ph, pw = pad8(height), pad8(width)
        self.linear = dev.empty((3, height, width))
        self.opsin = dev.empty((3, ph, pw))
        self.preprocessed = dev.empty((3, ph, pw))
        self.aq: dict = {}
        self.cfl: dict = {}
        self.costs = dev.empty((7, ph // 8, pw // 8))
        self.cells = dev.empty((ph // 8, pw // 8), np.uint8)
        nbytes = library().gjxl_amd_initial_quant_field_scratch_bytes(pw, ph)
        self.aq_scratch = dev.empty((nbytes,), np.uint8)
        self.session = AqSession(dev, width, height)
        self.butteraugli = Butteraugli(dev, width, height)
        self.tokenizer = Tokenizer(dev, width, height)
        self.token_count = 0
        s = self.session
        store = ((s.bw + 31) // 32) * ((s.bh + 31) // 32) * 65536
        self.quantized = s.buffer("quantized", (3, store), np.int32)
        self.raw_quant = s.buffer("raw_quant", (s.bh, s.bw), np.int32)
        self.field = s.buffer("field", (s.bh, s.bw))
        self.adjusted = s.buffer("adjusted_field", (s.bh, s.bw))
        self.block_distance = s.buffer("block_distance", (s.bh, s.bw))
        self.reconstructed_linear = s.buffer("linear", (3, height, width))
        self.score_words = s.buffer("scores", (64,), np.uint32)
        self.score_bits = self.butteraugli.buffer("score")
        self.distance_map = self.butteraugli.buffer("map")

    def close(self) -> None:
        self.session.close()
        self.butteraugli.close()
        self.tokenizer.close()

    def run(self, srgb: DeviceArray, bytes_per_pixel: int = 3,
            mark: Callable[[str], None] | None = None) -> None:
        dev, w, h, t = self.dev, self.width, self.height, self.target
        mark = mark or (lambda name: None)
        dev.srgb8_to_linear_and_opsin(srgb, w, h, bytes_per_pixel=bytes_per_pixel,
                                      linear=self.linear, opsin=self.opsin)
        mark("srgb_to_linear_and_opsin")
        dev.gaborish_inverse(self.opsin, out=self.preprocessed)
        mark("gaborish_inverse")
        dev.initial_quant_field(self.opsin, t, 1.0, out=self.aq, scratch=self.aq_scratch)
        mark("initial_quant_field")
        dev.initial_cfl(self.preprocessed, out=self.cfl)
        mark("initial_cfl")
        dev.ac_strategy_costs(self.preprocessed, self.aq["quant_field"],
                              self.aq["pixel_mask"], self.cfl["y_to_x"],
                              self.cfl["y_to_b"], t, costs=self.costs)
        mark("ac_strategy_costs")
        dev.ac_strategy_search(self.costs, t, cells=self.cells)
        mark("ac_strategy_search")
        self.session.prepare(self.opsin, self.preprocessed, self.cells,
                             self.aq["quant_field"], t)
        mark("aq_prepare")
        self.butteraugli.prepare(self.linear)
        mark("butteraugli_prepare")
        # the AQ loop, call by call (the same device work as
        # AqSession.run, with a mark after each part)
        dev.copy(self.field, self.adjusted)
        for it in range(self.iterations + 1):
            self.session.evaluate(self.field)
            mark("aq_evaluate")
            self.butteraugli.compare(self.reconstructed_linear)
            mark("butteraugli_compare")
            dev.copy(self.score_words, self.score_bits, 4, dst_offset=4 * it)
            reduce_distance_map(dev, self.distance_map, self.cells, self.block_distance)
            if it < self.iterations:
                self.session.update(self.field, self.block_distance, it, t)
            mark("aq_reduce_update")
        # the reference fails the loop on an invalid block distance
        self.session.check()
        # the effort-7 coefficient orders and block-context map, chosen on
        # the host from the final frame's statistics (waits for them)
        self.statistics = self.tokenizer.frame_statistics(self.quantized, self.cells,
                                                          self.raw_quant)
        self.entropy_inputs = entropy_inputs(self.statistics, 7)
        self.tokenizer.set_entropy_inputs(self.entropy_inputs)
        mark("entropy_inputs")
        # AC tokens of the last evaluation's frame (waits for the count)
        self.token_count = self.tokenizer.tokenize(self.quantized, self.cells,
                                                   self.raw_quant)
        mark("ac_tokenize")

    def scores(self) -> list[float]:
        words = self.session.buffer("scores", (self.iterations + 1,), np.uint32).numpy()
        return [float(v) for v in words.view(np.float32)]
