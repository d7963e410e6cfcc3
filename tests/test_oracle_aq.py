"""CPU: pin the oracle's AQ-evaluation pieces against the reference's own
compiled sources (quantization.cpp, chroma_from_luma.cpp, dc_conversion.cpp,
vardct/dc_quantization.cpp, dc_smoothing.cpp, epf.cpp, loop_filter.cpp), and
the orchestration around them (reconstruction.cpp, adaptive_quantization.cpp,
behind the reference's toolchain gate) by round-trip properties."""
import ctypes

import numpy as np
import pytest

from gjxl_amd import synthetic
from oracle import binding

F = np.float32
STRATEGIES = (0, 4, 5, 6, 7, 10, 11)


def _bits(a):
    return np.asarray(a).view(np.uint32)


@pytest.fixture(scope="module")
def scene():
    w, h = 200, 136
    srgb = synthetic.generate(w, h, 4)
    fe = binding.frontend(srgb)
    setup = binding.aq_setup(fe, w, h)
    ev = binding.aq_evaluate(fe, setup, setup["adjusted"], w, h)
    return w, h, fe, setup, ev


def test_quantizer_from_field_equals_reference(ref, scene):
    binding._declare_aq()
    _, _, fe, setup, ev = scene
    field = setup["adjusted"]
    bh, bw = field.shape
    raw = np.empty(field.shape, np.int32)
    params = np.empty(2, np.uint32)
    assert ref.ref_quantizer_from_field(setup["quant_dc"], binding.ptr(field),
                                        bw, bh, binding.ptr(raw),
                                        binding.ptr(params)) == 0
    assert np.array_equal(raw, ev["field_raw"])
    assert params[0] == int(ev["q"][2]) and params[1] == int(ev["q"][3])


def test_final_cfl_equals_reference(ref, scene):
    _, _, fe, setup, ev = scene
    pre = fe["pre"]
    _, ph, pw = pre.shape
    tiles = ev["y_to_x"].shape
    rx = np.empty(tiles, np.int8)
    rb = np.empty(tiles, np.int8)
    assert ref.ref_final_cfl(binding.ptr(pre), pw, ph,
                             binding.ptr(np.ascontiguousarray(fe["cells"])),
                             binding.ptr(ev["field_raw"]), int(ev["q"][2]),
                             int(ev["q"][3]), 1, binding.ptr(rx),
                             binding.ptr(rb)) == 0
    assert np.array_equal(rx, ev["y_to_x"]) and np.array_equal(rb, ev["y_to_b"])


@pytest.mark.parametrize("strategy", STRATEGIES)
def test_code_transform_equals_reference(ref, strategy):
    binding._declare_aq()
    cw, ch = binding.COVERED[strategy]
    n = cw * ch * 64
    rng = np.random.default_rng(strategy)
    q = np.array([0.0782, 12.79, 5122, 17, 1.8e-4, 1.47e-3, 2.94e-3,
                  5442, 680, 340], F)
    # a real quantizer (global scale 5122, quant dc 17)
    _, q = binding.quantizer_from_field(np.full((4, 4), 0.4, F), 1.3448377)
    for trial in range(8):
        pix = rng.normal(0, 0.05 * (trial + 1), (3, n)).astype(F)
        coef = np.empty_like(pix)
        for k in range(3):
            binding.oracle().oracle_forward_dct(strategy, binding.ptr(pix[k]),
                                                binding.ptr(coef[k]))
        raw = int(rng.integers(1, 60))
        want_coef = coef.copy()
        want_q = np.empty((3, n), np.int32)
        want_raw = ctypes.c_int32()
        assert ref.ref_code_transform(strategy, int(q[2]), int(q[3]), raw,
                                      1.25, 1.5625, 0.1, 1.05,
                                      binding.ptr(want_coef), binding.ptr(want_q),
                                      ctypes.byref(want_raw)) == 0
        got_coef = coef.copy()
        got_q = np.empty((3, n), np.int32)
        dc = np.empty((3, 16), F)
        got_raw = ctypes.c_int32()
        binding.oracle().oracle_code_transform(
            strategy, binding.ptr(q), raw, 1.25, 1.5625, 0.1, 1.05,
            binding.ptr(got_coef), binding.ptr(got_q), binding.ptr(dc),
            ctypes.byref(got_raw))
        assert got_raw.value == want_raw.value
        assert np.array_equal(got_q, want_q)
        assert np.array_equal(_bits(got_coef), _bits(want_coef))


@pytest.mark.parametrize("strategy", STRATEGIES)
def test_dc_conversion_equals_reference(ref, strategy):
    binding._declare_aq()
    cw, ch = binding.COVERED[strategy]
    n = cw * ch * 64
    coef = np.random.default_rng(1).normal(0, 1, n).astype(F)
    want = np.empty(16, F)
    got = np.empty(16, F)
    assert ref.ref_llf_to_dc(strategy, binding.ptr(coef), binding.ptr(want)) == 0
    binding.oracle().oracle_llf_to_dc(cw, ch, binding.ptr(coef), binding.ptr(got))
    assert np.array_equal(_bits(got[:cw * ch]), _bits(want[:cw * ch]))
    back_w = coef.copy()
    back_g = coef.copy()
    assert ref.ref_dc_to_llf(strategy, binding.ptr(want), binding.ptr(back_w)) == 0
    binding.oracle().oracle_dc_to_llf(cw, ch, binding.ptr(got), binding.ptr(back_g))
    assert np.array_equal(_bits(back_g), _bits(back_w))
    # DC -> LLF -> DC is the identity up to float rounding
    assert np.allclose(back_g[:1], coef[:1], atol=1e-5)


def test_quantize_dc_equals_reference(ref, scene):
    binding._declare_aq()
    _, _, fe, setup, ev = scene
    rng = np.random.default_rng(2)
    for bw, bh in ((25, 17), (300, 40), (3, 260)):
        dc = rng.normal(0, 0.2, (3, bh, bw)).astype(F)
        dc[1] = np.cumsum(dc[1], axis=1) * F(0.1)  # correlated luma
        q = ev["q"]
        want_q = np.empty((3, bh, bw), np.int32)
        want_r = np.empty((3, bh, bw), F)
        assert ref.ref_quantize_dc(binding.ptr(dc), bw, bh, int(q[2]), int(q[3]),
                                   binding.ptr(want_q), binding.ptr(want_r)) == 0
        got_q = np.empty_like(want_q)
        got_r = np.empty_like(want_r)
        binding.oracle().oracle_quantize_dc(binding.ptr(dc), bw, bh,
                                            binding.ptr(q[4:7]), binding.ptr(q[7:10]),
                                            1, binding.ptr(got_q), binding.ptr(got_r))
        assert np.array_equal(got_q, want_q)
        assert np.array_equal(_bits(got_r), _bits(want_r))


def test_smooth_dc_and_sigma_equal_reference(ref, scene):
    _, _, fe, setup, ev = scene
    q = ev["q"]
    dc = ev["dc_rec"]
    _, bh, bw = dc.shape
    want = np.empty_like(dc)
    assert ref.ref_smooth_dc(binding.ptr(dc), bw, bh, int(q[2]), int(q[3]),
                             binding.ptr(want)) == 0
    got = np.empty_like(dc)
    binding.oracle().oracle_smooth_dc(binding.ptr(dc), bw, bh, binding.ptr(q[4:7]),
                                      binding.ptr(got))
    assert np.array_equal(_bits(got), _bits(want))
    sig = np.empty((bh, bw), F)
    assert ref.ref_epf_inverse_sigma(binding.ptr(np.ascontiguousarray(fe["cells"])),
                                     bw, bh, binding.ptr(ev["raw"]), int(q[2]),
                                     int(q[3]), binding.ptr(sig)) == 0
    assert np.array_equal(_bits(sig), _bits(ev["inv_sigma"]))


def test_loop_filters_equal_reference(ref, scene):
    w, h, fe, setup, ev = scene
    cropped = np.ascontiguousarray(ev["reconstructed"][:, :h, :w])
    want = np.empty_like(cropped)
    assert ref.ref_loop_filters(binding.ptr(cropped), w, h,
                                binding.ptr(ev["inv_sigma"]), binding.ptr(want)) == 0
    assert np.array_equal(_bits(ev["filtered"]), _bits(want))


def test_evaluation_round_trip_properties(scene):
    w, h, fe, setup, ev = scene
    # raw quant only rises at anchors (adjustment), never below the field's
    cells = fe["cells"]
    anchors = (cells & 1) == 1
    assert np.all(ev["raw"][~anchors] == ev["field_raw"][~anchors])
    # reconstruction tracks the coded (Gaborish-inverted) input in XYB, and
    # a finer field (a larger global scale) makes it converge: a dropped
    # dequantisation, DC or CfL term would leave a floor here
    err = np.abs(ev["reconstructed"] - fe["pre"])
    assert float(err.mean()) < 0.05
    fine = binding.aq_evaluate(fe, setup, setup["adjusted"] * F(16), w, h)
    fine_err = np.abs(fine["reconstructed"] - fe["pre"])
    assert float(fine_err.mean()) < 0.25 * float(err.mean())
    assert float(fine_err.max()) < 0.05
    # and the filtered, colour-converted result approaches the source
    coarse_lin = float(np.abs(ev["linear"] - fe["linear"]).mean())
    fine_lin = float(np.abs(fine["linear"] - fe["linear"]).mean())
    assert coarse_lin < 0.06 and fine_lin < 0.5 * coarse_lin
    # final CfL maps are within the int8 range used by the format
    assert ev["y_to_x"].dtype == np.int8


def _composed_frame(ref, fe, setup, ev):
    """ComputeQuantizedCoefficients and ReconstructQuantizedCoefficients
    (src/codec/reconstruction.cpp:386-760,790-936) composed here, transform
    by transform in raster order of the anchors, from the reference's own
    compiled pieces: ForwardDctCpu, ConvertLowFrequenciesToDc,
    SelectAdjustedAcQuantization + the Y/X/B quantisation with CfL
    (ref_code_transform), QuantizeDcCoefficients, SmoothDcCoefficients,
    DequantizeAcBlock, ConvertDcToLowFrequencies and InverseDctCpu.  The
    quantizer and the final CfL map are the evaluation's (each pinned by its
    own test above)."""
    binding._declare_aq()
    cells = np.ascontiguousarray(fe["cells"])
    pre = fe["pre"]
    bh, bw = cells.shape
    q = ev["q"]
    gs, qdc = int(q[2]), int(q[3])
    mult = (F(setup["x_mult"]), F(1.0), F(setup["b_mult"]))
    fx, fb = ev["y_to_x"], ev["y_to_b"]
    raw = ev["field_raw"].copy()
    dc = np.zeros((3, bh, bw), F)
    gw = (bw + 31) // 32
    store = np.zeros((3, binding.store_size(bw, bh)), np.int32)
    used = {}
    anchors = []
    for by in range(bh):
        for bx in range(bw):
            c = int(cells[by, bx])
            if not c & 1:
                continue
            s = c >> 1
            cw, ch = binding.COVERED[s]
            n = cw * ch * 64
            coef = np.empty((3, n), F)
            for k in range(3):
                pix = np.ascontiguousarray(pre[k, by * 8:by * 8 + ch * 8, bx * 8:bx * 8 + cw * 8])
                assert ref.ref_forward_dct(s, binding.ptr(pix), binding.ptr(coef[k])) == 0
            block_dc = np.empty(16, F)
            assert ref.ref_llf_to_dc(s, binding.ptr(coef[1]), binding.ptr(block_dc)) == 0
            dc[1, by:by + ch, bx:bx + cw] = block_dc[:cw * ch].reshape(ch, cw)
            cfl_x = F(fx[by // 8, bx // 8]) / F(84)
            cfl_b = F(1) + F(fb[by // 8, bx // 8]) / F(84)
            quant = np.empty((3, n), np.int32)
            decided = ctypes.c_int32()
            assert ref.ref_code_transform(s, gs, qdc, int(raw[by, bx]), float(mult[0]),
                                          float(mult[2]), float(cfl_x), float(cfl_b),
                                          binding.ptr(coef), binding.ptr(quant),
                                          ctypes.byref(decided)) == 0
            raw[by, bx] = decided.value
            for k in (0, 2):
                assert ref.ref_llf_to_dc(s, binding.ptr(coef[k]), binding.ptr(block_dc)) == 0
                dc[k, by:by + ch, bx:bx + cw] = block_dc[:cw * ch].reshape(ch, cw)
            group = (by // 32) * gw + bx // 32
            off = used.get(group, 0)
            store[:, group * 65536 + off:group * 65536 + off + n] = quant
            used[group] = off + n
            anchors.append((bx, by, s, group * 65536 + off, cfl_x, cfl_b))
    dc_q = np.empty((3, bh, bw), np.int32)
    dc_rec = np.empty((3, bh, bw), F)
    assert ref.ref_quantize_dc(binding.ptr(dc), bw, bh, gs, qdc, binding.ptr(dc_q),
                               binding.ptr(dc_rec)) == 0
    smoothed = np.empty_like(dc_rec)
    assert ref.ref_smooth_dc(binding.ptr(dc_rec), bw, bh, gs, qdc, binding.ptr(smoothed)) == 0
    rec = np.zeros((3, bh * 8, bw * 8), F)
    for bx, by, s, slot, cfl_x, cfl_b in anchors:
        cw, ch = binding.COVERED[s]
        n = cw * ch * 64
        deq = np.empty((3, n), F)
        for k in range(3):
            qk = np.ascontiguousarray(store[k, slot:slot + n])
            assert ref.ref_dequantize(s, gs, qdc, int(raw[by, bx]), k, float(mult[k]),
                                      binding.ptr(qk), binding.ptr(deq[k])) == 0
        deq[0] = deq[0] + cfl_x * deq[1]
        deq[2] = deq[2] + cfl_b * deq[1]
        for k in range(3):
            block_dc = np.ascontiguousarray(smoothed[k, by:by + ch, bx:bx + cw]).reshape(-1)
            coef = np.ascontiguousarray(deq[k])
            assert ref.ref_dc_to_llf(s, binding.ptr(block_dc), binding.ptr(coef)) == 0
            pix = np.empty(n, F)
            assert ref.ref_inverse_dct(s, binding.ptr(coef), binding.ptr(pix)) == 0
            rec[k, by * 8:by * 8 + ch * 8, bx * 8:bx * 8 + cw * 8] = pix.reshape(ch * 8, cw * 8)
    return {"raw": raw, "quantized": store, "dc_q": dc_q, "dc_rec": dc_rec,
            "reconstructed": rec}


@pytest.mark.parametrize("w,h,seed,kinds", [(200, 136, 4, 3), (520, 300, 11, 6)])
def test_frame_coding_composed_from_reference_pieces(ref, w, h, seed, kinds):
    """The oracle's frame-level coding and reconstruction
    (oracle_code_coefficients, oracle_reconstruct) equal the reference's
    per-transform pieces composed in the reference's order, on frames with
    three strategies and (520x300) six of the seven and four AC groups."""
    srgb = synthetic.generate(w, h, seed, "textured")
    fe = binding.frontend(srgb)
    setup = binding.aq_setup(fe, w, h)
    ev = binding.aq_evaluate(fe, setup, setup["adjusted"], w, h)
    used = set(int(v) >> 1 for v in fe["cells"][fe["cells"] & 1 == 1].ravel())
    assert len(used) >= kinds, used
    want = _composed_frame(ref, fe, setup, ev)
    assert np.array_equal(ev["raw"], want["raw"])
    assert np.array_equal(ev["quantized"], want["quantized"])
    assert np.array_equal(ev["dc_q"], want["dc_q"])
    assert np.array_equal(_bits(ev["dc_rec"]), _bits(want["dc_rec"]))
    assert np.array_equal(_bits(ev["reconstructed"]), _bits(want["reconstructed"]))
