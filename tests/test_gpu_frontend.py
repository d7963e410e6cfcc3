"""GPU parity: Gaborish inverse and the initial quant field, bit-exact
against the oracle (pixel mask included: exercises the device log1pf)."""
import numpy as np
import pytest

from gjxl_amd import synthetic
from oracle import binding

pytestmark = pytest.mark.gpu


def _bits(a):
    return a.view(np.uint32)


def _opsin(w, h, seed):
    srgb = np.ascontiguousarray(synthetic.generate(max(w, 16), max(h, 16), seed)[:h, :w])
    return binding.linear_to_padded_opsin(binding.srgb8_to_linear(srgb))


@pytest.mark.parametrize("w,h,seed", [(512, 512, 1), (8, 8, 2), (24, 16, 3),
                                      (75, 41, 4), (130, 66, 5)])
def test_gaborish_inverse_bit_exact(device, w, h, seed):
    ops = _opsin(w, h, seed)
    got = device.gaborish_inverse(device.array(ops), (1.0, 0.7, 1.15)).numpy()
    want = binding.gaborish_inverse(ops, (1.0, 0.7, 1.15))
    assert np.array_equal(_bits(got), _bits(want))


@pytest.mark.parametrize("w,h,seed,target", [(512, 512, 1, 1.0), (8, 8, 2, 1.0),
                                             (24, 16, 3, 1.35), (75, 41, 4, 2.5),
                                             (264, 136, 6, 0.5)])
def test_initial_quant_field_bit_exact(device, w, h, seed, target):
    ops = _opsin(w, h, seed)
    got = device.initial_quant_field(device.array(ops), target, 1.0)
    want = binding.initial_quant_field(ops, target, 1.0)
    for name in ("quant_field", "strategy_mask", "pixel_mask"):
        assert np.array_equal(_bits(got[name].numpy()), _bits(want[name])), name


def test_frontend_full_size_4032x3024(device):
    srgb = synthetic.generate(**synthetic.WORKLOADS["4032x3024"])
    lin = binding.srgb8_to_linear(srgb)
    ops = binding.linear_to_padded_opsin(lin)
    d_ops = device.array(ops)
    gab = device.gaborish_inverse(d_ops).numpy()
    assert np.array_equal(_bits(gab), _bits(binding.gaborish_inverse(ops)))
    got = device.initial_quant_field(d_ops, 1.0, 1.0)
    want = binding.initial_quant_field(ops, 1.0, 1.0)
    for name in ("quant_field", "strategy_mask", "pixel_mask"):
        assert np.array_equal(_bits(got[name].numpy()), _bits(want[name])), name
