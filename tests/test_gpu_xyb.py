"""GPU parity: sRGB8 -> linear -> padded XYB, bit-exact against the oracle."""
import numpy as np
import pytest

from gjxl_amd import synthetic
from oracle import binding

pytestmark = pytest.mark.gpu


def _run(device, srgb, bpp=3, stride=None):
    h, w = srgb.shape[:2]
    if stride is not None:
        padded = np.zeros((h, stride), np.uint8)
        padded[:, : w * bpp] = srgb.reshape(h, w * bpp)
        buf = device.array(padded)
    else:
        buf = device.array(np.ascontiguousarray(srgb))
    lin, ops = device.srgb8_to_linear_and_opsin(buf, w, h, bytes_per_pixel=bpp,
                                                row_stride=stride)
    return lin.numpy(), ops.numpy()


def _expect(srgb):
    lin = binding.srgb8_to_linear(srgb)
    return lin, binding.linear_to_padded_opsin(lin)


def _same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32),
                                                 b.view(np.uint32))


@pytest.mark.parametrize("w,h,seed", [(512, 512, 1), (37, 13, 5), (8, 8, 2),
                                      (1, 1, 3), (257, 129, 4)])
def test_srgb_to_opsin_bit_exact(device, w, h, seed):
    srgb = synthetic.generate(max(w, 16), max(h, 16), seed)[:h, :w]
    srgb = np.ascontiguousarray(srgb)
    lin, ops = _run(device, srgb)
    elin, eops = _expect(srgb)
    assert _same_bits(lin, elin)
    assert _same_bits(ops, eops)


def test_rgba_with_row_padding(device):
    rgb = np.ascontiguousarray(synthetic.generate(64, 40, 9)[:, :61])
    rgba = np.concatenate([rgb, np.full(rgb.shape[:2] + (1,), 255, np.uint8)], -1)
    lin, ops = _run(device, rgba, bpp=4, stride=61 * 4 + 12)
    elin, eops = _expect(rgb)
    assert _same_bits(lin, elin) and _same_bits(ops, eops)


def test_linear_entry_bit_exact(device):
    rng = np.random.default_rng(11)
    lin = rng.uniform(0, 1, (3, 45, 70)).astype(np.float32)
    ops = device.linear_rgb_to_padded_opsin(device.array(lin), 70, 45).numpy()
    assert _same_bits(ops, binding.linear_to_padded_opsin(lin))


def test_full_size_4032x3024(device):
    srgb = synthetic.generate(**synthetic.WORKLOADS["4032x3024"])
    lin, ops = _run(device, srgb)
    elin, eops = _expect(srgb)
    assert _same_bits(lin, elin) and _same_bits(ops, eops)
