"""GPU parity: initial CfL, AC-strategy candidate costs and the greedy merge.

CfL maps, candidate costs and strategy grids must be identical to the
oracle, bit for bit.  The cost's loss^(1/8) (src/codec/ac_strategy.cpp:979)
is the reference's powf on both sides: the host libm in the oracle and its
device port (gjxl_amd/csrc/libm_port.h) on the GPU.
"""
import numpy as np
import pytest

from gjxl_amd import synthetic
from oracle import binding

pytestmark = pytest.mark.gpu


def _inputs(w, h, seed, recipe="basic"):
    srgb = np.ascontiguousarray(
        synthetic.generate(max(w, 16), max(h, 16), seed, recipe)[:h, :w])
    ops = binding.linear_to_padded_opsin(binding.srgb8_to_linear(srgb))
    pre = binding.gaborish_inverse(ops)
    aq = binding.initial_quant_field(ops)
    return ops, pre, aq


def _check_costs(got, want):
    nan_g, nan_w = np.isnan(got), np.isnan(want)
    assert np.array_equal(nan_g, nan_w), "evaluated-anchor sets differ"
    g, w = got[~nan_g], want[~nan_w]
    differ = np.count_nonzero(g.view(np.uint32) != w.view(np.uint32))
    assert differ == 0, f"{differ} of {g.size} costs differ"


@pytest.mark.parametrize("w,h,seed", [(512, 512, 1), (200, 136, 4),
                                      (72, 40, 7), (8, 8, 3), (136, 264, 9)])
def test_ac_search_parity(device, w, h, seed):
    ops, pre, aq = _inputs(w, h, seed)
    d_pre = device.array(pre)
    cfl = device.initial_cfl(d_pre)
    want_x, want_b = binding.initial_cfl(pre)
    assert np.array_equal(cfl["y_to_x"].numpy(), want_x)
    assert np.array_equal(cfl["y_to_b"].numpy(), want_b)
    costs = device.ac_strategy_costs(
        d_pre, device.array(aq["quant_field"]), device.array(aq["pixel_mask"]),
        cfl["y_to_x"], cfl["y_to_b"])
    want_costs = binding.candidate_costs(pre, aq["quant_field"],
                                         aq["pixel_mask"], want_x, want_b)
    got_costs = costs.numpy()
    _check_costs(got_costs, want_costs)
    cells = device.ac_strategy_search(costs).numpy()
    # the device merge equals the oracle merge over the same cost table ...
    assert np.array_equal(cells, binding.search_from_costs(got_costs))
    # ... and the oracle merge over the oracle's own costs
    assert np.array_equal(cells, binding.search_from_costs(want_costs))


def test_ac_search_full_size_4032x3024(device):
    srgb = synthetic.generate(**synthetic.WORKLOADS["4032x3024"])
    ops = binding.linear_to_padded_opsin(binding.srgb8_to_linear(srgb))
    d_ops = device.array(ops)
    d_pre = device.gaborish_inverse(d_ops)
    aq = device.initial_quant_field(d_ops)
    cfl = device.initial_cfl(d_pre)
    costs = device.ac_strategy_costs(d_pre, aq["quant_field"], aq["pixel_mask"],
                                     cfl["y_to_x"], cfl["y_to_b"])
    cells = device.ac_strategy_search(costs).numpy()
    got_costs = costs.numpy()
    # merge parity at full size over the device's costs
    assert np.array_equal(cells, binding.search_from_costs(got_costs))
    # CfL parity at full size
    pre = d_pre.numpy()
    want_x, want_b = binding.initial_cfl(pre)
    assert np.array_equal(cfl["y_to_x"].numpy(), want_x)
    assert np.array_equal(cfl["y_to_b"].numpy(), want_b)
    # costs of a strip of tiles against the oracle (the whole image is
    # minutes of single-threaded CPU work)
    strip = np.ascontiguousarray(pre[:, :128, :])
    aq_np = {k: v.numpy() for k, v in aq.items()}
    want = binding.candidate_costs(
        strip, np.ascontiguousarray(aq_np["quant_field"][:16]),
        np.ascontiguousarray(aq_np["pixel_mask"][:128]),
        np.ascontiguousarray(want_x[:2]), np.ascontiguousarray(want_b[:2]))
    _check_costs(got_costs[:, :16], want)
    kinds = np.unique(cells >> 1)
    assert len(kinds) >= 4, kinds
