"""CPU: the proposal network's oracle (oracle/proposal.py) pinned by what the
mathematics fixes -- central differences of the loss in binary64, a delta
input, the softmax closed forms, Adam's first step, a DCT basis pattern --
and bfloat16 rounding against torch's conversion."""
import numpy as np
import pytest
import torch

from oracle import proposal as P


def test_bf16_rounding_matches_torch_and_ties_to_even():
    assert P.bf16(np.float32(1.0)) == 1.0
    assert P.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0                  # tie -> even (down)
    assert P.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6   # tie -> even (up)
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(4096) * 10.0 ** rng.uniform(-6, 6, 4096)).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(P.bf16(x), want)


def test_conv_of_a_delta_is_the_weights():
    rng = np.random.default_rng(1)
    w = rng.standard_normal((9, 32, 16)).astype(np.float32)
    x = np.zeros((1, 6, 7, 16), np.float32)
    x[0, 2, 3, 5] = 1.0
    out = P.conv(x, w, None)
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        # out[q] = sum_t W[t] x[q + off(t)]: the delta at (2, 3) reaches
        # q = (2 - dy, 3 - dx) through tap t
        assert np.allclose(out[0, 2 - dy, 3 - dx], w[t, :, 5])
    assert np.count_nonzero(np.abs(out).sum(axis=-1)) == 9


def _problem(seed, n=2, h=5, w=6):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, h, w, P.FEATURES))
    params = P.init_params(seed).astype(np.float64)
    # biases away from zero so that few ReLUs sit at their kink
    layers = P.unpack(params)
    for wgt, b in layers:
        b[:] = rng.uniform(-0.3, 0.3, b.shape)
    params = np.concatenate([np.concatenate([wgt.reshape(-1), b]) for wgt, b in layers])
    labels = rng.choice(np.array(P.STRATEGIES + (P.IGNORE,), np.int64), size=(n, h, w))
    targets = rng.uniform(-1, 1, (n, h, w))
    return x, params, labels, targets


@pytest.mark.parametrize("seed", [3, 4])
def test_backward_equals_central_differences(seed):
    x, params, labels, targets = _problem(seed)
    qf_weight = 0.7
    ins, outs = P.forward(x, params, rounding=False)
    _, _, _, g = P.loss_and_output_grad(outs[-1], labels, targets, qf_weight, rounding=False)
    grads = P.backward(ins, outs, params, g, rounding=False)
    assert grads.shape == params.shape
    rng = np.random.default_rng(seed + 100)
    offsets = P.param_offsets() + [params.size]
    picks = []
    for L in range(P.LAYERS):  # weights and biases of every layer
        lo, hi = offsets[L], offsets[L + 1]
        picks += list(rng.integers(lo, hi - P.OUT, 6)) + list(rng.integers(hi - P.OUT, hi, 2))
    for i in picks:
        h = 1e-6 * max(1.0, abs(params[i]))
        up, dn = params.copy(), params.copy()
        up[i] += h
        dn[i] -= h
        fd = (P.loss_total(x, up, labels, targets, qf_weight) -
              P.loss_total(x, dn, labels, targets, qf_weight)) / (2 * h)
        assert abs(fd - grads[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, grads[i])


def test_softmax_closed_forms():
    head = np.zeros((2, 3, P.OUT))
    labels = np.array([[0, 5, P.IGNORE], [4, 10, 11]])
    targets = np.zeros((2, 3))
    ce, se, correct, g = P.loss_and_output_grad(head, labels, targets, 1.0, rounding=False)
    assert np.isclose(ce, 5 * np.log(27.0))
    assert se == 0.0
    live = labels != P.IGNORE
    assert np.allclose(g[..., :27].sum(axis=-1)[live], 0.0)
    assert np.allclose(g[~live], 0.0)
    assert np.isclose(g[0, 1, 5], 1.0 / 27 - 1.0)
    head[..., P.QF_CHANNEL] = 0.5
    _, se, _, g = P.loss_and_output_grad(head, labels, targets, 2.0, rounding=False)
    assert np.isclose(se, 5 * 0.25) and np.isclose(g[0, 0, P.QF_CHANNEL], 2.0)


def test_adam_first_step_is_a_signed_step():
    rng = np.random.default_rng(5)
    p = rng.standard_normal(100).astype(np.float32)
    g = rng.standard_normal(100).astype(np.float32)
    z = np.zeros(100, np.float32)
    new, m, v = P.adam(p, z, z, g, 1, lr=1e-3)
    assert np.allclose(new, p - 1e-3 * np.sign(g), atol=1e-8)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_features_of_a_basis_pattern():
    d = P.dct8_basis()
    opsin = np.zeros((3, 16, 8), np.float32)
    opsin[0, :8] = 0.01                               # flat X block
    opsin[1, :8] = 0.5 + 0.2 * np.outer(d[0], d[1])   # Y: DC + (v=0, u=1), band 0
    opsin[2, 8:] = 0.3 * np.outer(d[3], d[4])         # B (second block): u+v = 7, band 2
    qf = np.array([[0.8], [1.25]], np.float32)
    f = P.features(opsin, qf)
    assert f.shape == (2, 1, 16)
    assert np.isclose(f[0, 0, 0], P.bf16(np.float32(0.16)), rtol=0)
    assert np.isclose(f[0, 0, 1], P.bf16(np.float32(0.5)), rtol=2 ** -7)
    assert np.isclose(f[0, 0, 3 + 3], P.bf16(np.float32(np.log1p(100 * 0.04))), rtol=2 ** -7)
    # other bands hold only the rounding of the binary32 pattern
    assert abs(f[0, 0, 3 + 4]) < 1e-6 and abs(f[0, 0, 3 + 5]) < 1e-6
    assert abs(f[0, 0, 3 + 0]) < 1e-6
    assert np.isclose(f[1, 0, 3 + 8], P.bf16(np.float32(np.log1p(100 * 0.09))), rtol=2 ** -7)
    assert np.isclose(f[1, 0, 12], P.bf16(np.float32(np.log(1.25))), rtol=2 ** -7)
    assert np.all(f[..., 13:] == 0)


def test_decide_blocks_and_ranks():
    head = np.full((16, 8, P.OUT), -5.0, np.float32)
    # rows 0-7: confident DCT16x16, DCT8 second
    head[:8, :, 4] = 6.0
    head[:8, :, 0] = 2.0
    # rows 8-15: confident DCT32x32 but one unsure block
    head[8:, :, 5] = 6.0
    head[12, 3, 5] = -5.0
    allow = P.decide(head, threshold=0.9, top_k=2, p_min=0.2)
    assert np.all(allow[:8] == (1 << 0) | (1 << 3))
    unsure = np.zeros((8, 8), bool)
    unsure[4, 3] = True
    assert np.all(allow[8:][unsure] == 0x7F)
    # DCT32x32 (stage 6) and the runner-up, where the tie ranks stage 0 first
    assert np.all(allow[8:][~unsure] == (1 << 0) | (1 << 6))
    # equal probabilities (1/7 < p_min) rank to the lower stage: top 2 = stages 0, 1
    head2 = np.full((8, 8, P.OUT), 0.0, np.float32)
    allow2 = P.decide(head2, threshold=0.0, top_k=2, p_min=0.5)
    assert np.all(allow2 == 0b11)


def test_warm_start_bounds():
    head = np.zeros((2, 2, P.OUT), np.float32)
    head[..., P.QF_CHANNEL] = [[0.0, 5.0], [-5.0, np.log(2.0)]]
    adj = np.full((2, 2), 1.0, np.float32)
    f = P.warm_start(head, adj, 0.5, 4.0)
    assert f[0, 0] == 1.0 and f[0, 1] == 4.0 and f[1, 0] == np.float32(0.5)
    assert np.isclose(f[1, 1], 2.0, rtol=1e-6)
    g = P.warm_start(head, adj, 0.1, 4.0, offset=np.log(1.5))
    assert np.isclose(g[0, 0], 1.5, rtol=1e-6) and np.isclose(g[1, 1], 3.0, rtol=1e-6)
