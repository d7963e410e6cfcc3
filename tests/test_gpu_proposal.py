"""GPU: the learned proposal's kernels against the oracle (oracle/proposal.py),
and its training on the device.

Tolerances (floating point; the kernels accumulate bfloat16 products in
binary32 on MFMA in an order of their own, so a stored bfloat16 activation
may differ from the oracle's by one unit in the last place and carry that
on):
* features: within 2^-6 relative + 1e-3 absolute (one bfloat16 step);
* forward head: max error within 3 % of the head's largest magnitude, and
  the mean error within 0.3 %;
* gradients, per layer: relative L2 error under 3 %, cosine above 0.999;
* Adam: relative 1e-5 against the oracle's step on the device's gradients;
* decisions: equal, except blocks whose probability lies within 1e-4 of
  the threshold, p_min or another stage's probability;
* warm start: relative 1e-6.
"""
import ctypes

import numpy as np
import pytest

import gjxl_amd
from gjxl_amd import proposal as prop
from oracle import proposal as P

pytestmark = pytest.mark.gpu


def _to_bits(x):
    return (P.bf16(x).view(np.uint32) >> 16).astype(np.uint16)


def _from_bits(b):
    return (b.astype(np.uint32) << 16).view(np.float32)


def _params(seed):
    params = prop.init_parameters(seed)
    rng = np.random.default_rng(seed + 7)
    layers = P.unpack(params.copy())
    for w, b in layers:
        b[:] = rng.uniform(-0.2, 0.2, b.shape)
    return P.pack(layers)


def _problem(seed, n=2, h=13, w=21):
    rng = np.random.default_rng(seed)
    x = _to_bits(rng.standard_normal((n, h, w, P.FEATURES)).astype(np.float32))
    labels = rng.choice(np.array(P.STRATEGIES + (P.IGNORE,), np.uint8), size=(n, h, w))
    targets = rng.uniform(-1, 1, (n, h, w)).astype(np.float32)
    return x, labels, targets


def test_features_match_oracle(device):
    img = gjxl_amd.synthetic.generate(200, 136, seed=21)
    srgb = device.array(img)
    _, opsin = device.srgb8_to_linear_and_opsin(srgb, 200, 136)
    pw, ph = 200, 136
    field = device.initial_quant_field(opsin, 1.0)["quant_field"]
    f = prop.features(device, opsin, pw, ph, field).numpy()
    want = P.features(opsin.numpy().reshape(3, ph, pw), field.numpy().reshape(ph // 8, pw // 8))
    got = _from_bits(f)
    assert got.shape == want.shape
    assert np.all(np.abs(got - want) <= 2.0 ** -6 * np.abs(want) + 1e-3)
    assert np.all(got[..., 13:] == 0)


def test_forward_matches_oracle(device):
    x, _, _ = _problem(1)
    n, h, w, _ = x.shape
    params = _params(1)
    net = prop.ProposalNet(device, n * h * w, params)
    feats = device.array(x)
    net.forward(feats, n, h, w)
    head = net.head(n * h * w).numpy().reshape(n, h, w, P.OUT)
    _, outs = P.forward(_from_bits(x), params)
    want = outs[-1]
    scale = np.abs(want).max()
    err = np.abs(head - want)
    assert err.max() <= 0.03 * scale, (err.max(), scale)
    assert err.mean() <= 0.003 * scale
    # the stored activations of the first layer: at most one bfloat16 step
    act = _from_bits(net.activations().numpy()[0, :n * h * w]).reshape(n, h, w, P.OUT)
    assert np.all(np.abs(act - outs[0]) <= 2.0 ** -7 * np.abs(outs[0]) + 1e-6)
    net.close()


@pytest.mark.parametrize("seed", [2, 3])
def test_backward_matches_oracle(device, seed):
    x, labels, targets = _problem(seed)
    n, h, w, _ = x.shape
    params = _params(seed)
    net = prop.ProposalNet(device, n * h * w, params)
    feats, lab, tgt = device.array(x), device.array(labels), device.array(targets)
    net.forward(feats, n, h, w)
    ce, se, correct = net.backward(lab, tgt, 0.5)
    ins, outs = P.forward(_from_bits(x), params)
    ce0, se0, correct0, g = P.loss_and_output_grad(outs[-1], labels, targets, 0.5)
    assert ce == pytest.approx(ce0, rel=1e-2) and se == pytest.approx(se0, rel=2e-2, abs=1e-3)
    assert abs(correct - correct0) <= 2
    want = P.backward(ins, outs, params, g)
    got = net.gradients()
    offsets = P.param_offsets() + [params.size]
    for L in range(P.LAYERS):
        a, b = got[offsets[L]:offsets[L + 1]], want[offsets[L]:offsets[L + 1]]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
        assert rel < 0.03 and cos > 0.999, (L, rel, cos)
    net.close()


def test_adam_matches_oracle(device):
    x, labels, targets = _problem(4)
    n, h, w, _ = x.shape
    params = _params(4)
    net = prop.ProposalNet(device, n * h * w, params)
    # the features stay alive until the backward has read them
    feats, lab, tgt = device.array(x), device.array(labels), device.array(targets)
    net.forward(feats, n, h, w)
    net.backward(lab, tgt, 1.0, sums=False)
    grads = net.gradients()
    net.adam(1e-3, grad_scale=0.25)
    got = net.parameters()
    want, m, v = P.adam(params, np.zeros_like(params), np.zeros_like(params), grads, 1,
                        lr=1e-3, scale=0.25)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-9)
    # a second step from the device's state
    net.forward(feats, n, h, w)
    net.backward(lab, tgt, 1.0, sums=False)
    g2 = net.gradients()
    net.adam(1e-3, grad_scale=0.25)
    want2, _, _ = P.adam(got, m, v, g2, 2, lr=1e-3, scale=0.25)
    assert np.allclose(net.parameters(), want2, rtol=1e-5, atol=1e-9)
    net.close()


def _decide_margin(head, threshold, p_min):
    z = head[..., list(P.STRATEGIES)].astype(np.float64)
    z -= z.max(axis=-1, keepdims=True)
    p = np.exp(z)
    p /= p.sum(axis=-1, keepdims=True)
    srt = np.sort(p, axis=-1)
    gaps = np.diff(srt, axis=-1).min(axis=-1)
    near = (np.abs(p.max(axis=-1) - threshold) < 1e-4) | (np.abs(p - p_min).min(axis=-1) < 1e-4)
    return near | (gaps < 1e-4)


def test_decide_and_warm_start_match_oracle(device):
    rng = np.random.default_rng(5)
    bh, bw = 37, 45
    head = (rng.standard_normal((bh, bw, P.OUT)) * 2.0).astype(np.float32)
    # a few confident regions
    head[:8, :16, 4] += 12.0
    head[16:24, 8:24, 0] += 12.0
    d_head = device.array(head.reshape(-1, P.OUT))
    for threshold, top_k, p_min in ((0.8, 2, 0.1), (0.0, 1, 1.1), (0.5, 3, 0.05)):
        got = prop.decide(device, d_head, bw, bh, threshold, top_k, p_min).numpy()
        want = P.decide(head, threshold, top_k, p_min)
        near = _decide_margin(head, threshold, p_min)
        assert np.array_equal(got[~near], want[~near]), (threshold, top_k, p_min)
        assert (got[:8, :16] != 0x7F).all()
    adjusted = rng.uniform(0.3, 3.0, (bh, bw)).astype(np.float32)
    for offset in (0.0, 0.1):
        field = prop.warm_start(device, d_head, device.array(adjusted), bh * bw, 2.0, 0.4,
                                2.5, offset=offset).numpy().reshape(bh, bw)
        want = P.warm_start(head, adjusted, 0.4, 2.5, offset=offset)
        assert np.allclose(field, want, rtol=1e-6, atol=0)


def test_training_lowers_the_loss(device):
    """Forty steps of the device's own training on labels of plain runs."""
    feats, labels, targets = prop.collect(device, 4, 512, 512, seed0=50, workers=4)
    assert feats.shape == (16, 32, 32, 16)  # four 32 x 32-block crops per image
    assert set(np.unique(labels)) <= set(P.STRATEGIES)
    params, history = prop.train(device, feats, labels, targets, epochs=10, batch=4,
                                 learning_rate=3e-3)
    first, last = history[0], history[-1]
    loss0 = first["cross_entropy"] + first["field_mse"]
    loss1 = last["cross_entropy"] + last["field_mse"]
    print(f"training loss {loss0:.4f} -> {loss1:.4f}, accuracy {last['accuracy']:.3f}",
          flush=True)
    assert loss1 < 0.6 * loss0
    assert np.isfinite(params).all()


def _bits(tokens):
    """Adaptive-free entropy of the token stream given its contexts (bits):
    the size an ideal coder with per-context static statistics reaches."""
    v = tokens["values"].astype(np.uint64)
    c = tokens["contexts"].astype(np.uint64)
    _, cnt = np.unique((c << np.uint64(32)) | v, return_counts=True)
    _, ccnt = np.unique(c, return_counts=True)
    return float(-(cnt * np.log2(cnt)).sum() + (ccnt * np.log2(ccnt)).sum())


def _quality_and_bits(path):
    d = path.view("block_distance").numpy().astype(np.float64)
    return float(np.mean(d ** 3) ** (1.0 / 3.0)), _bits(path.tokens())


def _rate_excess(plain, q, b):
    """b over the plain curve's rate at quality q (log-rate interpolated
    linearly in the 3-norm), minus one."""
    qs = np.array([p[0] for p in plain])
    lb = np.log(np.array([p[1] for p in plain]))
    return float(b / np.exp(np.interp(q, qs, lb)) - 1.0)


def test_trained_network_gains_over_none(device):
    """The committed network against the same learned run without one (zero
    parameters: the loop starts from the adjusted field and every stage
    stays allowed), on a held-out image (training seeds are 1000-1434), both
    stopping after one or two evaluations.  Rate excess is measured against
    the plain path's own rate-quality curve (distances 0.9-1.2, three
    evaluations) at equal 3-norm of the block distances."""
    w, h = 1024, 768
    img = gjxl_amd.synthetic.generate(w, h, seed=4321, recipe="textured")
    path = gjxl_amd.NativePlainPath(device, w, h)
    plain = []
    for dist in (0.9, 1.0, 1.1, 1.2):
        path.reconfigure(w, h, distance=dist)
        path.run(img)
        plain.append(_quality_and_bits(path))
    path.reconfigure(w, h, distance=1.0)
    path.run(img)
    bw, bh = path.bw, path.bh
    feats, labels, targets = (device.empty((bh, bw, 16), np.uint16),
                              device.empty((bh, bw), np.uint8), device.empty((bh, bw), np.float32))
    path.proposal_sample(feats, labels, targets)
    f, lab, tgt = feats.numpy()[None], labels.numpy()[None], targets.numpy()[None]
    trained = prop.load_parameters()
    zero = np.zeros_like(trained)
    cls_trained = prop.evaluate(device, trained, f, lab, tgt, batch=1)
    cls_zero = prop.evaluate(device, zero, f, lab, tgt, batch=1)
    print(f"held-out classification: trained {cls_trained}, no network {cls_zero}", flush=True)
    assert cls_trained["top2"] >= 0.8 and cls_trained["top1"] >= 0.6
    assert cls_trained["field_mse"] < 0.6 * cls_trained["field_mse_zero"]
    excess = {}
    for name, params in (("trained", trained), ("none", zero)):
        net = prop.ProposalNet(device, bw * bh, params)
        for evaluations in (1, 2):
            path.set_proposal(net, evaluations=evaluations, judge_ratio=100.0)
            path.run(img)
            assert path.frame.evaluations == evaluations
            q, b = _quality_and_bits(path)
            excess[name, evaluations] = _rate_excess(plain, q, b)
        path.set_proposal(None)
        net.close()
    print(f"rate excess at equal quality: {excess}", flush=True)
    # the trained warm start costs less rate than none at every stopping point
    for evaluations in (1, 2):
        assert excess["trained", evaluations] < excess["none", evaluations]
    assert excess["trained", 2] < 0.01
    path.close()


def test_learned_run_defaults_and_judge(device):
    """Default options: the judged 3-norm is reported and bounded, the
    allowed stages always include DCT8, and a run with an unreachable judge
    falls back to the plain loop's three evaluations."""
    w, h = 512, 512
    img = gjxl_amd.synthetic.generate(**gjxl_amd.synthetic.WORKLOADS["512x512"])
    path = gjxl_amd.NativePlainPath(device, w, h)
    net = prop.ProposalNet(device, path.bw * path.bh, prop.load_parameters())
    path.set_proposal(net)
    assert path.run(img) > 0
    allow = path.view("proposal_allow").numpy()
    assert (allow & 1).all() and (allow <= 0x7F).all()
    assert path.frame.evaluations in (2, 3)
    norm = path.frame.proposal_norm
    assert 0.0 < norm and (path.frame.evaluations == 3 or norm <= 0.7)
    # the judged norm is the 3-norm of the last evaluation's block distances
    d = path.view("block_distance").numpy().astype(np.float64)
    assert norm == pytest.approx(np.mean(d ** 3) ** (1.0 / 3.0), rel=1e-6)
    path.set_proposal(net, judge_ratio=1e-3)
    path.run(img)
    assert path.frame.evaluations == 3
    with pytest.raises(gjxl_amd.GjxlError, match="INVALID_ARGUMENT"):
        path.set_proposal(net, evaluations=4)
    small = prop.ProposalNet(device, 16, prop.load_parameters())
    with pytest.raises(gjxl_amd.GjxlError, match="fewer positions"):
        path.set_proposal(small)
    with pytest.raises(gjxl_amd.GjxlError, match="INVALID_ARGUMENT"):
        path.set_proposal(net, final_compare=2)
    path.set_proposal(None)
    path.run(img)
    assert path.frame.evaluations == 3 and not path.frame.proposal_allow
    small.close()
    net.close()
    path.close()


def test_network_on_another_handle(device):
    """A network created on another device handle (its own stream) runs on
    the path's streams: the learned run's tokens equal those with the
    network on the path's own handle."""
    w, h = 512, 512
    img = gjxl_amd.synthetic.generate(**gjxl_amd.synthetic.WORKLOADS["512x512"])
    path = gjxl_amd.NativePlainPath(device, w, h)
    params = prop.load_parameters()
    own = prop.ProposalNet(device, path.bw * path.bh, params)
    path.set_proposal(own, evaluations=2, judge_ratio=100.0)
    path.run(img)
    want = path.tokens()
    other_device = gjxl_amd.Device(0)
    other = prop.ProposalNet(other_device, path.bw * path.bh, params)
    for final_compare in (1, 0):
        print(f"network on another handle, final_compare {final_compare}", flush=True)
        path.set_proposal(other, evaluations=2, judge_ratio=100.0, final_compare=final_compare)
        path.run(img)
        assert _same_tokens(path.tokens(), want)
    path.set_proposal(None)
    other.close()
    other_device.close()
    own.close()
    path.close()


def _same_tokens(a, b):
    return all(np.array_equal(a[k], b[k]) for k in ("values", "contexts", "group_offsets"))


def test_unjudged_final_pass_keeps_the_tokens(device):
    """final_compare 0: the last pass quantizes without reconstruction or a
    comparison, and the tokens equal those of the same run comparing that
    pass too (two passes, then one, which also skips the reference side);
    a plain run afterwards still gives the plain tokens."""
    w, h = 512, 512
    img = gjxl_amd.synthetic.generate(**gjxl_amd.synthetic.WORKLOADS["512x512"])
    path = gjxl_amd.NativePlainPath(device, w, h)
    path.run(img)
    plain = path.tokens()
    net = prop.ProposalNet(device, path.bw * path.bh, prop.load_parameters())
    for evaluations in (2, 1):
        print(f"unjudged final pass: {evaluations} passes", flush=True)
        path.set_proposal(net, evaluations=evaluations, judge_ratio=100.0)
        path.run(img)
        assert path.frame.evaluations == evaluations
        compared = path.tokens()
        path.set_proposal(net, evaluations=evaluations, judge_ratio=100.0, final_compare=0)
        assert path.run(img) > 0
        assert path.frame.evaluations == evaluations - 1
        assert np.isnan(path.frame.proposal_norm)
        assert _same_tokens(path.tokens(), compared)
    # a caller's struct from before final_compare: its fields over the
    # defaults (a compared last pass); a shorter one is refused
    opts = prop.ProposalOptions.defaults()
    opts.evaluations, opts.judge_ratio, opts.final_compare = 2, 100.0, 0
    opts.struct_size = prop.ProposalOptions.final_compare.offset
    lib = gjxl_amd.library()
    gjxl_amd._check(lib.gjxl_amd_plain_path_set_proposal(path.handle, net.handle,
                                                         ctypes.byref(opts)))
    path.run(img)
    assert path.frame.evaluations == 2
    opts.struct_size = prop.ProposalOptions.judge_ratio.offset
    with pytest.raises(gjxl_amd.GjxlError, match="too small"):
        gjxl_amd._check(lib.gjxl_amd_plain_path_set_proposal(path.handle, net.handle,
                                                             ctypes.byref(opts)))
    path.set_proposal(None)
    path.run(img)
    assert _same_tokens(path.tokens(), plain)
    net.close()
    path.close()
