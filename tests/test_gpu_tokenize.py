"""GPU parity of the AC tokenizer (gjxl_amd_tokenize): token values,
contexts, group offsets and populations are bit-exact with the oracle
(oracle/tokenize.c, pinned to tests/ac_group_test.cpp's known answers) on the
same frames, from single transforms to the 4032x3024 grid, with natural and
custom orders, the default and a thresholded block-context map."""
import numpy as np
import pytest

import gjxl_amd
from oracle import binding
from tests import token_frames
from tests.token_frames import fnv_tokens, poison_llf

pytestmark = pytest.mark.gpu

COVERED = token_frames.COVERED


def _tokenize(device, cells, q, raw, tokenizer=None, populations=False):
    bh, bw = cells.shape
    tok = tokenizer or gjxl_amd.Tokenizer(device, bw * 8, bh * 8)
    print(f"tokenize {bw}x{bh} blocks, populations={populations}", flush=True)
    n = tok.tokenize(device.array(np.ascontiguousarray(q)),
                     device.array(np.ascontiguousarray(cells)),
                     device.array(np.ascontiguousarray(raw)), populations)
    out = tok.tokens(n)
    if populations:
        out["histogram"] = tok.histogram()
    if tokenizer is None:
        tok.close()
    return out


def _same(got, want, histogram=False):
    assert got["values"].size == want["values"].size
    assert np.array_equal(got["group_offsets"], want["group_offsets"])
    bad = np.flatnonzero((got["values"] != want["values"]) |
                         (got["contexts"] != want["contexts"]))
    assert bad.size == 0, (
        f"{bad.size} tokens differ, first at {bad[0]}: "
        f"({got['contexts'][bad[0]]}, {got['values'][bad[0]]}) vs "
        f"({want['contexts'][bad[0]]}, {want['values'][bad[0]]})")
    if histogram:
        assert np.array_equal(got["histogram"], want["histogram"])


@pytest.mark.parametrize("strategy,count,digest", [
    (0, 192, 0x4bc215f68f1ece9d), (4, 759, 0xb4eb38be028e0165),
    (5, 3027, 0x7fa649fd5ea7a5d5), (6, 381, 0xd61b35ea9efaa9e2),
    (7, 381, 0xd61b35ea9efaa9e2), (10, 1515, 0x85c67d0bcd6d70bc),
    (11, 1515, 0x85c67d0bcd6d70bc)])
def test_reference_token_goldens(device, strategy, count, digest):
    """tests/ac_group_test.cpp:130-219 on the device."""
    order = gjxl_amd.natural_order(strategy)
    n, cov = order.size, order.size // 64
    cw, ch = COVERED[strategy]
    q = np.zeros((3, 65536), np.int32)
    for c in range(3):
        q[c, :n] = poison_llf(strategy, 1000 + 100 * c)
        q[c, order[cov]] = -(c + 1)
        q[c, order[cov + (n - cov) // 2]] = c + 5
        q[c, order[n - 1]] = -(c + 10)
    cells = np.full((ch, cw), strategy << 1, np.uint8)
    cells[0, 0] |= 1
    got = _tokenize(device, cells, q, np.ones((ch, cw), np.int32))
    assert got["values"].size == count
    assert fnv_tokens(got["contexts"], got["values"]) == digest


def test_multiblock_prediction(device):
    """tests/ac_group_test.cpp:461-543: contexts 81, 4, 5 at tokens 0, 6, 18."""
    cells = np.full((3, 4), 1, np.uint8)
    cells[0, 0:2] = 7 << 1
    cells[0, 0] |= 1
    cells[1:3, 0] = 6 << 1
    cells[1, 0] |= 1
    q = np.zeros((3, 65536), np.int32)
    offset, k = 0, 0
    for y in range(3):
        for x in range(4):
            if not cells[y, x] & 1:
                continue
            st = int(cells[y, x] >> 1)
            order = gjxl_amd.natural_order(st)
            for c in range(3):
                q[c, offset:offset + order.size] = poison_llf(st, 700 + c)
                q[c, offset + order[order.size // 64]] = 100 + 10 * k + c
            offset += order.size
            k += 1
    got = _tokenize(device, cells, q, np.ones((3, 4), np.int32))
    assert got["values"].size == 60
    assert (got["contexts"][0], got["contexts"][6], got["contexts"][18]) == (81, 4, 5)
    _same(got, binding.tokenize_frame(q, cells, np.ones((3, 4), np.int32)))


@pytest.mark.parametrize("bw,bh,seed,density", [
    (1, 1, 0, 0.3), (5, 3, 1, 0.2), (32, 32, 2, 0.1), (33, 65, 3, 0.05),
    (40, 36, 4, 0.0), (64, 64, 5, 1.0), (100, 37, 6, 0.02), (504, 378, 7, 0.03)])
def test_random_frames(device, bw, bh, seed, density):
    rng = np.random.default_rng(seed)
    cells, q, raw = token_frames.frame(rng, bw, bh, density=density)
    _same(_tokenize(device, cells, q, raw), binding.tokenize_frame(q, cells, raw))


def test_populations(device):
    rng = np.random.default_rng(11)
    cells, q, raw = token_frames.frame(rng, 70, 45, density=0.1, spread=5000)
    _same(_tokenize(device, cells, q, raw, populations=True),
          binding.tokenize_frame(q, cells, raw, histogram=True), histogram=True)


def test_custom_orders_and_thresholds(device):
    rng = np.random.default_rng(12)
    cells, q, raw = token_frames.frame(rng, 96, 80, density=0.08)
    table = binding.natural_order_table()
    orders = {}
    mask = 0
    for family, st in ((0, 0), (2, 4), (3, 5), (4, 6), (6, 10)):
        mask |= 1 << family
        for c in range(3):
            nat = binding.natural_order(st)
            cov = nat.size // 64
            custom = nat.copy()
            custom[cov:] = rng.permutation(nat[cov:])
            orders[(family, c)] = custom
            for s2 in ((6, 7) if st == 6 else (10, 11) if st == 10 else (st,)):
                binding.order_slice(table, s2, c)[:] = custom
    thresholds = (20, 90, 200)
    labels = rng.integers(0, 9, size=3 * 13 * 4).astype(np.uint8)
    labels[:9] = np.arange(9)  # every label used
    bh, bw = cells.shape
    tok = gjxl_amd.Tokenizer(device, bw * 8, bh * 8)
    tok.set_orders(mask, orders)
    tok.set_context_map(thresholds, labels, 9)
    got = _tokenize(device, cells, q, raw, tokenizer=tok, populations=True)
    want = binding.tokenize_frame(q, cells, raw, table, thresholds, labels, 9, histogram=True)
    _same(got, want, histogram=True)
    # back to the natural orders and the default map
    tok.set_orders(0)
    tok.set_context_map((), binding.DEFAULT_CONTEXT_MAP, 4)
    _same(_tokenize(device, cells, q, raw, tokenizer=tok),
          binding.tokenize_frame(q, cells, raw))
    tok.close()


def test_invalid_inputs_are_rejected(device):
    tok = gjxl_amd.Tokenizer(device, 16, 16)
    cells = np.ones((2, 2), np.uint8)
    q = np.zeros((3, 65536), np.int32)
    with pytest.raises(gjxl_amd.GjxlError):  # raw quant outside [1, 256]
        _tokenize(device, cells, q, np.zeros((2, 2), np.int32), tokenizer=tok)
    bad = cells.copy()
    bad[0, 0] = (4 << 1) | 1  # a 16x16 over three DCT8 anchors
    with pytest.raises(gjxl_amd.GjxlError):
        _tokenize(device, bad, q, np.ones((2, 2), np.int32), tokenizer=tok)
    nat = gjxl_amd.natural_order(0)
    broken = nat.copy()
    broken[0], broken[1] = broken[1], broken[0]  # LLF prefix lost
    with pytest.raises(gjxl_amd.GjxlError):
        tok.set_orders(1, {(0, c): broken for c in range(3)})
    with pytest.raises(gjxl_amd.GjxlError):  # a label beyond num_contexts
        tok.set_context_map((), binding.DEFAULT_CONTEXT_MAP, 3)
    # the tokenizer still works after the rejections
    _same(_tokenize(device, cells, q, np.ones((2, 2), np.int32), tokenizer=tok),
          binding.tokenize_frame(q, cells, np.ones((2, 2), np.int32)))
    tok.close()


@pytest.mark.parametrize("size", [(64, 64), (200, 136)])
def test_plain_path_frame(device, size):
    """The last AQ evaluation's frame of the device path, tokenized on the
    device and by the oracle from the same quantized store."""
    from gjxl_amd import pipeline, synthetic

    w, h = size
    srgb = synthetic.generate(w, h, seed=3)
    path = pipeline.PlainPath(device, w, h)
    path.run(device.array(srgb))
    s = path.session
    q = s.buffer("quantized", (3, binding.store_size(s.bw, s.bh)), np.int32)
    raw = s.buffer("raw_quant", (s.bh, s.bw), np.int32)
    tok = gjxl_amd.Tokenizer(device, w, h)
    n = tok.tokenize(q, path.cells, raw)
    got = tok.tokens(n)
    want = binding.tokenize_frame(q.numpy(), path.cells.numpy(), raw.numpy())
    _same(got, want)
    tok.close()
    path.close()


@pytest.mark.parametrize("bw,bh,seed,weights", [
    (40, 40, 1, None), (70, 33, 2, None), (64, 64, 3, [1, 0, 0, 0, 0, 0, 0, 0]),
    (128, 96, 4, None), (504, 378, 5, None), (504, 378, 6, [1, 0, 0, 0, 0, 0, 0, 0])])
def test_frame_statistics_and_effort7_tokens(device, bw, bh, seed, weights):
    """The device's frame statistics equal the oracle's; the orders and map
    chosen from them, and the tokens with them, equal the oracle's effort-7
    tokens (504x378 blocks: the 4032x3024 grid, above the 8192-block raw-quant
    split; DCT8-only grids take the sampled counts)."""
    rng = np.random.default_rng(seed)
    cells, q, raw = token_frames.frame(rng, bw, bh, density=0.2, weights=weights)
    print(f"frame statistics {bw}x{bh} seed {seed}", flush=True)
    tok = gjxl_amd.Tokenizer(device, bw * 8, bh * 8)
    d_q, d_cells, d_raw = (device.array(np.ascontiguousarray(a)) for a in (q, cells, raw))
    stats = tok.frame_statistics(d_q, d_cells, d_raw).arrays()
    want = binding.effort7_tokens(q, cells, raw)
    ws = want["statistics"]
    assert stats["present_mask"] == ws["present_mask"]
    assert np.array_equal(stats["zero_counts"], ws["zero_counts"])
    assert np.array_equal(stats["sampled"], ws["sampled"])
    assert np.array_equal(stats["raw_counts"], ws["raw_counts"])
    inputs = gjxl_amd.entropy_inputs(tok.frame_statistics(d_q, d_cells, d_raw), 7)
    assert inputs.used_order_mask == want["used_order_mask"]
    assert np.array_equal(inputs.order_table(), want["orders"])
    tok.set_entropy_inputs(inputs)
    n = tok.tokenize(d_q, d_cells, d_raw)
    _same(tok.tokens(n), want)
    tok.close()


def test_frame_statistics_reject_raw_quant_out_of_range(device):
    rng = np.random.default_rng(7)
    cells, q, raw = token_frames.frame(rng, 40, 40)
    raw[3, 5] = 0
    tok = gjxl_amd.Tokenizer(device, 320, 320)
    with pytest.raises(gjxl_amd.GjxlError, match="raw quantization"):
        tok.frame_statistics(device.array(q), device.array(cells), device.array(raw))
    tok.close()
