"""Pins the AC tokenization oracle (oracle/tokenize.c) to the reference's own
known answers, tests/ac_group_test.cpp: FNV-1a hashes of the natural orders
(:56-126), of the token streams of a poisoned single-transform group per
strategy (:130-219), first-token block contexts (:221-271), custom orders per
family (:334-398), zero / dense / signed-extreme blocks (:400-459) and
multi-block prediction with coefficient offsets (:461-543).  What those leave
open (quant-field thresholds, whole frames) is pinned by hand-worked cases
and by decoding the stream back to the coefficients."""
import numpy as np
import pytest

from oracle import binding as b

from tests.token_frames import COVERED, fnv_tokens, fnv_values, poison_llf

FAMILY = {0: 0, 4: 2, 5: 3, 6: 4, 7: 4, 10: 6, 11: 6}


def single(strategy, storage, orders=None, raw=1, **kw):
    """Tokens of one group holding one transform at the origin."""
    cw, ch = COVERED[strategy]
    n = cw * ch * 64
    q = np.zeros((3, 65536), np.int32)
    q[:, :n] = storage
    cells = np.full((ch, cw), strategy << 1, np.uint8)
    cells[0, 0] |= 1
    return b.tokenize_frame(q, cells, np.full((ch, cw), raw, np.int32), orders, **kw)


@pytest.mark.parametrize("strategy,size,digest,mid,mid_value", [
    (0, 64, 0x3429f64e9a8101d3, 32, 35), (4, 256, 0x08f226b939195a15, 128, 135),
    (5, 1024, 0x4e30e44c269548f1, 512, 527), (6, 128, 0x6100fcbe75b9416b, 64, 15),
    (7, 128, 0x6100fcbe75b9416b, 64, 15), (10, 512, 0x14407aca0189e559, 256, 31),
    (11, 512, 0x14407aca0189e559, 256, 31)])
def test_natural_order_goldens(strategy, size, digest, mid, mid_value):
    order = b.natural_order(strategy)
    assert order.size == size and fnv_values(order) == digest
    assert order[mid] == mid_value and order[-1] == size - 1
    assert np.array_equal(np.sort(order), np.arange(size))
    cw, ch = COVERED[strategy]
    lw, lh = max(cw, ch), min(cw, ch)
    for y in range(lh):
        for x in range(lw):
            assert order[y * lw + x] == y * lw * 8 + x


@pytest.mark.parametrize("strategy,count,digest", [
    (0, 192, 0x4bc215f68f1ece9d), (4, 759, 0xb4eb38be028e0165),
    (5, 3027, 0x7fa649fd5ea7a5d5), (6, 381, 0xd61b35ea9efaa9e2),
    (7, 381, 0xd61b35ea9efaa9e2), (10, 1515, 0x85c67d0bcd6d70bc),
    (11, 1515, 0x85c67d0bcd6d70bc)])
def test_token_stream_goldens(strategy, count, digest):
    order = b.natural_order(strategy)
    n = order.size
    cov = n // 64
    storage = np.zeros((3, n), np.int32)
    for c in range(3):
        storage[c] = poison_llf(strategy, 1000 + 100 * c)
        storage[c, order[cov]] = -(c + 1)
        storage[c, order[cov + (n - cov) // 2]] = c + 5
        storage[c, order[n - 1]] = -(c + 10)
    t = single(strategy, storage)
    assert t["values"].size == count
    assert fnv_tokens(t["contexts"], t["values"]) == digest
    assert int(t["contexts"].max()) < 1980


@pytest.mark.parametrize("strategy,expected", [
    (0, (80, 82, 82)), (4, (80, 82, 82)), (5, (80, 82, 82)), (6, (81, 83, 83)),
    (7, (81, 83, 83)), (10, (81, 83, 83)), (11, (81, 83, 83))])
def test_first_token_block_contexts(strategy, expected):
    n = COVERED[strategy][0] * COVERED[strategy][1] * 64
    t = single(strategy, np.zeros((3, n), np.int32))
    assert tuple(t["contexts"]) == expected and not t["values"].any()


def test_zero_dense_and_signed_extremes():
    s = np.zeros((3, 64), np.int32)
    s[:, 0] = np.iinfo(np.int32).max  # the LLF is not counted
    t = single(0, s)
    assert list(zip(t["contexts"], t["values"])) == [(80, 0), (82, 0), (82, 0)]
    s = np.ones((3, 64), np.int32)
    s[:, 0] = -777
    t = single(0, s)
    assert t["values"].size == 192
    for k, ctx in ((0, 80), (64, 82), (128, 82)):
        assert (t["contexts"][k], t["values"][k]) == (ctx, 63)
    s = np.zeros((3, 64), np.int32)
    s[0, 1] = np.iinfo(np.int32).min
    s[0, 8] = np.iinfo(np.int32).max
    t = single(0, s)
    assert list(zip(t["contexts"].tolist(), t["values"].tolist())) == [
        (80, 0), (82, 2), (1127, 0xFFFFFFFF), (1067, 0xFFFFFFFE), (82, 0)]


def test_multiblock_prediction_and_offsets():
    # 4x3 blocks: 8x16 at (0,0), 16x8 at (0,1), DCT8 elsewhere
    cells = np.full((3, 4), 1, np.uint8)
    cells[0, 0:2] = 7 << 1
    cells[0, 0] |= 1
    cells[1:3, 0] = 6 << 1
    cells[1, 0] |= 1
    anchors = [(y, x) for y in range(3) for x in range(4) if cells[y, x] & 1]
    q = np.zeros((3, 65536), np.int32)
    offset = 0
    for k, (y, x) in enumerate(anchors):
        st = cells[y, x] >> 1
        order = b.natural_order(st)
        cov = order.size // 64
        for c in range(3):
            p = poison_llf(st, 700 + c)
            q[c, offset:offset + order.size] = p
            q[c, offset + order[cov]] = 100 + 10 * k + c
        offset += order.size
    t = b.tokenize_frame(q, cells, np.ones((3, 4), np.int32))
    ctx, val = t["contexts"], t["values"]
    assert val.size == len(anchors) * 6
    assert ctx[0] == 81 and ctx[6] == 4 and ctx[18] == 5

    def pack(v):
        return (v << 1) ^ (0 if v >= 0 else 0xFFFFFFFF)
    for k in range(len(anchors)):
        assert val[6 * k + 1] == pack(101 + 10 * k)
        assert val[6 * k + 3] == pack(100 + 10 * k)
        assert val[6 * k + 5] == pack(102 + 10 * k)


@pytest.mark.parametrize("strategy", sorted(COVERED))
def test_custom_order_per_family(strategy):
    natural = b.natural_order(strategy)
    llf = natural.size // 64
    table = b.natural_order_table()
    for c in range(3):
        sl = b.order_slice(table, strategy, c)
        sl[llf], sl[llf + 1] = sl[llf + 1], sl[llf]
    storage = np.zeros((3, natural.size), np.int32)
    for c in range(3):
        storage[c, natural[llf]] = 10 + c
        storage[c, natural[llf + 1]] = 20 + c
    nat = single(strategy, storage)
    cus = single(strategy, storage, orders=table)
    assert nat["values"].size == cus["values"].size == 9
    # the custom scan visits the two coefficients in the other order
    for k in range(3):
        assert list(nat["values"][3 * k + 1:3 * k + 3]) == list(cus["values"][3 * k + 2:3 * k:-1])
    assert not np.array_equal(nat["values"], cus["values"])
    # the table's other strategies keep their natural orders
    for other in COVERED:
        if FAMILY[other] != FAMILY[strategy]:
            assert np.array_equal(b.order_slice(table, other, 0), b.natural_order(other))


def test_quant_field_thresholds_select_the_segment():
    """BlockContextValidated (ac_group.cpp:115-131): the segment counts the
    thresholds below raw quant; the map is indexed [(row * 13 + family) *
    (thresholds + 1) + segment]."""
    thresholds = (4, 9)
    cmap = np.zeros(3 * 13 * 3, np.uint8)
    # row 0 (Y), family 0: segments 0, 1, 2 -> contexts 0, 1, 2
    cmap[0:3] = (0, 1, 2)
    # rows 1, 2 (X, B), family 0: -> 3, 4, 5 and 5, 4, 3
    cmap[13 * 3:13 * 3 + 3] = (3, 4, 5)
    cmap[26 * 3:26 * 3 + 3] = (5, 4, 3)
    for raw, seg in ((1, 0), (4, 0), (5, 1), (9, 1), (10, 2), (256, 2)):
        t = single(0, np.zeros((3, 64), np.int32), raw=raw, thresholds=thresholds,
                   context_map=cmap, num_contexts=6)
        # 32 predicted nonzeros: bucket 20
        assert list(t["contexts"]) == [20 * 6 + cmap[seg], 20 * 6 + cmap[39 + seg],
                                       20 * 6 + cmap[78 + seg]]
    with pytest.raises(ValueError):  # raw quant outside [1, 256]
        single(0, np.zeros((3, 64), np.int32), raw=0)
    with pytest.raises(ValueError):  # a label beyond num_contexts
        single(0, np.zeros((3, 64), np.int32), context_map=cmap, thresholds=thresholds,
               num_contexts=5)


def test_malformed_grid_is_rejected():
    cells = np.full((2, 2), 1, np.uint8)
    cells[0, 0] = (4 << 1) | 1  # a 16x16 overlapping three DCT8 anchors
    with pytest.raises(ValueError):
        b.tokenize_frame(np.zeros((3, 65536), np.int32), cells, np.ones((2, 2), np.int32))
    cells = np.zeros((1, 1), np.uint8)  # a cell no transform covers
    with pytest.raises(ValueError):
        b.tokenize_frame(np.zeros((3, 65536), np.int32), cells, np.ones((1, 1), np.int32))


def test_token_symbols():
    b._declare_tokenize()
    sym = b.oracle().oracle_token_symbol
    # HybridUint(4, 2, 0): tokens below 16 are literal, then exponent and
    # two mantissa bits
    assert [sym(v) for v in (0, 15, 16, 19, 20, 31, 32, 48, 0xFFFFFFFF)] == [
        0, 15, 16, 16, 17, 19, 20, 22, 127]


def random_frame(rng, bw, bh, density=0.2):
    """A tiled strategy grid and sparse quantized coefficients."""
    cells = np.zeros((bh, bw), np.uint8)
    for y in range(bh):
        for x in range(bw):
            if cells[y, x]:
                continue
            choices = [s for s, (cw, ch) in COVERED.items()
                       if x % cw == 0 and y % ch == 0 and x + cw <= bw and y + ch <= bh
                       and (x // 32) == ((x + cw - 1) // 32) and (y // 32) == ((y + ch - 1) // 32)
                       and not cells[y:y + ch, x:x + cw].any()]
            s = int(rng.choice(choices))
            cw, ch = COVERED[s]
            cells[y:y + ch, x:x + cw] = s << 1
            cells[y, x] |= 1
    gw, gh = (bw + 31) // 32, (bh + 31) // 32
    q = rng.integers(-40, 41, size=(3, gw * gh * 65536)).astype(np.int32)
    q[rng.random(q.shape) > density] = 0
    raw = rng.integers(1, 257, size=(bh, bw)).astype(np.int32)
    return cells, q, raw


def decode(tokens, cells, orders):
    """Inverse of the stream: per transform and channel, the nonzero count,
    then coefficients along the scan until that many nonzeros are seen."""
    bh, bw = cells.shape
    gw = (bw + 31) // 32
    vals = tokens["values"].tolist()
    out = {}
    pos = 0
    for g in range(len(tokens["group_offsets"]) - 1):
        gx0, gy0 = (g % gw) * 32, (g // gw) * 32
        assert pos == tokens["group_offsets"][g]
        for y in range(gy0, min(bh, gy0 + 32)):
            for x in range(gx0, min(bw, gx0 + 32)):
                if not cells[y, x] & 1:
                    continue
                st = int(cells[y, x] >> 1)
                for c in (1, 0, 2):
                    order = b.order_slice(orders, st, c)
                    cov = order.size // 64
                    nnz = vals[pos]
                    pos += 1
                    seen, scan = 0, cov
                    coef = np.zeros(order.size, np.int64)
                    while seen < nnz:
                        u = vals[pos]
                        pos += 1
                        v = (u >> 1) ^ -(u & 1)
                        coef[order[scan]] = v
                        seen += v != 0
                        scan += 1
                    out[(y, x, c)] = coef
    assert pos == len(vals)
    return out


@pytest.mark.parametrize("bw,bh,seed", [(5, 3, 0), (40, 36, 1), (33, 65, 2)])
def test_stream_decodes_to_the_coefficients(bw, bh, seed):
    rng = np.random.default_rng(seed)
    cells, q, raw = random_frame(rng, bw, bh)
    orders = b.natural_order_table()
    if seed == 2:  # a custom order for every family
        for st in COVERED:
            for c in range(3):
                sl = b.order_slice(orders, st, c)
                cov = sl.size // 64
                sl[cov:] = rng.permutation(sl[cov:])
    t = b.tokenize_frame(q, cells, raw, orders)
    dec = decode(t, cells, orders)
    gw = (bw + 31) // 32
    store = q.shape[1]
    used = {}
    for y in range(bh):
        for x in range(bw):
            if not cells[y, x] & 1:
                continue
            st = int(cells[y, x] >> 1)
            cw, ch = COVERED[st]
            n = cw * ch * 64
            g = (y // 32) * gw + x // 32
            off = g * 65536 + used.get(g, 0)
            used[g] = used.get(g, 0) + n
            lw, lh, W = max(cw, ch), min(cw, ch), max(cw, ch) * 8
            for c in range(3):
                want = q[c, off:off + n].astype(np.int64).copy()
                for yy in range(lh):
                    want[yy * W:yy * W + lw] = 0
                assert np.array_equal(dec[(y, x, c)], want), (y, x, c)
    assert store == q.shape[1]


def test_histogram_counts_every_token():
    rng = np.random.default_rng(5)
    cells, q, raw = random_frame(rng, 20, 12)
    t = b.tokenize_frame(q, cells, raw, histogram=True)
    h = t["histogram"]
    assert h.sum() == t["values"].size
    b._declare_tokenize()
    sym = np.array([b.oracle().oracle_token_symbol(int(v)) for v in t["values"]])
    expect = np.zeros_like(h)
    np.add.at(expect, (t["contexts"].astype(np.int64), sym), 1)
    assert np.array_equal(h, expect)
