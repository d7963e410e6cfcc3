"""CPU: the effort-7 tokenization inputs -- sample decisions, frame
statistics, coefficient orders, block-context map, order tokens.

The oracle (oracle/entropy_inputs.c) is checked against:
* the scalar restatement the reference's own test carries for the orders
  (tests/coefficient_order_test.cpp:32-118), rewritten here in numpy
  (coefficient-major zero counts, keys uint64(float32(zeros) * 1/sqrt(n) +
  0.1f), a stable sort of the natural order after the LLF prefix, the
  xorshift128+ sample with its seeds), plus that test's cases: the
  small-frame cutoff and constant populations keeping the natural order;
* the reference's default block-context map (tests/block_context_map_test.
  cpp:15-39: four contexts, the per-strategy lookup);
* for the adaptive map, an independent restatement of BuildAdaptiveMap
  (src/codestream/block_context_map.cpp:126-200) and the invariants of
  ValidateSimpleBlockContextMap (:281-320);
* for the order tokens, decoding the Lehmer code back to the order.

The product's host code (gjxl_amd_entropy_inputs_from_statistics,
gjxl_amd_coefficient_order_tokens: no device calls) is compared with the
oracle on the same statistics.
"""
import ctypes

import numpy as np
import pytest

import gjxl_amd
from oracle import binding
from token_frames import frame

FAMILIES = gjxl_amd.FAMILIES  # (family, strategy, offset, size, llf)


def xorshift_decisions(bw, bh):
    """tests/coefficient_order_test.cpp:47-77, per block in group order."""
    mask = (1 << 64) - 1
    a, b = 0x94D049BB133111EB, 0xBF58476D1CE4E5B9
    out = np.zeros((bh, bw), np.uint8)
    for gy in range(0, bh, 32):
        for gx in range(0, bw, 32):
            for y in range(gy, min(gy + 32, bh)):
                for x in range(gx, min(gx + 32, bw)):
                    bits = (a + b) & mask
                    old_b = b
                    a = (a ^ (a << 23)) & mask
                    b = a ^ old_b ^ (a >> 18) ^ (old_b >> 5)
                    a = old_b
                    out[y, x] = (bits >> 32) <= ((1 << 32) - 1) // 2
    return out


def brute_statistics(quantized, cells, raw, decisions):
    """Coefficient-major recount over the anchors in group order."""
    bh, bw = cells.shape
    gw = (bw + 31) // 32
    zeros = np.zeros((3, 1984), np.int64)
    sampled = np.zeros((3, 64), np.int64)
    present = 0
    offsets = {f: off for f, _, off, _, _ in FAMILIES}
    family_of = {0: 0, 6: 4, 7: 4, 4: 2, 10: 6, 11: 6, 5: 3}
    for gy in range((bh + 31) // 32):
        for gx in range(gw):
            base = (gy * gw + gx) * 65536
            at = 0
            for y in range(gy * 32, min(gy * 32 + 32, bh)):
                for x in range(gx * 32, min(gx * 32 + 32, bw)):
                    c = cells[y, x]
                    if not c & 1:
                        continue
                    st = int(c >> 1)
                    f = family_of[st]
                    n = {0: 64, 4: 256, 5: 1024}.get(st, 128 if st in (6, 7) else 512)
                    present |= 1 << f
                    for ch in range(3):
                        seg = quantized[ch, base + at:base + at + n]
                        zeros[ch, offsets[f]:offsets[f] + n] += seg == 0
                        if st == 0 and decisions[y, x]:
                            sampled[ch] += seg == 0
                    at += n
    raw_counts = np.zeros((13, 256), np.int64)
    for y in range(bh):
        for x in range(bw):
            raw_counts[family_of[int(cells[y, x] >> 1)], raw[y, x] - 1] += 1
    return zeros, sampled, present, raw_counts


def numpy_orders(stats, sampled_behaviour=True):
    """tests/coefficient_order_test.cpp:79-118 on population counts."""
    bh_ok = not (stats["bw"] < 5 and stats["bh"] < 5)
    used, table = 0, binding.natural_order_table()
    if not bh_ok:
        return used, table
    sample = sampled_behaviour and stats["present_mask"] == 1
    for f, st, off, n, llf in FAMILIES:
        if not stats["present_mask"] >> f & 1:
            continue
        natural = binding.natural_order(st)
        scale = np.float32(1.0) / np.sqrt(np.float32(n))
        custom = []
        for c in range(3):
            counts = stats["sampled"][c] if sample else stats["zero_counts"][c, off:off + n]
            keys = (counts.astype(np.float32) * scale + np.float32(0.1)).astype(np.uint64)
            tail = natural[llf:]
            order = np.concatenate([natural[:llf], tail[np.argsort(keys[tail], kind="stable")]])
            custom.append(order)
        if any(not np.array_equal(o, natural) for o in custom):
            used |= 1 << f
            for stage, s in enumerate(binding.STAGE_STRATEGIES):
                if {0: 0, 6: 4, 7: 4, 4: 2, 10: 6, 11: 6, 5: 3}[s] == f:
                    for c in range(3):
                        binding.order_slice(table, s, c)[:] = custom[c]
    return used, table


def adaptive_map(raw_counts, blocks):
    """BuildAdaptiveMap restated (block_context_map.cpp:105-200)."""
    if blocks < 1024:
        return [], binding.DEFAULT_CONTEXT_MAP.copy(), 4
    thresholds = []
    if blocks >= 8192:
        totals = raw_counts.sum(axis=0)
        q = int(np.argmax(np.cumsum(totals) > blocks // 2))
        if q != 0:
            thresholds = [q]
    segs = len(thresholds) + 1
    cells = 13 * segs
    count = [0] * cells
    for f in range(13):
        for q in range(256):
            seg = sum(1 for t in thresholds if q + 1 > t)
            count[f * segs + seg] += int(raw_counts[f, q])
    remap = list(range(cells))
    clusters = list(range(cells))
    desired = min(max(blocks // 2048, 2), 9)
    while len(clusters) > desired:
        clusters = sorted(clusters, key=lambda c: -count[c])  # stable
        dst, src = clusters[-2], clusters[-1]
        count[dst] += count[src]
        count[src] = 0
        remap[src] = dst
        clusters.pop()
    for i in range(cells):
        while remap[remap[i]] != remap[i]:
            remap[i] = remap[remap[i]]
    labels, nxt = {}, 0
    for i in range(cells):
        if remap[i] not in labels:
            labels[remap[i]] = nxt
            nxt += 1
        remap[i] = labels[remap[i]]
    chroma = min(max(blocks // 3072, 1), 5)
    cmap = remap + [nxt + min(remap[i % cells], chroma - 1) for i in range(cells, 3 * cells)]
    return thresholds, np.array(cmap, np.uint8), max(cmap) + 1


def test_sample_decisions_match_the_reference_test_generator():
    for bw, bh in [(40, 40), (70, 33), (5, 90)]:
        assert np.array_equal(binding.order_sample_decisions(bw, bh), xorshift_decisions(bw, bh))


@pytest.mark.parametrize("bw,bh,seed,weights", [
    (40, 40, 1, None), (70, 33, 2, None), (64, 64, 3, [1, 0, 0, 0, 0, 0, 0, 0]),
    (36, 20, 4, [0, 1, 1, 1, 1, 1, 1, 1])])
def test_frame_statistics_recount(bw, bh, seed, weights):
    rng = np.random.default_rng(seed)
    cells, q, raw = frame(rng, bw, bh, density=0.3, weights=weights)
    stats = binding.frame_statistics(q, cells, raw)
    zeros, sampled, present, raw_counts = brute_statistics(
        q, cells, raw, xorshift_decisions(bw, bh))
    assert np.array_equal(stats["zero_counts"], zeros)
    assert np.array_equal(stats["sampled"], sampled)
    assert stats["present_mask"] == present
    assert np.array_equal(stats["raw_counts"], raw_counts)
    assert stats["raw_counts"].sum() == bw * bh


@pytest.mark.parametrize("bw,bh,seed,weights", [
    (40, 40, 1, None), (70, 33, 2, None), (64, 64, 3, [1, 0, 0, 0, 0, 0, 0, 0]),
    (36, 20, 4, [0, 1, 1, 1, 1, 1, 1, 1]), (8, 4, 5, None)])
def test_orders_match_the_reference_test_restatement(bw, bh, seed, weights):
    rng = np.random.default_rng(seed)
    cells, q, raw = frame(rng, bw, bh, density=0.3, weights=weights)
    stats = binding.frame_statistics(q, cells, raw)
    for sampled_behaviour in (True, False):
        used, table = binding.coefficient_orders(stats, sampled_behaviour)
        want_used, want_table = numpy_orders(stats, sampled_behaviour)
        assert used == want_used
        assert np.array_equal(table, want_table)
    assert used != 0  # random populations change the order


def test_small_frames_and_constant_populations_keep_natural_orders():
    rng = np.random.default_rng(9)
    cells, q, raw = frame(rng, 4, 4, density=0.5)
    used, table = binding.coefficient_orders(binding.frame_statistics(q, cells, raw))
    assert used == 0 and np.array_equal(table, binding.natural_order_table())
    # all-zero and all-nonzero coefficients: equal keys, natural order
    for fill in (0, 3):
        cells, q, raw = frame(rng, 40, 40, density=0.0)
        q[:] = fill
        used, table = binding.coefficient_orders(binding.frame_statistics(q, cells, raw))
        assert used == 0 and np.array_equal(table, binding.natural_order_table())


def test_default_block_context_map_below_1024_blocks():
    # tests/block_context_map_test.cpp:15-39: four contexts; DCT8 family row
    # 0 for Y, the strategy lookups of kSimpleBlockContextMap
    model = binding.block_context_map(np.ones((13, 256), np.uint32), 1023)
    assert model["num_contexts"] == 4 and model["thresholds"].size == 0
    assert np.array_equal(model["context_map"], binding.DEFAULT_CONTEXT_MAP)


@pytest.mark.parametrize("blocks,seed", [(1024, 1), (4096, 2), (8191, 3), (8192, 4),
                                          (190512, 5), (30000, 6)])
def test_adaptive_block_context_map(blocks, seed):
    rng = np.random.default_rng(seed)
    raw_counts = np.zeros((13, 256), np.uint32)
    fam = rng.choice([0, 2, 3, 4, 6], size=blocks, p=[0.6, 0.1, 0.05, 0.15, 0.1])
    q = np.clip(rng.normal(40, 15, size=blocks), 1, 256).astype(int)
    np.add.at(raw_counts, (fam, q - 1), 1)
    model = binding.block_context_map(raw_counts, blocks)
    thr, cmap, nctx = adaptive_map(raw_counts, blocks)
    assert list(model["thresholds"]) == list(thr)
    assert np.array_equal(model["context_map"], cmap)
    assert model["num_contexts"] == nctx
    # ValidateSimpleBlockContextMap: every label used, canonical, <= 16
    lib = binding.oracle()
    t = np.ascontiguousarray(model["thresholds"], np.uint32)
    assert lib.oracle_validate_context_map(
        binding.ptr(t) if t.size else None, t.size, binding.ptr(model["context_map"]),
        nctx) == 0
    assert nctx <= 16
    cells = 13 * (len(thr) + 1)
    luma = model["context_map"][:cells]
    assert len(set(luma.tolist())) == min(max(blocks // 2048, 2), 9)
    assert (blocks >= 8192) == (len(thr) == 1)


def decode_order_tokens(used, ctx, val):
    """Lehmer decode of TokenizeSimpleCoefficientOrders' stream."""
    table = binding.natural_order_table()
    i = 0
    for f, st, off, n, llf in sorted(FAMILIES):
        if not used >> f & 1:
            continue
        natural = binding.natural_order(st)
        for c in range(3):
            count = int(val[i])
            assert ctx[i] == min(1 + int(n).bit_length() - 1, 7)
            i += 1
            lehmer = [0] * n
            last = 0
            for k in range(llf, llf + count):
                assert ctx[i] == (0 if last == 0 else min(last.bit_length(), 7))
                lehmer[k] = int(val[i])
                last = lehmer[k]
                i += 1
            remaining = list(range(n))
            perm = [remaining.pop(code) for code in lehmer]
            order = natural[perm]
            for s in binding.STAGE_STRATEGIES:
                if {0: 0, 6: 4, 7: 4, 4: 2, 10: 6, 11: 6, 5: 3}[s] == f:
                    binding.order_slice(table, s, c)[:] = order
    assert i == len(val)
    return table


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_order_tokens_decode_to_the_orders(seed):
    rng = np.random.default_rng(seed)
    cells, q, raw = frame(rng, 48, 40, density=0.25)
    used, table = binding.coefficient_orders(binding.frame_statistics(q, cells, raw))
    ctx, val = binding.order_tokens(used, table)
    assert np.array_equal(decode_order_tokens(used, ctx, val), table)


def _product_statistics(stats):
    s = gjxl_amd.FrameStatistics()
    s.block_width, s.block_height = stats["bw"], stats["bh"]
    s.present_mask = stats["present_mask"]
    np.ctypeslib.as_array(s.zero_counts)[...] = stats["zero_counts"]
    np.ctypeslib.as_array(s.sampled_dct8_zero_counts)[...] = stats["sampled"]
    np.ctypeslib.as_array(s.raw_quant_counts)[...] = stats["raw_counts"]
    return s


@pytest.mark.parametrize("bw,bh,seed,weights", [
    (40, 40, 1, None), (70, 33, 2, None), (64, 64, 3, [1, 0, 0, 0, 0, 0, 0, 0]),
    (128, 72, 4, None), (8, 4, 5, None)])
def test_product_host_code_matches_the_oracle(bw, bh, seed, weights):
    """Host-only entry points of the library (no device is touched)."""
    rng = np.random.default_rng(seed)
    cells, q, raw = frame(rng, bw, bh, density=0.3, weights=weights)
    stats = binding.frame_statistics(q, cells, raw)
    inputs = gjxl_amd.entropy_inputs(_product_statistics(stats), 7)
    used, table = binding.coefficient_orders(stats, True)
    assert inputs.used_order_mask == used
    assert np.array_equal(inputs.order_table(), table)
    model = binding.block_context_map(stats["raw_counts"], bw * bh)
    thr, cmap, nctx = inputs.context_model()
    assert list(thr) == list(model["thresholds"])
    assert np.array_equal(cmap, model["context_map"]) and nctx == model["num_contexts"]
    ctx, val = gjxl_amd.order_tokens(inputs)
    want_ctx, want_val = binding.order_tokens(used, table)
    assert np.array_equal(ctx, want_ctx) and np.array_equal(val, want_val)


def test_product_rejects_inconsistent_statistics():
    rng = np.random.default_rng(3)
    cells, q, raw = frame(rng, 40, 40)
    s = _product_statistics(binding.frame_statistics(q, cells, raw))
    out = gjxl_amd.EntropyInputs()
    sentinel = bytes(out)
    s.raw_quant_counts[0][0] += 1  # counts no longer cover the frame
    rc = gjxl_amd.library().gjxl_amd_entropy_inputs_from_statistics(
        ctypes.byref(s), 7, ctypes.byref(out))
    assert rc == 1 and bytes(out) == sentinel
