"""CPU: the sparse AC format oracle -- round trips, the read rule on a
hand-made layout, and layouts its validator must reject."""
import numpy as np

from oracle import sparse_ac as S


def test_round_trip_and_validity():
    rng = np.random.default_rng(0)
    for count in (1, 63, 64, 65, 200, 4096 + 7):
        for dtype in (np.int8, np.int16, np.int32):
            x = rng.integers(-3, 4, count).astype(dtype)
            x[rng.random(count) < 0.6] = 0
            m, o, v = S.pack(x)
            assert S.validate(m, o, v, count) is None
            assert np.array_equal(S.read(m, o, v, count, dtype), x)


def test_read_rule_on_a_hand_made_layout():
    # two words; payloads of word 1 first (ranges may come in any order)
    masks = np.array([0b1010, 1 << 63], np.uint64)
    offsets = np.array([1, 0], np.uint32)
    values = np.array([9, 5, 7], np.int32)
    x = S.read(masks, offsets, values, 128, np.int32)
    want = np.zeros(128, np.int32)
    want[1], want[3], want[127] = 5, 7, 9
    assert np.array_equal(x, want) and S.validate(masks, offsets, values, 128) is None


def test_validator_rejects_broken_layouts():
    x = np.zeros(130, np.int32)
    x[[0, 5, 64, 129]] = [1, -2, 3, 4]
    m, o, v = S.pack(x)
    assert S.validate(m, o, v, 130) is None
    o2 = o.copy()
    o2[1] = 0
    assert S.validate(m, o2, v, 130) == "overlap"
    m2 = m.copy()
    m2[-1] |= np.uint64(1) << np.uint64(5)
    assert S.validate(m2, o, v, 130) == "tail mask"
    assert S.validate(m, o, v[:-1], 130) == "payload range"
    v2 = v.copy()
    v2[0] = 0
    assert S.validate(m, o, v2, 130) == "zero payload"
