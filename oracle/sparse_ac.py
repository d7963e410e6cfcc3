"""The sparse AC read-back format.  TEST INFRASTRUCTURE ONLY: tests/ check
gjxl_amd_sparse_ac with it; the product never imports it.

* `read` restates SparseCoefficientSpan::Read (src/codec/
  sparse_coefficients.h:96-102): value i is values[offsets[i / 64] +
  popcount(masks[i / 64] & ((1 << i % 64) - 1))] when its mask bit is set,
  else zero.
* `validate` restates SparseAcStorage::Validate (:120-167): shape, no mask
  bits past the count, payload ranges inside the payload, not overlapping,
  covering it exactly, and no zero payload.
* `pack` is one valid layout (word ranges in word order), the reference's
  with a single work group.

Pinned by round trips and by layouts the validator must reject
(tests/test_oracle_sparse_ac.py).
"""
from __future__ import annotations

import numpy as np


def pack(source: np.ndarray):
    count = source.size
    words = (count + 63) // 64
    padded = np.zeros(words * 64, source.dtype)
    padded[:count] = source
    nz = (padded != 0).reshape(words, 64)
    weights = (np.uint64(1) << np.arange(64, dtype=np.uint64))
    masks = (nz.astype(np.uint64) * weights).sum(axis=1, dtype=np.uint64)
    counts = nz.sum(axis=1)
    offsets = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint32)
    return masks, offsets, padded[padded != 0].copy()


def popcount64(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.uint64)
    out = np.zeros(x.shape, np.int64)
    for shift in range(0, 64, 8):
        out += np.array([bin(b).count("1") for b in range(256)])[(x >> np.uint64(shift)) & np.uint64(255)]
    return out


def read(masks, offsets, values, count: int, dtype) -> np.ndarray:
    i = np.arange(count, dtype=np.uint64)
    word = (i // np.uint64(64)).astype(np.int64)
    flag = np.uint64(1) << (i % np.uint64(64))
    m = masks[word]
    present = (m & flag) != 0
    below = popcount64(m & (flag - np.uint64(1)))
    out = np.zeros(count, dtype)
    at = offsets[word].astype(np.int64) + below
    out[present] = values[at[present]]
    return out


def validate(masks, offsets, values, count: int) -> str | None:
    """None when the layout is valid, else what is wrong."""
    words = (count + 63) // 64
    if count == 0 or masks.size != words or offsets.size != words or values.size > count:
        return "shape"
    if count % 64 and int(masks[-1]) >> (count % 64):
        return "tail mask"
    covered = np.zeros(values.size, bool)
    counts = popcount64(masks)
    for w in range(words):
        at, n = int(offsets[w]), int(counts[w])
        if at > values.size or n > values.size - at:
            return "payload range"
        if covered[at:at + n].any():
            return "overlap"
        covered[at:at + n] = True
    if int(counts.sum()) != values.size or not covered.all():
        return "coverage"
    if (values == 0).any():
        return "zero payload"
    return None
