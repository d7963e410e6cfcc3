"""GPU: the sparse AC read-back (gjxl_amd_sparse_ac) against the format's
oracle (oracle/sparse_ac.py): every layout is valid (ranges inside the
payload, no overlap, exact coverage, no zero payload, no bits past the
count), reads back to the source exactly, and keeps each work group's words
contiguous -- for 1-, 2- and 4-byte values, ragged counts and the plain
path's quantized store."""
import numpy as np
import pytest

import gjxl_amd
from gjxl_amd import synthetic
from oracle import sparse_ac as S

pytestmark = pytest.mark.gpu


def _check(device, x):
    src = device.array(x)
    masks, offsets, values, total = gjxl_amd.sparse_ac(device, src)
    count = x.size
    assert total == np.count_nonzero(x)
    assert S.validate(masks, offsets, values, count) is None
    assert np.array_equal(S.read(masks, offsets, values, count, x.dtype), x)
    # the four words of one work group take one contiguous range
    counts = S.popcount64(masks)
    for g in range(0, masks.size, 4):
        w = np.arange(g, min(g + 4, masks.size))
        assert np.array_equal(offsets[w[1:]], offsets[w[:-1]] + counts[w[:-1]])
    src.free()
    return total


@pytest.mark.parametrize("dtype", [np.int8, np.int16, np.int32])
def test_sparse_ac_random(device, dtype):
    rng = np.random.default_rng(np.dtype(dtype).itemsize)
    for count in (1, 63, 64, 65, 1000, 256 * 5 + 3, 1 << 20):
        for density in (0.0, 0.05, 0.5, 1.0):
            x = rng.integers(-100, 100, count).astype(dtype)
            x[x == 0] = 1
            x[rng.random(count) >= density] = 0
            _check(device, x)


def test_sparse_ac_of_the_plain_path_store(device):
    for spec in ("512x512", "4032x3024"):
        w, h = synthetic.WORKLOADS[spec]["width"], synthetic.WORKLOADS[spec]["height"]
        path = gjxl_amd.NativePlainPath(device, w, h)
        path.run(synthetic.generate(**synthetic.WORKLOADS[spec]), read_tokens=False)
        store = path.view("quantized")
        x = store.numpy().reshape(-1)
        total = _check(device, x)
        dense = x.nbytes
        sparse = ((x.size + 63) // 64) * 12 + 4 * total
        print(f"{spec}: dense {dense / 1e6:.1f} MB, sparse {sparse / 1e6:.1f} MB", flush=True)
        assert sparse < dense
        path.close()


def test_sparse_ac_refuses_bad_widths(device):
    src = device.empty((64,), np.int32)
    with pytest.raises(gjxl_amd.GjxlError, match="INVALID_ARGUMENT"):
        gjxl_amd._check(gjxl_amd.library().gjxl_amd_sparse_ac(
            device.handle, src.ptr, 64, 3, src.ptr, src.ptr, src.ptr, src.ptr))
    masks, offsets, values, total = gjxl_amd.sparse_ac(device, src, count=0)
    assert total == 0 and masks.size == 0
