"""GPU: the batch driver (gjxl_amd_batch_*).

Images of uneven extents run on three workers at once; each worker creates
its plain path at the bounding extent of the call and re-targets it to every
image it takes.  Every result equals a lone run of the
C++ plain path on the same image (tokens, group offsets, scores, orders and
block-context map), the 512x512 result equals the oracle's digests, request
order is kept, a bad request fails alone, and a shut-down driver refuses
further calls (VarDctBatchEncoder's contract, src/codestream/
batch_workflow.h:47-109).
"""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

import gjxl_amd
from gjxl_amd import synthetic

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
from make_plain_path_hashes import digest  # noqa: E402

pytestmark = pytest.mark.gpu

GOLDEN = json.loads(
    (Path(__file__).resolve().parent / "golden" / "plain_path_hashes.json").read_text())


def _images():
    out = [np.ascontiguousarray(synthetic.generate(256, 200, seed=3)),
           synthetic.generate(**synthetic.WORKLOADS["512x512"]),
           np.ascontiguousarray(synthetic.generate(256, 200, seed=4)),
           np.ascontiguousarray(synthetic.generate(136, 72, seed=5)[:66, :130]),
           np.ascontiguousarray(synthetic.generate(80, 48, seed=7)[:45, :77])]
    return out


def test_batch_matches_lone_runs(device):
    images = _images()
    print(f"batch of {len(images)} images on 3 workers", flush=True)
    batch = gjxl_amd.Batch(0, 3)
    assert batch.max_in_flight == 3
    results = batch.encode(images)
    for i, img in enumerate(images):
        r = results[i]
        assert r.status == 0, r.message
        h, w = img.shape[:2]
        assert (r.width, r.height) == (w, h)
        assert r.ready_nanoseconds == r.queue_nanoseconds + r.service_nanoseconds
        lone = gjxl_amd.NativePlainPath(device, w, h)
        lone.run(img)
        want = lone.tokens()
        got = r.tokens()
        for k in ("values", "contexts", "group_offsets"):
            assert np.array_equal(got[k], want[k]), (i, k)
        assert r.scores_list() == lone.scores()
        a, b = r.entropy(), lone.entropy_inputs()
        assert a.used_order_mask == b.used_order_mask
        assert np.array_equal(a.order_table(), b.order_table())
        assert all(np.array_equal(x, y) for x, y in zip(a.context_model(), b.context_model()))
        lone.close()
    g = GOLDEN["512x512"]
    t = results[1].tokens()
    assert digest(t["values"]) == g["token_values"]
    assert digest(t["contexts"]) == g["token_contexts"]
    batch.free_results(results, len(images))
    assert results[0].token_count == 0  # freed and zeroed
    batch.close()


def test_batch_failure_is_per_image_and_shutdown_refuses():
    images = _images()[3:]
    batch = gjxl_amd.Batch(0, 2)
    reqs_ok = batch.encode(images)
    assert all(reqs_ok[i].status == 0 for i in range(len(images)))
    batch.free_results(reqs_ok, len(images))
    # a 2-byte pixel layout is rejected by the path; the other image runs
    bad = np.zeros((8, 8, 2), np.uint8)
    results = batch.encode([bad, images[0]])
    assert results[0].status == 1 and b"RGB8" in results[0].message
    assert results[1].status == 0 and results[1].token_count > 0
    batch.free_results(results, 2)
    batch.shutdown()
    sentinel = (gjxl_amd.BatchResult * 1)()
    sentinel[0].width = 12345
    with pytest.raises(gjxl_amd.GjxlError, match="UNAVAILABLE"):
        batch.encode(images[:1], results=sentinel)
    assert sentinel[0].width == 12345
    batch.close()


def test_tile_queue_groups_match_default_mode(device):
    """Group mode (gjxl_amd_batch_set_tile_queue): images in groups of three,
    largest first, their candidate costs through one tile-queue graph per
    group; every result equals the default mode's, in request order, on a
    tuned call and on the call after it, and a bad request still fails
    alone."""
    images = _images() + [np.ascontiguousarray(synthetic.generate(520, 344, seed=8))]
    batch = gjxl_amd.Batch(0, 3)
    plain = batch.encode(images)
    want = [(plain[i].tokens(), plain[i].scores_list()) for i in range(len(images))]
    batch.free_results(plain, len(images))
    batch.set_tile_queue(True, tune_repetitions=1)
    for call in range(2):
        print(f"tile-queue call {call}", flush=True)
        got = batch.encode(images)
        for i in range(len(images)):
            r = got[i]
            assert r.status == 0, r.message
            assert (r.width, r.height) == images[i].shape[1::-1]
            t = r.tokens()
            for k in ("values", "contexts", "group_offsets"):
                assert np.array_equal(t[k], want[i][0][k]), (call, i, k)
            assert r.scores_list() == want[i][1]
        batch.free_results(got, len(images))
    bad = np.zeros((8, 8, 2), np.uint8)
    got = batch.encode([images[3], bad, images[4]])
    assert got[0].status == 0 and got[2].status == 0
    assert got[1].status == 1 and b"RGB8" in got[1].message
    batch.free_results(got, 3)
    batch.set_tile_queue(False)
    batch.close()
