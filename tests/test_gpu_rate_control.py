"""GPU: the target-size search with the device plain path as its encoder
(gjxl_amd.rate_control.search_plain_path).  The size of each encode is the
entropy of its tokens given their contexts (ANS is not built); the search is
the reference's widest-interval exploration of [0.01, 10], which tries the
lower and upper targets, then 5.005, 2.5075, 7.5025, ...

* A size reached at the first midpoint is met exactly on the third encode.
* A size between the tried targets (the one of distance 1.5) ends the search
  with the closest candidate tried, as the closest-absolute policy selects.
* Sizes fall as the target rises.
"""
import numpy as np
import pytest

import gjxl_amd
from gjxl_amd import rate_control as rc
from gjxl_amd import synthetic

pytestmark = pytest.mark.gpu


def _size_at(path, img, distance):
    path.reconfigure(512, 512, distance=distance)
    path.run(img)
    return rc.token_bytes(path.tokens())


def test_search_with_the_device_path(device):
    img = synthetic.generate(**synthetic.WORKLOADS["512x512"])
    path = gjxl_amd.NativePlainPath(device, 512, 512)
    midpoint = float(np.float32(np.float32(0.01) + np.float32(0.5) *
                                (np.float32(10.0) - np.float32(0.01))))
    goal = _size_at(path, img, midpoint)
    result, tried = rc.search_plain_path(path, img, goal, maximum_attempts=8)
    print(f"goal {goal} B at {midpoint}: tried {tried}", flush=True)
    assert result.target_size_met and result.attempts == 3
    assert result.selected_target == np.float32(midpoint)
    assert int(result.selected.encoded_bytes) == goal and result.selected.resident == 1

    goal = _size_at(path, img, 1.5)
    result, tried = rc.search_plain_path(path, img, goal, selection=rc.CLOSEST_ABSOLUTE,
                                         maximum_attempts=12)
    print(f"goal {goal} B at 1.5: tried {[(round(t, 4), s) for t, s in tried]}", flush=True)
    assert result.attempts == len(tried) == 12 and result.search_exhausted
    errors = [abs(s - goal) for _, s in tried]
    assert abs(int(result.selected.encoded_bytes) - goal) == min(errors)
    sizes = sorted(tried)
    assert all(a[1] > b[1] for a, b in zip(sizes, sizes[1:]))  # sizes fall with the target
    path.close()
