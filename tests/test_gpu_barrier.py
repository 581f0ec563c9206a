"""The device-wide barrier probe: a cooperative launch whose workgroups meet at a bounded barrier ends
normally and reports positive, finite costs beside the graph-replayed launch spacing; requests the
probe cannot honour (more workgroups than fit on the device at once, an unknown fence mode) are
refused on the host before anything is launched."""
import math

import pytest

from rfnetwork_amd import _lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fences", [0, 1, 2])
def test_barrier_probe_runs(hip_ready, fences):
    print(f"[case] barrier probe fences={fences}", flush=True)
    b, l = _lib.grid_barrier_cost(256, fences, 200)
    assert math.isfinite(b) and math.isfinite(l) and b > 0 and l > 0


def test_barrier_probe_refusals(hip_ready):
    print("[case] barrier probe refusals", flush=True)
    with pytest.raises(ValueError):
        _lib.grid_barrier_cost(1 << 24, 1, 10)
    with pytest.raises(ValueError):
        _lib.grid_barrier_cost(256, 3, 10)
    with pytest.raises(ValueError):
        _lib.grid_barrier_cost(256, 1, 0)
