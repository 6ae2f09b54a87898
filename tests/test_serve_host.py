"""Host logic of the serving path (no GPU): the admission plan of
fcs_serve_plan and the synthetic viewer traffic."""
import math

import numpy as np
import pytest

import fcs
from fcs import serve


def test_plan_admits_whole_volumes_by_demand_density():
    # densities 0.01, 0.015, 0.02: v2 then v1 fit whole, v0 gets nothing left
    p, pool = serve.plan([100, 200, 50], [1.0, 3.0, 1.0], 250)
    assert list(p) == [0, 2, 2] and list(pool) == [0, 200, 50]


def test_plan_splits_remainder_by_demand():
    # v0 whole (100), 150 left split 1:2 between v1 and v2
    p, pool = serve.plan([100, 1000, 1000], [10.0, 1.0, 2.0], 250)
    assert list(p) == [2, 1, 1]
    assert list(pool) == [100, 50, 100]


def test_plan_zero_demand_and_exact_fit():
    p, pool = serve.plan([100, 150, 7], [1.0, 1.0, 0.0], 250)
    assert list(p) == [2, 2, 0] and list(pool) == [100, 150, 0]


def test_plan_ties_keep_index_order():
    p, pool = serve.plan([100, 100, 100], [1.0, 1.0, 1.0], 250)
    assert list(p) == [2, 2, 1] and list(pool) == [100, 100, 50]


def test_plan_everything_fits_in_200_gb():
    # the four serving volumes at 32^3 bricks, u32: 36.6 GiB of finest labels,
    # 41.8 GiB with every coarser LOD of the pyramid (37,449 labels per brick)
    full = [(s // 32) ** 3 * 37449 * 4 for s in (256, 512, 1024, 2048)]
    assert abs(sum(full) / 2**30 - 41.79) < 0.01
    p, pool = serve.plan(full, [1, 1, 1, 1], 200 * 10**9)
    assert list(p) == [2, 2, 2, 2] and list(pool) == full


def test_plan_rejects_negative_demand():
    with pytest.raises(fcs.FcsError):
        serve.plan([1, 2], [1.0, -1.0], 10)


def test_poisson_arrivals_rate():
    rng = np.random.default_rng(7)
    t = serve.poisson_arrivals(2000.0, 5.0, rng)
    assert np.all(np.diff(t) > 0) and t[0] >= 0 and t[-1] < 5.0
    # count ~ Poisson(10000): within 5 sigma
    assert abs(len(t) - 10000) < 5 * math.sqrt(10000)
    gaps = np.diff(t)
    assert abs(gaps.mean() - 1 / 2000) < 0.05 / 2000
    # exponential gaps: coefficient of variation 1
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.05


class _StubVolume:
    """Volume stand-in with the two members make_trace reads."""

    def __init__(self, grid):
        self.info = {"grid_size": grid}
        self.calls = []

    def orbit_camera(self, yaw, pitch, radius, fov):
        self.calls.append((yaw, pitch, radius, fov))
        return fcs.Camera((radius, 0, 0), (0, 0, 1), (1, 0, 0), (0, 1, 0))


def test_trace_sessions_resolutions_and_walk_bounds():
    vols = [_StubVolume(256), _StubVolume(512)]
    reqs, meta = serve.make_trace(vols, n_sessions=8, rate=500.0, duration=2.0, seed=3)
    n = len(meta["arrival_s"])
    assert n > 800
    assert np.all(np.diff(meta["arrival_s"]) >= 0)
    assert np.array_equal(meta["volume"], meta["session"] % 2)
    for i in (0, n // 2, n - 1):
        assert reqs[i].volume == meta["volume"][i] and reqs[i].session == meta["session"][i]
        assert (reqs[i].w, reqs[i].h) in serve.RESOLUTIONS
    counts = np.bincount([serve.RESOLUTIONS.index((int(w), int(h))) for w, h in zip(meta["w"], meta["h"])])
    assert counts.min() > n / 3 - 4 * math.sqrt(n)
    for v in vols:
        diag = math.sqrt(3) * v.info["grid_size"]
        for yaw, pitch, r, fov in v.calls:
            assert -0.45 * math.pi <= pitch <= 0.45 * math.pi
            assert 0.8 * diag - 1e-6 <= r <= 1.5 * diag + 1e-6
    # the same seed gives the same trace
    reqs2, meta2 = serve.make_trace([_StubVolume(256), _StubVolume(512)], n_sessions=8, rate=500.0, duration=2.0,
                                    seed=3)
    assert np.array_equal(meta["arrival_s"], meta2["arrival_s"])
    assert reqs2[n - 1].cam.position[0] == reqs[n - 1].cam.position[0]
