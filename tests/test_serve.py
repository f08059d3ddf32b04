"""Served stream, host side: the request stream's statistics, job sizes and the device
footprint estimate (no GPU)."""
import math

import numpy as np

from pannde_amd import serve


def test_stream_is_seeded_and_follows_its_distributions():
    a = serve.request_stream(seed=3, n=4000, rate=5.0)
    b = serve.request_stream(seed=3, n=4000, rate=5.0)
    assert a == b
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 0.2) < 0.02  # Poisson at 5 per second
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.08  # exponential gaps
    cells = np.array([r.cells for r in a])
    assert cells.min() >= 0.5e6 and cells.max() <= 8e6
    # log-uniform: the log of the size is uniform over [log 0.5M, log 8M]
    u = (np.log(cells) - math.log(0.5e6)) / math.log(16)
    assert abs(u.mean() - 0.5) < 0.02
    steps = np.array([r.nsteps for r in a])
    assert steps.min() >= 1000 and steps.max() <= 4000
    assert all(0.5e-3 <= r.crack_length <= 5e-3 and 0.2e-3 <= r.crack_depth <= 2e-3 for r in a)
    assert all(0 <= r.crack_angle < 180 for r in a)
    frac = np.mean([r.spot_check for r in a])
    assert 0.04 < frac < 0.06
    assert {r.kind for r in a} == {"bracket", "plate"}


def test_job_sizes_track_the_request():
    for kind in ("bracket", "plate"):
        for target in (0.5e6, 2e6, 8e6):
            r = serve.Request(0, 0.0, kind, int(target), 1000, 1e-3, 5e-4, 30.0, (0.5, 0.5), 1, False)
            d = serve.mesh_dims(r)
            n = 2 * d["leg"] * d["width"] * d["height"] if kind == "bracket" else d["nx"] * d["ny"] * 64
            assert 0.8 * target < n < 1.2 * target, (kind, target, n)


def test_small_jobs_build_with_their_cracks():
    for kind in ("bracket", "plate"):
        r = serve.Request(0, 0.0, kind, 40_000, 120, 1.0e-3, 4e-4, 75.0, (0.3, 0.6), 11, False)
        c = serve.build_case(r)
        assert c.nsteps == 120 and c.meta["n_dup"] > 0 and c.meta["crack"].angle_deg == 75.0
        assert c.sensors.shape[0] > 0 and len(c.force) == 1


def test_device_footprint_estimate():
    from pannde_amd import rsg
    e32 = serve.estimate_device_bytes(1_000_000, 1_050_000, rsg.RSG_FP32)
    e64 = serve.estimate_device_bytes(1_000_000, 1_050_000, rsg.RSG_FP64)
    assert 250e6 < e32 < e64 < 500e6
