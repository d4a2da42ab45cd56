"""Host-side checks of the learned warm-start driver (lcdrsimulator_amd/warmstart.py):
the training distribution and the per-droplet parameters.  No GPU work."""
import math

import numpy as np
import pytest

from oracle import oracle as O


@pytest.fixture(scope="module")
def ws():
    import lcdrsimulator_amd as q
    q.load()
    from lcdrsimulator_amd import warmstart
    return warmstart


def test_droplet_params_box_radius_and_stable_dt(ws):
    p = ws.droplet_params(40.0, 300.0, 1e-4)
    assert p.dims == (88, 88, 88) and p.radius_index == 40.0
    assert (p.L1, p.L2, p.L3) == ws.ELASTIC and p.W == 1e-4 and p.boundary_custom == 1 and p.boundary_S == 0.5
    # dt is the stability window's dt_max, restated by the oracle's estimator
    ref = O.dt_estimate(1e-8, 1e-8, 1e-8, 0.1, p.a, p.b, p.c, p.T_star, 1, 0.0, *ws.ELASTIC, math.nan, 0.5,
                        300.0, 300.0)["dt_max"]
    assert p.dt == pytest.approx(ref, rel=1e-12)
    assert ws.droplet_params(88.0, 300.0, 1e-4, N=192).dims == (192, 192, 192)


def test_training_set_covers_the_stated_ranges(ws):
    t = ws.training_set(256)
    assert len(t) == 256
    R = np.array([p.radius_index for p, _, _ in t])
    T = np.array([p.T_high for p, _, _ in t])
    W = np.array([p.W for p, _, _ in t])
    modes = np.array([m for _, m, _ in t])
    assert R.min() >= 40 and R.max() <= 120 and np.all(R == np.round(R))
    assert all(p.Nx == int(2 * r + 8) for (p, _, _), r in zip(t, R))
    assert T.min() >= 295 and T.max() <= 305 and T.std() > 2.0
    assert W.min() >= 1e-5 and W.max() <= 1e-3
    lw = np.log10(W)
    assert abs(lw.mean() + 4.0) < 0.15  # log-uniform over two decades
    assert np.count_nonzero(modes == 0) == np.count_nonzero(modes == 1) == 128
    seeds = [s for _, _, s in t]
    assert len(set(seeds)) == 256


def test_training_set_matched_fixes_temperature_and_anchoring(ws):
    t = ws.training_set(16, T=300.0, W=1e-4)
    assert {p.T_high for p, _, _ in t} == {300.0} and {p.W for p, _, _ in t} == {1e-4}
    assert len({p.radius_index for p, _, _ in t}) > 4  # radii still drawn
    # same seed -> same draws as the unmatched set (radii agree)
    u = ws.training_set(16)
    assert [p.radius_index for p, _, _ in t] == [p.radius_index for p, _, _ in u]


def test_snapshot_pairs_cover_every_gap_apart_pair():
    from lcdrsimulator_amd import warmstart as ws
    for n, stride, gap in ((7, 500, 500), (13, 250, 500), (2, 250, 500), (5, 250, 50), (4, 100, 100), (3, 7, 5)):
        steps, inputs, pairs = ws.snapshot_pairs(n, stride, gap)
        assert steps == sorted(set(steps)) and len(inputs) == len(steps)
        want = [(k, k + gap) for k in range(0, n * stride, stride) if gap % stride or k + gap < n * stride]
        assert [(steps[a], steps[b]) for a, b in pairs] == want
        assert all(inputs[a] for a, _ in pairs)
    assert ws.snapshot_pairs(7, 500, 500)[2] == [(m, m + 1) for m in range(6)]
    assert ws.snapshot_pairs(2, 250, 500)[2] == []
    with pytest.raises(ValueError):
        ws.snapshot_pairs(7, 300, 500)
