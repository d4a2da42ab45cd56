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


# ---------------------------------------------------------------------------
# The warm-start drivers' host logic against stand-ins for the solver and the
# network (no GPU): schedule, retirement, back-off, step accounting.
# ---------------------------------------------------------------------------
class _FakeSolver:
    """F_b(t) = -1 - a_b exp(-t / tau_b) decays with the plain steps a box takes while active; a
    kept learned step multiplies a_b by `jump`.  Every call is logged."""

    def __init__(self, taus, a0=1e-2):
        self.t = [0] * len(taus)
        self.a = [a0] * len(taus)
        self.tau = list(taus)
        self.active = [True] * len(taus)
        self.log = []

    def F(self, b):
        import math as m
        return -1.0 - self.a[b] * m.exp(-self.t[b] / self.tau[b])

    def step(self, n):
        self.log.append(("step", n))
        for b, on in enumerate(self.active):
            if on:
                self.t[b] += n

    def observables(self, box=0, **kw):
        self.log.append(("energy", box))
        return {"total": self.F(box), "defects": 0, "plaquettes": 0}

    def energies(self, boxes):
        self.log.append(("energies", tuple(boxes)))
        return np.array([self.F(b) for b in boxes])

    def set_active(self, b, on):
        self.active[b] = bool(on)

    def set_active_many(self, boxes, on):
        self.log.append(("retire" if not on else "resume", tuple(boxes)))
        for b in boxes:
            self.active[b] = bool(on)


class _FakeNet:
    """Keeps a box's proposal when decide(box, call) says so, shrinking its amplitude."""

    def __init__(self, decide, jump=0.25):
        self.decide, self.jump, self.calls = decide, jump, []

    def _one(self, s, b):
        ok = self.decide(b, len([c for c in self.calls if c[0] == b]))
        self.calls.append((b, ok))
        F_before = s.F(b)
        if ok:
            s.a[b] *= self.jump
        return {"accepted": int(ok), "scale": 1.0, "ms_network": 1.0, "F_before": F_before,
                "F_after": s.F(b) if ok else F_before * (1.0 - 1e-3)}

    def apply(self, s, mode, scales=None):
        s.log.append(("apply", 0))
        return self._one(s, 0)

    def apply_boxes(self, s, boxes, mode, scales):
        s.log.append(("apply_boxes", tuple(boxes)))
        return [self._one(s, b) for b in boxes]


def test_relax_warm_cycles_never_test_across_a_jump(ws):
    s = _FakeSolver([300.0])
    r = ws.relax_warm(s, _FakeNet(lambda b, k: True), tol=1e-6, every=100, max_steps=100000)
    assert r["converged"] and r["steps"] % 200 == 0
    kinds = [e[0] for e in s.log]
    cycles = [kinds[i:i + 5] for i in range(0, len(kinds), 5)]
    assert all(c == ["apply", "step", "energy", "step", "energy"] for c in cycles)
    assert len(cycles) == r["jumps_accepted"] + r["jumps_rejected"] == r["steps"] // 200
    plain = ws.relax_plain(_FakeSolver([300.0]), tol=1e-6, every=100, max_steps=100000)
    assert plain["converged"] and r["steps"] < plain["steps"]


def test_relax_batch_plain_retires_each_box_at_its_own_test(ws):
    taus = [100.0, 300.0, 900.0]
    s = _FakeSolver(taus)
    r = ws.relax_batch(s, 3, None, tol=1e-6, every=100, max_steps=100000)
    assert r["converged"] == 3 and r["steps"][0] < r["steps"][1] < r["steps"][2] == r["total_steps"]
    for b, tau in enumerate(taus):  # the same box alone, with relax_plain's test
        alone = ws.relax_plain(_FakeSolver([tau]), tol=1e-6, every=100, max_steps=100000)
        assert alone["steps"] == r["steps"][b]
    assert not any(s.active)


def test_relax_batch_backs_off_after_refusals(ws):
    """Box 0's steps are always refused, box 1's always kept: box 0 is tried at cycles 0, 2, 5,
    10, ... (it sits out 1, 2, 4, ... cycles), box 1 every cycle until it is retired."""
    s = _FakeSolver([2000.0, 2000.0])
    net = _FakeNet(lambda b, k: b == 1)
    r = ws.relax_batch(s, 2, net, tol=1e-6, every=100, max_steps=6000)
    def attempts(log, box):  # cycle index (two plain steps per cycle) of each try of `box`
        out, nsteps = [], 0
        for e in log:
            if e[0] == "step":
                nsteps += 1
            elif e[0] == "apply_boxes" and box in e[1]:
                out.append(nsteps // 2)
        return out
    cyc0, cyc1 = attempts(s.log, 0), attempts(s.log, 1)
    assert cyc0[:5] == [0, 2, 5, 10, 19]
    assert cyc1 == list(range(len(cyc1))) and r["steps"][1] == 200 * len(cyc1)
    assert r["jumps_accepted"] == len(cyc1) and r["jumps_rejected"] == len(cyc0)
    without = _FakeSolver([2000.0, 2000.0])
    ws.relax_batch(without, 2, _FakeNet(lambda b, k: b == 1), tol=1e-6, every=100, max_steps=6000,
                   backoff=False)
    assert attempts(without.log, 0) == list(range(30))  # every cycle of the 6000 steps


def test_plain_tail_stops_learned_steps(ws):
    """With tail_tol, once a cycle's |dF/F| is below it no learned step is proposed again: the
    plain solver alone makes the last approach, in the single-box and the batched driver."""
    s = _FakeSolver([400.0])
    r = ws.relax_warm(s, _FakeNet(lambda b, k: False), tol=1e-6, every=100, max_steps=100000, tail_tol=1e-4)
    s0 = _FakeSolver([400.0])
    r0 = ws.relax_warm(s0, _FakeNet(lambda b, k: False), tol=1e-6, every=100, max_steps=100000)
    assert r["converged"] and r["steps"] == r0["steps"]  # refused steps change nothing
    assert r["jumps_rejected"] < r0["jumps_rejected"]
    # the proposals stop at the first cycle whose |dF/F| is below 1e-4 and never resume
    kinds = [e[0] for e in s.log]
    last_apply = max(i for i, k in enumerate(kinds) if k == "apply")
    assert "apply" not in kinds[last_apply + 1:]
    sb = _FakeSolver([400.0, 800.0])
    rb = ws.relax_batch(sb, 2, _FakeNet(lambda b, k: False), tol=1e-6, every=100, max_steps=100000,
                        backoff=False, tail_tol=1e-4)
    sb0 = _FakeSolver([400.0, 800.0])
    rb0 = ws.relax_batch(sb0, 2, _FakeNet(lambda b, k: False), tol=1e-6, every=100, max_steps=100000,
                         backoff=False)
    assert rb["steps"] == rb0["steps"] and rb["jumps_rejected"] < rb0["jumps_rejected"]


def test_short_cycles_test_from_the_energy_after_the_step(ws):
    """Short cycles: a learned step, `every` plain steps, one energy.  The test interval starts at the
    check's F_after of a kept step, and a cycle whose step was kept never ends the run."""
    every_kept = _FakeNet(lambda b, k: True, jump=1e-9)
    s = _FakeSolver([100.0])
    r = ws.relax_warm(s, every_kept, tol=1e-6, every=100, max_steps=3000, short=True)
    assert not r["converged"] and r["steps"] == 3000 and r["jumps_accepted"] == 30
    kinds = [e[0] for e in s.log]
    assert kinds[0] == "energies"
    assert all(kinds[i:i + 3] == ["apply", "step", "energies"] for i in range(1, len(kinds), 3))
    # with the tail rule: the first kept step leaves |dF/F| from its F_after far below 1e-5, so the next
    # cycle proposes nothing and converges (an interval starting before the step would not)
    s = _FakeSolver([100.0])
    r = ws.relax_warm(s, _FakeNet(lambda b, k: True, jump=1e-9), tol=1e-6, every=100, max_steps=3000,
                      short=True, tail_tol=1e-5)
    assert r["converged"] and r["steps"] == 200 and r["jumps_accepted"] == 1 and r["jumps_rejected"] == 0


def test_short_batch_cycles_match_the_plain_test_for_refused_boxes(ws):
    """Batched short cycles: box 0's steps are kept (it converges two cycles in, by the tail rule),
    box 1's are refused, so its test is relax_plain's on the same 100-step intervals."""
    s = _FakeSolver([100.0, 700.0])
    r = ws.relax_batch(s, 2, _FakeNet(lambda b, k: b == 0, jump=1e-9), tol=1e-6, every=100, max_steps=100000,
                       tail_tol=1e-5, short=True)
    assert r["converged"] == 2 and r["steps"][0] == 200
    alone = ws.relax_plain(_FakeSolver([700.0]), tol=1e-6, every=100, max_steps=100000)
    assert r["steps"][1] == alone["steps"]
    assert ("retire", (0,)) in s.log and not any(s.active)
