"""GPU coverage of the two batched / learned workloads at their stated sizes.

* The 64-droplet quench-rate ladder: radii R = 16..48 voxels, each droplet in
  its own (2R+8)^3 box (28.6 M voxels in one context), temperature ramps
  310 -> 300 K over 1e3..1e5 steps (8 rates x 8 seeds), a quarter of the
  droplets with strong anchoring and the rest weak with W log-uniform in
  1e-5..1e-3 J/m^2, L1 = 6e-12 N, isotropic-noise starts.
* The learned warm start at its application size: 192^3 box, R = 88,
  T = 300 K, W = 1e-4 J/m^2 (custom shell order 0.5), L1/L2/L3.

Tolerances (FP32 stencil vs the FP64 oracle, stated per check):
  ladder, 20 ramped steps: max |Q_gpu - Q_oracle| <= 2e-3 * max |Q_oracle - Q_0|
    (the change over the 20 steps; a temperature read one step off moves it
    by several percent) and <= 2e-6 absolute;
  graph replay of the autotuned batch: bit-exact against plain launches;
  network forward on a crop of the 192^3 field: the forward tolerance of the
    network tests, 2e-3 * max |dQ|;
  refused proposal: the field is restored bit for bit.
"""
import math

import numpy as np
import pytest

from oracle import surrogate as S

pytestmark = pytest.mark.gpu

FWD_TOL = 2e-3
HORIZON = 500.0
COMPARED = (0, 31, 32, 63)  # R = 16 strong, R = 32 weak, R = 32 strong, R = 48 weak


def _T_at(p, it):
    """compute_T_current (QSR.cu:2679-2687) for the ramp protocol without pre-equilibration."""
    t = it - p.pre_equil_iters
    if it < p.pre_equil_iters or t <= 0:
        return p.T_high
    if t >= p.ramp_iters:
        return p.T_low
    return p.T_high + (p.T_low - p.T_high) * (t / p.ramp_iters)


@pytest.fixture(scope="module")
def ladder(qsr):
    import bench
    params = bench.ladder_params(qsr)
    assert len(params) == 64 and sum(p.n for p in params) == 28615680
    radii = [p.radius_index for p in params]
    assert min(radii) == 16.0 and max(radii) == 48.0
    assert all(p.Nx == 2 * int(p.radius_index) + 8 for p in params)
    assert sum(p.W == 0.0 for p in params) == 16 and all(1e-5 <= p.W <= 1e-3 for p in params if p.W)
    assert sorted({p.ramp_iters for p in params}) == [int(r) for r in np.logspace(3, 5, 8)]
    fields = [qsr.init_field(p, 2, 0.5, 1e-3, 1000 + b) for b, p in enumerate(params)]
    return params, fields


def test_ladder_droplets_follow_the_oracle_through_their_ramps(qsr, orc, ladder):
    """All 64 droplets stepped together for 20 steps of their ramps in one context; four of them
    (R = 16, 32, 32, 48; strong and weak anchoring; ramps of 1e3 and 1e5 steps) against the FP64
    oracle stepped with the same T(iter) and equilibrium shell order S_eq(T(iter)) every step."""
    params, fields = ladder
    nsteps = 20
    print(f"[case] ladder 64 boxes, {nsteps} ramped steps", flush=True)
    with qsr.Solver(params) as s:
        for b, Q in enumerate(fields):
            s.upload(Q, box=b)
        s.step(nsteps)
        got = {b: s.download(box=b) for b in COMPARED}
        last = [s.download(box=b) for b in range(64)]
    for b in range(64):
        assert np.isfinite(last[b]).all(), b
    for b in COMPARED:
        p = params[b]
        ph = orc.phys(T=p.T_high, T_star=p.T_star, kappa=p.kappa, L1=p.L1, W=p.W)
        od = orc.DropletOracle(orc.Grid(p.Nx, p.Ny, p.Nz), ph, gamma=p.gamma, radius_index=p.radius_index)
        Q0 = fields[b].astype(np.float32).astype(np.float64)
        Q = Q0
        for it in range(nsteps):
            T = _T_at(p, it)
            S_shell = orc.shell_order(False, 0.0, p.a, p.b, p.c, T, p.T_star)
            Q = od.step(Q, S_shell, p.dt, T=T)
        err = np.abs(got[b] - Q).max()
        moved = np.abs(Q - Q0).max()
        print(f"  box {b}: R={p.radius_index:g} W={p.W:.2e} ramp={p.ramp_iters} T_end={_T_at(p, nsteps):.4f}: "
              f"max err {err:.3e} of max change {moved:.3e}")
        assert moved > 0
        assert err <= 2e-3 * moved and err <= 2e-6


def test_ladder_energy_descends_in_every_droplet_at_constant_temperature(qsr, ladder):
    """At a fixed T = 300 K below the transition every droplet's free energy (FP64, energies only)
    falls between checks 25 steps apart, in all 64 boxes, and the fields stay finite.  The first
    check follows 5 steps: the strong shell is set by the first step, not by the upload."""
    import dataclasses
    params, fields = ladder
    const = [dataclasses.replace(p, protocol=qsr.PROTOCOL_CONSTANT, T_high=300.0) for p in params]
    print("[case] ladder 64 boxes at 300 K, energy descent", flush=True)
    with qsr.Solver(const) as s:
        for b, Q in enumerate(fields):
            s.upload(Q, box=b)
        s.step(5)
        F = [s.energies(list(range(64)))]
        for _ in range(4):
            s.step(25)
            F.append(s.energies(list(range(64))))
        Fin = [s.download(box=b) for b in (0, 63)]
    F = np.array(F)
    assert np.isfinite(F).all()
    falls = (F[1:] < F[:-1]).all(axis=0)
    print(f"  F(100)/F(0) min {np.min(F[-1] / F[0]):.3e} max {np.max(F[-1] / F[0]):.3e}; "
          f"{int(falls.sum())} of 64 descend at every check")
    assert falls.all(), np.nonzero(~falls)[0]
    assert all(np.isfinite(Q).all() for Q in Fin)


def test_ladder_autotuned_graph_replay_is_bit_exact(qsr, ladder):
    """The autotuned tile shape and hipGraph-captured step loop on the whole ladder: 40 steps replayed
    from graphs of 16 steps equal 40 plain launches bit for bit, in every box."""
    params, fields = ladder
    print("[case] ladder 64 boxes, autotune + graph replay", flush=True)
    with qsr.Solver(params) as s:
        for b, Q in enumerate(fields):
            s.upload(Q, box=b)
        tune = s.autotune(3)
        for b, Q in enumerate(fields):
            s.upload(Q, box=b)
            s.set_iteration(0, box=b)
        s.step(40)
        plain = [s.download(box=b) for b in range(64)]
        for b, Q in enumerate(fields):
            s.upload(Q, box=b)
            s.set_iteration(0, box=b)
        s.step_graph(40, 16)
        for b in range(64):
            np.testing.assert_array_equal(s.download(box=b), plain[b])
    print(f"  tuning {tune}")


# ---------------------------------------------------------------------------
# The learned warm start at 192^3, R = 88.
# ---------------------------------------------------------------------------
def _increment(Q, Q1):
    """u = Q1 - Q in FP32 on the voxels a step changed (box faces excluded) and its RMS s."""
    step = np.any(Q1 != Q, axis=-1)
    step[0, :, :] = step[-1, :, :] = step[:, 0, :] = step[:, -1, :] = step[:, :, 0] = step[:, :, -1] = False
    u = np.where(step[..., None], Q1.astype(np.float32) - Q.astype(np.float32), np.float32(0))
    s = math.sqrt(float((u.astype(np.float64) ** 2).sum()) / (5.0 * step.sum()))
    return u, s, step


@pytest.fixture(scope="module")
def warm192(qsr):
    from lcdrsimulator_amd import warmstart as ws
    p = ws.droplet_params(88.0, 300.0, 1e-4, N=192)
    assert (p.Nx, p.radius_index, p.W, p.L2, p.L3) == (192, 88.0, 1e-4, 3e-12, 1e-12)
    return p


def test_warm_start_apply_at_192_matches_the_oracle_network_on_a_crop(qsr, warm192):
    """Forced proposal on the 192^3 droplet under increment input: on a block of 8 x 8 rows
    (all 192 k) that crosses the droplet surface, Q_new - Q = N u (1 + y) with y the FP64 oracle
    network (BF16 rounding points emulated) evaluated on the crop with a 3-voxel halo, so the
    crop's zero padding cannot reach the block."""
    p = warm192
    params = S.random_params(31, scale=1.0).astype(np.float32)
    print("[case] 192^3 R=88 forced increment proposal vs oracle crop", flush=True)
    with qsr.Solver(p) as s, qsr.Surrogate(seed=None) as net:
        net.params = params
        net.set_input(qsr.SURR_INPUT_INCREMENT, HORIZON)
        s.upload(qsr.init_field(p, 1, 0.5, 0.0, 0))
        s.step(10)
        Q = s.download()
        s.step(1)
        Q1 = s.download()
        s.upload(Q)
        r = net.apply(s, qsr.SURR_FORCE)
        Qn = s.download()
    assert r["accepted"] == 1
    u, srms, step = _increment(Q, Q1)
    i0, j0, h = 92, 12, 3
    sl = (slice(i0 - h, i0 + 8 + h), slice(j0 - h, j0 + 8 + h))
    x = (u[sl] * np.float32(1.0 / srms)).astype(np.float64)
    y = S.forward(params.astype(np.float64), x[None], emulate=True)["y"][0]
    blk = (slice(h, h + 8), slice(h, h + 8))
    ub = u[sl][blk].astype(np.float64)
    dq_want = HORIZON * ub * (1.0 + y[blk])
    mask = step[sl][blk]
    dq_got = (Qn - Q)[sl][blk]
    err = np.abs(np.where(mask[..., None], dq_got - dq_want, 0.0)).max()
    scale = np.abs(dq_want[mask]).max()
    print(f"  block: {int(mask.sum())} changeable voxels, max |dQ| {scale:.3e}, max err {err:.3e}")
    assert mask.sum() > 2000 and (~mask).sum() > 1000  # inside and outside the droplet
    assert err <= FWD_TOL * scale + 1e-6
    np.testing.assert_array_equal(Qn[~step], Q[~step])


def test_warm_start_at_192_keeps_a_good_step_and_restores_a_refused_one(qsr, warm192):
    """Checked proposals on the 192^3 droplet: the zero network (the increment extrapolated over
    N = 500 steps, the warm start's untrained form) is kept early in the relaxation and lowers F
    without raising the residual; a large random network is refused and the field is restored bit
    for bit, after which the solver keeps stepping."""
    p = warm192
    print("[case] 192^3 R=88 checked apply, keep and refuse", flush=True)
    with qsr.Solver(p) as s, qsr.Surrogate(seed=None) as zero, qsr.Surrogate(seed=None) as bad:
        zero.params = np.zeros(S.NPARAM, dtype=np.float32)
        zero.set_input(qsr.SURR_INPUT_INCREMENT, HORIZON)
        bad.params = S.random_params(21, scale=30.0).astype(np.float32)
        bad.set_input(qsr.SURR_INPUT_INCREMENT, HORIZON)
        s.upload(qsr.init_field(p, 1, 0.5, 0.0, 0))
        s.step(100)
        r = zero.apply(s, qsr.SURR_CHECKED, scales=[1.0, 0.5, 0.25])
        print(f"  zero network: kept {r['accepted']} at {r['scale']}, F {r['F_before']:.6e} -> {r['F_after']:.6e}, "
              f"res {r['res_before']:.3e} -> {r['res_after']:.3e}")
        assert r["accepted"] == 1
        assert r["F_after"] < r["F_before"] and r["res_after"] <= r["res_before"]
        s.step(20)
        Qa = s.download()
        rb = bad.apply(s, qsr.SURR_CHECKED, scales=[1.0, 0.5])
        print(f"  random network: kept {rb['accepted']} after {rb['tried']} tries, "
              f"F {rb['F_before']:.6e} -> {rb['F_after']:.6e}, res {rb['res_before']:.3e} -> {rb['res_after']:.3e}")
        assert rb["accepted"] == 0 and rb["tried"] == 2
        np.testing.assert_array_equal(s.download(), Qa)
        s.step(5)
        assert np.isfinite(s.download()).all()
