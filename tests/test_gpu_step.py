"""GPU parity of the fused FP32 step against the FP64 oracle (QSR.cu step order).

Each case uploads the same float32-representable field to the HIP path and the
oracle, advances both and compares.  Tolerances (FP32 stencil vs FP64 oracle):
one step: max |dQ| <= 2e-6 absolute; many steps: see test_gpu_relax.py.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEP_TOL = 2e-6


def _f32(Q):
    return Q.astype(np.float32).astype(np.float64)


def _case(q, orc, dims, L=(0.0, 0.0, 0.0), kappa=6.5e-12, W=0.0, R=-1.0, ctr=(math.nan,) * 3,
          d=(1e-8, 1e-8, 1e-8), T=300.0, Ts=307.0, init=0, custom=None, cap=0.0, seed=12345):
    L1, L2, L3 = L
    p = q.Params(Nx=dims[0], Ny=dims[1], Nz=dims[2], dx=d[0], dy=d[1], dz=d[2], T_star=Ts, T_high=T,
                 kappa=kappa, L1=L1, L2=L2, L3=L3, W=W, radius_index=R, center_index=ctr, q_norm_cap=cap,
                 boundary_custom=0 if custom is None else 1, boundary_S=0.0 if custom is None else custom)
    g = orc.Grid(*dims, *d)
    ph = orc.phys(T=T, T_star=Ts, kappa=kappa, L1=L1, L2=L2, L3=L3, W=W)
    dto = orc.dt_estimate(*d, 0.1, ph.a, ph.b, ph.c, Ts, 1, kappa, L1, L2, L3, math.nan, 0.5, T, T)
    p.dt = dto["dt_max"]
    od = orc.DropletOracle(g, ph, gamma=0.1, radius_index=R, center=ctr, q_norm_cap=cap)
    Q0 = od.init_director(0.5, init, seed) if init in (0, 1) else od.init_isotropic(1e-2, seed)
    Q0 = _f32(Q0)
    S_shell = orc.shell_order(custom is not None, custom or 0.0, ph.a, ph.b, ph.c, T, Ts)
    return p, od, Q0, S_shell


CASES = {
    "l1_strong_64": dict(dims=(24, 24, 24), L=(6e-12, 0, 0)),
    "kappa_strong_odd": dict(dims=(19, 23, 29), kappa=6.5e-12),
    "l1_weak_custom": dict(dims=(24, 20, 30), L=(6e-12, 0, 0), W=1e-4, custom=0.5),
    "l2_strong": dict(dims=(20, 21, 22), L=(6e-12, 3e-12, 0)),
    "l2_l3_weak": dict(dims=(22, 26, 70), L=(6e-12, 3e-12, 1e-12), W=1e-4, custom=0.5),
    "l3_only_weak": dict(dims=(18, 18, 18), L=(0, 0, 2e-12), W=1e-3),
    "aniso_spacing_l2": dict(dims=(17, 25, 33), L=(6e-12, 3e-12, 1e-12), d=(1e-8, 1.2e-8, 0.9e-8), W=2e-4),
    "offcentre_radius": dict(dims=(30, 30, 30), L=(6e-12, 0, 0), R=9.0, ctr=(12.0, 15.5, 16.0), W=1e-4),
    "touching_faces_weak_l2": dict(dims=(16, 16, 16), L=(6e-12, 3e-12, 0), W=1e-3),
    "cap": dict(dims=(20, 20, 20), L=(6e-12, 0, 0), cap=0.5 * math.sqrt(2 / 3)),
    "isotropic_init_quench_T": dict(dims=(24, 24, 24), L=(6e-12, 0, 0), init=2, T=295.0),
    "radial_init": dict(dims=(21, 21, 21), L=(6e-12, 2e-12, 0), init=1),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_single_step_matches_oracle(qsr, orc, name):
    p, od, Q0, S_shell = _case(qsr, orc, **CASES[name])
    print(f"[case] {name} dims={p.dims}", flush=True)
    with qsr.Solver(p) as s:
        s.upload(Q0)
        back = s.download()
        assert np.array_equal(back, Q0)
        for nsteps in (1, 2):
            s.step(1)
            Qg = s.download()
            Qo = od.step(Q0 if nsteps == 1 else Qo, S_shell, p.dt)
            err = np.abs(Qg - Qo).max()
            moved = np.abs(Qo - Q0).max()
            print(f"  step {nsteps}: max|Qgpu-Qorc| = {err:.3e}  (max|dQ| = {moved:.3e})")
            assert moved > 0
            assert err <= STEP_TOL


def test_static_voxels_untouched(qsr, orc):
    p, od, Q0, S_shell = _case(qsr, orc, (24, 24, 24), L=(6e-12, 0, 0))
    with qsr.Solver(p) as s:
        s.upload(Q0)
        s.step(3)
        Qg = s.download()
    Qo = od.step(Q0, S_shell, p.dt, steps=3)
    outside = np.abs(Q0).sum(-1) == 0
    assert np.array_equal(Qg[outside & (np.abs(Qo).sum(-1) == 0)], Qo[outside & (np.abs(Qo).sum(-1) == 0)])
    assert np.abs(Qg - Qo).max() < 4 * STEP_TOL


@pytest.mark.parametrize("noise", [0.0, 1.049e-3])
@pytest.mark.parametrize("dims", [(16, 16, 16), (20, 13, 70)])
def test_periodic_step_matches_oracle(qsr, orc, noise, dims):
    T, Ts, kappa, gamma = 301.0, 307.0, 6e-12, 0.1
    p = qsr.Params(Nx=dims[0], Ny=dims[1], Nz=dims[2], periodic=1, T_star=Ts, T_high=T, kappa=kappa,
                   gamma=gamma, dt=1.5e-10, noise_sigma=noise, seed=4242)
    g = orc.Grid(*dims)
    rng = np.random.default_rng(7)
    Q0 = _f32(rng.normal(0, 0.05, dims + (5,)))
    A, B, Cc = orc.bulk_ABC(p.a, p.b, p.c, T, Ts, 1)
    print(f"[case] periodic dims={dims} noise={noise}", flush=True)
    with qsr.Solver(p) as s:
        s.upload(Q0)
        Qo = Q0
        for it in range(3):
            s.step(1)
            Qo = orc.periodic_step(Qo, g, A, B, Cc, kappa, gamma, p.dt, noise, 4242, it)
            err = np.abs(s.download() - Qo).max()
            print(f"  step {it}: err {err:.3e}")
            assert err <= STEP_TOL


def test_batched_boxes_match_individual(qsr, orc):
    cases = [CASES["l1_weak_custom"], CASES["l2_l3_weak"], CASES["kappa_strong_odd"]]
    built = [_case(qsr, orc, **c, seed=100 + t) for t, c in enumerate(cases)]
    params = [b[0] for b in built]
    with qsr.Solver(params) as s:
        for t, b in enumerate(built):
            s.upload(b[2], box=t)
        s.step(2)
        for t, (p, od, Q0, S_shell) in enumerate(built):
            Qo = od.step(Q0, S_shell, p.dt, steps=2)
            err = np.abs(s.download(box=t) - Qo).max()
            print(f"  box {t}: err {err:.3e}")
            assert err <= 2 * STEP_TOL


def test_graph_steps_equal_plain_steps(qsr, orc):
    p, od, Q0, S_shell = _case(qsr, orc, **CASES["l2_l3_weak"])
    with qsr.Solver(p) as a, qsr.Solver(p) as b:
        a.upload(Q0)
        b.upload(Q0)
        a.step(37)
        for _ in range(3):
            b.step_graph(12, chunk=4)
        b.step(1)
        assert np.array_equal(a.download(), b.download())


@pytest.mark.parametrize("name", ["l2_l3_weak", "l1_weak_custom"])
def test_autotune_keeps_state_and_results(qsr, orc, name):
    p, od, Q0, S_shell = _case(qsr, orc, **CASES[name])
    with qsr.Solver(p) as a, qsr.Solver(p) as b:
        a.upload(Q0)
        b.upload(Q0)
        tune = a.autotune(2)
        print(f"  autotune {tune}")
        assert np.array_equal(a.download(), Q0)
        a.step(5)
        b.step(5)
        assert np.array_equal(a.download(), b.download())


@pytest.mark.parametrize("name,J", [("l1_strong_64", 25), ("l1_weak_custom", 25), ("l2_l3_weak", 7),
                                    ("touching_faces_weak_l2", 3)])
def test_imex_step_matches_oracle(qsr, orc, name, J):
    """Semi-implicit step: RHS = Q - (dt/gamma) mu - alpha lap Q, J Jacobi sweeps, strong BC (QSR.cu:2721-2753)."""
    p, od, Q0, S_shell = _case(qsr, orc, **CASES[name])
    L_stab = p.L1 if p.L1 else p.kappa
    p.dt *= 3.0  # the stabiliser allows a step beyond the explicit window
    print(f"[case] imex {name} J={J}", flush=True)
    with qsr.Solver(p) as s:
        s.upload(Q0)
        s.step_imex(2, L_stab, J)
        Qg = s.download()
    Qo = od.step_imex(Q0, S_shell, p.dt, L_stab, J, steps=2)
    err = np.abs(Qg - Qo).max()
    print(f"  max|Qgpu-Qorc| = {err:.3e}  (max|dQ| = {np.abs(Qo - Q0).max():.3e})")
    assert err <= 4 * STEP_TOL


@pytest.mark.parametrize("scale", [100.0, 1e4])
def test_imex_stiff_stabiliser_matches_oracle(qsr, orc, scale):
    """A stiff stabiliser (alpha/dx^2 ~ 3e-2 and ~3) keeps the Jacobi iterate moving for many sweeps:
    a converged-sweep skip taken too early would show against the oracle.  In FP32 a converged
    iterate can sit in a last-bit limit cycle, so more sweeps agree to a few ulp, not bit for bit."""
    p, od, Q0, S_shell = _case(qsr, orc, **CASES["l1_weak_custom"])
    L_stab = p.L1 * scale
    p.dt *= 3.0
    runs = {}
    for J in (25, 40):
        with qsr.Solver(p) as s:
            s.upload(Q0)
            s.step_imex(2, L_stab, J)
            runs[J] = s.download()
    Qo = od.step_imex(Q0, S_shell, p.dt, L_stab, 25, steps=2)
    err = np.abs(runs[25] - Qo).max()
    print(f"  scale {scale}: max|Qgpu-Qorc| = {err:.3e}  (max|dQ| = {np.abs(Qo - Q0).max():.3e})  "
          f"J25 vs J40 {np.abs(runs[25] - runs[40]).max():.3e}")
    assert err <= 4 * STEP_TOL
    if scale == 100.0:  # converged by 25 sweeps (contraction ~0.15 per sweep); at 1e4 it is not
        assert np.abs(runs[25] - runs[40]).max() <= 4 * STEP_TOL


def test_imex_converged_sweeps_are_exact(qsr, orc):
    """At the reference stabiliser (L_stab = L1) the FP32 sweeps reach a fixed point or a last-bit
    two-cycle within a few sweeps; later sweeps are skipped and must leave the result bit-identical to
    running them (same parity of J), while the first sweeps must not be skipped (sweep 2 still changes
    the field)."""
    p, od, Q0, S_shell = _case(qsr, orc, **CASES["l2_l3_weak"])
    p.dt *= 3.0
    runs = {}
    for J in (1, 2, 25, 41):
        with qsr.Solver(p) as s:
            s.upload(Q0)
            s.step_imex(3, p.L1, J)
            runs[J] = s.download()
    assert not np.array_equal(runs[1], runs[2])
    assert np.array_equal(runs[25], runs[41])
    Qo = od.step_imex(Q0, S_shell, p.dt, p.L1, 25, steps=3)
    assert np.abs(runs[25] - Qo).max() <= 4 * STEP_TOL


def test_holding_boxes_together_equals_holding_them_one_by_one(qsr, orc):
    """qsr_set_boxes_active on a list gives the fields and iteration counts of one qsr_set_box_active
    call per box, bit for bit, in an L1 / kappa batch (held and resumed mid-run)."""
    cases = [CASES["l1_weak_custom"], CASES["kappa_strong_odd"], CASES["l1_strong_64"], CASES["offcentre_radius"]]
    built = [_case(qsr, orc, **c, seed=400 + t) for t, c in enumerate(cases)]
    out = []
    for together in (True, False):
        with qsr.Solver([b[0] for b in built]) as s:
            for t, b in enumerate(built):
                s.upload(b[2], box=t)
            s.step(3)
            if together:
                s.set_active_many([2, 0], False)
            else:
                s.set_active(2, False)
                s.set_active(0, False)
            s.step(5)
            if together:
                s.set_active_many([0, 2, 3], True)  # box 3 is active already
            else:
                s.set_active(0, True)
                s.set_active(2, True)
            s.step(4)
            out.append(([s.download(box=t) for t in range(len(built))], [s.info(t)["iteration"] for t in range(4)]))
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b)
    assert out[0][1] == out[1][1] == [7, 12, 7, 12]


def test_held_box_is_frozen_and_others_continue(qsr, orc):
    """A held box keeps its field and iteration count while the rest of the batch steps on; resumed, it
    continues from where it stopped (quench-rate ladders retire finished droplets this way)."""
    cases = [CASES["l1_weak_custom"], CASES["kappa_strong_odd"], CASES["l2_strong"]]
    built = [_case(qsr, orc, **c, seed=300 + t) for t, c in enumerate(cases)]
    with qsr.Solver([b[0] for b in built]) as s:
        for t, b in enumerate(built):
            s.upload(b[2], box=t)
        s.step(2)
        s.set_active(1, False)
        s.step(3)           # odd count: the current buffer flips relative to the held copy
        held = s.download(box=1)
        assert s.info(1)["iteration"] == 2 and s.info(0)["iteration"] == 5
        s.step_graph(4, chunk=2)
        assert np.array_equal(s.download(box=1), held)
        s.set_active(1, True)
        s.step(1)
        for t, (p, od, Q0, S_shell) in enumerate(built):
            n = 3 if t == 1 else 10
            Qo = od.step(Q0, S_shell, p.dt, steps=n)
            err = np.abs(s.download(box=t) - Qo).max()
            print(f"  box {t}: {n} steps, err {err:.3e}")
            assert err <= 4 * STEP_TOL
        assert s.info(1)["iteration"] == 3 and s.info(2)["iteration"] == 10


@pytest.mark.parametrize("name", ["l2_l3_weak", "aniso_spacing_l2", "l1_weak_custom", "isotropic_init_quench_T"])
def test_tile_shapes_agree(qsr, orc, name):
    """Both tile shapes (64 k x 4 rows, 32 k x 8 rows) and several plane counts give the same field,
    bit for bit (every quantity a tile recomputes at its edges is rounded the same way), matching
    the oracle."""
    p, od, Q0, S_shell = _case(qsr, orc, **CASES[name])
    Qo = od.step(Q0, S_shell, p.dt, steps=3)
    got = {}
    for variant, ch in ((0, 8), (0, 32), (1, 8), (1, 16), (1, 18), (0, 14)):
        with qsr.Solver(p) as s:
            s.set_tuning(variant, ch)
            s.upload(Q0)
            s.step(3)
            got[(variant, ch)] = s.download()
        err = np.abs(got[(variant, ch)] - Qo).max()
        print(f"  variant {variant} ch {ch}: err {err:.3e}")
        assert err <= 4 * STEP_TOL
    ref = got[(0, 8)]
    for key, Q in got.items():
        assert np.array_equal(Q, ref), (key, np.abs(Q - ref).max())


# --- two steps per launch (temporal blocking of the L1 / kappa step) ---------------------

PAIR_CASES = ["l1_strong_64", "kappa_strong_odd", "l1_weak_custom", "offcentre_radius", "cap",
              "isotropic_init_quench_T"]


def _pair_vs_single(qsr, params, fields, n, graph_chunk=0):
    """Advance a context that pairs its steps and one that does not by n steps; return both
    downloads per box and whether the first one paired."""
    with qsr.Solver(params) as a, qsr.Solver(params) as b:
        b.set_step_fusion(False)
        for t, Q0 in enumerate(fields):
            a.upload(Q0, box=t)
            b.upload(Q0, box=t)
        active = a.step_fusion()["active"]
        assert not b.step_fusion()["active"]
        if graph_chunk:
            a.step_graph(n, chunk=graph_chunk)
        else:
            a.step(n)
        b.step(n)
        nb = len(fields)
        return [a.download(box=t) for t in range(nb)], [b.download(box=t) for t in range(nb)], active


@pytest.mark.parametrize("name", PAIR_CASES)
def test_step_pairs_equal_single_steps(qsr, orc, name):
    """Pairs of steps in one launch give the single-step field bit for bit (odd count: the last
    step runs alone)."""
    p, od, Q0, S_shell = _case(qsr, orc, **CASES[name])
    print(f"[case] pairs {name}", flush=True)
    (A,), (Bs,), active = _pair_vs_single(qsr, p, [Q0], 7)
    print(f"  paired={active} max|diff|={np.abs(A - Bs).max():.3e}")
    assert np.array_equal(A, Bs)


def test_step_pairs_are_used_and_match_oracle(qsr, orc):
    """A droplet clear of its box faces pairs its steps; six paired steps match the oracle."""
    p, od, Q0, S_shell = _case(qsr, orc, **CASES["offcentre_radius"])
    with qsr.Solver(p) as s:
        info = s.step_fusion()
        print(f"[case] pairs oracle offcentre_radius {info}", flush=True)
        assert info["active"] and info["tile_voxels"] > 0
        s.upload(Q0)
        s.step(6)
        Qo = od.step(Q0, S_shell, p.dt, steps=6)
        err = np.abs(s.download() - Qo).max()
        print(f"  6 paired steps: err {err:.3e}")
        assert err <= 3 * STEP_TOL


def test_step_pairs_follow_the_temperature_ramp(qsr, orc):
    """Each step of a pair takes its own row of the quench ramp (T, S_shell change every step)."""
    c = dict(CASES["offcentre_radius"])
    p, od, Q0, S_shell = _case(qsr, orc, **c)
    p.protocol = qsr.PROTOCOL_RAMP
    p.T_high, p.T_low, p.pre_equil_iters, p.ramp_iters = 300.0, 290.0, 1, 6
    print("[case] pairs ramp", flush=True)
    (A,), (Bs,), active = _pair_vs_single(qsr, p, [Q0], 9)
    assert active
    assert np.array_equal(A, Bs)


@pytest.mark.parametrize("dims", [(16, 16, 16), (20, 13, 70)])
def test_step_pairs_periodic(qsr, orc, dims):
    """Periodic box without noise: pairs equal single steps and the oracle's periodic step."""
    T, Ts, kappa, gamma = 301.0, 307.0, 6e-12, 0.1
    p = qsr.Params(Nx=dims[0], Ny=dims[1], Nz=dims[2], periodic=1, T_star=Ts, T_high=T, kappa=kappa,
                   gamma=gamma, dt=1.5e-10, seed=4242)
    g = orc.Grid(*dims)
    Q0 = _f32(np.random.default_rng(7).normal(0, 0.05, dims + (5,)))
    print(f"[case] pairs periodic dims={dims}", flush=True)
    (A,), (Bs,), active = _pair_vs_single(qsr, p, [Q0], 4)
    assert active
    assert np.array_equal(A, Bs)
    A_, B_, C_ = orc.bulk_ABC(p.a, p.b, p.c, T, Ts, 1)
    Qo = Q0
    for it in range(4):
        Qo = orc.periodic_step(Qo, g, A_, B_, C_, kappa, gamma, p.dt, 0.0, 4242, it)
    assert np.abs(A - Qo).max() <= 2 * STEP_TOL


def test_step_pairs_batch_and_graph(qsr, orc):
    """A batch of droplets with different protocols: paired steps, plain or replayed from graphs
    of 2, 4 and 6 steps, equal single steps."""
    built = [_case(qsr, orc, **CASES[n], seed=300 + t) for t, n in enumerate(["offcentre_radius", "cap"])]
    params = [b[0] for b in built]
    params[1].protocol = qsr.PROTOCOL_RAMP
    params[1].T_high, params[1].T_low, params[1].ramp_iters = 300.0, 292.0, 10
    fields = [b[2] for b in built]
    for chunk in (0, 2, 4, 6):
        print(f"[case] pairs batch graph chunk={chunk}", flush=True)
        A, Bs, active = _pair_vs_single(qsr, params, fields, 12, graph_chunk=chunk)
        for a, b in zip(A, Bs):
            assert np.array_equal(a, b)
