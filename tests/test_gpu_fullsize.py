"""GPU parity at the production sizes, against the FP64 oracle on the same inputs.

* The production relaxation box: 256^3, droplet radius 120, L1/L2/L3 = 6e-12/3e-12/1e-12,
  Rapini-Papoular W = 1e-4 with custom shell order 0.5, T = 300 K, seeded random-director
  start (S0 = 0.5, seed 12345), for both stencil tile shapes.
* The periodic bulk quench box: 256^3, kappa = 6e-12, with and without Philox noise.
* Two steps per launch at 256^3: the L1 droplet (R = 120, strong anchoring) against single steps
  (bit for bit) and the oracle, and the noise-free periodic box against the oracle.

Tolerances (FP32 stencil vs FP64 oracle):
  max |Q_gpu - Q_oracle| <= 2e-6 per step taken (the single-step bound of the small cases);
  RMS <= 1e-7 per step; FP64 energy of the same field, GPU vs oracle: relative 1e-9;
  energy of the GPU result vs the oracle's result: relative 1e-6.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEP_TOL = 2e-6
N = 256


def _f32(Q):
    return Q.astype(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def production(orc):
    L1, L2, L3, W = 6e-12, 3e-12, 1e-12, 1e-4
    ph = orc.phys(T=300.0, T_star=307.0, kappa=0.0, L1=L1, L2=L2, L3=L3, W=W)
    od = orc.DropletOracle(orc.Grid(N, N, N), ph, gamma=0.1, radius_index=120.0)
    dt = orc.dt_estimate(1e-8, 1e-8, 1e-8, 0.1, ph.a, ph.b, ph.c, 307.0, 1, 0.0, L1, L2, L3, math.nan, 0.5,
                         300.0, 300.0)["dt_max"]
    Q0 = _f32(od.init_director(0.5, 0, 12345))
    steps = 3
    Qo = od.step(Q0, 0.5, dt, steps=steps)
    return dict(od=od, dt=dt, Q0=Q0, Qo=Qo, steps=steps, L=(L1, L2, L3), W=W)


@pytest.mark.parametrize("variant", [0, 1])
def test_production_box_matches_oracle(qsr, production, variant):
    P = production
    L1, L2, L3 = P["L"]
    p = qsr.Params(Nx=N, Ny=N, Nz=N, T_star=307.0, T_high=300.0, kappa=0.0, L1=L1, L2=L2, L3=L3, W=P["W"],
                   gamma=0.1, radius_index=120.0, boundary_custom=1, boundary_S=0.5, dt=P["dt"])
    print(f"[case] 256^3 R=120 L1/L2/L3 weak, tile variant {variant}, {P['steps']} steps", flush=True)
    with qsr.Solver(p) as s:
        s.upload(P["Q0"])
        s.set_tuning(variant, 16)
        s.step(P["steps"])
        Qg = s.download()
        obs_g = s.observables(want_defects=False, want_xi=False)
        s.upload(P["Qo"])
        obs_o_gpu = s.observables(want_defects=False, want_xi=False)
    d = np.abs(Qg - P["Qo"])
    err, rms = d.max(), math.sqrt(float((d ** 2).mean()))
    Fo = sum(P["od"].energy(P["Qo"], 0.5))
    Fg_ref = sum(P["od"].energy(Qg, 0.5))
    print(f"  max {err:.3e} rms {rms:.3e}; F gpu {obs_g['total']:.12e} oracle {Fo:.12e}")
    assert err <= P["steps"] * STEP_TOL
    assert rms <= P["steps"] * 1e-7
    # the device's FP64 energy of a field equals the oracle's of the same field
    assert abs(obs_g["total"] - Fg_ref) <= 1e-9 * abs(Fg_ref)
    assert abs(obs_o_gpu["total"] - Fo) <= 1e-9 * abs(Fo)
    assert abs(obs_g["total"] - Fo) <= 1e-6 * abs(Fo)


@pytest.mark.parametrize("noise", [0.0, 1.049e-3])
def test_periodic_quench_box_matches_oracle(qsr, orc, noise):
    T, Ts, kappa, gamma, dt = 301.0, 307.0, 6e-12, 0.1, 1.5e-10
    p = qsr.Params(Nx=N, Ny=N, Nz=N, periodic=1, T_star=Ts, T_high=T, kappa=kappa, gamma=gamma, dt=dt,
                   noise_sigma=noise, seed=7)
    g = orc.Grid(N, N, N)
    Q0 = _f32(np.random.default_rng(3).normal(0.0, 1e-3, (N, N, N, 5)))
    A, B, Cc = orc.bulk_ABC(p.a, p.b, p.c, T, Ts, 1)
    print(f"[case] 256^3 periodic noise={noise}", flush=True)
    with qsr.Solver(p) as s:
        s.upload(Q0)
        Qo = Q0
        for it in range(2):
            s.step(1)
            Qo = orc.periodic_step(Qo, g, A, B, Cc, kappa, gamma, dt, noise, 7, it)
        Qg = s.download()
    err = np.abs(Qg - Qo).max()
    print(f"  max {err:.3e}")
    assert err <= 2 * STEP_TOL


def test_l1_droplet_pairs_at_256(qsr, orc):
    """The L1 droplet at 256^3 (R = 120, strong anchoring: the 64^3 relaxation's physics at the
    production size) runs its steps in pairs: four paired steps equal four single steps bit for bit
    and match the FP64 oracle."""
    L1 = 6e-12
    ph = orc.phys(T=300.0, T_star=307.0, kappa=0.0, L1=L1, L2=0.0, L3=0.0, W=0.0)
    od = orc.DropletOracle(orc.Grid(N, N, N), ph, gamma=0.1, radius_index=120.0)
    dt = orc.dt_estimate(1e-8, 1e-8, 1e-8, 0.1, ph.a, ph.b, ph.c, 307.0, 1, 0.0, L1, 0.0, 0.0, math.nan, 0.5,
                         300.0, 300.0)["dt_max"]
    S_shell = orc.shell_order(False, 0.0, ph.a, ph.b, ph.c, 300.0, 307.0)
    Q0 = _f32(od.init_director(0.5, 0, 12345))
    p = qsr.Params(Nx=N, Ny=N, Nz=N, T_star=307.0, T_high=300.0, kappa=0.0, L1=L1, W=0.0, gamma=0.1,
                   radius_index=120.0, dt=dt)
    print("[case] 256^3 R=120 L1 strong, 4 steps in pairs vs single", flush=True)
    with qsr.Solver(p) as a, qsr.Solver(p) as b:
        assert a.step_fusion()["active"]
        b.set_step_fusion(False)
        a.upload(Q0)
        b.upload(Q0)
        a.step(4)
        b.step(4)
        Qa, Qb = a.download(), b.download()
    assert np.array_equal(Qa, Qb)
    Qo = od.step(Q0, S_shell, dt, steps=4)
    err = np.abs(Qa - Qo).max()
    print(f"  max {err:.3e}")
    assert err <= 4 * STEP_TOL


def test_periodic_box_pairs_at_256(qsr, orc):
    """The noise-free periodic box at 256^3 runs two steps per launch: they match the oracle."""
    T, Ts, kappa, gamma, dt = 301.0, 307.0, 6e-12, 0.1, 1.5e-10
    p = qsr.Params(Nx=N, Ny=N, Nz=N, periodic=1, T_star=Ts, T_high=T, kappa=kappa, gamma=gamma, dt=dt, seed=7)
    g = orc.Grid(N, N, N)
    Q0 = _f32(np.random.default_rng(3).normal(0.0, 1e-3, (N, N, N, 5)))
    A, B, Cc = orc.bulk_ABC(p.a, p.b, p.c, T, Ts, 1)
    print("[case] 256^3 periodic, 2 steps in one launch", flush=True)
    with qsr.Solver(p) as s:
        assert s.step_fusion()["active"]
        s.upload(Q0)
        s.step(2)
        Qg = s.download()
    Qo = Q0
    for it in range(2):
        Qo = orc.periodic_step(Qo, g, A, B, Cc, kappa, gamma, dt, 0.0, 7, it)
    err = np.abs(Qg - Qo).max()
    print(f"  max {err:.3e}")
    assert err <= 2 * STEP_TOL
