"""End-to-end GPU parity: multi-step relaxation and every log-point observable.

Tolerances (FP32 stencil vs FP64 oracle, stated per check):
  * observables of the same field: relative 1e-9 for energies / averages,
    defect and plaquette counts identical;
  * 64^3 droplet (R = 28, strong anchoring, L1 only) after 2000 steps:
    max |Q_gpu - Q_oracle| <= 1e-4, RMS <= 1e-5, |F_gpu/F_oracle - 1| <= 1e-5
    at every log point, 2D defect count identical at every log point.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _f32(Q):
    return Q.astype(np.float32).astype(np.float64)


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


OBS_CASES = {
    "strong_l1": dict(dims=(40, 40, 40), L=(6e-12, 0, 0), W=0.0),
    "weak_l2l3": dict(dims=(36, 38, 40), L=(6e-12, 3e-12, 1e-12), W=1e-4, custom=0.5),
    "kappa_offc": dict(dims=(33, 30, 35), L=(0, 0, 0), W=5e-4, R=11.0, ctr=(15.0, 14.5, 17.0)),
}


@pytest.mark.parametrize("name", sorted(OBS_CASES))
@pytest.mark.parametrize("init", [0, 1])
def test_observables_same_field(qsr, orc, name, init):
    c = OBS_CASES[name]
    dims, (L1, L2, L3), W = c["dims"], c["L"], c["W"]
    R, ctr = c.get("R", -1.0), c.get("ctr", (math.nan,) * 3)
    custom = c.get("custom")
    p = qsr.Params(Nx=dims[0], Ny=dims[1], Nz=dims[2], T_star=307.0, T_high=300.0, kappa=6.5e-12, L1=L1, L2=L2,
                   L3=L3, W=W, radius_index=R, center_index=ctr, dt=1e-10,
                   boundary_custom=0 if custom is None else 1, boundary_S=custom or 0.0)
    ph = orc.phys(T=300.0, T_star=307.0, kappa=6.5e-12, L1=L1, L2=L2, L3=L3, W=W)
    od = orc.DropletOracle(orc.Grid(*dims), ph, radius_index=R, center=ctr)
    Q = _f32(od.init_director(0.8, init, 99))
    rng = np.random.default_rng(3)
    Q = _f32(Q + (np.abs(Q).sum(-1, keepdims=True) > 0) * rng.normal(0, 0.05, Q.shape))
    S_shell = orc.shell_order(custom is not None, custom or 0.0, ph.a, ph.b, ph.c, 300.0, 307.0)
    print(f"[case] obs {name} init={init}", flush=True)
    with qsr.Solver(p) as s:
        s.upload(Q)
        g = s.observables()
        info = s.info()
    o = od.observables(Q, S_shell)
    assert info["shell_voxels"] == od.shell_count
    assert _rel(info["effective_thickness"], od.thickness) < 1e-14
    for k in ("bulk", "elastic", "total", "avg_S", "max_S", "radiality"):
        print(f"  {k}: gpu {g[k]:.12e} oracle {o[k]:.12e}")
        assert _rel(g[k], o[k]) < 1e-9, k
    if W != 0.0:
        assert _rel(g["anchoring"], o["anchoring"]) < 1e-9
    assert (g["defects"], g["plaquettes"]) == (o["defects"], o["plaquettes"])
    assert g["xi_edges"] == o["xi_edges"]
    assert _rel(g["xi_grad_proxy"], o["xi_grad_proxy"]) < 1e-9
    if math.isnan(o["checker_rel"]):
        assert math.isnan(g["checker_rel"])
    else:
        assert _rel(g["checker_rel"], o["checker_rel"]) < 1e-9


def test_periodic_observables_same_field(qsr, orc):
    dims = (24, 20, 28)
    p = qsr.Params(Nx=dims[0], Ny=dims[1], Nz=dims[2], periodic=1, T_star=307.0, T_high=298.0, kappa=6e-12,
                   dt=1e-10)
    rng = np.random.default_rng(11)
    # smooth-ish ordered field with defects: random directors low-pass filtered
    n = rng.normal(size=dims + (3,))
    for ax in range(3):
        n = n + np.roll(n, 1, ax) + np.roll(n, -1, ax)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    S = 0.6
    Q = np.stack([S * (n[..., 0] ** 2 - 1 / 3), S * n[..., 0] * n[..., 1], S * n[..., 0] * n[..., 2],
                  S * (n[..., 1] ** 2 - 1 / 3), S * n[..., 1] * n[..., 2]], -1)
    Q = _f32(Q)
    A, B, Cc = orc.bulk_ABC(p.a, p.b, p.c, 298.0, 307.0, 1)
    with qsr.Solver(p) as s:
        s.upload(Q)
        g = s.observables()
    o = orc.periodic_observables(Q, orc.Grid(*dims), A, B, Cc, 6e-12)
    print({k: (g.get(k), o[k]) for k in o if k in g})
    for k in ("bulk", "elastic", "total", "avg_S", "max_S", "xi_grad_proxy", "defect_line_density"):
        assert _rel(g[k], o[k]) < 1e-9, k
    assert g["defects"] == o["defects"] and g["plaquettes"] == o["plaquettes_used"]
    assert o["defects"] > 0


@pytest.mark.slow
def test_config0_relaxation_2000_steps(qsr, orc):
    """64^3 box, droplet R = 28 voxels, dx = 10 nm, L1 = 6e-12, strong radial anchoring with the equilibrium
    shell order, T = 300 K, T* = 307 K, random directors S0 = 0.5 (seed 12345), dt = dt_max, 2000 steps."""
    dims, R = (64, 64, 64), 28.0
    ph = orc.phys(T=300.0, T_star=307.0, kappa=0.0, L1=6e-12)
    od = orc.DropletOracle(orc.Grid(*dims), ph, gamma=0.1, radius_index=R)
    dto = orc.dt_estimate(1e-8, 1e-8, 1e-8, 0.1, ph.a, ph.b, ph.c, 307.0, 1, 0.0, 6e-12, 0, 0, math.nan, 0.5,
                          300.0, 300.0)
    dt = dto["dt_max"]
    assert abs(dt / 1.60372e-10 - 1) < 1e-5
    S_shell = orc.shell_order(False, 0.0, ph.a, ph.b, ph.c, 300.0, 307.0)
    p = qsr.Params(Nx=64, Ny=64, Nz=64, T_star=307.0, T_high=300.0, kappa=0.0, L1=6e-12, radius_index=R, dt=dt)
    Q0 = qsr.init_field(p, 0, 0.5, 0.0, 12345)
    assert np.array_equal(Q0, od.init_director(0.5, 0, 12345)) or np.abs(Q0 - od.init_director(0.5, 0, 12345)).max() < 1e-15
    Q0 = _f32(Q0)
    Qo = Q0
    worst = dict(dF=0.0)
    with qsr.Solver(p) as s:
        s.upload(Q0)
        for blk in range(20):
            s.step(100)
            Qo = od.step(Qo, S_shell, dt, steps=100)
            g = s.observables()
            o = od.observables(Qo, S_shell)
            dF = _rel(g["total"], o["total"])
            worst["dF"] = max(worst["dF"], dF)
            print(f"  iter {100 * (blk + 1)}: F gpu {g['total']:.9e} orc {o['total']:.9e} rel {dF:.2e}  "
                  f"defects {g['defects']}/{o['defects']}", flush=True)
            assert dF <= 1e-5
            assert g["defects"] == o["defects"]
        Qg = s.download()
    err = np.abs(Qg - Qo)
    print(f"  final max|dQ| {err.max():.3e} rms {math.sqrt((err ** 2).mean()):.3e} worst dF {worst['dF']:.2e}")
    assert err.max() <= 1e-4
    assert math.sqrt((err ** 2).mean()) <= 1e-5


def test_energy_only_observables_equal_the_full_pass(qsr):
    """skip_order (want_order=False) drops the nematic pass and, with defects and xi off, the
    mid-plane pass: the energies are the full pass's to the last bit, the order columns NaN."""
    p = qsr.Params(Nx=48, Ny=44, Nz=52, T_star=307.0, T_high=300.0, kappa=0.0, L1=6e-12, L2=3e-12, L3=1e-12,
                   W=1e-4, boundary_custom=1, boundary_S=0.5, dt=1.5e-10)
    with qsr.Solver(p) as s:
        s.upload(qsr.init_field(p, 0, 0.5, seed=21))
        s.step(5)
        full = s.observables()
        fast = s.observables(want_defects=False, want_xi=False, want_order=False)
    for k in ("bulk", "elastic", "anchoring", "total"):
        assert fast[k] == full[k], k
    assert math.isnan(fast["avg_S"]) and math.isnan(fast["max_S"]) and math.isnan(fast["radiality"])
    assert not math.isnan(full["avg_S"]) and full["max_S"] > 0
