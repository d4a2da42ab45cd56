"""`QSR_hip --config` end to end: the key = value contract, stdout lines that
tools parse, output files and their formats, exit codes (QSR.cu:1910-3503).

CPU tests cover what the driver does before it needs a device (shell
diagnostics, guards).  GPU tests run sim_mode 1, 2 and 3 on small droplets
and compare the files the driver writes against the oracle stepping the same
seeded field with the same protocol.  Files hold %.6e values (6 significant
digits), so Q agrees to |dQ| <= 2e-5 after tens of FP32 steps (FP32 step error
~1e-7 per step plus 5e-7 print rounding), energies to 2e-5 relative.
"""
import math
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "lcdrsimulator_amd", "lib", "QSR_hip")

# The regexes tools/weak_anchor_mesh_check.py:316-323 applies to solver stdout.
SHELL_PATTERNS = {
    "voxels": r"Shell voxels\s*=\s*([0-9]+)",
    "volume": r"Shell volume\s*=\s*([0-9.eE+-]+)",
    "area": r"Estimated surface area\s*=\s*([0-9.eE+-]+)",
    "nominal": r"Nominal shell thickness\s*=\s*([0-9.eE+-]+)",
    "effective": r"Effective shell thickness\s*=\s*([0-9.eE+-]+)",
    "scale": r"Weak anchoring scale W/.*=\s*([0-9.eE+-]+)",
}

QUENCH_HEADER = ("iteration,time_s,dt_s,T_K,bulk,elastic,anchoring,total,radiality,avg_S,max_S,checker_rel,"
                 "defect_density_per_plaquette,defect_plaquettes_used,xi_grad_proxy,xi_grad_edges_used")
SINGLE_HEADER = ("iteration,bulk,elastic,anchoring,total,radiality,time,avg_S,max_S,defect_density_per_plaquette,"
                 "defect_plaquettes_used,xi_grad_proxy,xi_grad_edges_used")


def _need_cli():
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")


def run_cli(tmp_path, cfg, timeout=300):
    _need_cli()
    path = tmp_path / "run.cfg"
    path.write_text("".join(f"{k} = {v}\n" for k, v in cfg.items()))
    out = subprocess.run([CLI, "--config", str(path)], cwd=tmp_path, capture_output=True, text=True,
                         timeout=timeout)
    print(out.stdout[-3000:])
    print(out.stderr[-2000:])
    return out


def read_q(path, dims):
    a = np.loadtxt(path)
    assert a.shape == (np.prod(dims), 12)
    ijk = a[:, :3].astype(int)
    Q = np.zeros(tuple(dims) + (5,))
    Q[ijk[:, 0], ijk[:, 1], ijk[:, 2]] = a[:, [3, 4, 5, 7, 8]]
    # symmetric and traceless as written (QSR.cu:1818-1832)
    assert np.array_equal(a[:, 4], a[:, 6]) and np.array_equal(a[:, 5], a[:, 9]) and np.array_equal(a[:, 8], a[:, 10])
    return Q, a


def read_csv(path):
    lines = open(path).read().splitlines()
    rows = [list(map(float, ln.split(","))) for ln in lines[1:] if ln]
    return lines[0], np.array(rows)


BASE = dict(Nx=20, Ny=20, Nz=20, dx=1e-8, dy=1e-8, dz=1e-8, T_star=307.0, xi_guard_enabled="false",
            init_mode=0, random_seed=2024, L1=6e-12, W=1e-4, gamma=0.1)


def _oracle_for(cfg):
    dims = (cfg["Nx"], cfg["Ny"], cfg["Nz"])
    grid = O.Grid(*dims, cfg["dx"], cfg["dy"], cfg["dz"])
    ph = O.phys(T=cfg.get("T", 300.0), T_star=cfg["T_star"], kappa=0.0, L1=cfg["L1"], L2=cfg.get("L2", 0.0),
                L3=cfg.get("L3", 0.0), W=cfg["W"])
    od = O.DropletOracle(grid, ph, gamma=cfg["gamma"])
    Q0 = od.init_director(0.9, cfg["init_mode"], cfg["random_seed"]) if cfg["init_mode"] != 2 else None
    return dims, ph, od, Q0


def _f32(Q):
    return Q.astype(np.float32).astype(np.float64)


def test_cli_shell_diagnostics_match_oracle(tmp_path):
    """The weak-anchoring shell block parses with the mesh-check regexes and agrees with the oracle's mask."""
    cfg = dict(BASE, Nx=24, Ny=22, Nz=26, iterations=1, print_freq=1)
    out = run_cli(tmp_path, cfg)
    vals = {k: float(re.search(p, out.stdout).group(1)) for k, p in SHELL_PATTERNS.items()}
    grid = O.Grid(24, 22, 26)
    g = O.geometry(grid)
    _, count = O.shell_mask(grid, g)
    assert vals["voxels"] == count
    assert math.isclose(vals["volume"], count * 1e-24, rel_tol=1e-5)
    eff = O.shell_thickness(g, count)
    assert math.isclose(vals["effective"], eff, rel_tol=1e-5)
    assert math.isclose(vals["area"], count * 1e-24 / eff, rel_tol=1e-5)
    assert math.isclose(vals["nominal"], 2 * g.shell_half_thickness, rel_tol=1e-5)
    assert math.isclose(vals["scale"], 1e-4 / eff, rel_tol=1e-5)
    if out.returncode != 0:  # no device: the run stops at context creation, after the shell block
        assert "no HIP device" in out.stderr


def test_cli_guards_abort_with_status_1(tmp_path):
    out = run_cli(tmp_path, dict(Nx=16, Ny=16, Nz=16))  # default xi guard: xi/dh ~ 0.15 < 2
    assert out.returncode == 1
    assert "Aborting this run due to correlation-length guard." in out.stdout
    out = run_cli(tmp_path, dict(BASE, L1=-1e-12))
    assert out.returncode == 1
    assert "[ERROR] Unstable elastic parameters" in out.stdout
    out = run_cli(tmp_path, dict(BASE, use_frank_map="true", K1=6e-12, K2=4e-12, K3=12e-12))
    assert out.returncode == 1
    assert "[WARNING] Potentially unstable elastic parameters" in out.stdout
    assert re.search(r"Mapped L1=\S+, L2=\S+, L3=\S+ \(Q=S\(nn-I/3\), S_ref=", out.stdout)


@pytest.mark.gpu
def test_cli_quench_ramp_imex_matches_oracle(tmp_path):
    """sim_mode 3, ramp protocol through T*, semi-implicit default, GIF snapshots (QSR.cu:2412-2983)."""
    cfg = dict(BASE, sim_mode=3, protocol=2, T_high=309.0, T_low=301.0, pre_equil_iters=4, ramp_iters=16,
               total_iters=30, logFreq=10, snapshot_mode=1, snapshotFreq=10, out_dir="qout", jacobi_iters=9,
               boundary_order_mode="custom", boundary_S=0.5, S0=0.9)
    out = run_cli(tmp_path, cfg)
    assert out.returncode == 0
    dims, ph, od, Q0 = _oracle_for(cfg)
    w = O.dt_estimate(1e-8, 1e-8, 1e-8, 0.1, ph.a, ph.b, ph.c, 307.0, 1, 0.0, 6e-12, 0.0, 0.0, math.nan, 0.9,
                      309.0, 301.0, True, 6e-12, 9)
    dt = w["dt_max"]
    head, log = read_csv(tmp_path / "qout" / "quench_log.dat")
    assert head == QUENCH_HEADER
    assert log[:, 0].tolist() == [0, 10, 20, 29]
    Q = _f32(Q0)
    it_T = lambda it: 309.0 if it < 4 else (301.0 if it - 4 >= 16 else 309.0 + (301.0 - 309.0) * (it - 4) / 16)
    rows = {}
    for it in range(30):
        T = it_T(it)
        Q = od.step_imex(Q, 0.5, dt, 6e-12, 9, T=T)
        if it in (0, 10, 20, 29):
            od_T = O.DropletOracle(od.grid, O.phys(T=T, T_star=307.0, kappa=0.0, L1=6e-12, W=1e-4), gamma=0.1)
            rows[it] = (T, od_T.observables(Q, 0.5))
    for r in log:
        it = int(r[0])
        T, o = rows[it]
        assert math.isclose(r[1], dt * (it + 1), rel_tol=1e-5) and math.isclose(r[2], dt, rel_tol=1e-5)
        assert math.isclose(r[3], T, rel_tol=1e-9, abs_tol=1e-9)
        for col, key in ((4, "bulk"), (5, "elastic"), (6, "anchoring"), (7, "total")):
            assert abs(r[col] - o[key]) <= 2e-5 * abs(o["total"]) + 1e-30, (it, key, r[col], o[key])
        assert abs(r[8] - o["radiality"]) < 2e-5 and abs(r[9] - o["avg_S"]) < 2e-5 and abs(r[10] - o["max_S"]) < 2e-5
    Qf, _ = read_q(tmp_path / "qout" / "Qtensor_output_final.dat", dims)
    assert np.abs(Qf - Q).max() < 2e-5
    snaps = sorted(os.listdir(tmp_path / "qout"))
    assert "nematic_field_iter_0.dat" in snaps and "nematic_field_iter_29.dat" in snaps
    nem = open(tmp_path / "qout" / "nematic_field_final.dat").read().splitlines()
    assert nem[0] == "# i j k S nx ny nz" and len(nem) == 1 + np.prod(dims)
    assert re.fullmatch(r"\d+ \d+ \d+( -?\d\.\d{6}e[+-]\d\d){4}", nem[1])


@pytest.mark.gpu
def test_cli_quench_guard_aborts(tmp_path):
    """The fixed-dt instability guard ends the run with status 1 and no final state (QSR.cu:2890-2903, 2967-2972)."""
    cfg = dict(BASE, sim_mode=3, protocol=1, T_high=300.0, T_low=300.0, total_iters=12, logFreq=5,
               snapshot_mode=0, out_dir="qout", S_abort=0.2, S0=0.9, init_mode=1)
    out = run_cli(tmp_path, cfg)
    assert out.returncode == 1
    assert "[InstabilityGuard] Detected runaway/aliasing at iter 0" in out.stderr
    assert not (tmp_path / "qout" / "Qtensor_output_final.dat").exists()


@pytest.mark.gpu
def test_cli_single_temperature_matches_oracle(tmp_path):
    """sim_mode 1: explicit steps, log every print_freq, snapshots, final files (QSR.cu:3126-3476)."""
    cfg = dict(BASE, sim_mode=1, T=300.0, iterations=25, print_freq=10, L2=3e-12, S0=0.9)
    out = run_cli(tmp_path, cfg)
    assert out.returncode == 0
    dims, ph, od, Q0 = _oracle_for(cfg)
    head, log = read_csv(tmp_path / "free_energy_vs_iteration.dat")
    assert head == SINGLE_HEADER and log[:, 0].tolist() == [0, 10, 20]
    S_shell = O.shell_order(False, 0.0, ph.a, ph.b, ph.c, 300.0, 307.0)
    w = O.dt_estimate(1e-8, 1e-8, 1e-8, 0.1, ph.a, ph.b, ph.c, 307.0, 1, 0.0, 6e-12, 3e-12, 0.0, math.nan, 0.9,
                      300.0, 300.0)
    Q = _f32(Q0)
    for it in range(25):
        Q = od.step(Q, S_shell, w["dt_max"])
        if it % 10 == 0:
            o = od.observables(Q, S_shell)
            r = log[it // 10]
            assert abs(r[4] - o["total"]) <= 2e-5 * abs(o["total"])
            assert abs(r[7] - o["avg_S"]) < 2e-5
            assert r[10] == o["plaquettes"]
            assert (tmp_path / "output" / f"nematic_field_iter_{it}.dat").exists()
    Qf, _ = read_q(tmp_path / "Qtensor_output_final.dat", dims)
    assert np.abs(Qf - Q).max() < 2e-5
    assert (tmp_path / "nematic_field_final.dat").exists()


@pytest.mark.gpu
def test_cli_temperature_sweep_matches_oracle(tmp_path):
    """sim_mode 2: the field carries over from one temperature to the next (QSR.cu:2985-3124)."""
    cfg = dict(BASE, sim_mode=2, T_start=300.0, T_end=302.0, T_step=1.0, iterations=12, print_freq=5, S0=0.9)
    out = run_cli(tmp_path, cfg)
    assert out.returncode == 0
    dims, ph, od, Q0 = _oracle_for(cfg)
    head, summ = read_csv(tmp_path / "output_temp_sweep" / "summary.dat")
    assert head == "temperature,final_energy_including_anchoring,average_S"
    assert summ[:, 0].tolist() == [300.0, 301.0, 302.0]
    w = O.dt_estimate(1e-8, 1e-8, 1e-8, 0.1, ph.a, ph.b, ph.c, 307.0, 1, 0.0, 6e-12, 0.0, 0.0, math.nan, 0.9,
                      300.0, 302.0)
    Q = _f32(Q0)
    for t, T in enumerate((300.0, 301.0, 302.0)):
        S_shell = O.shell_order(False, 0.0, ph.a, ph.b, ph.c, T, 307.0)
        odT = O.DropletOracle(od.grid, O.phys(T=T, T_star=307.0, kappa=0.0, L1=6e-12, W=1e-4), gamma=0.1)
        Q = odT.step(Q, S_shell, w["dt_max"], steps=12)
        o = odT.observables(Q, S_shell)
        assert abs(summ[t, 1] - o["total"]) <= 2e-5 * abs(o["total"])
        assert abs(summ[t, 2] - o["avg_S"]) < 2e-5
        Qf, _ = read_q(tmp_path / "output_temp_sweep" / f"T_{T:.6f}" / "Qtensor_output_final.dat", dims)
        assert np.abs(Qf - Q).max() < 4e-5


@pytest.mark.gpu
def test_cli_warm_start_resumes_run(tmp_path):
    """init_qtensor_file: 15 steps, save, resume from the saved Qtensor file for 15 more steps ==
    30 uninterrupted steps, up to the 7 significant digits of the file."""
    cfg = dict(BASE, sim_mode=1, T=300.0, iterations=30, print_freq=10, L2=3e-12, S0=0.9,
               save_iteration_snapshots="false")
    runs = {}
    for name, extra in (("full", {}), ("first", dict(iterations=15))):
        d = tmp_path / name
        d.mkdir()
        out = run_cli(d, dict(cfg, **extra))
        assert out.returncode == 0
        runs[name] = d / "Qtensor_output_final.dat"
    d = tmp_path / "resume"
    d.mkdir()
    out = run_cli(d, dict(cfg, iterations=15, init_qtensor_file=str(runs["first"])))
    assert out.returncode == 0 and "[Init] Warm start from" in out.stdout
    dims = (cfg["Nx"], cfg["Ny"], cfg["Nz"])
    full, _ = read_q(runs["full"], dims)
    resumed, _ = read_q(d / "Qtensor_output_final.dat", dims)
    assert np.abs(full - resumed).max() < 2e-6
