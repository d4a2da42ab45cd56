#!/usr/bin/env python3
"""Throughput bench of the fused FP32 LdG step on one MI355X.

Workload (the production relaxation): a 256^3 box, droplet radius 120 voxels,
dx = 10 nm, bulk a = 0.044e6, b = 1.413e6, c = 1.153e6, T* = 307 K, T = 300 K,
L1/L2/L3 = 6e-12 / 3e-12 / 1e-12 N, Rapini-Papoular anchoring W = 1e-4 J/m^2
with custom shell order 0.5, gamma = 0.1 Pa s, dt = dt_max of the stability
window; seeded random-director start (S0 = 0.5, seed 12345).  A "step" is one
explicit-Euler iteration of the whole box (fused stencil launch + box-face
launch).  Inputs are resident in HBM when the timed region starts.

Prints one JSON line:
* value = droplet voxel-updates/s (the voxels a step changes, over the whole
  timed step); the box-equivalent and launched-tile rates are separate fields;
* roofline = algorithmic bytes of the fused stencil (40 B per droplet voxel:
  read 5 FP32, write 5 FP32) / its average launch time from hipEvents on the
  context stream, so value x 40 B / 8 TB/s ~ roofline.frac; tile_frac counts
  the launched tiles' voxels instead; traffic and the compute figure come from
  the committed rocprofv3 summaries named beside them;
* config.time_to_solution = the second half of the metric: the 64 quench-ladder
  droplets relaxed to |dF/F| < 1e-6 batched and warm-started against one plain
  context per droplet;
* cpu_baseline = the reference's own QSR.cpp step functions (oracle/_ref,
  OpenMP) timed on this host on a bounded sample of the same box.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_VOXEL = 40.0
METRIC = "Q-voxel updates/s & % HBM peak; time-to-ΔF/F<1e-6 batched/surrogate vs plain"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


WORKLOADS = ("droplet_l1l2l3_256", "periodic_noise_256", "periodic_noise_256x8", "droplet_l1_64", "droplet_l1_256",
             "ladder_64", "ladder_relax_64", "surrogate_192")
MFMA_BF16_PEAK_TFLOPS = 2500.0
FP32_VALU_PEAK_TFLOPS = 78.6  # 256 CUs x 4 SIMD x 16 lanes x 2 flop x 2.4 GHz: FP32 without packed issue
NET_FLOP_PER_VOXEL = 72576.0  # 2 * 27 * (5*32 + 32*32 + 32*5) multiply-adds of the 5-32-32-5 network


def ladder_params(q, n_drops=64, seed=2024):
    """Quench-rate ladder: droplet radius R = 16 + 32 i / 63 voxels in a (2R+8)^3 box, ramp 310 -> 300 K over
    ramp_iters log-spaced 1e3..1e5 (8 rates x 8 seeds), a quarter with strong anchoring and the rest weak with
    W log-uniform in 1e-5..1e-3 J/m^2, L1 = 6e-12 N, isotropic start with 1e-3 noise."""
    import numpy as np
    rng = np.random.default_rng(seed)
    rates = np.logspace(3, 5, 8)
    w = q.dt_estimate(1e-8, 1e-8, 1e-8, 0.1, 0.044e6, 1.413e6, 1.153e6, 307.0, 1, 0.0, 6e-12, 0, 0, math.nan, 0.5,
                      310.0, 300.0)
    out = []
    for i in range(n_drops):
        R = int(round(16 + 32 * i / (n_drops - 1)))
        N = 2 * R + 8
        W = 0.0 if i % 4 == 0 else float(10 ** rng.uniform(-5, -3))
        out.append(q.Params(Nx=N, Ny=N, Nz=N, dx=1e-8, dy=1e-8, dz=1e-8, T_star=307.0, kappa=0.0, L1=6e-12, W=W,
                            gamma=0.1, radius_index=float(R), protocol=q.PROTOCOL_RAMP, T_high=310.0, T_low=300.0,
                            ramp_iters=int(rates[i % 8]), seed=i // 8, dt=w["dt_max"]))
    return out


def workload_params(q, n=256, R=120.0, workload="droplet_l1l2l3_256"):
    if workload == "periodic_noise_256":
        # periodic bulk quench: 256^3, T ramp 310 -> 295 K, kappa = 6e-12, noise 1e-3 * S_eq(295 K)
        S295 = q.S_eq(0.044e6, 1.413e6, 1.153e6, 295.0, 307.0, 1)
        p = q.Params(Nx=n, Ny=n, Nz=n, dx=1e-8, dy=1e-8, dz=1e-8, T_star=307.0, kappa=6e-12, gamma=0.1,
                     periodic=1, protocol=q.PROTOCOL_RAMP, T_high=310.0, T_low=295.0, ramp_iters=20000,
                     noise_sigma=1e-3 * S295, seed=0)
        w = q.dt_estimate(p.dx, p.dy, p.dz, p.gamma, p.a, p.b, p.c, p.T_star, 1, p.kappa, 0, 0, 0, math.nan, 0.5,
                          310.0, 295.0)
        p.dt = w["dt_max"]
        return p
    if workload in ("droplet_l1_64", "droplet_l1_256"):
        # the config-0 physics (L1 = 6e-12, strong equilibrium anchoring) at 64^3, R = 28 or 256^3, R = 120
        big = workload == "droplet_l1_256"
        m = 256 if big else 64
        p = q.Params(Nx=m, Ny=m, Nz=m, dx=1e-8, dy=1e-8, dz=1e-8, T_star=307.0, T_high=300.0, kappa=0.0,
                     L1=6e-12, gamma=0.1, radius_index=120.0 if big else 28.0)
        w = q.dt_estimate(p.dx, p.dy, p.dz, p.gamma, p.a, p.b, p.c, p.T_star, 1, 0.0, p.L1, 0, 0, math.nan, 0.5,
                          300.0, 300.0)
        p.dt = w["dt_max"]
        return p
    p = q.Params(Nx=n, Ny=n, Nz=n, dx=1e-8, dy=1e-8, dz=1e-8, a=0.044e6, b=1.413e6, c=1.153e6, T_star=307.0,
                 T_high=300.0, kappa=0.0, L1=6e-12, L2=3e-12, L3=1e-12, W=1e-4, gamma=0.1, radius_index=R,
                 boundary_custom=1, boundary_S=0.5, seed=12345)
    w = q.dt_estimate(p.dx, p.dy, p.dz, p.gamma, p.a, p.b, p.c, p.T_star, 1, p.kappa, p.L1, p.L2, p.L3, math.nan,
                      0.5, p.T_high, p.T_high)
    p.dt = w["dt_max"]
    return p


def profiled(tag, key, workload, variant, planes):
    """A figure recorded in the committed counter summaries under profiles/ as a line
    `<tag> <workload> variant=V planes=P <key>=X`, for the same workload and tuning (variant
    x2: the two-step kernel); (None, None) if that combination was not profiled."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None, None
    for fn in sorted(os.listdir(pdir)):
        if not fn.endswith(".md"):
            continue
        for line in open(os.path.join(pdir, fn)):
            if not line.startswith(tag):
                continue
            f = line.split()
            kv = dict(x.split("=", 1) for x in f[2:])
            if f[1] == workload and kv["variant"] == str(variant) and int(kv["planes"]) == planes:
                return float(kv[key]), "profiles/" + fn
    return None, None


def measured_traffic(workload, variant, planes):
    """HBM bytes per fused-stencil launch (FETCH_SIZE and WRITE_SIZE passes, each scaled by a
    plain-copy calibration)."""
    return profiled("bench-traffic:", "bytes", workload, variant, planes)


def cpu_baseline(p, budget_s=15.0):
    """Reference QSR.cpp step functions (oracle/_ref) on the same 256^3 box."""
    import ctypes as C

    import numpy as np

    from oracle import oracle as O
    R = O.ref_lib()
    kind = "reference"
    if R is None:
        return None
    Q = np.zeros(p.dims + (5,))
    rng = np.random.default_rng(1)
    Q[...] = rng.normal(0, 0.2, Q.shape)
    params = (C.c_double * 9)(p.a, p.b, p.c, p.T_high, p.T_star, p.L1, p.L2, p.L3, p.W)
    dp = C.POINTER(C.c_double)
    args = lambda s: (Q.ctypes.data_as(dp), p.Nx, p.Ny, p.Nz, C.c_double(p.dx), C.c_double(p.dy),
                      C.c_double(p.dz), params, C.c_double(p.kappa), 1, C.c_double(p.dt), C.c_double(p.gamma),
                      C.c_double(0.5), s)
    t1 = R.ref_time_steps(*args(1))
    steps = max(1, min(20, int(budget_s / max(t1, 1e-3))))
    t = R.ref_time_steps(*args(steps))
    cores = R.ref_max_threads()
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            model = next(ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": p.n * steps / t, "unit": "voxel-updates/s", "cores": int(cores), "kind": kind,
            "sample": f"QSR.cpp computeLaplacian + computeChemicalPotential(L1,L2,L3) + applyWeakAnchoringPenalty "
                      f"+ updateQTensor + applyBoundaryConditions, {steps} steps of the {p.Nx}^3 box "
                      f"(first step {t1:.2f} s excluded), OpenMP threads={cores} on {model}"}


def describe(p, workload):
    if workload == "periodic_noise_256":
        return (f"periodic bulk LdG quench {p.Nx}^3, kappa=6e-12, T ramp 310->295 K, Philox noise sigma="
                f"{p.noise_sigma:.4e}, dt={p.dt:.5e}")
    if workload in ("droplet_l1_64", "droplet_l1_256"):
        return (f"LdG droplet relaxation {p.Nx}^3, R={p.radius_index:g}, L1=6e-12, strong anchoring, "
                f"dt={p.dt:.5e}")
    return (f"LdG droplet relaxation {p.Nx}^3, R={p.radius_index:g}, L1/L2/L3=6e-12/3e-12/1e-12, "
            f"weak anchoring W=1e-4 custom S=0.5, explicit Euler dt={p.dt:.5e}")


def time_to_converge(q, p, args):
    """Plain relaxation from the seeded start: explicit steps with an FP64 free-energy check every
    relax_every steps, stopping once |dF/F| < relax_tol between checks or after relax_max_steps
    (the production relaxation's early-stop test, energy part).  Wall time includes the checks."""
    Q0 = q.init_field(p, 0, 0.5, 0.0, 12345)
    with q.Solver(p) as s:
        s.upload(Q0)
        del Q0
        if args.tuning:
            v, ch = (int(x) for x in args.tuning.split(","))
            s.set_tuning(v, ch)
        else:
            s.autotune(10)
        prev, rel, n, converged = None, math.nan, 0, False
        t0 = time.perf_counter()
        while n < args.relax_max_steps:
            k = min(args.relax_every, args.relax_max_steps - n)
            s.step(k)
            n += k
            F = s.observables(want_defects=False, want_xi=False, want_order=False)["total"]
            if prev is not None and prev != 0.0:
                rel = abs((F - prev) / prev)
                if rel < args.relax_tol:
                    converged = True
                    break
            prev = F
        secs = time.perf_counter() - t0
        final = s.observables()
    log(f"relaxation: {n} steps in {secs:.3f} s, converged={converged}, last |dF/F|={rel:.3e}")
    return {"steps": n, "seconds": secs, "converged": converged, "last_rel_dF": rel, "tol": args.relax_tol,
            "log_every": args.relax_every, "final_total_F": final["total"], "final_defects": final["defects"],
            "final_plaquettes": final["plaquettes"]}


def run_ladder(q, args):
    """64 droplets in one batched context (one launch per step, or per two steps where they pair, over all
    boxes) against the same droplets stepped one context at a time; value = batched voxel-updates/s over the
    voxels a step changes (the box-voxel and launched-tile rates are separate fields)."""
    params = ladder_params(q)
    fields = [q.init_field(p, 2, 0.5, 1e-3, 1000 + b) for b, p in enumerate(params)]
    total = sum(p.n for p in params)
    log(f"ladder: {len(params)} boxes, {total} voxels, dt={params[0].dt:.5e}")
    with q.Solver(params) as s:
        for b, Q in enumerate(fields):
            s.upload(Q, box=b)
        if args.tuning:
            v, ch = (int(x) for x in args.tuning.split(","))
            s.set_tuning(v, ch)
            tune = {"variant": v, "planes_per_tile": ch, "fixed": True}
        else:
            tune = s.autotune(10)
        if args.no_fusion or args.pair_planes:
            s.set_step_fusion(not args.no_fusion, args.pair_planes)
        log(f"tuning: {tune}")
        advance = (lambda n: s.step_graph(n, args.graph)) if args.graph else s.step
        advance(args.warmup)
        t0 = time.perf_counter()
        advance(args.steps)
        wall = time.perf_counter() - t0
        main_ms, face_ms = s.profile_steps(min(100, args.steps))
        work = s.work()
        fusion = s.step_fusion()
    log(f"batched: {1e3 * wall / args.steps:.4f} ms/step")
    spl = 2 if fusion["active"] else 1  # steps per stencil launch
    tile_vox = fusion["tile_voxels"] if fusion["active"] else work["tile_voxels"]
    solvers = []
    try:
        for p, Q in zip(params, fields):
            sv = q.Solver(p)
            solvers.append(sv)
            sv.upload(Q)
            sv.autotune(5)
            sv.step(args.warmup)
        t0 = time.perf_counter()
        for sv in solvers:
            sv.step(args.steps)
        wall_plain = time.perf_counter() - t0
    finally:
        for sv in solvers:
            sv.close()
    log(f"one context per droplet: {1e3 * wall_plain / args.steps:.4f} ms/step")
    active = work["active_voxels"]
    value = active * args.steps / wall
    # roofline on the algorithmic bytes of the voxels a step changes (40 B each per step), per stencil launch
    achieved = BYTES_PER_VOXEL * active * spl / (main_ms * 1e-3) / 1e9
    tile_achieved = BYTES_PER_VOXEL * tile_vox * spl / (main_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC, "value": value, "unit": "voxel-updates/s (droplet voxels)", "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded isotropic start, reference init_mode 2)",
        "config": {"workload": "quench-rate ladder: 64 droplets R=16..48 in (2R+8)^3 boxes, ramp 310->300 K over "
                               "1e3..1e5 steps, strong/weak anchoring, L1=6e-12, one batched launch per two steps",
                   "box_voxels": total, "tile_voxels": tile_vox, "active_voxels": active,
                   "box_voxel_updates_per_s": total * args.steps / wall,
                   "tile_voxel_updates_per_s": tile_vox * args.steps / wall,
                   "graph_chunk": args.graph, "autotune": tune, "steps_per_launch": spl, "step_fusion": fusion,
                   "fused_stencil_ms": main_ms, "stencil_ms_per_step": main_ms / spl, "face_ms": face_ms,
                   "plain_voxel_updates_per_s": active * args.steps / wall_plain,
                   "batched_over_plain": wall_plain / wall},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "tile_frac": tile_achieved / HBM_PEAK_GBS},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


STOP_RULE = ("energy part of QSR.cu:3322-3335: |dF/F| < tol between FP64 energy checks `every` steps apart; "
             "the defect-density and 0.5 tau_align gates are not applied")


def ladder_relax(q, args, surr_traj=None, trained=True):
    """Time to |dF/F| < relax_tol (FP64 energy of every droplet every relax_every steps) for the 64 ladder
    droplets relaxed at T = 300 K from radial starts (S0 = 0.5): one batched context that retires each droplet
    once its test passes (qsr_set_box_active), plain and warm-started (one checked learned step per active
    droplet per cycle, qsr_surr_apply_boxes; the network trained on ``surr_traj`` trajectories of the warm-start
    training set, and the zero network beside it), against one context per droplet relaxed plainly in turn.
    Returns the figures of every leg."""
    import dataclasses

    import numpy as np

    from lcdrsimulator_amd import warmstart as ws
    surr_traj = args.surr_traj if surr_traj is None else surr_traj
    params = []
    for p in ladder_params(q):
        w = q.dt_estimate(p.dx, p.dy, p.dz, p.gamma, p.a, p.b, p.c, p.T_star, 1, 0.0, p.L1, 0, 0, math.nan, 0.5,
                          300.0, 300.0)
        params.append(dataclasses.replace(p, protocol=q.PROTOCOL_CONSTANT, T_high=300.0, dt=w["dt_max"]))
    fields = [q.init_field(p, 1, 0.5, 0.0, 0) for p in params]
    n, tol, every, nmax = len(params), args.relax_tol, args.relax_every, args.relax_max_steps or 200000

    def energy(s, b=0):
        return s.observables(box=b, want_defects=False, want_xi=False, want_order=False)["total"]

    with q.Solver(params) as s:
        for b, Q in enumerate(fields):
            s.upload(Q, box=b)
        tune = s.autotune(5)
        for b, Q in enumerate(fields):
            s.upload(Q, box=b)
            s.set_iteration(0, box=b)
        rb = ws.relax_batch(s, n, None, tol, every, nmax)
        t_batched, steps_b, done, n_conv = rb["seconds"], rb["steps"], rb["total_steps"], rb["converged"]
        final_b = [s.observables(box=b, want_xi=False) for b in range(n)]
    log(f"ladder relaxation batched: {t_batched:.3f} s, {done} steps, {n_conv} of {n} converged")
    # batched and warm-started: the trained network and the zero network
    warm, tr = {}, None
    with q.Surrogate(seed=3) as net, q.Surrogate(seed=None) as zero:
        nets = []
        if trained and surr_traj > 0:
            ws.zero_output_layer(net)
            net.set_input(q.SURR_INPUT_INCREMENT, 500.0)
            tr = ws.train(net, ws.training_set(surr_traj), gap=args.surr_gap, n_snap=13, stride=250, group=16,
                          iters_per_group=args.surr_iters, lr=5e-4, increment=True, first=args.surr_first, log=log)
            if args.surr_shrink != 1.0:
                ws.scale_output_layer(net, args.surr_shrink)
            nets.append(("trained", net))
        zero.params = np.zeros(zero.nparams, dtype=np.float32)
        zero.set_input(q.SURR_INPUT_INCREMENT, 500.0)
        nets.append(("zero", zero))
        cycles = {"both": (False, True), "long": (False,), "short": (True,)}[args.warm_cycles]
        for short in cycles:
            for name, nt in nets:
                with q.Solver(params) as s:
                    for b, Q in enumerate(fields):
                        s.upload(Q, box=b)
                    s.set_tuning(tune["variant"], tune["planes_per_tile"])
                    tail = 1e-5 if args.warm_tail_tol is None else (args.warm_tail_tol or None)
                    r = ws.relax_batch(s, n, nt, tol, every, nmax, tail_tol=tail, short=short)
                    r["final"] = [s.observables(box=b, want_xi=False) for b in range(n)]
                warm[(name, short)] = r
                log(f"ladder relaxation batched + {name} network ({'short' if short else 'long'} cycles): "
                    f"{r['seconds']:.3f} s, {r['converged']} of {n} converged, {r['jumps_accepted']} steps kept, "
                    f"{r['jumps_rejected']} refused")
    t_seq, same_defects, max_rel_dF, steps_s = 0.0, 0, 0.0, []
    for b, (p, Q) in enumerate(zip(params, fields)):
        with q.Solver(p) as s1:
            s1.upload(Q)
            s1.autotune(3)
            s1.upload(Q)
            s1.set_iteration(0)
            prevF, k = None, 0
            t0 = time.perf_counter()
            while k < nmax:
                s1.step(every)
                k += every
                F = energy(s1)
                if prevF is not None and abs((F - prevF) / prevF) < tol:
                    break
                prevF = F
            t_seq += time.perf_counter() - t0
            o = s1.observables(want_xi=False)
        steps_s.append(k)
        same_defects += int(o["defects"] == final_b[b]["defects"])
        max_rel_dF = max(max_rel_dF, abs(o["total"] - final_b[b]["total"]) / abs(o["total"]))
    log(f"ladder relaxation one context per droplet: {t_seq:.3f} s")

    def worst(r):
        fin = r["final"]
        rel = [(a["total"] - c["total"]) / abs(c["total"]) for a, c in zip(fin, final_b)]
        b = int(np.argmax(np.abs(rel)))
        return {"droplet": b, "radius": params[b].radius_index, "W": params[b].W,
                "F_plain": final_b[b]["total"], "F_warm": fin[b]["total"], "rel": rel[b],
                "steps_plain": steps_b[b], "steps_warm": r["steps"][b]}

    def vs_plain(r):
        fin = r["final"]
        return {"seconds": r["seconds"], "converged": r["converged"], "steps_per_droplet": r["steps"],
                "speedup_vs_sequential": t_seq / r["seconds"], "speedup_vs_batched_plain": t_batched / r["seconds"],
                "steps_kept": r["jumps_accepted"], "steps_refused": r["jumps_rejected"],
                "network_phase_ms": r["network_ms"], "apply_s": r["apply_s"],
                "same_defects_as_batched_plain": int(sum(int(a["defects"]) == int(c["defects"])
                                                         for a, c in zip(fin, final_b))),
                "max_rel_dF_vs_batched_plain": max(abs(a["total"] - c["total"]) / abs(c["total"])
                                                   for a, c in zip(fin, final_b)),
                "defects_per_droplet": [int(a["defects"]) for a in fin],
                "worst_droplet": worst(r)}
    out = {"workload": "relaxation of the 64 quench-ladder droplets (R=16..48 in (2R+8)^3 boxes, strong/weak "
                       "anchoring, L1=6e-12) at T=300 K from radial starts to |dF/F|<" f"{tol:g} between checks every "
                       f"{every} steps", "stop_rule": STOP_RULE, "droplets": n,
           "sequential_plain_s": t_seq, "batched_plain_s": t_batched, "autotune": tune,
           "converged_batched": n_conv, "steps_per_droplet_batched": steps_b,
           "steps_per_droplet_sequential": steps_s, "total_steps_batched": done,
           "same_steps": int(sum(a == c for a, c in zip(steps_b, steps_s))),
           "same_defects": same_defects, "max_rel_dF_final": max_rel_dF,
           "defects_per_droplet_batched": [int(o["defects"]) for o in final_b],
           "batched_plain_speedup": t_seq / t_batched,
           "warm_cycles": {"long": "a checked learned step and 2 x %d plain steps, the test between the two "
                                   "energies (keys warm_trained, warm_zero_network)" % every,
                           "short": "a checked learned step and %d plain steps, the test from the energy right "
                                    "after the step (or before it, when refused) to the energy %d plain steps "
                                    "later, no convergence in a cycle whose step was kept (keys ending in "
                                    "_short_cycles)" % (every, every)}}
    for (name, short), r in warm.items():
        key = ("warm_trained" if name == "trained" else "warm_zero_network") + ("_short_cycles" if short else "")
        out[key] = vs_plain(r)
    legs = [k for k in out if k.startswith("warm_") and isinstance(out[k], dict) and "seconds" in out[k]]
    if legs:
        out["fastest_warm_leg"] = min(legs, key=lambda k: out[k]["seconds"])
    if tr is not None:
        out["training"] = {k: v for k, v in tr.items() if k != "loss"}
    return out


def run_ladder_relax(q, args):
    """The ladder time-to-solution legs as their own line: value = sequential plain time / batched and
    warm-started (trained network) time."""
    r = ladder_relax(q, args)
    lead = r[r["fastest_warm_leg"]]
    out = {
        "metric": METRIC, "value": lead["speedup_vs_sequential"],
        "unit": "x (time to |dF/F|<1e-6: one context per droplet, plain / batched and warm-started, the fastest "
                "warm-started leg, named in config.fastest_warm_leg)",
        "n_gpus": 1, "steps": r["total_steps_batched"], "warmup": 0,
        "ms_per_step": 1e3 * r["batched_plain_s"] / max(r["total_steps_batched"], 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (radial start S0 = 0.5 per droplet)", "config": r,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def run_periodic_seeds(q, args):
    """Eight seeds of the periodic noisy quench (256^3 each, T 310 -> 295 K, Philox noise) in one batched
    context, one launch per step over all boxes, with the FP64 log (energies, 3-plane defect count, xi) of
    every box every 100 steps inside the timed region; value = voxel-updates/s over the 8 boxes."""
    import numpy as np
    base = workload_params(q, args.size, args.radius, "periodic_noise_256")
    params = []
    for seed in range(8):
        p = workload_params(q, args.size, args.radius, "periodic_noise_256")
        p.seed = seed
        params.append(p)
    total = sum(p.n for p in params)
    log(f"periodic seeds: 8 boxes of {base.dims}, {total} voxels, dt={base.dt:.5e}")
    with q.Solver(params) as s:
        for b in range(8):
            s.upload(np.random.default_rng(100 + b).normal(0.0, 1e-3, base.dims + (5,)), box=b)
        if args.tuning:
            v, ch = (int(x) for x in args.tuning.split(","))
            s.set_tuning(v, ch)
            tune = {"variant": v, "planes_per_tile": ch, "fixed": True}
        else:
            tune = s.autotune(10)
        log(f"tuning: {tune}")
        s.step(args.warmup)
        every = 100
        n_logs = max(1, args.steps // every)
        t0 = time.perf_counter()
        logs = []
        for _ in range(n_logs):
            s.step(every)
            logs.append([s.observables(box=b) for b in range(8)])
        wall = time.perf_counter() - t0
        t1 = time.perf_counter()
        s.step(n_logs * every)
        wall_steps = time.perf_counter() - t1
        main_ms, _ = s.profile_steps(50)
        work = s.work()
    steps = n_logs * every
    value = total * steps / wall
    achieved = BYTES_PER_VOXEL * work["tile_voxels"] / (main_ms * 1e-3) / 1e9
    last = logs[-1]
    log(f"8 seeds: {1e3 * wall / steps:.4f} ms/step with logs, {1e3 * wall_steps / steps:.4f} ms/step without")
    out = {
        "metric": METRIC, "value": value, "unit": "voxel-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded isotropic noise start per box)",
        "config": {"workload": describe(base, "periodic_noise_256") + ", 8 seeds batched, FP64 log every 100 steps",
                   "box_voxels": total, "tile_voxels": work["tile_voxels"], "autotune": tune,
                   "fused_stencil_ms": main_ms, "ms_per_step_without_logs": 1e3 * wall_steps / steps,
                   "log_overhead_frac": (wall - wall_steps) / wall,
                   "final_T_step": steps + args.warmup,
                   "final_defects_per_seed": [int(o["defects"]) for o in last],
                   "final_total_F_per_seed": [o["total"] for o in last]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def run_network_pass(q, args):
    """The learned step's network pass alone (pack + three convolutions + apply) on the 192^3, R = 88
    droplet, untrained weights: device time per pass from hipEvents on the context stream."""
    import numpy as np

    from lcdrsimulator_amd import warmstart as ws
    p = ws.droplet_params(88.0, 300.0, 1e-4, N=192)
    with q.Solver(p) as s, q.Surrogate(seed=3) as net:
        s.upload(q.init_field(p, 1, 0.5, 0.0, 12345))
        for _ in range(max(1, args.warmup)):
            net.apply(s, q.SURR_PROPOSE)
        rs = [net.apply(s, q.SURR_PROPOSE) for _ in range(args.steps)]
    net_ms = float(np.median([r["ms_network"] for r in rs]))
    nvox = int(rs[0]["net_voxels"])
    # matrix-core rate over the voxels the pass computes (its tiles: those within the network's reach
    # of a voxel a step can change)
    achieved = NET_FLOP_PER_VOXEL * nvox / (net_ms * 1e-3) / 1e12
    log(f"network pass at 192^3: median {net_ms:.4f} ms over {args.steps} passes, {nvox} voxels computed")
    out = {
        "metric": METRIC, "value": p.n / (net_ms * 1e-3),
        "unit": "box voxels/s through the learned step's network (every voxel a step can change gets its output)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": net_ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (fp32 accumulate)",
        "data": "synthetic (seeded radial droplet, untrained weights)",
        "config": {"workload": "network pass of the learned step: 5-32-32-5 channels, 3x3x3, 192^3 box, R=88",
                   "flop_per_voxel": NET_FLOP_PER_VOXEL, "box_voxels": p.n, "computed_voxels": nvox},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / MFMA_BF16_PEAK_TFLOPS, "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def run_surrogate(q, args):
    """Learned warm start (192^3, R = 88, T = 300 K, W = 1e-4, L1/L2/L3): train the modulating network on
    plain-solver trajectories (increment input, horizon 500 steps), then relax the same start plainly and
    warm-started; value = plain time / warm time to |dF/F| < 1e-6 between checks 100 steps apart (radial
    start; the random-director start and the zero network, i.e. the plain step extrapolated, are reported
    too)."""
    import numpy as np

    from lcdrsimulator_amd import warmstart as ws
    if args.surr_net_only:
        return run_network_pass(q, args)
    trajs = ws.training_set(args.surr_traj, T=300.0 if args.surr_matched else None,
                            W=1e-4 if args.surr_matched else None)
    horizon = 500
    with q.Surrogate(seed=3) as net, q.Surrogate(seed=None) as zero:
        ws.zero_output_layer(net)
        net.set_input(q.SURR_INPUT_INCREMENT, float(horizon))
        tr = ws.train(net, trajs, gap=args.surr_gap, n_snap=13, stride=250, group=16, iters_per_group=args.surr_iters,
                      lr=5e-4, increment=True, first=args.surr_first, log=log)
        if args.surr_shrink != 1.0:
            ws.scale_output_layer(net, args.surr_shrink)
        zero.params = np.zeros(zero.nparams, dtype=np.float32)
        zero.set_input(q.SURR_INPUT_INCREMENT, float(horizon))
        p = ws.droplet_params(88.0, 300.0, 1e-4, N=192)
        tail = args.warm_tail_tol or None  # off unless asked for
        res = {m: ws.compare(p, net, m, max_steps=200000, tail_tol=tail, log=log) for m in (1, 0)}
        res_zero = {m: ws.compare(p, zero, m, max_steps=200000, tail_tol=tail, log=log) for m in (1, 0)}
        with q.Solver(p) as s:
            s.upload(q.init_field(p, 1, 0.5, 0.0, 12345))
            s.step(10)
            rs = [net.apply(s, q.SURR_PROPOSE) for _ in range(5)]
    net_ms = float(np.median([r["ms_network"] for r in rs]))
    achieved = NET_FLOP_PER_VOXEL * int(rs[0]["net_voxels"]) / (net_ms * 1e-3) / 1e12
    r1 = res[1]

    def summary(r):
        return {"plain_s": r["plain"]["seconds"], "warm_s": r["warm"]["seconds"], "speedup": r["speedup"],
                "plain_steps": r["plain"]["steps"], "warm_plain_steps": r["warm"]["steps"],
                "jumps_accepted": r["warm"]["jumps_accepted"], "jumps_rejected": r["warm"]["jumps_rejected"],
                "rel_dF_final": r["rel_dF_final"], "defects_plain": r["plain"]["defects"],
                "defects_warm": r["warm"]["defects"], "same_defects": r["same_defects"]}
    out = {
        "metric": METRIC, "value": r1["speedup"], "unit": "x (time to |dF/F|<1e-6: plain / warm-started)",
        "n_gpus": 1, "steps": r1["plain"]["steps"], "warmup": 0, "ms_per_step": 1e3 * r1["plain"]["seconds"] /
        r1["plain"]["steps"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 step, bf16 network (fp32 accumulate)",
        "data": "synthetic (seeded radial and random-director droplets; training triples from the plain solver)",
        "config": {"workload": "learned warm start: 192^3, R=88, T=300 K, W=1e-4 custom S=0.5, L1/L2/L3=6e-12/3e-12/"
                               "1e-12; network modulating the explicit increment, trained on (Q_k, u_k, Q_k+500) "
                               "of R 40-120, " +
                               ("T 300 K, W 1e-4 (matched to the application)" if args.surr_matched else
                                "T 295-305 K, W 1e-5..1e-3"),
                   "training": {k: v for k, v in tr.items() if k != "loss"},
                   "train_loss_first_last": [float(tr["loss"][:100].mean()), float(tr["loss"][-100:].mean())],
                   "radial_start": summary(r1), "random_director_start": summary(res[0]),
                   "zero_network": {"radial_start": summary(res_zero[1]),
                                    "random_director_start": summary(res_zero[0])},
                   "network_ms_192": net_ms},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / MFMA_BF16_PEAK_TFLOPS, "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--radius", type=float, default=120.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--graph", type=int, default=0, help="replay steps from a hipGraph of this many steps")
    ap.add_argument("--tuning", default="", metavar="VARIANT,PLANES",
                    help="fixed stencil variant and planes per tile instead of autotuning (profiling runs)")
    ap.add_argument("--relax-max-steps", type=int, default=20000,
                    help="time-to-solution leg (headline workload): relax from the start state until the free "
                         "energy changes by less than --relax-tol between log points, at most this many steps "
                         "(0 skips the leg)")
    ap.add_argument("--tts-traj", type=int, default=16,
                    help="headline line: run the ladder time-to-solution legs (batched and warm-started vs one "
                         "plain context per droplet) with a network trained on this many trajectories "
                         "(0: zero network only; -1 skips the legs)")
    ap.add_argument("--relax-every", type=int, default=100)
    ap.add_argument("--relax-tol", type=float, default=1e-6)
    ap.add_argument("--imex", type=int, default=0, metavar="J",
                    help="time the semi-implicit quench step (RHS + J Jacobi sweeps, L_stab = L1 or kappa) "
                         "instead of the explicit step")
    ap.add_argument("--surr-traj", type=int, default=256, help="surrogate_192: training trajectories")
    ap.add_argument("--surr-net-only", action="store_true",
                    help="surrogate_192: time only the network pass (--steps passes after --warmup)")
    ap.add_argument("--surr-matched", action="store_true",
                    help="surrogate_192: train at the application's T and W instead of their ranges")
    ap.add_argument("--warm-tail-tol", type=float, default=None,
                    help="ladder_relax_64 and surrogate_192: no learned step in a cycle after one whose |dF/F| was "
                         "below this, so the plain solver makes the last approach (0: off; default 1e-5 for the "
                         "ladder's radial starts, off for surrogate_192, whose random-director start coarsens through "
                         "long stretches of small |dF/F| where the learned steps are what speeds it up)")
    ap.add_argument("--warm-cycles", choices=("both", "long", "short"), default="both",
                    help="ladder legs: warm-started cycles of a learned step and 2 x relax_every plain steps with "
                         "the test between the two energies (long), of a learned step and relax_every plain steps "
                         "tested from the energy right after the step, never converging in a cycle whose step was "
                         "kept (short), or both, each with the trained and the zero network")
    ap.add_argument("--surr-first", type=int, default=500,
                    help="training snapshots start this many plain steps into each trajectory (from step 0, the "
                         "network's first jumps from radial starts reshape the droplet cores)")
    ap.add_argument("--surr-gap", type=int, default=500,
                    help="training target Q(k + gap) in plain steps (a multiple of 250); the proposal stays "
                         "500 u (1 + y), so a longer gap teaches the network to jump further than 500 steps go")
    ap.add_argument("--surr-shrink", type=float, default=1.0,
                    help="scale the trained network's output layer by this factor before use (0: the zero network)")
    ap.add_argument("--surr-iters", type=int, default=1000,
                    help="surrogate_192: Adam iterations per group of 16 trajectories")
    ap.add_argument("--no-fusion", action="store_true",
                    help="run every step on its own (no two-step launches for L1-only contexts)")
    ap.add_argument("--pair-planes", type=int, default=0,
                    help="planes per tile of the two-step kernel (default: autotuned, or 16 with --tuning)")
    ap.add_argument("--workload", default=WORKLOADS[0], choices=WORKLOADS,
                    help="droplet_l1l2l3_256 is the headline line; the others are the periodic noisy quench, "
                         "the 64^3 L1 droplet, the 64-droplet quench ladder and the learned warm start")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU slab decomposition is a design note only)", file=sys.stderr)
        return 2
    import numpy as np

    import lcdrsimulator_amd as q
    q.load()
    if q.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    if args.workload == "ladder_64":
        return run_ladder(q, args)
    if args.workload == "surrogate_192":
        return run_surrogate(q, args)
    if args.workload == "periodic_noise_256x8":
        return run_periodic_seeds(q, args)
    if args.workload == "ladder_relax_64":
        return run_ladder_relax(q, args)
    p = workload_params(q, args.size, args.radius, args.workload)
    log(f"box {p.dims} R={p.radius_index} dt={p.dt:.5e}")
    if p.periodic:
        Q0 = np.random.default_rng(7).normal(0.0, 1e-3, p.dims + (5,))
    else:
        Q0 = q.init_field(p, 0, 0.5, 0.0, 12345)
    with q.Solver(p) as s:
        s.upload(Q0)
        del Q0
        if args.tuning:
            v, ch = (int(x) for x in args.tuning.split(","))
            s.set_tuning(v, ch)
            tune = {"variant": v, "planes_per_tile": ch, "fixed": True}
        else:
            tune = s.autotune(10)
        if args.no_fusion or args.pair_planes:
            s.set_step_fusion(not args.no_fusion, args.pair_planes)
        log(f"tuning: {tune}")
        L_stab = p.L1 if p.L1 else p.kappa

        def advance(n):
            if args.imex:
                s.step_imex(n, L_stab, args.imex)
            elif args.graph:
                s.step_graph(n, args.graph)
            else:
                s.step(n)

        advance(args.warmup)
        log(f"warmup {args.warmup} steps done")
        t0 = time.perf_counter()
        advance(args.steps)
        wall = time.perf_counter() - t0
        ev_ms = s.last_step_ms()
        log(f"timed {args.steps} steps: wall {wall:.3f} s, event {ev_ms:.4f} ms/step")
        main_ms, face_ms = s.profile_steps(min(100, args.steps))
        work = s.work()
        fusion = s.step_fusion()
        obs = s.observables()
        log(f"stencil launch {main_ms:.4f} ms ({'two steps' if fusion['active'] else 'one step'}), faces "
            f"{face_ms:.4f} ms; F={obs['total']:.6e}")
    # a context that pairs its steps runs the two-step kernel: one launch = two steps over its tiles
    spl = 2 if fusion["active"] else 1
    kern = ("x2", fusion["planes_per_tile"]) if fusion["active"] else (int(tune["variant"]), int(tune["planes_per_tile"]))
    tile_vox = fusion["tile_voxels"] if fusion["active"] else work["tile_voxels"]
    relax = None
    if args.workload == WORKLOADS[0] and args.relax_max_steps > 0 and not args.imex:
        relax = time_to_converge(q, p, args)
    tts = None
    if args.workload == WORKLOADS[0] and args.tts_traj >= 0 and not args.imex:
        tts = ladder_relax(q, args, surr_traj=args.tts_traj)
    ms_per_step = 1e3 * wall / args.steps
    # one voxel update = one droplet voxel stepped (what the reference's update kernel changes per step)
    value = work["active_voxels"] * args.steps / wall
    # algorithmic bytes per launch: 40 B per droplet voxel and step, times the steps of a launch
    algo_bytes = BYTES_PER_VOXEL * work["active_voxels"] * spl
    tile_bytes = BYTES_PER_VOXEL * tile_vox * spl
    achieved = algo_bytes / (main_ms * 1e-3) / 1e9
    hbm_bytes, hbm_src = measured_traffic(args.workload, *kern)
    valu_busy, valu_src = profiled("bench-valu:", "busy", args.workload, *kern)
    flop_vox, flop_src = profiled("bench-flop:", "per_tile_voxel", args.workload, *kern)
    compute = None
    if flop_vox is not None:
        tflops = flop_vox * work["tile_voxels"] / (main_ms * 1e-3) / 1e12
        compute = {"bound": "valu", "achieved": tflops, "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                   "frac": tflops / FP32_VALU_PEAK_TFLOPS, "flop_per_tile_voxel": flop_vox,
                   "flop_per_droplet_voxel": flop_vox * work["tile_voxels"] / work["active_voxels"],
                   "source": flop_src + " (executed FP32 lane operations, FMA = 2; peak = non-packed FP32 issue)"}
    cpu = None if (args.no_cpu or args.workload != WORKLOADS[0]) else cpu_baseline(p)
    if relax is not None and cpu is not None:
        # the CPU figure is a rate on a bounded sample; the full relaxation is not run on the CPU
        relax["cpu_seconds_estimate"] = relax["steps"] * work["box_voxels"] / cpu["value"]
    out = {
        "metric": METRIC, "value": value, "unit": "voxel-updates/s (droplet voxels)", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded random-director droplet, reference init_mode 0)",
        "config": {"workload": describe(p, args.workload) + (f", semi-implicit J={args.imex}" if args.imex else ""),
                   "box_voxels": work["box_voxels"], "active_voxels": work["active_voxels"],
                   "tile_voxels": tile_vox, "graph_chunk": args.graph, "autotune": tune,
                   "steps_per_launch": spl, "step_fusion": fusion,
                   "box_voxel_updates_per_s": work["box_voxels"] * args.steps / wall,
                   "tile_voxel_updates_per_s": tile_vox * args.steps / wall,
                   "fused_stencil_ms": main_ms, "stencil_ms_per_step": main_ms / spl, "face_ms": face_ms,
                   "event_ms_per_step": ev_ms,
                   "final_total_F": obs["total"], "algorithmic_bytes_per_launch": algo_bytes,
                   "tile_bytes_per_launch": tile_bytes,
                   "valu_busy": None if valu_busy is None else {"frac": valu_busy, "source": valu_src},
                   "relaxation": relax, "time_to_solution": tts},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": hbm_bytes,
                     "traffic_source": None if hbm_src is None else
                     hbm_src + " (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, same workload and tuning)",
                     "tile_frac": tile_bytes / (main_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "compute": compute},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
