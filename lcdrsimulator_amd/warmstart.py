"""Learned warm start for the LdG droplet relaxation.

No counterpart in the reference: QSR.cu relaxes with the plain loop only
(sim_mode 1, QSR.cu:3126-3476; early stop on |dF/F| between log points,
QSR.cu:3322-3335).  Here that loop stays both the fallback and the arbiter, and
a learned step (:class:`lcdrsimulator_amd.Surrogate`) may be interleaved:

* training data: trajectories of the plain solver itself, droplets of radius
  R in [40, 120] voxels in (2R+8)^3 boxes, T in [295, 305] K, Rapini-Papoular
  W log-uniform in [1e-5, 1e-3] J/m^2, half random-director and half radial
  starts (S0 = 0.5), snapshots every ``gap`` steps; pairs (Q_k, Q_{k+gap});
* training: the network's own forward / backward / Adam kernels on random
  crops of the pairs;
* application (:func:`relax_warm`): cycles of a checked learned step (kept
  only if neither the free energy nor the explicit-Euler residual rises; a
  ladder of step fractions is tried in turn) and plain steps, with the
  reference's energy test |dF/F| < tol between two log points ``every`` steps
  apart that no jump separates.

Both relaxations report wall time on the same context, the final energy and
the mid-plane defect count, so the bench can state what the network gains.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import qsr as q

ELASTIC = (6e-12, 3e-12, 1e-12)  # L1, L2, L3 (N)


def droplet_params(R: float, T: float, W: float, N: int | None = None, elastic=ELASTIC) -> q.Params:
    """A droplet of radius R voxels (box N^3, default 2R+8) at temperature T, weak anchoring W with
    shell order 0.5, dx = 10 nm, gamma = 0.1, dt = the stability window's dt_max."""
    N = N or int(2 * R + 8)
    p = q.Params(Nx=N, Ny=N, Nz=N, dx=1e-8, dy=1e-8, dz=1e-8, T_star=307.0, T_high=T, kappa=0.0,
                 L1=elastic[0], L2=elastic[1], L3=elastic[2], W=W, radius_index=float(R), boundary_custom=1,
                 boundary_S=0.5, gamma=0.1)
    w = q.dt_estimate(p.dx, p.dy, p.dz, p.gamma, p.a, p.b, p.c, p.T_star, 1, 0.0, *elastic, math.nan, 0.5, T, T)
    p.dt = w["dt_max"]
    return p


def training_set(n_traj: int = 256, seed: int = 2025, T: float | None = None, W: float | None = None):
    """(params, init mode, init seed) of the training trajectories; T and W are drawn from their
    ranges unless fixed (the network sees only Q, not the temperature or the anchoring)."""
    rng = np.random.default_rng(seed)
    out = []
    for t in range(n_traj):
        R = float(rng.integers(40, 121))
        Tt = float(rng.uniform(295.0, 305.0))
        Wt = float(10.0 ** rng.uniform(-5.0, -3.0))
        out.append((droplet_params(R, Tt if T is None else T, Wt if W is None else W), t % 2, 1000 + t))
    return out


def snapshot_pairs(n_snap: int, stride: int, gap: int):
    """Snapshot steps and the index pairs (m, m') of the pairs (Q_k, Q_{k+gap}), k = 0, stride,
    ... (n_snap input times): returns (steps, input flags, pairs).  Where gap is a multiple of
    stride, the inputs are their own targets' inputs (n_snap steps in all)."""
    ks = [m * stride for m in range(n_snap)]
    if gap % stride == 0:
        d = gap // stride
        ks = ks[:max(n_snap, d + 1)]
        steps = ks
        pairs = [(m, m + d) for m in range(n_snap - d)]
        inputs = [m < n_snap - d for m in range(len(steps))]
        return steps, inputs, pairs
    if gap > stride:
        raise ValueError("gap must be a multiple of the stride or shorter than it")
    steps = sorted(set(ks) | {k + gap for k in ks})
    idx = {t: i for i, t in enumerate(steps)}
    pairs = [(idx[k], idx[k + gap]) for k in ks]
    inputs = [t in set(ks) for t in steps]
    return steps, inputs, pairs


def train(net: q.Surrogate, trajectories, gap: int = 500, n_snap: int = 7, group: int = 32,
          iters_per_group: int = 1000, lr: float = 3e-4, batch: int = 8, crop: int = 32, weight_floor: float = 1.0,
          stride: int | None = None, increment: bool = False, first: int = 0, log=None) -> dict:
    """Generate (Q_k, Q_{k+gap}) pairs with the plain solver, ``group`` trajectories per batched
    context (snapshots every ``stride`` steps, default ``gap``; with ``increment`` each also records
    one explicit step's increment, the network's input under increment input), and train on each
    group's pairs (the dataset is cleared between groups, so device memory holds one group's
    snapshots).  With ``first`` the snapshots start that many plain steps into each trajectory.
    Returns timings and the loss history."""
    stride = stride or gap
    t_gen = t_train = 0.0
    hist = []
    n_pairs = 0
    for g0 in range(0, len(trajectories), group):
        chunk = trajectories[g0:g0 + group]
        ta = time.perf_counter()
        with q.Solver([p for p, _, _ in chunk]) as s:
            for b, (p, mode, seed) in enumerate(chunk):
                s.upload(q.init_field(p, mode, 0.5, 0.0, seed), box=b)
            steps, inputs, idx_pairs = snapshot_pairs(n_snap, stride, gap)
            steps = [first + k for k in steps]
            ids, t = [], 0
            for t_next, is_input in zip(steps, inputs):
                if t_next > t:
                    s.step(t_next - t)
                    t = t_next
                ids.append([net.dataset_add(s, box=b, increment=increment and is_input) for b in range(len(chunk))])
        pairs = [(ids[m][b], ids[m2][b]) for b in range(len(chunk)) for m, m2 in idx_pairs]
        tb = time.perf_counter()
        h = net.train(pairs, iters=iters_per_group, batch=batch, crop=crop, lr=lr, seed=g0,
                      weight_floor=weight_floor)
        net.dataset_clear()
        tc = time.perf_counter()
        t_gen += tb - ta
        t_train += tc - tb
        n_pairs += len(pairs)
        hist.append(h)
        if log:
            log(f"warm start training: group {g0 // group}: {len(chunk)} trajectories, {len(pairs)} pairs, "
                f"generate {tb - ta:.1f} s, train {tc - tb:.1f} s, loss {h[:50].mean():.3e} -> {h[-50:].mean():.3e}")
    return {"trajectories": len(trajectories), "pairs": n_pairs, "generate_s": t_gen, "train_s": t_train,
            "loss": np.concatenate(hist) if hist else np.zeros(0)}


def _energy(s: q.Solver) -> float:
    return s.observables(want_defects=False, want_xi=False, want_order=False)["total"]


def relax_plain(s: q.Solver, tol: float = 1e-6, every: int = 100, max_steps: int = 200000) -> dict:
    """Plain explicit steps with the free energy checked every ``every`` steps, until |dF/F| < tol
    between two checks (the energy part of QSR.cu:3322-3335) or max_steps."""
    n, prev, rel, conv = 0, None, math.nan, False
    t0 = time.perf_counter()
    while n < max_steps:
        s.step(every)
        n += every
        F = _energy(s)
        if prev is not None:
            rel = abs((F - prev) / prev)
            if rel < tol:
                conv = True
                break
        prev = F
    secs = time.perf_counter() - t0
    return {"steps": n, "seconds": secs, "converged": conv, "last_rel_dF": rel}


def relax_warm(s: q.Solver, net: q.Surrogate, tol: float = 1e-6, every: int = 100, max_steps: int = 200000,
               scales=(1.0, 0.5, 0.25), tail_tol: float | None = None, hold: int = 0, short: bool = False) -> dict:
    """Cycles of one checked learned step (the ladder ``scales``: the first proposal that
    raises neither F nor the residual is kept, else the field is left as it was) and 2 x ``every``
    plain steps with the energy checked after each half; converged when |dF/F| < tol between
    the two checks of a cycle, which no jump separates (the energy part of QSR.cu:3322-3335).
    With ``tail_tol``, a cycle whose predecessor's |dF/F| was below it takes no learned step (the
    plain solver alone makes the last approach; a plateau of a coarsening run only pauses them).
    No learned step is tried in the first ``hold`` plain steps.  With ``short`` a cycle is the
    learned step and ``every`` plain steps: the test interval runs from the energy after the learned
    step (the check's F_after, or F_before when it was refused) to the energy ``every`` steps later,
    and the run converges only in a cycle whose learned step was not kept.  That halves the plain
    steps per learned step; on coarsening (random-director) starts it also speeds up the coarsening
    itself, so there the defect count moves away from the plain run's.  ``steps`` counts plain
    steps only."""
    if short:
        return _relax_warm_short(s, net, tol, every, max_steps, scales, tail_tol, hold)
    n, att, acc, net_ms = 0, 0, 0, 0.0
    kept = []
    conv, rel = False, math.nan
    t0 = time.perf_counter()
    tail = False
    while n < max_steps:
        if not tail and n >= hold:
            r = net.apply(s, q.SURR_CHECKED, scales=list(scales))
            att += 1
            net_ms += r["ms_network"]
            if r["accepted"]:
                acc += 1
                kept.append(r["scale"])
        s.step(every)
        Fa = _energy(s)
        s.step(every)
        Fb = _energy(s)
        n += 2 * every
        rel = abs((Fb - Fa) / Fa)
        if rel < tol:
            conv = True
            break
        tail = tail_tol is not None and rel < tail_tol
    return {"steps": n, "seconds": time.perf_counter() - t0, "converged": conv, "last_rel_dF": rel,
            "jumps_accepted": acc, "jumps_rejected": att - acc, "scales_kept": kept, "network_ms": net_ms}


def _relax_warm_short(s, net, tol, every, max_steps, scales, tail_tol, hold) -> dict:
    n, att, acc, net_ms = 0, 0, 0, 0.0
    kept_scales = []
    conv, rel = False, math.nan
    t0 = time.perf_counter()
    tail = False
    prev = s.energies([0])[0]
    while n < max_steps:
        kept = False
        F0 = prev
        if not tail and n >= hold:
            r = net.apply(s, q.SURR_CHECKED, scales=list(scales))
            att += 1
            net_ms += r["ms_network"]
            if r["accepted"]:
                acc += 1
                kept = True
                kept_scales.append(r["scale"])
                F0 = r["F_after"]
        s.step(every)
        n += every
        F = s.energies([0])[0]
        rel = abs((F - F0) / F0)
        prev = F
        if rel < tol and not kept:
            conv = True
            break
        tail = tail_tol is not None and rel < tail_tol
    return {"steps": n, "seconds": time.perf_counter() - t0, "converged": conv, "last_rel_dF": rel,
            "jumps_accepted": acc, "jumps_rejected": att - acc, "scales_kept": kept_scales, "network_ms": net_ms}


def relax_batch(s: q.Solver, n_boxes: int, net: q.Surrogate | None = None, tol: float = 1e-6, every: int = 100,
                max_steps: int = 200000, scales=(1.0, 0.5, 0.25), backoff: bool = True,
                tail_tol: float | None = None, hold: int = 0, short: bool = False) -> dict:
    """Relax every box of a batched solver, retiring each (qsr_set_box_active) once its energy
    test passes.  Plain (no ``net``): the test of :func:`relax_plain` per box, checks every
    ``every`` steps.  Warm-started: the cycles of :func:`relax_warm`, with one checked learned step
    per active box per cycle (:meth:`Surrogate.apply_boxes`); with ``backoff`` a box whose step was
    refused sits out the next 1, 2, 4, ... cycles (reset when one is kept).  With ``tail_tol`` a box
    whose |dF/F| over the last cycle's test interval was below it takes no learned step in the next
    cycle: the plain solver alone makes the last approach to the test.  No learned step is tried in
    the first ``hold`` plain steps.  With ``short`` a cycle is the learned step and ``every`` plain
    steps (see :func:`relax_warm`).  ``steps`` per box counts plain steps."""
    if net is not None and short:
        return _relax_batch_short(s, n_boxes, net, tol, every, max_steps, scales, backoff, tail_tol, hold)
    active = list(range(n_boxes))
    steps = [0] * n_boxes
    prev = [None] * n_boxes
    wait, gap = [0] * n_boxes, [0] * n_boxes
    tail = [False] * n_boxes
    done, acc, att, net_ms, apply_s = 0, 0, 0, 0.0, 0.0
    t0 = time.perf_counter()

    def energies(boxes):
        return dict(zip(boxes, s.energies(boxes)))

    while active and done < max_steps:
        if net is None:
            s.step(every)
            done += every
            F = energies(active)
        else:
            sel = [b for b in active if wait[b] == 0 and not tail[b]] if done >= hold else []
            for b in active:
                wait[b] = max(0, wait[b] - 1)
            ta = time.perf_counter()
            res = net.apply_boxes(s, sel, q.SURR_CHECKED, scales) if sel else []
            apply_s += time.perf_counter() - ta
            for b, r in zip(sel, res):
                if r["accepted"]:
                    gap[b] = 0
                elif backoff:
                    gap[b] = max(1, 2 * gap[b])
                    wait[b] = gap[b]
            att += len(res)
            acc += sum(r["accepted"] for r in res)
            net_ms += res[0]["ms_network"] if res else 0.0  # each result holds the whole phase's time
            s.step(every)
            Fa = energies(active)
            s.step(every)
            done += 2 * every
            F = energies(active)
        still, done_now = [], []
        for b in active:
            ref = prev[b] if net is None else Fa[b]
            if net is not None and tail_tol is not None:
                tail[b] = abs((F[b] - ref) / ref) < tail_tol
            if ref is not None and abs((F[b] - ref) / ref) < tol:
                steps[b] = done
                done_now.append(b)
            else:
                still.append(b)
            prev[b] = F[b]
        if done_now:
            s.set_active_many(done_now, False)
        active = still
    return {"seconds": time.perf_counter() - t0, "steps": steps, "total_steps": done,
            "converged": n_boxes - len(active), "jumps_accepted": acc, "jumps_rejected": att - acc,
            "network_ms": net_ms, "apply_s": apply_s}


def _relax_batch_short(s, n_boxes, net, tol, every, max_steps, scales, backoff, tail_tol, hold) -> dict:
    """relax_batch's short cycles: one checked learned step per selected box, then ``every`` plain
    steps; a box's test interval runs from its energy after the learned step (the check's F_after,
    or F_before when the step was refused or not tried) to its energy ``every`` plain steps later,
    and a box converges only in a cycle whose learned step was not kept."""
    active = list(range(n_boxes))
    steps = [0] * n_boxes
    wait, gap = [0] * n_boxes, [0] * n_boxes
    tail = [False] * n_boxes
    done, acc, att, net_ms, apply_s = 0, 0, 0, 0.0, 0.0
    t0 = time.perf_counter()
    prev = dict(zip(active, s.energies(active)))
    while active and done < max_steps:
        sel = [b for b in active if wait[b] == 0 and not tail[b]] if done >= hold else []
        for b in active:
            wait[b] = max(0, wait[b] - 1)
        ta = time.perf_counter()
        res = net.apply_boxes(s, sel, q.SURR_CHECKED, scales) if sel else []
        apply_s += time.perf_counter() - ta
        F0 = {b: prev[b] for b in active}
        kept = set()
        for b, r in zip(sel, res):
            if r["accepted"]:
                gap[b] = 0
                F0[b] = r["F_after"]
                kept.add(b)
            else:
                F0[b] = r["F_before"]
                if backoff:
                    gap[b] = max(1, 2 * gap[b])
                    wait[b] = gap[b]
        att += len(res)
        acc += len(kept)
        net_ms += res[0]["ms_network"] if res else 0.0
        s.step(every)
        done += every
        F = dict(zip(active, s.energies(active)))
        still, done_now = [], []
        for b in active:
            rel = abs((F[b] - F0[b]) / F0[b])
            if tail_tol is not None:
                tail[b] = rel < tail_tol
            if rel < tol and b not in kept:
                steps[b] = done
                done_now.append(b)
            else:
                still.append(b)
            prev[b] = F[b]
        if done_now:
            s.set_active_many(done_now, False)
        active = still
    return {"seconds": time.perf_counter() - t0, "steps": steps, "total_steps": done,
            "converged": n_boxes - len(active), "jumps_accepted": acc, "jumps_rejected": att - acc,
            "network_ms": net_ms, "apply_s": apply_s}


def zero_output_layer(net: q.Surrogate):
    """Zero the last layer's weights and bias (W3, b3: the tail of the flat parameter array), so
    that a modulating network starts as the plain step extrapolated and training moves it from
    there."""
    p = net.params
    p[-(27 * 32 * 5 + 5):] = 0.0
    net.params = p


def scale_output_layer(net: q.Surrogate, alpha: float):
    """Scale the last layer's weights and bias by alpha, so the modulation y becomes alpha * y:
    alpha = 0 is the zero network (the plain step extrapolated), alpha = 1 the trained one."""
    p = net.params
    p[-(27 * 32 * 5 + 5):] *= float(alpha)
    net.params = p


def compare(p: q.Params, net: q.Surrogate, init_mode: int, seed: int = 12345, tol: float = 1e-6, every: int = 100,
            max_steps: int = 200000, scales=(1.0, 0.5, 0.25), tail_tol: float | None = None, hold: int = 0,
            log=None, short: bool = False) -> dict:
    """Plain relaxation and warm-started relaxation of the same start on one context."""
    Q0 = q.init_field(p, init_mode, 0.5, 0.0, seed)
    out = {}
    with q.Solver(p) as s:
        s.upload(Q0)
        s.autotune(5)
        for name in ("plain", "warm"):
            s.upload(Q0)
            s.set_iteration(0)
            if name == "plain":
                r = relax_plain(s, tol, every, max_steps)
            else:
                r = relax_warm(s, net, tol, every, max_steps, scales, tail_tol, hold, short)
            o = s.observables()
            r.update({"final_F": o["total"], "defects": int(o["defects"]), "plaquettes": int(o["plaquettes"])})
            out[name] = r
            if log:
                log(f"warm start comparison, init mode {init_mode}, {name}: {r['steps']} steps, {r['seconds']:.3f} s, "
                    f"F {o['total']:.9e}, defects {o['defects']}")
    pl, wa = out["plain"], out["warm"]
    out["speedup"] = pl["seconds"] / wa["seconds"] if wa["seconds"] > 0 else math.nan
    out["rel_dF_final"] = abs(wa["final_F"] - pl["final_F"]) / abs(pl["final_F"])
    out["same_defects"] = wa["defects"] == pl["defects"]
    return out
