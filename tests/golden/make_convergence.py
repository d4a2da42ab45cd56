"""Write the FP64 oracle's relaxation of the 64^3 droplet to the full stop rule.

The workload is the 2000-step relaxation case (64^3, R = 28 voxels, dx = 10 nm,
L1 = 6e-12 N, strong radial anchoring at the equilibrium shell order,
T = 300 K, T* = 307 K, random directors S0 = 0.5 from seed 12345, rounded to
FP32 as the GPU run receives them, dt = dt_max) continued until the stop rule
of the single-temperature loop holds (QSR.cu:3322-3335): iteration > 100,
|dF/F| < 1e-6 between log points 100 steps apart, and the mid-plane defect
density changed by less than 1e-4 (valid slice).  The 0.5 tau_align time gate
is off (the CLI's early_stop_min_alignment_time = false): at dt = 1.6e-10 s it
would need about 4e8 steps.

Runs in the build container with the oracle (tests/golden/ holds the result,
data only): per log point the iteration, F, the defect count and plaquettes;
the final field's mid-plane (z = 32) and, after the stop, the largest
|dQ| of one more step (how far the update is above the FP32 spacing of Q).

    python tests/golden/make_convergence.py
"""
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

OUT = os.path.join(HERE, "config0_convergence.npz")
TOL, DEFECT_EPS, EVERY, MAX_STEPS = 1e-6, 1e-4, 100, 400000


def main():
    dims, R = (64, 64, 64), 28.0
    ph = O.phys(T=300.0, T_star=307.0, kappa=0.0, L1=6e-12)
    od = O.DropletOracle(O.Grid(*dims), ph, gamma=0.1, radius_index=R)
    dt = O.dt_estimate(1e-8, 1e-8, 1e-8, 0.1, ph.a, ph.b, ph.c, 307.0, 1, 0.0, 6e-12, 0, 0, math.nan, 0.5,
                       300.0, 300.0)["dt_max"]
    S = O.shell_order(False, 0.0, ph.a, ph.b, ph.c, 300.0, 307.0)
    Q = od.init_director(0.5, 0, 12345).astype(np.float32).astype(np.float64)
    its, Fs, ds, us = [], [], [], []
    prevF = prevD = None
    it = 0
    t0 = time.time()
    stop = -1
    while it < MAX_STEPS:
        Q = od.step(Q, S, dt, steps=EVERY)
        it += EVERY
        o = od.observables(Q, S)
        F, D, used = o["total"], o["defect_density"], o["plaquettes"]
        its.append(it - 1)  # the reference logs iteration it-1 after it steps (first row: iter 0 after 1 step)
        Fs.append(F)
        ds.append(o["defects"])
        us.append(used)
        if it % 5000 == 0:
            print(f"iter {it}: F {F:.10e} defects {o['defects']} t {time.time() - t0:.0f} s", flush=True)
        if prevF is not None and it - 1 > 100:
            e_conv = abs((F - prevF) / prevF) < TOL
            d_conv = used > 0 and math.isfinite(prevD) and math.isfinite(D) and abs(D - prevD) < DEFECT_EPS
            if e_conv and d_conv:
                stop = it - 1
                break
        prevF, prevD = F, (D if used > 0 else math.nan)
    Q1 = od.step(Q, S, dt, steps=1)
    dq = float(np.abs(Q1 - Q).max())
    print(f"stop at log iteration {stop} ({it} steps), F {Fs[-1]:.12e}, defects {ds[-1]}, "
          f"next-step max |dQ| {dq:.3e}, {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(OUT, iteration=np.array(its), F=np.array(Fs), defects=np.array(ds),
                        plaquettes=np.array(us), stop_iteration=np.array(stop), steps=np.array(it),
                        dt=np.array(dt), mid_plane=Q[:, :, 32, :].copy(), next_step_max_dQ=np.array(dq))
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
