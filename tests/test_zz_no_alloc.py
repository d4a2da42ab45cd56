"""qsr_energies_noalloc (no device allocation; partials in the library's fixed
scratch array) against qsr_energies through the C ABI: bit-exact totals on two
calls in a row, the second on a smaller input, and the stated limits."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _droplets(qsr, specs):
    from lcdrsimulator_amd import warmstart as ws
    return [ws.droplet_params(R, T, W) for R, T, W in specs]


def _call(qsr, s, fn, boxes):
    L = qsr.load()
    bx = np.ascontiguousarray(boxes, dtype=np.int32)
    out = np.full(bx.size, np.nan)
    rc = getattr(L, fn)(s._h, bx.ctypes.data_as(C.POINTER(C.c_int)), bx.size, out.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, out


def test_noalloc_energies_equal_the_allocating_call(qsr):
    big = _droplets(qsr, [(40.0, 300.0, 1e-4), (24.0, 298.0, 0.0), (31.0, 303.0, 1e-3), (18.0, 300.0, 1e-5)])
    small = _droplets(qsr, [(9.0, 300.0, 1e-4), (7.0, 301.0, 0.0)])
    print("[case] noalloc energies: 4 boxes up to R=40, then 2 boxes up to R=9", flush=True)
    with qsr.Solver(big) as s:
        for b, p in enumerate(big):
            s.upload(qsr.init_field(p, b % 2, 0.5, seed=b), box=b)
        s.step(13)
        order = [2, 0, 3, 1]
        rc_a, ref = _call(qsr, s, "qsr_energies", order)
        rc_b, got = _call(qsr, s, "qsr_energies_noalloc", order)
        assert rc_a == 0 and rc_b == 0
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(s.energies_noalloc([1]), ref[[3]])
    with qsr.Solver(small) as s:
        for b, p in enumerate(small):
            s.upload(qsr.init_field(p, 1, 0.5, seed=7 + b), box=b)
        s.step(4)
        rc_a, ref = _call(qsr, s, "qsr_energies", [1, 0])
        rc_b, got = _call(qsr, s, "qsr_energies_noalloc", [1, 0])
        assert rc_a == 0 and rc_b == 0
        np.testing.assert_array_equal(got, ref)
        assert np.all(np.isfinite(got)) and np.all(got != 0.0)
        # the totals are the observables' energies
        o = s.observables(box=0, want_defects=False, want_xi=False, want_order=False)
        assert got[1] == o["total"]
        assert s.energies_noalloc([]).size == 0


def test_noalloc_energies_reject_what_they_cannot_hold(qsr):
    p = _droplets(qsr, [(9.0, 300.0, 1e-4)])[0]
    per = qsr.Params(Nx=16, Ny=16, Nz=16, periodic=1, T_star=307.0, T_high=300.0, kappa=6e-12, dt=1e-10)
    with qsr.Solver([p, per]) as s:
        s.upload(qsr.init_field(p, 1, 0.5, seed=1), box=0)
        s.upload(np.zeros((16, 16, 16, 5)), box=1)
        with pytest.raises(qsr.QSRError, match="droplet boxes only"):
            s.energies_noalloc([0, 1])
        with pytest.raises(qsr.QSRError, match="at most 4096"):
            s.energies_noalloc([0] * 4097)
        with pytest.raises(qsr.QSRError):
            s.energies_noalloc([2])
        np.testing.assert_array_equal(s.energies_noalloc([0]), s.energies([0]))
