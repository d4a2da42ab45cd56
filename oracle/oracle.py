"""ctypes wrapper over the FP64 oracle (``libqsr_oracle.so``) and the reference
harness (``_ref/libqsr_ref.so``).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` as the checker.  The product path
(``lcdrsimulator_amd``) never imports this module.

Arrays follow the reference interchange format: ``Q`` is float64 of shape
``(Nx, Ny, Nz, 5)`` with components (Qxx, Qxy, Qxz, Qyy, Qyz), C-contiguous, so
the flat index is ``k + Nz*(j + Ny*i)`` (QSR.cu:286).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)


class Geom(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "dx", "dy", "dz", "center_x", "center_y", "center_z", "radius_x", "radius_y", "radius_z",
        "min_spacing", "min_radius", "shell_half_thickness", "core_radius")]


class Phys(C.Structure):
    _fields_ = [("a", C.c_double), ("b", C.c_double), ("c", C.c_double), ("T", C.c_double),
                ("T_star", C.c_double), ("modechoice", C.c_int), ("kappa", C.c_double),
                ("L1", C.c_double), ("L2", C.c_double), ("L3", C.c_double), ("W", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "libqsr_oracle.so"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "libqsr_oracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


def ref_lib():
    """The reference's own QSR.cpp functions, or None when oracle/_ref is not built."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libqsr_ref.so")
        if not os.path.exists(path):
            return None
        _REF = C.CDLL(path)
        _REF.ref_time_steps.restype = C.c_double
    return _REF


def _declare(L):
    L.orc_sample.restype = None
    L.orc_shell_mask.restype = C.c_size_t
    for name in ("orc_surface_area", "orc_shell_thickness", "orc_bulk_fS", "orc_S_eq", "orc_A_NI",
                 "orc_A_spinN", "orc_T_from_A", "orc_shell_order", "orc_radiality", "orc_checker"):
        getattr(L, name).restype = C.c_double
    L.orc_bulk_fS.argtypes = [C.c_double] * 4
    L.orc_S_eq.argtypes = [C.c_double] * 3
    L.orc_A_NI.argtypes = [C.c_double] * 2
    L.orc_A_spinN.argtypes = [C.c_double] * 2
    L.orc_T_from_A.argtypes = [C.c_double] * 3
    L.orc_shell_order.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_double, C.c_int]
    L.orc_shell_thickness.argtypes = [C.POINTER(Geom), C.c_size_t]
    L.orc_checker.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.POINTER(Geom), C.c_int, C.c_double]


def _p(a):
    return a.ctypes.data_as(_dp)


@dataclass
class Grid:
    Nx: int
    Ny: int
    Nz: int
    dx: float = 1e-8
    dy: float = 1e-8
    dz: float = 1e-8

    @property
    def n(self):
        return self.Nx * self.Ny * self.Nz

    @property
    def dims(self):
        return (self.Nx, self.Ny, self.Nz)


def geometry(grid: Grid, radius_index: float = -1.0, center=(math.nan, math.nan, math.nan)) -> Geom:
    g = Geom()
    lib().orc_make_geometry(grid.Nx, grid.Ny, grid.Nz, C.c_double(grid.dx), C.c_double(grid.dy),
                            C.c_double(grid.dz), C.c_double(radius_index), C.c_double(center[0]),
                            C.c_double(center[1]), C.c_double(center[2]), C.byref(g))
    return g


def sample(g: Geom, i, j, k):
    out = (C.c_double * 5)()
    lib().orc_sample(C.byref(g), i, j, k, out)
    return tuple(out)


def shell_mask(grid: Grid, g: Geom):
    m = np.zeros(grid.dims, dtype=np.uint8)
    cnt = lib().orc_shell_mask(C.byref(g), grid.Nx, grid.Ny, grid.Nz, m.ctypes.data_as(_u8p))
    return m, int(cnt)


def shell_thickness(g: Geom, count: int) -> float:
    return lib().orc_shell_thickness(C.byref(g), count)


def phys(a=0.044e6, b=1.413e6, c=1.153e6, T=300.0, T_star=308.0, modechoice=1, kappa=6.5e-12,
         L1=0.0, L2=0.0, L3=0.0, W=0.0) -> Phys:
    return Phys(a, b, c, T, T_star, modechoice, kappa, L1, L2, L3, W)


def bulk_ABC(a, b, c, T, T_star, mode=1):
    out = (C.c_double * 3)()
    lib().orc_bulk_ABC(C.c_double(a), C.c_double(b), C.c_double(c), C.c_double(T), C.c_double(T_star),
                       C.c_int(mode), out)
    return tuple(out)


def S_eq(A, B, Cc):
    return lib().orc_S_eq(A, B, Cc)


def shell_order(custom, custom_S, a, b, c, T, T_star, mode=1):
    return lib().orc_shell_order(int(custom), custom_S, a, b, c, T, T_star, mode)


DT_KEYS = ("min_dx", "S_scale", "L_eff_total", "D_total", "dt_diff_total", "L_eff_explicit",
           "dt_diff", "bulk_rate", "dt_bulk", "dt_max")


def dt_estimate(dx, dy, dz, gamma, a, b, c, T_star, mode, kappa, L1, L2, L3, S_ref_used, S0, T_a,
                T_b, use_semi_implicit=False, L_stab=0.0, jacobi_iters=0):
    out = (C.c_double * 10)()
    args = [C.c_double(v) for v in (dx, dy, dz, gamma, a, b, c, T_star)]
    args += [C.c_int(mode)] + [C.c_double(v) for v in (kappa, L1, L2, L3, S_ref_used, S0, T_a, T_b)]
    args += [C.c_int(int(use_semi_implicit)), C.c_double(L_stab), C.c_int(jacobi_iters), out]
    lib().orc_dt_estimate(*args)
    return dict(zip(DT_KEYS, out))


def _q(Q):
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    assert Q.ndim == 4 and Q.shape[3] == 5
    return Q


def laplacian(Q, grid: Grid):
    Q = _q(Q)
    lap = np.zeros_like(Q)
    lib().orc_laplacian(_p(Q), _p(lap), *grid.dims, C.c_double(grid.dx), C.c_double(grid.dy),
                        C.c_double(grid.dz))
    return lap


def divergence(Q, grid: Grid):
    Q = _q(Q)
    D = np.zeros((3,) + grid.dims)
    lib().orc_divergence(_p(Q), _p(D[0]), _p(D[1]), _p(D[2]), *grid.dims, C.c_double(grid.dx),
                         C.c_double(grid.dy), C.c_double(grid.dz))
    return D


def chempot(Q, grid: Grid, p: Phys, with_l3=True):
    """mu = molecular field after the chemical-potential (and L3) kernels."""
    Q = _q(Q)
    lap = laplacian(Q, grid)
    D = divergence(Q, grid) if (p.L2 != 0.0 or p.L3 != 0.0) else np.zeros((3,) + grid.dims)
    mu = np.zeros_like(Q)
    d = [C.c_double(grid.dx), C.c_double(grid.dy), C.c_double(grid.dz)]
    lib().orc_chempot(_p(Q), _p(mu), _p(lap), *grid.dims, *d, C.byref(p), _p(D[0]), _p(D[1]), _p(D[2]))
    if with_l3 and p.L3 != 0.0:
        lib().orc_chempot_l3(_p(Q), _p(mu), *grid.dims, *d, C.byref(p), _p(D[0]), _p(D[1]), _p(D[2]))
    return mu


def update(Q, mu, grid: Grid, g: Geom, shell, dt, gamma, W, cap=0.0):
    """One pass of the explicit update alone (orc_update, QSR.cu:697-732) on a copy of Q."""
    Q = _q(Q).copy()
    mu = _q(mu)
    shell = np.ascontiguousarray(shell, dtype=np.uint8)
    lib().orc_update(_p(Q), _p(mu), *grid.dims, C.c_double(dt), shell.ctypes.data_as(_u8p), C.byref(g),
                     C.c_double(gamma), C.c_double(W), C.c_double(cap))
    return Q


class DropletOracle:
    """Stateful FP64 droplet solver in the QSR.cu step order (oracle)."""

    def __init__(self, grid: Grid, p: Phys, gamma=0.1, radius_index=-1.0,
                 center=(math.nan, math.nan, math.nan), q_norm_cap=0.0):
        self.grid, self.p, self.gamma, self.q_norm_cap = grid, p, gamma, q_norm_cap
        self.g = geometry(grid, radius_index, center)
        self.shell, self.shell_count = shell_mask(grid, self.g)
        self.thickness = shell_thickness(self.g, self.shell_count)
        self.work = np.zeros(23 * grid.n)

    def init_director(self, S0, mode, seed):
        Q = np.zeros(self.grid.dims + (5,))
        lib().orc_init_director(_p(Q), *self.grid.dims, C.c_double(S0), C.c_int(mode), C.byref(self.g),
                                C.c_longlong(seed))
        return Q

    def init_isotropic(self, amp, seed):
        Q = np.zeros(self.grid.dims + (5,))
        lib().orc_init_isotropic(_p(Q), *self.grid.dims, C.c_double(amp), C.byref(self.g),
                                 C.c_longlong(seed))
        return Q

    def step(self, Q, S_shell, dt, steps=1, T=None):
        Q = _q(Q).copy()
        p = self.p
        if T is not None:
            p = Phys(p.a, p.b, p.c, T, p.T_star, p.modechoice, p.kappa, p.L1, p.L2, p.L3, p.W)
        for _ in range(steps):
            lib().orc_step(_p(Q), *self.grid.dims, C.byref(self.g), C.byref(p),
                           self.shell.ctypes.data_as(_u8p), C.c_double(self.thickness),
                           C.c_double(S_shell), C.c_double(dt), C.c_double(self.gamma),
                           C.c_double(self.q_norm_cap), _p(self.work))
        return Q

    def step_imex(self, Q, S_shell, dt, L_stab, jacobi_iters, steps=1, T=None):
        Q = _q(Q).copy()
        p = self.p
        if T is not None:
            p = Phys(p.a, p.b, p.c, T, p.T_star, p.modechoice, p.kappa, p.L1, p.L2, p.L3, p.W)
        for _ in range(steps):
            lib().orc_step_imex(_p(Q), *self.grid.dims, C.byref(self.g), C.byref(p),
                                self.shell.ctypes.data_as(_u8p), C.c_double(self.thickness),
                                C.c_double(S_shell), C.c_double(dt), C.c_double(self.gamma),
                                C.c_double(L_stab), C.c_int(jacobi_iters), _p(self.work))
        return Q

    def energy(self, Q, S_shell, T=None):
        Q = _q(Q)
        p = self.p
        if T is not None:
            p = Phys(p.a, p.b, p.c, T, p.T_star, p.modechoice, p.kappa, p.L1, p.L2, p.L3, p.W)
        out = (C.c_double * 3)()
        lib().orc_energy(_p(Q), *self.grid.dims, C.byref(self.g), C.byref(p),
                         self.shell.ctypes.data_as(_u8p), C.c_double(S_shell), C.c_double(self.thickness), out)
        return tuple(out)

    def observables(self, Q, S_shell, z_slice=None, S_thr=0.1, cutoff=0.25, xi_S_thr=None):
        """Every quench_log column the log pass computes (QSR.cu:2787-2874)."""
        Q = _q(Q)
        Nx, Ny, Nz = self.grid.dims
        zs = Nz // 2 if z_slice is None else z_slice
        bulk, el, anch = self.energy(Q, S_shell)
        rad = lib().orc_radiality(_p(Q), Nx, Ny, Nz, C.byref(self.g))
        am = (C.c_double * 2)()
        lib().orc_avg_max_S(_p(Q), Nx, Ny, Nz, C.byref(self.g), am)
        chk = lib().orc_checker(_p(Q), Nx, Ny, Nz, C.byref(self.g), zs, am[1])
        nd, nu = C.c_uint(), C.c_uint()
        lib().orc_defect2d(_p(Q), Nx, Ny, Nz, zs, C.c_double(S_thr), C.c_double(cutoff), C.byref(nd),
                           C.byref(nu))
        xi, ne = C.c_double(), C.c_uint()
        lib().orc_xi2d(_p(Q), Nx, Ny, Nz, zs, C.c_double(S_thr if xi_S_thr is None else xi_S_thr),
                       C.byref(xi), C.byref(ne))
        return dict(bulk=bulk, elastic=el, anchoring=anch, total=bulk + el + anch, radiality=rad,
                    avg_S=am[0], max_S=am[1], checker_rel=chk, defects=nd.value, plaquettes=nu.value,
                    defect_density=(nd.value / nu.value) if nu.value else math.nan,
                    xi_grad_proxy=xi.value, xi_edges=ne.value)


def nematic(Q):
    q = np.ascontiguousarray(Q, dtype=np.float64).reshape(-1, 5)
    out = np.zeros((q.shape[0], 4))
    o = (C.c_double * 4)()
    for t in range(q.shape[0]):
        lib().orc_nematic(q[t].ctypes.data_as(_dp), o)
        out[t] = tuple(o)
    return out


def mt19937(seed, n):
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_mt19937_stream(C.c_uint32(seed), out.ctypes.data_as(_u32p), C.c_size_t(n))
    return out


def canonical(seed, n):
    out = np.zeros(n)
    lib().orc_canonical_stream(C.c_uint32(seed), _p(out), C.c_size_t(n))
    return out


def normal(seed, mean, std, n):
    out = np.zeros(n)
    lib().orc_normal_stream(C.c_uint32(seed), C.c_double(mean), C.c_double(std), _p(out), C.c_size_t(n))
    return out


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return tuple(o)


def noise5(seed, voxel, step):
    g = (C.c_double * 5)()
    lib().orc_noise5(C.c_uint64(seed), C.c_uint64(voxel), C.c_uint32(step), g)
    return tuple(g)


def periodic_step(Q, grid: Grid, A, B, Cc, kappa, gamma, dt, sigma=0.0, seed=0, step=0, cap=0.0):
    Q = _q(Q)
    out = np.zeros_like(Q)
    lib().orc_periodic_step(_p(Q), _p(out), *grid.dims, C.c_double(grid.dx), C.c_double(grid.dy),
                            C.c_double(grid.dz), C.c_double(A), C.c_double(B), C.c_double(Cc),
                            C.c_double(kappa), C.c_double(gamma), C.c_double(dt), C.c_double(sigma),
                            C.c_uint64(seed), C.c_uint32(step), C.c_double(cap))
    return out


PERIODIC_KEYS = ("bulk", "elastic", "total", "avg_S", "max_S", "defect_density", "plaquettes_used",
                 "defects", "xi_grad_proxy", "defect_line_density")


def periodic_observables(Q, grid: Grid, A, B, Cc, kappa, S_thr=0.1, cutoff=0.25):
    Q = _q(Q)
    out = (C.c_double * 10)()
    lib().orc_periodic_observables(_p(Q), *grid.dims, C.c_double(grid.dx), C.c_double(grid.dy),
                                   C.c_double(grid.dz), C.c_double(A), C.c_double(B), C.c_double(Cc),
                                   C.c_double(kappa), C.c_double(S_thr), C.c_double(cutoff), out)
    return dict(zip(PERIODIC_KEYS, out))
