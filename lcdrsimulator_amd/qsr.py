"""ctypes binding of libqsr_amd.so (include/qsr_amd.h).

The Python side of the engine: builds ``qsr_params`` records with the
reference's defaults (QSR.cu:1944-1949, 1958-2139), drives a context
(upload / step / observables / download) and exposes the host physics
(S_eq, the stability window, the seeded initial states).

Arrays use the reference interchange format: float64 ``(Nx, Ny, Nz, 5)`` with
components (Qxx, Qxy, Qxz, Qyy, Qyz), flat index ``k + Nz*(j + Ny*i)``.
There is no CPU fallback: every call goes to the HIP library and raises if
the library or a GPU is missing.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field, fields

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libqsr_amd.so")
_lib = None

PROTOCOL_CONSTANT, PROTOCOL_STEP, PROTOCOL_RAMP = 0, 1, 2


class QSRError(RuntimeError):
    pass


class _Params(C.Structure):
    _fields_ = [
        ("Nx", C.c_int32), ("Ny", C.c_int32), ("Nz", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("a", C.c_double), ("b", C.c_double), ("c", C.c_double), ("T_star", C.c_double),
        ("modechoice", C.c_int32),
        ("kappa", C.c_double), ("L1", C.c_double), ("L2", C.c_double), ("L3", C.c_double),
        ("W", C.c_double), ("gamma", C.c_double), ("dt", C.c_double), ("q_norm_cap", C.c_double),
        ("periodic", C.c_int32), ("radius_index", C.c_double), ("center_index", C.c_double * 3),
        ("protocol", C.c_int32), ("T_high", C.c_double), ("T_low", C.c_double),
        ("pre_equil_iters", C.c_int32), ("ramp_iters", C.c_int32),
        ("boundary_custom", C.c_int32), ("boundary_S", C.c_double),
        ("noise_sigma", C.c_double), ("seed", C.c_uint64),
    ]


class _Obs(C.Structure):
    _fields_ = [
        ("iteration", C.c_int64), ("T", C.c_double), ("S_shell", C.c_double),
        ("bulk", C.c_double), ("elastic", C.c_double), ("anchoring", C.c_double), ("total", C.c_double),
        ("radiality", C.c_double), ("avg_S", C.c_double), ("max_S", C.c_double), ("checker_rel", C.c_double),
        ("defects", C.c_uint64), ("plaquettes", C.c_uint64), ("defect_density", C.c_double),
        ("xi_grad_proxy", C.c_double), ("xi_edges", C.c_uint64), ("defect_line_density", C.c_double),
    ]


class _ObsOpts(C.Structure):
    _fields_ = [("z_slice", C.c_int32), ("S_threshold", C.c_double), ("charge_cutoff", C.c_double),
                ("xi_z_slice", C.c_int32), ("xi_S_threshold", C.c_double),
                ("want_defects", C.c_int32), ("want_xi", C.c_int32), ("skip_order", C.c_int32)]


class _BoxInfo(C.Structure):
    _fields_ = [("voxels", C.c_uint64), ("shell_voxels", C.c_uint64), ("active_voxels", C.c_uint64),
                ("shell_volume", C.c_double), ("surface_area", C.c_double), ("nominal_thickness", C.c_double),
                ("effective_thickness", C.c_double), ("S_eq", C.c_double), ("iteration", C.c_int64)]


EXPORTS = (
    "qsr_create", "qsr_create_batch", "qsr_destroy", "qsr_last_error", "qsr_device_count",
    "qsr_upload_q", "qsr_download_q", "qsr_download_nematic", "qsr_step", "qsr_step_graph",
    "qsr_step_imex", "qsr_observables", "qsr_box_info_get", "qsr_set_iteration", "qsr_last_step_ms",
    "qsr_step_work", "qsr_profile_steps", "qsr_autotune", "qsr_set_tuning", "qsr_set_step_fusion",
    "qsr_step_fusion_info", "qsr_S_eq", "qsr_dt_estimate", "qsr_init_field",
    "qsr_load_qtensor_file", "qsr_save_qtensor_file", "qsr_set_box_active", "qsr_set_boxes_active",
    "qsr_surr_num_params", "qsr_surr_create", "qsr_surr_destroy", "qsr_surr_init_params", "qsr_surr_set_params",
    "qsr_surr_get_params", "qsr_surr_save", "qsr_surr_load", "qsr_surr_eval", "qsr_surr_adam_step",
    "qsr_surr_dataset_add", "qsr_surr_dataset_size", "qsr_surr_dataset_clear", "qsr_surr_train", "qsr_surr_apply",
    "qsr_surr_dataset_add_increment", "qsr_surr_set_input", "qsr_surr_apply_ladder", "qsr_surr_eval_gain",
    "qsr_surr_apply_boxes", "qsr_energies", "qsr_energies_noalloc", "qsr_surr_set_tiles",
)

SURR_CHECKED, SURR_FORCE, SURR_PROPOSE = 0, 1, 2
SURR_INPUT_FIELD, SURR_INPUT_INCREMENT, SURR_INPUT_GAIN, SURR_INPUT_STEP = 0, 1, 2, 3
SURR_TILES_DROPLET, SURR_TILES_ALL = 0, 1


class _SurrResult(C.Structure):
    _fields_ = [("F_before", C.c_double), ("F_after", C.c_double), ("res_before", C.c_double),
                ("res_after", C.c_double), ("accepted", C.c_int32), ("ms_network", C.c_double),
                ("scale", C.c_double), ("tried", C.c_int32), ("net_voxels", C.c_int64)]


def load():
    """Load libqsr_amd.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise QSRError(f"{LIB_PATH} is missing: build it with `python -m lcdrsimulator_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, dp = C.c_void_p, C.POINTER(C.c_double)
    L.qsr_last_error.restype = C.c_char_p
    L.qsr_create.argtypes = [C.POINTER(_Params), C.c_int, C.POINTER(vp)]
    L.qsr_create_batch.argtypes = [C.POINTER(_Params), C.c_int, C.c_int, C.POINTER(vp)]
    L.qsr_destroy.argtypes = [vp]
    L.qsr_upload_q.argtypes = [vp, C.c_int, dp, C.c_size_t]
    L.qsr_download_q.argtypes = [vp, C.c_int, dp, C.c_size_t]
    L.qsr_download_nematic.argtypes = [vp, C.c_int, dp, C.c_size_t]
    L.qsr_step.argtypes = [vp, C.c_int]
    L.qsr_step_graph.argtypes = [vp, C.c_int, C.c_int]
    L.qsr_step_imex.argtypes = [vp, C.c_int, C.c_double, C.c_int]
    L.qsr_observables.argtypes = [vp, C.c_int, C.POINTER(_ObsOpts), C.POINTER(_Obs)]
    L.qsr_box_info_get.argtypes = [vp, C.c_int, C.POINTER(_BoxInfo)]
    L.qsr_set_iteration.argtypes = [vp, C.c_int, C.c_int64]
    L.qsr_last_step_ms.argtypes = [vp, dp]
    L.qsr_profile_steps.argtypes = [vp, C.c_int, dp, dp]
    L.qsr_autotune.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), dp]
    L.qsr_set_tuning.argtypes = [vp, C.c_int, C.c_int]
    L.qsr_set_step_fusion.argtypes = [vp, C.c_int, C.c_int]
    L.qsr_step_fusion_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_uint64)]
    L.qsr_set_box_active.argtypes = [vp, C.c_int, C.c_int]
    L.qsr_set_boxes_active.argtypes = [vp, C.POINTER(C.c_int), C.c_int, C.c_int]
    L.qsr_step_work.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.qsr_S_eq.restype = C.c_double
    L.qsr_S_eq.argtypes = [C.c_double] * 5 + [C.c_int]
    L.qsr_dt_estimate.argtypes = [C.c_double] * 8 + [C.c_int] + [C.c_double] * 8 + [C.c_int, C.c_double, C.c_int, dp]
    L.qsr_init_field.argtypes = [C.POINTER(_Params), C.c_int, C.c_double, C.c_double, C.c_int64, dp, C.c_size_t]
    L.qsr_load_qtensor_file.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, dp, C.c_size_t,
                                        C.POINTER(C.c_uint64)]
    L.qsr_save_qtensor_file.argtypes = [C.c_char_p, dp, C.c_int, C.c_int, C.c_int, C.c_size_t]
    L.qsr_device_count.argtypes = [C.POINTER(C.c_int)]
    fp = C.POINTER(C.c_float)
    L.qsr_surr_num_params.argtypes = [C.POINTER(C.c_size_t)]
    L.qsr_surr_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.qsr_surr_destroy.argtypes = [vp]
    L.qsr_surr_init_params.argtypes = [vp, C.c_uint64]
    L.qsr_surr_set_params.argtypes = [vp, fp, C.c_size_t]
    L.qsr_surr_get_params.argtypes = [vp, fp, C.c_size_t]
    L.qsr_surr_save.argtypes = [vp, C.c_char_p]
    L.qsr_surr_load.argtypes = [vp, C.c_char_p]
    L.qsr_surr_eval.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, fp, fp, fp, dp, fp, fp]
    L.qsr_surr_adam_step.argtypes = [vp, fp, C.c_double]
    L.qsr_surr_dataset_add.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int)]
    L.qsr_surr_dataset_size.argtypes = [vp, C.POINTER(C.c_int)]
    L.qsr_surr_dataset_clear.argtypes = [vp]
    L.qsr_surr_train.argtypes = [vp, C.POINTER(C.c_int), C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                 C.c_uint64, C.c_double, dp]
    L.qsr_surr_apply.argtypes = [vp, vp, C.c_int, C.POINTER(_SurrResult)]
    L.qsr_surr_dataset_add_increment.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int)]
    L.qsr_surr_set_input.argtypes = [vp, C.c_int, C.c_double]
    L.qsr_surr_set_tiles.argtypes = [vp, C.c_int]
    L.qsr_surr_apply_ladder.argtypes = [vp, vp, C.c_int, dp, C.c_int, C.POINTER(_SurrResult)]
    L.qsr_surr_eval_gain.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, fp, fp, fp, fp, dp, fp, fp]
    L.qsr_energies.argtypes = [vp, C.POINTER(C.c_int), C.c_int, dp]
    L.qsr_energies_noalloc.argtypes = [vp, C.POINTER(C.c_int), C.c_int, dp]
    L.qsr_surr_apply_boxes.argtypes = [vp, vp, C.c_int, dp, C.c_int, C.POINTER(C.c_int), C.c_int,
                                       C.POINTER(_SurrResult)]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise QSRError(f"qsr error {rc}: {load().qsr_last_error().decode()}")


def device_count() -> int:
    n = C.c_int(0)
    _check(load().qsr_device_count(C.byref(n)))
    return n.value


@dataclass
class Params:
    """One box.  Defaults are the reference's (QSR.cu:1944-1949, 2010-2116)."""
    Nx: int = 100
    Ny: int = 100
    Nz: int = 100
    dx: float = 1e-8
    dy: float = 1e-8
    dz: float = 1e-8
    a: float = 0.044e6
    b: float = 1.413e6
    c: float = 1.153e6
    T_star: float = 308.0
    modechoice: int = 1
    kappa: float = 6.5e-12
    L1: float = 0.0
    L2: float = 0.0
    L3: float = 0.0
    W: float = 0.0
    gamma: float = 0.1
    dt: float = 1e-10
    q_norm_cap: float = 0.0
    periodic: int = 0
    radius_index: float = -1.0
    center_index: tuple = (math.nan, math.nan, math.nan)
    protocol: int = PROTOCOL_CONSTANT
    T_high: float = 300.0
    T_low: float = 295.0
    pre_equil_iters: int = 0
    ramp_iters: int = 1000
    boundary_custom: int = 0
    boundary_S: float = 0.0
    noise_sigma: float = 0.0
    seed: int = 0

    def to_c(self) -> _Params:
        p = _Params()
        for f in fields(self):
            v = getattr(self, f.name)
            if f.name == "center_index":
                p.center_index = (C.c_double * 3)(*v)
            else:
                setattr(p, f.name, v)
        return p

    @property
    def dims(self):
        return (self.Nx, self.Ny, self.Nz)

    @property
    def n(self):
        return self.Nx * self.Ny * self.Nz


OBS_KEYS = [n for n, _ in _Obs._fields_]


class Solver:
    """A context of one box (or a batch of boxes) on one GPU."""

    def __init__(self, params, device: int = 0):
        L = load()
        self.params = list(params) if isinstance(params, (list, tuple)) else [params]
        arr = (_Params * len(self.params))(*[p.to_c() for p in self.params])
        h = C.c_void_p()
        if len(self.params) == 1:
            _check(L.qsr_create(arr, device, C.byref(h)))
        else:
            _check(L.qsr_create_batch(arr, len(self.params), device, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load().qsr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, Q, box: int = 0):
        p = self.params[box]
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        if Q.shape != p.dims + (5,):
            raise ValueError(f"Q shape {Q.shape} != {p.dims + (5,)}")
        _check(load().qsr_upload_q(self._h, box, Q.ctypes.data_as(C.POINTER(C.c_double)), p.n))

    def download(self, box: int = 0) -> np.ndarray:
        p = self.params[box]
        Q = np.empty(p.dims + (5,))
        _check(load().qsr_download_q(self._h, box, Q.ctypes.data_as(C.POINTER(C.c_double)), p.n))
        return Q

    def nematic(self, box: int = 0) -> np.ndarray:
        p = self.params[box]
        out = np.empty(p.dims + (4,))
        _check(load().qsr_download_nematic(self._h, box, out.ctypes.data_as(C.POINTER(C.c_double)), p.n))
        return out

    def step(self, n: int = 1):
        _check(load().qsr_step(self._h, int(n)))

    def step_graph(self, n: int, chunk: int = 64):
        _check(load().qsr_step_graph(self._h, int(n), int(chunk)))

    def step_imex(self, n: int, L_stab: float, jacobi_iters: int = 25):
        """Semi-implicit steps (explicit RHS + Jacobi sweeps, QSR.cu:2721-2749)."""
        _check(load().qsr_step_imex(self._h, int(n), float(L_stab), int(jacobi_iters)))

    def set_iteration(self, it: int, box: int = 0):
        _check(load().qsr_set_iteration(self._h, box, int(it)))

    def observables(self, box: int = 0, z_slice=-1, S_threshold=0.1, charge_cutoff=0.25, xi_z_slice=-1,
                    xi_S_threshold=-1.0, want_defects=True, want_xi=True, want_order=True) -> dict:
        """FP64 log-point observables; want_order=False skips <S>, max S and radiality (NaN)."""
        o = _ObsOpts(z_slice, S_threshold, charge_cutoff, xi_z_slice, xi_S_threshold, int(want_defects),
                     int(want_xi), int(not want_order))
        r = _Obs()
        _check(load().qsr_observables(self._h, box, C.byref(o), C.byref(r)))
        return {k: getattr(r, k) for k in OBS_KEYS}

    def energies(self, boxes) -> np.ndarray:
        """Total free energy of each listed box (observables' ``total``, energies only), with one
        host synchronisation for all."""
        bx = np.ascontiguousarray(boxes, dtype=np.int32).ravel()
        out = np.zeros(bx.size)
        _check(load().qsr_energies(self._h, bx.ctypes.data_as(C.POINTER(C.c_int)), bx.size,
                                   out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def energies_noalloc(self, boxes) -> np.ndarray:
        """:meth:`energies` without device allocation (droplet boxes only, at most 4096 per call):
        the library's fixed scratch array holds the partials."""
        bx = np.ascontiguousarray(boxes, dtype=np.int32).ravel()
        out = np.zeros(bx.size)
        _check(load().qsr_energies_noalloc(self._h, bx.ctypes.data_as(C.POINTER(C.c_int)), bx.size,
                                           out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def info(self, box: int = 0) -> dict:
        r = _BoxInfo()
        _check(load().qsr_box_info_get(self._h, box, C.byref(r)))
        return {k: getattr(r, k) for k, _ in _BoxInfo._fields_}

    def last_step_ms(self) -> float:
        v = C.c_double()
        _check(load().qsr_last_step_ms(self._h, C.byref(v)))
        return v.value

    def profile_steps(self, n: int):
        """(main-kernel ms, face-kernel ms) per launch, HIP events: n single-step launches,
        or n / 2 two-step launches when the context pairs its steps (see step_fusion)."""
        a, b = C.c_double(), C.c_double()
        _check(load().qsr_profile_steps(self._h, int(n), C.byref(a), C.byref(b)))
        return a.value, b.value

    def autotune(self, trial_steps: int = 10) -> dict:
        """Pick the fastest stencil variant / planes-per-tile on the resident field."""
        v, ch, ms = C.c_int(), C.c_int(), C.c_double()
        _check(load().qsr_autotune(self._h, int(trial_steps), C.byref(v), C.byref(ch), C.byref(ms)))
        return {"variant": v.value, "planes_per_tile": ch.value, "ms": ms.value}

    def set_tuning(self, variant: int, planes_per_tile: int):
        _check(load().qsr_set_tuning(self._h, int(variant), int(planes_per_tile)))

    def set_step_fusion(self, on: bool, planes_per_tile: int = 0):
        """Pair explicit steps in the two-step kernel where the context allows it (default on)."""
        _check(load().qsr_set_step_fusion(self._h, int(bool(on)), int(planes_per_tile)))

    def step_fusion(self) -> dict:
        """{"active", "planes_per_tile", "tile_voxels"} of the two-step kernel for this context."""
        a, ch, tv = C.c_int(), C.c_int(), C.c_uint64()
        _check(load().qsr_step_fusion_info(self._h, C.byref(a), C.byref(ch), C.byref(tv)))
        return {"active": bool(a.value), "planes_per_tile": ch.value, "tile_voxels": tv.value}

    def set_active(self, box: int, active: bool):
        """Hold (False) or resume (True) one box of a batch: held boxes are not stepped."""
        _check(load().qsr_set_box_active(self._h, int(box), int(bool(active))))

    def set_active_many(self, boxes, active: bool):
        """set_active for every listed box at once (one tile rebuild)."""
        bx = np.ascontiguousarray(boxes, dtype=np.int32).ravel()
        _check(load().qsr_set_boxes_active(self._h, bx.ctypes.data_as(C.POINTER(C.c_int)), bx.size,
                                           int(bool(active))))

    def work(self) -> dict:
        a, t, b = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(load().qsr_step_work(self._h, C.byref(a), C.byref(t), C.byref(b)))
        return {"active_voxels": a.value, "tile_voxels": t.value, "box_voxels": b.value}


def S_eq(a, b, c, T, T_star, modechoice=1) -> float:
    return load().qsr_S_eq(a, b, c, T, T_star, modechoice)


DT_KEYS = ("min_dx", "S_scale", "L_eff_total", "D_total", "dt_diff_total", "L_eff_explicit", "dt_diff",
           "bulk_rate", "dt_bulk", "dt_max")


def dt_estimate(dx, dy, dz, gamma, a, b, c, T_star, modechoice, kappa, L1, L2, L3, S_ref_used, S0, T_a, T_b,
                use_semi_implicit=False, L_stab=0.0, jacobi_iters=0) -> dict:
    out = (C.c_double * 10)()
    _check(load().qsr_dt_estimate(dx, dy, dz, gamma, a, b, c, T_star, modechoice, kappa, L1, L2, L3, S_ref_used,
                                  S0, T_a, T_b, int(use_semi_implicit), L_stab, jacobi_iters, out))
    return dict(zip(DT_KEYS, out))


def init_field(p: Params, mode: int, S0: float = 0.5, noise_amplitude: float = 1e-3, seed: int = 12345) -> np.ndarray:
    """Seeded initial state (QSR.cu:1841-1904): 0 random director, 1 radial, 2 isotropic noise."""
    Q = np.empty(p.dims + (5,))
    _check(load().qsr_init_field(C.byref(p.to_c()), mode, S0, noise_amplitude, seed,
                                 Q.ctypes.data_as(C.POINTER(C.c_double)), p.n))
    return Q


def load_qtensor(path, dims):
    """Read a Qtensor_output_*.dat file (QSR.cu:1818-1832 format) into an (Nx, Ny, Nz, 5) array,
    as QSRvis.py:2274-2304 reads it; returns (Q, rows_read)."""
    Nx, Ny, Nz = (int(x) for x in dims)
    Q = np.zeros((Nx, Ny, Nz, 5))
    rows = C.c_uint64()
    _check(load().qsr_load_qtensor_file(os.fsencode(str(path)), Nx, Ny, Nz, Q.ctypes.data_as(C.POINTER(C.c_double)),
                                        Nx * Ny * Nz, C.byref(rows)))
    return Q, rows.value


def save_qtensor(path, Q):
    """Write an (Nx, Ny, Nz, 5) field in the Qtensor_output_*.dat format (QSR.cu:1818-1832)."""
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    Nx, Ny, Nz = Q.shape[:3]
    _check(load().qsr_save_qtensor_file(os.fsencode(str(path)), Q.ctypes.data_as(C.POINTER(C.c_double)), Nx, Ny, Nz,
                                        Nx * Ny * Nz))


def _fptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


class Surrogate:
    """The learned warm-start step: a 5 -> 32 -> 32 -> 5 channel 3x3x3 convolution
    network on the GPU (BF16 matrix cores, FP32 accumulation) that predicts
    dQ = Q(k+N) - Q(k) from Q(k).  Trained with its own backward and Adam kernels
    on field pairs a Solver produced (dataset_add / train); apply() proposes the
    step to a single-box Solver and keeps it only if neither the free energy nor
    the explicit-Euler residual rises (SURR_CHECKED)."""

    def __init__(self, device: int = 0, seed: int | None = 1):
        L = load()
        h = C.c_void_p()
        _check(L.qsr_surr_create(device, C.byref(h)))
        self._h = h
        n = C.c_size_t()
        _check(L.qsr_surr_num_params(C.byref(n)))
        self.nparams = n.value
        if seed is not None:
            _check(L.qsr_surr_init_params(self._h, int(seed)))

    def close(self):
        if getattr(self, "_h", None):
            load().qsr_surr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def params(self) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        _check(load().qsr_surr_get_params(self._h, _fptr(out), self.nparams))
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32).ravel()
        _check(load().qsr_surr_set_params(self._h, _fptr(p), p.size))

    def save(self, path):
        _check(load().qsr_surr_save(self._h, os.fsencode(str(path))))

    def load(self, path):
        _check(load().qsr_surr_load(self._h, os.fsencode(str(path))))

    def eval(self, x5, target5=None, mask=None, want_pred=True, want_grad=False, modulator=None):
        """x5, target5: (n, X, Y, Z, 5); mask: (n, X, Y, Z).  Returns (loss, pred5, grad).
        With ``modulator`` g (same shape as x5) the prediction is g (1 + y), else y."""
        x5 = np.ascontiguousarray(x5, dtype=np.float32)
        nb, X, Y, Z, c5 = x5.shape
        if c5 != 5:
            raise ValueError("last axis must hold the 5 Q components")
        t5 = None if target5 is None else np.ascontiguousarray(target5, dtype=np.float32)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32)
        if t5 is not None and t5.shape != x5.shape:
            raise ValueError("target shape mismatch")
        if m is not None and m.shape != x5.shape[:4]:
            raise ValueError("mask shape mismatch")
        pred = np.empty_like(x5) if want_pred else None
        grad = np.empty(self.nparams, dtype=np.float32) if want_grad else None
        want_loss = t5 is not None or want_grad
        loss = C.c_double(0.0)
        if modulator is None:
            _check(load().qsr_surr_eval(self._h, nb, X, Y, Z, _fptr(x5), _fptr(t5), _fptr(m),
                                        C.byref(loss) if want_loss else None, _fptr(pred), _fptr(grad)))
        else:
            g = np.ascontiguousarray(modulator, dtype=np.float32)
            if g.shape != x5.shape:
                raise ValueError("modulator shape mismatch")
            _check(load().qsr_surr_eval_gain(self._h, nb, X, Y, Z, _fptr(x5), _fptr(g), _fptr(t5), _fptr(m),
                                             C.byref(loss) if want_loss else None, _fptr(pred), _fptr(grad)))
        return (loss.value if want_loss else None), pred, grad

    def adam_step(self, grad, lr: float):
        g = np.ascontiguousarray(grad, dtype=np.float32).ravel()
        if g.size != self.nparams:
            raise ValueError("gradient size mismatch")
        _check(load().qsr_surr_adam_step(self._h, _fptr(g), float(lr)))

    def set_input(self, kind: int, horizon: float = 1.0):
        """SURR_INPUT_FIELD: the network reads Q and predicts dQ = y.  SURR_INPUT_INCREMENT
        and SURR_INPUT_GAIN: it modulates u, one explicit step's increment of Q:
        dQ = horizon * u * (1 + y), reading u / s (s the RMS of u per component) or Q."""
        _check(load().qsr_surr_set_input(self._h, int(kind), float(horizon)))

    def set_tiles(self, mode: int):
        """SURR_TILES_DROPLET (the default): the application's network pass computes the tiles
        within its reach of a voxel a step can change; SURR_TILES_ALL: every tile of the box.
        The proposals are the same."""
        _check(load().qsr_surr_set_tiles(self._h, int(mode)))

    def dataset_add(self, solver: "Solver", box: int = 0, increment: bool = False) -> int:
        """Snapshot a box's field (and, with ``increment``, one explicit step's increment
        from it, which the input snapshot of a pair needs under increment input)."""
        i = C.c_int()
        fn = load().qsr_surr_dataset_add_increment if increment else load().qsr_surr_dataset_add
        _check(fn(self._h, solver._h, box, C.byref(i)))
        return i.value

    def dataset_size(self) -> int:
        i = C.c_int()
        _check(load().qsr_surr_dataset_size(self._h, C.byref(i)))
        return i.value

    def dataset_clear(self):
        _check(load().qsr_surr_dataset_clear(self._h))

    def train(self, pairs, iters: int, batch: int = 4, crop: int = 32, lr: float = 1e-3, seed: int = 7,
              weight_floor: float = 1.0) -> np.ndarray:
        """Adam on random crops of (input, target) snapshot pairs, each pair weighted by
        1 / max(mean square of its target, weight_floor); returns the per-iteration loss."""
        pr = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
        hist = np.zeros(int(iters))
        _check(load().qsr_surr_train(self._h, pr.ctypes.data_as(C.POINTER(C.c_int)), pr.shape[0], int(iters),
                                     int(batch), int(crop), float(lr), int(seed), float(weight_floor),
                                     hist.ctypes.data_as(C.POINTER(C.c_double))))
        return hist

    def apply(self, solver: "Solver", mode: int = SURR_CHECKED, scales=None) -> dict:
        """Propose Q + dQ (or, with ``scales``, Q + scales[i] dQ in turn, keeping the first
        that passes the check) on a single-box solver."""
        r = _SurrResult()
        if scales is None:
            _check(load().qsr_surr_apply(self._h, solver._h, int(mode), C.byref(r)))
        else:
            sc = np.ascontiguousarray(scales, dtype=np.float64).ravel()
            _check(load().qsr_surr_apply_ladder(self._h, solver._h, int(mode),
                                                sc.ctypes.data_as(C.POINTER(C.c_double)), sc.size, C.byref(r)))
        return {f: getattr(r, f) for f, _ in _SurrResult._fields_}

    def apply_boxes(self, solver: "Solver", boxes, mode: int = SURR_CHECKED, scales=(1.0,)) -> list:
        """The ladder of :meth:`apply` on each listed box of a batched solver, each kept or
        restored on its own; one result dict per listed box."""
        bx = np.ascontiguousarray(boxes, dtype=np.int32).ravel()
        sc = np.ascontiguousarray(scales, dtype=np.float64).ravel()
        res = (_SurrResult * max(bx.size, 1))()
        _check(load().qsr_surr_apply_boxes(self._h, solver._h, int(mode), sc.ctypes.data_as(C.POINTER(C.c_double)),
                                           sc.size, bx.ctypes.data_as(C.POINTER(C.c_int)), bx.size, res))
        return [{f: getattr(res[i], f) for f, _ in _SurrResult._fields_} for i in range(bx.size)]
