"""Write golden vectors from the reference's own QSR.cpp functions.

Runs in the build container only (it needs oracle/_ref/libqsr_ref.so, which
oracle/build_ref.sh compiles from the reference's QSR.cpp).  Inputs are seeded
synthetic fields; outputs are what QSR.cpp's computeLaplacian (QSR.cpp:178-282),
computeChemicalPotential (QSR.cpp:559-826), computeEnergyComponents
(QSR.cpp:410-557), calculateNematicField (QSR.cpp:10-61) and updateQTensor
(QSR.cpp:861-891) return for them.
The .npz holds data only (inputs and expected outputs).  computeEnergyComponents
sums with OpenMP reductions, so regenerated energies may differ from the
committed ones in the last bit; every other array is reproduced exactly.

    python tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

OUT = os.path.join(HERE, "qsr_cpp_vectors.npz")
GRIDS = [((9, 8, 7), (1e-8, 1e-8, 1e-8)), ((11, 10, 12), (1e-8, 1.1e-8, 0.9e-8))]
# (kappa, L1, L2, L3)
ELASTIC = [(6.5e-12, 0, 0, 0), (0, 6e-12, 0, 0), (0, 6e-12, 3e-12, 0), (0, 6e-12, 3e-12, 1e-12)]
BULK = dict(a=0.044e6, b=1.413e6, c=1.153e6, T=300.0, T_star=307.0)
# (dims, W): strong (W = 0, shell skipped) and weak anchoring (evolve band R + 1.5)
UPDATE = [((12, 11, 10), 0.0), ((12, 11, 10), 1e-4), ((9, 13, 11), 0.0)]
UPDATE_DT, UPDATE_GAMMA = 1.60372e-10, 0.1


def reference_shell(dims):
    """QSR.cpp's shell mask: |r - R| < 1 about (N/2), R = min(N)/2 (QSR.cpp:1088-1102)."""
    Nx, Ny, Nz = dims
    I, J, K = np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing="ij")
    r = np.sqrt((I - Nx / 2.0) ** 2 + (J - Ny / 2.0) ** 2 + (K - Nz / 2.0) ** 2)
    return (np.abs(r - min(dims) / 2.0) < 1.0).astype(np.uint8)


def nematic_inputs(rng):
    """Uniaxial, oblate, biaxial, near-degenerate and near-isotropic tensors."""
    qs = []
    for _ in range(24):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        S = rng.uniform(-0.5, 1.0)
        Q = S * (np.outer(n, n) - np.eye(3) / 3)
        qs.append([Q[0, 0], Q[0, 1], Q[0, 2], Q[1, 1], Q[1, 2]])
    for _ in range(16):
        qs.append(list(rng.normal(0, 0.3, 5)))
    for s in (1e-10, 5e-10, 2e-9, 1e-8):
        qs.append(list(rng.normal(0, s, 5)))
    qs.append([0.2, 0.0, 0.0, 0.2, 0.0])       # two equal largest eigenvalues
    qs.append([-0.1, 0.0, 0.0, -0.1, 0.0])     # prolate along z
    qs.append([0.0, 0.0, 0.0, 0.0, 0.0])
    return np.array(qs, dtype=np.float64)


def main():
    R = O.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libqsr_ref.so missing: run `bash oracle/build_ref.sh` first")
    dp = C.POINTER(C.c_double)
    rng = np.random.default_rng(20261203)
    data = {}
    for gi, (dims, d) in enumerate(GRIDS):
        Q = rng.normal(0, 0.3, dims + (5,))
        data[f"g{gi}_dims"] = np.array(dims)
        data[f"g{gi}_d"] = np.array(d)
        data[f"g{gi}_Q"] = Q
        lap = np.zeros_like(Q)
        R.ref_laplacian(Q.ctypes.data_as(dp), lap.ctypes.data_as(dp), *dims, *(C.c_double(v) for v in d))
        data[f"g{gi}_lap"] = lap
        for ei, (kap, L1, L2, L3) in enumerate(ELASTIC):
            pr = (C.c_double * 9)(BULK["a"], BULK["b"], BULK["c"], BULK["T"], BULK["T_star"], L1, L2, L3, 0.0)
            mu = np.zeros_like(Q)
            R.ref_chempot(Q.ctypes.data_as(dp), lap.ctypes.data_as(dp), mu.ctypes.data_as(dp), *dims,
                          *(C.c_double(v) for v in d), pr, C.c_double(kap), 1)
            data[f"g{gi}_e{ei}_mu"] = mu
            e = (C.c_double * 2)()
            R.ref_energy(Q.ctypes.data_as(dp), *dims, *(C.c_double(v) for v in d), pr, C.c_double(kap), 1, e)
            data[f"g{gi}_e{ei}_energy"] = np.array(list(e))
    data["elastic"] = np.array(ELASTIC)
    data["bulk"] = np.array([BULK[k] for k in ("a", "b", "c", "T", "T_star")])
    nq = nematic_inputs(rng)
    nf = np.zeros((nq.shape[0], 4))
    R.ref_nematic(nq.ctypes.data_as(dp), C.c_size_t(nq.shape[0]), nf.ctypes.data_as(dp))
    data["nematic_in"] = nq
    data["nematic_out"] = nf
    # updateQTensor: Q -= (dt/gamma) mu inside the reference's evolve mask, strong shell skipped.
    # Drawn after every other input so the earlier vectors do not change.
    for ui, (dims, W) in enumerate(UPDATE):
        Q = rng.normal(0, 0.3, dims + (5,))
        mu = rng.normal(0, 1e6, dims + (5,))
        shell = reference_shell(dims)
        out = Q.copy()
        R.ref_update(out.ctypes.data_as(dp), mu.ctypes.data_as(dp), *dims, C.c_double(UPDATE_DT),
                     C.c_double(UPDATE_GAMMA), C.c_double(W), shell.ctypes.data_as(C.POINTER(C.c_ubyte)))
        data[f"u{ui}_dims"] = np.array(dims)
        data[f"u{ui}_W"] = np.array(W)
        data[f"u{ui}_Q"] = Q
        data[f"u{ui}_mu"] = mu
        data[f"u{ui}_shell"] = shell
        data[f"u{ui}_out"] = out
    data["update_dt_gamma"] = np.array([UPDATE_DT, UPDATE_GAMMA])
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
