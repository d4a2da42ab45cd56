"""Q-tensor text files: the writer produces the saveQTensorToFile format
(QSR.cu:1818-1832) and the warm-start reader reads files the way the
reference's own reader does (QSRvis.py:2274-2304).  CPU only."""
import os
import re

import numpy as np
import pytest

import lcdrsimulator_amd as q

ROW = re.compile(r"\d+ \d+ \d+( -?\d\.\d{6}e[+-]\d\d){9}")


def _full(Qv):
    xx, xy, xz, yy, yz = Qv
    return [xx, xy, xz, xy, yy, yz, xz, yz, -(xx + yy)]


def test_round_trip_and_format(tmp_path):
    rng = np.random.default_rng(3)
    dims = (5, 7, 6)
    Q = rng.normal(0, 0.3, dims + (5,))
    path = tmp_path / "Qtensor_output_final.dat"
    q.save_qtensor(path, Q)
    lines = path.read_text().splitlines()
    assert len(lines) == np.prod(dims)
    assert all(ROW.fullmatch(ln) for ln in lines)
    # i-major order, z fastest, as the reference writes
    assert lines[0].startswith("0 0 0 ") and lines[1].startswith("0 0 1 ") and lines[6].startswith("0 1 0 ")
    cols = np.array([[float(x) for x in ln.split()] for ln in lines])
    assert np.array_equal(cols[:, 4], cols[:, 6]) and np.allclose(cols[:, 11], -(cols[:, 3] + cols[:, 7]), atol=2e-6)
    back, rows = q.load_qtensor(path, dims)
    assert rows == np.prod(dims)
    assert np.abs(back - Q).max() <= 5e-7 * np.abs(Q).max() + 1e-12  # 7 significant digits in the file


def test_reader_follows_qsrvis_rules(tmp_path):
    """Comments skipped, rows in any order, off-diagonals averaged, rows outside the grid ignored,
    missing voxels zero."""
    dims = (3, 3, 3)
    rng = np.random.default_rng(4)
    want = np.zeros(dims + (5,))
    rows = []
    for i in range(3):
        for j in range(3):
            for k in range(3):
                if (i, j, k) == (1, 1, 1):
                    continue  # missing voxel
                v = rng.normal(0, 0.2, 5)
                full = _full(v)
                full[1] += 0.01   # Qxy != Qyx: the reader takes the mean
                full[3] -= 0.01
                want[i, j, k] = v
                rows.append(f"{i} {j} {k} " + " ".join(f"{x:.9e}" for x in full))
    rows.append("7 0 0 " + " ".join(["1.0"] * 9))  # outside the grid
    rng.shuffle(rows)
    text = "# Qtensor snapshot\n" + "\n".join(rows[:10]) + "\n   # indented comment\n" + "\n".join(rows[10:]) + "\n"
    path = tmp_path / "q.dat"
    path.write_text(text)
    got, nrows = q.load_qtensor(path, dims)
    assert nrows == 27  # 26 in the grid + 1 outside
    assert np.allclose(got, want, rtol=0, atol=1e-9)  # 10 significant digits written
    assert np.all(got[1, 1, 1] == 0)


def test_reader_errors(tmp_path):
    with pytest.raises(q.QSRError, match="cannot open"):
        q.load_qtensor(tmp_path / "missing.dat", (2, 2, 2))
    bad = tmp_path / "bad.dat"
    bad.write_text("0 0 0 1 2 3 4 5\n")
    with pytest.raises(q.QSRError, match="fewer than 12"):
        q.load_qtensor(bad, (2, 2, 2))


def test_large_file_parallel_blocks(tmp_path):
    """A file larger than one parse block (64 MiB) is read whole and identically."""
    dims = (96, 96, 96)
    rng = np.random.default_rng(5)
    Q = rng.normal(0, 0.3, dims + (5,)).astype(np.float32).astype(np.float64)
    path = tmp_path / "big.dat"
    q.save_qtensor(path, Q)
    assert os.path.getsize(path) > (64 << 20)
    back, rows = q.load_qtensor(path, dims)
    assert rows == np.prod(dims)
    assert np.abs(back - Q).max() <= 5e-7 * np.abs(Q).max()
