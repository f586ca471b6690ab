"""Writes tests/golden/io/*: small LAFEM files written by the reference's own
SparseMatrixCSR::write_out / DenseVector::write_out (oracle/_ref, built from the
reference checkout by oracle/Makefile), for the file-format tests to run where the
reference build is absent.

    python tests/golden/make_io_fixtures.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.pyoracle import Csr, Reference  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "io")
FM = {"exp": 0, "dv": 1, "mtx": 2, "csr": 4}


def fixture_matrix():
    """7 x 5 matrix, every row non-empty, values of mixed sign and magnitude."""
    rng = np.random.default_rng(20240611)
    rows, cols = 7, 5
    rp, col, val = [0], [], []
    for r in range(rows):
        c = np.sort(rng.choice(cols, size=int(rng.integers(1, cols + 1)), replace=False))
        col.extend(int(x) for x in c)
        val.extend(float(v) for v in rng.standard_normal(c.size) * 10.0 ** rng.integers(-3, 4, c.size))
        rp.append(len(col))
    return Csr(rows, cols, np.array(val), np.array(col, np.uint64), np.array(rp, np.uint64))


def fixture_symmetric():
    """6 x 6 symmetric matrix (1D Laplacian plus a corner coupling)."""
    n = 6
    A = np.diag(np.full(n, 2.0)) + np.diag(np.full(n - 1, -1.0), 1) + np.diag(np.full(n - 1, -1.0), -1)
    A[0, n - 1] = A[n - 1, 0] = 0.125
    rp, col, val = [0], [], []
    for r in range(n):
        nz = np.nonzero(A[r])[0]
        col.extend(int(c) for c in nz)
        val.extend(float(A[r, c]) for c in nz)
        rp.append(len(col))
    return Csr(n, n, np.array(val), np.array(col, np.uint64), np.array(rp, np.uint64))


def fixture_vector():
    rng = np.random.default_rng(7)
    return rng.standard_normal(9) * 10.0 ** rng.integers(-5, 6, 9)


def main():
    ref = Reference()
    os.makedirs(OUT, exist_ok=True)
    A, S, v = fixture_matrix(), fixture_symmetric(), fixture_vector()
    ref.csr_write(FM["csr"], os.path.join(OUT, "matrix.csr"), A)
    ref.csr_write(FM["mtx"], os.path.join(OUT, "matrix_general.mtx"), A)
    ref.csr_write(FM["csr"], os.path.join(OUT, "symmetric.csr"), S)
    ref.csr_write(FM["mtx"], os.path.join(OUT, "symmetric_lower.mtx"), S, symmetric=True)
    ref.dv_write(FM["dv"], os.path.join(OUT, "vector.dv"), v)
    ref.dv_write(FM["mtx"], os.path.join(OUT, "vector.mtx"), v)
    ref.dv_write(FM["exp"], os.path.join(OUT, "vector.exp"), v)
    for name in sorted(os.listdir(OUT)):
        print(name, os.path.getsize(os.path.join(OUT, name)))


if __name__ == "__main__":
    main()
