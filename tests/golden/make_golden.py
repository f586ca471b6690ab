"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Runs only where the reference checkout is present (it loads
oracle/_ref/libfeatref.so, built by `make -C oracle ref` from the reference's
sources).  Every fixture is data: inputs and the reference's outputs.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.pyoracle import Reference, Csr, csr_transpose, COARSE_JACOBI  # noqa: E402


def q1_levels_from_product(dim, lf, lc):
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "feat3amd", os.path.join(ROOT, "feat3.doxytest_amd", "__init__.py"))
    F = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(F)
    levels = []
    for lvl in range(lf, lc - 1, -1):
        q = F.q1_problem(dim, lvl)
        lv = dict(A=Csr(q["n"], q["n"], q["val"], q["col"], q["rowptr"]), filt=q["filt"])
        if lvl > lc:
            P = Csr(q["n"], q["pcols"], q["pval"], q["pcol"], q["prowptr"])
            lv["P"] = P
            lv["R"] = csr_transpose(P)
        levels.append(lv)
    return levels, F


def main():
    R = Reference()
    out = {}
    # xorshift64* uniform values (random.hpp) from the default and another seed
    out["rng_default_m1p1"] = R.rng_fill(28054777172512, 64, -1.0, 1.0)
    out["rng_seed17_0p1"] = R.rng_fill(17, 64, 0.0, 1.0)

    # FD pointstar fixture (pointstar_factory.hpp) and the basic solver runs
    A, q2 = R.pointstar_fd(17, 2)
    out["fd_val"], out["fd_col"], out["fd_rowptr"], out["fd_q2bubble"] = A.val, A.col, A.rowptr, q2
    b = R.csr_apply(A, q2)
    x0 = np.zeros(A.rows)
    for name, kind, pc, mi in [("cg", "pcg", 0, 100), ("pcgjac", "pcg", 1, 100),
                               ("richjac", "richardson", 1, 1500)]:
        r = R.single_solve(kind, A, x0, b, precond=pc, apply=True, max_iter=mi)
        out[f"fd_{name}_iters"] = np.array([r["iters"]])
        out[f"fd_{name}_x"] = r["x"]
        out[f"fd_{name}_hist"] = r["hist"]

    # CSR apply known answers of sparse_matrix_csr-test.cpp:358-452 at n = 1..512
    for n in [1, 2, 8, 64, 512]:
        rows = []
        for i in range(n):
            for j in (i - 1, i, i + 1):
                if 0 <= j < n:
                    rows.append((i, j, 2.0 if i == j else -1.0))
        val = np.array([v for _, _, v in rows])
        col = np.array([j for _, j, _ in rows], np.uint64)
        rp = np.zeros(n + 1, np.uint64)
        for i, _, _ in rows:
            rp[i + 1] += 1
        rp = np.cumsum(rp).astype(np.uint64)
        T = Csr(n, n, val, col, rp)
        x = (np.arange(n) % 100) * 1.234
        y = 2.0 - (np.arange(n) % 42)
        for tag, a in [("0", 0.0), ("m1", -1.0), ("s", 4711.1)]:
            out[f"kat{n}_a{tag}"] = R.csr_apply(T, x, y, a, 1.0)
        out[f"kat{n}_ax"] = R.csr_apply(T, x)
        out[f"kat{n}_trs"] = R.csr_apply(T, x, y, 4711.1, 1.0, transposed=True)

    # Gauss-Jordan inverse (util/math.hpp:1169) of a fixed random 9x9 matrix
    rng = np.random.default_rng(123)
    M = rng.uniform(-1, 1, (9, 9)) + 4 * np.eye(9)
    inv, det = R.invert_matrix(M)
    out["inv_in"], out["inv_out"], out["inv_det"] = M, inv, np.array([det])

    # Reference FE assembly of Q1 on level 3 (2D) and 2 (3D): matrix, filter,
    # vertex coordinates, prolongation from the level below
    for dim, lvl in [(2, 3), (3, 2)]:
        q = R.q1_poisson(dim, lvl)
        p = f"fe{dim}d_l{lvl}_"
        out[p + "val"], out[p + "col"], out[p + "rowptr"] = q["A"].val, q["A"].col, q["A"].rowptr
        out[p + "filt"], out[p + "coords"] = q["filt"], q["coords"]
        out[p + "pval"], out[p + "pcol"], out[p + "prowptr"] = q["P"].val, q["P"].col, q["P"].rowptr
        qc = R.q1_poisson(dim, lvl - 1, with_prol=False)
        out[p + "coarse_coords"] = qc["coords"]

    # Reference PCG(MG-V(4,4) Richardson-Jacobi 0.5) on the product's C1 hierarchy
    levels, F = q1_levels_from_product(2, 7, 1)
    n = levels[0]["A"].rows
    b = F.q1_rhs_sine(2, 7, n)
    b[levels[0]["filt"].astype(np.int64)] = 0.0
    r = R.mg_pcg_solve(levels, np.zeros(n), b, coarse=COARSE_JACOBI)
    out["c1_rhs"] = b
    out["c1_iters"] = np.array([r["iters"]])
    out["c1_hist"] = r["hist"]
    out["c1_x"] = r["x"]

    path = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes,", len(out), "arrays")


if __name__ == "__main__":
    main()
