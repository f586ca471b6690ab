"""Shared builders for the tests (inputs only; the oracle is the checker)."""
import numpy as np

from oracle.pyoracle import Csr, csr_transpose


def random_csr(rng, rows, cols, min_len=0, max_len=30, long_rows=(), sort=True):
    """Random CSR with ragged, empty and optionally very long rows."""
    lens = rng.integers(min_len, max_len + 1, size=rows)
    for r, ln in long_rows:
        lens[r] = ln
    rp = np.zeros(rows + 1, np.uint64)
    rp[1:] = np.cumsum(lens)
    col = np.empty(int(rp[-1]), np.uint64)
    for r in range(rows):
        c = rng.choice(cols, size=int(lens[r]), replace=bool(lens[r] > cols))
        if sort:
            c = np.sort(c)
        col[int(rp[r]):int(rp[r + 1])] = c
    val = rng.standard_normal(col.size)
    A = Csr(rows, cols, val, col, rp)
    validate_csr(A)
    return A


def validate_csr(A):
    """Host-side shape check before any kernel sees the matrix."""
    rp = A.rowptr.astype(np.int64)
    assert rp.size == A.rows + 1 and rp[0] == 0 and np.all(np.diff(rp) >= 0)
    assert rp[-1] == A.col.size == A.val.size
    assert A.col.size == 0 or int(A.col.max()) < A.cols


def q1_levels(F, dim, lf, lc):
    """Structured Q1 hierarchy from the product's setup code (finest first)."""
    levels = []
    for lvl in range(lf, lc - 1, -1):
        q = F.q1_problem(dim, lvl)
        lv = dict(q)
        lv["A"] = Csr(q["n"], q["n"], q["val"], q["col"], q["rowptr"])
        if lvl > lc:
            P = Csr(q["n"], q["pcols"], q["pval"], q["pcol"], q["prowptr"])
            lv["P"] = P
            lv["R"] = csr_transpose(P)
        levels.append(lv)
    return levels
