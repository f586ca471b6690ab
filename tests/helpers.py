"""Shared builders for the tests (inputs only; the oracle is the checker)."""
import numpy as np

from oracle.pyoracle import Bcsr, Csr, csr_transpose


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


def stokes_levels(F, lf, lc):
    """Q2/P1dc lid-driven cavity hierarchy from the product's setup code (finest first)."""
    levels = []
    for lvl in range(lf, lc - 1, -1):
        s = F.stokes_problem(lvl)
        nv, np_ = s["nv"], s["np"]
        lv = dict(level=lvl, nv=nv, np=np_, filt=s["filt"], filt_val=s["filt_val"], prim=s["prim"],
                  dual=s["dual"],
                  A=Bcsr(nv, nv, 2, 2, *s["A"]), B=Bcsr(nv, np_, 2, 1, *s["B"]), D=Bcsr(np_, nv, 1, 2, *s["D"]))
        if lvl > lc:
            pv, pp = s["Pv"], s["Pp"]
            lv["Pv"] = Csr(nv, pv[3], pv[0], pv[1], pv[2])
            lv["Pp"] = Csr(np_, pp[3], pp[0], pp[1], pp[2])
            lv["Rv"] = csr_transpose(lv["Pv"])
            lv["Rp"] = csr_transpose(lv["Pp"])
        levels.append(lv)
    return levels


def stokes_rhs(lv, oracle):
    """Lid-driven cavity: zero force, filter_rhs; x0 = filter_sol(0)."""
    n = 2 * lv["nv"] + lv["np"]
    b = oracle.stokes_filter(lv, np.zeros(n), "rhs")
    x0 = oracle.stokes_filter(lv, np.zeros(n), "sol")
    return x0, b
