"""The SOR / SSOR oracles, pinned.

* The generic sequential SOR / SSOR restatement equals the reference's own
  SORPrecond / SSORPrecond<generic> (oracle/_ref) exactly; so does the ILU(0)
  restatement (ILUPrecond<generic>, p = 0), whose factors also satisfy
  (L U)_ij = a_ij on the pattern of A (to rounding).

* With one colour per row (processed in row order) the multicolour SOR sweep is the
  reference's generic sequential SOR: compared with the reference's own
  SORPrecond<generic> (oracle/_ref), exactly.
* For a real colouring the oracle's SOR / SSOR equal the closed forms
      SOR:  y = (D/omega + L_e)^{-1} x
      SSOR: y = omega (2 - omega) (D + omega U_e)^{-1} (D - omega U_l) (D + omega L_e)^{-1} x
  with L_e / U_e the lower / upper entries whose colour is processed earlier and U_l the
  upper entries of colours processed later (dense solves, relative 1e-12).
* The colouring is proper (no coupling inside a colour) and ordered by ascending row count.
"""
import numpy as np
import pytest

from oracle.pyoracle import Csr
from tests.helpers import random_csr


def dominant(rng, n, max_len=9):
    """Random sorted CSR with a dominant positive diagonal."""
    A = random_csr(rng, n, n, min_len=1, max_len=max_len)
    rp, col, val = [0], [], []
    for r in range(n):
        c = set(int(x) for x in A.col[int(A.rowptr[r]):int(A.rowptr[r + 1])]) | {r}
        c = sorted(c)
        v = rng.standard_normal(len(c))
        v[c.index(r)] = 2.0 + np.abs(v).sum()
        col.extend(c)
        val.extend(v.tolist())
        rp.append(len(col))
    return Csr(n, n, np.array(val), np.array(col, np.uint64), np.array(rp, np.uint64))


def dense(A):
    M = np.zeros((A.rows, A.cols))
    for r in range(A.rows):
        for k in range(int(A.rowptr[r]), int(A.rowptr[r + 1])):
            M[r, int(A.col[k])] = A.val[k]
    return M


def q1_csr(feat, dim, level):
    q = feat.q1_problem(dim, level)
    return Csr(q["n"], q["n"], q["val"], q["col"], q["rowptr"])


@pytest.mark.parametrize("seed,n", [(1, 40), (2, 200)])
def test_row_colors_sor_is_reference_generic_sor(oracle, reference, seed, n):
    rng = np.random.default_rng(seed)
    A = dominant(rng, n)
    x = rng.standard_normal(n)
    rank, nc = oracle.color_ranks(A, row_colors=True)
    assert nc == n and np.array_equal(rank, np.arange(n))
    for omega in (1.0, 0.7):
        y = oracle.mcsor_apply(A, rank, nc, omega, x)
        assert np.array_equal(y, reference.precond_apply("sor", A, omega, x))


def test_row_colors_sor_q1(feat, oracle, reference):
    A = q1_csr(feat, 2, 3)
    x = np.random.default_rng(3).standard_normal(A.rows)
    rank, nc = oracle.color_ranks(A, row_colors=True)
    assert np.array_equal(oracle.mcsor_apply(A, rank, nc, 1.2, x), reference.precond_apply("sor", A, 1.2, x))


def check_coloring(A, rank, nc):
    counts = np.bincount(rank, minlength=nc)
    assert np.all(counts > 0) and np.all(np.diff(counts) >= 0)      # ascending row counts
    for r in range(A.rows):
        for k in range(int(A.rowptr[r]), int(A.rowptr[r + 1])):
            c = int(A.col[k])
            assert c == r or rank[c] != rank[r]


@pytest.mark.parametrize("case", ["random", "q1_2d", "q1_3d"])
def test_multicolor_closed_forms(feat, oracle, case):
    rng = np.random.default_rng(17)
    A = dominant(rng, 120) if case == "random" else q1_csr(feat, 2 if case == "q1_2d" else 3, 3 if case == "q1_2d" else 2)
    rank, nc = oracle.color_ranks(A)
    check_coloring(A, rank, nc)
    if case == "q1_3d":
        assert nc == 8
    if case == "q1_2d":
        assert nc == 4
    M = dense(A)
    n = A.rows
    i, j = np.indices((n, n))
    D = np.diag(np.diag(M))
    earlier = rank[j] < rank[i]
    Le = np.where((j < i) & earlier, M, 0.0)
    Ue = np.where((j > i) & earlier, M, 0.0)
    Ul = np.where((j > i) & ~earlier & (rank[j] != rank[i]), M, 0.0)
    x = rng.standard_normal(n)
    for omega in (1.0, 1.3):
        y = oracle.mcsor_apply(A, rank, nc, omega, x)
        want = np.linalg.solve(D / omega + Le, x)
        assert np.linalg.norm(y - want) <= 1e-12 * np.linalg.norm(want)
        y = oracle.mcsor_apply(A, rank, nc, omega, x, ssor=True)
        yf = np.linalg.solve(D + omega * Le, x)
        want = omega * (2 - omega) * np.linalg.solve(D + omega * Ue, (D - omega * Ul) @ yf)
        assert np.linalg.norm(y - want) <= 1e-12 * np.linalg.norm(want)


def test_planted_error_is_caught(oracle):
    """A colour order reversed (a plausible slip) changes the SOR result."""
    rng = np.random.default_rng(5)
    A = dominant(rng, 60)
    rank, nc = oracle.color_ranks(A)
    x = rng.standard_normal(A.rows)
    y = oracle.mcsor_apply(A, rank, nc, 1.0, x)
    y_rev = oracle.mcsor_apply(A, (nc - 1 - rank).astype(np.uint32), nc, 1.0, x)
    assert not np.allclose(y, y_rev)


@pytest.mark.parametrize("case", ["random", "q1_2d", "q1_3d"])
def test_generic_sor_ssor_is_reference(feat, oracle, reference, case):
    rng = np.random.default_rng(31)
    A = dominant(rng, 150) if case == "random" else q1_csr(feat, 2 if case == "q1_2d" else 3, 4 if case == "q1_2d" else 2)
    x = rng.standard_normal(A.rows)
    for omega in (1.0, 0.75, 1.4):
        assert np.array_equal(oracle.sor_generic_apply(A, omega, x), reference.precond_apply("sor", A, omega, x))
        assert np.array_equal(oracle.sor_generic_apply(A, omega, x, ssor=True),
                              reference.precond_apply("ssor", A, omega, x))


@pytest.mark.parametrize("case", ["random", "q1_2d", "q1_3d"])
def test_ilu0_is_reference(feat, oracle, reference, case):
    rng = np.random.default_rng(41)
    A = dominant(rng, 150) if case == "random" else q1_csr(feat, 2 if case == "q1_2d" else 3, 4 if case == "q1_2d" else 2)
    f = oracle.ilu0_factor(A)
    for seed in range(3):
        x = np.random.default_rng(seed).standard_normal(A.rows)
        assert np.array_equal(oracle.ilu0_apply(A, f, x), reference.precond_apply("ilu", A, 0, x))
    # ILU(0) property: (I+L)(D+U) reproduces A on A's pattern
    l, u, dinv, rl, ru = f
    n = A.rows
    Lm, Um = np.eye(n), np.diag(1.0 / dinv)
    jl = ju = 0
    for i in range(n):
        for k in range(int(A.rowptr[i]), int(A.rowptr[i + 1])):
            c = int(A.col[k])
            if c < i:
                Lm[i, c] = l[jl]
                jl += 1
            elif c > i:
                Um[i, c] = u[ju]
                ju += 1
    LU, M = Lm @ Um, dense(A)
    mask = M != 0
    assert np.allclose(LU[mask], M[mask], rtol=1e-12, atol=1e-12 * np.abs(M).max())


def brute_level_pattern(A, p):
    """Level-of-fill pattern by the textbook rule on the dense index grid (lev(i,j) = 0 on
    A's pattern, then lev(i,j) = min(lev(i,j), lev(i,k) + lev(k,j) + 1) for k < min(i, j),
    rows in order): the entries of level <= p, as a set of (i, j) with i != j."""
    n = A.rows
    INF = 1 << 30
    lev = np.full((n, n), INF, dtype=np.int64)
    for r in range(n):
        for k in range(int(A.rowptr[r]), int(A.rowptr[r + 1])):
            lev[r, int(A.col[k])] = 0
    for i in range(n):
        for k in range(i):
            if lev[i, k] > p:
                continue
            for j in range(k + 1, n):
                if lev[k, j] <= p:
                    lev[i, j] = min(lev[i, j], lev[i, k] + lev[k, j] + 1)
    return {(i, j) for i in range(n) for j in range(n) if i != j and lev[i, j] <= p}


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("case", ["random", "q1_2d", "q1_3d"])
def test_ilup_is_reference(feat, oracle, reference, case, p):
    """ILU(p) restated (level-p pattern, factorisation with zero fill values, solves): the
    reference's own ILUPrecond<generic> with p fill levels gives the same bits."""
    rng = np.random.default_rng(43)
    A = dominant(rng, 120) if case == "random" else q1_csr(feat, 2 if case == "q1_2d" else 3, 3 if case == "q1_2d" else 2)
    f = oracle.ilup_factor(A, p)
    for seed in range(3):
        x = np.random.default_rng(seed).standard_normal(A.rows)
        assert np.array_equal(oracle.ilup_apply(f, x), reference.precond_apply("ilu", A, p, x))


@pytest.mark.parametrize("case", ["random", "q1_2d"])
def test_ilup_pattern_and_closed_forms(feat, oracle, case):
    """Without the reference build: the level-p pattern equals the textbook level-of-fill rule
    on the dense grid; p = 0 is ILU(0); a large p fills the whole LU pattern, so
    (I+L)(D+U) = A exactly up to rounding (complete LU of a diagonally dominant matrix)."""
    rng = np.random.default_rng(47)
    A = dominant(rng, 60, max_len=5) if case == "random" else q1_csr(feat, 2, 3)
    n = A.rows
    for p in (1, 2):
        rl, cl, ru, cu = oracle.ilup_pattern(A, p)
        got = {(i, int(c)) for i in range(n) for c in cl[int(rl[i]):int(rl[i + 1])]}
        got |= {(i, int(c)) for i in range(n) for c in cu[int(ru[i]):int(ru[i + 1])]}
        assert got == brute_level_pattern(A, p)
        for i in range(n):   # sorted rows
            assert np.all(np.diff(cl[int(rl[i]):int(rl[i + 1])].astype(np.int64)) > 0)
            assert np.all(np.diff(cu[int(ru[i]):int(ru[i + 1])].astype(np.int64)) > 0)
    f0, g = oracle.ilu0_factor(A), oracle.ilup_factor(A, 0)
    assert np.array_equal(f0[0], g["l"]) and np.array_equal(f0[1], g["u"]) and np.array_equal(f0[2], g["d"])
    f = oracle.ilup_factor(A, n)
    Lm, Um = np.eye(n), np.diag(1.0 / f["d"])
    for i in range(n):
        for j in range(int(f["rl"][i]), int(f["rl"][i + 1])):
            Lm[i, int(f["cl"][j])] = f["l"][j]
        for j in range(int(f["ru"][i]), int(f["ru"][i + 1])):
            Um[i, int(f["cu"][j])] = f["u"][j]
    M = dense(A)
    assert np.allclose(Lm @ Um, M, rtol=0, atol=1e-12 * np.abs(M).max())
