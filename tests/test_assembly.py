"""Structured Q1 setup of the host layer against the reference's FE assembly
(golden vectors from tests/golden/make_golden.py) and closed forms."""
import numpy as np
import pytest

from oracle.pyoracle import Csr


def perm_from_coords(coords, level, dim):
    m = 2 ** level + 1
    c = np.rint(coords * (m - 1)).astype(np.int64)
    return c[:, 0] + m * c[:, 1] + (m * m * c[:, 2] if dim == 3 else 0)


@pytest.mark.parametrize("dim,level", [(2, 3), (3, 2)])
def test_q1_matches_reference_fe_assembly(feat, golden, dim, level):
    p = f"fe{dim}d_l{level}_"
    q = feat.q1_problem(dim, level)
    n = q["n"]
    ref = Csr(n, n, golden[p + "val"], golden[p + "col"], golden[p + "rowptr"])
    perm = perm_from_coords(golden[p + "coords"], level, dim)
    Ar = np.zeros((n, n))
    Ar[np.ix_(perm, perm)] = ref.todense()
    Am = Csr(n, n, q["val"], q["col"], q["rowptr"]).todense()
    # values equal to the cubature sums within a few ulps; identical pattern
    assert np.abs(Am - Ar).max() <= 4 * np.finfo(float).eps * np.abs(Ar).max()
    assert q["val"].size == golden[p + "val"].size
    # unit filter = boundary nodes
    assert np.array_equal(np.sort(perm[golden[p + "filt"].astype(np.int64)]), q["filt"].astype(np.int64))
    # prolongation: the reference's 2-level pattern stores extra (near-)zeros
    permc = perm_from_coords(golden[p + "coarse_coords"], level - 1, dim)
    Pr = np.zeros((n, q["pcols"]))
    Pr[np.ix_(perm, permc)] = Csr(n, q["pcols"], golden[p + "pval"], golden[p + "pcol"],
                                  golden[p + "prowptr"]).todense()
    Pm = Csr(n, q["pcols"], q["pval"], q["pcol"], q["prowptr"]).todense()
    assert np.abs(Pm - Pr).max() <= 1e-14


def test_q1_prolongation_golden_9x4(feat):
    # assembly/grid_transfer-test.cpp:118-145: level 0 -> 1 on the unit square
    q = feat.q1_problem(2, 1)
    P = Csr(9, 4, q["pval"], q["pcol"], q["prowptr"]).todense()
    # reference dof order: 4 vertices, 4 edge midpoints (bottom, top, left, right), centre
    m = 3
    ref_nodes = [(0, 0), (2, 0), (0, 2), (2, 2), (1, 0), (1, 2), (0, 1), (2, 1), (1, 1)]
    rows = [x + m * y for x, y in ref_nodes]
    cnodes = [(0, 0), (1, 0), (0, 1), (1, 1)]
    cols = [x + 2 * y for x, y in cnodes]
    d0, d1, d2, d4 = 0.0, 1.0, 0.5, 0.25
    data_ref = np.array([d1, d0, d0, d0, d0, d1, d0, d0, d0, d0, d1, d0, d0, d0, d0, d1,
                         d2, d2, d0, d0, d0, d0, d2, d2, d2, d0, d2, d0, d0, d2, d0, d2,
                         d4, d4, d4, d4]).reshape(9, 4)
    assert np.array_equal(P[np.ix_(rows, cols)], data_ref)


@pytest.mark.parametrize("dim,level", [(2, 4), (3, 3)])
def test_q1_closed_forms(feat, dim, level):
    q = feat.q1_problem(dim, level)
    n, m = q["n"], 2 ** level + 1
    A = Csr(n, n, q["val"], q["col"], q["rowptr"])
    h = 1.0 / (m - 1)
    # interior stencil: 2D centre 8/3, neighbours -1/3; 3D centre 8h/3, faces 0 (stored),
    # edges -h/6, corners -h/12
    c = (m // 2) * (1 + m + (m * m if dim == 3 else 0))
    row = {int(A.col[i]): A.val[i] for i in range(int(A.rowptr[c]), int(A.rowptr[c + 1]))}
    assert len(row) == 3 ** dim
    for j, v in row.items():
        d = np.abs(np.array([j % m, (j // m) % m, j // (m * m)])[:dim] -
                   np.array([c % m, (c // m) % m, c // (m * m)])[:dim]).sum()
        if dim == 2:
            assert v == (8.0 / 3.0 if d == 0 else -1.0 / 3.0)
        else:
            assert v == {0: 8 * h / 3, 1: 0.0, 2: -h / 6, 3: -h / 12}[d]
    # interior row sums vanish; the matrix is symmetric
    inner = np.ones(n, bool)
    inner[q["filt"].astype(np.int64)] = False
    rs = np.add.reduceat(q["val"], q["rowptr"][:-1].astype(np.int64))
    assert np.abs(rs[inner]).max() < 1e-14
    D = A.todense()
    assert np.array_equal(D, D.T)
    # prolongation interpolates constants and linear functions exactly
    P = Csr(n, q["pcols"], q["pval"], q["pcol"], q["prowptr"]).todense()
    assert np.array_equal(P @ np.ones(q["pcols"]), np.ones(n))
    mc = 2 ** (level - 1) + 1
    xc = (np.arange(q["pcols"]) % mc) / (mc - 1)
    xf = (np.arange(n) % m) / (m - 1)
    assert np.allclose(P @ xc, xf, rtol=0, atol=1e-15)
    # closed-form sizes: n = (2^L+1)^d, nnz = (3(2^L+1)-2)^d
    assert n == m ** dim and A.nnz == (3 * m - 2) ** dim


def test_sine_rhs_closed_form(feat):
    # b_i = d*pi^2 * prod sin(pi x_k) * 2(1-cos(pi h))/(pi^2 h); compare with midpoint quadrature
    level, dim = 6, 2
    q = feat.q1_problem(dim, level)
    b = feat.q1_rhs_sine(dim, level, q["n"])
    m = 2 ** level + 1
    h = 1.0 / (m - 1)
    i = m // 3
    # fine 1D quadrature of int sin(pi x) phi_i(x) dx
    t = np.linspace(-h, h, 20001)
    w = np.sin(np.pi * (i * h + t)) * (1 - np.abs(t) / h)
    one_d = np.trapezoid(w, t) if hasattr(np, "trapezoid") else np.trapz(w, t)
    assert abs(b[i + m * i] - 2 * np.pi ** 2 * one_d ** 2) < 1e-9 * abs(b[i + m * i])


def test_random_rhs_matches_reference_rng(feat, golden):
    assert np.array_equal(feat.rhs_random(28054777172512, 64), golden["rng_default_m1p1"])
