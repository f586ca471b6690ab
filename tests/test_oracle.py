"""Pins the CPU oracle (oracle/feat_oracle.c) before it is trusted.

Each check is chosen so that a plausible slip (dropped term, wrong sign or
index, transposed operand, wrong rounding) fails it:
  * golden vectors written by the reference's own code (tests/golden/),
  * the reference's known-answer tests restated (closed forms),
  * brute force (dense products) and invariants,
  * where oracle/_ref is built here, bit-for-bit agreement with the reference.
"""
import numpy as np
import pytest

from oracle.pyoracle import Csr, COARSE_JACOBI, COARSE_DIRECT, COARSE_SMOOTHER
from tests.helpers import random_csr, q1_levels

EPS = np.finfo(np.float64).eps


def fd_matrix(golden):
    n = golden["fd_rowptr"].size - 1
    return Csr(n, n, golden["fd_val"], golden["fd_col"], golden["fd_rowptr"])


# ---------------------------------------------------------------- RNG
def test_rng_matches_reference_golden(oracle, golden):
    # kernel/util/random.hpp:84-110, 259-280 via DenseVector(Random&, n, a, b)
    assert np.array_equal(oracle.rng_fill(28054777172512, 64, -1.0, 1.0), golden["rng_default_m1p1"])
    assert np.array_equal(oracle.rng_fill(17, 64, 0.0, 1.0), golden["rng_seed17_0p1"])


def test_rng_xorshift_closed_form(oracle):
    # xorshift64*: x ^= x>>12; x ^= x<<25; x ^= x>>27; return x * 2685821657736338717
    x = 28054777172512
    exp = []
    for _ in range(4):
        x ^= x >> 12
        x ^= (x << 25) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 27
        exp.append((x * 2685821657736338717) & 0xFFFFFFFFFFFFFFFF)
    assert [int(v) for v in oracle.rng_next(28054777172512, 4)] == exp


# ---------------------------------------------------------------- CSR apply
@pytest.mark.parametrize("n", [1, 2, 8, 64, 512])
def test_csr_apply_kat(oracle, golden, n):
    # sparse_matrix_csr-test.cpp:358-452: tridiag(-1,2,-1), x_i=(i%100)*1.234, y_i=2-(i%42)
    rows = []
    for i in range(n):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n:
                rows.append((i, j, 2.0 if i == j else -1.0))
    rp = np.zeros(n + 1, np.uint64)
    for i, _, _ in rows:
        rp[i + 1] += 1
    A = Csr(n, n, np.array([v for *_, v in rows]), np.array([j for _, j, _ in rows], np.uint64),
            np.cumsum(rp).astype(np.uint64))
    i = np.arange(n)
    x = (i % 100) * 1.234
    y = 2.0 - (i % 42)
    # closed-form A*x of the reference test
    ax = np.where(i == 0, 1.234 * (2 * (i % 100) - ((i + 1) % 100)),
                  np.where(i == n - 1, 1.234 * (2 * (i % 100) - ((i - 1) % 100)),
                           1.234 * (2 * (i % 100) - ((i - 1) % 100) - ((i + 1) % 100))))
    if n == 1:
        ax = np.zeros(1)
    assert np.allclose(oracle.csr_apply(A, x), ax, atol=1e-2, rtol=0)
    assert np.allclose(oracle.csr_apply(A, x, y, 4711.1, 1.0), y + 4711.1 * ax, atol=5e-2, rtol=0)
    # bit-for-bit against the reference's own generic apply
    assert np.array_equal(oracle.csr_apply(A, x), golden[f"kat{n}_ax"])
    for tag, a in [("0", 0.0), ("m1", -1.0), ("s", 4711.1)]:
        assert np.array_equal(oracle.csr_apply(A, x, y, a, 1.0), golden[f"kat{n}_a{tag}"])
    assert np.array_equal(oracle.csr_apply(A, x, y, 4711.1, 1.0, transposed=True), golden[f"kat{n}_trs"])


def test_csr_apply_brute_force(oracle):
    rng = np.random.default_rng(5)
    A = random_csr(rng, 77, 53, 0, 12)
    D = A.todense()
    x = rng.standard_normal(53)
    y = rng.standard_normal(77)
    assert np.allclose(oracle.csr_apply(A, x, y, -0.7, 1.0), y - 0.7 * D @ x, rtol=1e-13, atol=1e-13)
    xt = rng.standard_normal(77)
    yt = rng.standard_normal(53)
    assert np.allclose(oracle.csr_apply(A, xt, yt, 2.5, 1.0, transposed=True), yt + 2.5 * D.T @ xt,
                       rtol=1e-12, atol=1e-12)
    # |b| < eps: y is not read (NaN in y must not propagate)
    ynan = np.full(77, np.nan)
    assert np.all(np.isfinite(oracle.csr_apply(A, x, ynan, 1.0, 0.0)))


def test_bcsr_apply_brute_force(oracle):
    rng = np.random.default_rng(6)
    for bh, bw in [(2, 2), (2, 1), (1, 2), (3, 3)]:
        S = random_csr(rng, 31, 23, 0, 6)
        val = rng.standard_normal(S.nnz * bh * bw)
        D = np.zeros((31 * bh, 23 * bw))
        for r in range(31):
            for i in range(int(S.rowptr[r]), int(S.rowptr[r + 1])):
                c = int(S.col[i])
                D[r * bh:(r + 1) * bh, c * bw:(c + 1) * bw] += val[i * bh * bw:(i + 1) * bh * bw].reshape(bh, bw)
        x = rng.standard_normal(23 * bw)
        y = rng.standard_normal(31 * bh)
        r = oracle.bcsr_apply(bh, bw, 31, val, S.col, S.rowptr, x, y, 1.5, 1.0)
        assert np.allclose(r, y + 1.5 * D @ x, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- BLAS-1
def test_dot_closed_form(oracle):
    # dense_vector-test.cpp:366: a_i=(i+1)/n, b_i=1/(i+1) => a.b = 1
    for n in [1, 7, 100, 10000]:
        a = (np.arange(n) + 1.0) / n
        b = 1.0 / (np.arange(n) + 1.0)
        assert abs(oracle.dot(a, b) - 1.0) <= EPS ** 0.8
        assert abs(oracle.norm2(b) - np.sqrt(np.sum(b * b))) <= 1e-14


def test_blas1_definitions(oracle):
    from fractions import Fraction
    rng = np.random.default_rng(7)
    x, y, z = rng.standard_normal((3, 300))
    # axpy is one correctly rounded a*x+y per element (exact rational arithmetic)
    fma_ref = np.array([float(Fraction(0.37) * Fraction(a) + Fraction(b)) for a, b in zip(x, y)])
    assert np.array_equal(oracle.axpy(0.37, x, y), fma_ref)
    assert np.array_equal(oracle.scale(x, 0.3), x * 0.3)
    assert np.array_equal(oracle.component_product(x, y), x * y)
    assert np.array_equal(oracle.component_invert(x, 0.7), 0.7 / x)
    assert abs(oracle.triple_dot(x, y, z) - np.sum(x * y * z)) < 1e-11


def test_blas1_bitwise_vs_reference(oracle, reference):
    rng = np.random.default_rng(8)
    x, y, z = rng.standard_normal((3, 4097))
    for alias in (None, "x", "y"):
        # r == x rounds twice (arch/axpy_generic.hpp:36-43), the others once
        assert np.array_equal(oracle.axpy(0.37, x, y, alias=alias), reference.axpy(0.37, x, y, alias=alias)), alias
    assert np.array_equal(oracle.axpy(0.37, x, y, alias="x"), x * 0.37 + y)
    assert oracle.dot(x, y) == reference.dot(x, y)
    assert oracle.dot(x, x) == reference.dot(x, x)
    assert oracle.norm2(x) == reference.norm2(x)
    assert np.array_equal(oracle.component_invert(x, 0.7), reference.component_invert(x, 0.7))


def test_diag_index_and_filters(oracle):
    rng = np.random.default_rng(9)
    A = random_csr(rng, 40, 40, 0, 8)
    d = oracle.diag_index(A)
    for r in range(40):
        cols = A.col[int(A.rowptr[r]):int(A.rowptr[r + 1])]
        hit = np.nonzero(cols == r)[0]
        assert d[r] == (A.rowptr[r] + hit[0] if hit.size else A.nnz)
    v = rng.standard_normal(40)
    idx = np.array([3, 0, 39], np.uint64)
    w = oracle.unit_filter_def(v, idx)
    assert np.all(w[[0, 3, 39]] == 0) and np.array_equal(np.delete(w, [0, 3, 39]), np.delete(v, [0, 3, 39]))
    w = oracle.unit_filter_rhs(v, np.array([1.0, 2.0, 3.0]), idx)
    assert w[3] == 1.0 and w[0] == 2.0 and w[39] == 3.0
    # mean filter: result has zero dual mean
    prim = np.ones(40)
    dual = rng.uniform(0.5, 1.5, 40)
    vol = float(prim @ dual)
    f = oracle.mean_filter(v, prim, dual, vol)
    assert abs(f @ prim) < 1e-12 * np.abs(v).sum()


# ---------------------------------------------------------------- invert_matrix
def test_invert_matrix_identity_and_golden(oracle, golden):
    # math-test.cpp:319-420: |I - A^-1 A| <= eps^0.8 for random n = 1..9
    rng = np.random.default_rng(10)
    for n in range(1, 10):
        M = rng.uniform(-1, 1, (n, n)) + n * np.eye(n)
        inv, det = oracle.invert_matrix(M)
        assert np.abs(np.eye(n) - inv @ M).max() <= EPS ** 0.8
        assert abs(det - np.linalg.det(M)) <= 1e-10 * abs(det)
    inv, det = oracle.invert_matrix(golden["inv_in"])
    assert np.array_equal(inv, golden["inv_out"]) and det == golden["inv_det"][0]


# ---------------------------------------------------------------- solvers
def test_fd_solvers_pinned_counts(oracle, golden):
    # basic_solver-test.cpp:130-190: CG 28, PCG-JAC 28, Richardson-JAC 1175 (+-2), |x-x_ref|^2 <= sqrt(eps)
    A = fd_matrix(golden)
    xref = golden["fd_q2bubble"]
    b = oracle.csr_apply(A, xref)
    x0 = np.zeros(A.rows)
    runs = {"cg": oracle.pcg(A, x0, b, precond=0, apply=True),
            "pcgjac": oracle.pcg(A, x0, b, precond=1, apply=True),
            "richjac": oracle.richardson(A, x0, b, precond=1, apply=True, max_iter=1500)}
    for name, ref_it in [("cg", 28), ("pcgjac", 28), ("richjac", 1175)]:
        r = runs[name]
        assert r["status"] == "success"
        assert abs(r["iters"] - ref_it) <= 2
        assert np.sum((r["x"] - xref) ** 2) <= np.sqrt(EPS)
        # reference run recorded by tests/golden/make_golden.py
        assert r["iters"] == int(golden[f"fd_{name}_iters"][0])
        assert np.allclose(r["x"], golden[f"fd_{name}_x"], rtol=0, atol=1e-14)
        assert np.allclose(r["hist"], golden[f"fd_{name}_hist"], rtol=1e-13, atol=0)


def test_fd_cg_closed_form_rate(oracle, golden):
    # CG contraction bound from the FD eigenvalues 2d(1 -+ cos(pi/(m+1))) (pointstar_factory.hpp:388-410)
    A = fd_matrix(golden)
    m, d = 17, 2
    lmin, lmax = 2 * d * (1 - np.cos(np.pi / (m + 1))), 2 * d * (1 + np.cos(np.pi / (m + 1)))
    kappa = lmax / lmin
    rate = (np.sqrt(kappa) - 1) / (np.sqrt(kappa) + 1)
    b = oracle.csr_apply(A, golden["fd_q2bubble"])
    r = oracle.pcg(A, np.zeros(A.rows), b, precond=0, apply=True)
    # A-norm error bound implies the iteration count cannot exceed the bound's count by much
    bound = np.log(np.sqrt(EPS) / 2) / np.log(rate) + np.log(np.sqrt(kappa))
    assert r["iters"] <= bound + 2


def test_poisson_dirichlet_l5_pinned(oracle, feat):
    # applications/CMakeLists.txt:140: Q1 2D level 5..0, PCG-MG V(4,4) Jacobi 0.7,
    # coarse = the smoother, tol 1e-8 -> 4 +- 1 iterations
    levels = q1_levels(feat, 2, 5, 0)
    n = levels[0]["n"]
    b = feat.q1_rhs_sine(2, 5, n)
    b[levels[0]["filt"].astype(np.int64)] = 0.0
    r = oracle.mg_pcg_solve(levels, np.zeros(n), b, omega_jac=0.7, coarse=COARSE_SMOOTHER)
    assert r["status"] == "success" and abs(r["iters"] - 4) <= 1


def test_mg_pcg_c1_golden(oracle, feat, golden):
    levels = q1_levels(feat, 2, 7, 1)
    n = levels[0]["n"]
    b = golden["c1_rhs"]
    r = oracle.mg_pcg_solve(levels, np.zeros(n), b, coarse=COARSE_JACOBI)
    # the reference's own solver on the same matrices, recorded bit for bit
    assert r["iters"] == int(golden["c1_iters"][0])
    assert np.allclose(r["hist"], golden["c1_hist"], rtol=4 * EPS, atol=0)
    assert np.array_equal(r["x"], golden["c1_x"])
    # exact coarse solve: with one interior coarse dof it equals Jacobi(1) up to one rounding
    d = oracle.mg_pcg_solve(levels, np.zeros(n), b, coarse=COARSE_DIRECT)
    assert d["iters"] == r["iters"]
    assert np.allclose(d["x"], r["x"], rtol=0, atol=1e-14 * np.abs(r["x"]).max())


def test_mg_invariants(oracle, feat):
    # monotone defect and exact solve of the coarse problem: a 2-level MG on a
    # level-2 grid with direct coarse solve converges with rate well below 1
    levels = q1_levels(feat, 2, 4, 2)
    n = levels[0]["n"]
    b = feat.rhs_random(3, n)
    b[levels[0]["filt"].astype(np.int64)] = 0.0
    r = oracle.mg_pcg_solve(levels, np.zeros(n), b, coarse=COARSE_DIRECT)
    h = r["hist"]
    assert r["status"] == "success" and np.all(np.diff(h) < 0)
    assert (h[-1] / h[0]) ** (1.0 / r["iters"]) < 0.2
    # the solution solves the filtered system
    A = levels[0]["A"]
    res = b - oracle.csr_apply(A, r["x"])
    res[levels[0]["filt"].astype(np.int64)] = 0.0
    assert np.linalg.norm(res) <= 1e-8 * np.linalg.norm(b) * 1.0001


def test_mg_pcg_bitwise_vs_reference(oracle, reference, feat):
    levels = q1_levels(feat, 3, 4, 1)
    n = levels[0]["n"]
    b = feat.rhs_random(28054777172512, n)
    b[levels[0]["filt"].astype(np.int64)] = 0.0
    for coarse in (COARSE_JACOBI, COARSE_SMOOTHER):
        o = oracle.mg_pcg_solve(levels, np.zeros(n), b, coarse=coarse)
        r = reference.mg_pcg_solve(levels, np.zeros(n), b, coarse=coarse)
        assert o["iters"] == r["iters"]
        assert np.array_equal(o["x"], r["x"])
        # the solution is bit-identical; the reference's own norm2 contracts
        # differently depending on where the compiler inlines it, so the
        # defect norms agree to the last bit or two
        assert np.allclose(o["hist"], r["hist"], rtol=4 * EPS, atol=0)


@pytest.mark.parametrize("bs", [1, 2, 3])
def test_unit_filter_blocked_vs_reference(oracle, reference, bs):
    # unit_filter_blocked_generic.hpp:24-60 with and without ign_nans; NaNs planted
    # in the filter values, so both branches differ
    rng = np.random.default_rng(40 + bs)
    nb, ue = 50, 17
    idx = np.sort(rng.choice(nb, ue, replace=False)).astype(np.uint64)
    vals = rng.standard_normal(ue * bs)
    vals[::4] = np.nan
    v = rng.standard_normal(nb * bs)
    for rhs in (True, False):
        for ign in (False, True):
            o = oracle.unit_filter_blocked(rhs, bs, v, vals, idx, ign)
            r = reference.unit_filter_blocked(rhs, bs, v, vals, idx, ign)
            assert np.array_equal(o, r, equal_nan=True)
            # definition: only the filtered components change
            touched = np.zeros(nb * bs, bool)
            for i, b in enumerate(idx):
                for j in range(bs):
                    if not (ign and np.isnan(vals[i * bs + j])):
                        touched[int(b) * bs + j] = True
                        exp = vals[i * bs + j] if rhs else 0.0
                        assert o[int(b) * bs + j] == exp or (np.isnan(exp) and np.isnan(o[int(b) * bs + j]))
            assert np.array_equal(o[~touched], v[~touched])
