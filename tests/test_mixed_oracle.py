"""Pins the oracle's FP32 / FP16 arithmetic (mixed-precision V-cycle, FP32 CSR apply).

* FP32 CSR apply and the FP32 V-cycle inside FP64 PCG are compared bit for bit
  with the reference's own float code (oracle/_ref: SparseMatrixCSR<float> apply,
  PCG<double> around ConvertPrecond<double, float> around a float MultiGrid with
  float JacobiPrecond / Richardson / Transfer).
* The FP16-valued form is pinned in two parts: its value rounding against numpy's
  float16 conversion (an independent round-to-nearest-even implementation), and its
  cycle against the FP32 cycle on matrices whose values FP16 holds exactly (there the
  two must agree bit for bit, so a misplaced scale or conversion shows).
* Each pin is shown to catch a planted slip of the kind it is there for.
"""
import numpy as np
import pytest

from oracle.pyoracle import Csr, COARSE_DIRECT, COARSE_JACOBI, COARSE_SMOOTHER
from tests.helpers import q1_levels, random_csr

EPS = np.finfo(np.float64).eps


def _kat_tridiag(n):
    # sparse_matrix_csr-test.cpp:358-452 pattern: tridiag(-1, 2, -1)
    rp = np.zeros(n + 1, np.uint64)
    cols, vals = [], []
    for i in range(n):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n:
                cols.append(j)
                vals.append(2.0 if i == j else -1.0)
        rp[i + 1] = len(cols)
    return Csr(n, n, np.array(vals), np.array(cols, np.uint64), rp)


def _f32_cases(feat):
    rng = np.random.default_rng(41)
    yield random_csr(rng, 97, 61, 0, 40, long_rows=((5, 300),))
    yield _kat_tridiag(64)
    lv = q1_levels(feat, 3, 3, 3)[0]
    yield lv["A"]


def _nonfused_rowsum32(A, x):
    """A planted slip: every product rounded before it is added (no fma)."""
    r = np.zeros(A.rows, np.float32)
    v32 = A.val.astype(np.float32)
    for row in range(A.rows):
        s = np.float32(0.0)
        for i in range(int(A.rowptr[row]), int(A.rowptr[row + 1])):
            s = np.float32(s + np.float32(v32[i] * x[int(A.col[i])]))
        r[row] = s
    return r


def test_csr_apply_f32_bitwise_vs_reference(oracle, reference, feat):
    rng = np.random.default_rng(3)
    for A in _f32_cases(feat):
        x = rng.standard_normal(A.cols).astype(np.float32)
        y = rng.standard_normal(A.rows).astype(np.float32)
        xt = rng.standard_normal(A.rows).astype(np.float32)
        yt = rng.standard_normal(A.cols).astype(np.float32)
        for a, b in [(1.0, 0.0), (-1.0, 1.0), (4711.1, 1.0), (0.0, 1.0), (0.37, -2.5)]:
            o = oracle.csr_apply_f32(A, x, y, a, b)
            r = reference.csr_apply_f32(A, x, y, a, b)
            assert np.array_equal(o.view(np.uint32), r.view(np.uint32)), (a, b)
            ot = oracle.csr_apply_f32(A, xt, yt, a if a else 1.0, b, transposed=True)
            rt = reference.csr_apply_f32(A, xt, yt, a if a else 1.0, b, transposed=True)
            assert np.array_equal(ot.view(np.uint32), rt.view(np.uint32)), (a, b)


def test_csr_apply_f32_pin_catches_unfused_sums(reference):
    # the pin above would fail an oracle that rounds each product before the add
    rng = np.random.default_rng(8)
    A = random_csr(rng, 200, 150, 10, 40)
    x = rng.standard_normal(150).astype(np.float32)
    r = reference.csr_apply_f32(A, x)
    assert not np.array_equal(_nonfused_rowsum32(A, x), r)


def test_component_invert_f32_vs_reference(reference):
    rng = np.random.default_rng(9)
    d = rng.uniform(0.1, 3.0, 1000).astype(np.float32)
    r = reference.component_invert_f32(d, 0.5)
    assert np.array_equal(r, np.float32(0.5) / d)


@pytest.mark.parametrize("dim,lf", [(2, 5), (3, 4)])
@pytest.mark.parametrize("coarse", [COARSE_JACOBI, COARSE_SMOOTHER, COARSE_DIRECT])
def test_mixed_vcycle_bitwise_vs_reference(oracle, reference, feat, dim, lf, coarse):
    levels = q1_levels(feat, dim, lf, 1)
    n = levels[0]["n"]
    b = feat.rhs_random(28054777172512, n)
    b[levels[0]["filt"].astype(np.int64)] = 0.0
    o = oracle.mg_pcg_solve(levels, np.zeros(n), b, coarse=coarse, precision="f32")
    r = reference.mg_pcg_solve_mixed(levels, np.zeros(n), b, coarse=coarse)
    assert o["status"] == r["status"] == "success"
    assert o["iters"] == r["iters"]
    assert np.array_equal(o["x"], r["x"])
    # defect norms: the reference's norm2 contracts by inlining site (see test_oracle.py)
    assert np.allclose(o["hist"], r["hist"], rtol=4 * EPS, atol=0)
    # and the FP32 cycle really ran: the FP64 oracle gives another x
    o64 = oracle.mg_pcg_solve(levels, np.zeros(n), b, coarse=coarse)
    assert not np.array_equal(o64["x"], o["x"])


def _shifted_levels(levels, seed):
    """Each level's diagonal raised by a random 0..10 % (still SPD): diagonal
    values with no short binary expansion, where FP32 and FP64 quotients part."""
    rng = np.random.default_rng(seed)
    out = []
    for lv in levels:
        A = lv["A"]
        val = A.val.copy()
        for r in range(A.rows):
            for i in range(int(A.rowptr[r]), int(A.rowptr[r + 1])):
                if int(A.col[i]) == r:
                    val[i] *= 1.0 + 0.1 * rng.random()
        d = dict(lv)
        d["A"] = Csr(A.rows, A.cols, val, A.col, A.rowptr)
        out.append(d)
    return out


@pytest.mark.parametrize("coarse", [COARSE_JACOBI, COARSE_DIRECT])
def test_mixed_vcycle_bitwise_vs_reference_shifted_diagonal(oracle, reference, feat, coarse):
    levels = _shifted_levels(q1_levels(feat, 3, 4, 1), 5)
    n = levels[0]["n"]
    b = feat.rhs_random(17, n)
    b[levels[0]["filt"].astype(np.int64)] = 0.0
    o = oracle.mg_pcg_solve(levels, np.zeros(n), b, coarse=coarse, precision="f32")
    r = reference.mg_pcg_solve_mixed(levels, np.zeros(n), b, coarse=coarse)
    assert o["iters"] == r["iters"]
    assert np.array_equal(o["x"], r["x"])


def test_mixed_pin_catches_fp64_rounded_diagonal(oracle, reference, feat):
    # planted slip: omega / a_ii formed in FP64 and rounded once (the form the
    # oracle had before it was pinned); on the Q1 levels themselves both forms
    # happen to round alike (0.5 / (8/3), 0.5 / (1/6) ...), so the shifted
    # diagonals of the test above carry the check
    levels = _shifted_levels(q1_levels(feat, 3, 4, 1), 5)
    n = levels[0]["n"]
    b = feat.rhs_random(17, n)
    b[levels[0]["filt"].astype(np.int64)] = 0.0
    o = oracle.mg_pcg_solve(levels, np.zeros(n), b, coarse=COARSE_JACOBI, precision="f32",
                            planted_slip=True)
    r = reference.mg_pcg_solve_mixed(levels, np.zeros(n), b, coarse=COARSE_JACOBI)
    assert not np.array_equal(o["x"], r["x"])


def _f16_numpy(val):
    """FP16 value rounding by numpy (independent of the oracle's F16C path)."""
    v32 = np.asarray(val, np.float64).astype(np.float32)
    mx = np.max(np.abs(v32)) if v32.size else np.float32(0)
    scale = np.float32(1.0)
    if mx > 0 and np.isfinite(mx):
        _, e = np.frexp(mx)
        scale = np.float32(np.ldexp(1.0, int(e)))
    inv = np.float32(1.0) / scale
    return (v32 * inv).astype(np.float16).astype(np.float32), float(scale)


def test_f16_values_vs_numpy(oracle):
    rng = np.random.default_rng(12)
    cases = [
        rng.standard_normal(5000),
        rng.standard_normal(3000) * 1e-3,
        # ties: halfway between two FP16 numbers in [0.5, 1) (spacing 2^-11) at scale 1
        np.array([0.5 + (2 * k + 1) * 2.0 ** -12 for k in range(64)] + [0.75]),
        # FP16 subnormals and underflow after scaling, signed zeros
        np.array([1.0, 3e-5, -3e-5, 6e-8, -6e-8, 1e-9, 0.0, -0.0, 2.0 ** -24, 2.0 ** -25]),
        # the 3D Q1 stencil values at level 7 (h = 1/128)
        np.array([8.0 / 384, -1.0 / 768, -1.0 / 1536, 0.0]),
    ]
    for val in cases:
        o, so = oracle.f16_values(val)
        e, se = _f16_numpy(val)
        assert so == se
        assert np.array_equal(o.view(np.uint32), e.view(np.uint32))
    # the scale puts the largest magnitude in [1/2, 1)
    _, s = oracle.f16_values(np.array([3.0, -5.5]))
    assert s == 8.0


def _scaled_levels(levels, s):
    out = []
    for lv in levels:
        A = lv["A"]
        d = dict(lv)
        d["A"] = Csr(A.rows, A.cols, A.val * s, A.col, A.rowptr)
        out.append(d)
    return out


def test_f16_cycle_equals_f32_on_exact_values(oracle, reference, feat):
    # 3 * (2D Q1 Laplace) has integer entries (8, -1 and boundary sums) and the Q1
    # transfers hold 1, 1/2, 1/4: FP16 stores every value exactly, so the FP16-valued
    # cycle must equal the FP32 cycle (pinned to the reference above) bit for bit
    levels = _scaled_levels(q1_levels(feat, 2, 5, 1), 3.0)
    for lv in levels:
        v, _ = oracle.f16_values(lv["A"].val)
        assert np.array_equal(v * np.float32(_f16_numpy(lv["A"].val)[1]), lv["A"].val.astype(np.float32))
    n = levels[0]["n"]
    b = feat.rhs_random(28054777172512, n)
    b[levels[0]["filt"].astype(np.int64)] = 0.0
    h = oracle.mg_pcg_solve(levels, np.zeros(n), b, coarse=COARSE_JACOBI, precision="f16")
    f = oracle.mg_pcg_solve(levels, np.zeros(n), b, coarse=COARSE_JACOBI, precision="f32")
    r = reference.mg_pcg_solve_mixed(levels, np.zeros(n), b, coarse=COARSE_JACOBI)
    assert h["iters"] == f["iters"] == r["iters"]
    assert np.array_equal(h["x"], f["x"]) and np.array_equal(f["x"], r["x"])


def test_f16_cycle_rounds_inexact_values(oracle, feat):
    # on the unscaled matrix (8/3, -1/3 are not FP16 numbers) the FP16 form differs
    # from the FP32 form but converges the same way
    levels = q1_levels(feat, 2, 5, 1)
    n = levels[0]["n"]
    b = feat.rhs_random(28054777172512, n)
    b[levels[0]["filt"].astype(np.int64)] = 0.0
    h = oracle.mg_pcg_solve(levels, np.zeros(n), b, coarse=COARSE_JACOBI, precision="f16")
    f = oracle.mg_pcg_solve(levels, np.zeros(n), b, coarse=COARSE_JACOBI, precision="f32")
    assert not np.array_equal(h["x"], f["x"])
    assert abs(h["iters"] - f["iters"]) <= 1
    assert np.linalg.norm(h["x"] - f["x"]) <= 1e-8 * np.linalg.norm(f["x"])
