"""The oracle on the 3D Q1 variable-coefficient problem (level 6, log-normal sigma =
1.5 per cell, anisotropy diag(1, 1, 10)): PCG(MG-V(4,4) Richardson(Jacobi 0.5)).

* Where oracle/_ref is built, the oracle's solve equals the reference's own
  PCG / MultiGrid / Richardson / JacobiPrecond on the same matrices (Jacobi coarse
  solve): the same iterations, the same solution bits.
* The problem's rounding sensitivity, which sets the GPU test's bar: one-ulp
  changes of 1,000 right-hand-side entries move the defect history by more than
  1 % while the solution after the same number of iterations moves by less than
  1e-10 (relative).
"""
import numpy as np
import pytest

from oracle.pyoracle import COARSE_DIRECT, COARSE_JACOBI
from tests.test_gpu_c5 import c5_levels, SEED


@pytest.fixture(scope="module")
def c5cpu(feat):
    levels = c5_levels(feat)
    n = levels[0]["n"]
    b = feat.rhs_random(SEED + 2000, n)
    b[levels[0]["filt"].astype(np.int64)] = 0.0
    return levels, b


def test_c5_oracle_equals_reference(oracle, reference, c5cpu):
    levels, b = c5cpu
    n = levels[0]["n"]
    o = oracle.mg_pcg_solve(levels, np.zeros(n), b, coarse=COARSE_JACOBI, max_iter=200)
    r = reference.mg_pcg_solve(levels, np.zeros(n), b, coarse=COARSE_JACOBI, max_iter=200)
    assert o["status"] == r["status"] == "success"
    assert o["iters"] == r["iters"]
    assert np.array_equal(o["x"].view(np.uint64), r["x"].view(np.uint64))
    assert np.allclose(o["hist"], r["hist"], rtol=1e-15, atol=0)


def test_c5_rounding_sensitivity(oracle, c5cpu):
    levels, b = c5cpu
    n = levels[0]["n"]
    o = oracle.mg_pcg_solve(levels, np.zeros(n), b, coarse=COARSE_DIRECT, max_iter=200)
    rng = np.random.default_rng(1)
    bp = b.copy()
    idx = rng.integers(0, n, 1000)
    bp[idx] = np.nextafter(bp[idx], 2.0)
    bp[levels[0]["filt"].astype(np.int64)] = 0.0
    p = oracle.mg_pcg_solve(levels, np.zeros(n), bp, coarse=COARSE_DIRECT, max_iter=o["iters"])
    ho, hp = o["hist"] / o["hist"][0], p["hist"] / p["hist"][0]
    m = min(ho.size, hp.size)
    assert np.max(np.abs(ho[:m] - hp[:m]) / ho[:m]) > 1e-2
    assert np.linalg.norm(p["x"] - o["x"]) <= 1e-10 * np.linalg.norm(o["x"])
