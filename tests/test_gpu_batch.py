"""The batch of uneven Q1 diffusion MG-PCG solves on the GPU against the oracle,
instance by instance (same method, same stopping rule): equal iteration counts,
solutions within 1e-10, and bit-identical results with and without the
recorded-graph replay and with and without autotuned launch shapes."""
import numpy as np
import pytest

from oracle.pyoracle import Csr, csr_transpose, COARSE_JACOBI
from tests.test_gpu_lafem import dev, host

pytestmark = pytest.mark.gpu

SEED = 28054777172512


def instance_levels(feat, L, coef):
    levels = []
    for k in range(L, 0, -1):
        q = feat.q1_problem(2, k)
        lv = dict(A=Csr(q["n"], q["n"], feat.q1_diffusion(k, L, coef), q["col"], q["rowptr"]), filt=q["filt"])
        if k > 1:
            P = Csr(q["n"], q["pcols"], q["pval"], q["pcol"], q["prowptr"])
            lv["P"], lv["R"] = P, csr_transpose(P)
        levels.append(lv)
    return levels


def batch_rhs(feat, bs):
    return feat.batch_rhs(bs, SEED)


@pytest.fixture(scope="module")
def small(feat, torch_gpu):
    bs = feat.BatchSolver(SEED, ninst=12, lmin=3, lmax=6)
    yield bs
    bs.close()


def test_batch_vs_oracle(feat, oracle, torch_gpu, small):
    torch = torch_gpu
    bs = small
    b = batch_rhs(feat, bs)
    x = torch.zeros(bs.n, dtype=torch.float64, device="cuda")
    r = bs.solve(x, dev(torch, b))
    xg = host(x)
    for i, L in enumerate(bs.levels):
        L = int(L)
        lo, hi = int(bs.off[i]), int(bs.off[i + 1])
        coef = feat.lognormal_cells(SEED + 1000 + i, 4 ** L, 1.0)
        o = oracle.mg_pcg_solve(instance_levels(feat, L, coef), np.zeros(hi - lo), b[lo:hi], coarse=COARSE_JACOBI,
                                coarse_omega=1.0)
        assert r["status"][i] == o["status"] == "success", i
        assert r["iters"][i] == o["iters"], (i, r["iters"][i], o["iters"])
        assert np.linalg.norm(xg[lo:hi] - o["x"]) <= 1e-10 * np.linalg.norm(o["x"]), i
        assert np.allclose(r["hist"][i, :o["iters"] + 1], o["hist"], rtol=1e-8)


def test_batch_graph_and_tuning_bitwise(feat, torch_gpu, small):
    torch = torch_gpu
    b = dev(torch, batch_rhs(feat, small))
    outs = []
    for kw in (dict(), dict(graphs=False), dict(autotune=False)):
        bs = small if not kw else feat.BatchSolver(SEED, ninst=12, lmin=3, lmax=6, **kw)
        for rep in range(2):
            x = torch.zeros(bs.n, dtype=torch.float64, device="cuda")
            r = bs.solve(x, b)
            outs.append((host(x), r["iters"].copy()))
        if kw:
            bs.close()
    for xo, it in outs[1:]:
        assert np.array_equal(xo, outs[0][0]) and np.array_equal(it, outs[0][1])
    tunes, us = small.tuning()
    assert len(tunes) > 0 and all(0 <= t < 32 for t in tunes)


def test_batch_full_64(feat, torch_gpu):
    # the full batch: 64 instances, levels 3..8, all converge to 1e-8
    torch = torch_gpu
    bs = feat.BatchSolver(SEED, ninst=64, lmin=3, lmax=8)
    b = batch_rhs(feat, bs)
    x = torch.zeros(bs.n, dtype=torch.float64, device="cuda")
    r = bs.solve(x, dev(torch, b))
    bs.close()
    assert all(s == "success" for s in r["status"])
    assert np.all(r["def_final"] <= 1e-8 * r["def_init"])
    assert set(int(v) for v in bs.levels) >= {3, 8}
