"""The batch of uneven Q1 diffusion MG-PCG solves on the GPU against the oracle,
instance by instance (same method, same stopping rule): equal iteration counts,
solutions within 1e-10, and bit-identical results with and without the
recorded-graph replay, with and without autotuned launch shapes, with and
without the fused coarse tail (levels 1..T of every instance in one launch) and
with and without the tiled smoother calls on the structured grids."""
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


def _solve(torch, bs, b):
    x = torch.zeros(bs.n, dtype=torch.float64, device="cuda")
    r = bs.solve(x, b)
    return host(x), r


@pytest.mark.parametrize("ell", [False, True])
@pytest.mark.parametrize("ninst,lmin,lmax,tail,depth", [
    (12, 3, 6, 1, 1),   # only level 1 in the tail: every instance enters it from level 2
    (12, 3, 6, 3, 3),   # instances whose finest level is 3 run entirely inside the tail
    (12, 3, 6, 8, 5),   # lowered to the deepest tail one workgroup holds (level 5: 1089 rows)
    (6, 2, 4, 5, 4),    # the whole hierarchy of every instance in the tail
])
def test_batch_coarse_tail_bitwise(feat, torch_gpu, ninst, lmin, lmax, tail, depth, ell):
    # the tail in its grid form (stencils, bilinear transfers, matrices prefetched) and in
    # its ELL form, against the block-diagonal SELL-64 launches
    torch = torch_gpu
    ref = feat.BatchSolver(SEED, ninst=ninst, lmin=lmin, lmax=lmax, tail_levels=0, tiled=False)
    bs = feat.BatchSolver(SEED, ninst=ninst, lmin=lmin, lmax=lmax, tail_levels=tail, tiled=False, tail_ell=ell)
    try:
        assert ref.tail_levels == 0 and bs.tail_levels == depth and bs.tail_grid == (not ell)
        b = dev(torch, batch_rhs(feat, ref))
        x0, r0 = _solve(torch, ref, b)
        for rep in range(2):
            x1, r1 = _solve(torch, bs, b)
            assert np.array_equal(x1, x0), rep
            assert np.array_equal(r1["iters"], r0["iters"])
            for i, it in enumerate(r0["iters"]):  # entries past an instance's last iteration are not written
                assert np.array_equal(r1["hist"][i, :it + 1], r0["hist"][i, :it + 1]), i
        assert all(s == "success" for s in r0["status"])
    finally:
        bs.close()
        ref.close()


@pytest.mark.parametrize("tail,prol_arrays,rest", [(5, False, True), (0, False, True), (5, True, True),
                                                  (5, False, False)])
def test_batch_tiled_smoother_bitwise(feat, torch_gpu, tail, prol_arrays, rest):
    # tiled smoother calls on the levels above the tail (tail 0: levels 2..6, down to the
    # 5 x 5 grid of level 2, one clipped tile) against the SELL-64 launches; the
    # prolongation formed from the node position, or read from its ELL arrays; the
    # restriction inside the tiled pre-smoothing, or a launch of its own
    torch = torch_gpu
    ref = feat.BatchSolver(SEED, ninst=12, lmin=3, lmax=6, tail_levels=tail, tiled=False)
    bs = feat.BatchSolver(SEED, ninst=12, lmin=3, lmax=6, tail_levels=tail, tiled_prol_arrays=prol_arrays,
                          tiled_restriction=rest)
    try:
        assert ref.tiled_levels == []
        assert bs.tiled_levels == ([6] if tail else [2, 3, 4, 5, 6])
        b = dev(torch, batch_rhs(feat, ref))
        x0, r0 = _solve(torch, ref, b)
        for rep in range(2):
            x1, r1 = _solve(torch, bs, b)
            assert np.array_equal(x1, x0), rep
            assert np.array_equal(r1["iters"], r0["iters"])
            for i, it in enumerate(r0["iters"]):
                assert np.array_equal(r1["hist"][i, :it + 1], r0["hist"][i, :it + 1]), i
    finally:
        bs.close()
        ref.close()


def test_batch_full_64(feat, torch_gpu):
    # the full batch: 64 instances, levels 3..8, all converge to 1e-8; the fused coarse
    # tail (levels 1..5) and the tiled smoother calls (levels 6..8) give the same bits as
    # the block-diagonal SELL-64 launches
    torch = torch_gpu
    bs = feat.BatchSolver(SEED, ninst=64, lmin=3, lmax=8)
    ref = feat.BatchSolver(SEED, ninst=64, lmin=3, lmax=8, tail_levels=0, tiled=False)
    try:
        assert bs.tail_levels == 5 and bs.tail_grid and bs.tiled_levels == [6, 7, 8]
        b = dev(torch, batch_rhs(feat, bs))
        xg, r = _solve(torch, bs, b)
        x0, r0 = _solve(torch, ref, b)
    finally:
        bs.close()
        ref.close()
    assert all(s == "success" for s in r["status"])
    assert np.all(r["def_final"] <= 1e-8 * r["def_init"])
    assert set(int(v) for v in bs.levels) >= {3, 8}
    assert np.array_equal(xg, x0) and np.array_equal(r["iters"], r0["iters"])
