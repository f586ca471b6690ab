"""GPU parity of the Stokes path (SaddlePointMatrix, system filter, transfers,
AmaVanka factorisation and apply, Richardson(Schwarz(AmaVanka)) smoother,
Vanka multigrid) against the oracle.  Per-row arithmetic (saddle-point and
Vanka SpMVs, transfers, the Vanka matrix itself) is bit-exact; anything that
passes through the pressure mean filter carries one global dot product whose
summation order differs, so it is compared to a relative 1e-12."""
import numpy as np
import pytest

from tests.helpers import stokes_levels, stokes_rhs
from tests.test_gpu_lafem import dev, host

pytestmark = pytest.mark.gpu

LF, LC = 4, 1


@pytest.fixture(scope="module")
def stk(feat, torch_gpu):
    s = feat.Stokes(LF, LC)
    yield s
    s.close()


@pytest.fixture(scope="module")
def lvls(feat):
    return stokes_levels(feat, LF, LC)


def _rand(n, seed):
    return np.random.default_rng(seed).standard_normal(n)


def test_vanka_matrix_bitwise(stk, lvls, oracle):
    for i, lv in enumerate(lvls):
        skip = i == len(lvls) - 1
        g = stk.vanka(i)
        o = oracle.vanka(lv, 1.0, skip)
        assert g["macros"] == o["macros"] and g["stride"] == o["stride"] == 21
        for bg, bo in zip(g["blocks"], o["blocks"]):
            nnz = int(bo["rowptr"][-1])
            assert np.array_equal(bg["rowptr"], bo["rowptr"])
            assert np.array_equal(bg["col"][:nnz], bo["col"][:nnz])
            assert np.array_equal(bg["val"], bo["val"]), (i, np.max(np.abs(bg["val"] - bo["val"])))
        if skip:
            assert np.array_equal(g["mask"], o["mask"])


def test_vanka_init_numeric_repeatable(stk, lvls, oracle):
    # a second factorisation on the same matrices gives the same bits
    before = stk.vanka(0)
    stk.init_numeric()
    after = stk.vanka(0)
    for a, b in zip(before["blocks"], after["blocks"]):
        assert np.array_equal(a["val"], b["val"])


def test_saddle_apply_bitwise(stk, lvls, oracle, torch_gpu):
    torch = torch_gpu
    for i, lv in enumerate(lvls[:2]):
        n = 2 * lv["nv"] + lv["np"]
        x, y = _rand(n, 1 + i), _rand(n, 7 + i)
        xd, yd = dev(torch, x), dev(torch, y)
        r = torch.empty(n, dtype=torch.float64, device="cuda")
        stk.apply(i, r, xd)
        assert np.array_equal(host(r), oracle.saddle_apply(lv, x))
        for alpha in (-1.0, 0.7):
            stk.apply(i, r, xd, yd, alpha)
            assert np.array_equal(host(r), oracle.saddle_apply(lv, x, y, alpha))


def test_transfer_bitwise(stk, lvls, oracle, torch_gpu):
    torch = torch_gpu
    f, c = lvls[0], lvls[1]
    nf, nc = 2 * f["nv"] + f["np"], 2 * c["nv"] + c["np"]
    xc, xf = _rand(nc, 3), _rand(nf, 4)
    rf = torch.empty(nf, dtype=torch.float64, device="cuda")
    rc = torch.empty(nc, dtype=torch.float64, device="cuda")
    xcd, xfd = dev(torch, xc), dev(torch, xf)
    stk.transfer(0, True, rf, xcd)
    assert np.array_equal(host(rf), oracle.stokes_transfer(f, c, xc, True))
    stk.transfer(0, False, rc, xfd)
    assert np.array_equal(host(rc), oracle.stokes_transfer(f, c, xf, False))


@pytest.mark.parametrize("kind", ["rhs", "sol", "def", "cor"])
def test_filters(stk, lvls, oracle, torch_gpu, kind):
    torch = torch_gpu
    lv = lvls[0]
    nu = 2 * lv["nv"]
    x = _rand(nu + lv["np"], 5)
    # second vector: pressure mean ~ 1e-18, where DenseVector::axpy skips the update
    tiny = x.copy()
    tiny[nu:] -= np.dot(tiny[nu:], lv["prim"]) * lv["dual"] / np.dot(lv["prim"], lv["dual"])
    for v in (x, tiny):
        vd = dev(torch, v)
        stk.filter(0, vd, kind)
        g, o = host(vd), oracle.stokes_filter(lv, v, kind)
        assert np.array_equal(g[:nu], o[:nu])
        assert np.allclose(g[nu:], o[nu:], rtol=1e-12, atol=1e-15 * np.max(np.abs(o[nu:])))


def test_vanka_apply(stk, lvls, oracle, torch_gpu):
    torch = torch_gpu
    lv = lvls[0]
    nu = 2 * lv["nv"]
    d = _rand(nu + lv["np"], 6)
    c = torch.empty(d.size, dtype=torch.float64, device="cuda")
    dd = dev(torch, d)
    stk.vanka_apply(0, c, dd, 1)
    o = oracle.vanka(lv, 1.0, False, d=d)["c"]
    g = host(c)
    assert np.array_equal(g[:nu], o[:nu])
    assert np.allclose(g[nu:], o[nu:], rtol=1e-12, atol=1e-14 * np.max(np.abs(o[nu:])))


def test_smoother(stk, lvls, oracle, torch_gpu):
    torch = torch_gpu
    lv = lvls[0]
    rhs = oracle.stokes_filter(lv, _rand(2 * lv["nv"] + lv["np"], 8), "def")
    sol = torch.empty(rhs.size, dtype=torch.float64, device="cuda")
    rd = dev(torch, rhs)
    stk.smooth(0, sol, rd)
    o = oracle.stokes_smooth(lv, rhs, 5)
    assert np.linalg.norm(host(sol) - o) <= 1e-12 * np.linalg.norm(o)


@pytest.mark.parametrize("lf,lc", [(3, 1), (4, 1), (5, 1)])
def test_stokes_solve_vs_oracle(feat, oracle, torch_gpu, lf, lc):
    torch = torch_gpu
    lv = stokes_levels(feat, lf, lc)
    x0, b = stokes_rhs(lv[0], oracle)
    o = oracle.stokes_mg_solve(lv, x0, b, tol_rel=1e-8)
    s = feat.Stokes(lf, lc)
    x = dev(torch, x0)
    bd = dev(torch, b)
    r = s.solve(x, bd)
    s.close()
    assert r["status"] == o["status"] == "success"
    assert r["iters"] == o["iters"]
    xg = host(x)
    assert np.linalg.norm(xg - o["x"]) <= 1e-10 * np.linalg.norm(o["x"])
    assert np.allclose(r["hist"], o["hist"], rtol=1e-8)
    # FP64 residual of the GPU solution, with the oracle's operator and filter
    res = oracle.stokes_filter(lv[0], oracle.saddle_apply(lv[0], xg, b, -1.0), "def")
    assert np.linalg.norm(res) <= 1e-8 * o["def_init"] * 1.01


def test_stokes_graph_replay_matches_eager(feat, oracle, torch_gpu):
    """Each outer iteration replayed as a recorded HIP graph gives bit-identical iterates
    and defect histories to eager launches, over repeated solves (the second one replays
    the recording) and a new solution buffer (re-recorded)."""
    torch = torch_gpu
    lv = stokes_levels(feat, 4, 1)
    x0, b = stokes_rhs(lv[0], oracle)
    bd = dev(torch, b)
    runs = []
    for graphs in (True, False):
        s = feat.Stokes(4, 1, graphs=graphs)
        out = []
        for case in range(3):
            x = dev(torch, x0)
            r = s.solve(x, bd)
            out.append((r["iters"], np.asarray(r["hist"]), host(x)))
        s.close()
        runs.append(out)
    for (ig, hg, xg), (ie, he, xe) in zip(*runs):
        assert ig == ie
        assert np.array_equal(hg, he)
        assert np.array_equal(xg, xe)


@pytest.mark.parametrize("lf,tail_dofs", [(2, 0), (3, 0), (3, 1000), (4, 1000), (6, 0)])
def test_stokes_coarse_fusion_bitwise(feat, oracle, torch_gpu, lf, tail_dofs):
    """The coarsest levels of the V-cycle in one workgroup (feat_stokes_tail_f64: levels 2..1 by
    default, 3..1 with levels of up to 1,000 unknowns admitted; at lf = 2 and lf = 3 with 1,000
    the whole cycle) give bit-identical iterates and defect histories to the same levels as
    separate launches, eagerly and replayed from the recorded iteration, over two solves in a
    row."""
    torch = torch_gpu
    lv = stokes_levels(feat, lf, lf)
    x0, b = stokes_rhs(lv[0], oracle)
    bd = dev(torch, b)
    runs = {}
    for fusion in (True, False):
        for graphs in (True, False):
            s = feat.Stokes(lf, 1, coarse_fusion=fusion, graphs=graphs, tail_max_level_dofs=tail_dofs)
            out = []
            for _ in range(2):
                x = dev(torch, x0)
                r = s.solve(x, bd)
                out.append((r["status"], r["iters"], np.asarray(r["hist"]), host(x)))
            s.close()
            runs[(fusion, graphs)] = out
    ref = runs[(False, False)]
    assert ref[0][0] == "success"
    for key, out in runs.items():
        for (sa, ia, ha, xa), (sb, ib, hb, xb) in zip(out, ref):
            assert sa == sb and ia == ib, key
            assert np.array_equal(ha, hb), key
            assert np.array_equal(xa, xb), key
