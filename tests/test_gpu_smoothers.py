"""SOR / SSOR sweeps on the GPU through the C-ABI.

Bar: in level-set order (the default) y equals the generic sequential SOR / SSOR of the
oracle and of the reference's own SORPrecond / SSORPrecond<generic> exactly; in the
multicolour order the colouring equals the oracle's and y equals the oracle's multicolour
sweep exactly (as values: the kernel leaves out lower entries of colours not yet processed,
which the reference multiplies by y = 0, so only the sign of an exact zero can differ).
"""
import numpy as np
import pytest

from oracle.pyoracle import Csr
from tests.test_smoothers_oracle import dominant, q1_csr

pytestmark = pytest.mark.gpu


def upload(feat, torch, A):
    val = torch.from_numpy(np.ascontiguousarray(A.val, np.float64)).cuda()
    col = torch.from_numpy(np.ascontiguousarray(A.col, np.uint64).view(np.int64)).cuda()
    rp = torch.from_numpy(np.ascontiguousarray(A.rowptr, np.uint64).view(np.int64)).cuda()
    return (val, col, rp), feat.CsrF64(A.rows, A.cols, int(A.val.size), val.data_ptr(), col.data_ptr(), rp.data_ptr())


def cases(feat):
    rng = np.random.default_rng(23)
    return {"random": dominant(rng, 3000, max_len=15), "q1_2d_l6": q1_csr(feat, 2, 6), "q1_3d_l4": q1_csr(feat, 3, 4)}


def levels(A, upper=False):
    lev = np.zeros(A.rows, np.int64)
    order = range(A.rows - 1, -1, -1) if upper else range(A.rows)
    for i in order:
        c = A.col[int(A.rowptr[i]):int(A.rowptr[i + 1])].astype(np.int64)
        dep = c[c > i] if upper else c[c < i]
        lev[i] = lev[dep].max() + 1 if dep.size else 0
    return lev


@pytest.mark.parametrize("name", ["random", "q1_2d_l6", "q1_3d_l4"])
def test_level_order_is_generic_sor_ssor(feat, oracle, reference, handle, torch_gpu, name):
    print(f"level-set sor/ssor {name}", flush=True)
    torch = torch_gpu
    k, _ = feat.libs()
    A = cases(feat)[name]
    keep, desc = upload(feat, torch, A)
    plan = feat.SorPlan(handle, desc)
    for bwd in (False, True):
        rank, rows = plan.colors(backward=bwd)
        assert np.array_equal(rank, levels(A, upper=bwd))
    if name == "q1_3d_l4":
        assert plan.ncolors == 7 * 16 + 1
    rng = np.random.default_rng(4)
    x = rng.standard_normal(A.rows)
    xd = torch.from_numpy(x).cuda()
    y = torch.empty_like(xd)
    for omega in (1.0, 0.8, 1.3):
        plan.sor(y, xd, omega)
        torch.cuda.synchronize()
        want = oracle.sor_generic_apply(A, omega, x)
        assert np.array_equal(y.cpu().numpy(), want)
        if name != "random":
            assert np.array_equal(want, reference.precond_apply("sor", A, omega, x))
        plan.ssor(y, xd, omega)
        feat.check(k.feat_scale_f64(handle.h, feat.ptr(y), feat.ptr(y), omega * (2.0 - omega), A.rows), "scale")
        torch.cuda.synchronize()
        want = oracle.sor_generic_apply(A, omega, x, ssor=True)
        assert np.array_equal(y.cpu().numpy(), want)
        if name != "random":
            assert np.array_equal(want, reference.precond_apply("ssor", A, omega, x))
    # new values, same pattern: init_numeric re-reads them
    keep[0].mul_(0.5)
    A2 = Csr(A.rows, A.cols, A.val * 0.5, A.col, A.rowptr)
    plan.init_numeric()
    plan.sor(y, xd, 1.1)
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), oracle.sor_generic_apply(A2, 1.1, x))
    plan.close()


@pytest.mark.parametrize("name", ["random", "q1_2d_l6", "q1_3d_l4"])
def test_multicolor_sor_ssor_vs_oracle(feat, oracle, handle, torch_gpu, name):
    print(f"multicolour sor/ssor {name}", flush=True)
    torch = torch_gpu
    k, _ = feat.libs()
    A = cases(feat)[name]
    keep, desc = upload(feat, torch, A)
    plan = feat.SorPlan(handle, desc, multicolor=True)
    rank, rows = plan.colors()
    orank, onc = oracle.color_ranks(A)
    assert plan.ncolors == onc and np.array_equal(rank, orank)
    assert np.array_equal(plan.colors(backward=True)[0], orank)
    assert rows.tolist() == np.bincount(orank, minlength=onc).tolist()
    if name == "q1_3d_l4":
        assert onc == 8
    rng = np.random.default_rng(4)
    x = rng.standard_normal(A.rows)
    xd = torch.from_numpy(x).cuda()
    y = torch.empty_like(xd)
    for omega in (1.0, 0.8):
        plan.sor(y, xd, omega)
        torch.cuda.synchronize()
        assert np.array_equal(y.cpu().numpy(), oracle.mcsor_apply(A, orank, onc, omega, x))
        plan.ssor(y, xd, omega)
        feat.check(k.feat_scale_f64(handle.h, feat.ptr(y), feat.ptr(y), omega * (2.0 - omega), A.rows), "scale")
        torch.cuda.synchronize()
        assert np.array_equal(y.cpu().numpy(), oracle.mcsor_apply(A, orank, onc, omega, x, ssor=True))
    plan.close()


def test_sor_rejects_bad_input(feat, handle, torch_gpu):
    torch = torch_gpu
    A = Csr(3, 3, np.array([1.0, 2.0, 3.0]), np.array([0, 2, 1], np.uint64), np.array([0, 1, 2, 3], np.uint64))
    keep, desc = upload(feat, torch, A)      # row 1 has no diagonal entry
    with pytest.raises(feat.FeatError):
        feat.SorPlan(handle, desc)


def np_krylov(oracle, A, filt, prec, b, tol, maxit, pcg=True, omega_rich=1.0):
    """pcg.hpp:188-288 / richardson.hpp:160-226 in numpy, preconditioner prec(r)."""
    fl = filt.astype(np.int64)

    def filt_def(v):
        v[fl] = 0.0
        return v

    x = np.zeros_like(b)
    r = filt_def(b - oracle.csr_apply(A, x))
    d0 = np.linalg.norm(r)
    hist = [d0]
    if not pcg:
        for _ in range(maxit):
            x = x + omega_rich * filt_def(prec(r))
            r = filt_def(b - oracle.csr_apply(A, x))
            hist.append(np.linalg.norm(r))
            if hist[-1] <= tol * d0:
                break
        return x, np.array(hist)
    p = filt_def(prec(r))
    gamma = r @ p
    for _ in range(maxit):
        q = filt_def(oracle.csr_apply(A, p))
        alpha = gamma / (q @ p)
        x = x + alpha * p
        r = r - alpha * q
        hist.append(np.linalg.norm(r))
        if hist[-1] <= tol * d0:
            break
        z = filt_def(prec(r))
        gamma2, gamma = gamma, r @ z
        p = z + (gamma / gamma2) * p
    return x, np.array(hist)


@pytest.mark.parametrize("solver,precond,omega,dim,level", [("pcg", "ssor", 1.0, 2, 5), ("pcg", "ssor", 1.2, 3, 3),
                                                            ("richardson", "sor", 1.0, 2, 4),
                                                            ("richardson", "ssor_mc", 1.0, 2, 4),
                                                            ("pcg", "ilu", 0.0, 2, 5), ("pcg", "ilu", 0.0, 3, 3),
                                                            ("pcg", "jacobi", 0.8, 2, 5)])
def test_krylov_with_device_precond(feat, oracle, torch_gpu, solver, precond, omega, dim, level):
    """The solver tree around the device preconditioners follows the numpy restatement
    (iteration count equal, defect history and solution to 1e-10); for ILU, omega is the
    fill-in level p (0 here)."""
    print(f"krylov {solver}-{precond} omega={omega} dim={dim} level={level}", flush=True)
    torch = torch_gpu
    q = feat.q1_problem(dim, level)
    A = Csr(q["n"], q["n"], q["val"], q["col"], q["rowptr"])
    b = np.random.default_rng(9).uniform(-1, 1, A.rows)
    b[q["filt"].astype(np.int64)] = 0.0
    rank, nc = oracle.color_ranks(A)
    ilu = oracle.ilu0_factor(A)
    inv_diag = omega / np.array([q["val"][int(q["rowptr"][i]) + int(np.searchsorted(
        q["col"][int(q["rowptr"][i]):int(q["rowptr"][i + 1])], i))] for i in range(A.rows)])
    prec = {"ssor": lambda r: oracle.sor_generic_apply(A, omega, r, ssor=True),
            "sor": lambda r: oracle.sor_generic_apply(A, omega, r),
            "ssor_mc": lambda r: oracle.mcsor_apply(A, rank, nc, omega, r, ssor=True),
            "ilu": lambda r: oracle.ilu0_apply(A, ilu, r),
            "jacobi": lambda r: inv_diag * r}[precond]
    maxit = 2000
    x_ref, h_ref = np_krylov(oracle, A, q["filt"], prec, b, 1e-8, maxit, pcg=(solver == "pcg"))
    s = feat.KrylovSolver(solver, A, q["filt"], precond, omega_pre=omega, tol_rel=1e-8, max_iter=maxit)
    x = torch.zeros(A.rows, dtype=torch.float64, device="cuda")
    res = s.solve(x, torch.from_numpy(b).cuda(), hist_cap=maxit + 1)
    s.close()
    assert res["status"] == "success"
    assert res["iters"] == len(h_ref) - 1
    assert np.allclose(res["hist"], h_ref, rtol=1e-10, atol=0)
    xs = x.cpu().numpy()
    assert np.linalg.norm(xs - x_ref) <= 1e-10 * np.linalg.norm(x_ref)


@pytest.mark.parametrize("name", ["random", "q1_2d_l6", "q1_3d_l4"])
def test_ilu0_vs_oracle_and_reference(feat, oracle, reference, handle, torch_gpu, name):
    print(f"ilu0 {name}", flush=True)
    torch = torch_gpu
    A = cases(feat)[name]
    keep, desc = upload(feat, torch, A)
    plan = feat.IluPlan(handle, desc)
    l, u, d, rl, ru = oracle.ilu0_factor(A)
    assert (plan.nnz_l, plan.nnz_u) == (l.size, u.size)
    assert plan.levels_l == int(levels(A).max()) + 1 and plan.levels_u == int(levels(A, upper=True).max()) + 1
    gl, gu, gd = plan.factors()
    assert np.array_equal(gl, l) and np.array_equal(gu, u) and np.array_equal(gd, d)
    x = np.random.default_rng(6).standard_normal(A.rows)
    xd = torch.from_numpy(x).cuda()
    y = torch.empty_like(xd)
    plan.apply(y, xd)
    torch.cuda.synchronize()
    want = oracle.ilu0_apply(A, (l, u, d, rl, ru), x)
    assert np.array_equal(y.cpu().numpy(), want)
    if name != "random":
        assert np.array_equal(want, reference.precond_apply("ilu", A, 0, x))
    plan.apply(xd, xd)        # in place, as solve_il / solve_du allow
    torch.cuda.synchronize()
    assert np.array_equal(xd.cpu().numpy(), want)
    plan.close()


@pytest.mark.parametrize("p", [1, 2])
@pytest.mark.parametrize("name", ["random", "q1_2d_l6", "q1_3d_l4"])
def test_ilup_vs_oracle(feat, oracle, handle, torch_gpu, name, p):
    """ILU(p): the level-p pattern, the factors on it (zeros at fill-in) and both solves equal
    the oracle bit for bit (the oracle's ILU(p) is pinned to the reference's ILUPrecond)."""
    print(f"ilu({p}) {name}", flush=True)
    torch = torch_gpu
    A = cases(feat)[name]
    keep, desc = upload(feat, torch, A)
    plan = feat.IluPlan(handle, desc, p=p)
    f = oracle.ilup_factor(A, p)
    assert (plan.nnz_l, plan.nnz_u) == (f["l"].size, f["u"].size)
    assert f["l"].size > int(np.sum(A.col < np.repeat(np.arange(A.rows), np.diff(A.rowptr.astype(np.int64)))))
    gl, gu, gd = plan.factors()
    assert np.array_equal(gl, f["l"]) and np.array_equal(gu, f["u"]) and np.array_equal(gd, f["d"])
    x = np.random.default_rng(7).standard_normal(A.rows)
    xd = torch.from_numpy(x).cuda()
    y = torch.empty_like(xd)
    plan.apply(y, xd)
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), oracle.ilup_apply(f, x))
    plan.close()


def test_pcg_ilup_iterations(feat, oracle, torch_gpu):
    """PCG with ILU(0), ILU(1) and ILU(2) follows the numpy restatement around the oracle's
    ILU(p) (same iterations, history to 1e-10).  (On this 27-point stencil, whose face
    couplings are stored zeros, more fill does not lower the iteration count.)"""
    torch = torch_gpu
    q = feat.q1_problem(3, 4)
    A = Csr(q["n"], q["n"], q["val"], q["col"], q["rowptr"])
    b = np.random.default_rng(10).uniform(-1, 1, A.rows)
    b[q["filt"].astype(np.int64)] = 0.0
    its = []
    for p in (0, 1, 2):
        f = oracle.ilup_factor(A, p)
        x_ref, h_ref = np_krylov(oracle, A, q["filt"], lambda r: oracle.ilup_apply(f, r), b, 1e-8, 500, pcg=True)
        s = feat.KrylovSolver("pcg", A, q["filt"], "ilu", omega_pre=float(p), tol_rel=1e-8, max_iter=500)
        x = torch.zeros(A.rows, dtype=torch.float64, device="cuda")
        res = s.solve(x, torch.from_numpy(b).cuda(), hist_cap=501)
        s.close()
        assert res["status"] == "success" and res["iters"] == len(h_ref) - 1, (p, res["iters"], len(h_ref) - 1)
        assert np.allclose(res["hist"], h_ref, rtol=1e-10, atol=0)
        its.append(res["iters"])
    print(f"PCG-ILU(p) iterations, 3D level 4: {its}")
