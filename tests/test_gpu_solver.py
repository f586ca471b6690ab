"""GPU parity of the fused solver kernels and of PCG(MG-V(Richardson(Jacobi)))
against the oracle.

Bars: fused kernels bit-exact except their dot/norm outputs (1e-13 relative);
solver iteration counts equal to the oracle's within +-1; solution error
|x - x_oracle| / |x_oracle| <= 1e-11; final relative defect <= tol.
"""
import numpy as np
import pytest

from oracle.pyoracle import Csr, COARSE_JACOBI, COARSE_DIRECT, COARSE_SMOOTHER
from tests.helpers import q1_levels
from tests.test_gpu_lafem import dev, host, upload_csr

pytestmark = pytest.mark.gpu

SOL_TOL = 1e-11


def mask_of(torch, feat, handle, n, idx):
    k, _ = feat.libs()
    m = torch.zeros((n + 31) // 32, dtype=torch.int32, device="cuda")
    idd = dev(torch, np.asarray(idx, np.int64))
    feat.check(k.feat_unit_filter_mask_u64(handle.h, feat.ptr(m), n, feat.ptr(idd), len(idx)), "mask")
    return m


def q1_level(feat, dim, lvl):
    q = feat.q1_problem(dim, lvl)
    return q, Csr(q["n"], q["n"], q["val"], q["col"], q["rowptr"])


def test_fused_defect_masked_prol(feat, handle, torch_gpu, oracle):
    torch = torch_gpu
    k, _ = feat.libs()
    P = feat.ptr
    keep, ws = feat.reduce_ws(torch, "cuda")
    q, A = q1_level(feat, 3, 4)
    n = q["n"]
    rng = np.random.default_rng(21)
    x, rhs, w = rng.standard_normal((3, n))
    v, c, r = upload_csr(torch, A)
    desc = feat.csr_desc(v, c, r, n, n)
    m = mask_of(torch, feat, handle, n, q["filt"])
    scal = torch.zeros(4, dtype=torch.float64, device="cuda")
    d = torch.empty(n, dtype=torch.float64, device="cuda")
    xd, rd, wd = dev(torch, x), dev(torch, rhs), dev(torch, w)
    feat.check(k.feat_csr_defect_f64(handle.h, desc, P(d), P(rd), P(xd), P(m), P(scal), ws), "defect")
    handle.sync()
    ref = oracle.unit_filter_def(oracle.csr_apply(A, x, rhs, -1.0, 1.0), q["filt"])
    assert np.array_equal(host(d), ref)
    assert abs(host(scal)[0] - ref @ ref) <= 1e-13 * (ref @ ref)
    # masked apply with norm and dot
    feat.check(k.feat_csr_apply_masked_f64(handle.h, desc, P(d), P(xd), P(m), P(scal[1:]), P(wd), P(scal[2:]), ws), "masked")
    handle.sync()
    ref = oracle.unit_filter_def(oracle.csr_apply(A, x), q["filt"])
    assert np.array_equal(host(d), ref)
    s = host(scal)
    assert abs(s[1] - ref @ ref) <= 1e-13 * (ref @ ref)
    assert abs(s[2] - ref @ w) <= 1e-13 * np.abs(ref * w).sum()
    # prolongation + correction
    Pm = Csr(n, q["pcols"], q["pval"], q["pcol"], q["prowptr"])
    pv, pc, pr = upload_csr(torch, Pm)
    pdesc = feat.csr_desc(pv, pc, pr, n, q["pcols"])
    xc = rng.standard_normal(q["pcols"])
    sol = dev(torch, x)
    xcd = dev(torch, xc)
    feat.check(k.feat_csr_prol_add_f64(handle.h, pdesc, P(sol), P(xcd), P(m)), "prol")
    handle.sync()
    cor = oracle.unit_filter_def(oracle.csr_apply(Pm, xc), q["filt"])
    assert np.array_equal(host(sol), oracle.axpy(1.0, cor, x))


@pytest.mark.parametrize("steps", [1, 2, 4])
def test_fused_richardson_jacobi(feat, handle, torch_gpu, oracle, steps):
    torch = torch_gpu
    k, _ = feat.libs()
    P = feat.ptr
    keep, ws = feat.reduce_ws(torch, "cuda")
    q, A = q1_level(feat, 2, 6)
    n = q["n"]
    rng = np.random.default_rng(22)
    rhs = oracle.unit_filter_def(rng.standard_normal(n), q["filt"])
    v, c, r = upload_csr(torch, A)
    desc = feat.csr_desc(v, c, r, n, n)
    m = mask_of(torch, feat, handle, n, q["filt"])
    omega_j, omega_r = 0.5, 0.9
    inv = torch.empty(n, dtype=torch.float64, device="cuda")
    feat.check(k.feat_csr_extract_diag_f64_u64(handle.h, P(inv), P(v), P(c), P(r), n), "diag")
    feat.check(k.feat_component_invert_f64(handle.h, P(inv), P(inv), omega_j, n), "inv")
    rd = dev(torch, rhs)
    nrm = torch.tensor([rhs @ rhs], dtype=torch.float64, device="cuda")
    x = torch.empty(n, dtype=torch.float64, device="cuda")
    tmp = torch.empty(n, dtype=torch.float64, device="cuda")
    dout = torch.empty(n, dtype=torch.float64, device="cuda")
    # oracle: Richardson::apply with min_iter = max_iter = steps (richardson.hpp:160-226)
    o = oracle.richardson(A, np.zeros(n), rhs, filt=q["filt"], precond=1, omega_jac=omega_j,
                          omega_rich=omega_r, apply=True, max_iter=steps, min_iter=steps)
    assert o["iters"] == steps
    xo = o["x"]
    do = oracle.unit_filter_def(oracle.csr_apply(A, xo, rhs, -1.0, 1.0), q["filt"])
    # pre-smoothing form: x and the final defect
    feat.check(k.feat_rj_smooth_f64(handle.h, desc, P(inv), P(m), steps, omega_r, P(rd), P(nrm), P(x), P(tmp),
                                    P(dout), None, None, None, ws), "smooth pre")
    handle.sync()
    assert np.array_equal(host(x), xo)
    assert np.array_equal(host(dout), do)
    # plain form
    feat.check(k.feat_rj_smooth_f64(handle.h, desc, P(inv), P(m), steps, omega_r, P(rd), P(nrm), P(x), P(tmp),
                                    None, None, None, None, ws), "smooth plain")
    handle.sync()
    assert np.array_equal(host(x), xo)
    # post-smoothing form: acc += x, with dot(acc, w)
    base = rng.standard_normal(n)
    w = rng.standard_normal(n)
    acc = dev(torch, base)
    wd = dev(torch, w)
    dot = torch.zeros(1, dtype=torch.float64, device="cuda")
    feat.check(k.feat_rj_smooth_f64(handle.h, desc, P(inv), P(m), steps, omega_r, P(rd), P(nrm), P(x), P(tmp),
                                    None, P(acc), P(wd), P(dot), ws), "smooth acc")
    handle.sync()
    ref = oracle.axpy(1.0, xo, base)
    assert np.array_equal(host(acc), ref)
    assert abs(host(dot)[0] - ref @ w) <= 1e-13 * np.abs(ref * w).sum()
    # skip rule: a zero right-hand side leaves x = 0 and d = rhs (iterative.hpp:761-766)
    zero = torch.zeros(n, dtype=torch.float64, device="cuda")
    z2 = torch.zeros(1, dtype=torch.float64, device="cuda")
    feat.check(k.feat_rj_smooth_f64(handle.h, desc, P(inv), P(m), steps, omega_r, P(zero), P(z2), P(x), P(tmp),
                                    P(dout), None, None, None, ws), "smooth skip")
    handle.sync()
    assert not host(x).any() and not host(dout).any()


def c1_problem(feat):
    levels = q1_levels(feat, 2, 7, 1)
    n = levels[0]["n"]
    b = feat.q1_rhs_sine(2, 7, n)
    b[levels[0]["filt"].astype(np.int64)] = 0.0
    return levels, b


@pytest.mark.parametrize("generic", [False, True])
def test_mgpcg_c1_vs_reference_golden(feat, torch_gpu, golden, generic):
    torch = torch_gpu
    levels, b = c1_problem(feat)
    assert np.array_equal(b, golden["c1_rhs"])
    cfg = feat.mgpcg_cfg(coarse="jacobi", force_generic=generic)
    s = feat.MgPcg(cfg, levels=levels)
    x = torch.zeros(b.size, dtype=torch.float64, device="cuda")
    r = s.solve(x, dev(torch, b))
    s.close()
    ref_it = int(golden["c1_iters"][0])
    assert r["status"] == "success" and abs(r["iters"] - ref_it) <= 1
    xr = golden["c1_x"]
    assert np.linalg.norm(host(x) - xr) <= SOL_TOL * np.linalg.norm(xr)
    assert r["def_final"] <= 1e-8 * r["def_init"]
    assert np.allclose(r["hist"][:ref_it + 1], golden["c1_hist"][:r["iters"] + 1], rtol=1e-9)


@pytest.mark.parametrize("dim,lf,lc,coarse", [
    (2, 7, 1, "direct"),
    (2, 5, 0, "smoother"),
    (3, 5, 1, "direct"),
    (3, 5, 1, "jacobi"),
    (3, 6, 2, "direct"),
])
def test_mgpcg_vs_oracle(feat, torch_gpu, oracle, dim, lf, lc, coarse):
    torch = torch_gpu
    levels = q1_levels(feat, dim, lf, lc)
    n = levels[0]["n"]
    b = feat.rhs_random(28054777172512, n)
    b[levels[0]["filt"].astype(np.int64)] = 0.0
    ocoarse = {"direct": COARSE_DIRECT, "jacobi": COARSE_JACOBI, "smoother": COARSE_SMOOTHER}[coarse]
    o = oracle.mg_pcg_solve(levels, np.zeros(n), b, coarse=ocoarse)
    for generic in (False, True):
        s = feat.MgPcg(feat.mgpcg_cfg(coarse=coarse, force_generic=generic), levels=levels)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        r = s.solve(x, dev(torch, b))
        s.close()
        assert r["status"] == "success"
        assert abs(r["iters"] - o["iters"]) <= 1, (generic, r["iters"], o["iters"])
        assert np.linalg.norm(host(x) - o["x"]) <= SOL_TOL * np.linalg.norm(o["x"]), generic
        assert r["def_final"] <= 1e-8 * r["def_init"]


def test_mgpcg_q1_builder_matches_level_upload(feat, torch_gpu):
    # the host layer's own hierarchy builder gives the same solve as uploaded levels
    torch = torch_gpu
    levels = q1_levels(feat, 3, 5, 1)
    n = levels[0]["n"]
    b = dev(torch, np.where(np.isin(np.arange(n), levels[0]["filt"]), 0.0, feat.rhs_random(5, n)))
    outs = []
    for kw in (dict(levels=levels), dict(q1=(3, 5, 1))):
        s = feat.MgPcg(feat.mgpcg_cfg(), **kw)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        outs.append((s.solve(x, b), host(x)))
        s.close()
    assert outs[0][0]["iters"] == outs[1][0]["iters"]
    assert np.array_equal(outs[0][1], outs[1][1])


def test_mgpcg_failure_status(feat, torch_gpu):
    # non-finite right-hand side -> Status::aborted (iterative.hpp:761); max_iter cap -> max_iter
    torch = torch_gpu
    s = feat.MgPcg(feat.mgpcg_cfg(max_iter=2), q1=(2, 5, 1))
    n = 33 * 33
    b = torch.zeros(n, dtype=torch.float64, device="cuda")
    b[100] = float("nan")
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert s.solve(x, b)["status"] == "aborted"
    b = dev(torch, feat.rhs_random(9, n))
    x.zero_()
    r = s.solve(x, b)
    s.close()
    assert r["status"] == "max_iter" and r["iters"] == 2
