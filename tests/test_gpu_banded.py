"""Banded plans of Q1 matrices with variable coefficients on the GPU.  Such matrices have
no pair dictionary; their rows are the clipped 3^dim-point stencils of a structured grid,
so the plan keeps only the values (slot-major) and forms each column from the node
position.  The row sums run in the CSR order: every result is bit-identical to the plain
SELL-64 plan, and r = A x to the oracle's CSR loop (FP64, FP32)."""
import numpy as np
import pytest

from oracle.pyoracle import Csr
from tests.test_gpu_lafem import dev, host, upload_csr
from tests.test_gpu_solver import mask_of, _same_bits

pytestmark = pytest.mark.gpu


def varcoef_level(feat, dim, lvl, seed):
    q = feat.q1_problem(dim, lvl)
    if dim == 2:
        val = feat.q1_diffusion(lvl, lvl, feat.lognormal_cells(seed, 4 ** lvl, 1.5))
    else:
        val = feat.q1_diffusion3d(lvl, lvl, feat.lognormal_cells(seed, 8 ** lvl, 1.5), 10.0)
    return q, Csr(q["n"], q["n"], val, q["col"], q["rowptr"])


@pytest.mark.parametrize("fp", ["f64", "f32", "f16"])
@pytest.mark.parametrize("dim,lvl", [(2, 5), (3, 4), (3, 2), (2, 6)])
def test_banded_plan_bitwise(feat, handle, torch_gpu, oracle, dim, lvl, fp):
    print(f"banded plan dim={dim} level={lvl} {fp}", flush=True)
    torch = torch_gpu
    k, _ = feat.libs()
    P = feat.ptr
    q, A = varcoef_level(feat, dim, lvl, 4100 + lvl)
    n = A.rows
    v, c, r = upload_csr(torch, A)
    band = feat.SellPlan(handle, v, c, r, n, n, fp=fp)
    plain = feat.SellPlan(handle, v, c, r, n, n, fp=fp, banded=False)
    try:
        assert band.m.format == feat.FMT_BANDED and band.m.slice_words == 2 ** lvl + 1
        assert band.m.nnz == 3 ** dim * n
        assert plain.m.format == feat.FMT_SELL
        rng = np.random.default_rng(lvl + 10 * dim)
        x, rhs, y = rng.standard_normal((3, n))
        keep, ws = feat.reduce_ws(torch, "cuda")
        if fp == "f64":
            xd, yd, rd = dev(torch, x), dev(torch, y), dev(torch, rhs)
            m = mask_of(torch, feat, handle, n, q["filt"])
            outs = []
            for pl in (band, plain):
                o = torch.empty(n, dtype=torch.float64, device="cuda")
                feat.check(k.feat_matrix_apply_f64(handle.h, pl.m, P(o), -0.75, P(xd), 1.5, P(yd)), "apply")
                d = torch.empty(n, dtype=torch.float64, device="cuda")
                scal = torch.zeros(4, dtype=torch.float64, device="cuda")
                feat.check(k.feat_defect_f64(handle.h, pl.m, P(d), P(rd), P(xd), P(m), P(scal), ws), "defect")
                handle.sync()
                outs.append((host(o), host(d), host(scal)))
            _same_bits(outs[0][0], outs[1][0])
            _same_bits(outs[0][0], oracle.csr_apply(A, x, y, -0.75, 1.5))
            _same_bits(outs[0][1], outs[1][1])
            _same_bits(outs[0][2], outs[1][2])
        else:
            x32 = x.astype(np.float32)
            xd = dev(torch, x32)
            outs = []
            for pl in (band, plain):
                o = torch.empty(n, dtype=torch.float32, device="cuda")
                feat.check(k.feat_apply_masked_f32(handle.h, pl.m, P(o), P(xd), None, None, ws), "apply f32")
                handle.sync()
                outs.append(host(o))
            _same_bits(outs[0], outs[1])
            if fp == "f32":
                _same_bits(outs[0], oracle.csr_apply_f32(A, x32))
    finally:
        band.close()
        plain.close()


def test_banded_plan_rejects_other_patterns(feat, handle, torch_gpu):
    # a Q1 matrix with one entry removed from an interior row, and the constant-coefficient
    # Q1 matrix (pair dictionary), keep their SELL-64 plans
    torch = torch_gpu
    q, A = varcoef_level(feat, 3, 3, 77)
    n = A.rows
    rp = A.rowptr.astype(np.int64)
    row = n // 2
    drop = rp[row] + 1
    col = np.delete(A.col, drop)
    val = np.delete(A.val, drop)
    rp2 = rp.copy()
    rp2[row + 1:] -= 1
    B = Csr(n, n, val, col, rp2.astype(np.uint64))
    v, c, r = upload_csr(torch, B)
    p = feat.SellPlan(handle, v, c, r, n, n)
    assert p.m.format == feat.FMT_SELL
    p.close()
    qc = feat.q1_problem(3, 3)
    C0 = Csr(qc["n"], qc["n"], qc["val"], qc["col"], qc["rowptr"])
    v, c, r = upload_csr(torch, C0)
    p = feat.SellPlan(handle, v, c, r, C0.rows, C0.cols)
    assert p.m.format == feat.FMT_SELL and p.m.index in (feat.IDX_DICT, feat.IDX_DIAG)
    p.close()


@pytest.mark.parametrize("prec", ["f64", "f32", "f16"])
def test_banded_mgpcg_bitwise(feat, torch_gpu, prec):
    # the variable-coefficient hierarchy (3D, levels 5..1): PCG + MG V(4,4) on banded plans
    # gives the same iterates as on plain SELL-64 plans, in every precision of the V-cycle
    print(f"banded MG-PCG {prec}", flush=True)
    torch = torch_gpu
    lv = feat.diffusion3d_levels(5, 1, 4242, 1.5, 10.0)
    n = lv[0]["n"]
    bh = feat.rhs_random(4343, n, -1.0, 1.0)
    bh[lv[0]["filt"].astype(np.int64)] = 0.0
    b = dev(torch, bh)
    res = []
    try:
        for flags in (0, feat.SELL_NO_BANDED):
            feat.runtime_sell_flags(flags)
            s = feat.MgPcg(feat.mgpcg_cfg(coarse="direct", max_iter=300, precision=prec, graphs=False), levels=lv)
            x = torch.zeros(n, dtype=torch.float64, device="cuda")
            r = s.solve(x, b)
            s.close()
            res.append((host(x), r["iters"], r["status"]))
    finally:
        feat.runtime_sell_flags(0)
    assert res[0][2] == res[1][2] == "success"
    assert res[0][1] == res[1][1]
    assert np.array_equal(res[0][0], res[1][0])


@pytest.mark.parametrize("dim,lf", [(2, 7), (3, 5)])
def test_f16_cycle_dlinear_transfers_bitwise(feat, torch_gpu, dim, lf):
    # the FP16-valued V-cycle on matrix-free d-linear transfers (grid forms on) and on its
    # FP16 transfer plans (off): the same iterates
    from tests.helpers import q1_levels
    torch = torch_gpu
    levels = q1_levels(feat, dim, lf, 1)
    n = levels[0]["n"]
    bh = feat.rhs_random(28054777172512, n)
    bh[levels[0]["filt"].astype(np.int64)] = 0.0
    b = dev(torch, bh)
    res = []
    try:
        for grid in (True, False):
            feat.runtime_grid_forms(grid)
            s = feat.MgPcg(feat.mgpcg_cfg(coarse="direct", precision="f16"), levels=levels)
            x = torch.zeros(n, dtype=torch.float64, device="cuda")
            r = s.solve(x, b)
            s.close()
            res.append((host(x), r["iters"], r["status"]))
    finally:
        feat.runtime_grid_forms(True)
    assert res[0][2] == res[1][2] == "success"
    assert res[0][1] == res[1][1]
    assert np.array_equal(res[0][0], res[1][0])
