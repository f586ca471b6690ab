"""Multicolour SOR / SSOR on the GPU through the C-ABI, against the oracle.

Bar: the colouring equals the oracle's; y equals the oracle's multicolour sweep
exactly (as values: the kernel leaves out lower entries of colours not yet processed,
which the reference multiplies by y = 0, so only the sign of an exact zero can differ);
with one colour per row the GPU SOR equals the reference's generic SORPrecond.
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


@pytest.mark.parametrize("name", ["random", "q1_2d_l6", "q1_3d_l4"])
def test_sor_ssor_vs_oracle(feat, oracle, handle, torch_gpu, name):
    print(f"multicolour sor/ssor {name}", flush=True)
    torch = torch_gpu
    k, _ = feat.libs()
    A = cases(feat)[name]
    keep, desc = upload(feat, torch, A)
    plan = feat.SorPlan(handle, desc)
    rank, rows = plan.colors()
    orank, onc = oracle.color_ranks(A)
    assert plan.ncolors == onc and np.array_equal(rank, orank)
    assert rows.tolist() == np.bincount(orank, minlength=onc).tolist()
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
    # new values, same pattern: init_numeric re-reads them
    keep[0].mul_(0.5)
    A2 = Csr(A.rows, A.cols, A.val * 0.5, A.col, A.rowptr)
    plan.init_numeric()
    plan.sor(y, xd, 1.1)
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), oracle.mcsor_apply(A2, orank, onc, 1.1, x))
    plan.close()


def test_row_colors_sor_is_reference_generic(feat, reference, handle, torch_gpu):
    print("row-colour sor vs reference generic", flush=True)
    torch = torch_gpu
    A = q1_csr(feat, 2, 3)
    keep, desc = upload(feat, torch, A)
    plan = feat.SorPlan(handle, desc, row_colors=True)
    assert plan.ncolors == A.rows
    x = np.random.default_rng(8).standard_normal(A.rows)
    xd = torch.from_numpy(x).cuda()
    y = torch.empty_like(xd)
    plan.sor(y, xd, 1.2)
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), reference.precond_apply("sor", A, 1.2, x))
    plan.close()


def test_sor_rejects_bad_input(feat, handle, torch_gpu):
    torch = torch_gpu
    A = Csr(3, 3, np.array([1.0, 2.0, 3.0]), np.array([0, 2, 1], np.uint64), np.array([0, 1, 2, 3], np.uint64))
    keep, desc = upload(feat, torch, A)      # row 1 has no diagonal entry
    with pytest.raises(feat.FeatError):
        feat.SorPlan(handle, desc)
