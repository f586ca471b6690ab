"""GPU assembly of the structured Q1 hierarchy, through the C-ABI.

Bar: bit-exact against the host assembly (stiffness values, columns and row
pointers, boundary list, prolongation) and against the host transpose of the
prolongation for the restriction; the host assembly itself is pinned to the
reference's assembled matrices by the oracle tests.
"""
import numpy as np
import pytest

from oracle.pyoracle import Csr, csr_transpose

pytestmark = pytest.mark.gpu

CASES = [(2, 1), (2, 2), (2, 5), (2, 9), (3, 1), (3, 2), (3, 4), (3, 6)]


def host(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("dim,level", CASES)
def test_q1_device_assembly_bitwise(feat, handle, torch_gpu, dim, level):
    print(f"q1 device assembly dim={dim} level={level}", flush=True)
    q = feat.q1_problem(dim, level)
    d = feat.q1_problem_device(torch_gpu, handle, dim, level)
    torch_gpu.cuda.synchronize()
    assert d["n"] == q["n"]
    assert np.array_equal(host(d["rowptr"]).view(np.uint64), q["rowptr"])
    assert np.array_equal(host(d["col"]).view(np.uint64), q["col"])
    assert np.array_equal(host(d["val"]).view(np.uint64), q["val"].view(np.uint64))
    assert np.array_equal(host(d["filt"]).view(np.uint64), q["filt"])
    P = Csr(q["n"], q["pcols"], q["pval"], q["pcol"], q["prowptr"])
    assert d["pcols"] == P.cols
    assert np.array_equal(host(d["prowptr"]).view(np.uint64), P.rowptr)
    assert np.array_equal(host(d["pcol"]).view(np.uint64), P.col)
    assert np.array_equal(host(d["pval"]).view(np.uint64), P.val.view(np.uint64))
    R = csr_transpose(P)
    assert np.array_equal(host(d["rrowptr"]).view(np.uint64), R.rowptr.astype(np.uint64))
    assert np.array_equal(host(d["rcol"]).view(np.uint64), R.col.astype(np.uint64))
    assert np.array_equal(host(d["rval"]).view(np.uint64), R.val.view(np.uint64))


def test_q1_device_assembly_level0(feat, handle, torch_gpu):
    q = feat.q1_problem(3, 0)
    d = feat.q1_problem_device(torch_gpu, handle, 3, 0)
    torch_gpu.cuda.synchronize()
    assert "pval" not in d
    assert np.array_equal(host(d["val"]).view(np.uint64), q["val"].view(np.uint64))
    assert np.array_equal(host(d["filt"]).view(np.uint64), q["filt"])


@pytest.mark.parametrize("dim,lf,lc", [(2, 6, 2), (3, 4, 1)])
def test_mgpcg_device_vs_host_assembly(feat, torch_gpu, dim, lf, lc):
    """The hierarchy assembled on the device solves exactly like the uploaded host one."""
    print(f"mgpcg device vs host assembly dim={dim} levels {lf}..{lc}", flush=True)
    torch = torch_gpu
    cfg = feat.mgpcg_cfg(nu_pre=4, nu_post=4, omega_jac=0.5, omega_rich=1.0, coarse="direct",
                         tol_rel=1e-8, max_iter=100)
    q = feat.q1_problem(dim, lf)
    b_host = feat.rhs_random(7, q["n"], -1.0, 1.0)
    b_host[q["filt"].astype(np.int64)] = 0.0
    b = torch.from_numpy(b_host).cuda()
    out = []
    for host_assembly in (False, True):
        s = feat.MgPcg(cfg, q1=(dim, lf, lc), host_assembly=host_assembly)
        x = torch.zeros(q["n"], dtype=torch.float64, device="cuda")
        r = s.solve(x, b)
        torch.cuda.synchronize()
        s.close()
        assert r["status"] == "success"
        out.append((r, x.cpu().numpy()))
    (rd, xd), (rh, xh) = out
    assert rd["iters"] == rh["iters"]
    assert np.array_equal(rd["hist"].view(np.uint64), rh["hist"].view(np.uint64))
    assert np.array_equal(xd.view(np.uint64), xh.view(np.uint64))


def _csr_host_transpose(val, col, ptr, rows, cols):
    return csr_transpose(Csr(rows, cols, val, col, ptr))


@pytest.mark.parametrize("level", [0, 1, 2, 4, 6, 7])
def test_stokes_device_assembly_bitwise(feat, handle, torch_gpu, level):
    """Q2/P1dc level: A, B, D, filters, prolongations bit-identical to the host assembly,
    restrictions bit-identical to the host transpose of the prolongations."""
    print(f"stokes device assembly level={level}", flush=True)
    S = feat.stokes_problem(level)
    d = feat.stokes_problem_device(torch_gpu, handle, level)
    torch_gpu.cuda.synchronize()
    assert (d["nv"], d["np"]) == (S["nv"], S["np"])
    u = lambda t: host(t).view(np.uint64)
    for key in ("A", "B", "D"):
        for hv, dv in zip(S[key], d[key]):
            assert np.array_equal(u(dv), hv.view(np.uint64)), key
    for key in ("filt", "filt_val", "prim", "dual"):
        assert np.array_equal(u(d[key]), S[key].view(np.uint64)), key
    if level == 0:
        assert "Pv" not in d
        return
    for key, rows in (("Pv", S["nv"]), ("Pp", S["np"])):
        hv, hc, hp, cols = S[key]
        dv, dc, dp, dcols = d[key]
        assert dcols == cols
        assert np.array_equal(u(dv), hv.view(np.uint64)) and np.array_equal(u(dc), hc) and np.array_equal(u(dp), hp), key
        R = _csr_host_transpose(hv, hc, hp, rows, cols)
        rv, rc, rp, rcols = d["R" + key[1]]
        assert rcols == rows
        assert np.array_equal(u(rv), R.val.view(np.uint64)), key
        assert np.array_equal(u(rc), R.col.astype(np.uint64)) and np.array_equal(u(rp), R.rowptr.astype(np.uint64)), key
