"""The fused Stokes smoothing step, feat_tuple_apply_post_f64: one launch of the tuple apply,
the MeanFilter passes on its pressure part and the gated smoother update.  Bar: bit-identical
to feat_tuple_apply_f64, feat_mean_filter_passes_f64 and feat_axpy_gated_f64 called in a row,
on the saddle-point matrix and on the AmaVanka matrix of levels 3, 4 and 6 (grids of a few to
several hundred workgroups, so the last-arriving workgroup that runs the filter varies), with
the gate open and closed, omega below eps, zero to two passes, and the counter left at zero
for the next call."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import Csr, csr_transpose
from tests.test_gpu_lafem import dev, host

pytestmark = pytest.mark.gpu


class Tuple(C.Structure):
    _fields_ = [("nv", C.c_uint64), ("np", C.c_uint64), ("val", C.c_void_p * 4), ("col", C.c_void_p * 4),
                ("ptr", C.c_void_p * 4)]


def _tuple(torch, nv, np_, blocks):
    """blocks: four (val, col, rowptr) host triples or None; returns (keep-alive, Tuple)."""
    keep, t = [], Tuple(nv, np_)
    for k, b in enumerate(blocks):
        if b is None:
            continue
        v, c, p = (dev(torch, a.astype(np.float64)) if i == 0 else dev(torch, a.astype(np.int64))
                   for i, a in enumerate(b))
        keep += [v, c, p]
        t.val[k], t.col[k], t.ptr[k] = v.data_ptr(), c.data_ptr(), p.data_ptr()
    return keep, t


@pytest.fixture(scope="module", params=[3, 4, 6])
def level(request, feat, torch_gpu, handle):
    torch = torch_gpu
    lvl = request.param
    k, _ = feat.libs()
    S = feat.stokes_problem(lvl)
    nv, np_ = S["nv"], S["np"]
    stk = feat.Stokes(lvl, lvl - 1)
    V = stk.vanka(0)
    stk.close()
    vb = [(b["val"][:int(b["rowptr"][-1]) * s], b["col"][:int(b["rowptr"][-1])], b["rowptr"])
          for b, s in zip(V["blocks"], (4, 2, 2, 1))]
    kk, K = _tuple(torch, nv, np_, [S["A"], S["B"], S["D"], None])
    kv, Vt = _tuple(torch, nv, np_, vb)
    nodes = dev(torch, S["filt"].astype(np.int64))
    vmask = torch.zeros((nv + 31) // 32, dtype=torch.int32, device="cuda")
    feat.check(k.feat_unit_filter_mask_u64(handle.h, vmask.data_ptr(), nv, nodes.data_ptr(), nodes.numel()),
               "mask")
    vol = 0.0
    for a, b in zip(S["prim"], S["dual"]):
        vol = vol + a * b
    keep_ws, ws = feat.reduce_ws(torch, "cuda")
    # transfers to level lvl - 1 and that level's filter
    Sc = feat.stokes_problem(lvl - 1)
    nvc, npc = Sc["nv"], Sc["np"]
    pv, pp = S["Pv"], S["Pp"]
    Pv = Csr(nv, pv[3], pv[0], pv[1], pv[2])
    Pp = Csr(np_, pp[3], pp[0], pp[1], pp[2])
    tr = {}
    for name, A in (("pv", Pv), ("pp", Pp), ("rv", csr_transpose(Pv)), ("rp", csr_transpose(Pp))):
        tens = (dev(torch, A.val.astype(np.float64)), dev(torch, A.col.astype(np.int64)),
                dev(torch, A.rowptr.astype(np.int64)))
        tr[name] = (A.rows, A.cols, A.val.size, tens)
    cnodes = dev(torch, Sc["filt"].astype(np.int64))
    cmask = torch.zeros((nvc + 31) // 32, dtype=torch.int32, device="cuda")
    feat.check(k.feat_unit_filter_mask_u64(handle.h, cmask.data_ptr(), nvc, cnodes.data_ptr(), cnodes.numel()),
               "mask")
    cvol = 0.0
    for a, b in zip(Sc["prim"], Sc["dual"]):
        cvol = cvol + a * b
    return dict(k=k, nv=nv, np=np_, K=K, V=Vt, keep=(kk, kv, nodes, keep_ws, cnodes), vmask=vmask, nodes=nodes,
                prim=dev(torch, S["prim"]), dual=dev(torch, S["dual"]), vol=vol, ws=ws, cnt=keep_ws[1],
                nvc=nvc, npc=npc, tr=tr, cmask=cmask, cnodes=cnodes, cprim=dev(torch, Sc["prim"]),
                cdual=dev(torch, Sc["dual"]), cvol=cvol)


def _separate(L, h, M, r, x, y, alpha, wd, wa, passes, xu, omega, gate):
    k = L["k"]
    n = 2 * L["nv"] + L["np"]
    yp = y.data_ptr() if y is not None else None
    assert k.feat_tuple_apply_f64(h, C.byref(M), r.data_ptr(), x.data_ptr(), yp, alpha, L["vmask"].data_ptr()) == 0
    if passes:
        assert k.feat_mean_filter_passes_f64(h, r.data_ptr() + 16 * L["nv"], wd.data_ptr(), wa.data_ptr(), L["vol"],
                                             L["np"], passes, 0) == 0
    if xu is not None:
        assert k.feat_axpy_gated_f64(h, xu.data_ptr(), omega, r.data_ptr(), n,
                                     gate.data_ptr() if gate is not None else None) == 0


def _fused(L, h, M, r, x, y, alpha, wd, wa, passes, xu, omega, gate, x_zero=0):
    k = L["k"]
    yp = y.data_ptr() if y is not None else None
    assert k.feat_tuple_apply_post_f64(h, C.byref(M), r.data_ptr(), x.data_ptr(), yp, alpha,
                                       L["vmask"].data_ptr(), wd.data_ptr(), wa.data_ptr(), L["vol"], passes, 0,
                                       xu.data_ptr() if xu is not None else None, omega,
                                       gate.data_ptr() if gate is not None else None, x_zero, C.byref(L["ws"])) == 0


@pytest.mark.parametrize("case", ["vanka_update", "vanka_gate_closed", "vanka_tiny_omega", "vanka_no_update",
                                  "defect", "apply_no_filter"])
def test_fused_step_bitwise(level, feat, torch_gpu, handle, case):
    torch = torch_gpu
    L = level
    n = 2 * L["nv"] + L["np"]
    rng = np.random.default_rng(sum(case.encode()) + n)
    x = dev(torch, rng.standard_normal(n))
    y = dev(torch, rng.standard_normal(n))
    x_up = rng.standard_normal(n)
    gate_open = dev(torch, np.array([0.3]))
    gate_shut = dev(torch, np.array([1e-40]))
    cfg = dict(vanka_update=(L["V"], None, 1.0, L["dual"], L["prim"], 2, True, 0.7, gate_open),
               vanka_gate_closed=(L["V"], None, 1.0, L["dual"], L["prim"], 2, True, 0.7, gate_shut),
               vanka_tiny_omega=(L["V"], None, 1.0, L["dual"], L["prim"], 2, True, 1e-17, gate_open),
               vanka_no_update=(L["V"], None, 1.0, L["dual"], L["prim"], 2, False, 0.0, None),
               defect=(L["K"], y, -1.0, L["prim"], L["dual"], 1, False, 0.0, None),
               apply_no_filter=(L["K"], y, 0.7, L["prim"], L["dual"], 0, True, -0.3, None))[case]
    M, yy, alpha, wd, wa, passes, with_upd, omega, gate = cfg
    outs = []
    for fn in (_separate, _fused):
        r = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
        xu = dev(torch, x_up) if with_upd else None
        for _ in range(2):   # twice in a row: the counter must be back at zero
            fn(L, handle.h, M, r, x, yy, alpha, wd, wa, passes, xu, omega, gate)
        torch.cuda.synchronize()
        outs.append((host(r), host(xu) if xu is not None else None))
    (rs, xs), (rf, xf) = outs
    assert np.array_equal(rs, rf), np.max(np.abs(rs - rf))
    if with_upd:
        assert np.array_equal(xs, xf)
        if case in ("vanka_gate_closed", "vanka_tiny_omega"):
            assert np.array_equal(xf, x_up)   # the update is skipped
        else:
            assert not np.array_equal(xf, x_up)
    assert int(L["cnt"][0].item()) == 0


def test_fused_step_rejects_aliases(level, feat, torch_gpu, handle):
    torch = torch_gpu
    L = level
    n = 2 * L["nv"] + L["np"]
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    r = torch.zeros(n, dtype=torch.float64, device="cuda")
    k = L["k"]
    common = (L["vmask"].data_ptr(), L["dual"].data_ptr(), L["prim"].data_ptr(), L["vol"], 2, 0)
    rc = k.feat_tuple_apply_post_f64(handle.h, C.byref(L["V"]), r.data_ptr(), x.data_ptr(), None, 1.0, *common,
                                     x.data_ptr(), 0.7, None, 0, C.byref(L["ws"]))
    assert rc != 0
    rc = k.feat_tuple_apply_post_f64(handle.h, C.byref(L["V"]), r.data_ptr(), x.data_ptr(), None, 1.0, *common,
                                     None, 0.7, None, 0, None)
    assert rc != 0


@pytest.mark.parametrize("case", ["open", "gate_closed", "tiny_omega"])
def test_fused_step_zero_start(level, feat, torch_gpu, handle, case):
    """x_upd_zero: the first smoother update writes omega c (or zeros where it is skipped) into
    an unread x_upd -- bit-identical to formatting x_upd and updating it."""
    torch = torch_gpu
    L = level
    n = 2 * L["nv"] + L["np"]
    rng = np.random.default_rng(11 + n)
    d = dev(torch, rng.standard_normal(n))
    gate = dev(torch, np.array([1e-40 if case == "gate_closed" else 0.4]))
    omega = 1e-17 if case == "tiny_omega" else 0.7
    outs = []
    for fused in (False, True):
        r = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
        xu = torch.full((n,), np.nan if fused else 0.0, dtype=torch.float64, device="cuda")
        if fused:
            _fused(L, handle.h, L["V"], r, d, None, 1.0, L["dual"], L["prim"], 2, xu, omega, gate, x_zero=1)
        else:
            _separate(L, handle.h, L["V"], r, d, None, 1.0, L["dual"], L["prim"], 2, xu, omega, gate)
        torch.cuda.synchronize()
        outs.append((host(r), host(xu)))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    if case != "open":
        assert not np.any(outs[1][1])


def _transfer_separate(L, h, Vn, Pn, x, nvx, r, nvr, mask_nodes, wd, wa, vol, npr, xu):
    k = L["k"]
    rows_v, cols_v, nnz_v, (vv, vc, vp) = L["tr"][Vn]
    rows_p, cols_p, nnz_p, (pv, pc, pp) = L["tr"][Pn]
    assert k.feat_csrsb2_apply_f64_u64(h, r.data_ptr(), x.data_ptr(), vv.data_ptr(), vc.data_ptr(), vp.data_ptr(),
                                       rows_v) == 0
    rp = r.data_ptr() + 16 * nvr
    assert k.feat_csr_apply_f64_u64(h, rp, 1.0, x.data_ptr() + 16 * nvx, 0.0, rp, pv.data_ptr(), pc.data_ptr(),
                                    pp.data_ptr(), rows_p, cols_p, nnz_p, 0) == 0
    assert k.feat_unit_filter_blocked_def_f64_u64(h, 2, r.data_ptr(), None, mask_nodes.data_ptr(),
                                                  mask_nodes.numel(), 0) == 0
    assert k.feat_mean_filter_passes_f64(h, rp, wd.data_ptr(), wa.data_ptr(), vol, npr, 1, 0) == 0
    if xu is not None:
        assert k.feat_axpy_f64(h, xu.data_ptr(), 1.0, r.data_ptr(), xu.data_ptr(), 2 * nvr + npr) == 0


class CsrDesc(C.Structure):
    _fields_ = [("rows", C.c_uint64), ("cols", C.c_uint64), ("nnz", C.c_uint64), ("val", C.c_void_p),
                ("col", C.c_void_p), ("rowptr", C.c_void_p)]


def _transfer_fused(L, h, Vn, Pn, x, nvx, r, vmask, wd, wa, vol, xu):
    k = L["k"]
    d = []
    for name in (Vn, Pn):
        rows, cols, nnz, (v, c, p) = L["tr"][name]
        d.append(CsrDesc(rows, cols, nnz, v.data_ptr(), c.data_ptr(), p.data_ptr()))
    assert k.feat_stokes_transfer_post_f64(h, C.byref(d[0]), C.byref(d[1]), x.data_ptr(), nvx, r.data_ptr(),
                                           vmask.data_ptr(), wd.data_ptr(), wa.data_ptr(), vol, 1,
                                           xu.data_ptr() if xu is not None else None, C.byref(L["ws"])) == 0


@pytest.mark.parametrize("kind", ["rest", "prol"])
def test_fused_transfer_bitwise(level, feat, torch_gpu, handle, kind):
    """feat_stokes_transfer_post_f64 against the separate calls: restriction + filter_def of the
    coarse level, prolongation + filter_cor of the fine level + sol += cor; twice in a row (the
    counter returns to zero)."""
    torch = torch_gpu
    L = level
    nf = 2 * L["nv"] + L["np"]
    nc = 2 * L["nvc"] + L["npc"]
    rng = np.random.default_rng(5 + nf + (kind == "prol"))
    if kind == "rest":
        x = dev(torch, rng.standard_normal(nf))
        args = ("rv", "rp", x, L["nv"], nc, L["nvc"], L["cnodes"], L["cmask"], L["cprim"], L["cdual"], L["cvol"],
                L["npc"], None)
    else:
        x = dev(torch, rng.standard_normal(nc))
        args = ("pv", "pp", x, L["nvc"], nf, L["nv"], L["nodes"], L["vmask"], L["dual"], L["prim"], L["vol"],
                L["np"], rng.standard_normal(nf))
    Vn, Pn, x, nvx, nr, nvr, nodes, vmask, wd, wa, vol, npr, xu0 = args
    outs = []
    for fused in (False, True):
        r = torch.full((nr,), np.nan, dtype=torch.float64, device="cuda")
        xu = dev(torch, xu0) if xu0 is not None else None
        for _ in range(2):
            if xu is not None:
                xu.copy_(dev(torch, xu0))
            if fused:
                _transfer_fused(L, handle.h, Vn, Pn, x, nvx, r, vmask, wd, wa, vol, xu)
            else:
                _transfer_separate(L, handle.h, Vn, Pn, x, nvx, r, nvr, nodes, wd, wa, vol, npr, xu)
        torch.cuda.synchronize()
        outs.append((host(r), host(xu) if xu is not None else None))
    assert np.array_equal(outs[0][0], outs[1][0]), np.max(np.abs(outs[0][0] - outs[1][0]))
    if xu0 is not None:
        assert np.array_equal(outs[0][1], outs[1][1])
    assert int(L["cnt"][0].item()) == 0
