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
    return dict(k=k, nv=nv, np=np_, K=K, V=Vt, keep=(kk, kv, nodes, keep_ws), vmask=vmask,
                prim=dev(torch, S["prim"]), dual=dev(torch, S["dual"]), vol=vol, ws=ws, cnt=keep_ws[1])


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


def _fused(L, h, M, r, x, y, alpha, wd, wa, passes, xu, omega, gate):
    k = L["k"]
    yp = y.data_ptr() if y is not None else None
    assert k.feat_tuple_apply_post_f64(h, C.byref(M), r.data_ptr(), x.data_ptr(), yp, alpha,
                                       L["vmask"].data_ptr(), wd.data_ptr(), wa.data_ptr(), L["vol"], passes, 0,
                                       xu.data_ptr() if xu is not None else None, omega,
                                       gate.data_ptr() if gate is not None else None, C.byref(L["ws"])) == 0


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
                                     x.data_ptr(), 0.7, None, C.byref(L["ws"]))
    assert rc != 0
    rc = k.feat_tuple_apply_post_f64(handle.h, C.byref(L["V"]), r.data_ptr(), x.data_ptr(), None, 1.0, *common,
                                     None, 0.7, None, None)
    assert rc != 0
