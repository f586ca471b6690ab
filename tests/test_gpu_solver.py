"""GPU parity of the fused solver kernels and of PCG(MG-V(Richardson(Jacobi)))
against the oracle.

Bars: fused kernels bit-exact except their dot/norm outputs (1e-13 relative);
solver iteration counts equal to the oracle's within +-1; solution error
|x - x_oracle| / |x_oracle| <= 1e-11; final relative defect <= tol.
"""
import ctypes

import numpy as np
import pytest

from oracle.pyoracle import Csr, COARSE_JACOBI, COARSE_DIRECT, COARSE_SMOOTHER
from tests.helpers import q1_levels
from tests.test_gpu_lafem import dev, host, upload_csr

pytestmark = pytest.mark.gpu

SOL_TOL = 1e-11
C_byref = ctypes.byref


def mask_of(torch, feat, handle, n, idx):
    k, _ = feat.libs()
    m = torch.zeros((n + 31) // 32, dtype=torch.int32, device="cuda")
    idd = dev(torch, np.asarray(idx, np.int64))
    feat.check(k.feat_unit_filter_mask_u64(handle.h, feat.ptr(m), n, feat.ptr(idd), len(idx)), "mask")
    return m


def q1_level(feat, dim, lvl):
    q = feat.q1_problem(dim, lvl)
    return q, Csr(q["n"], q["n"], q["val"], q["col"], q["rowptr"])


def exec_desc(feat, handle, torch, A, fmt):
    """(keep-alive, descriptor) of A on the device: its CSR arrays or a SELL-64 plan."""
    v, c, r = upload_csr(torch, A)
    if fmt == "csr":
        return (v, c, r), feat.csr_matrix(v, c, r, A.rows, A.cols)
    plan = feat.SellPlan(handle, v, c, r, A.rows, A.cols)
    return (v, c, r, plan), plan.m


@pytest.mark.parametrize("fmt", ["csr", "sell"])
def test_fused_defect_masked_prol(feat, handle, torch_gpu, oracle, fmt):
    torch = torch_gpu
    k, _ = feat.libs()
    P = feat.ptr
    keep, ws = feat.reduce_ws(torch, "cuda")
    q, A = q1_level(feat, 3, 4)
    n = q["n"]
    rng = np.random.default_rng(21)
    x, rhs, w = rng.standard_normal((3, n))
    keepA, desc = exec_desc(feat, handle, torch, A, fmt)
    m = mask_of(torch, feat, handle, n, q["filt"])
    scal = torch.zeros(4, dtype=torch.float64, device="cuda")
    d = torch.empty(n, dtype=torch.float64, device="cuda")
    xd, rd, wd = dev(torch, x), dev(torch, rhs), dev(torch, w)
    feat.check(k.feat_defect_f64(handle.h, desc, P(d), P(rd), P(xd), P(m), P(scal), ws), "defect")
    handle.sync()
    ref = oracle.unit_filter_def(oracle.csr_apply(A, x, rhs, -1.0, 1.0), q["filt"])
    assert np.array_equal(host(d), ref)
    assert abs(host(scal)[0] - ref @ ref) <= 1e-13 * (ref @ ref)
    # masked apply with norm and dot
    feat.check(k.feat_apply_masked_f64(handle.h, desc, P(d), P(xd), P(m), P(scal[1:]), P(wd), P(scal[2:]), ws), "masked")
    handle.sync()
    ref = oracle.unit_filter_def(oracle.csr_apply(A, x), q["filt"])
    assert np.array_equal(host(d), ref)
    s = host(scal)
    assert abs(s[1] - ref @ ref) <= 1e-13 * (ref @ ref)
    assert abs(s[2] - ref @ w) <= 1e-13 * np.abs(ref * w).sum()
    # prolongation + correction
    Pm = Csr(n, q["pcols"], q["pval"], q["pcol"], q["prowptr"])
    keepP, pdesc = exec_desc(feat, handle, torch, Pm, fmt)
    xc = rng.standard_normal(q["pcols"])
    sol = dev(torch, x)
    xcd = dev(torch, xc)
    feat.check(k.feat_prol_add_f64(handle.h, pdesc, P(sol), P(xcd), P(m)), "prol")
    handle.sync()
    cor = oracle.unit_filter_def(oracle.csr_apply(Pm, xc), q["filt"])
    assert np.array_equal(host(sol), oracle.axpy(1.0, cor, x))


@pytest.mark.parametrize("fmt", ["csr", "sell"])
@pytest.mark.parametrize("steps", [1, 2, 4])
def test_fused_richardson_jacobi(feat, handle, torch_gpu, oracle, steps, fmt):
    torch = torch_gpu
    k, _ = feat.libs()
    P = feat.ptr
    keep, ws = feat.reduce_ws(torch, "cuda")
    q, A = q1_level(feat, 2, 6)
    n = q["n"]
    rng = np.random.default_rng(22)
    rhs = oracle.unit_filter_def(rng.standard_normal(n), q["filt"])
    keepA, desc = exec_desc(feat, handle, torch, A, fmt)
    v, c, r = keepA[:3]
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


@pytest.mark.parametrize("seed,rows,cols,maxlen,band", [
    (31, 5000, 5000, 30, 100),       # narrow band -> i16 column deltas
    (32, 3000, 70000, 25, None),     # wide columns -> u32 absolute indices
    (33, 700, 700, 0, None),         # empty rows only
    (34, 129, 129, 40, 20),          # last slice holds one row
])
def test_sell_plan_apply_bitwise(feat, handle, torch_gpu, oracle, seed, rows, cols, maxlen, band):
    torch = torch_gpu
    k, _ = feat.libs()
    P = feat.ptr
    rng = np.random.default_rng(seed)
    from tests.helpers import random_csr
    A = random_csr(rng, rows, cols, maxlen // 2, maxlen)
    if band is not None:
        # keep columns within +-band of the row (clipped), still sorted per row
        rr = np.repeat(np.arange(rows), np.diff(A.rowptr.astype(np.int64)))
        A.col = np.clip(rr + (A.col.astype(np.int64) % (2 * band + 1)) - band, 0, cols - 1).astype(np.uint64)
        for r in range(rows):
            a, b = int(A.rowptr[r]), int(A.rowptr[r + 1])
            A.col[a:b] = np.sort(A.col[a:b])
    x = rng.standard_normal(cols)
    y = rng.standard_normal(rows)
    keep, desc = exec_desc(feat, handle, torch, A, "sell")
    if rows and A.nnz:
        assert desc.format == feat.FMT_SELL
        assert desc.index == (feat.IDX_D16 if band is not None else feat.IDX_U32)
    xd, yd = dev(torch, x), dev(torch, y)
    rd = torch.empty(rows, dtype=torch.float64, device="cuda")
    for a, b in [(1.0, 0.0), (-1.0, 1.0), (2.5, -0.7)]:
        feat.check(k.feat_matrix_apply_f64(handle.h, desc, P(rd), a, P(xd), b, P(yd)), "sell apply")
        handle.sync()
        assert np.array_equal(host(rd), oracle.csr_apply(A, x, y, a, b)), (a, b)
    # FP32 plan: values rounded once, fmaf in CSR order (the oracle's f32 apply)
    v, c, r = keep[:3]
    p32 = feat.SellPlan(handle, v, c, r, rows, cols, fp="f32")
    x32 = dev(torch, x.astype(np.float32))
    r32 = torch.empty(rows, dtype=torch.float32, device="cuda")
    keep_ws, ws = feat.reduce_ws(torch, "cuda")
    feat.check(k.feat_apply_masked_f32(handle.h, p32.m, P(r32), P(x32), None, None, ws), "sell f32")
    handle.sync()
    ref = oracle.csr_apply_f32(A, x.astype(np.float32))
    assert np.array_equal(host(r32), ref)
    p32.close()


def stencil_csr(rng, rows, cols, offsets, values, p_keep):
    """Rows holding a random subset of `offsets` (col - row, clipped to the matrix),
    values drawn from `values`: at most len(offsets) * len(values) distinct pairs."""
    from oracle.pyoracle import Csr
    rp, col = [0], []
    for r in range(rows):
        c = sorted({r + o for o in offsets if 0 <= r + o < cols and rng.random() < p_keep})
        col += c
        rp.append(len(col))
    val = rng.choice(np.asarray(values, np.float64), size=len(col))
    return Csr(rows, cols, val, np.asarray(col, np.uint64), np.asarray(rp, np.uint64))


@pytest.mark.parametrize("dim,level", [(2, 1), (2, 5), (3, 1), (3, 4)])
def test_dlinear_transfer_descriptor_bitwise(feat, handle, torch_gpu, oracle, dim, level):
    """The matrix-free d-linear transfer descriptors (FEAT_FMT_DLINEAR) equal the assembled
    Q1 prolongation from level-1 and its transpose row for row, bit for bit, in FP64 and
    FP32, through the plain apply and the masked apply with its norm."""
    print(f"dlinear transfer dim={dim} level={level}", flush=True)
    torch = torch_gpu
    k, _ = feat.libs()
    P_ = feat.ptr
    q = feat.q1_problem(dim, level)
    n, nc = q["n"], q["pcols"]
    Pm = Csr(n, nc, q["pval"], q["pcol"], q["prowptr"])
    from oracle.pyoracle import csr_transpose
    Rm = csr_transpose(Pm)
    n1 = 2 ** level + 1
    rng = np.random.default_rng(5 + level)
    keep_ws, ws = feat.reduce_ws(torch, "cuda")
    for transpose, M in ((0, Pm), (1, Rm)):
        d = feat.Matrix()
        feat.check(k.feat_matrix_dlinear_f64(dim, n1, transpose, C_byref(d)), "dlinear f64")
        assert d.format == feat.FMT_DLINEAR and d.rows == M.rows and d.cols == M.cols
        x = rng.standard_normal(M.cols)
        xd = dev(torch, x)
        rd = torch.empty(M.rows, dtype=torch.float64, device="cuda")
        feat.check(k.feat_matrix_apply_f64(handle.h, C_byref(d), P_(rd), 1.0, P_(xd), 0.0, None), "apply")
        handle.sync()
        assert np.array_equal(host(rd), oracle.csr_apply(M, x)), transpose
        nrm = torch.zeros(1, dtype=torch.float64, device="cuda")
        feat.check(k.feat_apply_masked_f64(handle.h, C_byref(d), P_(rd), P_(xd), None, P_(nrm), None, None, ws), "masked")
        handle.sync()
        want = oracle.csr_apply(M, x)
        assert np.array_equal(host(rd), want)
        assert abs(host(nrm)[0] - float(want @ want)) <= 1e-13 * float(want @ want)
        d32 = feat.Matrix()
        feat.check(k.feat_matrix_dlinear_f32(dim, n1, transpose, C_byref(d32)), "dlinear f32")
        x32 = dev(torch, x.astype(np.float32))
        r32 = torch.empty(M.rows, dtype=torch.float32, device="cuda")
        feat.check(k.feat_apply_masked_f32(handle.h, C_byref(d32), P_(r32), P_(x32), None, None, ws), "masked f32")
        handle.sync()
        assert np.array_equal(host(r32), oracle.csr_apply_f32(M, x.astype(np.float32))), transpose
    # sizes that are no such grid pair are refused
    d = feat.Matrix()
    assert k.feat_matrix_dlinear_f64(dim, n1 + 1, 0, C_byref(d)) != 0
    assert k.feat_matrix_dlinear_f64(4, n1, 0, C_byref(d)) != 0


@pytest.mark.parametrize("diag", [True, False])
@pytest.mark.parametrize("case", ["q1", "stencil", "partial", "wide", "skewed", "many_pairs", "many_offsets",
                                  "no_dict", "q1_2d_odd", "singles", "wide_rect", "tall_rect"])
def test_sell_dict_plan_bitwise(feat, handle, torch_gpu, oracle, case, diag):
    """Pair-dictionary SELL plans (one byte per entry, in entry order or by diagonal)
    reproduce the CSR loop bit for bit in FP64 and FP32; matrices with more than 256
    distinct (col - row, value) pairs, or dict=False, get the plain plan, with 1-byte
    value codes where at most 256 distinct values occur (not with dict=False); the
    diagonal layout applies to at most 32 distinct offsets with little slot padding."""
    print(f"dict plan {case} diag={diag}", flush=True)
    torch = torch_gpu
    k, _ = feat.libs()
    P = feat.ptr
    rng = np.random.default_rng(77)
    offs = [-130, -129, -128, -2, -1, 0, 1, 2, 128, 129, 130, 4000]
    vals = [0.25, -1.0, 3.5, -0.125, 2.0 / 3.0, 1e-300, -7.0]
    want = feat.IDX_DICT
    if case == "q1":
        A = q1_level(feat, 3, 4)[1]
    elif case == "stencil":
        A = stencil_csr(rng, 5000, 5000, offs, vals, 0.8)
    elif case == "partial":
        A = stencil_csr(rng, 129, 140, offs[:9], vals, 0.5)      # last slice holds one row, empty rows
    elif case == "wide":
        # offsets spread over +-1000
        A = stencil_csr(rng, 3000, 3000, [-1000 + 100 * i for i in range(21)], vals[:3], 0.7)
    elif case == "skewed":
        # one slice of 40-entry rows, the rest 4: slices keep their own offsets
        from oracle.pyoracle import Csr
        n = 2000
        rp, col = [0], []
        for r in range(n):
            offs_r = range(-20, 20) if r < 64 else (-1, 0, 1, 2)
            col += [r + o for o in offs_r if 0 <= r + o < n]
            rp.append(len(col))
        A = Csr(n, n, rng.choice(np.asarray(vals[:2]), size=len(col)), np.asarray(col, np.uint64),
                np.asarray(rp, np.uint64))
    elif case == "many_pairs":
        A = stencil_csr(rng, 4000, 4000, list(range(-20, 20)), [float(v) for v in range(1, 9)], 1.0)  # 320 pairs
        want = feat.IDX_D16
    elif case == "many_offsets":
        from tests.helpers import random_csr
        A = random_csr(rng, 3000, 70000, 10, 20)
        A.val = rng.choice(np.asarray(vals), size=A.nnz)           # few values, 1000s of offsets
        want = feat.IDX_U32
    elif case == "q1_2d_odd":
        A = q1_level(feat, 2, 5)[1]                                # 1089 rows: a partial last chunk
    elif case == "wide_rect":
        # a 9-point pattern whose columns run beyond the rows
        A = stencil_csr(rng, 700, 3000, [-1, 0, 1, 899, 900, 901, 2199, 2200, 2201], vals, 0.95)
    elif case == "tall_rect":
        A = stencil_csr(rng, 3000, 700, [-2300, -1, 0, 1], vals, 0.9)
    elif case == "singles":
        A = stencil_csr(rng, 3000, 3000, [-300, -7, 0, 5, 300], vals, 0.9)   # no consecutive offsets
    else:
        A = q1_level(feat, 3, 4)[1]
        want = feat.IDX_D16
    # the diagonal layout: offsets covered greedily by at most 10 triples (c-1, c, c+1)
    offs_all = sorted({int(c) - r for r in range(A.rows) for c in A.col[int(A.rowptr[r]):int(A.rowptr[r + 1])]})
    ntrip, i = 0, 0
    while i < len(offs_all):
        centre = offs_all[i] + 1
        ntrip += 1
        while i < len(offs_all) and offs_all[i] <= centre + 1:
            i += 1
    nslot = 3 * ntrip
    if want == feat.IDX_DICT and diag and ntrip <= 10 and 64 * ((A.rows + 63) // 64) * nslot <= A.nnz + A.nnz // 4 + 64 * nslot:
        want = feat.IDX_DIAG
    if case in ("q1", "q1_2d_odd", "wide_rect") and diag:
        assert want == feat.IDX_DIAG
    v, c, r = upload_csr(torch, A)
    plan = feat.SellPlan(handle, v, c, r, A.rows, A.cols, dict=(case != "no_dict"), diag=diag)
    assert plan.m.format == feat.FMT_SELL and plan.m.index == want, (plan.m.index, want)
    want_vk = feat.VAL_CODE8 if case in ("many_pairs", "many_offsets") else feat.VAL_NATIVE
    assert plan.m.value_kind == want_vk, (plan.m.value_kind, want_vk)
    if want == feat.IDX_DICT:
        # one code-word stride for all slices unless that would pad by more than a quarter
        assert (plan.m.slice_words == 0) == (case in ("skewed", "tall_rect")), plan.m.slice_words
    x = rng.standard_normal(A.cols)
    y = rng.standard_normal(A.rows)
    xd, yd = dev(torch, x), dev(torch, y)
    rd = torch.empty(A.rows, dtype=torch.float64, device="cuda")
    for tune in (0, 1, 3):
        plan.m.tune = tune
        for a, b in [(1.0, 0.0), (-1.0, 1.0)]:
            feat.check(k.feat_matrix_apply_f64(handle.h, plan.m, P(rd), a, P(xd), b, P(yd)), "dict apply")
            handle.sync()
            assert np.array_equal(host(rd), oracle.csr_apply(A, x, y, a, b)), (tune, a, b)
    plan.close()
    p32 = feat.SellPlan(handle, v, c, r, A.rows, A.cols, fp="f32", dict=(case != "no_dict"), diag=diag)
    assert p32.m.index == want and p32.m.value_kind == want_vk
    x32 = dev(torch, x.astype(np.float32))
    r32 = torch.empty(A.rows, dtype=torch.float32, device="cuda")
    keep_ws, ws = feat.reduce_ws(torch, "cuda")
    feat.check(k.feat_apply_masked_f32(handle.h, p32.m, P(r32), P(x32), None, None, ws), "dict f32")
    handle.sync()
    assert np.array_equal(host(r32), oracle.csr_apply_f32(A, x.astype(np.float32)))
    p32.close()


@pytest.mark.parametrize("with_nan", [False, True])
def test_sell_value_codes(feat, handle, torch_gpu, oracle, with_nan):
    """Plain SELL-64 plans of a matrix with few distinct values (a Q1 prolongation: 1, 1/2,
    1/4, 1/8) store 1-byte value codes and reproduce the CSR loop bit for bit in FP64 and
    FP32; a NaN entry turns the codes off (its payload may convert differently on host and
    device) and the plan keeps full values."""
    print(f"value codes nan={with_nan}", flush=True)
    torch = torch_gpu
    k, _ = feat.libs()
    P = feat.ptr
    q = feat.q1_problem(3, 4)
    A = Csr(q["n"], q["pcols"], q["pval"].copy(), q["pcol"], q["prowptr"])
    if with_nan:
        A.val[A.nnz // 2] = np.nan
    v, c, r = upload_csr(torch, A)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(A.cols)
    xd = dev(torch, x)
    rd = torch.empty(A.rows, dtype=torch.float64, device="cuda")
    plan = feat.SellPlan(handle, v, c, r, A.rows, A.cols)
    assert plan.m.index in (feat.IDX_U32, feat.IDX_D16)
    assert plan.m.value_kind == (feat.VAL_NATIVE if with_nan else feat.VAL_CODE8)
    feat.check(k.feat_matrix_apply_f64(handle.h, plan.m, P(rd), 1.0, P(xd), 0.0, None), "apply")
    handle.sync()
    assert np.array_equal(host(rd), oracle.csr_apply(A, x), equal_nan=True)
    plan.close()
    p32 = feat.SellPlan(handle, v, c, r, A.rows, A.cols, fp="f32")
    assert p32.m.value_kind == (feat.VAL_NATIVE if with_nan else feat.VAL_CODE8)
    x32 = dev(torch, x.astype(np.float32))
    r32 = torch.empty(A.rows, dtype=torch.float32, device="cuda")
    keep_ws, ws = feat.reduce_ws(torch, "cuda")
    feat.check(k.feat_apply_masked_f32(handle.h, p32.m, P(r32), P(x32), None, None, ws), "apply f32")
    handle.sync()
    assert np.array_equal(host(r32), oracle.csr_apply_f32(A, x.astype(np.float32)), equal_nan=True)
    p32.close()


@pytest.mark.parametrize("nvals,specials", [(256, False), (257, False), (256, True), (257, True)])
def test_sell_value_codes_boundary(feat, handle, torch_gpu, oracle, nvals, specials):
    """Exactly 256 distinct values get codes 0..255, a 257th keeps full values; +0.0 and
    -0.0 are distinct keys, and +-Inf are values like any other.  Both forms reproduce
    the CSR loop bit for bit in FP64 and FP32 (values exact in FP32, so both precisions
    see the same distinct count).  Random columns keep the pair dictionary out."""
    print(f"value-code boundary nvals={nvals} specials={specials}", flush=True)
    torch = torch_gpu
    k, _ = feat.libs()
    P = feat.ptr
    rng = np.random.default_rng(nvals + 7 * specials)
    table = (np.arange(nvals) - 1000) / 8.0     # no zero among them
    if specials:
        table[:4] = [0.0, -0.0, np.inf, -np.inf]
    rows, cols, per = 700, 900, 5
    rp = (np.arange(rows + 1) * per).astype(np.uint64)
    col = np.concatenate([np.sort(rng.choice(cols, per, replace=False)) for _ in range(rows)]).astype(np.uint64)
    idx = np.concatenate([np.arange(nvals), rng.integers(0, nvals, rows * per - nvals)])
    val = table[rng.permutation(idx)]
    A = Csr(rows, cols, val, col, rp)
    v, c, r = upload_csr(torch, A)
    x = rng.standard_normal(cols).astype(np.float32).astype(np.float64)
    xd = dev(torch, x)
    rd = torch.empty(rows, dtype=torch.float64, device="cuda")
    want = feat.VAL_CODE8 if nvals <= 256 else feat.VAL_NATIVE
    plan = feat.SellPlan(handle, v, c, r, rows, cols)
    assert plan.m.index in (feat.IDX_U32, feat.IDX_D16)
    assert plan.m.value_kind == want
    feat.check(k.feat_matrix_apply_f64(handle.h, plan.m, P(rd), 1.0, P(xd), 0.0, None), "apply")
    handle.sync()
    _same_bits(host(rd), oracle.csr_apply(A, x))
    plan.close()
    p32 = feat.SellPlan(handle, v, c, r, rows, cols, fp="f32")
    assert p32.m.value_kind == want
    x32 = dev(torch, x.astype(np.float32))
    r32 = torch.empty(rows, dtype=torch.float32, device="cuda")
    keep_ws, ws = feat.reduce_ws(torch, "cuda")
    feat.check(k.feat_apply_masked_f32(handle.h, p32.m, P(r32), P(x32), None, None, ws), "apply f32")
    handle.sync()
    _same_bits(host(r32), oracle.csr_apply_f32(A, x.astype(np.float32)))
    p32.close()


def _same_bits(a, b):
    """Bit equality, with NaN only required in the same places (the payload and sign
    of a NaN made by Inf - Inf differ between the host CPU and the GPU)."""
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb)
    u = np.uint64 if a.dtype == np.float64 else np.uint32
    assert np.array_equal(a[~na].view(u), b[~nb].view(u))


def test_sell_plan_rejects_long_rows(feat, handle, torch_gpu):
    # a 70000-entry row exceeds the 16-bit row length; the plan refuses and the CSR stays usable
    torch = torch_gpu
    from oracle.pyoracle import Csr
    n = 70001
    lens = np.r_[70000, np.ones(n - 1, np.int64)]
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    col = np.concatenate([np.arange(70000), np.arange(1, n)]).astype(np.uint64)
    A = Csr(n, n, np.ones(col.size), col, rp)
    v, c, r = upload_csr(torch, A)
    with pytest.raises(feat.FeatError):
        feat.SellPlan(handle, v, c, r, n, n)


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


@pytest.mark.parametrize("dim,lf,lc,coarse,csr_only", [
    (2, 7, 1, "direct", False),
    (2, 5, 0, "smoother", False),
    (3, 5, 1, "direct", False),
    (3, 5, 1, "jacobi", True),
    (3, 6, 2, "direct", False),
    (3, 6, 1, "direct", True),
])
def test_mgpcg_vs_oracle(feat, torch_gpu, oracle, dim, lf, lc, coarse, csr_only):
    torch = torch_gpu
    levels = q1_levels(feat, dim, lf, lc)
    n = levels[0]["n"]
    b = feat.rhs_random(28054777172512, n)
    b[levels[0]["filt"].astype(np.int64)] = 0.0
    ocoarse = {"direct": COARSE_DIRECT, "jacobi": COARSE_JACOBI, "smoother": COARSE_SMOOTHER}[coarse]
    o = oracle.mg_pcg_solve(levels, np.zeros(n), b, coarse=ocoarse)
    for generic in (False, True):
        s = feat.MgPcg(feat.mgpcg_cfg(coarse=coarse, force_generic=generic, csr_only=csr_only), levels=levels)
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


@pytest.mark.parametrize("precision", ["f32", "f16"])
@pytest.mark.parametrize("dim,lf,lc,coarse", [(2, 7, 1, "direct"), (3, 5, 1, "direct"), (3, 6, 2, "jacobi"),
                                              (2, 6, 0, "smoother")])
def test_mixed_precision_vs_oracle(feat, torch_gpu, oracle, dim, lf, lc, coarse, precision):
    """FP32 V-cycle (FP32 or FP16 matrix values) inside FP64 PCG: iterations within
    +-1 of the oracle's same method, solution within 1e-10 (FP32) / 1e-9 (FP16) of the
    FP64 oracle solution, FP64-verified relative residual below the tolerance."""
    torch = torch_gpu
    levels = q1_levels(feat, dim, lf, lc)
    n = levels[0]["n"]
    b = feat.rhs_random(28054777172512, n)
    filt = levels[0]["filt"].astype(np.int64)
    b[filt] = 0.0
    oc = {"direct": COARSE_DIRECT, "jacobi": COARSE_JACOBI, "smoother": COARSE_SMOOTHER}[coarse]
    o64 = oracle.mg_pcg_solve(levels, np.zeros(n), b, coarse=oc)
    o32 = oracle.mg_pcg_solve(levels, np.zeros(n), b, coarse=oc, precision=precision)
    s = feat.MgPcg(feat.mgpcg_cfg(coarse=coarse, precision=precision), levels=levels)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    r = s.solve(x, dev(torch, b))
    s.close()
    xg = host(x)
    assert r["status"] == "success"
    assert abs(r["iters"] - o32["iters"]) <= 1
    tol = 1e-10 if precision == "f32" else 1e-9
    assert np.linalg.norm(xg - o32["x"]) <= tol * np.linalg.norm(o32["x"])
    assert np.linalg.norm(xg - o64["x"]) <= tol * np.linalg.norm(o64["x"])
    res = b - oracle.csr_apply(levels[0]["A"], xg)
    res[filt] = 0.0
    assert np.linalg.norm(res) <= 1.01e-8 * np.linalg.norm(b)


@pytest.mark.parametrize("precision,coarse", [("f64", "direct"), ("f64", "jacobi"), ("f64", "smoother"),
                                              ("f32", "direct"), ("f16", "jacobi")])
def test_mgpcg_graph_replay_matches_eager(feat, torch_gpu, precision, coarse):
    """The loop replayed as recorded HIP graphs gives bit-identical iterates, histories
    and iteration counts to eager launches, over repeated solves: new right-hand side in
    the same buffers, a non-zero initial guess, and a new solution buffer (re-recorded)."""
    torch = torch_gpu
    dim, lf, lc = (3, 5, 1) if coarse != "smoother" else (2, 6, 0)
    levels = q1_levels(feat, dim, lf, lc)
    n = levels[0]["n"]
    filt = levels[0]["filt"].astype(np.int64)
    solvers = [feat.MgPcg(feat.mgpcg_cfg(coarse=coarse, precision=precision, graphs=g), levels=levels)
               for g in (True, False)]
    b = torch.zeros(n, dtype=torch.float64, device="cuda")
    xs = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(2)]
    x_other = torch.zeros(n, dtype=torch.float64, device="cuda")
    try:
        for case in range(4):
            bh = feat.rhs_random(1000 + case, n)
            bh[filt] = 0.0
            b.copy_(torch.from_numpy(bh))
            x0 = np.zeros(n) if case != 2 else 0.1 * feat.rhs_random(77, n)
            out = []
            for s, x in zip(solvers, xs):
                xx = x_other if (case == 3 and s is solvers[0]) else x
                xx.copy_(torch.from_numpy(x0))
                r = s.solve(xx, b)
                out.append((r, host(xx)))
            (rg, xg), (re_, xe) = out
            assert rg["status"] == re_["status"] == "success", case
            assert rg["iters"] == re_["iters"], case
            assert np.array_equal(rg["hist"][:rg["iters"] + 1], re_["hist"][:re_["iters"] + 1]), case
            assert np.array_equal(xg, xe), case
    finally:
        for s in solvers:
            s.close()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_mgpcg_dictionary_vs_plain_plans(feat, torch_gpu, precision):
    """Entry-order pair-dictionary and plain SELL-64 plans (and whatever launch shapes the
    autotuner picks for each) give bit-identical iterates and iteration counts; the
    diagonal dictionary plan gives the same iteration count and iterates equal to 1e-13:
    its SpMV rows are bit-identical (test_sell_dict_plan_bitwise), but it sums dots
    and norms over chunks of 248 rows instead of 256, which rounds differently."""
    torch = torch_gpu
    levels = q1_levels(feat, 3, 5, 1)
    n = levels[0]["n"]
    filt = levels[0]["filt"].astype(np.int64)
    bh = feat.rhs_random(4242, n)
    bh[filt] = 0.0
    b = torch.from_numpy(bh).cuda()
    out = []
    for flags in (feat.SELL_NO_DIAG, feat.SELL_NO_DICT, 0):
        feat.runtime_sell_flags(flags)
        try:
            s = feat.MgPcg(feat.mgpcg_cfg(precision=precision), levels=levels)
        finally:
            feat.runtime_sell_flags(0)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        r = s.solve(x, b)
        s.close()
        assert r["status"] == "success"
        out.append((r["iters"], host(x)))
    assert out[0][0] == out[1][0] == out[2][0]
    assert np.array_equal(out[0][1], out[1][1])
    assert np.linalg.norm(out[2][1] - out[0][1]) <= 1e-13 * np.linalg.norm(out[0][1])


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("dim,lf,lc,first", [(3, 5, 1, 2), (2, 7, 1, 3), (3, 7, 1, 4)])
def test_coarse_fusion_bitwise(feat, torch_gpu, dim, lf, lc, first, precision):
    """The coarse levels run as one LDS-resident workgroup (feat_mg_tail_f64, or
    feat_mg_tail_f32 in the FP32 cycle of the mixed multigrid) give the solution and
    defect history of the separate launches bit for bit, in the grid form (the common
    stencil of the free rows and the d-linear transfers formed in the kernel) and from
    the plans and CSR transfers.  `first` is the first fused level: the finest level
    with at most 1,024 rows."""
    print(f"coarse fusion {precision} dim={dim} levels {lf}..{lc}", flush=True)
    torch = torch_gpu
    q = feat.q1_problem(dim, lf)
    b = feat.rhs_random(11, q["n"], -1.0, 1.0)
    b[q["filt"].astype(np.int64)] = 0.0
    bd = torch.from_numpy(b).cuda()
    out = []
    for fusion, grid in ((True, True), (True, False), (False, True)):
        cfg = feat.mgpcg_cfg(nu_pre=4, nu_post=4, omega_jac=0.5, omega_rich=1.0, coarse="direct",
                             tol_rel=1e-8, max_iter=100, graphs=False, coarse_fusion=fusion,
                             precision=precision)
        feat.runtime_grid_forms(grid)
        try:
            s = feat.MgPcg(cfg, q1=(dim, lf, lc))
            assert s.coarse_fusion_level() == (first if fusion else 0)
            # every smoothed fused level (solver indices first .. lf - lc - 1) has the grid form
            assert s.coarse_fusion_grid() == ((lf - lc - first) if (fusion and grid) else 0)
            x = torch.zeros(q["n"], dtype=torch.float64, device="cuda")
            r = s.solve(x, bd)
            torch.cuda.synchronize()
            assert s.coarse_fusion_grid() == ((lf - lc - first) if (fusion and grid) else 0)
        finally:
            feat.runtime_grid_forms(True)
        s.close()
        assert r["status"] == "success"
        out.append((r, x.cpu().numpy()))
    r0, x0 = out[-1]
    for r1, x1 in out[:-1]:
        assert r1["iters"] == r0["iters"]
        assert np.array_equal(r1["hist"].view(np.uint64), r0["hist"].view(np.uint64))
        assert np.array_equal(x1.view(np.uint64), x0.view(np.uint64))


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("nu_pre,nu_post", [(1, 1), (3, 2), (2, 5)])
def test_coarse_fusion_bitwise_step_counts(feat, torch_gpu, nu_pre, nu_post, precision):
    """The fused coarse levels with one smoothing step (the first-iterate-only
    branch), odd step counts (the ping-pong parity of the iterates) and unequal
    pre/post counts equal the separate launches bit for bit (3D level 5)."""
    print(f"coarse fusion {precision} V({nu_pre},{nu_post})", flush=True)
    torch = torch_gpu
    q = feat.q1_problem(3, 5)
    b = feat.rhs_random(23, q["n"], -1.0, 1.0)
    b[q["filt"].astype(np.int64)] = 0.0
    bd = torch.from_numpy(b).cuda()
    out = []
    for fusion in (True, False):
        cfg = feat.mgpcg_cfg(nu_pre=nu_pre, nu_post=nu_post, omega_jac=0.5, omega_rich=1.0, coarse="direct",
                             tol_rel=1e-8, max_iter=200, graphs=False, coarse_fusion=fusion,
                             precision=precision)
        s = feat.MgPcg(cfg, q1=(3, 5, 1))
        assert s.coarse_fusion_level() == (2 if fusion else 0)
        x = torch.zeros(q["n"], dtype=torch.float64, device="cuda")
        r = s.solve(x, bd)
        torch.cuda.synchronize()
        s.close()
        assert r["status"] == "success"
        out.append((r, x.cpu().numpy()))
    (r1, x1), (r0, x0) = out
    assert r1["iters"] == r0["iters"]
    assert np.array_equal(r1["hist"].view(np.uint64), r0["hist"].view(np.uint64))
    assert np.array_equal(x1.view(np.uint64), x0.view(np.uint64))


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_numeric_reinit_rebuilds_fused_levels(feat, torch_gpu, precision):
    """New matrix values followed by the reference's re-initialisation order
    (hierarchy init_numeric, then the solver tree's) replace the coarse inverse,
    the Jacobi diagonals and the value codes: a solve with the fused coarse levels
    afterwards equals, bit for bit, fresh solvers built on the new values with and
    without the fusion (no stale addresses from before the re-init)."""
    print(f"numeric re-init {precision}", flush=True)
    torch = torch_gpu
    levels = q1_levels(feat, 3, 5, 1)
    new = [lv["val"] * s for lv, s in zip(levels, (1.0, 1.0, 1.25, 1.0, 1.5))]
    n = levels[0]["n"]
    b = feat.rhs_random(29, n, -1.0, 1.0)
    b[levels[0]["filt"].astype(np.int64)] = 0.0
    bd = torch.from_numpy(b).cuda()

    def cfg(fusion):
        return feat.mgpcg_cfg(nu_pre=4, nu_post=4, omega_jac=0.5, omega_rich=1.0, coarse="direct",
                              tol_rel=1e-8, max_iter=100, graphs=True, coarse_fusion=fusion,
                              precision=precision)

    def run(s):
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        r = s.solve(x, bd)
        torch.cuda.synchronize()
        assert r["status"] == "success"
        return r, x.cpu().numpy()

    s = feat.MgPcg(cfg(True), levels=levels)
    assert s.coarse_fusion_level() == 2
    r_old, x_old = run(s)
    for i in (2, 4):
        s.set_values(i, new[i])
    r_re, x_re = run(s)
    s.close()
    assert not np.array_equal(x_re, x_old)
    changed = [dict(lv, val=v) for lv, v in zip(levels, new)]
    for fusion in (True, False):
        f = feat.MgPcg(cfg(fusion), levels=changed)
        r_f, x_f = run(f)
        f.close()
        assert r_f["iters"] == r_re["iters"]
        assert np.array_equal(x_f.view(np.uint64), x_re.view(np.uint64))


@pytest.mark.parametrize("steps", [2, 3, 4])
@pytest.mark.parametrize("rhs_kind", ["random", "zero", "nan"])
def test_first_iterate_from_producer_bitwise(feat, handle, torch_gpu, steps, rhs_kind):
    """A smoother that starts from the first iterate its right-hand side's producer formed
    (feat_defect_first_f64 / feat_apply_masked_first_f64 / feat_pcg_update_first_f64, then
    feat_rj_smooth_from1_f64) gives the bits of the separate first step, including the skip
    rule on a zero or non-finite right-hand side (applied by the smoother's next step)."""
    torch = torch_gpu
    k, _ = feat.libs()
    P = feat.ptr
    keep, ws = feat.reduce_ws(torch, "cuda")
    q, A = q1_level(feat, 3, 4)
    n = q["n"]
    rng = np.random.default_rng(40 + steps)
    keepA, desc = exec_desc(feat, handle, torch, A, "sell")
    v, c, r = keepA[:3]
    m = mask_of(torch, feat, handle, n, q["filt"])
    omega = 0.8
    inv = torch.empty(n, dtype=torch.float64, device="cuda")
    feat.check(k.feat_csr_extract_diag_f64_u64(handle.h, P(inv), P(v), P(c), P(r), n), "diag")
    feat.check(k.feat_component_invert_f64(handle.h, P(inv), P(inv), 0.5, n), "inv")
    x0 = rng.standard_normal(n)
    rhs = rng.standard_normal(n)
    if rhs_kind == "zero":
        # x = 0, rhs = 0: the defect and the restricted vector are exactly zero
        x0[:] = 0.0
        rhs[:] = 0.0
    elif rhs_kind == "nan":
        rhs[7] = np.nan
    xd, rd = dev(torch, x0), dev(torch, rhs)

    def vecs(k_):
        return [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(k_)]

    def b1(x, tmp):
        return x if steps & 1 else tmp

    # post-smoothing: defect, then the smoother accumulating into sol
    sol0 = rng.standard_normal(n)
    outs = []
    for first in (False, True):
        d, cor, tmp = vecs(3)
        sol = dev(torch, sol0)
        nrm = torch.zeros(1, dtype=torch.float64, device="cuda")
        if first:
            feat.check(k.feat_defect_first_f64(handle.h, desc, P(d), P(rd), P(xd), P(m), P(nrm), P(inv), omega,
                                               P(b1(cor, tmp)), ws), "defect first")
            fn = k.feat_rj_smooth_from1_f64
        else:
            feat.check(k.feat_defect_f64(handle.h, desc, P(d), P(rd), P(xd), P(m), P(nrm), ws), "defect")
            fn = k.feat_rj_smooth_f64
        feat.check(fn(handle.h, desc, P(inv), P(m), steps, omega, P(d), P(nrm), P(cor), P(tmp), None, P(sol), None, None,
                      ws), "smooth")
        handle.sync()
        outs.append((host(d), host(sol)))
    for a, b in zip(*outs):
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(a[~np.isnan(a)].view(np.uint64), b[~np.isnan(b)].view(np.uint64))
    if rhs_kind == "zero":
        assert np.array_equal(outs[1][1], sol0)   # skipped: sol unchanged
    # pre-smoothing: restriction-like producer (masked apply), then the smoother with d_out
    outs = []
    for first in (False, True):
        rr, x, tmp, dout = vecs(4)
        nrm = torch.zeros(1, dtype=torch.float64, device="cuda")
        src = rd if rhs_kind == "nan" else xd
        if first:
            feat.check(k.feat_apply_masked_first_f64(handle.h, desc, P(rr), P(src), P(m), P(nrm), P(inv), omega,
                                                     P(b1(x, tmp)), ws), "masked first")
            fn = k.feat_rj_smooth_from1_f64
        else:
            feat.check(k.feat_apply_masked_f64(handle.h, desc, P(rr), P(src), P(m), P(nrm), None, None, ws), "masked")
            fn = k.feat_rj_smooth_f64
        feat.check(fn(handle.h, desc, P(inv), P(m), steps, omega, P(rr), P(nrm), P(x), P(tmp), P(dout), None, None, None,
                      ws), "smooth")
        handle.sync()
        outs.append((host(x), host(dout)))
    for a, b in zip(*outs):
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(a[~np.isnan(a)].view(np.uint64), b[~np.isnan(b)].view(np.uint64))
    # PCG update: x, r as the plain update, x1 = filter(omega * (inv * r_new)) in one rounding each
    p_, q_ = rng.standard_normal((2, n))
    sc = torch.tensor([0.7, 1.3], dtype=torch.float64, device="cuda")
    res = []
    for first in (False, True):
        xx, rr = dev(torch, x0), dev(torch, rhs)
        x1 = torch.zeros(n, dtype=torch.float64, device="cuda")
        nrm = torch.zeros(1, dtype=torch.float64, device="cuda")
        if first:
            feat.check(k.feat_pcg_update_first_f64(handle.h, P(xx), P(rr), P(dev(torch, p_)), P(dev(torch, q_)), P(sc),
                                                   P(sc[1:]), P(nrm), n, P(inv), P(m), omega, P(x1), ws), "update first")
        else:
            feat.check(k.feat_pcg_update_f64(handle.h, P(xx), P(rr), P(dev(torch, p_)), P(dev(torch, q_)), P(sc),
                                             P(sc[1:]), P(nrm), n, ws), "update")
        handle.sync()
        res.append((host(xx), host(rr), host(x1)))
    assert np.array_equal(res[0][0], res[1][0], equal_nan=True)
    assert np.array_equal(res[0][1], res[1][1], equal_nan=True)
    cexp = host(inv) * res[1][1]
    cexp[q["filt"].astype(np.int64)] = 0.0
    assert np.array_equal(res[1][2], omega * cexp, equal_nan=True)

