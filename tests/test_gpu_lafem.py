"""GPU parity of the LAFEM kernels against the oracle, through the C-ABI.

Bar: bit-exact where the per-element arithmetic is the oracle's (SpMV rows,
transposed SpMV, axpy, scale, component ops, filters, diag); dot/norm within
1e-13 relative (tree order vs sequential).
"""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import random_csr, validate_csr

pytestmark = pytest.mark.gpu


def dev(torch, a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda")


def host(t):
    return t.cpu().numpy()


def upload_csr(torch, A, idx=np.uint64, val=np.float64):
    validate_csr(A)
    ti = torch.int64 if idx == np.uint64 else torch.int32
    return (dev(torch, A.val.astype(val)), dev(torch, A.col.astype(np.int64)).to(ti),
            dev(torch, A.rowptr.astype(np.int64)).to(ti))


def csr_apply(feat, h, torch, A, x, y, a, b, tr=False, idx="u64", fp="f64", alias_ry=False):
    k, _ = feat.libs()
    itype = np.uint64 if idx == "u64" else np.uint32
    vtype = np.float64 if fp == "f64" else np.float32
    v, c, r = upload_csr(torch, A, itype, vtype)
    xd = dev(torch, x.astype(vtype))
    yd = dev(torch, y.astype(vtype))
    rd = yd if alias_ry else torch.empty(A.cols if tr else A.rows, dtype=yd.dtype, device="cuda")
    fn = getattr(k, f"feat_csr_apply_{fp}_{idx}")
    feat.check(fn(h.h, feat.ptr(rd), a, feat.ptr(xd), b, feat.ptr(yd), feat.ptr(v), feat.ptr(c),
                  feat.ptr(r), A.rows, A.cols, A.nnz, int(tr)), fn.__name__)
    h.sync()
    return host(rd)


@pytest.mark.parametrize("seed,rows,cols,maxlen,longrows", [
    (1, 1, 1, 3, ()),
    (2, 63, 40, 7, ()),
    (3, 1000, 700, 30, ()),
    (4, 3000, 3000, 40, ((5, 5000), (2999, 2500))),   # rows longer than one LDS chunk
    (5, 257, 1000, 0, ()),                            # all rows empty
])
def test_csr_apply_f64_bitwise(feat, handle, torch_gpu, oracle, seed, rows, cols, maxlen, longrows):
    rng = np.random.default_rng(seed)
    A = random_csr(rng, rows, cols, 0, maxlen, longrows)
    x = rng.standard_normal(cols)
    y = rng.standard_normal(rows)
    for a, b in [(1.0, 0.0), (-1.0, 1.0), (4711.1, 1.0), (2.5, -0.7)]:
        ref = oracle.csr_apply(A, x, y, a, b)
        for idx in ("u64", "u32"):
            got = csr_apply(feat, handle, torch_gpu, A, x, y, a, b, idx=idx)
            assert np.array_equal(got, ref), (a, b, idx)
    # r == y aliasing (sparse_matrix_csr.hpp:1799)
    got = csr_apply(feat, handle, torch_gpu, A, x, y, -1.0, 1.0, alias_ry=True)
    assert np.array_equal(got, oracle.csr_apply(A, x, y, -1.0, 1.0))
    # |b| < eps: y is not read
    got = csr_apply(feat, handle, torch_gpu, A, x, np.full(rows, np.nan), 1.0, 0.0)
    assert np.array_equal(got, oracle.csr_apply(A, x))


def test_csr_apply_f32_bitwise(feat, handle, torch_gpu, oracle):
    rng = np.random.default_rng(11)
    A = random_csr(rng, 2000, 1500, 0, 30)
    x = rng.standard_normal(1500).astype(np.float32)
    y = rng.standard_normal(2000).astype(np.float32)
    ref = oracle.csr_apply_f32(A, x, y, -1.0, 1.0)
    for idx in ("u32", "u64"):
        got = csr_apply(feat, handle, torch_gpu, A, x, y, -1.0, 1.0, idx=idx, fp="f32")
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("seed,rows,cols,maxlen,longrows", [
    (12, 800, 600, 20, ()),
    (13, 3000, 2000, 40, ((7, 1900),)),      # one long row: its entries spread over most columns
    (14, 500, 5000, 6, ()),                   # many empty columns
    (15, 100, 1, 1, ()),                      # one column
])
def test_csr_apply_transposed(feat, handle, torch_gpu, oracle, seed, rows, cols, maxlen, longrows):
    """The transposed apply adds each column's entries in the generic scatter's order:
    bit-identical to the oracle (pinned to the reference's generic code), for every
    (a, b), both index widths, FP32, r == y and |b| < eps (y not read)."""
    rng = np.random.default_rng(seed)
    A = random_csr(rng, rows, cols, 0, maxlen, longrows)
    x = rng.standard_normal(rows)
    y = rng.standard_normal(cols)
    for a, b in [(4711.1, 1.0), (1.0, 0.0), (-1.0, 1.0), (0.3, -2.5)]:
        ref = oracle.csr_apply(A, x, y, a, b, transposed=True)
        for idx in ("u64", "u32"):
            got = csr_apply(feat, handle, torch_gpu, A, x, y, a, b, tr=True, idx=idx)
            assert np.array_equal(got, ref), (a, b, idx)
    got = csr_apply(feat, handle, torch_gpu, A, x, y, -1.0, 1.0, tr=True, alias_ry=True)
    assert np.array_equal(got, oracle.csr_apply(A, x, y, -1.0, 1.0, transposed=True))
    got = csr_apply(feat, handle, torch_gpu, A, x, np.full(cols, np.nan), 2.0, 0.0, tr=True)
    assert np.array_equal(got, oracle.csr_apply(A, x, None, 2.0, 0.0, transposed=True))
    x32, y32 = x.astype(np.float32), y.astype(np.float32)
    ref32 = oracle.csr_apply_f32(A, x32, y32, -1.5, 1.0, transposed=True)
    got32 = csr_apply(feat, handle, torch_gpu, A, x32, y32, -1.5, 1.0, tr=True, idx="u32", fp="f32")
    assert np.array_equal(got32.view(np.uint32), ref32.view(np.uint32))


def test_csr_apply_transposed_equals_explicit_restriction(feat, handle, torch_gpu):
    # P^T x through the transposed apply equals R x with the explicit R = P^T (whose
    # rows list P's column entries in ascending row order): the same sums, bit for bit
    from oracle.pyoracle import Csr, csr_transpose
    q = feat.q1_problem(3, 4)
    P = Csr(q["n"], q["pcols"], q["pval"], q["pcol"], q["prowptr"])
    R = csr_transpose(P)
    x = np.random.default_rng(16).standard_normal(P.rows)
    z = np.zeros(P.cols)
    got = csr_apply(feat, handle, torch_gpu, P, x, z, 1.0, 0.0, tr=True)
    ref = csr_apply(feat, handle, torch_gpu, R, x, z, 1.0, 0.0)
    assert np.array_equal(got, ref)


def test_csr_apply_kat_golden(feat, handle, torch_gpu, golden):
    from oracle.pyoracle import Csr
    for n in [1, 2, 8, 64, 512]:
        rows = [(i, j, 2.0 if i == j else -1.0) for i in range(n) for j in (i - 1, i, i + 1) if 0 <= j < n]
        rp = np.zeros(n + 1, np.uint64)
        for i, _, _ in rows:
            rp[i + 1] += 1
        A = Csr(n, n, np.array([v for *_, v in rows]), np.array([j for _, j, _ in rows], np.uint64),
                np.cumsum(rp).astype(np.uint64))
        x = (np.arange(n) % 100) * 1.234
        y = 2.0 - (np.arange(n) % 42)
        for tag, a in [("0", 0.0), ("m1", -1.0), ("s", 4711.1)]:
            got = csr_apply(feat, handle, torch_gpu, A, x, y, a, 1.0)
            assert np.array_equal(got, golden[f"kat{n}_a{tag}"])


def test_csr_alias_rejected(feat, handle, torch_gpu):
    k, _ = feat.libs()
    x = torch_gpu.zeros(4, dtype=torch_gpu.float64, device="cuda")
    rp = torch_gpu.zeros(5, dtype=torch_gpu.int64, device="cuda")
    rc = k.feat_csr_apply_f64_u64(handle.h, feat.ptr(x), 1.0, feat.ptr(x), 0.0, None, None, None,
                                  feat.ptr(rp), 4, 4, 0, 0)
    assert rc == 3  # FEAT_ERR_ALIAS


@pytest.mark.parametrize("bh,bw", [(2, 2), (2, 1), (1, 2), (3, 3)])
def test_bcsr_apply_bitwise(feat, handle, torch_gpu, oracle, bh, bw):
    k, _ = feat.libs()
    rng = np.random.default_rng(13)
    S = random_csr(rng, 500, 400, 0, 9)
    val = rng.standard_normal(S.nnz * bh * bw)
    x = rng.standard_normal(400 * bw)
    y = rng.standard_normal(500 * bh)
    ref = oracle.bcsr_apply(bh, bw, 500, val, S.col, S.rowptr, x, y, -1.0, 1.0)
    vd, cd, rd = dev(torch_gpu, val), dev(torch_gpu, S.col.astype(np.int64)), dev(torch_gpu, S.rowptr.astype(np.int64))
    xd, yd = dev(torch_gpu, x), dev(torch_gpu, y)
    r = torch_gpu.empty(500 * bh, dtype=torch_gpu.float64, device="cuda")
    feat.check(k.feat_bcsr_apply_f64_u64(handle.h, bh, bw, feat.ptr(r), -1.0, feat.ptr(xd), 1.0, feat.ptr(yd),
                                         feat.ptr(vd), feat.ptr(cd), feat.ptr(rd), 500, 400, S.nnz), "bcsr")
    handle.sync()
    assert np.array_equal(host(r), ref)


def test_blas1(feat, handle, torch_gpu, oracle):
    k, _ = feat.libs()
    rng = np.random.default_rng(14)
    n = 100003
    x, y, z = rng.standard_normal((3, n))
    xd, yd, zd = dev(torch_gpu, x), dev(torch_gpu, y), dev(torch_gpu, z)
    r = torch_gpu.empty(n, dtype=torch_gpu.float64, device="cuda")
    P = feat.ptr
    # axpy, three aliasing branches (arch/axpy_generic.hpp:25-50)
    feat.check(k.feat_axpy_f64(handle.h, P(r), 0.37, P(xd), P(yd), n), "axpy")
    handle.sync()
    assert np.array_equal(host(r), oracle.axpy(0.37, x, y))
    for alias in ("x", "y"):
        t = xd.clone() if alias == "x" else yd.clone()
        xa = t if alias == "x" else xd
        ya = t if alias == "y" else yd
        feat.check(k.feat_axpy_f64(handle.h, P(t), 0.37, P(xa), P(ya), n), "axpy alias")
        handle.sync()
        assert np.array_equal(host(t), oracle.axpy(0.37, x, y, alias=alias)), alias
    feat.check(k.feat_scale_f64(handle.h, P(r), P(xd), 0.3, n), "scale")
    handle.sync()
    assert np.array_equal(host(r), oracle.scale(x, 0.3))
    feat.check(k.feat_component_product_f64(handle.h, P(r), P(xd), P(yd), n), "cprod")
    handle.sync()
    assert np.array_equal(host(r), oracle.component_product(x, y))
    feat.check(k.feat_component_invert_f64(handle.h, P(r), P(xd), 0.7, n), "cinv")
    handle.sync()
    assert np.array_equal(host(r), oracle.component_invert(x, 0.7))
    d = C.c_double()
    feat.check(k.feat_dot_f64_host(handle.h, C.byref(d), P(xd), P(yd), n), "dot")
    assert abs(d.value - oracle.dot(x, y)) <= 1e-13 * np.abs(x * y).sum()
    feat.check(k.feat_norm2_f64_host(handle.h, C.byref(d), P(xd), n), "norm2")
    assert abs(d.value - oracle.norm2(x)) <= 1e-14 * oracle.norm2(x)
    feat.check(k.feat_triple_dot_f64_host(handle.h, C.byref(d), P(xd), P(yd), P(zd), n), "tdot")
    assert abs(d.value - oracle.triple_dot(x, y, z)) <= 1e-13 * np.abs(x * y * z).sum()
    # dot closed form (dense_vector-test.cpp:366)
    a = (np.arange(n) + 1.0) / n
    b = 1.0 / (np.arange(n) + 1.0)
    ad, bd = dev(torch_gpu, a), dev(torch_gpu, b)   # keep both alive across the call
    feat.check(k.feat_dot_f64_host(handle.h, C.byref(d), P(ad), P(bd), n), "dot")
    assert abs(d.value - 1.0) <= np.finfo(float).eps ** 0.8
    # empty vectors
    feat.check(k.feat_dot_f64_host(handle.h, C.byref(d), P(xd), P(yd), 0), "dot0")
    assert d.value == 0.0
    # deterministic: repeated reductions are bit-identical
    vals = set()
    for _ in range(5):
        feat.check(k.feat_dot_f64_host(handle.h, C.byref(d), P(xd), P(yd), n), "dot")
        vals.add(d.value)
    assert len(vals) == 1


@pytest.mark.parametrize("bs", [1, 2, 3])
def test_unit_filter_blocked(feat, handle, torch_gpu, oracle, bs):
    """UnitFilterBlocked filter_rhs / filter_def through the C-ABI, with and without
    ign_nans (NaNs planted in the filter values), bit for bit against the oracle."""
    k, _ = feat.libs()
    P = feat.ptr
    rng = np.random.default_rng(60 + bs)
    nb = 5000
    idx = np.unique(rng.integers(0, nb, 900)).astype(np.uint64)
    vals = rng.standard_normal(idx.size * bs)
    vals[::7] = np.nan
    x = rng.standard_normal(nb * bs)
    idd = dev(torch_gpu, idx.astype(np.int64))
    svd = dev(torch_gpu, vals)
    for ign in (0, 1):
        xd = dev(torch_gpu, x)
        feat.check(k.feat_unit_filter_blocked_rhs_f64_u64(handle.h, bs, P(xd), P(svd), P(idd), idx.size, ign), "rhs")
        handle.sync()
        assert np.array_equal(host(xd), oracle.unit_filter_blocked(True, bs, x, vals, idx, ign), equal_nan=True)
        xd = dev(torch_gpu, x)
        feat.check(k.feat_unit_filter_blocked_def_f64_u64(handle.h, bs, P(xd), P(svd) if ign else None, P(idd),
                                                          idx.size, ign), "def")
        handle.sync()
        assert np.array_equal(host(xd), oracle.unit_filter_blocked(False, bs, x, vals, idx, ign))


def test_diag_filters_convert(feat, handle, torch_gpu, oracle):
    k, _ = feat.libs()
    P = feat.ptr
    rng = np.random.default_rng(15)
    A = random_csr(rng, 3000, 3000, 0, 12)
    v, c, r = upload_csr(torch_gpu, A)
    di = torch_gpu.empty(3000, dtype=torch_gpu.int64, device="cuda")
    feat.check(k.feat_csr_diag_index_u64(handle.h, P(di), P(c), P(r), 3000), "diag")
    dv = torch_gpu.empty(3000, dtype=torch_gpu.float64, device="cuda")
    feat.check(k.feat_csr_extract_diag_f64_u64(handle.h, P(dv), P(v), P(c), P(r), 3000), "extract_diag")
    handle.sync()
    assert np.array_equal(host(di).astype(np.uint64), oracle.diag_index(A))
    assert np.array_equal(host(dv), oracle.extract_diag(A))
    x = rng.standard_normal(3000)
    idx = np.unique(rng.integers(0, 3000, 500)).astype(np.uint64)
    sv = rng.standard_normal(idx.size)
    xd = dev(torch_gpu, x)
    idd = dev(torch_gpu, idx.astype(np.int64))
    svd = dev(torch_gpu, sv)
    feat.check(k.feat_unit_filter_rhs_f64_u64(handle.h, P(xd), P(svd), P(idd), idx.size), "rhs")
    handle.sync()
    assert np.array_equal(host(xd), oracle.unit_filter_rhs(x, sv, idx))
    feat.check(k.feat_unit_filter_def_f64_u64(handle.h, P(xd), P(idd), idx.size), "def")
    handle.sync()
    assert np.array_equal(host(xd), oracle.unit_filter_def(oracle.unit_filter_rhs(x, sv, idx), idx))
    # bit mask equals the index list
    m = torch_gpu.empty((3000 + 31) // 32, dtype=torch_gpu.int32, device="cuda")
    feat.check(k.feat_unit_filter_mask_u64(handle.h, P(m), 3000, P(idd), idx.size), "mask")
    handle.sync()
    bits = np.unpackbits(host(m).view(np.uint8), bitorder="little")[:3000]
    assert np.array_equal(np.nonzero(bits)[0].astype(np.uint64), idx)
    # mean filter (mean_filter.hpp:247-264)
    prim, dual = np.ones(3000), rng.uniform(0.5, 1.5, 3000)
    vol = float(prim @ dual)
    yd = dev(torch_gpu, x)
    primd, duald = dev(torch_gpu, prim), dev(torch_gpu, dual)
    feat.check(k.feat_mean_filter_f64(handle.h, P(yd), P(primd), P(duald), vol, 3000), "mean")
    handle.sync()
    assert np.allclose(host(yd), oracle.mean_filter(x, prim, dual, vol), rtol=0, atol=1e-14)
    # conversions
    f = torch_gpu.empty(3000, dtype=torch_gpu.float32, device="cuda")
    xd64 = dev(torch_gpu, x)
    feat.check(k.feat_convert_f64_f32(handle.h, P(f), P(xd64), 3000), "cvt")
    handle.sync()
    assert np.array_equal(host(f), x.astype(np.float32))
    g = torch_gpu.empty(3000, dtype=torch_gpu.float64, device="cuda")
    feat.check(k.feat_convert_f32_f64(handle.h, P(g), P(f), 3000), "cvt")
    handle.sync()
    assert np.array_equal(host(g), x.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("nbytes,off", [(16, 0), (8 * 1000003, 0), (4096 * 16 * 300 + 16, 0), (1000, 0), (4096, 4)])
def test_copy_device(feat, handle, torch_gpu, nbytes, off):
    # the kernel copy (16-byte aligned pointers and size) and its hipMemcpyAsync fallback
    torch = torch_gpu
    k, _ = feat.libs()
    src = torch.randint(0, 256, (nbytes + 64,), dtype=torch.uint8, device="cuda")
    dst = torch.zeros(nbytes + 64, dtype=torch.uint8, device="cuda")
    feat.check(k.feat_copy_device(handle.h, dst.data_ptr() + off, src.data_ptr() + off, nbytes), "copy")
    handle.sync()
    s, d = src.cpu().numpy(), dst.cpu().numpy()
    assert np.array_equal(d[off:off + nbytes], s[off:off + nbytes])
    assert not d[:off].any() and not d[off + nbytes:].any()
