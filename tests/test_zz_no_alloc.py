"""The allocation-free transposed CSR apply (feat_csr_apply_transposed_noalloc_f64_u64)
against the allocating entry point (feat_csr_apply_f64_u64 with transposed = 1), bit for
bit: two calls in a row, the second on a smaller matrix, every (a, b) branch; a matrix
beyond the fixed scratch is rejected before anything runs."""
import numpy as np
import pytest

from oracle.pyoracle import Csr
from tests.helpers import random_csr
from tests.test_gpu_lafem import dev, host, upload_csr

pytestmark = pytest.mark.gpu

FEAT_ERR_UNSUPPORTED = 5


def _both(feat, handle, torch, A, x, y, a, b):
    k, _ = feat.libs()
    v, c, r = upload_csr(torch, A)
    xd = dev(torch, x)
    yd = dev(torch, y)
    out = []
    for fn in ("alloc", "noalloc"):
        rd = torch.full((A.cols,), 7.0, dtype=torch.float64, device="cuda")
        if fn == "alloc":
            rc = k.feat_csr_apply_f64_u64(handle.h, feat.ptr(rd), a, feat.ptr(xd), b, feat.ptr(yd), feat.ptr(v),
                                          feat.ptr(c), feat.ptr(r), A.rows, A.cols, A.nnz, 1)
        else:
            rc = k.feat_csr_apply_transposed_noalloc_f64_u64(handle.h, feat.ptr(rd), a, feat.ptr(xd), b, feat.ptr(yd),
                                                             feat.ptr(v), feat.ptr(c), feat.ptr(r), A.rows, A.cols,
                                                             A.nnz)
        feat.check(rc, fn)
        out.append(host(rd))
    return out


def test_noalloc_transposed_matches_allocating(feat, handle, torch_gpu, oracle):
    torch = torch_gpu
    q = feat.q1_problem(3, 6)   # prolongation 274,625 x 35,937: about 5.5 MiB of sort buffers
    big = Csr(q["n"], q["pcols"], q["pval"], q["pcol"], q["prowptr"])
    small = random_csr(np.random.default_rng(3), 900, 410, 0, 9, [(5, 300)])
    rng = np.random.default_rng(4)
    for A in (big, small):   # the second call is the smaller one
        x = rng.standard_normal(A.rows)
        y = rng.standard_normal(A.cols)
        for a, b in [(1.0, 0.0), (-0.7, 1.0), (0.3, -2.5)]:
            ga, gn = _both(feat, handle, torch, A, x, y, a, b)
            assert np.array_equal(ga.view(np.uint64), gn.view(np.uint64)), (A.rows, a, b)
        assert np.array_equal(gn, oracle.csr_apply(A, x, y, 0.3, -2.5, transposed=True))


def test_noalloc_transposed_rejects_oversize(feat, handle, torch_gpu):
    torch = torch_gpu
    k, _ = feat.libs()
    n, cols = 2_000_000, 1000   # 2M entries need about 40 MB: beyond the 32 MiB scratch
    rp = torch.arange(n + 1, dtype=torch.int64, device="cuda")
    col = torch.remainder(torch.arange(n, dtype=torch.int64, device="cuda"), cols)
    val = torch.ones(n, dtype=torch.float64, device="cuda")
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    r = torch.full((cols,), 3.0, dtype=torch.float64, device="cuda")
    rc = k.feat_csr_apply_transposed_noalloc_f64_u64(handle.h, feat.ptr(r), 1.0, feat.ptr(x), 0.0, None, feat.ptr(val),
                                                     feat.ptr(col), feat.ptr(rp), n, cols, n)
    assert rc == FEAT_ERR_UNSUPPORTED
    assert b"32 MiB" in k.feat_last_error()
    torch.cuda.synchronize()
    assert torch.all(r == 3.0)
