"""Stokes part of the oracle (SaddlePointMatrix, UnitFilterBlocked + MeanFilter,
AmaVanka, transfers, Richardson(Schwarz(AmaVanka)) smoother, MG solve) pinned
bit-for-bit against the reference build, and the Q2/P1dc setup pinned by
closed forms."""
import numpy as np
import pytest

from helpers import stokes_levels, stokes_rhs


@pytest.fixture(scope="module")
def s31(feat):
    return stokes_levels(feat, 3, 1)


def _rand(lv, seed):
    return np.random.default_rng(seed).standard_normal(2 * lv["nv"] + lv["np"])


def test_q2p1dc_closed_forms(feat):
    for level in (1, 2, 3):
        lv = stokes_levels(feat, level, level)[0]
        N = 2 ** level
        m = 2 * N + 1
        h = 1.0 / N
        A, B, D = lv["A"].todense(), lv["B"].todense(), lv["D"].todense()
        nv = lv["nv"]
        a = np.arange(nv) % m * (h / 2)
        b = np.arange(nv) // m * (h / 2)
        # A is symmetric; constants are in its kernel (row sums of every element matrix vanish)
        assert np.allclose(A, A.T, atol=1e-14)
        for comp in (0, 1):
            u = np.zeros(2 * nv)
            u[comp::2] = 1.0
            assert np.max(np.abs(A @ u)) < 1e-13
        # u = (x^2, 0): weak form of -lap u - grad div u = (-4, 0) on interior test functions
        w1 = np.array([1 / 3, 4 / 3, 1 / 3]) * (h / 2)  # int of the 1D Q2 basis on one element
        lump1 = np.zeros(m)
        for e in range(N):
            lump1[2 * e:2 * e + 3] += w1
        lump = np.outer(lump1, lump1).ravel()         # int phi_I, lexicographic (y major)
        u = np.zeros(2 * nv)
        u[0::2] = a * a
        r = A @ u
        interior = (a > 0) & (a < 1) & (b > 0) & (b < 1)
        assert np.allclose(r[0::2][interior], -4.0 * lump[interior], rtol=1e-12, atol=1e-14)
        assert np.max(np.abs(r[1::2][interior])) < 1e-13
        # B = D^T and D * I(x, 0) = D * I(0, y) = -int psi_p = -dual
        assert np.array_equal(B, D.T)
        for comp, coord in ((0, a), (1, b)):
            u = np.zeros(2 * nv)
            u[comp::2] = coord
            assert np.allclose(D @ u, -lv["dual"], atol=1e-15)
        assert np.allclose(lv["dual"][0::3], h * h) and not lv["dual"][1::3].any()
        assert np.dot(lv["prim"], lv["dual"]) == 1.0


def test_q2p1dc_prolongation_exact(feat):
    f, c = stokes_levels(feat, 3, 2)
    for L in (f, c):
        N = 2 ** L["level"]
        m = 2 * N + 1
        L["x"] = np.arange(L["nv"]) % m / (2 * N)
        L["y"] = np.arange(L["nv"]) // m / (2 * N)
    Pv = f["Pv"].todense()
    for q in (lambda x, y: x * x, lambda x, y: x * y, lambda x, y: y * y - 3 * x + 1):
        assert np.allclose(Pv @ q(c["x"], c["y"]), q(f["x"], f["y"]), atol=1e-14)
    # P1dc: linear function 1 + 2x - 3y, dofs (centre value, (h/2) d/dx, (h/2) d/dy)
    def p1(L):
        N = 2 ** L["level"]
        e = np.arange(N * N)
        xc, yc = (e % N + 0.5) / N, (e // N + 0.5) / N
        v = np.zeros(3 * N * N)
        v[0::3], v[1::3], v[2::3] = 1 + 2 * xc - 3 * yc, 2 * (0.5 / N), -3 * (0.5 / N)
        return v
    assert np.allclose(f["Pp"].todense() @ p1(c), p1(f), atol=1e-15)


def test_stokes_apply_filters_transfers_vs_reference(feat, oracle, reference, s31):
    f, c = s31[0], s31[1]
    x, y = _rand(f, 1), _rand(f, 2)
    for kw in (dict(), dict(y=y, alpha=-1.0), dict(y=y, alpha=0.7)):
        assert np.array_equal(oracle.saddle_apply(f, x, **kw), reference.saddle_apply(f, x, **kw))
    # a pressure part with integral mean ~ 1e-18 exercises DenseVector::axpy's |alpha| < eps copy
    tiny = x.copy()
    tiny[2 * f["nv"]:] -= np.dot(tiny[2 * f["nv"]:], f["prim"]) * f["dual"] / np.dot(f["prim"], f["dual"])
    for v in (x, tiny):
        for kind in ("rhs", "sol", "def", "cor"):
            assert np.array_equal(oracle.stokes_filter(f, v, kind), reference.stokes_filter(f, v, kind)), kind
    xc = _rand(c, 3)
    for prol, v in ((True, xc), (False, x)):
        assert np.array_equal(oracle.stokes_transfer(f, c, v, prol), reference.stokes_transfer(f, c, v, prol))


@pytest.mark.parametrize("skip", [False, True])
def test_amavanka_vs_reference(feat, oracle, reference, s31, skip):
    for lv in s31:
        d = _rand(lv, 4)
        o = oracle.vanka(lv, 1.0, skip, d=d)
        r = reference.vanka(lv, 1.0, skip, d=d)
        for bo, br in zip(o["blocks"], r["blocks"]):
            nnz = int(bo["rowptr"][-1])
            assert np.array_equal(bo["rowptr"], br["rowptr"])
            assert np.array_equal(bo["col"][:nnz], br["col"][:nnz])
            assert np.array_equal(bo["val"], br["val"])
        assert np.array_equal(o["c"], r["c"])
        N = 2 ** lv["level"]
        assert o["macros"] == N * N and o["stride"] == 21
        if skip:
            assert np.array_equal(o["mask"], r["mask"])


def test_amavanka_sum_identity(feat, oracle):
    """Vanka matrix = diag(omega / #macros) * sum_m R_m^T (R_m K1 R_m^T)^-1 R_m with dense
    numpy inverses of the element patches (9 Q2 nodes x 2 + 3 P1dc dofs = 21) of the
    type-1 matrix K1 (unit rows in A, zero rows in B on the Dirichlet nodes)."""
    from oracle.pyoracle import Bcsr
    lv = stokes_levels(feat, 2, 2)[0]
    nv, np_ = lv["nv"], lv["np"]
    n = 2 * nv + np_
    A, B, D = lv["A"].todense(), lv["B"].todense(), lv["D"].todense()
    for i in lv["filt"].astype(int):
        A[2 * i:2 * i + 2, :] = 0.0
        A[2 * i, 2 * i] = A[2 * i + 1, 2 * i + 1] = 1.0
        B[2 * i:2 * i + 2, :] = 0.0
    K = np.zeros((n, n))
    K[:2 * nv, :2 * nv], K[:2 * nv, 2 * nv:], K[2 * nv:, :2 * nv] = A, B, D
    N, m = 4, 9
    Vd = np.zeros((n, n))
    cnt = np.zeros(n)
    omega = 0.8
    for ey in range(N):
        for ex in range(N):
            nodes = sorted((2 * ey + j) * m + 2 * ex + i for j in range(3) for i in range(3))
            e = ey * N + ex
            idx = [2 * q + c for q in nodes for c in (0, 1)] + [2 * nv + 3 * e + k for k in range(3)]
            Vd[np.ix_(idx, idx)] += np.linalg.inv(K[np.ix_(idx, idx)])
            cnt[idx] += 1
    Vd *= (omega / cnt)[:, None]
    o = oracle.vanka(lv, omega, False)
    blk = [Bcsr(b["rows"], {0: nv, 1: np_, 2: nv, 3: np_}[k], b["bh"], b["bw"], b["val"], b["col"], b["rowptr"])
           for k, b in enumerate(o["blocks"])]
    V = np.zeros((n, n))
    V[:2 * nv, :2 * nv], V[:2 * nv, 2 * nv:] = blk[0].todense(), blk[1].todense()
    V[2 * nv:, :2 * nv], V[2 * nv:, 2 * nv:] = blk[2].todense(), blk[3].todense()
    assert np.max(np.abs(V - Vd)) <= 1e-11 * np.max(np.abs(Vd))


def test_stokes_smoother_vs_reference(feat, oracle, reference, s31):
    f = s31[0]
    d = _rand(f, 5)
    for steps in (1, 5):
        assert np.array_equal(oracle.stokes_smooth(f, d, steps), reference.stokes_smooth(f, d, steps))


@pytest.mark.parametrize("lf,lc", [(2, 1), (3, 1), (4, 1)])
def test_stokes_mg_solve_vs_reference(feat, oracle, reference, lf, lc):
    lv = stokes_levels(feat, lf, lc)
    x0, b = stokes_rhs(lv[0], oracle)
    o = oracle.stokes_mg_solve(lv, x0, b, tol_rel=1e-8)
    r = reference.stokes_mg_solve(lv, x0, b, tol_rel=1e-8)
    assert o["status"] == r["status"] == "success"
    assert o["iters"] == r["iters"]
    assert np.array_equal(o["x"], r["x"])
    # defect norms: the reference's TupleVector norm contracts differently where it is inlined
    assert np.allclose(o["hist"], r["hist"], rtol=4 * np.finfo(float).eps, atol=0)
    assert np.all(np.diff(np.log(o["hist"])) < 0)


def _bcsr_dense(m, rows, cols, bh, bw):
    val, col, rp = m
    M = np.zeros((rows * bh, cols * bw))
    for r in range(rows):
        for k in range(int(rp[r]), int(rp[r + 1])):
            c = int(col[k])
            M[r * bh:(r + 1) * bh, c * bw:(c + 1) * bw] = val[k * bh * bw:(k + 1) * bh * bw].reshape(bh, bw)
    return M


@pytest.mark.parametrize("level", [1, 2, 3])
def test_stokes_matrices_equal_reference_assembly(feat, reference, level):
    """A, B and D of the product's Q2/P1dc assembly equal the reference's own assembly
    (DuDvOperatorBlocked<2> and GradPresDivVeloAssembler on a RefinedUnitCubeFactory mesh,
    as kernel/solver/amavanka-test.cpp:110-117 assembles them) after renumbering.  Velocity
    dofs: the reference numbers vertices, then edges, then cells (Lagrange2's uniform dof
    assignment); each is placed at its vertex, edge midpoint or cell centre and mapped to the
    lexicographic node there.  Pressure: the reference's P1dc basis of a cell is (1, x^, y^)
    with (x^, y^) the cell's reference coordinates, whose axes are the directions of its
    local edges v0->v1 and v0->v2; the product's is (1, x^, y^) along +x and +y.  Each cell's
    three columns are related by the signed permutation T read off its vertices; on the
    refined unit square T is the identity for every cell (asserted), so the pressure dofs
    map by cell centre alone."""
    r = reference.stokes_asm(level)
    o = feat.stokes_problem(level)
    N = 2 ** level
    m = 2 * N + 1
    h = 1.0 / N
    assert (r["nv"], r["np"]) == (o["nv"], o["np"]) == (m * m, 3 * N * N)
    xy = r["coords"]
    cells = r["cells"].astype(np.int64)
    vpos = np.concatenate([xy, xy[r["edges"].astype(np.int64)].mean(axis=1), xy[cells].mean(axis=1)])
    assert vpos.shape[0] == r["nv"]
    g = np.rint(vpos * 2 * N).astype(np.int64)
    assert np.allclose(g, vpos * 2 * N, atol=1e-9)
    vmap = g[:, 1] * m + g[:, 0]                      # reference velocity dof -> product node
    assert np.array_equal(np.sort(vmap), np.arange(m * m))
    cen = xy[cells].mean(axis=1)
    cmap = np.rint((cen[:, 1] - 0.5 * h) / h).astype(np.int64) * N + np.rint((cen[:, 0] - 0.5 * h) / h).astype(np.int64)
    assert np.array_equal(np.sort(cmap), np.arange(N * N))
    # per reference cell: T (3x3) with psi_ref = T psi_product
    T = np.zeros((N * N, 3, 3))
    T[:, 0, 0] = 1.0
    d1 = (xy[cells[:, 1]] - xy[cells[:, 0]]) / h
    d2 = (xy[cells[:, 2]] - xy[cells[:, 0]]) / h
    T[:, 1, 1:], T[:, 2, 1:] = d1, d2
    assert np.allclose(np.abs(T[:, 1:, 1:]).sum(axis=2), 1.0)   # axis-aligned cells: signed permutations
    assert np.array_equal(T, np.broadcast_to(np.eye(3), T.shape))  # every cell oriented +x, +y
    # dense matrices, reference renumbered to the product's velocity ordering
    nv, np_ = o["nv"], o["np"]
    Ar = _bcsr_dense(r["A"], nv, nv, 2, 2)
    Br = _bcsr_dense(r["B"], nv, np_, 2, 1)
    Dr = _bcsr_dense(r["D"], np_, nv, 1, 2)
    Ao = _bcsr_dense(o["A"], nv, nv, 2, 2)
    Bo = _bcsr_dense(o["B"], nv, np_, 2, 1)
    Do = _bcsr_dense(o["D"], np_, nv, 1, 2)
    vp = np.empty(2 * nv, np.int64)                   # product velocity row of each reference row
    vp[0::2], vp[1::2] = 2 * vmap, 2 * vmap + 1
    Ar_p = np.zeros_like(Ar)
    Ar_p[np.ix_(vp, vp)] = Ar
    scale = np.abs(Ao).max()
    assert np.abs(Ar_p - Ao).max() <= 1e-14 * scale
    # B: reference pressure columns = product columns combined by T per cell
    Bo_t = np.zeros_like(Bo)
    for cr in range(N * N):
        co = cmap[cr]
        Bo_t[:, 3 * cr:3 * cr + 3] = Bo[:, 3 * co:3 * co + 3] @ T[cr].T
    Br_p = np.zeros_like(Br)
    Br_p[vp, :] = Br
    assert np.abs(Br_p - Bo_t).max() <= 1e-14 * np.abs(Bo).max()
    Dr_p = np.zeros_like(Dr)
    Dr_p[:, vp] = Dr
    assert np.abs(Dr_p - Bo_t.T).max() <= 1e-14 * np.abs(Do).max()
    assert np.array_equal(Do, Bo.T)
