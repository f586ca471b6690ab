"""ctypes bindings of the CPU oracle (liboracle.so) and of the reference build
(_ref/libfeatref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; the product never imports it.
"""
import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_ORACLE = os.path.join(HERE, "liboracle.so")
LIB_REF = os.path.join(HERE, "_ref", "libfeatref.so")

_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p

# solver status codes (solver/base.hpp:48-64)
STATUS = {0: "undefined", 1: "progress", 2: "success", 3: "aborted", 4: "diverged",
          5: "max_iter", 6: "stagnated"}
COARSE_NONE, COARSE_SMOOTHER, COARSE_JACOBI, COARSE_DIRECT = 0, 1, 2, 3
# flags of orc_mg_pcg_solve_mixed (oracle/feat_oracle.h)
MIXED_F16, MIXED_SLIP_INV64 = 1, 2


@dataclass
class Csr:
    """CSR matrix with uint64 indices (reference SparseMatrixCSR<DT, Index>)."""
    rows: int
    cols: int
    val: np.ndarray
    col: np.ndarray
    rowptr: np.ndarray

    @property
    def nnz(self):
        return int(self.val.size)

    def todense(self):
        d = np.zeros((self.rows, self.cols), dtype=np.float64)
        for r in range(self.rows):
            for i in range(int(self.rowptr[r]), int(self.rowptr[r + 1])):
                d[r, int(self.col[i])] += float(self.val[i])
        return d


class _CsrS(C.Structure):
    _fields_ = [("rows", C.c_uint64), ("cols", C.c_uint64), ("nnz", C.c_uint64),
                ("val", _vp), ("col", _vp), ("rowptr", _vp)]


class _LevelS(C.Structure):
    _fields_ = [("A", _CsrS), ("filt", _vp), ("nfilt", C.c_uint64), ("P", _CsrS), ("R", _CsrS)]


class _IterCfgS(C.Structure):
    _fields_ = [("tol_rel", C.c_double), ("tol_abs", C.c_double), ("tol_abs_low", C.c_double),
                ("div_rel", C.c_double), ("div_abs", C.c_double), ("stag_rate", C.c_double),
                ("min_iter", C.c_uint64), ("max_iter", C.c_uint64),
                ("min_stag_iter", C.c_uint64)]


class _MgCfgS(C.Structure):
    _fields_ = [("nu_pre", C.c_int), ("nu_post", C.c_int), ("omega_jac", C.c_double),
                ("omega_rich", C.c_double), ("coarse", C.c_int), ("coarse_omega", C.c_double)]


class _ResultS(C.Structure):
    _fields_ = [("status", C.c_int), ("num_iter", C.c_uint64), ("def_init", C.c_double),
                ("def_final", C.c_double)]


@dataclass
class Bcsr:
    """Block CSR with row-major bh x bw blocks (reference SparseMatrixBCSR)."""
    rows: int
    cols: int
    bh: int
    bw: int
    val: np.ndarray
    col: np.ndarray
    rowptr: np.ndarray

    def todense(self):
        d = np.zeros((self.rows * self.bh, self.cols * self.bw))
        v = self.val.reshape(-1, self.bh, self.bw)
        for r in range(self.rows):
            for i in range(int(self.rowptr[r]), int(self.rowptr[r + 1])):
                c = int(self.col[i])
                d[r * self.bh:(r + 1) * self.bh, c * self.bw:(c + 1) * self.bw] += v[i]
        return d


class _BcsrS(C.Structure):
    _fields_ = [("rows", C.c_uint64), ("cols", C.c_uint64), ("bh", C.c_int), ("bw", C.c_int),
                ("val", _vp), ("col", _vp), ("rowptr", _vp)]


class _StokesLevelS(C.Structure):
    _fields_ = [("nv", C.c_uint64), ("np", C.c_uint64), ("A", _BcsrS), ("B", _BcsrS), ("D", _BcsrS),
                ("filt", _vp), ("filt_val", _vp), ("nfilt", C.c_uint64), ("prim", _vp), ("dual", _vp),
                ("Pv", _CsrS), ("Rv", _CsrS), ("Pp", _CsrS), ("Rp", _CsrS)]


class _StokesCfgS(C.Structure):
    _fields_ = [("steps", C.c_int), ("omega_rich", C.c_double), ("vanka_omega", C.c_double),
                ("coarse_skip_singular", C.c_int), ("omega_outer", C.c_double)]


def _arr(a, dt, keep):
    a = np.ascontiguousarray(a, dtype=dt)
    keep.append(a)
    return a.ctypes.data if a.size else None


def _bcsr_s(m, keep):
    return _BcsrS(m.rows, m.cols, m.bh, m.bw, _arr(m.val, np.float64, keep), _arr(m.col, np.uint64, keep),
                  _arr(m.rowptr, np.uint64, keep))


def _stokes_levels_s(levels):
    """levels: list of dicts {nv, np, A, B, D, filt, filt_val, prim, dual, Pv, Rv, Pp, Rp}."""
    keep = []
    arr = (_StokesLevelS * len(levels))()
    for i, lv in enumerate(levels):
        arr[i] = _StokesLevelS(lv["nv"], lv["np"], _bcsr_s(lv["A"], keep), _bcsr_s(lv["B"], keep),
                               _bcsr_s(lv["D"], keep), _arr(lv["filt"], np.uint64, keep),
                               _arr(lv["filt_val"], np.float64, keep), len(lv["filt"]),
                               _arr(lv["prim"], np.float64, keep), _arr(lv["dual"], np.float64, keep),
                               _csr_s(lv.get("Pv"), keep), _csr_s(lv.get("Rv"), keep),
                               _csr_s(lv.get("Pp"), keep), _csr_s(lv.get("Rp"), keep))
    return arr, keep


def _vanka_blocks(nv, np_, sizes):
    shapes = [(nv, 2, 2), (nv, 2, 1), (np_, 1, 2), (np_, 1, 1)]
    blocks = []
    for (rows, bh, bw), nnz in zip(shapes, sizes):
        blocks.append(dict(rows=rows, bh=bh, bw=bw, rowptr=np.zeros(rows + 1, np.uint64),
                           col=np.zeros(max(int(nnz), 1), np.uint64),
                           val=np.zeros(max(int(nnz), 1) * bh * bw)))
    ptrs = [(C.POINTER(C.c_uint64) * 4)(), (C.POINTER(C.c_uint64) * 4)(), (C.POINTER(C.c_double) * 4)()]
    for k, b in enumerate(blocks):
        ptrs[0][k] = b["rowptr"].ctypes.data_as(C.POINTER(C.c_uint64))
        ptrs[1][k] = b["col"].ctypes.data_as(C.POINTER(C.c_uint64))
        ptrs[2][k] = b["val"].ctypes.data_as(C.POINTER(C.c_double))
    return blocks, ptrs


def _csr_s(m, keep):
    if m is None:
        return _CsrS(0, 0, 0, None, None, None)
    val = np.ascontiguousarray(m.val, dtype=np.float64)
    col = np.ascontiguousarray(m.col, dtype=np.uint64)
    rp = np.ascontiguousarray(m.rowptr, dtype=np.uint64)
    keep += [val, col, rp]
    return _CsrS(m.rows, m.cols, val.size, val.ctypes.data, col.ctypes.data, rp.ctypes.data)


def _levels_s(levels):
    """levels: list of dicts {A, filt, P, R} with level 0 the finest."""
    keep = []
    arr = (_LevelS * len(levels))()
    for i, lv in enumerate(levels):
        f = np.ascontiguousarray(lv.get("filt", np.zeros(0, np.uint64)), dtype=np.uint64)
        keep.append(f)
        arr[i] = _LevelS(_csr_s(lv["A"], keep), f.ctypes.data if f.size else None, f.size,
                         _csr_s(lv.get("P"), keep), _csr_s(lv.get("R"), keep))
    return arr, keep


class Oracle:
    """The C oracle."""

    def __init__(self, path=LIB_ORACLE):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        L = C.CDLL(path)
        self.L = L
        L.orc_rng_next.restype = C.c_uint64
        L.orc_rng_next.argtypes = [C.POINTER(C.c_uint64)]
        L.orc_rng_fill_f64.argtypes = [C.POINTER(C.c_uint64), _f64p, C.c_uint64, C.c_double,
                                       C.c_double]
        L.orc_csr_apply_f64.argtypes = [_f64p, C.c_double, _f64p, C.c_double, _f64p, _f64p,
                                        _u64p, _u64p, C.c_uint64, C.c_uint64, C.c_int]
        L.orc_csr_apply_f32.argtypes = [_f32p, C.c_float, _f32p, C.c_float, _f32p, _f32p,
                                        _u64p, _u64p, C.c_uint64, C.c_uint64, C.c_int]
        L.orc_bcsr_apply_f64.argtypes = [C.c_int, C.c_int, _f64p, C.c_double, _f64p, C.c_double,
                                         _f64p, _f64p, _u64p, _u64p, C.c_uint64]
        L.orc_axpy_f64.argtypes = [_f64p, C.c_double, _f64p, _f64p, C.c_uint64]
        for fn in ("orc_dot_f64",):
            getattr(L, fn).restype = C.c_double
            getattr(L, fn).argtypes = [_f64p, _f64p, C.c_uint64]
        L.orc_triple_dot_f64.restype = C.c_double
        L.orc_triple_dot_f64.argtypes = [_f64p, _f64p, _f64p, C.c_uint64]
        L.orc_norm2_f64.restype = C.c_double
        L.orc_norm2_f64.argtypes = [_f64p, C.c_uint64]
        L.orc_scale_f64.argtypes = [_f64p, _f64p, C.c_double, C.c_uint64]
        L.orc_component_product_f64.argtypes = [_f64p, _f64p, _f64p, C.c_uint64]
        L.orc_component_invert_f64.argtypes = [_f64p, _f64p, C.c_double, C.c_uint64]
        L.orc_diag_index.argtypes = [_u64p, _u64p, _u64p, C.c_uint64]
        L.orc_extract_diag_f64.argtypes = [_f64p, _f64p, _u64p, _u64p, C.c_uint64]
        L.orc_unit_filter_rhs_f64.argtypes = [_f64p, _f64p, _u64p, C.c_uint64]
        L.orc_unit_filter_def_f64.argtypes = [_f64p, _u64p, C.c_uint64]
        L.orc_unit_filter_blocked_f64.argtypes = [C.c_int, C.c_int, _f64p, _vp, _u64p, C.c_uint64, C.c_int]
        L.orc_mean_filter_f64.argtypes = [_f64p, _f64p, _f64p, C.c_double, C.c_uint64]
        L.orc_invert_matrix_f64.restype = C.c_double
        L.orc_invert_matrix_f64.argtypes = [C.c_int, C.c_int, _f64p, _i32p]
        L.orc_iter_cfg_default.argtypes = [C.POINTER(_IterCfgS)]
        L.orc_mg_pcg_solve.restype = C.c_int
        L.orc_mg_pcg_solve.argtypes = [C.POINTER(_LevelS), C.c_int, C.POINTER(_MgCfgS),
                                       C.POINTER(_IterCfgS), _f64p, _f64p, _vp, C.POINTER(_ResultS)]
        L.orc_mg_pcg_solve_mixed.restype = C.c_int
        L.orc_mg_pcg_solve_mixed.argtypes = [C.POINTER(_LevelS), C.c_int, C.POINTER(_MgCfgS),
                                             C.POINTER(_IterCfgS), C.c_int, _f64p, _f64p, _vp,
                                             C.POINTER(_ResultS)]
        L.orc_f16_values.argtypes = [_f64p, C.c_uint64, _f32p, C.POINTER(C.c_float)]
        for fn in ("orc_pcg_solve", "orc_pcg_apply"):
            f = getattr(L, fn)
            f.restype = C.c_int
            f.argtypes = [C.POINTER(_CsrS), _vp, C.c_uint64, C.c_int, C.c_double,
                          C.POINTER(_IterCfgS), _f64p, _f64p, _vp, C.POINTER(_ResultS)]
        for fn in ("orc_richardson_solve", "orc_richardson_apply"):
            f = getattr(L, fn)
            f.restype = C.c_int
            f.argtypes = [C.POINTER(_CsrS), _vp, C.c_uint64, C.c_int, C.c_double, C.c_double,
                          C.POINTER(_IterCfgS), _f64p, _f64p, _vp, C.POINTER(_ResultS)]

        L.orc_saddle_apply.argtypes = [C.POINTER(_StokesLevelS), _f64p, _f64p, _vp, C.c_double]
        L.orc_stokes_filter.argtypes = [C.POINTER(_StokesLevelS), _f64p, C.c_int]
        L.orc_stokes_transfer.argtypes = [C.POINTER(_StokesLevelS), C.POINTER(_StokesLevelS), C.c_int, _f64p, _f64p]
        L.orc_stokes_smooth.argtypes = [C.POINTER(_StokesLevelS), C.c_int, C.c_double, C.c_double, _f64p, _f64p]
        L.orc_vanka_create.restype = _vp
        L.orc_vanka_create.argtypes = [C.POINTER(_StokesLevelS), C.c_double, C.c_int]
        L.orc_vanka_sizes.argtypes = [_vp, _u64p]
        L.orc_vanka_copy.argtypes = [_vp, C.POINTER(C.POINTER(C.c_uint64)), C.POINTER(C.POINTER(C.c_uint64)),
                                     C.POINTER(C.POINTER(C.c_double)), _vp]
        L.orc_vanka_apply.argtypes = [_vp, C.POINTER(_StokesLevelS), _f64p, _f64p]
        L.orc_vanka_free.argtypes = [_vp]
        L.orc_stokes_mg_solve.restype = C.c_int
        L.orc_stokes_mg_solve.argtypes = [C.POINTER(_StokesLevelS), C.c_int, C.POINTER(_StokesCfgS),
                                          C.POINTER(_IterCfgS), _f64p, _f64p, _vp, C.POINTER(_ResultS)]

    # -- learned smoother network
    def mlp_forward(self, p, X, round_bf16=True):
        self._mlp_decl()
        X = np.ascontiguousarray(X, dtype=np.uint16)
        y = np.zeros(X.shape[0], np.float32)
        self.L.orc_mlp_forward(np.ascontiguousarray(p, dtype=np.float32).ctypes.data, X.ctypes.data, X.shape[0],
                               int(round_bf16), y.ctypes.data)
        return y

    def mlp_grad(self, p, X, t, round_bf16=True):
        self._mlp_decl()
        X = np.ascontiguousarray(X, dtype=np.uint16)
        g = np.zeros(8448, np.float32)
        self.L.orc_mlp_grad(np.ascontiguousarray(p, dtype=np.float32).ctypes.data, X.ctypes.data,
                            np.ascontiguousarray(t, dtype=np.float32).ctypes.data, X.shape[0], int(round_bf16),
                            g.ctypes.data)
        return g

    def mlp_features(self, m, r, e, dinv, cfeat, nodes):
        self._mlp_decl()
        nodes = np.ascontiguousarray(nodes, dtype=np.uint64)
        X = np.zeros((nodes.size, 64), np.uint16)
        t = np.zeros(nodes.size, np.float32)
        self.L.orc_mlp_features(m, *(np.ascontiguousarray(a).ctypes.data for a in (r, e, dinv)),
                                np.ascontiguousarray(cfeat, dtype=np.float32).ctypes.data, nodes.ctypes.data,
                                nodes.size, X.ctypes.data, t.ctypes.data)
        return X, t

    def mlp_smooth(self, m, p, r, dinv, cfeat, omega, jacobi, x):
        self._mlp_decl()
        x = np.array(x, dtype=np.float64)
        self.L.orc_mlp_smooth(m, np.ascontiguousarray(p, dtype=np.float32).ctypes.data,
                              np.ascontiguousarray(r).ctypes.data, np.ascontiguousarray(dinv).ctypes.data,
                              np.ascontiguousarray(cfeat, dtype=np.float32).ctypes.data, omega, jacobi, x.ctypes.data)
        return x

    def _mlp_decl(self):
        L = self.L
        if getattr(self, "_mlp_ok", False):
            return
        L.orc_mlp_forward.argtypes = [_vp, _vp, C.c_uint64, C.c_int, _vp]
        L.orc_mlp_grad.argtypes = [_vp, _vp, _vp, C.c_uint64, C.c_int, _vp]
        L.orc_mlp_features.argtypes = [C.c_uint64, _vp, _vp, _vp, _vp, _vp, C.c_uint64, _vp, _vp]
        L.orc_mlp_smooth.argtypes = [C.c_uint64, _vp, _vp, _vp, _vp, C.c_double, C.c_double, _vp]
        self._mlp_ok = True

    # -- Stokes
    def saddle_apply(self, lv, x, y=None, alpha=1.0):
        arr, keep = _stokes_levels_s([lv])
        r = np.zeros(2 * lv["nv"] + lv["np"])
        yy = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
        self.L.orc_saddle_apply(arr, r, np.ascontiguousarray(x, dtype=np.float64),
                                None if yy is None else yy.ctypes.data, alpha)
        return r

    def stokes_filter(self, lv, v, kind):
        arr, keep = _stokes_levels_s([lv])
        v = np.array(v, dtype=np.float64)
        self.L.orc_stokes_filter(arr, v, {"rhs": 0, "sol": 1, "def": 2, "cor": 3}[kind])
        return v

    def stokes_transfer(self, fine, coarse, x, prol):
        arr, keep = _stokes_levels_s([fine, coarse])
        out = coarse if not prol else fine
        r = np.zeros(2 * out["nv"] + out["np"])
        self.L.orc_stokes_transfer(C.byref(arr[0]), C.byref(arr[1]), int(prol), r,
                                     np.ascontiguousarray(x, dtype=np.float64))
        return r

    def stokes_smooth(self, lv, rhs, steps=5, omega=0.7, vanka_omega=1.0):
        arr, keep = _stokes_levels_s([lv])
        sol = np.zeros(2 * lv["nv"] + lv["np"])
        self.L.orc_stokes_smooth(arr, steps, omega, vanka_omega, sol, np.ascontiguousarray(rhs, dtype=np.float64))
        return sol

    def vanka(self, lv, omega=1.0, skip_singular=False, d=None):
        """AmaVanka on the type-1 matrix: the four Vanka blocks, the macro mask and,
        if d is given, the smoother applied to d."""
        arr, keep = _stokes_levels_s([lv])
        h = self.L.orc_vanka_create(arr, omega, int(skip_singular))
        try:
            sz = np.zeros(6, np.uint64)
            self.L.orc_vanka_sizes(h, sz)
            blocks, ptrs = _vanka_blocks(lv["nv"], lv["np"], sz[1:5])
            mask = np.zeros(max(int(sz[0]), 1), np.int32)
            self.L.orc_vanka_copy(h, ptrs[0], ptrs[1], ptrs[2], mask.ctypes.data)
            out = dict(blocks=blocks, macros=int(sz[0]), stride=int(sz[5]),
                       mask=mask[:int(sz[0])] if skip_singular else None)
            if d is not None:
                c = np.zeros(2 * lv["nv"] + lv["np"])
                self.L.orc_vanka_apply(h, arr, c, np.ascontiguousarray(d, dtype=np.float64))
                out["c"] = c
        finally:
            self.L.orc_vanka_free(h)
        return out

    def stokes_mg_solve(self, levels, x0, b, steps=5, omega_rich=0.7, vanka_omega=1.0,
                        coarse_skip_singular=True, omega_outer=1.0, max_iter=100, **kw):
        arr, keep = _stokes_levels_s(levels)
        cfg = _StokesCfgS(steps, omega_rich, vanka_omega, int(coarse_skip_singular), omega_outer)
        it = self.iter_cfg(max_iter=max_iter, **kw)
        x = np.array(x0, dtype=np.float64)
        hist = np.full(max_iter + 1, np.nan)
        res = _ResultS()
        st = self.L.orc_stokes_mg_solve(arr, len(levels), C.byref(cfg), C.byref(it), x,
                                        np.ascontiguousarray(b, dtype=np.float64), hist.ctypes.data,
                                        C.byref(res))
        return dict(status=STATUS[st], iters=int(res.num_iter), x=x, hist=hist[:res.num_iter + 1],
                    def_init=res.def_init, def_final=res.def_final)

    # -- RNG
    def rng_fill(self, seed, n, a, b):
        s = C.c_uint64(seed)
        v = np.empty(n, dtype=np.float64)
        self.L.orc_rng_fill_f64(C.byref(s), v, n, a, b)
        return v

    def rng_next(self, seed, k):
        s = C.c_uint64(seed)
        return np.array([self.L.orc_rng_next(C.byref(s)) for _ in range(k)], dtype=np.uint64)

    # -- kernels (all return new arrays; inputs untouched)
    def color_ranks(self, A, row_colors=False):
        """Processing rank of each row's colour (greedy first fit, ascending row count)."""
        L = self.L
        L.orc_color_ranks.restype = C.c_uint32
        L.orc_color_ranks.argtypes = [C.c_uint64, _vp, _vp, C.c_int, _vp]
        rank = np.empty(A.rows, np.uint32)
        rp = np.ascontiguousarray(A.rowptr, np.uint64)
        col = np.ascontiguousarray(A.col, np.uint64)
        nc = L.orc_color_ranks(A.rows, rp.ctypes.data, col.ctypes.data, int(row_colors), rank.ctypes.data)
        return rank, int(nc)

    def mcsor_apply(self, A, rank, ncolors, omega, x, ssor=False):
        """y = M^{-1} x of the multicolour SOR (ssor=False) or SSOR sweep."""
        L = self.L
        L.orc_mcsor_apply.argtypes = [C.c_uint64, _vp, _vp, _vp, _vp, C.c_uint32, C.c_double, C.c_int, _vp, _vp]
        rp = np.ascontiguousarray(A.rowptr, np.uint64)
        col = np.ascontiguousarray(A.col, np.uint64)
        val = np.ascontiguousarray(A.val, np.float64)
        rank = np.ascontiguousarray(rank, np.uint32)
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(A.rows)
        L.orc_mcsor_apply(A.rows, rp.ctypes.data, col.ctypes.data, val.ctypes.data, rank.ctypes.data, ncolors,
                          omega, int(ssor), x.ctypes.data, y.ctypes.data)
        return y

    def sor_generic_apply(self, A, omega, x, ssor=False):
        """y = M^{-1} x of the generic sequential SOR / SSOR."""
        L = self.L
        L.orc_sor_generic_apply.argtypes = [C.c_uint64, _vp, _vp, _vp, C.c_double, C.c_int, _vp, _vp]
        rp = np.ascontiguousarray(A.rowptr, np.uint64)
        col = np.ascontiguousarray(A.col, np.uint64)
        val = np.ascontiguousarray(A.val, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(A.rows)
        L.orc_sor_generic_apply(A.rows, rp.ctypes.data, col.ctypes.data, val.ctypes.data, omega, int(ssor),
                                x.ctypes.data, y.ctypes.data)
        return y

    def ilu0_factor(self, A):
        """ILU(0) factors (l, u, d_inverted, rptr_l, rptr_u) in the CSR order of A."""
        L = self.L
        L.orc_ilu0_factor.argtypes = [C.c_uint64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        rp = np.ascontiguousarray(A.rowptr, np.uint64)
        col = np.ascontiguousarray(A.col, np.uint64)
        val = np.ascontiguousarray(A.val, np.float64)
        l, u, d = np.empty(max(val.size, 1)), np.empty(max(val.size, 1)), np.empty(A.rows)
        rl, ru = np.empty(A.rows + 1, np.uint64), np.empty(A.rows + 1, np.uint64)
        L.orc_ilu0_factor(A.rows, rp.ctypes.data, col.ctypes.data, val.ctypes.data, l.ctypes.data, u.ctypes.data,
                          d.ctypes.data, rl.ctypes.data, ru.ctypes.data)
        return l[:int(rl[-1])].copy(), u[:int(ru[-1])].copy(), d, rl, ru

    def ilu0_apply(self, A, f, x):
        L = self.L
        L.orc_ilu0_apply.argtypes = [C.c_uint64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        l, u, d, rl, ru = (np.ascontiguousarray(a) for a in f)
        rp = np.ascontiguousarray(A.rowptr, np.uint64)
        col = np.ascontiguousarray(A.col, np.uint64)
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(A.rows)
        L.orc_ilu0_apply(A.rows, rp.ctypes.data, col.ctypes.data, l.ctypes.data, u.ctypes.data, d.ctypes.data,
                         rl.ctypes.data, ru.ctypes.data, x.ctypes.data, y.ctypes.data)
        return y

    def ilup_pattern(self, A, p):
        """The level-p split pattern (rl, cl, ru, cu) of ILUCoreSymbolic::factorize_symbolic."""
        L = self.L
        L.orc_ilup_pattern.restype = C.c_int
        L.orc_ilup_pattern.argtypes = [C.c_uint64, _vp, _vp, C.c_int, C.c_uint64, C.c_uint64, _vp, _vp, _vp, _vp,
                                       C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        rp = np.ascontiguousarray(A.rowptr, np.uint64)
        col = np.ascontiguousarray(A.col, np.uint64)
        cap = max(col.size, 16)
        while True:
            rl, ru = np.empty(A.rows + 1, np.uint64), np.empty(A.rows + 1, np.uint64)
            cl, cu = np.empty(cap, np.uint64), np.empty(cap, np.uint64)
            nl, nu = C.c_uint64(), C.c_uint64()
            rc = L.orc_ilup_pattern(A.rows, rp.ctypes.data, col.ctypes.data, int(p), cap, cap, rl.ctypes.data,
                                    cl.ctypes.data, ru.ctypes.data, cu.ctypes.data, C.byref(nl), C.byref(nu))
            if rc == 0:
                return rl, cl[:nl.value].copy(), ru, cu[:nu.value].copy()
            cap *= 2

    def ilup_factor(self, A, p):
        """ILU(p) factors (l, u, d_inverted) on the level-p pattern, with the pattern."""
        L = self.L
        L.orc_ilu_factor_pattern.argtypes = [C.c_uint64] + [_vp] * 10
        rl, cl, ru, cu = self.ilup_pattern(A, p)
        rp = np.ascontiguousarray(A.rowptr, np.uint64)
        col = np.ascontiguousarray(A.col, np.uint64)
        val = np.ascontiguousarray(A.val, np.float64)
        l, u, d = np.empty(max(cl.size, 1)), np.empty(max(cu.size, 1)), np.empty(A.rows)
        L.orc_ilu_factor_pattern(A.rows, rp.ctypes.data, col.ctypes.data, val.ctypes.data, rl.ctypes.data,
                                 cl.ctypes.data, ru.ctypes.data, cu.ctypes.data, l.ctypes.data, u.ctypes.data,
                                 d.ctypes.data)
        return dict(l=l[:cl.size].copy(), u=u[:cu.size].copy(), d=d, rl=rl, cl=cl, ru=ru, cu=cu)

    def ilup_apply(self, f, x):
        L = self.L
        L.orc_ilu_apply_pattern.argtypes = [C.c_uint64] + [_vp] * 9
        n = f["d"].size
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(n)
        L.orc_ilu_apply_pattern(n, f["rl"].ctypes.data, f["cl"].ctypes.data, f["ru"].ctypes.data, f["cu"].ctypes.data,
                                np.ascontiguousarray(f["l"]).ctypes.data, np.ascontiguousarray(f["u"]).ctypes.data,
                                f["d"].ctypes.data, x.ctypes.data, y.ctypes.data)
        return y

    def csr_apply(self, A, x, y=None, a=1.0, b=0.0, transposed=False, r=None):
        nr = A.cols if transposed else A.rows
        if y is None:
            y = np.zeros(nr)
        if r is None:
            r = np.empty(nr)
        self.L.orc_csr_apply_f64(r, a, np.ascontiguousarray(x, np.float64), b,
                                 np.ascontiguousarray(y, np.float64), A.val, A.col, A.rowptr,
                                 A.rows, A.cols, int(transposed))
        return r

    def csr_apply_f32(self, A, x, y=None, a=1.0, b=0.0, transposed=False):
        nr = A.cols if transposed else A.rows
        y = np.zeros(nr, np.float32) if y is None else np.ascontiguousarray(y, np.float32)
        r = np.empty(nr, np.float32)
        self.L.orc_csr_apply_f32(r, a, np.ascontiguousarray(x, np.float32), b, y,
                                 np.ascontiguousarray(A.val, np.float32), A.col, A.rowptr,
                                 A.rows, A.cols, int(transposed))
        return r

    def f16_values(self, val):
        """(FP32 values of the FP16 numbers, power-of-two scale) as the FP16-valued
        form stores a matrix."""
        v = np.ascontiguousarray(val, np.float64)
        out = np.empty(max(v.size, 1), np.float32)
        sc = C.c_float()
        self.L.orc_f16_values(v, v.size, out, C.byref(sc))
        return out[:v.size], sc.value

    def bcsr_apply(self, bh, bw, rows, val, col, rowptr, x, y=None, a=1.0, b=0.0):
        y = np.zeros(rows * bh) if y is None else y
        r = np.empty(rows * bh)
        self.L.orc_bcsr_apply_f64(bh, bw, r, a, np.ascontiguousarray(x, np.float64), b,
                                  np.ascontiguousarray(y, np.float64),
                                  np.ascontiguousarray(val, np.float64),
                                  np.ascontiguousarray(col, np.uint64),
                                  np.ascontiguousarray(rowptr, np.uint64), rows)
        return r

    def axpy(self, a, x, y, alias=None):
        """r = a*x + y; alias 'x' or 'y' makes r share that operand's storage."""
        x = np.ascontiguousarray(x, np.float64).copy()
        y = np.ascontiguousarray(y, np.float64).copy()
        raw = self.L["orc_axpy_f64"]
        raw.argtypes = [_vp, C.c_double, _vp, _vp, C.c_uint64]
        r = x if alias == "x" else (y if alias == "y" else np.empty_like(y))
        raw(r.ctypes.data, a, x.ctypes.data, y.ctypes.data, y.size)
        return r

    def dot(self, x, y):
        return self.L.orc_dot_f64(x, y, x.size)

    def triple_dot(self, x, y, z):
        return self.L.orc_triple_dot_f64(x, y, z, x.size)

    def norm2(self, x):
        return self.L.orc_norm2_f64(x, x.size)

    def scale(self, x, s):
        r = np.empty_like(x)
        self.L.orc_scale_f64(r, x, s, x.size)
        return r

    def component_product(self, x, y):
        r = np.empty_like(x)
        self.L.orc_component_product_f64(r, x, y, x.size)
        return r

    def component_invert(self, x, s=1.0):
        r = np.empty_like(x)
        self.L.orc_component_invert_f64(r, x, s, x.size)
        return r

    def diag_index(self, A):
        d = np.empty(A.rows, np.uint64)
        self.L.orc_diag_index(d, A.col, A.rowptr, A.rows)
        return d

    def extract_diag(self, A):
        d = np.empty(A.rows)
        self.L.orc_extract_diag_f64(d, A.val, A.col, A.rowptr, A.rows)
        return d

    def unit_filter_blocked(self, rhs, bs, v, vals, idx, ign_nans=False):
        """UnitFilterBlocked filter_rhs (rhs=True) / filter_def of block size bs."""
        v = np.array(v, np.float64)
        idx = np.ascontiguousarray(idx, np.uint64)
        vals = None if vals is None else np.ascontiguousarray(vals, np.float64)
        self.L.orc_unit_filter_blocked_f64(int(rhs), bs, v, None if vals is None else vals.ctypes.data, idx,
                                           idx.size, int(ign_nans))
        return v

    def unit_filter_def(self, v, idx):
        v = v.copy()
        idx = np.ascontiguousarray(idx, np.uint64)
        self.L.orc_unit_filter_def_f64(v, idx, idx.size)
        return v

    def unit_filter_rhs(self, v, vals, idx):
        v = v.copy()
        idx = np.ascontiguousarray(idx, np.uint64)
        self.L.orc_unit_filter_rhs_f64(v, np.ascontiguousarray(vals, np.float64), idx, idx.size)
        return v

    def mean_filter(self, v, w_dot, w_add, volume):
        v = v.copy()
        self.L.orc_mean_filter_f64(v, w_dot, w_add, volume, v.size)
        return v

    def invert_matrix(self, a, stride=None):
        a = np.ascontiguousarray(a, np.float64).copy()
        n = a.shape[0]
        stride = stride or a.shape[1]
        p = np.zeros(n, np.int32)
        det = self.L.orc_invert_matrix_f64(n, stride, a.reshape(-1), p)
        return a, det

    # -- solvers
    def iter_cfg(self, **kw):
        c = _IterCfgS()
        self.L.orc_iter_cfg_default(C.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        return c

    def mg_pcg_solve(self, levels, x0, b, nu_pre=4, nu_post=4, omega_jac=0.5, omega_rich=1.0,
                     coarse=COARSE_DIRECT, coarse_omega=1.0, tol_rel=1e-8, max_iter=100,
                     precision="f64", planted_slip=False, **kw):
        """precision: "f64", "f32" or "f16" (FP16 values).  planted_slip (mixed forms
        only) turns on the oracle's planted rounding slip, for the test that shows
        the pin against the reference catches it."""
        arr, keep = _levels_s(levels)
        if precision == "f64":
            fn = self.L.orc_mg_pcg_solve
        else:
            flags = (MIXED_F16 if precision == "f16" else 0) | (MIXED_SLIP_INV64 if planted_slip else 0)
            fn = lambda *a: self.L.orc_mg_pcg_solve_mixed(*a[:4], flags, *a[4:])
        mg = _MgCfgS(nu_pre, nu_post, omega_jac, omega_rich, coarse, coarse_omega)
        it = self.iter_cfg(tol_rel=tol_rel, max_iter=max_iter, **kw)
        x = np.ascontiguousarray(x0, np.float64).copy()
        hist = np.full(max_iter + 1, np.nan)
        res = _ResultS()
        st = fn(arr, len(levels), C.byref(mg), C.byref(it), x, np.ascontiguousarray(b, np.float64),
                hist.ctypes.data, C.byref(res))
        return dict(x=x, status=STATUS[st], iters=int(res.num_iter), def_init=res.def_init,
                    def_final=res.def_final, hist=hist[:int(res.num_iter) + 1])

    def _single(self, fn, A, filt, args, x0, b, max_iter, **kw):
        keep = []
        a = _csr_s(A, keep)
        f = np.ascontiguousarray(filt if filt is not None else np.zeros(0), np.uint64)
        it = self.iter_cfg(max_iter=max_iter, **kw)
        x = np.ascontiguousarray(x0, np.float64).copy()
        hist = np.full(max_iter + 1, np.nan)
        res = _ResultS()
        st = fn(C.byref(a), f.ctypes.data if f.size else None, f.size, *args, C.byref(it), x,
                np.ascontiguousarray(b, np.float64), hist.ctypes.data, C.byref(res))
        return dict(x=x, status=STATUS[st], iters=int(res.num_iter), def_init=res.def_init,
                    def_final=res.def_final, hist=hist[:int(res.num_iter) + 1])

    def pcg(self, A, x0, b, filt=None, precond=0, omega=1.0, apply=False, max_iter=100, **kw):
        fn = self.L.orc_pcg_apply if apply else self.L.orc_pcg_solve
        return self._single(fn, A, filt, (precond, omega), x0, b, max_iter, **kw)

    def richardson(self, A, x0, b, filt=None, precond=1, omega_jac=1.0, omega_rich=1.0,
                   apply=False, max_iter=100, **kw):
        fn = self.L.orc_richardson_apply if apply else self.L.orc_richardson_solve
        return self._single(fn, A, filt, (precond, omega_jac, omega_rich), x0, b, max_iter, **kw)


class Reference:
    """The reference's own generic code, built by oracle/Makefile into _ref/."""

    def __init__(self, path=LIB_REF):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
        L = C.CDLL(path)
        self.L = L
        L.ref_init()
        L.ref_csr_apply_f64.argtypes = [_f64p, C.c_double, _f64p, C.c_double, _f64p, _f64p,
                                        _u64p, _u64p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int]
        L.ref_bcsr_apply_f64.argtypes = [C.c_int, C.c_int, _f64p, C.c_double, _f64p, C.c_double,
                                         _f64p, _f64p, _u64p, _u64p, C.c_uint64, C.c_uint64,
                                         C.c_uint64]
        L.ref_axpy_f64.argtypes = [_vp, C.c_double, _vp, _vp, C.c_uint64]
        L.ref_dot_f64.restype = C.c_double
        L.ref_dot_f64.argtypes = [_f64p, _f64p, C.c_uint64]
        L.ref_triple_dot_f64.restype = C.c_double
        L.ref_triple_dot_f64.argtypes = [_f64p, _f64p, _f64p, C.c_uint64]
        L.ref_norm2_f64.restype = C.c_double
        L.ref_norm2_f64.argtypes = [_f64p, C.c_uint64]
        L.ref_scale_f64.argtypes = [_f64p, _f64p, C.c_double, C.c_uint64]
        L.ref_component_product_f64.argtypes = [_f64p, _f64p, _f64p, C.c_uint64]
        L.ref_component_invert_f64.argtypes = [_f64p, _f64p, C.c_double, C.c_uint64]
        L.ref_diag_index.argtypes = [_u64p, _u64p, _u64p, C.c_uint64]
        L.ref_invert_matrix_f64.restype = C.c_double
        L.ref_invert_matrix_f64.argtypes = [C.c_int, C.c_int, _f64p, _i32p]
        L.ref_rng_fill_f64.argtypes = [C.c_uint64, _f64p, C.c_uint64, C.c_double, C.c_double]
        L.ref_pointstar_fd.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint64), _vp, _vp, _vp, _vp]
        L.ref_mg_pcg_solve.restype = C.c_int
        L.ref_mg_pcg_solve.argtypes = [C.POINTER(_LevelS), C.c_int, C.c_int, C.c_int, C.c_double,
                                       C.c_double, C.c_int, C.c_double, C.c_double, C.c_uint64,
                                       _f64p, _f64p, _vp, C.c_uint64, C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.ref_mg_pcg_solve_mixed.restype = C.c_int
        L.ref_mg_pcg_solve_mixed.argtypes = [C.POINTER(_LevelS), C.c_int, C.c_int, C.c_int, C.c_double,
                                             C.c_double, C.c_int, C.c_double, C.c_double, C.c_uint64,
                                             _f64p, _f64p, _vp, C.c_uint64, C.POINTER(C.c_uint64),
                                             C.POINTER(C.c_double)]
        L.ref_csr_apply_f32.argtypes = [_f32p, C.c_float, _f32p, C.c_float, _f32p, _f32p,
                                        _u64p, _u64p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int]
        L.ref_component_invert_f32.argtypes = [_f32p, _f32p, C.c_float, C.c_uint64]
        L.ref_unit_filter_blocked_f64.argtypes = [C.c_int, C.c_int, _f64p, _f64p, _u64p, C.c_uint64, C.c_int]
        L.ref_single_solve.restype = C.c_int
        L.ref_single_solve.argtypes = [C.c_int, C.POINTER(_CsrS), _vp, C.c_uint64, C.c_int,
                                       C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int,
                                       _f64p, _f64p, _vp, C.c_uint64, C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_double)]
        L.ref_csr_write.argtypes = [C.c_int, C.c_char_p, C.POINTER(_CsrS), C.c_int]
        L.ref_csr_read.argtypes = [C.c_int, C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_uint64), _vp, _vp, _vp, C.c_uint64]
        L.ref_dv_write.argtypes = [C.c_int, C.c_char_p, _vp, C.c_uint64]
        L.ref_dv_read.argtypes = [C.c_int, C.c_char_p, C.POINTER(C.c_uint64), _vp, C.c_uint64]
        L.ref_q1_create.restype = _vp
        L.ref_q1_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_char_p]
        L.ref_q1_sizes.argtypes = [_vp, _u64p]
        L.ref_q1_copy.argtypes = [_vp, _f64p, _u64p, _u64p, _u64p, _f64p, _f64p, _vp, _vp, _vp]
        L.ref_q1_free.argtypes = [_vp]
        L.ref_vanka_create.restype = _vp
        L.ref_vanka_create.argtypes = [C.POINTER(_StokesLevelS), C.c_double, C.c_int, _u64p]
        L.ref_vanka_copy.argtypes = [_vp, C.POINTER(_StokesLevelS), C.POINTER(C.POINTER(C.c_uint64)),
                                     C.POINTER(C.POINTER(C.c_uint64)), C.POINTER(C.POINTER(C.c_double)), _vp]
        L.ref_vanka_apply.argtypes = [_vp, C.POINTER(_StokesLevelS), _f64p, _f64p]
        L.ref_vanka_free.argtypes = [_vp]
        L.ref_saddle_apply.argtypes = [C.POINTER(_StokesLevelS), _f64p, _f64p, _vp, C.c_double]
        L.ref_stokes_filter.argtypes = [C.POINTER(_StokesLevelS), _f64p, C.c_int]
        L.ref_stokes_transfer.argtypes = [C.POINTER(_StokesLevelS), C.POINTER(_StokesLevelS), C.c_int, _f64p, _f64p]
        L.ref_stokes_smooth.argtypes = [C.POINTER(_StokesLevelS), C.c_int, C.c_double, C.c_double, _f64p, _f64p]
        L.ref_stokes_mg_solve.restype = C.c_int
        L.ref_stokes_mg_solve.argtypes = [C.POINTER(_StokesLevelS), C.c_int, C.c_int, C.c_double, C.c_double,
                                          C.c_int, C.c_double, C.c_double, C.c_uint64, _f64p, _f64p, _vp,
                                          C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_double),
                                          C.POINTER(C.c_double)]

    # -- Stokes
    def saddle_apply(self, lv, x, y=None, alpha=1.0):
        arr, keep = _stokes_levels_s([lv])
        r = np.zeros(2 * lv["nv"] + lv["np"])
        yy = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
        self.L.ref_saddle_apply(arr, r, np.ascontiguousarray(x, dtype=np.float64),
                                None if yy is None else yy.ctypes.data, alpha)
        return r

    def stokes_filter(self, lv, v, kind):
        arr, keep = _stokes_levels_s([lv])
        v = np.array(v, dtype=np.float64)
        self.L.ref_stokes_filter(arr, v, {"rhs": 0, "sol": 1, "def": 2, "cor": 3}[kind])
        return v

    def stokes_transfer(self, fine, coarse, x, prol):
        arr, keep = _stokes_levels_s([fine, coarse])
        out = coarse if not prol else fine
        r = np.zeros(2 * out["nv"] + out["np"])
        self.L.ref_stokes_transfer(C.byref(arr[0]), C.byref(arr[1]), int(prol), r,
                                     np.ascontiguousarray(x, dtype=np.float64))
        return r

    def stokes_smooth(self, lv, rhs, steps=5, omega=0.7, vanka_omega=1.0):
        arr, keep = _stokes_levels_s([lv])
        sol = np.zeros(2 * lv["nv"] + lv["np"])
        self.L.ref_stokes_smooth(arr, steps, omega, vanka_omega, sol, np.ascontiguousarray(rhs, dtype=np.float64))
        return sol

    def vanka(self, lv, omega=1.0, skip_singular=False, d=None):
        arr, keep = _stokes_levels_s([lv])
        sz = np.zeros(5, np.uint64)
        h = self.L.ref_vanka_create(arr, omega, int(skip_singular), sz)
        try:
            blocks, ptrs = _vanka_blocks(lv["nv"], lv["np"], sz[1:5])
            mask = np.zeros(max(int(sz[0]), 1), np.int32)
            self.L.ref_vanka_copy(h, arr, ptrs[0], ptrs[1], ptrs[2], mask.ctypes.data)
            out = dict(blocks=blocks, mask=mask[:int(sz[0])] if skip_singular else None)
            if d is not None:
                c = np.zeros(2 * lv["nv"] + lv["np"])
                self.L.ref_vanka_apply(h, arr, c, np.ascontiguousarray(d, dtype=np.float64))
                out["c"] = c
        finally:
            self.L.ref_vanka_free(h)
        return out

    def stokes_mg_solve(self, levels, x0, b, steps=5, omega_rich=0.7, vanka_omega=1.0,
                        coarse_skip_singular=True, omega_outer=1.0, tol_rel=1e-8, max_iter=100):
        arr, keep = _stokes_levels_s(levels)
        x = np.array(x0, dtype=np.float64)
        hist = np.full(max_iter + 1, np.nan)
        it, dfin, secs = C.c_uint64(), C.c_double(), C.c_double()
        st = self.L.ref_stokes_mg_solve(arr, len(levels), steps, omega_rich, vanka_omega,
                                        int(coarse_skip_singular), omega_outer, tol_rel, max_iter, x,
                                        np.ascontiguousarray(b, dtype=np.float64), hist.ctypes.data,
                                        hist.size, C.byref(it), C.byref(dfin), C.byref(secs))
        return dict(status=STATUS[st], iters=int(it.value), x=x, hist=hist[:it.value + 1],
                    def_final=dfin.value, seconds=secs.value)

    def stokes_asm(self, level):
        """The reference's own Q2/P1dc assembly of the unit square refined `level` times
        (DuDvOperatorBlocked<2>, GradPresDivVeloAssembler, cubature auto-degree:5): A, B, D as
        (val, col, rowptr) with row-major blocks, the mesh's vertex coordinates, edge->vertex
        and cell->vertex lists."""
        L = self.L
        L.ref_stokes_asm_create.restype = _vp
        L.ref_stokes_asm_create.argtypes = [C.c_int]
        L.ref_stokes_asm_sizes.argtypes = [_vp, _vp]
        L.ref_stokes_asm_copy.argtypes = [_vp] * 13
        L.ref_stokes_asm_free.argtypes = [_vp]
        h = L.ref_stokes_asm_create(int(level))
        z = np.zeros(8, np.uint64)
        L.ref_stokes_asm_sizes(h, z.ctypes.data)
        nv, np_, na, nb, nd, nvert, nedge, ncell = (int(v) for v in z)
        A = (np.empty(4 * na), np.empty(na, np.uint64), np.empty(nv + 1, np.uint64))
        B = (np.empty(2 * nb), np.empty(nb, np.uint64), np.empty(nv + 1, np.uint64))
        D = (np.empty(2 * nd), np.empty(nd, np.uint64), np.empty(np_ + 1, np.uint64))
        coords, edges, cells = np.empty(2 * nvert), np.empty(2 * nedge, np.uint64), np.empty(4 * ncell, np.uint64)
        L.ref_stokes_asm_copy(h, *(a.ctypes.data for a in A + B + D), coords.ctypes.data, edges.ctypes.data,
                              cells.ctypes.data)
        L.ref_stokes_asm_free(h)
        return dict(nv=nv, np=np_, A=A, B=B, D=D, coords=coords.reshape(-1, 2), edges=edges.reshape(-1, 2),
                    cells=cells.reshape(-1, 4))

    def precond_apply(self, kind, A, omega, x):
        """y = M^{-1} x of the reference's generic SORPrecond ("sor"), SSORPrecond ("ssor")
        or ILUPrecond ("ilu", level p = omega)."""
        L = self.L
        L.ref_precond_apply.argtypes = [C.c_int, C.POINTER(_CsrS), C.c_double, _vp, _vp]
        val = np.ascontiguousarray(A.val, np.float64)
        col = np.ascontiguousarray(A.col, np.uint64)
        rp = np.ascontiguousarray(A.rowptr, np.uint64)
        m = _CsrS(A.rows, A.cols, val.size, val.ctypes.data, col.ctypes.data, rp.ctypes.data)
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(A.rows)
        L.ref_precond_apply({"sor": 0, "ssor": 1, "ilu": 2}[kind], C.byref(m), omega, x.ctypes.data, y.ctypes.data)
        return y

    def csr_write(self, mode, path, A, symmetric=False):
        """SparseMatrixCSR::write_out (mode: the reference's FileMode value)."""
        val = np.ascontiguousarray(A.val, np.float64)
        col = np.ascontiguousarray(A.col, np.uint64)
        rp = np.ascontiguousarray(A.rowptr, np.uint64)
        m = _CsrS(A.rows, A.cols, val.size, val.ctypes.data, col.ctypes.data, rp.ctypes.data)
        self.L.ref_csr_write(mode, str(path).encode(), C.byref(m), int(symmetric))

    def csr_read(self, mode, path):
        """SparseMatrixCSR::read_from: Csr of host arrays."""
        rows, cols, nnz = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.L.ref_csr_read(mode, str(path).encode(), C.byref(rows), C.byref(cols), C.byref(nnz),
                            None, None, None, 0)
        cap = max(nnz.value, rows.value + 1)
        val, col, rp = np.empty(cap), np.empty(cap, np.uint64), np.empty(cap, np.uint64)
        self.L.ref_csr_read(mode, str(path).encode(), C.byref(rows), C.byref(cols), C.byref(nnz),
                            val.ctypes.data, col.ctypes.data, rp.ctypes.data, cap)
        return Csr(rows.value, cols.value, val[:nnz.value].copy(), col[:nnz.value].copy(),
                   rp[:rows.value + 1].copy())

    def dv_write(self, mode, path, v):
        v = np.ascontiguousarray(v, np.float64)
        self.L.ref_dv_write(mode, str(path).encode(), v.ctypes.data, v.size)

    def dv_read(self, mode, path):
        n = C.c_uint64()
        self.L.ref_dv_read(mode, str(path).encode(), C.byref(n), None, 0)
        v = np.empty(max(n.value, 1))
        self.L.ref_dv_read(mode, str(path).encode(), C.byref(n), v.ctypes.data, v.size)
        return v[:n.value].copy()

    def csr_apply(self, A, x, y=None, a=1.0, b=0.0, transposed=False):
        nr = A.cols if transposed else A.rows
        y = np.zeros(nr) if y is None else np.ascontiguousarray(y, np.float64)
        r = np.empty(nr)
        self.L.ref_csr_apply_f64(r, a, np.ascontiguousarray(x, np.float64), b, y, A.val, A.col,
                                 A.rowptr, A.rows, A.cols, A.nnz, int(transposed))
        return r

    def csr_apply_f32(self, A, x, y=None, a=1.0, b=0.0, transposed=False):
        """SparseMatrixCSR<float, Index> apply (arch/apply_generic.hpp:25-69 for float)."""
        nr = A.cols if transposed else A.rows
        y = np.zeros(nr, np.float32) if y is None else np.ascontiguousarray(y, np.float32)
        r = np.empty(nr, np.float32)
        self.L.ref_csr_apply_f32(r, a, np.ascontiguousarray(x, np.float32), b, y,
                                 np.ascontiguousarray(A.val, np.float32), A.col, A.rowptr,
                                 A.rows, A.cols, A.nnz, int(transposed))
        return r

    def unit_filter_blocked(self, rhs, bs, v, vals, idx, ign_nans=False):
        v = np.array(v, np.float64)
        idx = np.ascontiguousarray(idx, np.uint64)
        self.L.ref_unit_filter_blocked_f64(int(rhs), bs, v, np.ascontiguousarray(vals, np.float64), idx, idx.size,
                                           int(ign_nans))
        return v

    def component_invert_f32(self, x, s=1.0):
        x = np.ascontiguousarray(x, np.float32)
        r = np.empty_like(x)
        self.L.ref_component_invert_f32(r, x, s, x.size)
        return r

    def mg_pcg_solve_mixed(self, levels, x0, b, nu_pre=4, nu_post=4, omega_jac=0.5, omega_rich=1.0,
                           coarse=COARSE_JACOBI, coarse_omega=1.0, tol_rel=1e-8, max_iter=100):
        """PCG<double> preconditioned by ConvertPrecond around a float V-cycle of the
        reference (oracle/ref_driver.cpp ref_mg_pcg_solve_mixed); COARSE_DIRECT is the
        driver's FP64 dense-inverse coarse solver."""
        ref_coarse = {COARSE_SMOOTHER: 1, COARSE_JACOBI: 2, COARSE_DIRECT: 3}[coarse]
        arr, keep = _levels_s(levels)
        x = np.ascontiguousarray(x0, np.float64).copy()
        hist = np.full(max_iter + 2, np.nan)
        iters = C.c_uint64()
        dfin = C.c_double()
        st = self.L.ref_mg_pcg_solve_mixed(arr, len(levels), nu_pre, nu_post, omega_jac, omega_rich,
                                           ref_coarse, coarse_omega, tol_rel, max_iter, x,
                                           np.ascontiguousarray(b, np.float64), hist.ctypes.data,
                                           hist.size, C.byref(iters), C.byref(dfin))
        return dict(x=x, status=STATUS[st], iters=int(iters.value), def_final=dfin.value,
                    hist=hist[:int(iters.value) + 1])

    def bcsr_apply(self, bh, bw, rows, cols, val, col, rowptr, x, y=None, a=1.0, b=0.0):
        y = np.zeros(rows * bh) if y is None else np.ascontiguousarray(y, np.float64)
        r = np.empty(rows * bh)
        self.L.ref_bcsr_apply_f64(bh, bw, r, a, np.ascontiguousarray(x, np.float64), b, y,
                                  np.ascontiguousarray(val, np.float64),
                                  np.ascontiguousarray(col, np.uint64),
                                  np.ascontiguousarray(rowptr, np.uint64), rows, cols, col.size)
        return r

    def axpy(self, a, x, y, alias=None):
        """alias: None (r distinct), 'x' (r is x), 'y' (r is y)."""
        x = x.copy()
        y = y.copy()
        if alias == "x":
            self.L.ref_axpy_f64(x.ctypes.data, a, x.ctypes.data, y.ctypes.data, x.size)
            return x
        if alias == "y":
            self.L.ref_axpy_f64(y.ctypes.data, a, x.ctypes.data, y.ctypes.data, x.size)
            return y
        r = np.empty_like(x)
        self.L.ref_axpy_f64(r.ctypes.data, a, x.ctypes.data, y.ctypes.data, x.size)
        return r

    def dot(self, x, y):
        return self.L.ref_dot_f64(x, y, x.size)

    def triple_dot(self, x, y, z):
        return self.L.ref_triple_dot_f64(x, y, z, x.size)

    def norm2(self, x):
        return self.L.ref_norm2_f64(x, x.size)

    def scale(self, x, s):
        r = np.empty_like(x)
        self.L.ref_scale_f64(r, x, s, x.size)
        return r

    def component_product(self, x, y):
        r = np.empty_like(x)
        self.L.ref_component_product_f64(r, x, y, x.size)
        return r

    def component_invert(self, x, s=1.0):
        r = np.empty_like(x)
        self.L.ref_component_invert_f64(r, x, s, x.size)
        return r

    def diag_index(self, A):
        d = np.empty(A.rows, np.uint64)
        self.L.ref_diag_index(d, A.col, A.rowptr, A.rows)
        return d

    def invert_matrix(self, a):
        a = np.ascontiguousarray(a, np.float64).copy()
        n = a.shape[0]
        p = np.zeros(n, np.int32)
        det = self.L.ref_invert_matrix_f64(n, a.shape[1], a.reshape(-1), p)
        return a, det

    def rng_fill(self, seed, n, a, b):
        v = np.empty(n)
        self.L.ref_rng_fill_f64(seed, v, n, a, b)
        return v

    def pointstar_fd(self, m, d):
        n = C.c_uint64()
        nnz = C.c_uint64()
        self.L.ref_pointstar_fd(m, d, C.byref(n), C.byref(nnz), None, None, None, None)
        val = np.empty(nnz.value)
        col = np.empty(nnz.value, np.uint64)
        rp = np.empty(n.value + 1, np.uint64)
        q = np.empty(n.value)
        self.L.ref_pointstar_fd(m, d, C.byref(n), C.byref(nnz), val.ctypes.data, col.ctypes.data,
                                rp.ctypes.data, q.ctypes.data)
        return Csr(n.value, n.value, val, col, rp), q

    def mg_pcg_solve(self, levels, x0, b, nu_pre=4, nu_post=4, omega_jac=0.5, omega_rich=1.0,
                     coarse=COARSE_JACOBI, coarse_omega=1.0, tol_rel=1e-8, max_iter=100):
        if coarse == COARSE_DIRECT:
            raise ValueError("the reference's generic path has no dense direct coarse solver")
        ref_coarse = {COARSE_NONE: 0, COARSE_SMOOTHER: 1, COARSE_JACOBI: 2}[coarse]
        arr, keep = _levels_s(levels)
        x = np.ascontiguousarray(x0, np.float64).copy()
        hist = np.full(max_iter + 2, np.nan)
        iters = C.c_uint64()
        dfin = C.c_double()
        secs = C.c_double()
        st = self.L.ref_mg_pcg_solve(arr, len(levels), nu_pre, nu_post, omega_jac, omega_rich,
                                     ref_coarse, coarse_omega, tol_rel, max_iter, x,
                                     np.ascontiguousarray(b, np.float64), hist.ctypes.data,
                                     hist.size, C.byref(iters), C.byref(dfin), C.byref(secs))
        return dict(x=x, status=STATUS[st], iters=int(iters.value), def_final=dfin.value,
                    hist=hist[:int(iters.value) + 1], seconds=secs.value)

    def single_solve(self, kind, A, x0, b, filt=None, precond=1, omega_jac=1.0, omega_rich=1.0,
                     tol_rel=np.sqrt(np.finfo(float).eps), max_iter=100, apply=False):
        keep = []
        a = _csr_s(A, keep)
        f = np.ascontiguousarray(filt if filt is not None else np.zeros(0), np.uint64)
        x = np.ascontiguousarray(x0, np.float64).copy()
        hist = np.full(max_iter + 2, np.nan)
        iters = C.c_uint64()
        dfin = C.c_double()
        st = self.L.ref_single_solve({"pcg": 0, "richardson": 1}[kind], C.byref(a),
                                     f.ctypes.data if f.size else None, f.size, precond,
                                     omega_jac, omega_rich, tol_rel, max_iter, int(apply), x,
                                     np.ascontiguousarray(b, np.float64), hist.ctypes.data,
                                     hist.size, C.byref(iters), C.byref(dfin))
        return dict(x=x, status=STATUS[st], iters=int(iters.value), def_final=dfin.value,
                    hist=hist[:int(iters.value) + 1])

    def q1_poisson(self, dim, level, with_prol=True, cubature="gauss-legendre:2"):
        """FE-assembled Q1 Laplace matrix, unit filter, vertex coordinates, sine-bubble
        load vector and (optionally) the prolongation from level-1, all in the
        reference's own DOF numbering."""
        h = self.L.ref_q1_create(dim, level, int(with_prol), cubature.encode())
        try:
            s = np.zeros(5, np.uint64)
            self.L.ref_q1_sizes(h, s)
            n, nnz, nf, pn, pnnz = (int(v) for v in s)
            val = np.empty(nnz)
            col = np.empty(nnz, np.uint64)
            rp = np.empty(n + 1, np.uint64)
            filt = np.empty(nf, np.uint64)
            coords = np.empty(n * dim)
            rhs = np.empty(n)
            pval = np.empty(max(pnnz, 1))
            pcol = np.empty(max(pnnz, 1), np.uint64)
            prp = np.empty(n + 1, np.uint64)
            self.L.ref_q1_copy(h, val, col, rp, filt, coords, rhs, pval.ctypes.data,
                               pcol.ctypes.data, prp.ctypes.data)
        finally:
            self.L.ref_q1_free(h)
        out = dict(A=Csr(n, n, val, col, rp), filt=filt, coords=coords.reshape(n, dim), rhs=rhs)
        if pnnz:
            out["P"] = Csr(n, pn, pval[:pnnz], pcol[:pnnz], prp)
        return out


def csr_transpose(A):
    """Explicit CSR transpose with column-sorted rows (SparseMatrixCSR::transpose)."""
    counts = np.bincount(A.col.astype(np.int64), minlength=A.cols)
    rp = np.zeros(A.cols + 1, np.uint64)
    rp[1:] = np.cumsum(counts)
    rows = np.repeat(np.arange(A.rows, dtype=np.uint64), np.diff(A.rowptr.astype(np.int64)))
    # stable sort by column keeps the row order inside each new row
    order = np.argsort(A.col, kind="stable")
    return Csr(A.cols, A.rows, A.val[order].copy(), rows[order].copy(), rp)
