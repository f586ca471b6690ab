// feat3amd/stokes.hpp -- the Stokes part of the path on the device: the
// SaddlePointMatrix, the system filter TupleFilter<UnitFilterBlocked<2>,
// MeanFilter>, the system transfer (BWrappedCSR velocity, CSR pressure), the
// AmaVanka smoother and the multigrid solver of the Stokes benchmark
// (benchmarks/stokes_multigrid-bench.cpp:714-810):
//   Richardson(omega_outer) preconditioned by MG-V with
//   Richardson(omega_rich, SchwarzPrecond(AmaVanka)) smoothers (min = max steps)
//   and SchwarzPrecond(AmaVanka with skip_singular) as coarse solver.
// A Stokes vector is one DenseVector laid out [u (2 per node) | p].
// The AmaVanka macro graphs and the Vanka pattern are built on the host
// (init_symbolic, as the reference does); the numeric factorisation and every
// solve-time operation run in HIP kernels.
#pragma once

#include <feat3amd/solver.hpp>
#include <feat3amd/stokes_assembly.hpp>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <map>
#include <vector>

namespace feat3amd
{
  namespace LAFEM
  {
    typedef DenseVector<double, Index> StokesVector;

    /// host -> device copy of a block-CSR matrix
    struct BcsrDev
    {
      Index rows = 0, cols = 0;
      int bh = 1, bw = 1;
      DeviceArray<double> val;
      DeviceArray<std::uint64_t> col, ptr;
      static BcsrDev from_host(const Assembly::BcsrHost& m)
      {
        BcsrDev d;
        d.rows = m.rows;
        d.cols = m.cols;
        d.bh = m.bh;
        d.bw = m.bw;
        d.val = DeviceArray<double>::from_host(m.val.data(), m.val.size());
        d.col = DeviceArray<std::uint64_t>::from_host(m.col.data(), m.col.size());
        d.ptr = DeviceArray<std::uint64_t>::from_host(m.rowptr.data(), m.rowptr.size());
        return d;
      }
      static BcsrDev adopt(Index rows, Index cols, int bh, int bw, DeviceArray<double> val,
                           DeviceArray<std::uint64_t> col, DeviceArray<std::uint64_t> ptr)
      {
        BcsrDev d;
        d.rows = rows;
        d.cols = cols;
        d.bh = bh;
        d.bw = bw;
        d.val = val;
        d.col = col;
        d.ptr = ptr;
        return d;
      }
    };

    /// SaddlePointMatrix<BCSR<2,2>, BCSR<2,1>, BCSR<1,2>> (lafem/saddle_point_matrix.hpp)
    class SaddlePointMatrix
    {
      Index _nv = 0, _np = 0;
      BcsrDev _a, _b, _d;

    public:
      SaddlePointMatrix() = default;
      static SaddlePointMatrix from_host(const Assembly::StokesLevelHost& S)
      {
        SaddlePointMatrix m;
        m._nv = S.nv;
        m._np = S.np;
        m._a = BcsrDev::from_host(S.A);
        m._b = BcsrDev::from_host(S.B);
        m._d = BcsrDev::from_host(S.D);
        return m;
      }
      /// from blocks already on the device (feat_stokes_assemble_device)
      static SaddlePointMatrix from_device(Index nv, Index np, BcsrDev a, BcsrDev b, BcsrDev d)
      {
        SaddlePointMatrix m;
        m._nv = nv;
        m._np = np;
        m._a = a;
        m._b = b;
        m._d = d;
        return m;
      }
      Index nv() const { return _nv; }
      Index np() const { return _np; }
      Index size() const { return 2 * _nv + _np; }
      StokesVector create_vector() const { return StokesVector(size()); }
      feat_tuple_f64_t desc() const
      {
        feat_tuple_f64_t t{};
        t.nv = _nv;
        t.np = _np;
        const BcsrDev* b[3] = {&_a, &_b, &_d};
        for(int k = 0; k < 3; ++k)
        {
          t.val[k] = b[k]->val.get();
          t.col[k] = b[k]->col.get();
          t.ptr[k] = b[k]->ptr.get();
        }
        return t;
      }
      /// r = K x (saddle_point_matrix.hpp:446-451)
      void apply(StokesVector& r, const StokesVector& x) const
      {
        const feat_tuple_f64_t t = desc();
        check(feat_tuple_apply_f64(Runtime::handle(), &t, r.elements(), x.elements(), nullptr, 1.0, nullptr),
              "SaddlePointMatrix::apply");
      }
      /// r = y + alpha K x (saddle_point_matrix.hpp:486-491); vmask zeroes velocity rows after
      void apply(StokesVector& r, const StokesVector& x, const StokesVector& y, double alpha,
                 const std::uint32_t* vmask = nullptr) const
      {
        const feat_tuple_f64_t t = desc();
        check(feat_tuple_apply_f64(Runtime::handle(), &t, r.elements(), x.elements(), y.elements(), alpha, vmask),
              "SaddlePointMatrix::apply");
      }
    };

    /// One launch (two above 4,096 pressure rows): r = y + alpha M x (r = M x without y), velocity
    /// rows zeroed by vmask, then
    /// `passes` MeanFilter passes on r's pressure part (weights w_dot, w_add, volume), then
    /// x_upd += omega r gated by *gate (x_upd null: no update) -- bit-identical to
    /// feat_tuple_apply_f64, feat_mean_filter_passes_f64 and feat_axpy_gated_f64 in a row.
    /// (x_zero: x_upd holds zeros and is not read)
    inline void tuple_apply_post(const feat_tuple_f64_t& M, StokesVector& r, const StokesVector& x,
                                 const StokesVector* y, double alpha, const std::uint32_t* vmask, const double* w_dot,
                                 const double* w_add, double volume, int passes, double* x_upd, double omega,
                                 const double* gate, const feat_reduce_ws_t& ws, bool x_zero = false)
    {
      check(feat_tuple_apply_post_f64(Runtime::handle(), &M, r.elements(), x.elements(), y ? y->elements() : nullptr,
                                      alpha, vmask, w_dot, w_add, volume, passes, 0, x_upd, omega, gate,
                                      x_zero ? 1 : 0, &ws),
            "tuple apply + filter + update");
    }

    /// TupleFilter<UnitFilterBlocked<DT,IT,2>, MeanFilter<DT,IT>> on [u | p]
    /// (lafem/unit_filter_blocked.hpp:449-499, lafem/mean_filter.hpp:200-262)
    class StokesFilter
    {
      Index _nv = 0, _np = 0, _nf = 0;
      DeviceArray<std::uint64_t> _nodes;        // blocked node indices
      DeviceArray<double> _sval;                // two values per filtered node
      DeviceArray<std::uint32_t> _vmask;        // bit per velocity node
      DeviceArray<double> _prim, _dual;
      double _vol = 0.0;

    public:
      StokesFilter() = default;
      static StokesFilter from_host(const Assembly::StokesLevelHost& S)
      {
        StokesFilter f;
        f._nv = S.nv;
        f._np = S.np;
        f._nf = S.filt.size();
        f._nodes = DeviceArray<std::uint64_t>::from_host(S.filt.data(), f._nf);
        f._sval = DeviceArray<double>::from_host(S.filt_val.data(), S.filt_val.size());
        f._vmask = DeviceArray<std::uint32_t>((S.nv + 31) / 32);
        check(feat_unit_filter_mask_u64(Runtime::handle(), f._vmask.get(), S.nv, f._nodes.get(), f._nf), "StokesFilter mask");
        f._prim = DeviceArray<double>::from_host(S.prim.data(), S.np);
        f._dual = DeviceArray<double>::from_host(S.dual.data(), S.np);
        // MeanFilter volume = prim . dual (mean_filter.hpp:73), the generic sequential dot
        double v = 0.0;
        for(Index i = 0; i < S.np; ++i)
        {
          const double t = S.prim[i] * S.dual[i];
          v = v + t;
        }
        f._vol = v;
        return f;
      }
      /// from arrays already on the device (feat_stokes_assemble_device); prim and dual are
      /// read back once for the volume, summed as from_host sums it
      static StokesFilter from_device(Index nv, Index np, DeviceArray<std::uint64_t> nodes, DeviceArray<double> sval,
                                      DeviceArray<double> prim, DeviceArray<double> dual)
      {
        StokesFilter f;
        f._nv = nv;
        f._np = np;
        f._nf = nodes.size();
        f._nodes = nodes;
        f._sval = sval;
        f._vmask = DeviceArray<std::uint32_t>((nv + 31) / 32);
        check(feat_unit_filter_mask_u64(Runtime::handle(), f._vmask.get(), nv, f._nodes.get(), f._nf), "StokesFilter mask");
        f._prim = prim;
        f._dual = dual;
        std::vector<double> hp(np), hd(np);
        prim.to_host(hp.data());
        dual.to_host(hd.data());
        double v = 0.0;
        for(Index i = 0; i < np; ++i)
        {
          const double t = hp[i] * hd[i];
          v = v + t;
        }
        f._vol = v;
        return f;
      }
      const std::uint32_t* vmask() const { return _vmask.get(); }
      const double* prim() const { return _prim.get(); }
      const double* dual() const { return _dual.get(); }
      double volume() const { return _vol; }
      Index np() const { return _np; }
      double* pres(StokesVector& v) const { return v.elements() + 2 * _nv; }

      void unit_rhs(StokesVector& v) const
      {
        if(_nf)
          check(feat_unit_filter_blocked_rhs_f64_u64(Runtime::handle(), 2, v.elements(), _sval.get(), _nodes.get(), _nf, 0),
                "filter_rhs");
      }
      void unit_def(StokesVector& v) const
      {
        if(_nf)
          check(feat_unit_filter_blocked_def_f64_u64(Runtime::handle(), 2, v.elements(), nullptr, _nodes.get(), _nf, 0),
                "filter_def");
      }
      /// mean part: def/rhs subtract dual-weighted, cor/sol prim-weighted (passes: repeated calls)
      void mean_def(StokesVector& v) const
      {
        check(feat_mean_filter_passes_f64(Runtime::handle(), pres(v), _prim.get(), _dual.get(), _vol, _np, 1, 0), "mean filter_def");
      }
      void mean_cor(StokesVector& v, int passes = 1) const
      {
        check(feat_mean_filter_passes_f64(Runtime::handle(), pres(v), _dual.get(), _prim.get(), _vol, _np, passes, 0), "mean filter_cor");
      }
      void filter_rhs(StokesVector& v) const
      {
        unit_rhs(v);
        mean_def(v);
      }
      void filter_sol(StokesVector& v) const
      {
        unit_rhs(v);
        check(feat_mean_filter_passes_f64(Runtime::handle(), pres(v), _dual.get(), _prim.get(), _vol, _np, 1, 1), "mean filter_sol");
      }
      void filter_def(StokesVector& v) const
      {
        unit_def(v);
        mean_def(v);
      }
      void filter_cor(StokesVector& v) const
      {
        unit_def(v);
        mean_cor(v);
      }
    };

    /// TupleDiagMatrix<BWrappedCSR<2>, CSR> transfer (control/stokes_blocked.hpp:193-202)
    class StokesTransfer
    {
      struct Csr
      {
        Index rows = 0, cols = 0, nnz = 0;
        DeviceArray<double> val;
        DeviceArray<std::uint64_t> col, ptr;
        static Csr from_host(const Assembly::CsrHost& m)
        {
          Csr c;
          c.rows = m.rows;
          c.cols = m.cols;
          c.nnz = m.val.size();
          c.val = DeviceArray<double>::from_host(m.val.data(), m.val.size());
          c.col = DeviceArray<std::uint64_t>::from_host(m.col.data(), m.col.size());
          c.ptr = DeviceArray<std::uint64_t>::from_host(m.rowptr.data(), m.rowptr.size());
          return c;
        }
      };
      Csr _pv, _rv, _pp, _rp;
      Index _nv_f = 0, _nv_c = 0;

      static void apply(const Csr& mv, const Csr& mp, double* r, Index r_nv, const double* x, Index x_nv)
      {
        feat_handle_t h = Runtime::handle();
        check(feat_csrsb2_apply_f64_u64(h, r, x, mv.val.get(), mv.col.get(), mv.ptr.get(), mv.rows), "velocity transfer");
        check(feat_csr_apply_f64_u64(h, r + 2 * r_nv, 1.0, x + 2 * x_nv, 0.0, r + 2 * r_nv, mp.val.get(), mp.col.get(),
                                     mp.ptr.get(), mp.rows, mp.cols, mp.nnz, 0),
              "pressure transfer");
      }

    public:
      StokesTransfer() = default;
      /// fine: the level whose Pv/Pp prolongate from the next coarser level
      static StokesTransfer from_host(const Assembly::StokesLevelHost& fine)
      {
        StokesTransfer t;
        t._pv = Csr::from_host(fine.Pv);
        t._rv = Csr::from_host(Assembly::transpose(fine.Pv));
        t._pp = Csr::from_host(fine.Pp);
        t._rp = Csr::from_host(Assembly::transpose(fine.Pp));
        t._nv_f = fine.Pv.rows;
        t._nv_c = fine.Pv.cols;
        return t;
      }
      /// from the prolongations and restrictions already on the device (feat_stokes_assemble_device)
      static StokesTransfer from_device(Index nv_f, Index np_f, Index nv_c, Index np_c, std::uint64_t nnz_v,
                                        std::uint64_t nnz_p, DeviceArray<double> pv_val,
                                        DeviceArray<std::uint64_t> pv_col, DeviceArray<std::uint64_t> pv_ptr,
                                        DeviceArray<double> pp_val, DeviceArray<std::uint64_t> pp_col,
                                        DeviceArray<std::uint64_t> pp_ptr, DeviceArray<double> rv_val,
                                        DeviceArray<std::uint64_t> rv_col, DeviceArray<std::uint64_t> rv_ptr,
                                        DeviceArray<double> rp_val, DeviceArray<std::uint64_t> rp_col,
                                        DeviceArray<std::uint64_t> rp_ptr)
      {
        auto mk = [](Index rows, Index cols, std::uint64_t nnz, DeviceArray<double> v, DeviceArray<std::uint64_t> c,
                     DeviceArray<std::uint64_t> p) {
          Csr m;
          m.rows = rows;
          m.cols = cols;
          m.nnz = nnz;
          m.val = v;
          m.col = c;
          m.ptr = p;
          return m;
        };
        StokesTransfer t;
        t._pv = mk(nv_f, nv_c, nnz_v, pv_val, pv_col, pv_ptr);
        t._pp = mk(np_f, np_c, nnz_p, pp_val, pp_col, pp_ptr);
        t._rv = mk(nv_c, nv_f, nnz_v, rv_val, rv_col, rv_ptr);
        t._rp = mk(np_c, np_f, nnz_p, rp_val, rp_col, rp_ptr);
        t._nv_f = nv_f;
        t._nv_c = nv_c;
        return t;
      }
      /// Transfer::prol (lafem/transfer.hpp:255-259): fine = P coarse
      void prol(StokesVector& fine, const StokesVector& coarse) const
      {
        apply(_pv, _pp, fine.elements(), _nv_f, coarse.elements(), _nv_c);
      }
      /// Transfer::rest (lafem/transfer.hpp:241-245): coarse = R fine
      void rest(const StokesVector& fine, StokesVector& coarse) const
      {
        apply(_rv, _rp, coarse.elements(), _nv_c, fine.elements(), _nv_f);
      }
      /// coarse = filter_def(R fine) in one launch (filter: the coarse level's)
      void rest_filtered(const StokesVector& fine, StokesVector& coarse, const StokesFilter& cf,
                         const feat_reduce_ws_t& ws) const
      {
        const feat_csr_f64_t v = desc(0), p = desc(1);
        check(feat_stokes_transfer_post_f64(Runtime::handle(), &v, &p, fine.elements(), _nv_f, coarse.elements(),
                                            cf.vmask(), cf.prim(), cf.dual(), cf.volume(), 1, nullptr, &ws),
              "restriction + filter_def");
      }
      /// cor = filter_cor(P coarse), then sol += cor, in one launch (filter: the fine level's)
      void prol_filtered_add(StokesVector& cor, const StokesVector& coarse, const StokesFilter& ff, StokesVector& sol,
                             const feat_reduce_ws_t& ws) const
      {
        const feat_csr_f64_t v = desc(2), p = desc(3);
        check(feat_stokes_transfer_post_f64(Runtime::handle(), &v, &p, coarse.elements(), _nv_c, cor.elements(),
                                            ff.vmask(), ff.dual(), ff.prim(), ff.volume(), 1, sol.elements(), &ws),
              "prolongation + filter_cor + axpy");
      }
      /// device descriptors: 0 = Rv, 1 = Rp, 2 = Pv, 3 = Pp
      feat_csr_f64_t desc(int k) const
      {
        const Csr* m[4] = {&_rv, &_rp, &_pv, &_pp};
        const Csr& c = *m[k];
        return feat_csr_f64_t{c.rows, c.cols, c.nnz, c.val.get(), c.col.get(), c.ptr.get()};
      }
    };
  } // namespace LAFEM

  namespace Solver
  {
    /// AmaVanka (solver/amavanka.hpp:1190-1590) for the saddle-point matrix of one level.
    class AmaVanka
    {
      const LAFEM::SaddlePointMatrix& _matrix;
      const LAFEM::StokesFilter& _filter;
      double _omega;
      bool _skip_singular = false;
      // macro graphs (host build, device copy)
      std::vector<std::uint64_t> _vptr, _vidx, _pptr, _pidx, _dvptr, _dvidx, _dpptr, _dpidx;
      DeviceArray<std::uint64_t> _g[8];
      Index _nm = 0, _stride = 0;
      // Vanka matrix blocks (uu 2x2, up 2x1, pu 1x2, pp 1x1)
      LAFEM::BcsrDev _v[4];
      DeviceArray<double> _local;
      DeviceArray<int> _mask;

      static void transpose_graph(Index nm, const std::vector<std::uint64_t>& ptr, const std::vector<std::uint64_t>& idx,
                                  Index ndof, std::vector<std::uint64_t>& tptr, std::vector<std::uint64_t>& tidx)
      {
        tptr.assign(ndof + 1, 0);
        for(auto i : idx) ++tptr[i + 1];
        for(Index i = 0; i < ndof; ++i) tptr[i + 1] += tptr[i];
        tidx.resize(idx.size());
        std::vector<std::uint64_t> fill(tptr.begin(), tptr.end() - 1);
        for(Index m = 0; m < nm; ++m)
          for(auto k = ptr[m]; k < ptr[m + 1]; ++k) tidx[fill[idx[k]]++] = m;
      }

      // injectify_sorted(dof_macros, macro_dofs) pattern (amavanka.hpp:663-673)
      static Assembly::BcsrHost alloc(Index nrows, const std::vector<std::uint64_t>& dptr,
                                      const std::vector<std::uint64_t>& didx, const std::vector<std::uint64_t>& mptr,
                                      const std::vector<std::uint64_t>& midx, Index ncols, int bh, int bw)
      {
        Assembly::BcsrHost M;
        M.rows = nrows;
        M.cols = ncols;
        M.bh = bh;
        M.bw = bw;
        M.rowptr.assign(nrows + 1, 0);
        std::vector<std::uint64_t> buf;
        for(Index r = 0; r < nrows; ++r)
        {
          buf.clear();
          for(auto k = dptr[r]; k < dptr[r + 1]; ++k)
            for(auto j = mptr[didx[k]]; j < mptr[didx[k] + 1]; ++j) buf.push_back(midx[j]);
          std::sort(buf.begin(), buf.end());
          buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
          M.col.insert(M.col.end(), buf.begin(), buf.end());
          M.rowptr[r + 1] = M.col.size();
        }
        M.val.assign(M.col.size() * std::size_t(bh * bw), 0.0);
        return M;
      }

    public:
      AmaVanka(const LAFEM::SaddlePointMatrix& m, const LAFEM::StokesFilter& f, double omega = 1.0)
        : _matrix(m), _filter(f), _omega(omega) {}
      void set_skip_singular(bool s) { _skip_singular = s; }
      Index num_macros() const { return _nm; }
      Index stride() const { return _stride; }

      /// deduct_macro_dofs (amavanka.hpp:224-350) from the host pattern of B and D,
      /// transposed graphs and the Vanka pattern
      void init_symbolic(const Assembly::StokesLevelHost& S)
      {
        const Index np = S.np, nv = S.nv;
        const auto &rb = S.B.rowptr, &cb = S.B.col, &rd = S.D.rowptr, &cd = S.D.col;
        std::vector<int> mask(np, 0);
        _pptr.assign(1, 0);
        _pidx.clear();
        std::map<std::uint64_t, int> cnt;
        for(Index i = 0; i < np; ++i)
        {
          if(mask[i]) continue;
          cnt.clear();
          for(auto kd = rd[i]; kd < rd[i + 1]; ++kd)
            for(auto kb = rb[cd[kd]]; kb < rb[cd[kd] + 1]; ++kb) ++cnt[cb[kb]];
          int deg = 0;
          for(auto& e : cnt) deg = std::max(deg, e.second);
          for(auto& e : cnt)
            if(e.second == deg)
            {
              _pidx.push_back(e.first);
              mask[e.first] = 1;
            }
          _pptr.push_back(_pidx.size());
        }
        _nm = _pptr.size() - 1;
        _vptr.assign(1, 0);
        _vidx.clear();
        std::vector<std::uint64_t> buf;
        for(Index m = 0; m < _nm; ++m)
        {
          buf.clear();
          for(auto j = _pptr[m]; j < _pptr[m + 1]; ++j)
            for(auto k = rd[_pidx[j]]; k < rd[_pidx[j] + 1]; ++k) buf.push_back(cd[k]);
          std::sort(buf.begin(), buf.end());
          buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
          _vidx.insert(_vidx.end(), buf.begin(), buf.end());
          _vptr.push_back(_vidx.size());
        }
        transpose_graph(_nm, _vptr, _vidx, nv, _dvptr, _dvidx);
        transpose_graph(_nm, _pptr, _pidx, np, _dpptr, _dpidx);
        _stride = 0;
        for(Index m = 0; m < _nm; ++m)
          _stride = std::max<Index>(_stride, 2 * (_vptr[m + 1] - _vptr[m]) + (_pptr[m + 1] - _pptr[m]));
        const std::vector<std::uint64_t>* g[8] = {&_vptr, &_vidx, &_pptr, &_pidx, &_dvptr, &_dvidx, &_dpptr, &_dpidx};
        for(int k = 0; k < 8; ++k) _g[k] = DeviceArray<std::uint64_t>::from_host(g[k]->data(), g[k]->size());
        _v[0] = LAFEM::BcsrDev::from_host(alloc(nv, _dvptr, _dvidx, _vptr, _vidx, nv, 2, 2));
        _v[1] = LAFEM::BcsrDev::from_host(alloc(nv, _dvptr, _dvidx, _pptr, _pidx, np, 2, 1));
        _v[2] = LAFEM::BcsrDev::from_host(alloc(np, _dpptr, _dpidx, _vptr, _vidx, nv, 1, 2));
        _v[3] = LAFEM::BcsrDev::from_host(alloc(np, _dpptr, _dpidx, _pptr, _pidx, np, 1, 1));
        _local = DeviceArray<double>(_nm * _stride * _stride);
        _mask = DeviceArray<int>(_nm);
      }

      feat_tuple_f64_t vanka_desc() const
      {
        feat_tuple_f64_t t{};
        t.nv = _matrix.nv();
        t.np = _matrix.np();
        for(int k = 0; k < 4; ++k)
        {
          t.val[k] = _v[k].val.get();
          t.col[k] = _v[k].col.get();
          t.ptr[k] = _v[k].ptr.get();
        }
        return t;
      }

      /// init_numeric (amavanka.hpp:1478-1559) on the device
      void init_numeric()
      {
        feat_vanka_graph_t g{_nm, _matrix.nv(), _matrix.np(), _stride, _g[0].get(), _g[1].get(), _g[2].get(),
                             _g[3].get(), _g[4].get(), _g[5].get(), _g[6].get(), _g[7].get()};
        const feat_tuple_f64_t K = _matrix.desc(), V = vanka_desc();
        check(feat_vanka_numeric_f64(Runtime::handle(), &g, &K, _filter.vmask(), _omega, _skip_singular ? 1 : 0,
                                     _local.get(), _mask.get(), &V),
              "AmaVanka::init_numeric");
      }

      /// c = V d, then `cor_passes` filter_cor calls (1: AmaVanka::apply, amavanka.hpp:1562-1567;
      /// 2: SchwarzPrecond after it, schwarz_precond.hpp:155-158)
      void apply(LAFEM::StokesVector& c, const LAFEM::StokesVector& d, int cor_passes) const
      {
        const feat_tuple_f64_t V = vanka_desc();
        check(feat_tuple_apply_f64(Runtime::handle(), &V, c.elements(), d.elements(), nullptr, 1.0, _filter.vmask()),
              "AmaVanka::apply");
        _filter.mean_cor(c, cor_passes);
      }

      /// apply(c, d, cor_passes) followed by x_upd += omega c gated by *gate (the Richardson
      /// step, richardson.hpp:203-211), in one or two launches (x_upd null: no update)
      void apply_update(LAFEM::StokesVector& c, const LAFEM::StokesVector& d, int cor_passes, double* x_upd,
                        double omega, const double* gate, const feat_reduce_ws_t& ws, bool x_zero = false) const
      {
        LAFEM::tuple_apply_post(vanka_desc(), c, d, nullptr, 1.0, _filter.vmask(), _filter.dual(), _filter.prim(),
                                _filter.volume(), cor_passes, x_upd, omega, gate, ws, x_zero);
      }

      /// host copy of Vanka block k (0: uu, 1: up, 2: pu, 3: pp) and the macro mask
      void copy_block(int k, std::uint64_t* ptr, std::uint64_t* col, double* val) const
      {
        if(ptr) _v[k].ptr.to_host(ptr);
        if(col) _v[k].col.to_host(col);
        if(val) _v[k].val.to_host(val);
      }
      Index block_nnz(int k) const { return _v[k].col.size(); }
      void copy_mask(int* mask) const { _mask.to_host(mask); }
    };

    struct StokesConfig
    {
      int steps = 5;                   // smoothing steps (stokes_multigrid-bench.cpp:380)
      double omega_rich = 0.7;         // smoother damping (:381)
      double vanka_omega = 1.0;
      bool coarse_skip_singular = true;
      double omega_outer = 1.0;        // outer Richardson (:761)
      bool graphs = true;              // replay each outer iteration as a recorded HIP graph
      bool coarse_fusion = true;       // the coarsest levels in one workgroup (feat_stokes_tail_f64)
      // largest level in that workgroup: it applies a level-2 tuple (210 unknowns) in 4.7 us
      // against ~9 us for a launch of its own, but a level-3 tuple (770) in 12.5 us against 9.5 us
      std::uint64_t tail_max_level_dofs = 512;
      bool device_assembly = true;     // levels from feat_stokes_assemble_device (else the host assembly)
    };

    /// Richardson(MG-V(Richardson(SchwarzPrecond(AmaVanka)))) on a Q2/P1dc hierarchy.
    class StokesMultiGrid : public IterativeSolver<LAFEM::StokesVector>
    {
    public:
      struct Level
      {
        Assembly::StokesLevelHost host;
        LAFEM::SaddlePointMatrix K;
        LAFEM::StokesFilter F;
        LAFEM::StokesTransfer T;  // to the next coarser level
        std::unique_ptr<AmaVanka> V;
        LAFEM::StokesVector rhs, sol, def, cor, rdef, rcor;
        DeviceArray<double> gate;  // smoother's initial defect norm
      };

    private:
      std::vector<std::unique_ptr<Level>> _lv;
      StokesConfig _cfg;
      LAFEM::StokesVector _d, _c;
      DeviceArray<double> _scal;
      FusedWorkspace _fw;
      // levels _tail_start.. run in one workgroup (== number of levels: none do)
      std::size_t _tail_start = 0;
      feat_stokes_tail_t _tail{};
      // one outer iteration recorded as a HIP graph on a solver-owned stream
      struct Graph
      {
        feat_stream_t stream = nullptr;
        feat_handle_t h = nullptr;
        double* host = nullptr;
        feat_graph_t g = nullptr;
        const void* key_x = nullptr;
        const void* key_b = nullptr;
        bool failed = false;
        void drop()
        {
          feat_graph_destroy(g);
          g = nullptr;
        }
        ~Graph()
        {
          drop();
          if(h) feat_handle_destroy(h);
          if(stream) feat_stream_destroy(stream);
          if(host) feat_host_free(host);
        }
      };
      std::unique_ptr<Graph> _graph;

      Index n(const Level& L) const { return L.K.size(); }

      // Richardson::apply (richardson.hpp:160-226), min_iter = max_iter = steps; the last
      // step's defect is never read by the caller (the norm is skipped), so it is not formed
      void smooth(Level& L, LAFEM::StokesVector& sol, const LAFEM::StokesVector& rhs)
      {
        feat_handle_t h = Runtime::handle();
        // sol.format() is left to the first update, which writes omega c (or zeros) into sol
        check(feat_tuple_norm2_f64(h, rhs.elements(), 2 * L.K.nv(), L.K.np(), L.gate.get(), &_fw.ws), "smoother norm");
        // each step: c = filter_cor(V d) with sol += omega c, then d = filter_def(rhs - K sol),
        // one launch each on small levels
        const feat_tuple_f64_t K = L.K.desc();
        for(int s = 0; s < _cfg.steps; ++s)
        {
          const LAFEM::StokesVector& d = (s == 0) ? rhs : L.rdef;
          L.V->apply_update(L.rcor, d, 2, sol.elements(), _cfg.omega_rich, L.gate.get(), _fw.ws, s == 0);
          if(s + 1 < _cfg.steps)
            LAFEM::tuple_apply_post(K, L.rdef, sol, &rhs, -1.0, L.F.vmask(), L.F.prim(), L.F.dual(), L.F.volume(), 1,
                                    nullptr, 0.0, nullptr, _fw.ws);
        }
      }

      void defect(Level& L, LAFEM::StokesVector& d, const LAFEM::StokesVector& x, const LAFEM::StokesVector& b)
      {
        LAFEM::tuple_apply_post(L.K.desc(), d, x, &b, -1.0, L.F.vmask(), L.F.prim(), L.F.dual(), L.F.volume(), 1, nullptr,
                                0.0, nullptr, _fw.ws);
      }

      // MultiGrid::apply with the V-cycle (multigrid.hpp:1270-1350, 1526-1570, 1713-1980)
      void mg_apply(LAFEM::StokesVector& cor, const LAFEM::StokesVector& def)
      {
        const std::size_t nl = _lv.size();
        _lv[0]->rhs.copy(def);
        for(std::size_t i = 0; i + 1 < nl && i < _tail_start; ++i)
        {
          Level& L = *_lv[i];
          Level& C = *_lv[i + 1];
          smooth(L, L.sol, L.rhs);
          defect(L, L.def, L.sol, L.rhs);
          L.T.rest_filtered(L.def, C.rhs, C.F, _fw.ws);
        }
        const std::size_t t = _tail_start;
        if(t < nl)
        {
          // levels t.. (down, coarse solve, up) in one workgroup: rhs of level t in, its sol out
          check(feat_stokes_tail_f64(Runtime::handle(), &_tail, _lv[t]->rhs.elements(), _lv[t]->sol.elements()),
                "Stokes coarse levels");
        }
        else
        {
          Level& Lc = *_lv[nl - 1];
          Lc.V->apply_update(Lc.sol, Lc.rhs, 2, nullptr, 0.0, nullptr, _fw.ws);
        }
        for(std::size_t i = std::min(t, nl - 1); i-- > 0;)
        {
          Level& L = *_lv[i];
          L.T.prol_filtered_add(L.cor, _lv[i + 1]->sol, L.F, L.sol, _fw.ws);
          defect(L, L.def, L.sol, L.rhs);
          smooth(L, L.cor, L.def);
          L.sol.axpy(L.cor, L.sol, 1.0);
        }
        cor.copy(_lv[0]->sol);
      }

      // One level from feat_stokes_assemble_device: matrices, filter and transfers stay on
      // the device; the host keeps the B and D patterns that AmaVanka's symbolic phase walks.
      void _assemble_on_device(Level& L, int l, bool transfer)
      {
        std::uint64_t z[10];
        check(feat_stokes_device_sizes(l, z), "feat_stokes_device_sizes");
        const Index nv = z[0], np = z[1], na = z[2], nb = z[3], nd = z[4], nf = z[5], npv = z[6], npp = z[7];
        const Index nvc = z[8], npc = z[9];
        DeviceArray<double> av(4 * na), bv(2 * nb), dv(2 * nd), fv(2 * nf), pr(np), du(np);
        DeviceArray<std::uint64_t> ac(na), ap(nv + 1), bc(nb), bp(nv + 1), dc(nd), dp(np + 1), fn(nf);
        DeviceArray<double> pvv, ppv, rvv, rpv;
        DeviceArray<std::uint64_t> pvc, pvp, ppc, ppp, rvc, rvp, rpc, rpp;
        feat_stokes_level_dev_t o{};
        o.a_val = av.get();
        o.a_col = ac.get();
        o.a_ptr = ap.get();
        o.b_val = bv.get();
        o.b_col = bc.get();
        o.b_ptr = bp.get();
        o.d_val = dv.get();
        o.d_col = dc.get();
        o.d_ptr = dp.get();
        o.filt = fn.get();
        o.filt_val = fv.get();
        o.prim = pr.get();
        o.dual = du.get();
        if(transfer)
        {
          pvv = DeviceArray<double>(npv);
          pvc = DeviceArray<std::uint64_t>(npv);
          pvp = DeviceArray<std::uint64_t>(nv + 1);
          ppv = DeviceArray<double>(npp);
          ppc = DeviceArray<std::uint64_t>(npp);
          ppp = DeviceArray<std::uint64_t>(np + 1);
          rvv = DeviceArray<double>(npv);
          rvc = DeviceArray<std::uint64_t>(npv);
          rvp = DeviceArray<std::uint64_t>(nvc + 1);
          rpv = DeviceArray<double>(npp);
          rpc = DeviceArray<std::uint64_t>(npp);
          rpp = DeviceArray<std::uint64_t>(npc + 1);
          o.pv_val = pvv.get();
          o.pv_col = pvc.get();
          o.pv_ptr = pvp.get();
          o.pp_val = ppv.get();
          o.pp_col = ppc.get();
          o.pp_ptr = ppp.get();
          o.rv_val = rvv.get();
          o.rv_col = rvc.get();
          o.rv_ptr = rvp.get();
          o.rp_val = rpv.get();
          o.rp_col = rpc.get();
          o.rp_ptr = rpp.get();
        }
        check(feat_stokes_assemble_device(Runtime::handle(), l, &o), "feat_stokes_assemble_device");
        L.K = LAFEM::SaddlePointMatrix::from_device(nv, np, LAFEM::BcsrDev::adopt(nv, nv, 2, 2, av, ac, ap),
                                                    LAFEM::BcsrDev::adopt(nv, np, 2, 1, bv, bc, bp),
                                                    LAFEM::BcsrDev::adopt(np, nv, 1, 2, dv, dc, dp));
        L.F = LAFEM::StokesFilter::from_device(nv, np, fn, fv, pr, du);
        if(transfer)
          L.T = LAFEM::StokesTransfer::from_device(nv, np, nvc, npc, npv, npp, pvv, pvc, pvp, ppv, ppc, ppp, rvv, rvc,
                                                   rvp, rpv, rpc, rpp);
        // the patterns of B and D for AmaVanka::init_symbolic
        L.host = Assembly::StokesLevelHost{};
        L.host.level = l;
        L.host.nv = nv;
        L.host.np = np;
        L.host.B.rows = nv;
        L.host.B.cols = np;
        L.host.B.bh = 2;
        L.host.B.bw = 1;
        L.host.B.rowptr.resize(nv + 1);
        L.host.B.col.resize(nb);
        bp.to_host(L.host.B.rowptr.data());
        bc.to_host(L.host.B.col.data());
        L.host.D.rows = np;
        L.host.D.cols = nv;
        L.host.D.bh = 1;
        L.host.D.bw = 2;
        L.host.D.rowptr.resize(np + 1);
        L.host.D.col.resize(nd);
        dp.to_host(L.host.D.rowptr.data());
        dc.to_host(L.host.D.col.data());
      }

      // The coarsest levels whose vectors fit the fused workgroup's LDS, each of at most
      // StokesConfig::tail_max_level_dofs unknowns.
      void _tail_plan()
      {
        const std::size_t nl = _lv.size();
        _tail_start = nl;
        if(!_cfg.coarse_fusion) return;
        std::uint64_t dofs = 0, pres = 0, words = 0;
        std::size_t t = nl;
        while(t > 0 && nl - t < std::size_t(FEAT_STOKES_TAIL_MAX_LEVELS))
        {
          const Level& L = *_lv[t - 1];
          const std::uint64_t w = (L.K.nv() + 31) / 32;
          if(L.K.size() > _cfg.tail_max_level_dofs || dofs + L.K.size() > FEAT_STOKES_TAIL_MAX_DOFS ||
             pres + L.K.np() > FEAT_STOKES_TAIL_MAX_PRESSURE || words + w > FEAT_STOKES_TAIL_MAX_MASK_WORDS)
            break;
          dofs += L.K.size();
          pres += L.K.np();
          words += w;
          --t;
        }
        if(t == nl) return;
        _tail = feat_stokes_tail_t{};
        _tail.nlevels = int(nl - t);
        _tail.steps = _cfg.steps;
        _tail.omega = _cfg.omega_rich;
        for(std::size_t i = t; i < nl; ++i)
        {
          const Level& L = *_lv[i];
          feat_stokes_tail_level_t& d = _tail.lv[i - t];
          d.K = L.K.desc();
          d.V = L.V->vanka_desc();
          d.vmask = L.F.vmask();
          d.prim = L.F.prim();
          d.dual = L.F.dual();
          d.volume = L.F.volume();
          if(i + 1 < nl)
          {
            d.rv = L.T.desc(0);
            d.rp = L.T.desc(1);
            d.pv = L.T.desc(2);
            d.pp = L.T.desc(3);
          }
        }
        _tail_start = t;
      }

      double read_norm(const LAFEM::StokesVector& v)
      {
        double* dn = _scal.get();
        check(feat_tuple_norm2_f64(Runtime::handle(), v.elements(), 2 * _lv[0]->K.nv(), _lv[0]->K.np(), dn, &_fw.ws), "norm");
        double r = 0.0;
        check(feat_read_scalar_f64(Runtime::handle(), dn, &r), "read norm");
        return r;
      }

    public:
      explicit StokesMultiGrid(const StokesConfig& cfg = StokesConfig()) : _cfg(cfg) {}
      std::string name() const override { return "Richardson(MG-V(AmaVanka))"; }

      /// levels finest first; level i > 0 of the list must be level(i-1) - 1
      void build(int level_fine, int level_coarse)
      {
        _lv.clear();
        for(int l = level_fine; l >= level_coarse; --l)
        {
          auto L = std::make_unique<Level>();
          if(_cfg.device_assembly)
            _assemble_on_device(*L, l, l > level_coarse);
          else
          {
            L->host = Assembly::stokes_q2p1dc(l);
            L->K = LAFEM::SaddlePointMatrix::from_host(L->host);
            L->F = LAFEM::StokesFilter::from_host(L->host);
            if(l > level_coarse) L->T = LAFEM::StokesTransfer::from_host(L->host);
          }
          _lv.push_back(std::move(L));
        }
      }
      std::size_t num_levels() const { return _lv.size(); }
      Level& level(std::size_t i) { return *_lv[i]; }

      void init_symbolic() override
      {
        for(std::size_t i = 0; i < _lv.size(); ++i)
        {
          Level& L = *_lv[i];
          L.V = std::make_unique<AmaVanka>(L.K, L.F, _cfg.vanka_omega);
          L.V->set_skip_singular(i + 1 == _lv.size() && _cfg.coarse_skip_singular);
          L.V->init_symbolic(L.host);
          for(LAFEM::StokesVector* v : {&L.rhs, &L.sol, &L.def, &L.cor, &L.rdef, &L.rcor}) *v = L.K.create_vector();
          L.gate = DeviceArray<double>(1);
        }
        _tail_plan();
        _d = _lv[0]->K.create_vector();
        _c = _lv[0]->K.create_vector();
        _scal = DeviceArray<double>(4);
        _fw.create(4);
        if(_cfg.graphs && !_graph)
        {
          auto g = std::make_unique<Graph>();
          void* hp = nullptr;
          check(feat_stream_create(&g->stream), "Stokes graph: stream");
          check(feat_handle_create(&g->h, 0, g->stream), "Stokes graph: handle");
          check(feat_host_malloc(&hp, 8 * sizeof(double)), "Stokes graph: pinned scalar");
          g->host = static_cast<double*>(hp);
          _graph = std::move(g);
        }
      }
      void init_numeric() override
      {
        for(auto& L : _lv) L->V->init_numeric();
        if(_graph) _graph->drop();   // the recording holds the numeric data's addresses
      }

      /// one smoother application on level i (for tests)
      void smooth_level(std::size_t i, LAFEM::StokesVector& sol, const LAFEM::StokesVector& rhs) { smooth(*_lv[i], sol, rhs); }
      void mg_apply_public(LAFEM::StokesVector& cor, const LAFEM::StokesVector& def) { mg_apply(cor, def); }

      Status apply(LAFEM::StokesVector& cor, const LAFEM::StokesVector& def) override
      {
        _d.copy(def);
        cor.format();
        return _run(cor, def);
      }

      /// Richardson::correct (richardson.hpp:174-185)
      Status correct(LAFEM::StokesVector& x, const LAFEM::StokesVector& b) override
      {
        defect(*_lv[0], _d, x, b);
        return _run(x, b);
      }

    private:
      Status _run(LAFEM::StokesVector& x, const LAFEM::StokesVector& b)
      {
        if(_graph && !_graph->failed) return _run_graph(x, b);
        Level& L = *_lv[0];
        Status st = this->_set_initial_defect(read_norm(_d));
        while(st == Status::progress)
        {
          mg_apply(_c, _d);
          x.axpy(_c, x, _cfg.omega_outer);
          defect(L, _d, x, b);
          st = this->_set_new_defect([&]() { return read_norm(_d); });
        }
        this->_status = st;
        return st;
      }

      // The same loop with each outer iteration (V-cycle, update, defect, norm) recorded
      // once as a HIP graph on the solver's stream and replayed: identical kernels in the
      // identical order, so identical results; one host read per iteration as before.
      Status _run_graph(LAFEM::StokesVector& x, const LAFEM::StokesVector& b)
      {
        Graph& g = *_graph;
        Level& L = *_lv[0];
        feat_handle_t caller = Runtime::handle();
        if(g.key_x != x.elements() || g.key_b != b.elements())
        {
          g.drop();
          g.key_x = x.elements();
          g.key_b = b.elements();
        }
        Status st = this->_set_initial_defect(read_norm(_d));
        if(st == Status::progress && !g.g)
        {
          Runtime::Use use(g.h);
          bool ok = feat_graph_begin(g.h) == FEAT_OK;
          if(ok)
          {
            try
            {
              mg_apply(_c, _d);
              x.axpy(_c, x, _cfg.omega_outer);
              defect(L, _d, x, b);
              check(feat_tuple_norm2_f64(g.h, _d.elements(), 2 * L.K.nv(), L.K.np(), _scal.get(), &_fw.ws), "norm");
              check(feat_copy_scalar_to_host(g.h, _scal.get(), g.host), "read-back");
            }
            catch(const Error&)
            {
              ok = false;
            }
            feat_graph_t gr = nullptr;
            if(feat_graph_end(g.h, &gr) != FEAT_OK) ok = false;
            if(ok)
              g.g = gr;
            else
              feat_graph_destroy(gr);
          }
          if(!ok)
          {
            g.failed = true;   // recording is not available: this and later solves run eagerly
            g.drop();
            while(st == Status::progress)
            {
              mg_apply(_c, _d);
              x.axpy(_c, x, _cfg.omega_outer);
              defect(L, _d, x, b);
              st = this->_set_new_defect([&]() { return read_norm(_d); });
            }
            this->_status = st;
            return st;
          }
        }
        if(st == Status::progress) check(feat_handle_wait(g.h, caller), "Stokes: order after caller");
        while(st == Status::progress)
        {
          check(feat_graph_launch(g.h, g.g), "Stokes graph");
          check(feat_sync(g.h), "Stokes sync");
          st = this->_set_new_defect([&]() { return g.host[0]; });
        }
        check(feat_handle_wait(caller, g.h), "Stokes: order before caller");
        this->_status = st;
        return st;
      }
    };
  } // namespace Solver
} // namespace feat3amd
