// feat3amd/solver.hpp -- Solver::SolverBase lifecycle and the iterative
// solvers of the hot path, on device containers.
//
// Mirrors kernel/solver/{base,iterative,jacobi_precond,richardson,pcg,
// multigrid}.hpp: init_symbolic/init_numeric/apply/done, the same status
// codes, defaults, stopping rules and op order.  Two execution forms:
//  * the generic form composes container ops exactly like the reference
//    (host scalars, one synchronisation per reduction);
//  * the fused form is taken by PCG(MultiGrid(Richardson(Jacobi))) and keeps
//    every scalar on the device (one synchronisation per PCG iteration for the
//    convergence test) and uses the fused kernels of feat3amd.h.  It computes
//    the same per-element arithmetic; only the global-sum order of dots and
//    norms differs.
#pragma once

#include <feat3amd/lafem.hpp>

#include <cfloat>
#include <cmath>
#include <functional>
#include <limits>
#include <memory>
#include <string>
#include <cstring>
#include <vector>

namespace feat3amd
{
  namespace Solver
  {
    /// solver/base.hpp:48-64
    enum class Status { undefined = 0, progress, success, aborted, diverged, max_iter, stagnated };

    inline bool status_success(Status s)
    {
      return s == Status::success || s == Status::max_iter || s == Status::stagnated;
    }

    /// Device workspace shared by the fused kernels of one solver tree.
    struct FusedWorkspace
    {
      DeviceArray<double> partials;
      DeviceArray<std::uint32_t> counter;
      DeviceArray<double> scalars;
      feat_reduce_ws_t ws{};
      void create(Index nscal)
      {
        if(!partials.empty()) return;
        partials = DeviceArray<double>(2 * FEAT_REDUCE_MAX_BLOCKS);
        counter = DeviceArray<std::uint32_t>(16);
        scalars = DeviceArray<double>(nscal);
        std::vector<std::uint32_t> z(16, 0u);
        check(feat_copy(Runtime::handle(), counter.get(), z.data(), sizeof(std::uint32_t) * 16), "workspace");
        ws.partials = partials.get();
        ws.counter = counter.get();
      }
      double* scalar(Index i) const { return scalars.get() + i; }
    };

    /// SolverBase<Vector_> (solver/base.hpp)
    template<typename Vector_>
    class SolverBase
    {
    public:
      typedef Vector_ VectorType;
      typedef typename Vector_::DataType DataType;
      virtual ~SolverBase() {}
      virtual std::string name() const = 0;
      virtual void init_symbolic() {}
      virtual void init_numeric() {}
      virtual void done_numeric() {}
      virtual void done_symbolic() {}
      void init()
      {
        init_symbolic();
        init_numeric();
      }
      void done()
      {
        done_numeric();
        done_symbolic();
      }
      virtual Status apply(Vector_& vec_cor, const Vector_& vec_def) = 0;
    };

    /// IterativeSolver bookkeeping (solver/iterative.hpp:117-885)
    template<typename Vector_>
    class IterativeSolver : public SolverBase<Vector_>
    {
    public:
      typedef typename Vector_::DataType DataType;

    protected:
      // defaults of iterative.hpp:213-236
      DataType _tol_rel = std::sqrt(std::numeric_limits<DataType>::epsilon());
      DataType _tol_abs = DataType(1) / (std::numeric_limits<DataType>::epsilon() * std::numeric_limits<DataType>::epsilon());
      DataType _tol_abs_low = DataType(0);
      DataType _div_rel = DataType(1) / std::numeric_limits<DataType>::epsilon();
      DataType _div_abs = DataType(1) / (std::numeric_limits<DataType>::epsilon() * std::numeric_limits<DataType>::epsilon());
      DataType _stag_rate = DataType(0.95);
      Index _min_iter = 0, _max_iter = 100, _num_iter = 0, _min_stag_iter = 0, _num_stag_iter = 0;
      DataType _def_init = 0, _def_cur = 0, _def_prev = 0;
      Status _status = Status::undefined;
      std::vector<DataType> _hist;

    public:
      void set_tol_rel(DataType t) { _tol_rel = t; }
      void set_tol_abs(DataType t) { _tol_abs = t; }
      void set_div_rel(DataType t) { _div_rel = t; }
      void set_div_abs(DataType t) { _div_abs = t; }
      void set_min_iter(Index n) { _min_iter = n; }
      void set_max_iter(Index n) { _max_iter = n; }
      void set_min_stag_iter(Index n) { _min_stag_iter = n; }
      void set_stag_rate(DataType r) { _stag_rate = r; }
      Index get_num_iter() const { return _num_iter; }
      Index get_min_iter() const { return _min_iter; }
      Index get_max_iter() const { return _max_iter; }
      DataType get_def_initial() const { return _def_init; }
      DataType get_def_final() const { return _def_cur; }
      Status get_status() const { return _status; }
      const std::vector<DataType>& history() const { return _hist; }
      /// defect of a smoother is never computed (iterative.hpp:862-871)
      bool skips_defect() const { return !(_min_iter < _max_iter) && _min_stag_iter == 0; }

      virtual Status correct(Vector_& vec_sol, const Vector_& vec_rhs) = 0;

    protected:
      /// iterative.hpp:748-770
      Status _set_initial_defect(DataType def)
      {
        _def_init = _def_cur = _def_prev = def;
        _num_iter = 0;
        _num_stag_iter = 0;
        _hist.assign(1, def);
        if(!std::isfinite(def)) return Status::aborted;
        const DataType e = std::numeric_limits<DataType>::epsilon();
        if(def <= e * e) return Status::success;
        return Status::progress;
      }

      /// iterative.hpp:795-836
      Status _analyse_defect(Index it, DataType cur, DataType prev, bool check_stag)
      {
        if(!std::isfinite(cur)) return Status::aborted;
        if(cur > _div_abs || cur > _div_rel * _def_init) return Status::diverged;
        if(it < _min_iter) return Status::progress;
        if(cur <= _tol_abs && (cur <= _tol_rel * _def_init || cur <= _tol_abs_low)) return Status::success;
        if(it >= _max_iter) return Status::max_iter;
        if(check_stag && _min_stag_iter > 0)
        {
          if(cur >= _stag_rate * prev)
          {
            if(++_num_stag_iter >= _min_stag_iter) return Status::stagnated;
          }
          else
            _num_stag_iter = 0;
        }
        return Status::progress;
      }

      /// iterative.hpp:853-885; `norm` is evaluated only when the reference computes it
      Status _set_new_defect(const std::function<DataType()>& norm)
      {
        ++_num_iter;
        _def_prev = _def_cur;
        if(!skips_defect())
        {
          _def_cur = norm();
          _hist.push_back(_def_cur);
        }
        return _analyse_defect(_num_iter, _def_cur, _def_prev, true);
      }
    };

    /// JacobiPrecond (solver/jacobi_precond.hpp:136-164)
    template<typename Matrix_, typename Filter_>
    class JacobiPrecond : public SolverBase<typename Matrix_::VectorTypeL>
    {
    public:
      typedef typename Matrix_::VectorTypeL VectorType;
      typedef typename Matrix_::DataType DataType;

    private:
      const Matrix_& _matrix;
      const Filter_& _filter;
      DataType _omega;
      VectorType _inv_diag;

    public:
      JacobiPrecond(const Matrix_& m, const Filter_& f, DataType omega = DataType(1)) : _matrix(m), _filter(f), _omega(omega) {}
      std::string name() const override { return "Jacobi"; }
      void init_symbolic() override { _inv_diag = _matrix.create_vector_r(); }
      void init_numeric() override
      {
        _matrix.extract_diag(_inv_diag);
        _inv_diag.component_invert(_inv_diag, _omega);
      }
      Status apply(VectorType& cor, const VectorType& def) override
      {
        cor.component_product(_inv_diag, def);
        _filter.filter_cor(cor);
        return Status::success;
      }
      const VectorType& inv_diag() const { return _inv_diag; }
      const Matrix_& matrix() const { return _matrix; }
      const Filter_& filter() const { return _filter; }
      DataType omega() const { return _omega; }
    };

    template<typename Matrix_, typename Filter_>
    std::shared_ptr<JacobiPrecond<Matrix_, Filter_>> new_jacobi_precond(const Matrix_& m, const Filter_& f,
                                                                        typename Matrix_::DataType omega = 1)
    {
      return std::make_shared<JacobiPrecond<Matrix_, Filter_>>(m, f, omega);
    }

    /// PreconditionedIterativeSolver::_apply_precond (iterative.hpp:1067-1080)
    template<typename Vector_, typename Filter_>
    bool apply_precond(SolverBase<Vector_>* precond, Vector_& cor, const Vector_& def, const Filter_& filter)
    {
      if(precond) return status_success(precond->apply(cor, def));
      cor.copy(def);
      filter.filter_cor(cor);
      return true;
    }

    /// Richardson (solver/richardson.hpp:160-226)
    template<typename Matrix_, typename Filter_>
    class Richardson : public IterativeSolver<typename Matrix_::VectorTypeL>
    {
    public:
      typedef typename Matrix_::VectorTypeL VectorType;
      typedef typename Matrix_::DataType DataType;
      typedef SolverBase<VectorType> PrecondType;

    private:
      const Matrix_& _matrix;
      const Filter_& _filter;
      DataType _omega;
      std::shared_ptr<PrecondType> _precond;
      VectorType _vec_def, _vec_cor;

    public:
      Richardson(const Matrix_& m, const Filter_& f, DataType omega, std::shared_ptr<PrecondType> p = nullptr)
        : _matrix(m), _filter(f), _omega(omega), _precond(p) {}
      std::string name() const override { return "Richardson"; }
      void init_symbolic() override
      {
        if(_precond) _precond->init_symbolic();
        _vec_def = _matrix.create_vector_r();
        _vec_cor = _matrix.create_vector_r();
      }
      void init_numeric() override
      {
        if(_precond) _precond->init_numeric();
      }
      DataType omega() const { return _omega; }
      const Matrix_& matrix() const { return _matrix; }
      const Filter_& filter() const { return _filter; }

      /// The Jacobi preconditioner when this smoother can run fused
      /// (min_iter == max_iter, no stagnation test), else nullptr.
      JacobiPrecond<Matrix_, Filter_>* fusable_jacobi() const
      {
        if(!this->skips_defect() || this->_max_iter < 1) return nullptr;
        return dynamic_cast<JacobiPrecond<Matrix_, Filter_>*>(_precond.get());
      }

      Status apply(VectorType& vec_cor, const VectorType& vec_def) override
      {
        _vec_def.copy(vec_def);
        vec_cor.format();
        this->_status = _apply_intern(vec_cor, vec_def);
        return this->_status;
      }

      Status correct(VectorType& vec_sol, const VectorType& vec_rhs) override
      {
        _matrix.apply(_vec_def, vec_sol, vec_rhs, -DataType(1));
        _filter.filter_def(_vec_def);
        this->_status = _apply_intern(vec_sol, vec_rhs);
        return this->_status;
      }

    private:
      Status _apply_intern(VectorType& vec_sol, const VectorType& vec_rhs)
      {
        Status st = this->_set_initial_defect(_vec_def.norm2());
        while(st == Status::progress)
        {
          if(!apply_precond(_precond.get(), _vec_cor, _vec_def, _filter)) return Status::aborted;
          vec_sol.axpy(_vec_cor, vec_sol, _omega);
          _matrix.apply(_vec_def, vec_sol, vec_rhs, -DataType(1));
          _filter.filter_def(_vec_def);
          st = this->_set_new_defect([&]() { return _vec_def.norm2(); });
        }
        return st;
      }
    };

    template<typename Matrix_, typename Filter_>
    std::shared_ptr<Richardson<Matrix_, Filter_>> new_richardson(
      const Matrix_& m, const Filter_& f, typename Matrix_::DataType omega,
      std::shared_ptr<SolverBase<typename Matrix_::VectorTypeL>> p = nullptr)
    {
      return std::make_shared<Richardson<Matrix_, Filter_>>(m, f, omega, p);
    }

    /// Exact coarse solver: inverse of the filtered coarse matrix (Dirichlet
    /// rows/columns replaced by identity), inverted once on the host in FP64
    /// by Gauss-Jordan with diagonal pivoting (the scheme of util/math.hpp:1169)
    /// and applied on the device as a dense GEMV followed by filter_cor.
    template<typename Matrix_, typename Filter_>
    class DirectCoarseSolver : public SolverBase<typename Matrix_::VectorTypeL>
    {
    public:
      typedef typename Matrix_::VectorTypeL VectorType;

    private:
      const Matrix_& _matrix;
      const Filter_& _filter;
      std::vector<double> _host_a, _host_val;
      std::vector<std::uint64_t> _host_col, _host_rp, _host_filt;
      DeviceArray<double> _inv;

    public:
      DirectCoarseSolver(const Matrix_& m, const Filter_& f, std::vector<double> val, std::vector<std::uint64_t> col,
                         std::vector<std::uint64_t> rp, std::vector<std::uint64_t> filt)
        : _matrix(m), _filter(f), _host_val(std::move(val)), _host_col(std::move(col)), _host_rp(std::move(rp)),
          _host_filt(std::move(filt)) {}
      std::string name() const override { return "DirectCoarse"; }
      void init_numeric() override
      {
        const Index n = _matrix.rows();
        if(n > 4096) throw Error("DirectCoarseSolver: coarse matrix too large for a dense inverse");
        // the values as they are now (the pattern is fixed since init_symbolic)
        _host_val.resize(_matrix.used_elements());
        if(!_host_val.empty())
        {
          check(feat_copy(Runtime::handle(), _host_val.data(), _matrix.val(), sizeof(double) * _host_val.size()),
                "DirectCoarseSolver: values");
          check(feat_sync(Runtime::handle()), "DirectCoarseSolver: sync");
        }
        std::vector<char> f(n, 0);
        for(auto i : _host_filt) f[i] = 1;
        std::vector<double> a(n * n, 0.0);
        for(Index r = 0; r < n; ++r)
        {
          if(f[r])
          {
            a[r * n + r] = 1.0;
            continue;
          }
          for(Index i = _host_rp[r]; i < _host_rp[r + 1]; ++i)
            if(!f[_host_col[i]]) a[r * n + _host_col[i]] = _host_val[i];
        }
        invert(int(n), a.data());
        _inv = DeviceArray<double>::from_host(a.data(), n * n);
      }
      Status apply(VectorType& cor, const VectorType& def) override
      {
        check(feat_dense_gemv_f64(Runtime::handle(), int(_matrix.rows()), _inv.get(), (const double*)def.elements(),
                                  (double*)cor.elements(), _filter.mask()),
              "DirectCoarseSolver::apply");
        return Status::success;
      }
      const double* inverse() const { return _inv.get(); }

      /// Gauss-Jordan with diagonal pivoting; each update is one fused multiply-add.
      static void invert(int n, double* a)
      {
        if(n == 1)
        {
          a[0] = 1.0 / a[0];
          return;
        }
        std::vector<int> p(n);
        for(int i = 0; i < n; ++i) p[i] = i;
        for(int k = 0; k < n; ++k)
        {
          double piv = std::fabs(a[p[k] * n + p[k]]);
          int i = k;
          for(int j = k + 1; j < n; ++j)
          {
            const double av = std::fabs(a[p[j] * n + p[j]]);
            if(av > piv) { piv = av; i = j; }
          }
          if(i > k) std::swap(p[k], p[i]);
          const int pk = p[k] * n;
          const double inv = 1.0 / a[pk + p[k]];
          a[pk + p[k]] = 1.0;
          for(int j = 0; j < n; ++j) a[pk + j] *= inv;
          for(int r = 0; r < n; ++r)
          {
            if(r == p[k]) continue;
            const int ro = r * n;
            const double fac = a[ro + p[k]];
            a[ro + p[k]] = 0.0;
            for(int j = 0; j < n; ++j) a[ro + j] = std::fma(-a[pk + j], fac, a[ro + j]);
          }
        }
      }
    };

    enum class MultiGridCycle { V };

    /// MultiGridHierarchy (solver/multigrid.hpp:419-957); level 0 is the finest.
    /// Grid forms of a multigrid level and its transfers.  The level is an n1^dim grid
    /// (dim 2 or 3, lexicographic) and the next one has (n1 + 1) / 2 nodes per side.
    ///  * dlinear: the prolongation from the next level is d-linear injection (the coarse
    ///    nodes around each fine node, ascending, weight 1/2 per odd axis) and the
    ///    restriction its transpose.  The transfers then run as FEAT_FMT_DLINEAR
    ///    descriptors, and in the fused coarse levels as feat_mg_tail_level_t dlinear.
    ///  * sten (fused coarse levels only): every free row (not in the mask) holds the same
    ///    3^dim values at the neighbour offsets in CSR order (feat_mg_tail_level_t sten).
    /// Both are checked entry by entry on host copies of the matrices.
    struct TailGridForm
    {
      int dim = 0;
      std::uint32_t n1 = 0;
      std::vector<double> sten;   // empty: the rows come from the plan
      bool dlinear = false;
    };

    template<typename Matrix_>
    TailGridForm tail_grid_form(const Matrix_& A, const std::uint32_t* mask_dev, const Matrix_& P, const Matrix_& R,
                                Index n_next, bool want_stencil)
    {
      TailGridForm f;
      const Index n = A.rows();
      int dim = 0;
      std::int64_t n1 = 0;
      for(int d = 3; d >= 2 && !dim; --d)
      {
        const std::int64_t r = std::llround(std::pow(double(n), 1.0 / d));
        for(std::int64_t c = std::max<std::int64_t>(r - 1, 3); c <= r + 1; ++c)
        {
          const std::int64_t cn = d == 3 ? c * c * c : c * c, nc = (c + 1) / 2;
          const std::int64_t ccn = d == 3 ? nc * nc * nc : nc * nc;
          if((c & 1) && cn == std::int64_t(n) && ccn == std::int64_t(n_next))
          {
            dim = d;
            n1 = c;
            break;
          }
        }
      }
      if(!dim || !Runtime::grid_forms()) return f;
      f.dim = dim;
      f.n1 = std::uint32_t(n1);
      auto down = [](auto* dst, const auto* src, std::size_t count) {
        if(count) check(feat_copy(Runtime::handle(), dst, src, sizeof(*dst) * count), "grid form: copy");
      };
      auto fetch = [&](const Matrix_& M, std::vector<double>& v, std::vector<std::uint64_t>& c,
                       std::vector<std::uint64_t>& p) {
        p.resize(M.rows() + 1);
        down(p.data(), (const std::uint64_t*)M.row_ptr(), p.size());
        v.resize(M.used_elements());
        c.resize(M.used_elements());
        down(v.data(), (const double*)M.val(), v.size());
        down(c.data(), (const std::uint64_t*)M.col_ind(), c.size());
      };
      std::vector<double> av, pv, rv;
      std::vector<std::uint64_t> ac, ap, pc, pp, rc, rp;
      if(want_stencil) fetch(A, av, ac, ap);
      fetch(P, pv, pc, pp);
      fetch(R, rv, rc, rp);
      std::vector<std::uint32_t> mask(want_stencil && mask_dev ? (n + 31) / 32 : 0, 0u);
      down(mask.data(), mask_dev, mask.size());
      Runtime::synchronize();
      const std::int64_t nn = n1 * n1, nc = (n1 + 1) / 2;
      const int zl = dim == 3 ? -1 : 0, zh = dim == 3 ? 1 : 0;
      auto coord = [&](std::int64_t i, std::int64_t side, std::int64_t& x, std::int64_t& y, std::int64_t& z) {
        x = i % side;
        y = (i / side) % side;
        z = dim == 3 ? i / (side * side) : 0;
      };
      // the stencil of the free rows
      bool sten_ok = want_stencil;
      for(std::int64_t i = 0; i < std::int64_t(n) && sten_ok; ++i)
      {
        if(!mask.empty() && ((mask[i >> 5] >> (i & 31)) & 1u)) continue;
        std::int64_t x, y, z;
        coord(i, n1, x, y, z);
        if(x < 1 || x > n1 - 2 || y < 1 || y > n1 - 2 || (dim == 3 && (z < 1 || z > n1 - 2)))
        {
          sten_ok = false;
          break;
        }
        std::uint64_t e = ap[i];
        if(ap[i + 1] - e != (dim == 3 ? 27u : 9u))
        {
          sten_ok = false;
          break;
        }
        std::size_t u = 0;
        for(int dz = zl; dz <= zh && sten_ok; ++dz)
          for(int dy = -1; dy <= 1 && sten_ok; ++dy)
            for(int dx = -1; dx <= 1; ++dx, ++u, ++e)
            {
              if(std::int64_t(ac[e]) != i + dz * nn + dy * n1 + dx)
              {
                sten_ok = false;
                break;
              }
              if(f.sten.size() <= u)
                f.sten.push_back(av[e]);
              else if(std::memcmp(&f.sten[u], &av[e], sizeof(double)) != 0)
              {
                sten_ok = false;
                break;
              }
            }
      }
      if(!sten_ok || f.sten.empty()) f.sten.clear();
      // the transfers
      bool p_ok = P.rows() == n && rp.size() == std::size_t(n_next) + 1;
      for(std::int64_t i = 0; i < std::int64_t(n) && p_ok; ++i)
      {
        std::int64_t x, y, z;
        coord(i, n1, x, y, z);
        const int ox = int(x & 1), oy = int(y & 1), oz = int(z & 1);
        const double w = (ox ? 0.5 : 1.0) * (oy ? 0.5 : 1.0) * (oz ? 0.5 : 1.0);
        std::uint64_t e = pp[i];
        if(pp[i + 1] - e != std::uint64_t((1 + ox) * (1 + oy) * (1 + oz)))
        {
          p_ok = false;
          break;
        }
        for(int iz = 0; iz <= oz; ++iz)
          for(int iy = 0; iy <= oy; ++iy)
            for(int ix = 0; ix <= ox; ++ix, ++e)
              if(std::int64_t(pc[e]) != ((z >> 1) + iz) * nc * nc + ((y >> 1) + iy) * nc + (x >> 1) + ix || pv[e] != w)
                p_ok = false;
      }
      for(std::int64_t j = 0; j < std::int64_t(n_next) && p_ok; ++j)
      {
        std::int64_t X, Y, Z;
        coord(j, nc, X, Y, Z);
        std::uint64_t e = rp[j];
        for(int dz = zl; dz <= zh; ++dz)
          for(int dy = -1; dy <= 1; ++dy)
            for(int dx = -1; dx <= 1; ++dx)
            {
              const std::int64_t fx = 2 * X + dx, fy = 2 * Y + dy, fz = 2 * Z + dz;
              if(fx < 0 || fx >= n1 || fy < 0 || fy >= n1 || (dim == 3 && (fz < 0 || fz >= n1))) continue;
              if(e >= rp[j + 1] || std::int64_t(rc[e]) != (fz * n1 + fy) * n1 + fx ||
                 rv[e] != (dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0) * (dz ? 0.5 : 1.0))
                p_ok = false;
              ++e;
            }
        if(e != rp[j + 1]) p_ok = false;
      }
      f.dlinear = p_ok;
      return f;
    }

    template<typename Matrix_, typename Filter_, typename Transfer_>
    class MultiGridHierarchy
    {
    public:
      typedef typename Matrix_::VectorTypeL VectorType;
      typedef SolverBase<VectorType> SolverType;
      struct Level
      {
        const Matrix_* matrix = nullptr;
        const Filter_* filter = nullptr;
        const Transfer_* transfer = nullptr;
        std::shared_ptr<SolverType> pre, post, coarse;
        VectorType rhs, sol, def, cor, tmp;
      };
      std::vector<Level> levels;

      void push_level(const Matrix_& m, const Filter_& f, const Transfer_& t, std::shared_ptr<SolverType> pre,
                      std::shared_ptr<SolverType> post, std::shared_ptr<SolverType> /*peak*/ = nullptr)
      {
        Level l;
        l.matrix = &m;
        l.filter = &f;
        l.transfer = &t;
        l.pre = pre;
        l.post = post;
        levels.push_back(l);
      }
      void push_level(const Matrix_& m, const Filter_& f, std::shared_ptr<SolverType> coarse)
      {
        Level l;
        l.matrix = &m;
        l.filter = &f;
        l.coarse = coarse;
        levels.push_back(l);
      }
      Index size() const { return levels.size(); }

      void init_symbolic()
      {
        for(auto& l : levels)
        {
          l.rhs = l.matrix->create_vector_l();
          l.sol = l.matrix->create_vector_l();
          l.def = l.matrix->create_vector_l();
          l.cor = l.matrix->create_vector_l();
          l.tmp = l.matrix->create_vector_l();
          for(auto* s : {l.pre.get(), l.post.get(), l.coarse.get()})
            if(s) s->init_symbolic();
        }
      }
      void init_numeric()
      {
        for(auto& l : levels)
          for(auto* s : {l.pre.get(), l.post.get(), l.coarse.get()})
            if(s) s->init_numeric();
      }
      void init()
      {
        init_symbolic();
        init_numeric();
      }
      void done() {}
    };

    /// A preconditioner that runs on device scalars: cor = M^-1 def, where
    /// *def_nrm2 (device) holds |def|^2 and, when gamma != nullptr, the device
    /// scalar *gamma receives cor . def.  PCG uses it to keep its loop free of
    /// host round trips.
    template<typename Vector_>
    class FusedPrecond
    {
    public:
      virtual ~FusedPrecond() {}
      virtual bool fusable() const = 0;
      virtual void apply_fused(Vector_& cor, const Vector_& def, const double* def_nrm2, double* gamma) = 0;
      /// The first iterate of the preconditioner's first smoother, formed by the kernel
      /// that produces def (feat_pcg_update_first_f64): x1 = filter(omega * (inv ∘ def)).
      struct FirstIterate
      {
        const double* inv = nullptr;
        const uint32_t* mask = nullptr;
        double omega = 0.0;
        double* x1 = nullptr;
      };
      /// Where the first iterate goes when cor is the output; false: not supported.
      virtual bool first_iterate(const Vector_& /*cor*/, FirstIterate& /*fi*/) { return false; }
      /// apply_fused after the producer of def has written the first iterate.
      virtual void apply_fused_from1(Vector_& cor, const Vector_& def, const double* def_nrm2, double* gamma)
      {
        apply_fused(cor, def, def_nrm2, gamma);
      }
    };

    /// MultiGrid (solver/multigrid.hpp:984-2000), V-cycle with fixed coarse-grid correction.
    template<typename Matrix_, typename Filter_, typename Transfer_>
    class MultiGrid : public SolverBase<typename Matrix_::VectorTypeL>,
                      public FusedPrecond<typename Matrix_::VectorTypeL>
    {
    public:
      typedef typename Matrix_::VectorTypeL VectorType;
      typedef MultiGridHierarchy<Matrix_, Filter_, Transfer_> HierarchyType;
      typedef Richardson<Matrix_, Filter_> RichType;
      typedef JacobiPrecond<Matrix_, Filter_> JacType;

    private:
      std::shared_ptr<HierarchyType> _h;
      FusedWorkspace _fw;
      bool _fusable = false;
      bool _autotune = true;
      bool _tail_on = true, _tail_built = false;
      Index _tail_t = 0;   // first level of the fused coarse part; 0: none (level 0 is never in it)
      feat_mg_tail_t _tail{};
      std::vector<TailGridForm> _grid;               // per level: grid form if it is fused
      std::vector<DeviceArray<double>> _grid_sten;   // its stencil on the device (empty: none)
      std::vector<feat_matrix_f64_t> _dl_P, _dl_R;   // d-linear transfer descriptors (where dlinear)

      /// The grid form of every level that the fused coarse part could hold, with host copies
      /// of its matrices: at init_numeric, never inside a recorded graph.
      void _scan_grid()
      {
        const Index L = _h->size();
        _grid.assign(L, TailGridForm{});
        _grid_sten.assign(L, DeviceArray<double>());
        _dl_P.assign(L, feat_matrix_f64_t{});
        _dl_R.assign(L, feat_matrix_f64_t{});
        for(Index s = 0; s + 1 < L; ++s)
        {
          auto& l = _h->levels[s];
          if(!l.transfer || !l.filter) continue;
          const bool cand = s >= 1 && l.matrix->rows() <= FEAT_MG_TAIL_MAX_ROWS;
          _grid[s] = tail_grid_form(*l.matrix, l.filter->mask(), l.transfer->get_mat_prol(), l.transfer->get_mat_rest(),
                                    _h->levels[s + 1].matrix->rows(), cand);
          if(!_grid[s].sten.empty())
            _grid_sten[s] = DeviceArray<double>::from_host(_grid[s].sten.data(), _grid[s].sten.size());
          if(_grid[s].dlinear)
          {
            check(feat_matrix_dlinear_f64(_grid[s].dim, _grid[s].n1, 0, &_dl_P[s]), "MultiGrid d-linear prolongation");
            check(feat_matrix_dlinear_f64(_grid[s].dim, _grid[s].n1, 1, &_dl_R[s]), "MultiGrid d-linear restriction");
          }
        }
      }
      /// the transfer descriptors of level i: d-linear where checked, else the plans
      const feat_matrix_f64_t& _rest(Index i)
      {
        return (i < _grid.size() && _grid[i].dlinear) ? _dl_R[i] : _h->levels[i].transfer->get_mat_rest().exec();
      }
      const feat_matrix_f64_t& _prol(Index i)
      {
        return (i < _grid.size() && _grid[i].dlinear) ? _dl_P[i] : _h->levels[i].transfer->get_mat_prol().exec();
      }
      enum class Coarse { none, direct, jacobi, richardson, other } _coarse_kind = Coarse::other;

      /// The coarse levels that run as one LDS-resident workgroup (feat_mg_tail_f64): the
      /// longest run ending at the coarsest level, starting at level 1 or below, whose
      /// vectors fit the LDS budget, with diagonal-dictionary plans and a direct coarse solve.
      void _build_tail()
      {
        _tail_built = true;
        _tail_t = 0;
        const Index L = _h->size(), c = L - 1;
        if(!_tail_on || !_fusable || _coarse_kind != Coarse::direct || L < 3 ||
           _h->levels[c].matrix->rows() > FEAT_MG_TAIL_MAX_ROWS)
          return;
        uint64_t doubles = 2 * _h->levels[c].matrix->rows();
        Index t = c;
        for(Index s = c; s-- > 1;)
        {
          const auto& l = _h->levels[s];
          const feat_matrix_f64_t& A = l.matrix->exec();
          const uint64_t need = doubles + 5 * l.matrix->rows() + 256;   // vectors and the pair table
          if(A.index != FEAT_IDX_DIAG || A.slice_words > 8 || need * sizeof(double) > FEAT_MG_TAIL_MAX_LDS ||
             c - s + 1 > FEAT_MG_TAIL_MAX_LEVELS || l.matrix->rows() > FEAT_MG_TAIL_MAX_ROWS)
            break;
          doubles = need;
          t = s;
        }
        if(t == c) return;
        _tail = feat_mg_tail_t{};
        _tail.nlevels = int(c - t + 1);
        for(Index s = t; s <= c; ++s)
        {
          auto& l = _h->levels[s];
          feat_mg_tail_level_t& d = _tail.lv[s - t];
          d.n = l.matrix->rows();
          d.mask = l.filter->mask();
          if(s == c) break;
          const feat_matrix_f64_t& A = l.matrix->exec();
          d.table = static_cast<const double*>(A.val);
          d.codes = static_cast<const uint32_t*>(A.col);
          d.words = A.slice_words;
          auto* pre = static_cast<RichType*>(l.pre.get());
          auto* post = static_cast<RichType*>(l.post.get());
          d.inv_diag = (const double*)pre->fusable_jacobi()->inv_diag().elements();
          if(post->fusable_jacobi()->inv_diag().elements() != pre->fusable_jacobi()->inv_diag().elements()) return;
          const Matrix_& R = l.transfer->get_mat_rest();
          const Matrix_& P = l.transfer->get_mat_prol();
          d.r_val = (const double*)R.val();
          d.r_col = (const uint64_t*)R.col_ind();
          d.r_ptr = (const uint64_t*)R.row_ptr();
          d.p_val = (const double*)P.val();
          d.p_col = (const uint64_t*)P.col_ind();
          d.p_ptr = (const uint64_t*)P.row_ptr();
          d.pre_steps = int(pre->get_max_iter());
          d.post_steps = int(post->get_max_iter());
          d.pre_omega = pre->omega();
          d.post_omega = post->omega();
          // grid form found at init_numeric (no copies here: this may run inside a recorded graph)
          if(s < _grid.size() && (_grid[s].dlinear || _grid_sten[s].size()))
          {
            d.dim = _grid[s].dim;
            d.n1 = _grid[s].n1;
            d.dlinear = _grid[s].dlinear ? 1 : 0;
            d.sten = _grid_sten[s].size() ? _grid_sten[s].get() : nullptr;
          }
        }
        _tail.coarse_inverse = static_cast<DirectCoarseSolver<Matrix_, Filter_>*>(_h->levels[c].coarse.get())->inverse();
        _tail_t = t;
      }

    public:
      explicit MultiGrid(std::shared_ptr<HierarchyType> h, MultiGridCycle = MultiGridCycle::V) : _h(h) {}
      std::string name() const override { return "MultiGrid-V"; }
      HierarchyType& hierarchy() { return *_h; }
      /// pick the SELL-64 launch shape of every level's matrices at init_symbolic (default on;
      /// results are bit-identical for every shape)
      void set_autotune(bool on) { _autotune = on; }
      /// run the coarse levels as one LDS-resident workgroup where they qualify (default on;
      /// the solution is bit-identical either way)
      void set_coarse_fusion(bool on)
      {
        _tail_on = on;
        _tail_built = false;
      }
      Index coarse_fusion_level() { if(!_tail_built) _build_tail(); return _tail_t; }
      /// fused levels in the grid form (stencil from the kernel, d-linear transfers)
      int coarse_fusion_grid_levels()
      {
        if(!_tail_built) _build_tail();
        int k = 0;
        for(int l = 0; _tail_t && l + 1 < _tail.nlevels; ++l) k += (_tail.lv[l].sten && _tail.lv[l].dlinear) ? 1 : 0;
        return k;
      }

      void init_symbolic() override
      {
        const Index L = _h->size();
        _fw.create(2 * L + 8);
        _tail_built = false;
        _fusable = true;
        for(Index i = 0; i + 1 < L; ++i)
        {
          auto* pre = dynamic_cast<RichType*>(_h->levels[i].pre.get());
          auto* post = dynamic_cast<RichType*>(_h->levels[i].post.get());
          if(!pre || !post || !pre->fusable_jacobi() || !post->fusable_jacobi()) _fusable = false;
        }
        auto* c = _h->levels[L - 1].coarse.get();
        if(!c)
          _coarse_kind = Coarse::none;
        else if(dynamic_cast<DirectCoarseSolver<Matrix_, Filter_>*>(c))
          _coarse_kind = Coarse::direct;
        else if(dynamic_cast<JacType*>(c))
          _coarse_kind = Coarse::jacobi;
        else if(auto* r = dynamic_cast<RichType*>(c))
          _coarse_kind = r->fusable_jacobi() ? Coarse::richardson : Coarse::other;
        else
          _coarse_kind = Coarse::other;
        if(_coarse_kind == Coarse::other) _fusable = false;
        _build_plans(true);
      }

      /// Execution plans (and launch shapes) of every level's matrices that have none,
      /// at init_symbolic and again at init_numeric for matrices whose values changed:
      /// plans allocate, so they are never built inside a recorded graph.
      void _build_plans(bool all = false)
      {
        auto prep = [&](const Matrix_& m) {
          const bool fresh = all || !m.has_plan();
          (void)m.exec();
          if(_autotune && fresh) (void)m.autotune();
        };
        const Index L = _h->size();
        for(Index i = 0; i < L; ++i)
        {
          prep(*_h->levels[i].matrix);
          if(i + 1 < L)
          {
            prep(_h->levels[i].transfer->get_mat_prol());
            prep(_h->levels[i].transfer->get_mat_rest());
          }
        }
      }
      /// multigrid.hpp:1235-1241: the hierarchy's init_numeric runs first (it may
      /// replace the coarse inverse and the Jacobi diagonals), so the fused coarse
      /// levels' descriptor, which holds their addresses, is built again on next use
      void init_numeric() override
      {
        _tail_built = false;
        _build_plans();
        _scan_grid();
      }
      void done_numeric() override { _tail_built = false; }
      void done_symbolic() override { _tail_built = false; }
      bool fusable() const override { return _fusable; }
      const FusedWorkspace& workspace() const { return _fw; }

      /// MultiGrid::apply (multigrid.hpp:1270-1330)
      Status apply(VectorType& vec_cor, const VectorType& vec_def) override
      {
        if(_fusable)
        {
          double* nrm = _fw.scalar(0);
          check(feat_dot_f64(Runtime::handle(), nrm, (const double*)vec_def.elements(), (const double*)vec_def.elements(),
                             vec_def.size()),
                "MultiGrid::apply norm");
          apply_fused(vec_cor, vec_def, nrm, nullptr);
          return Status::success;
        }
        auto& top = _h->levels[0];
        top.rhs.copy(vec_def);
        _apply_generic();
        vec_cor.copy(top.sol);
        return Status::success;
      }

      /// The pre-smoother's first ping-pong buffer on level i with output x (smooth_impl: the
      /// first iterate lives in tmp for an even number of steps, in x for an odd one); null
      /// where the level's smoother does not start from a producer-formed iterate (one step,
      /// or the level runs in the fused tail or is the coarsest).
      double* _x1_of(Index i, double* x, bool pre)
      {
        auto& l = _h->levels[i];
        auto* sm = static_cast<RichType*>((pre ? l.pre : l.post).get());
        const int steps = int(sm->get_max_iter());
        if(steps < 2) return nullptr;
        return (steps & 1) ? x : (double*)l.tmp.elements();
      }

    public:
      bool first_iterate(const VectorType& cor, typename FusedPrecond<VectorType>::FirstIterate& fi) override
      {
        if(!_fusable) return false;
        if(!_tail_built) _build_tail();
        const Index c = _h->size() - 1;
        const Index t = _tail_t ? _tail_t : c;
        if(t == 0) return false;   // level 0 runs in the fused tail or is the coarsest
        double* x1 = _x1_of(0, (double*)cor.elements(), true);
        if(!x1) return false;
        auto& l = _h->levels[0];
        auto* pre = static_cast<RichType*>(l.pre.get());
        fi.inv = (const double*)pre->fusable_jacobi()->inv_diag().elements();
        fi.mask = l.filter->mask();
        fi.omega = pre->omega();
        fi.x1 = x1;
        return true;
      }
      void apply_fused_from1(VectorType& cor, const VectorType& def, const double* def_nrm2, double* gamma) override
      {
        _apply_fused(cor, def, def_nrm2, gamma, true);
      }

      /// Fused V-cycle: cor = MG(def) where *def_nrm2 (device) holds |def|^2;
      /// when gamma != nullptr, *gamma = cor . def (device).
      void apply_fused(VectorType& cor, const VectorType& def, const double* def_nrm2, double* gamma) override
      {
        _apply_fused(cor, def, def_nrm2, gamma, false);
      }

    private:
      /// top_x1: level 0's pre-smoother starts from the iterate the producer of def formed.
      /// Every other smoother above the fused tail starts from the iterate formed by the
      /// kernel that produced its right-hand side: the restriction for pre-smoothing, the
      /// defect after the prolongation for post-smoothing.
      void _apply_fused(VectorType& cor, const VectorType& def, const double* def_nrm2, double* gamma, bool top_x1)
      {
        const Index L = _h->size();
        const Index c = L - 1;
        feat_handle_t hd = Runtime::handle();
        const feat_reduce_ws_t* ws = &_fw.ws;
        if(!_tail_built) _build_tail();
        const Index t = _tail_t ? _tail_t : c;   // levels below t run in the fused tail (when _tail_t)
        auto rhs_of = [&](Index i) -> const double* {
          return i == 0 ? (const double*)def.elements() : (const double*)_h->levels[i].rhs.elements();
        };
        auto sol_of = [&](Index i) -> double* {
          return i == 0 ? (double*)cor.elements() : (double*)_h->levels[i].sol.elements();
        };
        auto nrm_of = [&](Index i) -> const double* { return i == 0 ? def_nrm2 : _fw.scalar(2 + i); };
        // restriction leg (multigrid.hpp:1713-1840)
        bool x1_ready = top_x1;
        for(Index i = 0; i < t; ++i)
        {
          auto& l = _h->levels[i];
          auto* pre = static_cast<RichType*>(l.pre.get());
          auto* jac = pre->fusable_jacobi();
          const feat_matrix_f64_t& A = l.matrix->exec();
          auto smooth = x1_ready ? feat_rj_smooth_from1_f64 : feat_rj_smooth_f64;
          check(smooth(hd, &A, (const double*)jac->inv_diag().elements(), l.filter->mask(), int(pre->get_max_iter()),
                       pre->omega(), rhs_of(i), nrm_of(i), sol_of(i), (double*)l.tmp.elements(),
                       (double*)l.def.elements(), nullptr, nullptr, nullptr, ws),
                "MultiGrid pre-smoothing");
          const feat_matrix_f64_t& R = _rest(i);
          auto& lc = _h->levels[i + 1];
          // the next level's pre-smoother starts from the iterate the restriction forms
          double* x1 = (i + 1 < t) ? _x1_of(i + 1, sol_of(i + 1), true) : nullptr;
          if(x1)
          {
            auto* pc = static_cast<RichType*>(lc.pre.get());
            check(feat_apply_masked_first_f64(hd, &R, (double*)lc.rhs.elements(), (const double*)l.def.elements(),
                                              lc.filter->mask(), _fw.scalar(2 + i + 1),
                                              (const double*)pc->fusable_jacobi()->inv_diag().elements(), pc->omega(),
                                              x1, ws),
                  "MultiGrid restriction");
          }
          else
            check(feat_apply_masked_f64(hd, &R, (double*)lc.rhs.elements(), (const double*)l.def.elements(),
                                        lc.filter->mask(), _fw.scalar(2 + i + 1), nullptr, nullptr, ws),
                  "MultiGrid restriction");
          x1_ready = x1 != nullptr;
        }
        if(_tail_t)
          check(feat_mg_tail_f64(hd, &_tail, rhs_of(t), nrm_of(t), sol_of(t)), "MultiGrid fused coarse levels");
        // coarse solve (multigrid.hpp:1526-1589)
        else
        {
          auto& l = _h->levels[c];
          double* xs = sol_of(c);
          switch(_coarse_kind)
          {
          case Coarse::direct:
          {
            auto* d = static_cast<DirectCoarseSolver<Matrix_, Filter_>*>(l.coarse.get());
            check(feat_dense_gemv_f64(hd, int(l.matrix->rows()), d->inverse(), rhs_of(c), xs, l.filter->mask()),
                  "MultiGrid coarse direct");
            break;
          }
          case Coarse::jacobi:
          {
            // cor = filter_cor(inv_diag∘rhs): one step from zero with omega 1 is exact (fma(1,c,0) = c)
            auto* j = static_cast<JacType*>(l.coarse.get());
            const feat_matrix_f64_t& A = l.matrix->exec();
            check(feat_rj_smooth_f64(hd, &A, (const double*)j->inv_diag().elements(), l.filter->mask(), 1, 1.0,
                                     rhs_of(c), nullptr, xs, (double*)l.tmp.elements(), nullptr, nullptr, nullptr,
                                     nullptr, ws),
                  "MultiGrid coarse Jacobi");
            break;
          }
          case Coarse::richardson:
          {
            auto* r = static_cast<RichType*>(l.coarse.get());
            const feat_matrix_f64_t& A = l.matrix->exec();
            check(feat_rj_smooth_f64(hd, &A, (const double*)r->fusable_jacobi()->inv_diag().elements(), l.filter->mask(),
                                     int(r->get_max_iter()), r->omega(), rhs_of(c), nrm_of(c), xs,
                                     (double*)l.tmp.elements(), nullptr, nullptr, nullptr, nullptr, ws),
                  "MultiGrid coarse Richardson");
            break;
          }
          default:
            check(feat_copy(hd, xs, rhs_of(c), sizeof(double) * l.matrix->rows()), "MultiGrid coarse copy");
            if(l.filter->used_elements())
              check(feat_unit_filter_def_f64_u64(hd, xs, l.filter->indices(), l.filter->used_elements()),
                    "MultiGrid coarse filter");
            break;
          }
        }
        // prolongation leg (multigrid.hpp:1845-2000)
        for(Index k = t; k-- > 0;)
        {
          auto& l = _h->levels[k];
          auto* post = static_cast<RichType*>(l.post.get());
          auto* jac = post->fusable_jacobi();
          const feat_matrix_f64_t& P = _prol(k);
          check(feat_prol_add_f64(hd, &P, sol_of(k), sol_of(k + 1), l.filter->mask()), "MultiGrid prolongation");
          const feat_matrix_f64_t& A = l.matrix->exec();
          double* dn = _fw.scalar(2 + L + k);
          // the post-smoother starts from the iterate the defect kernel forms
          double* x1 = _x1_of(k, (double*)l.cor.elements(), false);
          if(x1)
            check(feat_defect_first_f64(hd, &A, (double*)l.def.elements(), rhs_of(k), sol_of(k), l.filter->mask(), dn,
                                        (const double*)jac->inv_diag().elements(), post->omega(), x1, ws),
                  "MultiGrid defect");
          else
            check(feat_defect_f64(hd, &A, (double*)l.def.elements(), rhs_of(k), sol_of(k), l.filter->mask(), dn, ws),
                  "MultiGrid defect");
          const bool top = (k == 0);
          auto smooth = x1 ? feat_rj_smooth_from1_f64 : feat_rj_smooth_f64;
          check(smooth(hd, &A, (const double*)jac->inv_diag().elements(), l.filter->mask(), int(post->get_max_iter()),
                       post->omega(), (const double*)l.def.elements(), dn, (double*)l.cor.elements(),
                       (double*)l.tmp.elements(), nullptr, sol_of(k),
                       top && gamma ? (const double*)def.elements() : nullptr, top ? gamma : nullptr, ws),
                "MultiGrid post-smoothing");
        }
      }

    private:
      /// Generic V-cycle, op for op as multigrid.hpp:1335-2000.
      void _apply_generic()
      {
        const Index L = _h->size();
        const Index c = L - 1;
        for(Index i = 0; i < c; ++i)
        {
          auto& l = _h->levels[i];
          auto& lc = _h->levels[i + 1];
          l.pre->apply(l.sol, l.rhs);
          l.matrix->apply(l.def, l.sol, l.rhs, -1.0);
          l.filter->filter_def(l.def);
          l.transfer->rest(l.def, lc.rhs);
          lc.filter->filter_def(lc.rhs);
        }
        auto& lc = _h->levels[c];
        if(lc.coarse)
          lc.coarse->apply(lc.sol, lc.rhs);
        else
        {
          lc.sol.copy(lc.rhs);
          lc.filter->filter_cor(lc.sol);
        }
        for(Index k = c; k-- > 0;)
        {
          auto& l = _h->levels[k];
          l.transfer->prol(l.cor, _h->levels[k + 1].sol);
          l.filter->filter_cor(l.cor);
          l.sol.axpy(l.cor, l.sol, 1.0);
          l.matrix->apply(l.def, l.sol, l.rhs, -1.0);
          l.filter->filter_def(l.def);
          l.post->apply(l.cor, l.def);
          l.sol.axpy(l.cor, l.sol);
        }
      }
    };

    /// PCG (solver/pcg.hpp:33-288)
    template<typename Matrix_, typename Filter_>
    class PCG : public IterativeSolver<typename Matrix_::VectorTypeL>
    {
    public:
      typedef typename Matrix_::VectorTypeL VectorType;
      typedef typename Matrix_::DataType DataType;
      typedef SolverBase<VectorType> PrecondType;
      typedef FusedPrecond<VectorType> MGType;

    private:
      const Matrix_& _matrix;
      const Filter_& _filter;
      std::shared_ptr<PrecondType> _precond;
      VectorType _vec_r, _vec_p, _vec_t;
      FusedWorkspace _fw;
      bool _force_generic = false;
      bool _use_graphs = true;

      /// Solver-owned stream, handle and recorded graphs of the fused loop: the
      /// first iteration (MG(p <- r), A*p, update, defect read-back) and every
      /// later one (MG(t <- r), direction, A*p, update, read-back).
      struct Graphs
      {
        feat_stream_t stream = nullptr;
        feat_handle_t h = nullptr;
        double* host = nullptr;
        feat_graph_t first = nullptr, loop = nullptr;
        const void* key_x = nullptr;
        const void* key_b = nullptr;
        bool failed = false;
        void drop()
        {
          feat_graph_destroy(first);
          feat_graph_destroy(loop);
          first = loop = nullptr;
        }
        ~Graphs()
        {
          drop();
          if(h) feat_handle_destroy(h);
          if(stream) feat_stream_destroy(stream);
          if(host) feat_host_free(host);
        }
      };
      std::unique_ptr<Graphs> _graphs;

    public:
      PCG(const Matrix_& m, const Filter_& f, std::shared_ptr<PrecondType> p = nullptr)
        : _matrix(m), _filter(f), _precond(p) {}
      std::string name() const override { return "PCG"; }
      void set_force_generic(bool g) { _force_generic = g; }
      /// Record the fused loop as HIP graphs (default) or launch every kernel eagerly.
      void set_use_graphs(bool g) { _use_graphs = g; }
      void init_symbolic() override
      {
        if(_precond) _precond->init_symbolic();
        _vec_r = _matrix.create_vector_r();
        _vec_p = _matrix.create_vector_r();
        _vec_t = _matrix.create_vector_r();
        _fw.create(8);
        (void)_matrix.exec();
        if(_use_graphs && !_graphs)
        {
          auto g = std::make_unique<Graphs>();
          void* hp = nullptr;
          check(feat_stream_create(&g->stream), "PCG graphs: stream");
          check(feat_handle_create(&g->h, 0, g->stream), "PCG graphs: handle");
          check(feat_host_malloc(&hp, 8 * sizeof(double)), "PCG graphs: pinned scalar");
          g->host = static_cast<double*>(hp);
          _graphs = std::move(g);
        }
      }
      void init_numeric() override
      {
        if(_precond) _precond->init_numeric();
        // recorded graphs hold the numeric data's addresses: record again on next use
        if(_graphs) _graphs->drop();
      }
      void done_symbolic() override
      {
        _graphs.reset();
        if(_precond) _precond->done_symbolic();
      }

      /// pcg.hpp:140-156
      Status apply(VectorType& vec_cor, const VectorType& vec_def) override
      {
        _vec_r.copy(vec_def);
        vec_cor.format();
        this->_status = _run(vec_cor, false, nullptr);
        return this->_status;
      }

      /// pcg.hpp:159-173 (Solver::solve -> correct, iterative.hpp:930-938)
      Status correct(VectorType& vec_sol, const VectorType& vec_rhs) override
      {
        this->_status = _run(vec_sol, true, &vec_rhs);
        return this->_status;
      }

    private:
      MGType* _fused_mg() const
      {
        if(_force_generic) return nullptr;
        auto* mg = dynamic_cast<MGType*>(_precond.get());
        return (mg && mg->fusable()) ? mg : nullptr;
      }

      Status _run(VectorType& x, bool correct, const VectorType* rhs)
      {
        if(MGType* mg = _fused_mg())
        {
          if(correct && _use_graphs && _graphs && !_graphs->failed) return _run_graphs(*mg, x, *rhs);
          return _run_fused(*mg, x, correct, rhs);
        }
        if(correct)
        {
          _matrix.apply(_vec_r, x, *rhs, -DataType(1));
          _filter.filter_def(_vec_r);
        }
        return _apply_intern_generic(x);
      }

      /// pcg.hpp:188-288, op for op.
      Status _apply_intern_generic(VectorType& vec_sol)
      {
        VectorType& r = _vec_r;
        VectorType& p = _vec_p;
        VectorType& q = _vec_t;
        VectorType& z = _vec_t;
        Status st = this->_set_initial_defect(r.norm2());
        if(st != Status::progress) return st;
        if(!apply_precond(_precond.get(), p, r, _filter)) return Status::aborted;
        DataType gamma = r.dot(p);
        while(st == Status::progress)
        {
          _matrix.apply(q, p);
          _filter.filter_def(q);
          DataType alpha = gamma / q.dot(p);
          vec_sol.axpy(p, vec_sol, alpha);
          r.axpy(q, r, -alpha);
          st = this->_set_new_defect([&]() { return r.norm2(); });
          if(st != Status::progress) return st;
          if(!apply_precond(_precond.get(), z, r, _filter)) return Status::aborted;
          DataType gamma2 = gamma;
          gamma = r.dot(z);
          DataType beta = gamma / gamma2;
          p.axpy(p, z, beta);
        }
        return st;
      }

      /// The fused loop recorded as two HIP graphs and replayed once per iteration
      /// on the solver's own stream (ordered after and before the caller's stream).
      /// Same kernels in the same order as _run_fused, so the results are identical.
      Status _run_graphs(MGType& mg, VectorType& x, const VectorType& rhs)
      {
        Graphs& g = *_graphs;
        feat_handle_t caller = Runtime::handle();
        if(g.key_x != x.elements() || g.key_b != rhs.elements())
        {
          g.drop();
          g.key_x = x.elements();
          g.key_b = rhs.elements();
        }
        const feat_reduce_ws_t* ws = &_fw.ws;
        double* s_rr = _fw.scalar(0);
        double* s_g = _fw.scalar(1);
        double* s_g2 = _fw.scalar(2);
        double* s_qp = _fw.scalar(3);
        const feat_matrix_f64_t& A = _matrix.exec();
        const Index n = _matrix.rows();
        double* r = (double*)_vec_r.elements();
        double* p = (double*)_vec_p.elements();
        double* t = (double*)_vec_t.elements();
        double* xs = (double*)x.elements();
        typename MGType::FirstIterate fi;
        const bool from1 = mg.first_iterate(_vec_t, fi);
        auto body_tail = [&]() {
          check(feat_apply_masked_f64(g.h, &A, t, p, _filter.mask(), nullptr, p, s_qp, ws), "PCG A*p");
          if(from1)
            check(feat_pcg_update_first_f64(g.h, xs, r, p, t, s_g, s_qp, s_rr, n, fi.inv, fi.mask, fi.omega, fi.x1, ws),
                  "PCG update");
          else
            check(feat_pcg_update_f64(g.h, xs, r, p, t, s_g, s_qp, s_rr, n, ws), "PCG update");
          check(feat_copy_scalar_to_host(g.h, s_rr, g.host), "PCG read-back");
        };
        auto record = [&](feat_graph_t& slot, const std::function<void()>& body) -> bool {
          if(slot) return true;
          Runtime::Use use(g.h);
          if(feat_graph_begin(g.h) != FEAT_OK) return false;
          bool ok = true;
          try
          {
            body();
          }
          catch(const Error&)
          {
            ok = false;
          }
          feat_graph_t gr = nullptr;
          if(feat_graph_end(g.h, &gr) != FEAT_OK) ok = false;
          if(!ok)
          {
            feat_graph_destroy(gr);
            return false;
          }
          slot = gr;
          return true;
        };
        check(feat_handle_wait(g.h, caller), "PCG: order after caller");
        check(feat_defect_f64(g.h, &A, r, (const double*)rhs.elements(), xs, _filter.mask(), s_rr, ws), "PCG defect");
        check(feat_copy_scalar_to_host(g.h, s_rr, g.host), "PCG read-back");
        check(feat_sync(g.h), "PCG sync");
        Status st = this->_set_initial_defect(std::sqrt(g.host[0]));
        if(st == Status::progress)
        {
          const bool ok = record(g.first, [&]() {
            mg.apply_fused(_vec_p, _vec_r, s_rr, s_g);
            body_tail();
          }) && record(g.loop, [&]() {
            if(from1)
              mg.apply_fused_from1(_vec_t, _vec_r, s_rr, s_g2);
            else
              mg.apply_fused(_vec_t, _vec_r, s_rr, s_g2);
            check(feat_pcg_direction_f64(g.h, p, t, s_g2, s_g, n), "PCG direction");
            body_tail();
          });
          if(!ok)
          {
            // recording is not available: run this and later solves eagerly
            g.failed = true;
            g.drop();
            check(feat_handle_wait(caller, g.h), "PCG: order before caller");
            return _run_fused(mg, x, true, &rhs);
          }
          feat_graph_t cur = g.first;
          while(st == Status::progress)
          {
            check(feat_graph_launch(g.h, cur), "PCG graph");
            check(feat_sync(g.h), "PCG sync");
            cur = g.loop;
            st = this->_set_new_defect([&]() { return std::sqrt(g.host[0]); });
          }
        }
        check(feat_handle_wait(caller, g.h), "PCG: order before caller");
        return st;
      }

      /// Device-scalar form of the same loop with the fused kernels.
      Status _run_fused(MGType& mg, VectorType& x, bool correct, const VectorType* rhs)
      {
        feat_handle_t hd = Runtime::handle();
        const feat_reduce_ws_t* ws = &_fw.ws;
        double* s_rr = _fw.scalar(0);
        double* s_g = _fw.scalar(1);
        double* s_g2 = _fw.scalar(2);
        double* s_qp = _fw.scalar(3);
        const feat_matrix_f64_t& A = _matrix.exec();
        const Index n = _matrix.rows();
        double* r = (double*)_vec_r.elements();
        double* p = (double*)_vec_p.elements();
        double* t = (double*)_vec_t.elements();
        if(correct)
          check(feat_defect_f64(hd, &A, r, (const double*)rhs->elements(), (const double*)x.elements(), _filter.mask(),
                                s_rr, ws),
                "PCG defect");
        else
          check(feat_dot_f64(hd, s_rr, r, r, n), "PCG initial norm");
        double rr = 0.0;
        check(feat_read_scalar_f64(hd, s_rr, &rr), "PCG read");
        Status st = this->_set_initial_defect(std::sqrt(rr));
        if(st != Status::progress) return st;
        mg.apply_fused(_vec_p, _vec_r, s_rr, s_g);
        // the update forms the first iterate of the next preconditioner call's first smoother
        typename MGType::FirstIterate fi;
        const bool from1 = mg.first_iterate(_vec_t, fi);
        while(st == Status::progress)
        {
          check(feat_apply_masked_f64(hd, &A, t, p, _filter.mask(), nullptr, p, s_qp, ws), "PCG A*p");
          if(from1)
            check(feat_pcg_update_first_f64(hd, (double*)x.elements(), r, p, t, s_g, s_qp, s_rr, n, fi.inv, fi.mask,
                                            fi.omega, fi.x1, ws),
                  "PCG update");
          else
            check(feat_pcg_update_f64(hd, (double*)x.elements(), r, p, t, s_g, s_qp, s_rr, n, ws), "PCG update");
          st = this->_set_new_defect([&]() {
            check(feat_read_scalar_f64(hd, s_rr, &rr), "PCG read");
            return std::sqrt(rr);
          });
          if(st != Status::progress) return st;
          if(from1)
            mg.apply_fused_from1(_vec_t, _vec_r, s_rr, s_g2);
          else
            mg.apply_fused(_vec_t, _vec_r, s_rr, s_g2);
          check(feat_pcg_direction_f64(hd, p, t, s_g2, s_g, n), "PCG direction");
        }
        return st;
      }
    };

    template<typename Matrix_, typename Filter_>
    std::shared_ptr<PCG<Matrix_, Filter_>> new_pcg(const Matrix_& m, const Filter_& f,
                                                   std::shared_ptr<SolverBase<typename Matrix_::VectorTypeL>> p = nullptr)
    {
      return std::make_shared<PCG<Matrix_, Filter_>>(m, f, p);
    }
  } // namespace Solver
} // namespace feat3amd
