// feat3amd/learned.hpp -- multigrid with a learned finest-level smoother on the
// 3D Q1 anisotropic variable-coefficient diffusion problem.
//
// No counterpart in the reference: the method around the network is the
// reference's geometric multigrid (solver/multigrid.hpp V-cycle, Richardson
// smoothing, restriction/prolongation, Jacobi coarse solve) used as the
// preconditioner of a flexible PCG (pcg.hpp:188-288 with the Polak-Ribiere beta,
// since the learned smoother makes the preconditioner nonlinear) with the plain
// FP64 defect norm as the stopping criterion.  On the finest level the
// smoother is either damped Jacobi (Richardson(JacobiPrecond)) or the learned map
//   x_i += omega * s_i * f(rho / s_i, log-coefficient patch),  rho = r / a_ii over
//   the node's 3x3x3 patch, s_i = max |rho|,
// with f the BF16 MLP of feat3amd.h trained on residual/error pairs that the FP64
// Jacobi multigrid itself produces on the error equation.  The
// learned smoother falls back to Jacobi for the rest of a solve as soon as an
// outer iteration contracts the FP64 defect less than the Jacobi multigrid does
// (its contraction measured while the training data is generated).
#pragma once

#include <feat3amd/batch_assembly.hpp>
#include <feat3amd/solver.hpp>

#include <chrono>
#include <cmath>
#include <memory>
#include <random>
#include <vector>

namespace feat3amd
{
  namespace Solver
  {
    struct LearnedConfig
    {
      int level_fine = 6, level_coarse = 1;
      double sigma = 1.5, kz = 10.0;
      std::uint64_t seed = 28054777172512ull;
      int nu = 4;                 // pre- and post-smoothing steps
      double omega_jac = 0.5;     // Jacobi damping (inv_diag = omega / a_ii)
      double omega_learned = 0.5; // learned smoother damping
      double tol_rel = 1e-8;
      Index max_iter = 200;
      // training data and optimisation
      int n_init = 4, n_snap = 3;
      Index batch = 8192;
      float lr = 1e-3f;
    };

    class LearnedMG
    {
      typedef LAFEM::SparseMatrixCSR<double, Index> Csr;
      typedef LAFEM::DenseVector<double, Index> Vec;

      struct Level
      {
        Csr A, P, R;   // P from the next coarser level (not on the coarsest)
        Vec inv, rhs, sol, def, cor, tmp;
        DeviceArray<std::uint32_t> mask;
      };

      LearnedConfig _cfg;
      std::vector<std::unique_ptr<Level>> _lv;  // 0 = finest
      std::vector<double> _coef;
      Index _m = 0, _n = 0;
      Vec _dinv, _b, _r, _c, _xs;
      DeviceArray<float> _cfeat;
      FusedWorkspace _fw;
      // network
      DeviceArray<float> _p, _gm, _gv, _grad, _partials, _t;
      DeviceArray<std::uint16_t> _wbf, _X;
      DeviceArray<std::uint64_t> _snode;
      DeviceArray<std::uint32_t> _ssnap;
      Index _partials_cap = 0;
      int _adam_step = 0;
      // training data
      DeviceArray<double> _R, _E;
      Index _nsnap = 0;
      double _rho_jacobi = 1.0;
      std::vector<float> _loss;
      DeviceArray<std::uint64_t> _bnd;
      std::vector<std::uint64_t> _bnd_host;

      static std::vector<std::uint32_t> unit_mask(Index n, const std::vector<std::uint64_t>& idx)
      {
        std::vector<std::uint32_t> m((n + 31) / 32, 0u);
        for(auto i : idx) m[i >> 5] |= 1u << (i & 31);
        return m;
      }

      Level& L0() { return *_lv[0]; }

      // finest-level learned smoothing: x = 0, `steps` steps of x += omega s f(patch(rhs - A x))
      void learned_smooth(Vec& x, const Vec& rhs, int steps)
      {
        feat_handle_t h = Runtime::handle();
        Level& L = L0();
        x.format();
        const feat_matrix_f64_t& A = L.A.exec();
        for(int s = 0; s < steps; ++s)
        {
          const double* d = rhs.elements();
          if(s > 0)
          {
            check(feat_defect_f64(h, &A, L.tmp.elements(), rhs.elements(), x.elements(), L.mask.get(), nullptr, &_fw.ws),
                  "learned defect");
            d = L.tmp.elements();
          }
          check(feat_mlp_smooth_f64(h, _m, _p.get(), _wbf.get(), d, _dinv.elements(), _cfeat.get(), _cfg.omega_learned,
                                    x.elements()),
                "learned smoothing");
        }
      }

      // V-cycle: cor = MG(def) (multigrid.hpp:1270-2000 with the fused kernels of solver.hpp).
      // cap_r/cap_e (optional): the finest level's post-smoothing input defect and, for the
      // homogeneous error equation with current iterate cap_x, the error -(cap_x + sol) there.
      void vcycle(Vec& cor, const Vec& def, bool learned, double* cap_r = nullptr, double* cap_e = nullptr,
                  const Vec* cap_x = nullptr)
      {
        feat_handle_t h = Runtime::handle();
        const feat_reduce_ws_t* ws = &_fw.ws;
        const std::size_t nl = _lv.size();
        auto rhs_of = [&](std::size_t i) -> const double* { return i == 0 ? def.elements() : _lv[i]->rhs.elements(); };
        auto sol_of = [&](std::size_t i) -> double* { return i == 0 ? cor.elements() : _lv[i]->sol.elements(); };
        for(std::size_t i = 0; i + 1 < nl; ++i)
        {
          Level& L = *_lv[i];
          const feat_matrix_f64_t& A = L.A.exec();
          if(i == 0 && learned)
          {
            Vec x = Vec::wrap(sol_of(0), _n), r = Vec::wrap(const_cast<double*>(rhs_of(0)), _n);
            learned_smooth(x, r, _cfg.nu);
            check(feat_defect_f64(h, &A, L.def.elements(), rhs_of(0), sol_of(0), L.mask.get(), nullptr, ws), "defect");
          }
          else
            check(feat_rj_smooth_f64(h, &A, L.inv.elements(), L.mask.get(), _cfg.nu, 1.0, rhs_of(i), nullptr, sol_of(i),
                                     L.tmp.elements(), L.def.elements(), nullptr, nullptr, nullptr, ws),
                  "pre-smoothing");
          const feat_matrix_f64_t& R = L.R.exec();
          check(feat_apply_masked_f64(h, &R, _lv[i + 1]->rhs.elements(), L.def.elements(), _lv[i + 1]->mask.get(),
                                      nullptr, nullptr, nullptr, ws),
                "restriction");
        }
        {
          Level& L = *_lv[nl - 1];
          const feat_matrix_f64_t& A = L.A.exec();
          // Jacobi with omega 1 on the coarsest level (its inverse diagonal holds 1/a_ii)
          check(feat_rj_smooth_f64(h, &A, L.cor.elements(), L.mask.get(), 1, 1.0, rhs_of(nl - 1), nullptr, sol_of(nl - 1),
                                   L.tmp.elements(), nullptr, nullptr, nullptr, nullptr, ws),
                "coarse Jacobi");
        }
        for(std::size_t i = nl - 1; i-- > 0;)
        {
          Level& L = *_lv[i];
          const feat_matrix_f64_t& P = L.P.exec();
          check(feat_prol_add_f64(h, &P, sol_of(i), sol_of(i + 1), L.mask.get()), "prolongation");
          const feat_matrix_f64_t& A = L.A.exec();
          check(feat_defect_f64(h, &A, L.def.elements(), rhs_of(i), sol_of(i), L.mask.get(), nullptr, ws), "defect");
          if(i == 0 && cap_r)
          {
            check(feat_copy(h, cap_r, L.def.elements(), _n * sizeof(double)), "capture");
            Vec e = Vec::wrap(cap_e, _n), sol = Vec::wrap(sol_of(0), _n);
            e.axpy(*cap_x, sol, 1.0);  // x + sol
            e.scale(e, -1.0);
          }
          if(i == 0 && learned)
          {
            learned_smooth(L.cor, L.def, _cfg.nu);
            Vec s = Vec::wrap(sol_of(0), _n);
            s.axpy(L.cor, s, 1.0);
          }
          else
            check(feat_rj_smooth_f64(h, &A, L.inv.elements(), L.mask.get(), _cfg.nu, 1.0, L.def.elements(), nullptr,
                                     L.cor.elements(), L.tmp.elements(), nullptr, sol_of(i), nullptr, nullptr, ws),
                  "post-smoothing");
        }
      }

      double defect_norm(Vec& r, const Vec& x, const Vec& b)
      {
        Level& L = L0();
        double* nrm = _fw.scalar(0);
        check(feat_defect_f64(Runtime::handle(), &L.A.exec(), r.elements(), b.elements(), x.elements(), L.mask.get(), nrm,
                              &_fw.ws),
              "defect");
        double v = 0.0;
        check(feat_read_scalar_f64(Runtime::handle(), nrm, &v), "read");
        return std::sqrt(v);
      }

    public:
      struct SolveInfo
      {
        Status status = Status::undefined;
        Index iters = 0;
        bool fell_back = false;
        Index fallback_iter = 0;
        double def_init = 0, def_final = 0, seconds = 0;
        std::vector<double> hist;
      };

      explicit LearnedMG(const LearnedConfig& cfg) : _cfg(cfg) {}

      Index size() const { return _n; }
      Index nodes_1d() const { return _m; }
      double rho_jacobi() const { return _rho_jacobi; }
      const std::vector<float>& loss_history() const { return _loss; }
      float* params() { return _p.get(); }
      std::uint16_t* params_bf16() { return _wbf.get(); }
      const Vec& rhs() const { return _b; }

      void init()
      {
        const int Lf = _cfg.level_fine, Lc = _cfg.level_coarse;
        _coef = Assembly::lognormal_cells(_cfg.seed + 1000u, std::uint64_t(1) << (3 * Lf), _cfg.sigma);
        _lv.clear();
        for(int k = Lf; k >= Lc; --k)
        {
          auto L = std::make_unique<Level>();
          const auto A = Assembly::q1_diffusion_3d(k, _coef, Lf, _cfg.kz);
          L->A = Csr::from_host(A.rows, A.cols, A.val.data(), A.col.data(), A.rowptr.data());
          if(k > Lc)
          {
            const auto P = Assembly::q1_prolongation(3, k);
            const auto R = Assembly::transpose(P);
            L->P = Csr::from_host(P.rows, P.cols, P.val.data(), P.col.data(), P.rowptr.data());
            L->R = Csr::from_host(R.rows, R.cols, R.val.data(), R.col.data(), R.rowptr.data());
          }
          const auto bnd = Assembly::q1_boundary(3, k);
          const auto mk = unit_mask(A.rows, bnd);
          L->mask = DeviceArray<std::uint32_t>::from_host(mk.data(), mk.size());
          for(Vec* v : {&L->inv, &L->rhs, &L->sol, &L->def, &L->cor, &L->tmp}) *v = Vec(A.rows);
          L->A.extract_diag(L->inv);
          L->cor.component_invert(L->inv, 1.0);             // coarsest level: 1 / a_ii
          L->inv.component_invert(L->inv, _cfg.omega_jac);  // omega / a_ii
          (void)L->A.exec();
          if(k > Lc)
          {
            (void)L->P.exec();
            (void)L->R.exec();
          }
          _lv.push_back(std::move(L));
        }
        _m = Assembly::q1_nodes_1d(Lf);
        _n = _m * _m * _m;
        _dinv = Vec(_n);
        L0().A.extract_diag(_dinv);
        _dinv.component_invert(_dinv, 1.0);
        _bnd_host = Assembly::q1_boundary(3, Lf);
        _bnd = DeviceArray<std::uint64_t>::from_host(_bnd_host.data(), _bnd_host.size());
        const auto cf = Assembly::node_log_coef_3d(_coef, Lf);
        _cfeat = DeviceArray<float>::from_host(cf.data(), cf.size());
        _fw.create(8);
        // right-hand side: uniform [-1,1], boundary rows zero
        std::vector<double> b(_n);
        Assembly::random_fill(_cfg.seed + 2000u, b.data(), _n, -1.0, 1.0);
        for(auto i : Assembly::q1_boundary(3, Lf)) b[i] = 0.0;
        _b = Vec::from_host(b);
        _r = Vec(_n);
        _c = Vec(_n);
        // network: He-style uniform initialisation from a fixed seed
        std::vector<float> p(FEAT_MLP_NPARAM, 0.0f);
        std::mt19937_64 g(_cfg.seed + 3000u);
        auto uni = [&](float a) { return float(std::uniform_real_distribution<double>(-a, a)(g)); };
        for(int o = 0; o < 64; ++o)
          for(int i = 0; i < FEAT_MLP_IN; ++i) p[o * 64 + i] = uni(std::sqrt(6.0f / FEAT_MLP_IN));
        for(int i = 0; i < 64 * 64; ++i) p[4160 + i] = uni(std::sqrt(6.0f / 64.0f));
        for(int i = 0; i < 64; ++i) p[8320 + i] = uni(std::sqrt(3.0f / 64.0f));
        set_params(p.data());
        _gm = DeviceArray<float>(FEAT_MLP_NPARAM_PAD);
        _gv = DeviceArray<float>(FEAT_MLP_NPARAM_PAD);
        std::vector<float> z(FEAT_MLP_NPARAM_PAD, 0.0f);
        check(feat_copy(Runtime::handle(), _gm.get(), z.data(), z.size() * 4), "adam m");
        check(feat_copy(Runtime::handle(), _gv.get(), z.data(), z.size() * 4), "adam v");
        _grad = DeviceArray<float>(FEAT_MLP_NPARAM_PAD);
        _partials_cap = 4 * 2048;
        _partials = DeviceArray<float>(_partials_cap * FEAT_MLP_NPARAM_PAD);
        _X = DeviceArray<std::uint16_t>(_cfg.batch * 64);
        _t = DeviceArray<float>(_cfg.batch);
        _snode = DeviceArray<std::uint64_t>(_cfg.batch);
        _ssnap = DeviceArray<std::uint32_t>(_cfg.batch);
        Runtime::synchronize();
      }

      void set_params(const float* host)
      {
        if(_p.empty())
        {
          _p = DeviceArray<float>(FEAT_MLP_NPARAM_PAD);
          _wbf = DeviceArray<std::uint16_t>(2 * 64 * 64);
        }
        check(feat_copy(Runtime::handle(), _p.get(), host, FEAT_MLP_NPARAM * sizeof(float)), "params");
        check(feat_mlp_to_bf16(Runtime::handle(), _p.get(), _wbf.get()), "params bf16");
        _adam_step = 0;
      }
      void get_params(float* host) const { DeviceArray<float>::wrap(_p.get(), FEAT_MLP_NPARAM).to_host(host); }

      /// Flexible PCG preconditioned by the V-cycle (Polak-Ribiere beta, so a nonlinear
      /// preconditioner is admissible; with the linear Jacobi V-cycle it is PCG in exact
      /// arithmetic).  learned = false: the Jacobi multigrid.  The FP64 defect norm decides
      /// convergence; with allow_fallback the learned smoother is dropped for the rest of the
      /// solve when an iteration contracts the defect less than rho_jacobi.
      SolveInfo solve(Vec& x, const Vec& b, bool learned, bool allow_fallback = true)
      {
        SolveInfo info;
        Runtime::synchronize();
        const auto t0 = std::chrono::steady_clock::now();
        Vec z(_n), p(_n), q(_n), rold(_n);
        double d = defect_norm(_r, x, b);
        info.def_init = d;
        info.hist.push_back(d);
        bool use_learned = learned;
        const double eps = 2.220446049250313e-16;
        info.status = std::isfinite(d) ? (d <= eps * eps ? Status::success : Status::progress) : Status::aborted;
        if(info.status == Status::progress)
        {
          vcycle(z, _r, use_learned);
          p.copy(z);
          double gamma = _r.dot(z);
          while(info.status == Status::progress)
          {
            L0().A.apply(q, p);
            check(feat_unit_filter_def_f64_u64(Runtime::handle(), q.elements(), _bnd.get(), _bnd.size()), "filter");
            const double alpha = gamma / p.dot(q);
            x.axpy(p, x, alpha);
            rold.copy(_r);
            _r.axpy(q, _r, -alpha);
            const double dn = _r.norm2();
            ++info.iters;
            info.hist.push_back(dn);
            if(use_learned && allow_fallback && (!std::isfinite(dn) || dn > _rho_jacobi * d))
            {
              use_learned = false;
              info.fell_back = true;
              info.fallback_iter = info.iters;
            }
            d = dn;
            if(!std::isfinite(d))
              info.status = Status::aborted;
            else if(d <= _cfg.tol_rel * info.def_init)
              info.status = Status::success;
            else if(info.iters >= _cfg.max_iter)
              info.status = Status::max_iter;
            if(info.status != Status::progress) break;
            vcycle(z, _r, use_learned);
            const double gnew = _r.dot(z);
            const double beta = (gnew - rold.dot(z)) / gamma;
            gamma = gnew;
            p.axpy(p, z, beta);
          }
        }
        Runtime::synchronize();
        info.def_final = d;
        info.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return info;
      }

      /// Residual/error pairs produced by the FP64 Jacobi multigrid itself on the error
      /// equation A e = 0 from random (white-noise) starts: the start (r = -A x, e = -x) and,
      /// per V-cycle, the finest level's post-smoothing input after the coarse-grid
      /// correction -- the rough error a fine-level smoother has to remove.  Also measures
      /// the Jacobi multigrid's contraction (flexible PCG on the right-hand side).
      void generate_data()
      {
        const Index n = _n;
        Vec xj(n, 0.0);
        _rho_jacobi = 1.0;
        const auto ref = solve(xj, _b, false, false);
        _rho_jacobi = std::pow(ref.def_final / ref.def_init, 1.0 / double(std::max<Index>(ref.iters, 1)));
        _nsnap = Index(_cfg.n_init) * Index(_cfg.n_snap);
        _R = DeviceArray<double>(_nsnap * n);
        _E = DeviceArray<double>(_nsnap * n);
        Vec zero(n, 0.0);
        Index k = 0;
        for(int s = 0; s < _cfg.n_init; ++s)
        {
          std::vector<double> x0(n);
          Assembly::random_fill(_cfg.seed + 4000u + std::uint64_t(s), x0.data(), n, -1.0, 1.0);
          for(Index b = 0; b < _bnd.size(); ++b) x0[_bnd_host[b]] = 0.0;
          Vec x = Vec::from_host(x0);
          {
            Vec r = Vec::wrap(_R.get() + k * n, n), e = Vec::wrap(_E.get() + k * n, n);
            (void)defect_norm(r, x, zero);
            e.scale(x, -1.0);
            ++k;
          }
          for(int it = 1; it < _cfg.n_snap; ++it, ++k)
          {
            (void)defect_norm(_r, x, zero);
            vcycle(_c, _r, false, _R.get() + k * n, _E.get() + k * n, &x);
            x.axpy(_c, x, 1.0);
          }
        }
        Runtime::synchronize();
      }

      /// Adam training on random (node, snapshot) samples; loss recorded every step
      void train(int steps, std::uint64_t seed)
      {
        if(_nsnap == 0) generate_data();
        feat_handle_t h = Runtime::handle();
        std::mt19937_64 g(seed);
        std::uniform_int_distribution<std::uint64_t> du(1, _m - 2);
        std::uniform_int_distribution<std::uint32_t> ds(0, std::uint32_t(_nsnap - 1));
        std::vector<std::uint64_t> nodes(_cfg.batch);
        std::vector<std::uint32_t> snaps(_cfg.batch);
        std::vector<float> lossv(FEAT_MLP_NPARAM_PAD);
        for(int st = 0; st < steps; ++st)
        {
          for(Index k = 0; k < _cfg.batch; ++k)
          {
            nodes[k] = (du(g) * _m + du(g)) * _m + du(g);
            snaps[k] = ds(g);
          }
          check(feat_copy(h, _snode.get(), nodes.data(), nodes.size() * 8), "samples");
          check(feat_copy(h, _ssnap.get(), snaps.data(), snaps.size() * 4), "samples");
          check(feat_mlp_features(h, _m, _R.get(), _E.get(), _dinv.elements(), _cfeat.get(), _snode.get(), _ssnap.get(),
                                  _cfg.batch, _X.get(), _t.get()),
                "features");
          check(feat_mlp_grad(h, _p.get(), _wbf.get(), _X.get(), _t.get(), _cfg.batch, _partials.get(), _partials_cap,
                              _grad.get()),
                "grad");
          float loss = 0.0f;
          check(feat_copy(h, &loss, _grad.get() + FEAT_MLP_NPARAM_PAD - 1, 4), "loss");
          check(feat_mlp_adam(h, _p.get(), _wbf.get(), _grad.get(), _gm.get(), _gv.get(), ++_adam_step, _cfg.lr, 0.9f,
                              0.999f, 1e-8f),
                "adam");
          Runtime::synchronize();
          _loss.push_back(loss / float(_cfg.batch));
        }
      }
    };
  } // namespace Solver
} // namespace feat3amd
