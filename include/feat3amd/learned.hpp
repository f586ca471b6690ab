// feat3amd/learned.hpp -- multigrid with a learned finest-level smoother on the
// 3D Q1 anisotropic variable-coefficient diffusion problem.
//
// No counterpart in the reference: the method around the network is the
// reference's geometric multigrid (solver/multigrid.hpp V-cycle, Richardson
// smoothing, restriction/prolongation, Jacobi coarse solve) used as the
// preconditioner of a truncated flexible CG (the PCG of pcg.hpp:188-288 with each new
// direction A-orthogonalised against the last fcg_m ones, since the learned smoother makes
// the preconditioner nonlinear) with the plain FP64 defect norm as the stopping criterion.  On the finest level the
// smoother is either damped Jacobi (Richardson(JacobiPrecond)) or damped Jacobi plus a
// learned correction,
//   x_i += omega * (omega_jac rho_i + s_i f(rho / s_i, log-coefficient patch)),
//   rho = r / a_ii over the node's 3x3x3 patch, s_i = max |rho|,
// with f the BF16 MLP of feat3amd.h (output layer initialised to zero, so the untrained
// smoother is Jacobi), each step taken with the energy-optimal length along its direction
// (so no smoothing step can increase the error's energy norm: the V-cycle stays
// energy-stable whatever the network outputs), trained on residual/error fields that the
// FP64 Jacobi multigrid
// itself produces on the error equation -- the input of every finest-level smoothing step,
// first with the Jacobi smoother, then (on-policy refresh) with the learned one -- to
// minimise the energy norm of the error one learned step leaves, relative to the energy
// it started from.  The learned smoother falls back to Jacobi for the rest of a solve as
// soon as its FP64 defect falls a factor 3 behind the Jacobi multigrid's defect history
// at the same iteration (recorded on the same right-hand side before training), or the
// defect grows past its initial value.
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
      double omega_learned = 1.0; // learned smoother damping (x += omega (omega_jac r/a_ii + s f))
      double tol_rel = 1e-8;
      Index max_iter = 200;
      int fcg_m = 4;              // flexible CG: previous directions each new one is A-orthogonalised to
      // training data (n_init random starts x n_snap V-cycles x 2 nu smoothing steps) and optimisation
      int n_init = 4, n_snap = 2;
      Index batch = 8192;
      float lr = 1e-3f;
      // learned steps per finest-level smoother call (1..nu); the remaining nu - learned_steps
      // are damped Jacobi steps (the fused smoother on the current defect)
      int learned_steps = 4;
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
      Vec _dinv, _b, _bval, _r, _c, _xs;
      DeviceArray<float> _cfeat;
      FusedWorkspace _fw;
      // network
      DeviceArray<float> _p, _gm, _gv, _grad, _partials, _t;
      DeviceArray<std::uint16_t> _wbf, _X;   // _X: inputs of every node of every snapshot
      DeviceArray<float> _p0;                 // zero network: the learned step is damped Jacobi
      DeviceArray<std::uint16_t> _wbf0;
      DeviceArray<float> _y;
      DeviceArray<double> _S, _Rho, _ep, _q;
      Vec _cv, _cq;   // learned step direction c and A c
      Vec _jy, _jt;   // Jacobi steps after the learned ones: their correction and scratch
      Index _partials_cap = 0;
      int _adam_step = 0;
      // training data
      DeviceArray<double> _R, _E;
      Index _nsnap = 0;
      int _refresh = 0;
      std::vector<double> _enorm;     // e^T A e of every snapshot
      double _rho_jacobi = 1.0;
      std::vector<double> _jac_rel;   // Jacobi multigrid defect history / initial defect
      // capture of the finest smoothing steps' inputs (training data generation)
      double *_cap_r = nullptr, *_cap_e = nullptr;
      Index _cap_next = 0, _cap_max = 0;
      const Vec* _cap_x = nullptr;
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

      // finest-level learned smoothing: x = 0, d = rhs, `steps` steps of
      //   c = omega (omega_jac rho + s f(patch(d))),  x += alpha c,  d -= alpha A c,
      // alpha = c.d / c.Ac (energy-optimal); net = false: zero network (Jacobi direction).
      // While capturing, every step's input defect and the error of the finest iterate,
      // -(cap_x + base + x), are stored (the homogeneous error equation with iterate cap_x,
      // base: the finest solution before these steps, nullptr for zero).
      void learned_smooth(Vec& x, const Vec& rhs, int steps, bool net, const Vec* base)
      {
        feat_handle_t h = Runtime::handle();
        Level& L = L0();
        x.format();
        Vec d = Vec::wrap(L.tmp.elements(), _n);
        d.copy(rhs);
        const feat_matrix_f64_t& A = L.A.exec();
        double *ca = _fw.scalar(4), *cb = _fw.scalar(5);
        for(int s = 0; s < steps; ++s)
        {
          if(_cap_r && _cap_next < _cap_max)
          {
            check(feat_copy(h, _cap_r + _cap_next * _n, d.elements(), _n * sizeof(double)), "capture");
            Vec e = Vec::wrap(_cap_e + _cap_next * _n, _n);
            e.axpy(x, *_cap_x, 1.0);
            if(base) e.axpy(*base, e, 1.0);
            e.scale(e, -1.0);
            ++_cap_next;
          }
          check(feat_mlp_direction_f64(h, _m, net ? _p.get() : _p0.get(), net ? _wbf.get() : _wbf0.get(), d.elements(),
                                       _dinv.elements(), _cfeat.get(), _cfg.omega_learned, _cfg.omega_jac, _cv.elements(),
                                       ca, &_fw.ws),
                "learned direction");
          check(feat_apply_masked_f64(h, &A, _cq.elements(), _cv.elements(), L.mask.get(), nullptr, _cv.elements(), cb,
                                      &_fw.ws),
                "learned A c");
          check(feat_mlp_step_f64(h, _n, x.elements(), d.elements(), _cv.elements(), _cq.elements(), ca, cb),
                "learned step");
          // the trained network on its first learned_steps steps only: the rest are damped
          // Jacobi steps from the current defect d (Richardson with the Jacobi preconditioner
          // on A y = d, then x += y -- continuing the smoother from x)
          if(net && !_cap_r && s + 1 == _cfg.learned_steps && s + 1 < steps)
          {
            check(feat_rj_smooth_f64(h, &A, L.inv.elements(), L.mask.get(), steps - s - 1, 1.0, d.elements(), nullptr,
                                     _jy.elements(), _jt.elements(), nullptr, nullptr, nullptr, nullptr, &_fw.ws),
                  "Jacobi steps after the learned ones");
            x.axpy(_jy, x, 1.0);
            break;
          }
        }
      }

      // V-cycle: cor = MG(def) (multigrid.hpp:1270-2000 with the fused kernels of solver.hpp).
      // Finest-level smoother: mode 0 damped Jacobi (fused kernels), 1 learned, 2 the learned
      // step with the zero network (damped Jacobi through the learned path, for capture).
      void vcycle(Vec& cor, const Vec& def, int mode)
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
          if(i == 0 && mode != 0)
          {
            Vec x = Vec::wrap(sol_of(0), _n), r = Vec::wrap(const_cast<double*>(rhs_of(0)), _n);
            learned_smooth(x, r, _cfg.nu, mode == 1, nullptr);
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
          if(i == 0 && mode != 0)
          {
            Vec s = Vec::wrap(sol_of(0), _n);
            learned_smooth(L.cor, L.def, _cfg.nu, mode == 1, &s);
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
      /// relative defect of the Jacobi multigrid after k iterations (recorded history,
      /// extended by its mean contraction)
      double jacobi_rel(Index k) const
      {
        if(_jac_rel.empty()) return 1.0;
        if(k < _jac_rel.size()) return _jac_rel[k];
        return _jac_rel.back() * std::pow(_rho_jacobi, double(k + 1 - _jac_rel.size()));
      }
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
        _cv = Vec(_n);
        _jy = Vec(_n);
        _jt = Vec(_n);
        _cq = Vec(_n);
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
        // a second right-hand side (seed + 2500) for model selection in fit()
        Assembly::random_fill(_cfg.seed + 2500u, b.data(), _n, -1.0, 1.0);
        for(auto i : Assembly::q1_boundary(3, Lf)) b[i] = 0.0;
        _bval = Vec::from_host(b);
        _r = Vec(_n);
        _c = Vec(_n);
        // network: He-style uniform initialisation from a fixed seed
        std::vector<float> p(FEAT_MLP_NPARAM, 0.0f);
        std::mt19937_64 g(_cfg.seed + 3000u);
        auto uni = [&](float a) { return float(std::uniform_real_distribution<double>(-a, a)(g)); };
        for(int o = 0; o < 64; ++o)
          for(int i = 0; i < FEAT_MLP_IN; ++i) p[o * 64 + i] = uni(std::sqrt(6.0f / FEAT_MLP_IN));
        for(int i = 0; i < 64 * 64; ++i) p[4160 + i] = uni(std::sqrt(6.0f / 64.0f));
        // output layer zero: the untrained learned smoother is exactly damped Jacobi
        set_params(p.data());
        _p0 = DeviceArray<float>(FEAT_MLP_NPARAM_PAD);
        _wbf0 = DeviceArray<std::uint16_t>(2 * 64 * 64);
        {
          std::vector<float> z0(FEAT_MLP_NPARAM_PAD, 0.0f);
          check(feat_copy(Runtime::handle(), _p0.get(), z0.data(), z0.size() * 4), "zero network");
          check(feat_mlp_to_bf16(Runtime::handle(), _p0.get(), _wbf0.get()), "zero network bf16");
        }
        _gm = DeviceArray<float>(FEAT_MLP_NPARAM_PAD);
        _gv = DeviceArray<float>(FEAT_MLP_NPARAM_PAD);
        std::vector<float> z(FEAT_MLP_NPARAM_PAD, 0.0f);
        check(feat_copy(Runtime::handle(), _gm.get(), z.data(), z.size() * 4), "adam m");
        check(feat_copy(Runtime::handle(), _gv.get(), z.data(), z.size() * 4), "adam v");
        _grad = DeviceArray<float>(FEAT_MLP_NPARAM_PAD);
        _partials_cap = 4 * 2048;
        _partials = DeviceArray<float>(_partials_cap * FEAT_MLP_NPARAM_PAD);
        _t = DeviceArray<float>(_n);
        _y = DeviceArray<float>(_n);
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

      /// Flexible CG preconditioned by the V-cycle: each new direction is made A-orthogonal
      /// to the last fcg_m ones (truncated FCG, so a nonlinear preconditioner is admissible;
      /// with the linear Jacobi V-cycle it is PCG in exact arithmetic).  learned = false: the Jacobi multigrid.  The FP64 defect norm decides
      /// convergence; with allow_fallback the learned smoother is dropped for the rest of the
      /// solve when its mean contraction falls behind rho_jacobi.
      SolveInfo solve(Vec& x, const Vec& b, bool learned, bool allow_fallback = true, Index max_iter = 0)
      {
        const Index iter_cap = max_iter ? max_iter : _cfg.max_iter;
        SolveInfo info;
        Runtime::synchronize();
        const auto t0 = std::chrono::steady_clock::now();
        // ring of the last m directions plus the new one
        const int mfcg = std::max(1, _cfg.fcg_m), slots = mfcg + 1;
        std::vector<Vec> P, Q;
        std::vector<double> dpq;
        for(int j = 0; j < slots; ++j)
        {
          P.emplace_back(_n);
          Q.emplace_back(_n);
        }
        dpq.assign(slots, 0.0);
        Vec z(_n);
        double d = defect_norm(_r, x, b);
        info.def_init = d;
        info.hist.push_back(d);
        bool use_learned = learned;
        const double eps = 2.220446049250313e-16;
        info.status = std::isfinite(d) ? (d <= eps * eps ? Status::success : Status::progress) : Status::aborted;
        Index k = 0;
        while(info.status == Status::progress)
        {
          vcycle(z, _r, use_learned ? 1 : 0);
          // p_k = z - sum over the last m directions of (z . A p_j / p_j . A p_j) p_j
          Vec& pk = P[k % slots];
          Vec& qk = Q[k % slots];
          const Index jlo = k >= Index(mfcg) ? k - Index(mfcg) : 0;
          std::vector<double> beta;
          for(Index j = jlo; j < k; ++j) beta.push_back(dpq[j % slots] > 0.0 ? z.dot(Q[j % slots]) / dpq[j % slots] : 0.0);
          pk.copy(z);
          for(Index j = jlo; j < k; ++j) pk.axpy(P[j % slots], pk, -beta[j - jlo]);
          L0().A.apply(qk, pk);
          check(feat_unit_filter_def_f64_u64(Runtime::handle(), qk.elements(), _bnd.get(), _bnd.size()), "filter");
          dpq[k % slots] = pk.dot(qk);
          const double alpha = dpq[k % slots] > 0.0 ? pk.dot(_r) / dpq[k % slots] : 0.0;
          x.axpy(pk, x, alpha);
          _r.axpy(qk, _r, -alpha);
          ++k;
          const double dn = _r.norm2();
          ++info.iters;
          info.hist.push_back(dn);
          // contraction check against the Jacobi multigrid's relative defect at the same
          // iteration (CG defect norms are not monotone, hence the factor 3); immediate
          // fallback on growth past the initial defect or a non-finite defect
          if(use_learned && allow_fallback &&
             (!std::isfinite(dn) || dn > info.def_init || (info.iters >= 2 && dn > 3.0 * info.def_init * jacobi_rel(info.iters))))
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
          else if(info.iters >= iter_cap)
            info.status = Status::max_iter;
        }
        Runtime::synchronize();
        info.def_final = d;
        info.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return info;
      }

      /// Residual/error pairs of the multigrid itself on the error equation A e = 0 from
      /// random (white-noise) starts: the input of every finest-level smoothing step (pre and
      /// post) of n_snap V-cycles per start, with the Jacobi smoother (on_policy = false) or
      /// the current learned one (on-policy refresh, new starts).  The first call also
      /// records the Jacobi multigrid's defect history on the right-hand side (flexible CG).
      void generate_data(bool on_policy = false)
      {
        const Index n = _n;
        if(_jac_rel.empty())
        {
          Vec xj(n, 0.0);
          const auto ref = solve(xj, _b, false, false);
          _rho_jacobi = std::pow(ref.def_final / ref.def_init, 1.0 / double(std::max<Index>(ref.iters, 1)));
          for(double d : ref.hist) _jac_rel.push_back(d / ref.def_init);
        }
        const Index nsnap = Index(_cfg.n_init) * Index(_cfg.n_snap) * 2 * Index(_cfg.nu);
        if(_nsnap != nsnap)
        {
          _nsnap = nsnap;
          _R = DeviceArray<double>(_nsnap * n);
          _E = DeviceArray<double>(_nsnap * n);
          _X = DeviceArray<std::uint16_t>(_nsnap * n * 64);
          _S = DeviceArray<double>(_nsnap * n);
          _Rho = DeviceArray<double>(_nsnap * n);
          _ep = DeviceArray<double>(n);
          _q = DeviceArray<double>(n);
        }
        Vec zero(n, 0.0);
        _cap_r = _R.get();
        _cap_e = _E.get();
        _cap_next = 0;
        _cap_max = _nsnap;
        for(int s = 0; s < _cfg.n_init; ++s)
        {
          std::vector<double> x0(n);
          Assembly::random_fill(_cfg.seed + 4000u + std::uint64_t(s) + 100u * std::uint64_t(_refresh), x0.data(), n, -1.0,
                                1.0);
          for(Index b = 0; b < _bnd.size(); ++b) x0[_bnd_host[b]] = 0.0;
          Vec x = Vec::from_host(x0);
          _cap_x = &x;
          for(int it = 0; it < _cfg.n_snap; ++it)
          {
            (void)defect_norm(_r, x, zero);
            vcycle(_c, _r, on_policy ? 1 : 2);
            x.axpy(_c, x, 1.0);
          }
          Runtime::synchronize();
        }
        _cap_r = _cap_e = nullptr;
        _cap_x = nullptr;
        ++_refresh;
        // network inputs of every node of every snapshot (fixed during training) and the
        // energy each snapshot starts from
        feat_handle_t h = Runtime::handle();
        Level& L = L0();
        _enorm.assign(_nsnap, 0.0);
        for(Index j = 0; j < _nsnap; ++j)
        {
          check(feat_mlp_features_full(h, _m, _R.get() + j * n, _dinv.elements(), _cfeat.get(), _X.get() + j * n * 64,
                                       _S.get() + j * n, _Rho.get() + j * n),
                "features");
          check(feat_apply_masked_f64(h, &L.A.exec(), _q.get(), _E.get() + j * n, L.mask.get(), nullptr, _E.get() + j * n,
                                      _fw.scalar(1), &_fw.ws),
                "A e");
          check(feat_read_scalar_f64(h, _fw.scalar(1), &_enorm[j]), "energy");
        }
        Runtime::synchronize();
      }

      /// energy left by one learned step (energy-optimal length) on snapshot k relative to the
      /// snapshot's energy: with c the step direction and r = A e,
      ///   ep = e - alpha c, alpha = c.r / c.Ac,  ep^T A ep / e^T A e = 1 - (c.r)^2 / (c.Ac e^T A e);
      /// with grad: the training target of its gradient (alpha is optimal, so the gradient at
      /// fixed alpha is the full one: -2 alpha s (r - alpha A c) / e^T A e per output)
      double step_energy(Index k, bool grad)
      {
        feat_handle_t h = Runtime::handle();
        Level& L = L0();
        const feat_matrix_f64_t& A = L.A.exec();
        const Index n = _n;
        const std::uint16_t* X = _X.get() + k * n * 64;
        const double *S = _S.get() + k * n, *Rho = _Rho.get() + k * n, *R = _R.get() + k * n;
        const double e0 = _enorm[k] > 0.0 ? _enorm[k] : 1.0;
        check(feat_mlp_infer(h, _p.get(), _wbf.get(), X, n, _y.get()), "infer");
        // ep = -c (e = 0), q = A ep = -A c, b = c.Ac, a' = ep.r = -c.r
        check(feat_mlp_energy(h, n, 0, _cfg.omega_jac, _y.get(), S, Rho, nullptr, nullptr, _ep.get(), nullptr), "energy 0");
        check(feat_apply_masked_f64(h, &A, _q.get(), _ep.get(), L.mask.get(), nullptr, _ep.get(), _fw.scalar(1), &_fw.ws),
              "A c");
        check(feat_dot_f64(h, _fw.scalar(2), _ep.get(), R, n), "c.r");
        double b = 0.0, am = 0.0;
        check(feat_read_scalar_f64(h, _fw.scalar(1), &b), "c.Ac");
        check(feat_read_scalar_f64(h, _fw.scalar(2), &am), "c.r");
        const double a = -am;
        const double alpha = (b > 0.0 && std::isfinite(a / b)) ? a / b : 0.0;
        if(grad)
        {
          // r_new = r - alpha A c = r + alpha q, into ep
          Vec rn = Vec::wrap(_ep.get(), n), q = Vec::wrap(_q.get(), n), r = Vec::wrap(const_cast<double*>(R), n);
          rn.axpy(q, r, alpha);
          check(feat_mlp_energy(h, n, 1, 2.0 * double(n) * alpha / e0, _y.get(), S, nullptr, nullptr, _ep.get(), nullptr,
                                _t.get()),
                "energy 1");
        }
        return b > 0.0 ? 1.0 - a * alpha / e0 : 1.0;
      }

      /// mean over the snapshots of the relative energy one learned step leaves
      double energy()
      {
        if(_nsnap == 0) generate_data();
        double sum = 0.0;
        for(Index k = 0; k < _nsnap; ++k) sum += step_energy(k, false);
        return sum / double(_nsnap);
      }

      struct FitInfo
      {
        int best_round = -1;          // -1: no round beat the Jacobi multigrid
        Index jacobi_iters = 0, best_iters = 0;
        std::vector<Index> round_iters;   // validation iterations per round (0: no convergence)
      };

      /// Training with model selection: `rounds` rounds of `steps` Adam steps (from the
      /// second round on, after an on-policy data refresh when on_policy), each followed by a
      /// validation solve (flexible CG, learned smoother, no fallback) on a second
      /// right-hand side; the parameters of the round with the fewest validation iterations
      /// are kept when it beats the Jacobi multigrid there, else the initial parameters.
      FitInfo fit(int rounds, int steps, bool on_policy, std::uint64_t seed)
      {
        if(_nsnap == 0) generate_data();
        FitInfo fi;
        Vec x(_n, 0.0);
        const auto jac = solve(x, _bval, false, false);
        fi.jacobi_iters = fi.best_iters = jac.iters;
        std::vector<float> best(FEAT_MLP_NPARAM);
        get_params(best.data());
        for(int r = 0; r < rounds; ++r)
        {
          if(r > 0 && on_policy) generate_data(true);
          train(steps, seed + std::uint64_t(r));
          x.format();
          const auto li = solve(x, _bval, true, false, 2 * jac.iters);
          const Index it = li.status == Status::success ? li.iters : 0;
          fi.round_iters.push_back(it);
          if(it > 0 && it < fi.best_iters)
          {
            fi.best_iters = it;
            fi.best_round = r;
            get_params(best.data());
          }
        }
        set_params(best.data());
        return fi;
      }

      /// Adam training of the learned correction on whole snapshot fields: one random
      /// snapshot per step, objective = the relative energy ep^T A ep / e^T A e of the error
      /// ep left by one learned smoothing step (ep = e - (omega_jac r/a_ii + s f)); its value
      /// is recorded
      void train(int steps, std::uint64_t seed)
      {
        if(_nsnap == 0) generate_data();
        feat_handle_t h = Runtime::handle();
        const Index n = _n;
        std::mt19937_64 g(seed);
        std::uniform_int_distribution<Index> ds(0, _nsnap - 1);
        for(int st = 0; st < steps; ++st)
        {
          const Index k = ds(g);
          const double rel = step_energy(k, true);
          check(feat_mlp_grad(h, _p.get(), _wbf.get(), _X.get() + k * n * 64, _t.get(), n, _partials.get(), _partials_cap,
                              _grad.get()),
                "grad");
          check(feat_mlp_adam(h, _p.get(), _wbf.get(), _grad.get(), _gm.get(), _gv.get(), ++_adam_step, _cfg.lr, 0.9f,
                              0.999f, 1e-8f),
                "adam");
          _loss.push_back(float(rel));
        }
      }
    };
  } // namespace Solver
} // namespace feat3amd
