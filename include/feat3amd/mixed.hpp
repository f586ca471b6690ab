// feat3amd/mixed.hpp -- mixed-precision multigrid preconditioner.
//
// The ConvertPrecond pattern (solver/convert_precond.hpp:100-119): the FP64
// defect is converted to FP32, an FP32 V-cycle is applied and the correction
// is converted back.  Here the V-cycle runs on FP32 SELL-64 plans of the same
// matrices (values rounded once), FP32 Jacobi diagonals (float(omega) divided by
// the FP32 diagonal in FP32, as JacobiPrecond<float> forms them,
// jacobi_precond.hpp:136-143) and FP32 vectors, while the coarse-grid solve stays
// FP64 (the dense inverse of DirectCoarseSolver applied in FP64).  Used as the
// preconditioner of an FP64 PCG, the outer residual, solution update and
// convergence test stay FP64: iterative refinement with Krylov acceleration.
#pragma once

#include <feat3amd/solver.hpp>

namespace feat3amd
{
  namespace Solver
  {
    template<typename Matrix_, typename Filter_, typename Transfer_>
    class MixedMultiGrid : public SolverBase<typename Matrix_::VectorTypeL>,
                           public FusedPrecond<typename Matrix_::VectorTypeL>
    {
    public:
      typedef typename Matrix_::VectorTypeL VectorType;
      typedef LAFEM::DenseVector<float, Index> Vector32;
      typedef MultiGridHierarchy<Matrix_, Filter_, Transfer_> HierarchyType;
      typedef Richardson<Matrix_, Filter_> RichType;
      typedef JacobiPrecond<Matrix_, Filter_> JacType;
      typedef DirectCoarseSolver<Matrix_, Filter_> DirectType;

    private:
      struct Level32
      {
        Vector32 rhs, sol, def, cor, tmp, inv_pre, inv_post;
        int pre_steps = 0, post_steps = 0;
        float omega_pre = 1.0f, omega_post = 1.0f;
      };
      std::shared_ptr<HierarchyType> _h;
      std::vector<Level32> _l;
      FusedWorkspace _fw;
      enum class Coarse { direct, jacobi, richardson, none } _coarse = Coarse::none;
      bool _ok = false;
      bool _half = false;
      bool _tail_on = true, _tail_built = false;
      Index _tail_t = 0;   // first level of the fused coarse part; 0: none
      feat_mg_tail_f32_t _tail{};
      std::vector<TailGridForm> _grid;              // per level: grid form if it is fused
      std::vector<DeviceArray<float>> _grid_sten;   // its FP32 stencil (rounded as the FP32 plan rounds)
      std::vector<feat_matrix_f32_t> _dl_P, _dl_R;  // d-linear transfer descriptors (where dlinear)

      /// The grid form of every level that the fused coarse part could hold (host copies of
      /// its matrices): at init_numeric, never inside a recorded graph.
      void _scan_grid()
      {
        const Index L = _h->size();
        _grid.assign(L, TailGridForm{});
        _grid_sten.assign(L, DeviceArray<float>());
        _dl_P.assign(L, feat_matrix_f32_t{});
        _dl_R.assign(L, feat_matrix_f32_t{});
        for(Index s = 0; s + 1 < L; ++s)
        {
          auto& lv = _h->levels[s];
          if(!lv.transfer || !lv.filter) continue;
          const bool cand = s >= 1 && lv.matrix->rows() <= FEAT_MG_TAIL_MAX_ROWS;
          _grid[s] = tail_grid_form(*lv.matrix, lv.filter->mask(), lv.transfer->get_mat_prol(),
                                    lv.transfer->get_mat_rest(), _h->levels[s + 1].matrix->rows(), cand);
          if(!_grid[s].sten.empty())
          {
            std::vector<float> s32(_grid[s].sten.begin(), _grid[s].sten.end());
            _grid_sten[s] = DeviceArray<float>::from_host(s32.data(), s32.size());
          }
          // the FP16-valued cycle too: its transfer plans store the d-linear weights w as w/2
          // with the exact scale 2, and power-of-two scaling commutes with rounding, so the
          // matrix-free rows give its row sums (outside the subnormal range)
          if(_grid[s].dlinear)
          {
            check(feat_matrix_dlinear_f32(_grid[s].dim, _grid[s].n1, 0, &_dl_P[s]), "MixedMultiGrid d-linear prolongation");
            check(feat_matrix_dlinear_f32(_grid[s].dim, _grid[s].n1, 1, &_dl_R[s]), "MixedMultiGrid d-linear restriction");
          }
        }
      }
      /// the FP32 transfer descriptors of level i: d-linear where checked, else the plans
      const feat_matrix_f32_t& _rest(Index i) const
      {
        return (i < _grid.size() && _grid[i].dlinear) ? _dl_R[i] : plan(_h->levels[i].transfer->get_mat_rest());
      }
      const feat_matrix_f32_t& _prol(Index i) const
      {
        return (i < _grid.size() && _grid[i].dlinear) ? _dl_P[i] : plan(_h->levels[i].transfer->get_mat_prol());
      }

      // execution plan of a level's matrix: FP32 values or FP16 values (FP32 arithmetic)
      const feat_matrix_f32_t& plan(const Matrix_& m) const { return _half ? m.exec16() : m.exec32(); }

      static Vector32 to32(const VectorType& v)
      {
        Vector32 r(v.size());
        check(feat_convert_f64_f32(Runtime::handle(), r.elements(), (const double*)v.elements(), v.size()), "to32");
        return r;
      }

      /// JacobiPrecond<float>::init_numeric on the FP32 copy of the matrix: its
      /// diagonal (each a_ii rounded once), then float(omega) / d_ii in FP32
      template<typename Jac_>
      static Vector32 inv_diag32(const Jac_& jac)
      {
        VectorType d(jac.matrix().rows());
        jac.matrix().extract_diag(d);
        Vector32 r = to32(d);
        check(feat_component_invert_f32(Runtime::handle(), r.elements(), r.elements(), float(jac.omega()), r.size()),
              "inv_diag32");
        return r;
      }

      /// The coarse levels that run as one LDS-resident workgroup (feat_mg_tail_f32): the
      /// longest run ending at the coarsest level, starting at level 1 or below, whose
      /// vectors fit the LDS budget, with FP32 diagonal-dictionary plans, FP32-valued transfer
      /// plans (an FP16-valued cycle uses FP32 plans where the matrix has a dictionary plan)
      /// and a direct coarse solve.
      void _build_tail()
      {
        _tail_built = true;
        _tail_t = 0;
        const Index L = _h->size(), c = L - 1;
        if(!_tail_on || _coarse != Coarse::direct || L < 3 || _h->levels[c].matrix->rows() > FEAT_MG_TAIL_MAX_ROWS)
          return;
        uint64_t words = 2 * _h->levels[c].matrix->rows();
        Index t = c;
        for(Index s = c; s-- > 1;)
        {
          const auto& l = _h->levels[s];
          const feat_matrix_f32_t& A = plan(*l.matrix);
          const bool f16 = plan(l.transfer->get_mat_rest()).value_kind == FEAT_VAL_F16 ||
                           plan(l.transfer->get_mat_prol()).value_kind == FEAT_VAL_F16;
          const uint64_t need = words + 5 * l.matrix->rows() + 256;   // vectors and the pair table
          if(f16 || A.index != FEAT_IDX_DIAG || A.slice_words > 8 || need * sizeof(float) > FEAT_MG_TAIL_MAX_LDS ||
             c - s + 1 > FEAT_MG_TAIL_MAX_LEVELS || l.matrix->rows() > FEAT_MG_TAIL_MAX_ROWS)
            break;
          words = need;
          t = s;
        }
        if(t == c) return;
        _tail = feat_mg_tail_f32_t{};
        _tail.nlevels = int(c - t + 1);
        for(Index s = t; s <= c; ++s)
        {
          auto& lv = _h->levels[s];
          feat_mg_tail_level_f32_t& d = _tail.lv[s - t];
          d.n = lv.matrix->rows();
          d.mask = lv.filter->mask();
          if(s == c) break;
          const feat_matrix_f32_t& A = plan(*lv.matrix);
          d.table = static_cast<const float*>(A.val);
          d.codes = static_cast<const uint32_t*>(A.col);
          d.words = A.slice_words;
          // one inverted diagonal serves both smoothers: their FP64 sources must be one vector
          auto* pre = static_cast<RichType*>(lv.pre.get());
          auto* post = static_cast<RichType*>(lv.post.get());
          if(post->fusable_jacobi()->inv_diag().elements() != pre->fusable_jacobi()->inv_diag().elements()) return;
          d.inv_diag = _l[s].inv_pre.elements();
          const Matrix_& R = lv.transfer->get_mat_rest();
          const Matrix_& P = lv.transfer->get_mat_prol();
          d.r_val = (const double*)R.val();
          d.r_col = (const uint64_t*)R.col_ind();
          d.r_ptr = (const uint64_t*)R.row_ptr();
          d.p_val = (const double*)P.val();
          d.p_col = (const uint64_t*)P.col_ind();
          d.p_ptr = (const uint64_t*)P.row_ptr();
          d.pre_steps = _l[s].pre_steps;
          d.post_steps = _l[s].post_steps;
          d.pre_omega = _l[s].omega_pre;
          d.post_omega = _l[s].omega_post;
          // grid form found at init_numeric (no copies here: this may run inside a recorded graph)
          if(s < _grid.size() && (_grid[s].dlinear || _grid_sten[s].size()))
          {
            d.dim = _grid[s].dim;
            d.n1 = _grid[s].n1;
            d.dlinear = _grid[s].dlinear ? 1 : 0;
            d.sten = _grid_sten[s].size() ? _grid_sten[s].get() : nullptr;
          }
        }
        _tail.coarse_inverse = static_cast<DirectType*>(_h->levels[c].coarse.get())->inverse();
        _tail_t = t;
      }

    public:
      /// half = true stores the matrices of the cycle as FP16 (FP32 arithmetic)
      explicit MixedMultiGrid(std::shared_ptr<HierarchyType> h, bool half = false) : _h(h), _half(half) {}
      std::string name() const override { return _half ? "MultiGrid-V(FP16/FP32)" : "MultiGrid-V(FP32)"; }
      bool fusable() const override { return _ok; }
      /// run the coarse levels as one LDS-resident workgroup where they qualify (default on)
      void set_coarse_fusion(bool on)
      {
        _tail_on = on;
        _tail_built = false;
      }
      /// first level of the fused coarse part (0: none)
      /// fused levels in the grid form (stencil from the kernel, d-linear transfers)
      int coarse_fusion_grid_levels()
      {
        if(!_tail_built) _build_tail();
        int k = 0;
        for(int l = 0; _tail_t && l + 1 < _tail.nlevels; ++l) k += (_tail.lv[l].sten && _tail.lv[l].dlinear) ? 1 : 0;
        return k;
      }
      Index coarse_fusion_level()
      {
        if(!_tail_built) _build_tail();
        return _tail_t;
      }

      void init_symbolic() override
      {
        const Index L = _h->size();
        _fw.create(2 * L + 8);
        _ok = true;
        _l.assign(L, Level32());
        for(Index i = 0; i < L; ++i)
        {
          auto& lv = _h->levels[i];
          const Index n = lv.matrix->rows();
          _l[i].rhs = Vector32(n);
          _l[i].sol = Vector32(n);
          _l[i].def = Vector32(n);
          _l[i].cor = Vector32(n);
          _l[i].tmp = Vector32(n);
          if(i + 1 < L)
          {
            auto* pre = dynamic_cast<RichType*>(lv.pre.get());
            auto* post = dynamic_cast<RichType*>(lv.post.get());
            if(!pre || !post || !pre->fusable_jacobi() || !post->fusable_jacobi()) _ok = false;
          }
        }
        auto* c = _h->levels[L - 1].coarse.get();
        if(!c)
          _coarse = Coarse::none;
        else if(dynamic_cast<DirectType*>(c))
          _coarse = Coarse::direct;
        else if(dynamic_cast<JacType*>(c))
          _coarse = Coarse::jacobi;
        else if(auto* r = dynamic_cast<RichType*>(c))
        {
          _coarse = Coarse::richardson;
          if(!r->fusable_jacobi()) _ok = false;
        }
        else
          _ok = false;
        if(!_ok) throw Error("MixedMultiGrid: needs Richardson(Jacobi) smoothers with fixed step counts");
      }

      /// Needs the hierarchy's init_numeric (the FP64 Jacobi diagonals) first.
      void init_numeric() override
      {
        const Index L = _h->size();
        for(Index i = 0; i < L; ++i)
        {
          auto& lv = _h->levels[i];
          // FP32 execution plans of the level's matrices (values rounded once)
          (void)plan(*lv.matrix);
          if(i + 1 < L)
          {
            (void)plan(lv.transfer->get_mat_prol());
            (void)plan(lv.transfer->get_mat_rest());
            auto* pre = static_cast<RichType*>(lv.pre.get());
            auto* post = static_cast<RichType*>(lv.post.get());
            _l[i].inv_pre = inv_diag32(*pre->fusable_jacobi());
            _l[i].inv_post = inv_diag32(*post->fusable_jacobi());
            _l[i].pre_steps = int(pre->get_max_iter());
            _l[i].post_steps = int(post->get_max_iter());
            _l[i].omega_pre = float(pre->omega());
            _l[i].omega_post = float(post->omega());
          }
        }
        auto& lc = _h->levels[L - 1];
        if(_coarse == Coarse::jacobi)
          _l[L - 1].inv_pre = inv_diag32(*static_cast<JacType*>(lc.coarse.get()));
        else if(_coarse == Coarse::richardson)
        {
          auto* r = static_cast<RichType*>(lc.coarse.get());
          _l[L - 1].inv_pre = inv_diag32(*r->fusable_jacobi());
          _l[L - 1].pre_steps = int(r->get_max_iter());
          _l[L - 1].omega_pre = float(r->omega());
        }
        _tail_built = false;   // the tail holds pointers to the vectors made here
        _scan_grid();
      }

      Status apply(VectorType& cor, const VectorType& def) override
      {
        double* nrm = _fw.scalar(0);
        check(feat_dot_f64(Runtime::handle(), nrm, (const double*)def.elements(), (const double*)def.elements(), def.size()),
              "MixedMultiGrid::apply norm");
        apply_fused(cor, def, nrm, nullptr);
        return Status::success;
      }

      void apply_fused(VectorType& cor, const VectorType& def, const double* def_nrm2, double* gamma) override
      {
        const Index L = _h->size();
        const Index c = L - 1;
        feat_handle_t hd = Runtime::handle();
        const feat_reduce_ws_t* ws = &_fw.ws;
        auto nrm_of = [&](Index i) -> const double* { return i == 0 ? def_nrm2 : _fw.scalar(2 + i); };
        if(!_tail_built) _build_tail();
        const Index t = _tail_t ? _tail_t : c;   // levels from t down run in the fused tail (when _tail_t)
        check(feat_convert_f64_f32(hd, _l[0].rhs.elements(), (const double*)def.elements(), def.size()), "to FP32");
        for(Index i = 0; i < t; ++i)
        {
          auto& lv = _h->levels[i];
          Level32& l = _l[i];
          const feat_matrix_f32_t& A = plan(*lv.matrix);
          check(feat_rj_smooth_f32(hd, &A, l.inv_pre.elements(), lv.filter->mask(), l.pre_steps, l.omega_pre,
                                   l.rhs.elements(), nrm_of(i), l.sol.elements(), l.tmp.elements(), l.def.elements(),
                                   nullptr, ws),
                "FP32 pre-smoothing");
          const feat_matrix_f32_t& R = _rest(i);
          check(feat_apply_masked_f32(hd, &R, _l[i + 1].rhs.elements(), l.def.elements(), _h->levels[i + 1].filter->mask(),
                                      _fw.scalar(2 + i + 1), ws),
                "FP32 restriction");
        }
        if(_tail_t)
          check(feat_mg_tail_f32(hd, &_tail, _l[t].rhs.elements(), nrm_of(t), _l[t].sol.elements()),
                "FP32 fused coarse levels");
        else
        {
          auto& lv = _h->levels[c];
          Level32& l = _l[c];
          const int n = int(lv.matrix->rows());
          switch(_coarse)
          {
          case Coarse::direct:
            check(feat_dense_gemv_mixed(hd, n, static_cast<DirectType*>(lv.coarse.get())->inverse(), l.rhs.elements(),
                                        l.sol.elements(), lv.filter->mask()),
                  "FP64 coarse solve");
            break;
          case Coarse::jacobi:
          {
            const feat_matrix_f32_t& A = plan(*lv.matrix);
            check(feat_rj_smooth_f32(hd, &A, l.inv_pre.elements(), lv.filter->mask(), 1, 1.0f, l.rhs.elements(), nullptr,
                                     l.sol.elements(), l.tmp.elements(), nullptr, nullptr, ws),
                  "FP32 coarse Jacobi");
            break;
          }
          case Coarse::richardson:
          {
            const feat_matrix_f32_t& A = plan(*lv.matrix);
            check(feat_rj_smooth_f32(hd, &A, l.inv_pre.elements(), lv.filter->mask(), l.pre_steps, l.omega_pre,
                                     l.rhs.elements(), nrm_of(c), l.sol.elements(), l.tmp.elements(), nullptr, nullptr, ws),
                  "FP32 coarse Richardson");
            break;
          }
          default:
            throw Error("MixedMultiGrid: coarse solver not supported");
          }
        }
        for(Index k = t; k-- > 0;)
        {
          auto& lv = _h->levels[k];
          Level32& l = _l[k];
          const feat_matrix_f32_t& P = _prol(k);
          check(feat_prol_add_f32(hd, &P, l.sol.elements(), _l[k + 1].sol.elements(), lv.filter->mask()), "FP32 prolongation");
          const feat_matrix_f32_t& A = plan(*lv.matrix);
          double* dn = _fw.scalar(2 + L + k);
          check(feat_defect_f32(hd, &A, l.def.elements(), l.rhs.elements(), l.sol.elements(), lv.filter->mask(), dn, ws),
                "FP32 defect");
          check(feat_rj_smooth_f32(hd, &A, l.inv_post.elements(), lv.filter->mask(), l.post_steps, l.omega_post,
                                   l.def.elements(), dn, l.cor.elements(), l.tmp.elements(), nullptr, l.sol.elements(), ws),
                "FP32 post-smoothing");
        }
        check(feat_convert_dot_f32_f64(hd, (double*)cor.elements(), _l[0].sol.elements(), (const double*)def.elements(),
                                       cor.size(), gamma, ws),
              "to FP64");
      }
    };
  } // namespace Solver
} // namespace feat3amd
