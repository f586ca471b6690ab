// host_capi.cpp -- extern "C" wrappers of the C++ host layer.
// Built by g++ (no HIP headers) into libfeat3amd_host.so, linked against the
// kernel library libfeat3amd.so; every device operation goes through feat3amd.h.
#include <feat3amd/assembly.hpp>
#include <feat3amd/batch.hpp>
#include <feat3amd/learned.hpp>
#include <feat3amd/lafem.hpp>
#include <feat3amd/mixed.hpp>
#include <feat3amd/solver.hpp>
#include <feat3amd/stokes.hpp>
#include <feat3amd/stokes_assembly.hpp>
#include <feat3amd/io.hpp>
#include <feat3amd/sor.hpp>
#include <feat3amd_host.h>

#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace feat3amd;
typedef LAFEM::DenseVector<double, Index> Vec;
typedef LAFEM::SparseMatrixCSR<double, Index> Csr;
typedef LAFEM::UnitFilter<double, Index> Filt;
typedef LAFEM::Transfer<Csr> Trans;
typedef Solver::MultiGridHierarchy<Csr, Filt, Trans> Hier;
typedef Solver::MultiGrid<Csr, Filt, Trans> MG;
typedef Solver::MixedMultiGrid<Csr, Filt, Trans> MG32;

namespace
{
  thread_local std::string g_host_err;

  template<typename F_>
  int guarded(F_&& f)
  {
    try
    {
      f();
      return FEAT_OK;
    }
    catch(const std::exception& e)
    {
      g_host_err = e.what();
      return FEAT_ERR_INVALID;
    }
  }

  struct LevelData
  {
    Csr A;
    Filt F;
    Trans T;
    std::vector<double> hval;
    std::vector<std::uint64_t> hcol, hrp, hfilt;
  };
}

struct feat_mgpcg_s
{
  std::vector<std::unique_ptr<LevelData>> levels;
  std::shared_ptr<Hier> hier;
  std::shared_ptr<MG> mg;
  std::shared_ptr<MG32> mg32;
  std::shared_ptr<Solver::PCG<Csr, Filt>> pcg;
};

extern "C" {

const char* feat_host_last_error(void) { return g_host_err.c_str(); }

int feat_runtime_init(int device, void* stream)
{
  return guarded([&]() { Runtime::initialize(device, stream); });
}

int feat_runtime_grid_forms(int on)
{
  return guarded([&]() { Runtime::grid_forms() = on != 0; });
}

int feat_runtime_sell_flags(int flags)
{
  return guarded([&]() { Runtime::sell_flags() = flags; });
}

int feat_q1_sizes(int dim, int level, std::uint64_t* sizes)
{
  return guarded([&]() {
    if(dim != 2 && dim != 3) throw Error("feat_q1_sizes: dim must be 2 or 3");
    const auto A = Assembly::q1_laplace(dim, level);
    sizes[0] = A.rows;
    sizes[1] = A.rowptr[A.rows];
    sizes[2] = Assembly::q1_boundary(dim, level).size();
    if(level > 0)
    {
      const auto P = Assembly::q1_prolongation(dim, level);
      sizes[3] = P.rowptr[P.rows];
      sizes[4] = P.cols;
    }
    else
      sizes[3] = sizes[4] = 0;
  });
}

int feat_q1_assemble(int dim, int level, double* val, std::uint64_t* col, std::uint64_t* rowptr, std::uint64_t* bnd,
                     double* pval, std::uint64_t* pcol, std::uint64_t* prowptr)
{
  return guarded([&]() {
    const auto A = Assembly::q1_laplace(dim, level);
    std::memcpy(val, A.val.data(), sizeof(double) * A.val.size());
    std::memcpy(col, A.col.data(), sizeof(std::uint64_t) * A.col.size());
    std::memcpy(rowptr, A.rowptr.data(), sizeof(std::uint64_t) * A.rowptr.size());
    const auto b = Assembly::q1_boundary(dim, level);
    std::memcpy(bnd, b.data(), sizeof(std::uint64_t) * b.size());
    if(level > 0 && pval)
    {
      const auto P = Assembly::q1_prolongation(dim, level);
      std::memcpy(pval, P.val.data(), sizeof(double) * P.val.size());
      std::memcpy(pcol, P.col.data(), sizeof(std::uint64_t) * P.col.size());
      std::memcpy(prowptr, P.rowptr.data(), sizeof(std::uint64_t) * P.rowptr.size());
    }
  });
}

int feat_q1_rhs_sine(int dim, int level, double* b)
{
  return guarded([&]() {
    const auto v = Assembly::q1_rhs_sine(dim, level);
    std::memcpy(b, v.data(), sizeof(double) * v.size());
  });
}

int feat_stokes_sizes(int level, std::uint64_t* sizes)
{
  return guarded([&]() {
    if(level < 0 || level > 12) throw Error("feat_stokes_sizes: level out of range");
    const auto S = Assembly::stokes_q2p1dc(level);
    sizes[0] = S.nv;
    sizes[1] = S.np;
    sizes[2] = S.A.rowptr[S.A.rows];
    sizes[3] = S.B.rowptr[S.B.rows];
    sizes[4] = S.D.rowptr[S.D.rows];
    sizes[5] = S.filt.size();
    sizes[6] = level > 0 ? S.Pv.rowptr[S.Pv.rows] : 0;
    sizes[7] = level > 0 ? S.Pp.rowptr[S.Pp.rows] : 0;
    sizes[8] = level > 0 ? S.Pv.cols : 0;
    sizes[9] = level > 0 ? S.Pp.cols : 0;
  });
}

int feat_stokes_assemble(int level, const feat_stokes_arrays_t* o)
{
  return guarded([&]() {
    if(!o) throw Error("feat_stokes_assemble: null output");
    const auto S = Assembly::stokes_q2p1dc(level);
    auto put = [](auto* dst, const auto& v) {
      if(dst) std::memcpy(dst, v.data(), sizeof(v[0]) * v.size());
    };
    put(o->a_val, S.A.val); put(o->a_col, S.A.col); put(o->a_ptr, S.A.rowptr);
    put(o->b_val, S.B.val); put(o->b_col, S.B.col); put(o->b_ptr, S.B.rowptr);
    put(o->d_val, S.D.val); put(o->d_col, S.D.col); put(o->d_ptr, S.D.rowptr);
    put(o->filt, S.filt); put(o->filt_val, S.filt_val);
    put(o->prim, S.prim); put(o->dual, S.dual);
    if(level > 0)
    {
      put(o->pv_val, S.Pv.val); put(o->pv_col, S.Pv.col); put(o->pv_ptr, S.Pv.rowptr);
      put(o->pp_val, S.Pp.val); put(o->pp_col, S.Pp.col); put(o->pp_ptr, S.Pp.rowptr);
    }
  });
}

struct feat_io_s
{
  Assembly::CsrHost A;
  std::vector<double> v;
  bool is_matrix = false;
};

int feat_csr_write_file(int mode, const char* path, std::uint64_t rows, std::uint64_t cols, const double* val,
                        const std::uint64_t* col, const std::uint64_t* rowptr, int symmetric)
{
  return guarded([&]() {
    if(!path || !rowptr) throw Error("feat_csr_write_file: null argument");
    Assembly::CsrHost A;
    A.rows = rows;
    A.cols = cols;
    A.rowptr.assign(rowptr, rowptr + rows + 1);
    const std::uint64_t nnz = A.rowptr[rows];
    if(nnz && (!val || !col)) throw Error("feat_csr_write_file: null arrays");
    A.val.assign(val, val + nnz);
    A.col.assign(col, col + nnz);
    LAFEM::write_csr(LAFEM::FileMode(mode), std::string(path), A, symmetric != 0);
  });
}

int feat_csr_read_file(int mode, const char* path, feat_io_t* out, std::uint64_t* rows, std::uint64_t* cols,
                       std::uint64_t* nnz)
{
  return guarded([&]() {
    if(!path || !out) throw Error("feat_csr_read_file: null argument");
    auto f = std::make_unique<feat_io_s>();
    f->A = LAFEM::read_csr(LAFEM::FileMode(mode), std::string(path));
    f->is_matrix = true;
    if(rows) *rows = f->A.rows;
    if(cols) *cols = f->A.cols;
    if(nnz) *nnz = f->A.val.size();
    *out = f.release();
  });
}

int feat_io_fetch_csr(feat_io_t f, double* val, std::uint64_t* col, std::uint64_t* rowptr)
{
  return guarded([&]() {
    if(!f || !f->is_matrix) throw Error("feat_io_fetch_csr: not a matrix");
    if(val) std::copy(f->A.val.begin(), f->A.val.end(), val);
    if(col) std::copy(f->A.col.begin(), f->A.col.end(), col);
    if(rowptr) std::copy(f->A.rowptr.begin(), f->A.rowptr.end(), rowptr);
  });
}

int feat_dv_write_file(int mode, const char* path, const double* v, std::uint64_t n)
{
  return guarded([&]() {
    if(!path || (n && !v)) throw Error("feat_dv_write_file: null argument");
    LAFEM::write_dv(LAFEM::FileMode(mode), std::string(path), v, n);
  });
}

int feat_dv_read_file(int mode, const char* path, feat_io_t* out, std::uint64_t* n)
{
  return guarded([&]() {
    if(!path || !out) throw Error("feat_dv_read_file: null argument");
    auto f = std::make_unique<feat_io_s>();
    f->v = LAFEM::read_dv(LAFEM::FileMode(mode), std::string(path));
    if(n) *n = f->v.size();
    *out = f.release();
  });
}

int feat_io_fetch_dv(feat_io_t f, double* v)
{
  return guarded([&]() {
    if(!f || f->is_matrix) throw Error("feat_io_fetch_dv: not a vector");
    if(v) std::copy(f->v.begin(), f->v.end(), v);
  });
}

int feat_io_free(feat_io_t f)
{
  return guarded([&]() { delete f; });
}

int feat_rhs_random(std::uint64_t seed, std::uint64_t n, double a, double b, double* out)
{
  return guarded([&]() { Assembly::random_fill(seed, out, n, a, b); });
}

static void build_solver(feat_mgpcg_s* s, const feat_mgpcg_cfg_t* cfg)
{
  const int nl = int(s->levels.size());
  if(cfg->csr_only)
    for(auto& L : s->levels)
    {
      L->A.set_sell(false);
      if(!L->T.get_mat_prol().empty())
      {
        L->T.get_mat_prol().set_sell(false);
        L->T.get_mat_rest().set_sell(false);
      }
    }
  s->hier = std::make_shared<Hier>();
  for(int i = 0; i < nl; ++i)
  {
    LevelData& L = *s->levels[i];
    auto jac = Solver::new_jacobi_precond(L.A, L.F, cfg->omega_jac);
    if(i + 1 < nl)
    {
      auto pre = Solver::new_richardson(L.A, L.F, cfg->omega_rich, jac);
      pre->set_min_iter(cfg->nu_pre);
      pre->set_max_iter(cfg->nu_pre);
      auto post = Solver::new_richardson(L.A, L.F, cfg->omega_rich, jac);
      post->set_min_iter(cfg->nu_post);
      post->set_max_iter(cfg->nu_post);
      s->hier->push_level(L.A, L.F, L.T, pre, post, pre);
    }
    else
    {
      std::shared_ptr<Solver::SolverBase<Vec>> cs;
      if(cfg->coarse == FEAT_COARSE_SMOOTHER)
      {
        auto r = Solver::new_richardson(L.A, L.F, cfg->omega_rich, jac);
        r->set_min_iter(cfg->nu_pre);
        r->set_max_iter(cfg->nu_pre);
        cs = r;
      }
      else if(cfg->coarse == FEAT_COARSE_JACOBI)
        cs = Solver::new_jacobi_precond(L.A, L.F, cfg->coarse_omega);
      else if(cfg->coarse == FEAT_COARSE_DIRECT)
        cs = std::make_shared<Solver::DirectCoarseSolver<Csr, Filt>>(L.A, L.F, L.hval, L.hcol, L.hrp, L.hfilt);
      s->hier->push_level(L.A, L.F, cs);
    }
  }
  if(cfg->precision == 1 || cfg->precision == 2)
  {
    s->mg32 = std::make_shared<MG32>(s->hier, cfg->precision == 2);
    s->mg32->set_coarse_fusion(cfg->no_coarse_fusion == 0);
    s->pcg = Solver::new_pcg(s->levels[0]->A, s->levels[0]->F, s->mg32);
  }
  else
  {
    s->mg = std::make_shared<MG>(s->hier);
    s->mg->set_autotune(cfg->no_autotune == 0);
    s->mg->set_coarse_fusion(cfg->no_coarse_fusion == 0);
    s->pcg = Solver::new_pcg(s->levels[0]->A, s->levels[0]->F, s->mg);
  }
  s->pcg->set_tol_rel(cfg->tol_rel);
  s->pcg->set_max_iter(cfg->max_iter);
  s->pcg->set_force_generic(cfg->force_generic != 0);
  s->pcg->set_use_graphs(cfg->no_graphs == 0);
  s->hier->init();
  s->pcg->init();
}

static std::unique_ptr<LevelData> make_level(const Assembly::CsrHost& A, const std::vector<std::uint64_t>& filt,
                                             const Assembly::CsrHost* P, bool keep_host)
{
  auto L = std::make_unique<LevelData>();
  L->A = Csr::from_host(A.rows, A.cols, A.val.data(), A.col.data(), A.rowptr.data());
  L->F = Filt(A.rows, filt.data(), nullptr, filt.size());
  if(P)
  {
    const auto R = Assembly::transpose(*P);
    L->T = Trans(Csr::from_host(P->rows, P->cols, P->val.data(), P->col.data(), P->rowptr.data()),
                 Csr::from_host(R.rows, R.cols, R.val.data(), R.col.data(), R.rowptr.data()));
  }
  if(keep_host)
  {
    L->hval = A.val;
    L->hcol = A.col;
    L->hrp = A.rowptr;
    L->hfilt = filt;
  }
  return L;
}

int feat_mgpcg_create(feat_mgpcg_t* out, const feat_level_host_t* lv, int nl, const feat_mgpcg_cfg_t* cfg)
{
  return guarded([&]() {
    if(!out || !lv || nl < 1 || !cfg) throw Error("feat_mgpcg_create: bad argument");
    auto s = std::make_unique<feat_mgpcg_s>();
    for(int i = 0; i < nl; ++i)
    {
      Assembly::CsrHost A;
      A.rows = A.cols = lv[i].rows;
      A.rowptr.assign(lv[i].rowptr, lv[i].rowptr + lv[i].rows + 1);
      A.val.assign(lv[i].val, lv[i].val + A.rowptr[A.rows]);
      A.col.assign(lv[i].col, lv[i].col + A.rowptr[A.rows]);
      std::vector<std::uint64_t> f(lv[i].filt, lv[i].filt + lv[i].nfilt);
      Assembly::CsrHost P;
      const bool hasP = (i + 1 < nl);
      if(hasP)
      {
        P.rows = lv[i].rows;
        P.cols = lv[i].pcols;
        P.rowptr.assign(lv[i].prowptr, lv[i].prowptr + P.rows + 1);
        P.val.assign(lv[i].pval, lv[i].pval + P.rowptr[P.rows]);
        P.col.assign(lv[i].pcol, lv[i].pcol + P.rowptr[P.rows]);
      }
      s->levels.push_back(make_level(A, f, hasP ? &P : nullptr, i + 1 == nl));
    }
    build_solver(s.get(), cfg);
    *out = s.release();
  });
}

// A structured Q1 level assembled on the device (stiffness, boundary, and the
// prolongation from level-1 with its transpose when `transfer`).
static std::unique_ptr<LevelData> make_level_device(int dim, int l, bool transfer)
{
  std::uint64_t sz[5];
  check(feat_q1_device_sizes(dim, l, sz), "feat_q1_device_sizes");
  auto L = std::make_unique<LevelData>();
  L->A = Csr::allocate(sz[0], sz[0], sz[1]);
  DeviceArray<std::uint64_t> bnd(sz[2]);
  Csr P, R;
  if(transfer)
  {
    P = Csr::allocate(sz[0], sz[4], sz[3]);
    R = Csr::allocate(sz[4], sz[0], sz[3]);
  }
  check(feat_q1_assemble_device(Runtime::handle(), dim, l, L->A.val_mut(), L->A.col_mut(), L->A.row_ptr_mut(),
                                bnd.get(), transfer ? P.val_mut() : nullptr, transfer ? P.col_mut() : nullptr,
                                transfer ? P.row_ptr_mut() : nullptr, transfer ? R.val_mut() : nullptr,
                                transfer ? R.col_mut() : nullptr, transfer ? R.row_ptr_mut() : nullptr),
        "feat_q1_assemble_device");
  L->F = Filt(sz[0], bnd);
  if(transfer) L->T = Trans(std::move(P), std::move(R));
  return L;
}

int feat_mgpcg_create_q1_ex(feat_mgpcg_t* out, int dim, int lf, int lc, const feat_mgpcg_cfg_t* cfg, int flags)
{
  return guarded([&]() {
    if(!out || !cfg || lf < lc || lc < 0 || (dim != 2 && dim != 3)) throw Error("feat_mgpcg_create_q1: bad argument");
    auto s = std::make_unique<feat_mgpcg_s>();
    for(int l = lf; l >= lc; --l)
    {
      if(l > lc && !(flags & FEAT_Q1_HOST_ASSEMBLY))
      {
        s->levels.push_back(make_level_device(dim, l, true));
        continue;
      }
      const auto A = Assembly::q1_laplace(dim, l);
      const auto f = Assembly::q1_boundary(dim, l);
      if(l > lc)
      {
        const auto P = Assembly::q1_prolongation(dim, l);
        s->levels.push_back(make_level(A, f, &P, false));
      }
      else  // the coarse solver factorises the host copy
        s->levels.push_back(make_level(A, f, nullptr, true));
    }
    build_solver(s.get(), cfg);
    *out = s.release();
  });
}

int feat_mgpcg_create_q1(feat_mgpcg_t* out, int dim, int lf, int lc, const feat_mgpcg_cfg_t* cfg)
{
  return feat_mgpcg_create_q1_ex(out, dim, lf, lc, cfg, 0);
}

int feat_mgpcg_solve(feat_mgpcg_t s, double* x, const double* b, feat_solve_result_t* res, double* hist,
                     std::uint64_t hist_cap)
{
  return guarded([&]() {
    if(!s || !x || !b || !res) throw Error("feat_mgpcg_solve: bad argument");
    const Index n = s->levels[0]->A.rows();
    Vec vx = Vec::wrap(x, n);
    Vec vb = Vec::wrap(const_cast<double*>(b), n);
    Solver::Status st = s->pcg->correct(vx, vb);
    Runtime::synchronize();
    res->status = int(st);
    res->iters = s->pcg->get_num_iter();
    res->def_init = s->pcg->get_def_initial();
    res->def_final = s->pcg->get_def_final();
    if(hist)
    {
      const auto& h = s->pcg->history();
      for(std::size_t i = 0; i < h.size() && i < hist_cap; ++i) hist[i] = h[i];
    }
  });
}

int feat_mgpcg_write_matrix(feat_mgpcg_t s, int lvl, int mode, const char* path, int symmetric)
{
  return guarded([&]() {
    if(!s || !path || lvl < 0 || lvl >= int(s->levels.size())) throw Error("feat_mgpcg_write_matrix: bad argument");
    LAFEM::write_out(s->levels[lvl]->A, LAFEM::FileMode(mode), path, symmetric != 0);
  });
}

int feat_mgpcg_matrix(feat_mgpcg_t s, int lvl, std::uint64_t* rows, std::uint64_t* nnz, const double** val,
                      const std::uint64_t** col, const std::uint64_t** rowptr)
{
  return guarded([&]() {
    if(!s || lvl < 0 || lvl >= int(s->levels.size())) throw Error("feat_mgpcg_matrix: bad level");
    const Csr& A = s->levels[lvl]->A;
    *rows = A.rows();
    *nnz = A.used_elements();
    *val = A.val();
    *col = A.col_ind();
    *rowptr = A.row_ptr();
  });
}

int feat_mgpcg_set_values(feat_mgpcg_t s, int lvl, const double* val_host)
{
  return guarded([&]() {
    if(!s || !val_host || lvl < 0 || lvl >= int(s->levels.size())) throw Error("feat_mgpcg_set_values: bad argument");
    Csr& A = s->levels[lvl]->A;
    check(feat_copy(Runtime::handle(), A.val_mut(), val_host, sizeof(double) * A.used_elements()),
          "feat_mgpcg_set_values: upload");
    A.values_changed();
    s->hier->init_numeric();
    s->pcg->init_numeric();
    check(feat_sync(Runtime::handle()), "feat_mgpcg_set_values: sync");
  });
}

struct feat_krylov_s
{
  Csr A;
  Filt F;
  std::shared_ptr<Solver::SolverBase<Vec>> pre;
  std::shared_ptr<Solver::IterativeSolver<Vec>> solver;
};

int feat_krylov_create(feat_krylov_t* out, int solver, std::uint64_t n, const double* val, const std::uint64_t* col,
                       const std::uint64_t* rowptr, const std::uint64_t* filt, std::uint64_t nfilt, int precond,
                       double omega_pre, double omega_rich, double tol_rel, std::uint64_t max_iter)
{
  return guarded([&]() {
    if(!out || !val || !col || !rowptr || (nfilt && !filt) || (solver != 0 && solver != 1))
      throw Error("feat_krylov_create: bad argument");
    auto s = std::make_unique<feat_krylov_s>();
    s->A = Csr::from_host(n, n, val, col, rowptr);
    s->F = Filt(n, filt, nullptr, nfilt);
    switch(precond)
    {
    case FEAT_PRECOND_NONE: break;
    case FEAT_PRECOND_JACOBI: s->pre = Solver::new_jacobi_precond(s->A, s->F, omega_pre); break;
    case FEAT_PRECOND_SOR: s->pre = Solver::new_sor_precond(s->A, s->F, omega_pre); break;
    case FEAT_PRECOND_SSOR: s->pre = Solver::new_ssor_precond(s->A, s->F, omega_pre); break;
    case FEAT_PRECOND_ILU:
      // ILUPrecond's fill-in level p travels in omega_pre (0 = ILU(0))
      if(!(omega_pre >= 0.0) || omega_pre != double(int(omega_pre))) throw Error("feat_krylov_create: ILU level p must be a whole number >= 0");
      s->pre = Solver::new_ilu_precond(s->A, s->F, int(omega_pre));
      break;
    case FEAT_PRECOND_SOR_MULTICOLOR:
      s->pre = std::make_shared<Solver::SORPrecond<Csr, Filt>>(s->A, s->F, omega_pre, FEAT_SOR_MULTICOLOR);
      break;
    case FEAT_PRECOND_SSOR_MULTICOLOR:
      s->pre = std::make_shared<Solver::SSORPrecond<Csr, Filt>>(s->A, s->F, omega_pre, FEAT_SOR_MULTICOLOR);
      break;
    default: throw Error("feat_krylov_create: unknown preconditioner");
    }
    if(solver == 0)
      s->solver = Solver::new_pcg(s->A, s->F, s->pre);
    else
      s->solver = Solver::new_richardson(s->A, s->F, omega_rich, s->pre);
    s->solver->set_tol_rel(tol_rel);
    s->solver->set_max_iter(max_iter);
    s->solver->init();
    *out = s.release();
  });
}

int feat_krylov_solve(feat_krylov_t s, double* x, const double* b, feat_solve_result_t* res, double* hist,
                      std::uint64_t hist_cap)
{
  return guarded([&]() {
    if(!s || !x || !b || !res) throw Error("feat_krylov_solve: bad argument");
    const Index n = s->A.rows();
    Vec vx = Vec::wrap(x, n);
    Vec vb = Vec::wrap(const_cast<double*>(b), n);
    Solver::Status st = s->solver->correct(vx, vb);
    Runtime::synchronize();
    res->status = int(st);
    res->iters = s->solver->get_num_iter();
    res->def_init = s->solver->get_def_initial();
    res->def_final = s->solver->get_def_final();
    if(hist)
    {
      const auto& h = s->solver->history();
      for(std::size_t i = 0; i < h.size() && i < hist_cap; ++i) hist[i] = h[i];
    }
  });
}

int feat_krylov_precond_apply(feat_krylov_t s, double* cor, const double* def)
{
  return guarded([&]() {
    if(!s || !cor || !def || !s->pre) throw Error("feat_krylov_precond_apply: bad argument");
    const Index n = s->A.rows();
    Vec vc = Vec::wrap(cor, n);
    Vec vd = Vec::wrap(const_cast<double*>(def), n);
    s->pre->apply(vc, vd);
    Runtime::synchronize();
  });
}

int feat_krylov_destroy(feat_krylov_t s)
{
  return guarded([&]() {
    if(s && s->solver) s->solver->done();
    delete s;
  });
}

int feat_mgpcg_coarse_fusion_grid(feat_mgpcg_t s, int* levels)
{
  return guarded([&]() {
    if(!s || !levels) throw Error("feat_mgpcg_coarse_fusion_grid: bad argument");
    *levels = s->mg32 ? s->mg32->coarse_fusion_grid_levels() : (s->mg ? s->mg->coarse_fusion_grid_levels() : 0);
  });
}

int feat_mgpcg_coarse_fusion_level(feat_mgpcg_t s, int* level)
{
  return guarded([&]() {
    if(!s || !level) throw Error("feat_mgpcg_coarse_fusion_level: bad argument");
    *level = int(s->mg32 ? s->mg32->coarse_fusion_level() : (s->mg ? s->mg->coarse_fusion_level() : 0));
  });
}

int feat_mgpcg_destroy(feat_mgpcg_t s)
{
  return guarded([&]() { delete s; });
}

} // extern "C"

struct feat_stokes_s
{
  std::unique_ptr<Solver::StokesMultiGrid> mg;
};

namespace
{
  Solver::StokesMultiGrid::Level& stokes_level(feat_stokes_t s, int lvl)
  {
    if(!s || lvl < 0 || std::size_t(lvl) >= s->mg->num_levels()) throw Error("feat_stokes: bad level");
    return s->mg->level(std::size_t(lvl));
  }
  LAFEM::StokesVector svec(const Solver::StokesMultiGrid::Level& L, const double* p)
  {
    return LAFEM::StokesVector::wrap(const_cast<double*>(p), L.K.size());
  }
}

extern "C" {

int feat_stokes_create(feat_stokes_t* out, int level_fine, int level_coarse, const feat_stokes_cfg_t* cfg)
{
  return guarded([&]() {
    if(!out || !cfg || level_coarse < 0 || level_fine < level_coarse || level_fine > 10)
      throw Error("feat_stokes_create: bad argument");
    Solver::StokesConfig c;
    c.steps = cfg->steps;
    c.omega_rich = cfg->omega_rich;
    c.vanka_omega = cfg->vanka_omega;
    c.coarse_skip_singular = cfg->coarse_skip_singular != 0;
    c.omega_outer = cfg->omega_outer;
    c.graphs = cfg->no_graphs == 0;
    c.coarse_fusion = cfg->no_coarse_fusion == 0;
    if(cfg->tail_max_level_dofs) c.tail_max_level_dofs = cfg->tail_max_level_dofs;
    auto s = std::make_unique<feat_stokes_s>();
    s->mg = std::make_unique<Solver::StokesMultiGrid>(c);
    s->mg->set_tol_rel(cfg->tol_rel);
    s->mg->set_max_iter(cfg->max_iter);
    s->mg->build(level_fine, level_coarse);
    s->mg->init();
    Runtime::synchronize();
    *out = s.release();
  });
}

int feat_stokes_solve(feat_stokes_t s, double* x, const double* b, feat_solve_result_t* res, double* hist,
                      std::uint64_t hist_cap)
{
  return guarded([&]() {
    if(!s || !x || !b || !res) throw Error("feat_stokes_solve: bad argument");
    auto& L = s->mg->level(0);
    LAFEM::StokesVector vx = svec(L, x), vb = svec(L, b);
    Solver::Status st = s->mg->correct(vx, vb);
    Runtime::synchronize();
    res->status = int(st);
    res->iters = s->mg->get_num_iter();
    res->def_init = s->mg->get_def_initial();
    res->def_final = s->mg->get_def_final();
    if(hist)
    {
      const auto& h = s->mg->history();
      for(std::size_t i = 0; i < h.size() && i < hist_cap; ++i) hist[i] = h[i];
    }
  });
}

int feat_stokes_init_numeric(feat_stokes_t s, double* seconds)
{
  return guarded([&]() {
    if(!s) throw Error("feat_stokes_init_numeric: null");
    Runtime::synchronize();
    const auto t0 = std::chrono::steady_clock::now();
    s->mg->init_numeric();
    Runtime::synchronize();
    if(seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

int feat_stokes_sizes_of(feat_stokes_t s, int lvl, std::uint64_t* out)
{
  return guarded([&]() {
    auto& L = stokes_level(s, lvl);
    out[0] = L.K.nv();
    out[1] = L.K.np();
    out[2] = L.V->num_macros();
    out[3] = L.V->stride();
    for(int k = 0; k < 4; ++k) out[4 + k] = L.V->block_nnz(k);
  });
}

int feat_stokes_vanka_copy(feat_stokes_t s, int lvl, std::uint64_t** ptr, std::uint64_t** col, double** val, int* mask)
{
  return guarded([&]() {
    auto& L = stokes_level(s, lvl);
    for(int k = 0; k < 4; ++k) L.V->copy_block(k, ptr ? ptr[k] : nullptr, col ? col[k] : nullptr, val ? val[k] : nullptr);
    if(mask) L.V->copy_mask(mask);
  });
}

int feat_stokes_apply(feat_stokes_t s, int lvl, double* r, const double* x, const double* y, double alpha)
{
  return guarded([&]() {
    auto& L = stokes_level(s, lvl);
    LAFEM::StokesVector vr = svec(L, r), vx = svec(L, x);
    if(y)
    {
      LAFEM::StokesVector vy = svec(L, y);
      L.K.apply(vr, vx, vy, alpha);
    }
    else
      L.K.apply(vr, vx);
    Runtime::synchronize();
  });
}

int feat_stokes_vanka_apply(feat_stokes_t s, int lvl, double* c, const double* d, int passes)
{
  return guarded([&]() {
    auto& L = stokes_level(s, lvl);
    LAFEM::StokesVector vc = svec(L, c), vd = svec(L, d);
    L.V->apply(vc, vd, passes);
    Runtime::synchronize();
  });
}

int feat_stokes_smooth(feat_stokes_t s, int lvl, double* sol, const double* rhs)
{
  return guarded([&]() {
    auto& L = stokes_level(s, lvl);
    LAFEM::StokesVector vs = svec(L, sol), vr = svec(L, rhs);
    s->mg->smooth_level(std::size_t(lvl), vs, vr);
    Runtime::synchronize();
  });
}

int feat_stokes_filter(feat_stokes_t s, int lvl, double* v, int kind)
{
  return guarded([&]() {
    auto& L = stokes_level(s, lvl);
    LAFEM::StokesVector vv = svec(L, v);
    switch(kind)
    {
    case 0: L.F.filter_rhs(vv); break;
    case 1: L.F.filter_sol(vv); break;
    case 2: L.F.filter_def(vv); break;
    case 3: L.F.filter_cor(vv); break;
    default: throw Error("feat_stokes_filter: bad kind");
    }
    Runtime::synchronize();
  });
}

int feat_stokes_transfer(feat_stokes_t s, int lvl, int prol, double* r, const double* x)
{
  return guarded([&]() {
    auto& F = stokes_level(s, lvl);
    auto& C = stokes_level(s, lvl + 1);
    if(prol)
    {
      LAFEM::StokesVector vr = svec(F, r), vx = svec(C, x);
      F.T.prol(vr, vx);
    }
    else
    {
      LAFEM::StokesVector vr = svec(C, r), vx = svec(F, x);
      F.T.rest(vx, vr);
    }
    Runtime::synchronize();
  });
}

int feat_stokes_destroy(feat_stokes_t s)
{
  return guarded([&]() { delete s; });
}

} // extern "C"

struct feat_batchsolver_s
{
  std::unique_ptr<Solver::BatchedMgPcg> b;
};

extern "C" {

int feat_batchsolver_create(feat_batchsolver_t* out, const feat_batch_cfg_t* cfg)
{
  return guarded([&]() {
    if(!out || !cfg || cfg->ninst < 1 || cfg->lmin < 1 || cfg->lmax < cfg->lmin || cfg->lmax > 11)
      throw Error("feat_batchsolver_create: bad argument");
    Solver::BatchConfig c;
    c.spec.seed = cfg->seed;
    c.spec.ninst = cfg->ninst;
    c.spec.lmin = cfg->lmin;
    c.spec.lmax = cfg->lmax;
    c.spec.sigma = cfg->sigma;
    c.tol_rel = cfg->tol_rel;
    c.max_iter = cfg->max_iter;
    c.autotune = cfg->autotune != 0;
    c.graphs = cfg->graphs != 0;
    c.tail_levels = cfg->tail_levels > 0 ? cfg->tail_levels : 0;
    c.tiled = cfg->no_tiled == 0;
    c.tiled_prol_arrays = cfg->tiled_prol_arrays != 0;
    c.tail_ell = cfg->tail_ell != 0;
    c.tiled_restriction = cfg->no_tiled_restriction == 0;
    auto s = std::make_unique<feat_batchsolver_s>();
    s->b = std::make_unique<Solver::BatchedMgPcg>(c);
    s->b->init();
    *out = s.release();
  });
}

int feat_batchsolver_layout(feat_batchsolver_t s, int* levels, std::uint64_t* off)
{
  return guarded([&]() {
    if(!s) throw Error("feat_batchsolver_layout: null");
    const auto& l = s->b->levels();
    const auto& o = s->b->offsets();
    if(levels) std::copy(l.begin(), l.end(), levels);
    if(off) std::copy(o.begin(), o.end(), off);
  });
}

int feat_batchsolver_solve(feat_batchsolver_t s, double* x, const double* b, int* status, std::uint64_t* iters,
                           double* def_init, double* def_final, double* hist)
{
  return guarded([&]() {
    if(!s || !x || !b) throw Error("feat_batchsolver_solve: bad argument");
    const auto R = s->b->solve(x, b);
    Runtime::synchronize();
    if(status) std::copy(R.status.begin(), R.status.end(), status);
    if(iters) std::copy(R.iters.begin(), R.iters.end(), iters);
    if(def_init) std::copy(R.def_init.begin(), R.def_init.end(), def_init);
    if(def_final) std::copy(R.def_final.begin(), R.def_final.end(), def_final);
    if(hist) std::copy(R.hist.begin(), R.hist.end(), hist);
  });
}

int feat_batchsolver_tail_levels(feat_batchsolver_t s, int* levels, int* grid)
{
  return guarded([&]() {
    if(!s || !levels) throw Error("feat_batchsolver_tail_levels: null argument");
    *levels = s->b->tail_depth();
    if(grid) *grid = s->b->tail_grid() ? 1 : 0;
  });
}

int feat_batchsolver_tiled_levels(feat_batchsolver_t s, int* levels, int cap, int* count)
{
  return guarded([&]() {
    if(!s || !count || (cap > 0 && !levels)) throw Error("feat_batchsolver_tiled_levels: null argument");
    const auto v = s->b->tiled_levels();
    *count = int(v.size());
    for(int i = 0; i < cap && i < int(v.size()); ++i) levels[i] = v[i];
  });
}

int feat_batchsolver_tuning(feat_batchsolver_t s, int* tunes, double* us, int cap, int* count)
{
  return guarded([&]() {
    if(!s) throw Error("feat_batchsolver_tuning: null");
    const auto t = s->b->tunes();
    const auto& u = s->b->tuned_us();
    if(count) *count = int(t.size());
    for(int i = 0; i < cap && i < int(t.size()); ++i)
    {
      if(tunes) tunes[i] = t[i];
      if(us) us[i] = i < int(u.size()) ? u[i] : -1.0;
    }
  });
}

int feat_batchsolver_destroy(feat_batchsolver_t s)
{
  return guarded([&]() { delete s; });
}

int feat_batch_levels(std::uint64_t seed, int ninst, int lmin, int lmax, int* out)
{
  return guarded([&]() {
    const auto l = Assembly::batch_levels(seed, ninst, lmin, lmax);
    std::copy(l.begin(), l.end(), out);
  });
}

int feat_lognormal_cells(std::uint64_t seed, std::uint64_t n, double sigma, double* out)
{
  return guarded([&]() {
    const auto c = Assembly::lognormal_cells(seed, n, sigma);
    std::copy(c.begin(), c.end(), out);
  });
}

int feat_q1_diffusion_assemble(int level, int level_c, const double* coef, double* val, std::uint64_t* col,
                               std::uint64_t* rowptr)
{
  return guarded([&]() {
    if(level < 1 || level_c < level || !coef) throw Error("feat_q1_diffusion_assemble: bad argument");
    std::vector<double> c(coef, coef + (std::size_t(1) << (2 * level_c)));
    const auto A = Assembly::q1_diffusion_2d(level, c, level_c);
    std::copy(A.val.begin(), A.val.end(), val);
    if(col) std::copy(A.col.begin(), A.col.end(), col);
    if(rowptr) std::copy(A.rowptr.begin(), A.rowptr.end(), rowptr);
  });
}

int feat_q1_diffusion3d_assemble(int level, int level_c, const double* coef, double kz, double* val,
                                 std::uint64_t* col, std::uint64_t* rowptr)
{
  return guarded([&]() {
    if(level < 1 || level_c < level || level_c > 8 || !coef || !val || !(kz > 0.0))
      throw Error("feat_q1_diffusion3d_assemble: bad argument");
    std::vector<double> c(coef, coef + (std::size_t(1) << (3 * level_c)));
    const auto A = Assembly::q1_diffusion_3d(level, c, level_c, kz);
    std::copy(A.val.begin(), A.val.end(), val);
    if(col) std::copy(A.col.begin(), A.col.end(), col);
    if(rowptr) std::copy(A.rowptr.begin(), A.rowptr.end(), rowptr);
  });
}

} // extern "C"

struct feat_learned_s
{
  std::unique_ptr<Solver::LearnedMG> mg;
  LAFEM::DenseVector<double, Index> x0;
};

extern "C" {

int feat_learned_create(feat_learned_t* out, const feat_learned_cfg_t* cfg)
{
  return guarded([&]() {
    if(!out || !cfg || cfg->level_coarse < 1 || cfg->level_fine < cfg->level_coarse || cfg->level_fine > 8 || cfg->batch < 1)
      throw Error("feat_learned_create: bad argument");
    Solver::LearnedConfig c;
    c.level_fine = cfg->level_fine;
    c.level_coarse = cfg->level_coarse;
    c.sigma = cfg->sigma;
    c.kz = cfg->kz;
    c.seed = cfg->seed;
    c.nu = cfg->nu;
    c.omega_jac = cfg->omega_jac;
    c.omega_learned = cfg->omega_learned;
    c.tol_rel = cfg->tol_rel;
    c.max_iter = cfg->max_iter;
    c.n_init = cfg->n_init;
    c.n_snap = cfg->n_snap;
    c.batch = cfg->batch;
    c.lr = cfg->lr;
    c.fcg_m = cfg->fcg_m > 0 ? cfg->fcg_m : 1;
    c.learned_steps = (cfg->learned_steps > 0 && cfg->learned_steps < cfg->nu) ? cfg->learned_steps : cfg->nu;
    auto s = std::make_unique<feat_learned_s>();
    s->mg = std::make_unique<Solver::LearnedMG>(c);
    s->mg->init();
    s->x0 = LAFEM::DenseVector<double, Index>(s->mg->size());
    *out = s.release();
  });
}

int feat_learned_generate(feat_learned_t s, int on_policy, double* rho)
{
  return guarded([&]() {
    if(!s) throw Error("feat_learned_generate: null");
    s->mg->generate_data(on_policy != 0);
    if(rho) *rho = s->mg->rho_jacobi();
  });
}

int feat_learned_train(feat_learned_t s, int steps, std::uint64_t seed, float* loss)
{
  return guarded([&]() {
    if(!s || steps < 0) throw Error("feat_learned_train: bad argument");
    const std::size_t before = s->mg->loss_history().size();
    s->mg->train(steps, seed);
    const auto& h = s->mg->loss_history();
    if(loss)
      for(int i = 0; i < steps; ++i) loss[i] = h[before + std::size_t(i)];
  });
}

int feat_learned_solve(feat_learned_t s, int learned, int allow_fallback, double* x, feat_learned_result_t* res,
                       double* hist, std::uint64_t hist_cap)
{
  return guarded([&]() {
    if(!s || !res) throw Error("feat_learned_solve: bad argument");
    LAFEM::DenseVector<double, Index> vx;
    if(x)
      vx = LAFEM::DenseVector<double, Index>::wrap(x, s->mg->size());
    else
    {
      s->x0.format();
      vx = s->x0;
    }
    const auto info = s->mg->solve(vx, s->mg->rhs(), learned != 0, allow_fallback != 0);
    res->status = int(info.status);
    res->iters = info.iters;
    res->fell_back = info.fell_back ? 1 : 0;
    res->fallback_iter = info.fallback_iter;
    res->def_init = info.def_init;
    res->def_final = info.def_final;
    res->seconds = info.seconds;
    if(hist)
      for(std::size_t i = 0; i < info.hist.size() && i < hist_cap; ++i) hist[i] = info.hist[i];
  });
}

int feat_learned_size(feat_learned_t s, std::uint64_t* n, std::uint64_t* m)
{
  return guarded([&]() {
    if(!s) throw Error("feat_learned_size: null");
    if(n) *n = s->mg->size();
    if(m) *m = s->mg->nodes_1d();
  });
}

int feat_learned_fit(feat_learned_t s, int rounds, int steps, int on_policy, uint64_t seed, int* best_round,
                     uint64_t* jacobi_iters, uint64_t* best_iters, uint64_t* round_iters)
{
  return guarded([&]() {
    if(!s || rounds < 1 || steps < 0) throw Error("feat_learned_fit: bad argument");
    const auto fi = s->mg->fit(rounds, steps, on_policy != 0, seed);
    if(best_round) *best_round = fi.best_round;
    if(jacobi_iters) *jacobi_iters = fi.jacobi_iters;
    if(best_iters) *best_iters = fi.best_iters;
    if(round_iters)
      for(size_t r = 0; r < fi.round_iters.size(); ++r) round_iters[r] = fi.round_iters[r];
  });
}

int feat_learned_energy(feat_learned_t s, double* energy)
{
  return guarded([&]() {
    if(!s || !energy) throw Error("feat_learned_energy: null");
    *energy = s->mg->energy();
  });
}

int feat_learned_params(feat_learned_t s, float* get, const float* set)
{
  return guarded([&]() {
    if(!s) throw Error("feat_learned_params: null");
    if(get) s->mg->get_params(get);
    if(set)
    {
      s->mg->set_params(set);
      Runtime::synchronize();
    }
  });
}

int feat_learned_param_ptrs(feat_learned_t s, float** params, std::uint16_t** wbf)
{
  return guarded([&]() {
    if(!s) throw Error("feat_learned_param_ptrs: null");
    if(params) *params = s->mg->params();
    if(wbf) *wbf = s->mg->params_bf16();
  });
}

int feat_learned_destroy(feat_learned_t s)
{
  return guarded([&]() { delete s; });
}

} // extern "C"
