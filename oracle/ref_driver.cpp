// ref_driver.cpp -- C entry points into the reference's own generic (CPU) code.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference's sources where they lie (the reference checkout's kernel/),
// using the reference's own FEAT_NO_CONFIG switch (kernel/base_header.hpp:19),
// into oracle/_ref/libfeatref.so.  Nothing of the reference is copied into
// this repository; this file only calls its public API.  Tests use the
// library to pin the C oracle (oracle/feat_oracle.c) against the real
// reference, and bench.py may time it as the CPU baseline.
#include <kernel/runtime.hpp>
#include <kernel/util/random.hpp>
#include <kernel/util/math.hpp>
#include <kernel/lafem/dense_vector.hpp>
#include <kernel/lafem/sparse_matrix_csr.hpp>
#include <kernel/lafem/sparse_matrix_bcsr.hpp>
#include <kernel/lafem/unit_filter.hpp>
#include <kernel/lafem/none_filter.hpp>
#include <kernel/lafem/transfer.hpp>
#include <kernel/lafem/pointstar_factory.hpp>
#include <kernel/lafem/arch/apply.hpp>
#include <kernel/lafem/arch/axpy.hpp>
#include <kernel/lafem/arch/dot_product.hpp>
#include <kernel/lafem/arch/norm.hpp>
#include <kernel/lafem/arch/scale.hpp>
#include <kernel/lafem/arch/component_product.hpp>
#include <kernel/lafem/arch/component_invert.hpp>
#include <kernel/lafem/arch/diagonal.hpp>
#include <kernel/lafem/arch/unit_filter_blocked.hpp>
#include <kernel/solver/pcg.hpp>
#include <kernel/solver/richardson.hpp>
#include <kernel/solver/jacobi_precond.hpp>
#include <kernel/solver/multigrid.hpp>
#include <kernel/solver/convert_precond.hpp>
#include <kernel/geometry/boundary_factory.hpp>
#include <kernel/geometry/conformal_mesh.hpp>
#include <kernel/geometry/common_factories.hpp>
#include <kernel/geometry/mesh_part.hpp>
#include <kernel/trafo/standard/mapping.hpp>
#include <kernel/space/lagrange1/element.hpp>
#include <kernel/cubature/dynamic_factory.hpp>
#include <kernel/analytic/common.hpp>
#include <kernel/assembly/symbolic_assembler.hpp>
#include <kernel/assembly/unit_filter_assembler.hpp>
#include <kernel/assembly/domain_assembler.hpp>
#include <kernel/assembly/domain_assembler_helpers.hpp>
#include <kernel/assembly/common_operators.hpp>
#include <kernel/assembly/common_functionals.hpp>
#include <kernel/assembly/grid_transfer.hpp>
#include <kernel/lafem/dense_vector_blocked.hpp>
#include <kernel/lafem/sparse_matrix_bwrappedcsr.hpp>
#include <kernel/lafem/saddle_point_matrix.hpp>
#include <kernel/lafem/tuple_vector.hpp>
#include <kernel/lafem/tuple_filter.hpp>
#include <kernel/lafem/tuple_diag_matrix.hpp>
#include <kernel/lafem/unit_filter_blocked.hpp>
#include <kernel/lafem/mean_filter.hpp>
#include <kernel/solver/amavanka.hpp>
#include <kernel/solver/sor_precond.hpp>
#include <kernel/solver/ssor_precond.hpp>
#include <kernel/solver/ilu_precond.hpp>

#include <chrono>
#include <deque>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

using namespace FEAT;
typedef double DT;
typedef Index IT;
typedef LAFEM::DenseVector<DT, IT> Vec;
typedef LAFEM::DenseVector<IT, IT> IVec;
typedef LAFEM::SparseMatrixCSR<DT, IT> Csr;
typedef LAFEM::UnitFilter<DT, IT> UFilter;
typedef LAFEM::Transfer<Csr> Trans;

namespace
{
  // The reference runtime must be initialised once per process.
  struct RuntimeInit
  {
    RuntimeInit()
    {
      static int argc = 1;
      static char name[] = "featref";
      static char* argv_data[] = {name, nullptr};
      static char** argv = argv_data;
      Runtime::initialize(argc, argv);
    }
  };
  void ensure_runtime()
  {
    static RuntimeInit init;
    (void)init;
  }

  Vec make_vec(const double* p, Index n)
  {
    Vec v(n);
    if(n) std::memcpy(v.elements(), p, sizeof(double) * n);
    return v;
  }

  IVec make_ivec(const std::uint64_t* p, Index n)
  {
    IVec v(n);
    for(Index i = 0; i < n; ++i) v.elements()[i] = IT(p[i]);
    return v;
  }
}

extern "C" {

// Mirrors orc_csr_t / orc_mg_level_t of oracle/feat_oracle.h.
struct ref_csr_t
{
  std::uint64_t rows, cols, nnz;
  const double* val;
  const std::uint64_t* col;
  const std::uint64_t* rowptr;
};

struct ref_level_t
{
  ref_csr_t A;
  const std::uint64_t* filt;
  std::uint64_t nfilt;
  ref_csr_t P;
  ref_csr_t R;
};

static Csr make_csr(const ref_csr_t& m)
{
  IVec ci = make_ivec(m.col, m.nnz);
  Vec va = make_vec(m.val, m.nnz);
  IVec rp = make_ivec(m.rowptr, m.rows + 1);
  return Csr(m.rows, m.cols, ci, va, rp);
}

static UFilter make_filter(Index n, const std::uint64_t* idx, std::uint64_t nf)
{
  UFilter f(n);
  for(std::uint64_t i = 0; i < nf; ++i)
    f.add(IT(idx[i]), DT(0));
  return f;
}

int ref_init() { ensure_runtime(); return 1; }

// ---------------- Arch kernels (generic backend) ----------------
void ref_csr_apply_f64(double* r, double a, const double* x, double b, const double* y,
                       const double* val, const std::uint64_t* col, const std::uint64_t* rowptr,
                       std::uint64_t rows, std::uint64_t cols, std::uint64_t nnz, int transposed)
{
  ensure_runtime();
  LAFEM::Arch::Apply::csr_generic(r, a, x, b, y, val, (const IT*)col, (const IT*)rowptr,
                                  Index(rows), Index(cols), Index(nnz), transposed != 0);
}

void ref_bcsr_apply_f64(int bh, int bw, double* r, double a, const double* x, double b,
                        const double* y, const double* val, const std::uint64_t* col,
                        const std::uint64_t* rowptr, std::uint64_t rows, std::uint64_t cols,
                        std::uint64_t nnz)
{
  ensure_runtime();
  const IT* c = (const IT*)col;
  const IT* rp = (const IT*)rowptr;
  if(bh == 2 && bw == 2)
    LAFEM::Arch::Apply::bcsr_generic<2, 2>(r, a, x, b, y, val, c, rp, rows, cols, nnz);
  else if(bh == 2 && bw == 1)
    LAFEM::Arch::Apply::bcsr_generic<2, 1>(r, a, x, b, y, val, c, rp, rows, cols, nnz);
  else if(bh == 1 && bw == 2)
    LAFEM::Arch::Apply::bcsr_generic<1, 2>(r, a, x, b, y, val, c, rp, rows, cols, nnz);
  else if(bh == 3 && bw == 3)
    LAFEM::Arch::Apply::bcsr_generic<3, 3>(r, a, x, b, y, val, c, rp, rows, cols, nnz);
}

void ref_axpy_f64(double* r, double a, const double* x, const double* y, std::uint64_t n)
{
  LAFEM::Arch::Axpy::value_generic(r, a, x, y, Index(n));
}

double ref_dot_f64(const double* x, const double* y, std::uint64_t n)
{
  return LAFEM::Arch::DotProduct::value_generic(x, y, Index(n));
}

double ref_triple_dot_f64(const double* x, const double* y, const double* z, std::uint64_t n)
{
  return LAFEM::Arch::TripleDotProduct::value_generic(x, y, z, Index(n));
}

double ref_norm2_f64(const double* x, std::uint64_t n)
{
  return LAFEM::Arch::Norm2::value_generic(x, Index(n));
}

void ref_scale_f64(double* r, const double* x, double s, std::uint64_t n)
{
  LAFEM::Arch::Scale::value_generic(r, x, s, Index(n));
}

void ref_component_product_f64(double* r, const double* x, const double* y, std::uint64_t n)
{
  LAFEM::Arch::ComponentProduct::value_generic(r, x, y, Index(n));
}

void ref_component_invert_f64(double* r, const double* x, double s, std::uint64_t n)
{
  LAFEM::Arch::ComponentInvert::value_generic(r, x, s, Index(n));
}

// Arch::UnitFilterBlocked::filter_rhs / filter_def (arch/unit_filter_blocked_generic.hpp:24-60)
void ref_unit_filter_blocked_f64(int rhs, int bs, double* v, const double* sv_el, const std::uint64_t* sv_idx,
                                 std::uint64_t ue, int ign_nans)
{
  const IT* ix = (const IT*)sv_idx;
  const bool ig = ign_nans != 0;
  switch(bs)
  {
  case 1:
    if(rhs) LAFEM::Arch::UnitFilterBlocked::filter_rhs_generic<1>(v, sv_el, ix, Index(ue), ig);
    else LAFEM::Arch::UnitFilterBlocked::filter_def_generic<1>(v, sv_el, ix, Index(ue), ig);
    break;
  case 2:
    if(rhs) LAFEM::Arch::UnitFilterBlocked::filter_rhs_generic<2>(v, sv_el, ix, Index(ue), ig);
    else LAFEM::Arch::UnitFilterBlocked::filter_def_generic<2>(v, sv_el, ix, Index(ue), ig);
    break;
  case 3:
    if(rhs) LAFEM::Arch::UnitFilterBlocked::filter_rhs_generic<3>(v, sv_el, ix, Index(ue), ig);
    else LAFEM::Arch::UnitFilterBlocked::filter_def_generic<3>(v, sv_el, ix, Index(ue), ig);
    break;
  default:
    break;
  }
}

void ref_diag_index(std::uint64_t* diag, const std::uint64_t* col, const std::uint64_t* rowptr,
                    std::uint64_t rows)
{
  LAFEM::Arch::Diagonal::csr_generic((IT*)diag, (const IT*)col, (const IT*)rowptr, Index(rows));
}

double ref_invert_matrix_f64(int n, int stride, double* a, int* p)
{
  return Math::invert_matrix(n, stride, a, p);
}

void ref_rng_fill_f64(std::uint64_t seed, double* v, std::uint64_t n, double a, double b)
{
  ensure_runtime();
  Random rng(seed);
  Vec vec(rng, Index(n), a, b);
  std::memcpy(v, vec.elements(), sizeof(double) * n);
}

// ---------------- FD pointstar fixture (lafem/pointstar_factory.hpp) ----------------
// Call once with val == nullptr to get the sizes.
int ref_pointstar_fd(std::uint64_t m, std::uint64_t d, std::uint64_t* n_out,
                     std::uint64_t* nnz_out, double* val, std::uint64_t* col,
                     std::uint64_t* rowptr, double* q2bubble)
{
  ensure_runtime();
  LAFEM::PointstarFactoryFD<DT, IT> psf(m, d);
  Csr a = psf.matrix_csr();
  *n_out = a.rows();
  *nnz_out = a.used_elements();
  if(val)
  {
    std::memcpy(val, a.val(), sizeof(double) * a.used_elements());
    for(Index i = 0; i < a.used_elements(); ++i) col[i] = a.col_ind()[i];
    for(Index i = 0; i <= a.rows(); ++i) rowptr[i] = a.row_ptr()[i];
    Vec q = psf.vector_q2_bubble();
    std::memcpy(q2bubble, q.elements(), sizeof(double) * q.size());
  }
  return 0;
}

// ---------------- solvers on caller-given matrices ----------------
} // extern "C"

// Records the defect norm of every iteration (IterativeSolver::_set_new_defect
// is virtual, solver/iterative.hpp:853).
template<typename Base_>
class Recorder : public Base_
{
public:
  std::vector<double> hist;
  template<typename... A_>
  explicit Recorder(A_&&... args) : Base_(std::forward<A_>(args)...) {}
  typedef typename Base_::VectorType VecT;
  virtual Solver::Status _set_initial_defect(const VecT& d, const VecT& s) override
  {
    Solver::Status st = Base_::_set_initial_defect(d, s);
    hist.assign(1, this->_def_init);
    return st;
  }
  virtual Solver::Status _set_new_defect(const VecT& d, const VecT& s) override
  {
    Solver::Status st = Base_::_set_new_defect(d, s);
    hist.push_back(this->_def_cur);
    return st;
  }
};

typedef Solver::PreconditionedIterativeSolver<Vec> PISolver;

extern "C" {

static void copy_hist(const std::vector<double>& h, double* out, std::uint64_t cap)
{
  if(!out) return;
  for(std::size_t i = 0; i < h.size() && i < cap; ++i) out[i] = h[i];
}

// coarse: 0 none (identity + filter), 1 Richardson-Jacobi smoother, 2 Jacobi(coarse_omega)
int ref_mg_pcg_solve(const ref_level_t* L, int nl, int nu_pre, int nu_post, double omega_jac,
                     double omega_rich, int coarse, double coarse_omega, double tol_rel,
                     std::uint64_t max_iter, double* x, const double* b, double* hist,
                     std::uint64_t hist_cap, std::uint64_t* iters, double* def_final,
                     double* solve_seconds)
{
  ensure_runtime();
  std::vector<Csr> A(nl), P(nl), R(nl);
  std::vector<UFilter> F;
  std::vector<Trans> T(nl);
  F.reserve(nl);
  for(int i = 0; i < nl; ++i)
  {
    A[i] = make_csr(L[i].A);
    F.push_back(make_filter(L[i].A.rows, L[i].filt, L[i].nfilt));
    if(i + 1 < nl)
      T[i] = Trans(make_csr(L[i].P), make_csr(L[i].R));
  }
  auto hierarchy = std::make_shared<Solver::MultiGridHierarchy<Csr, UFilter, Trans>>(nl);
  for(int i = 0; i < nl; ++i)
  {
    auto jac = Solver::new_jacobi_precond(A[i], F[i], omega_jac);
    if(i + 1 < nl)
    {
      auto pre = Solver::new_richardson(A[i], F[i], omega_rich, jac);
      pre->set_min_iter(nu_pre);
      pre->set_max_iter(nu_pre);
      auto post = Solver::new_richardson(A[i], F[i], omega_rich, jac);
      post->set_min_iter(nu_post);
      post->set_max_iter(nu_post);
      hierarchy->push_level(A[i], F[i], T[i], pre, post, pre);
    }
    else
    {
      std::shared_ptr<Solver::SolverBase<Vec>> cs;
      if(coarse == 1)
      {
        auto s = Solver::new_richardson(A[i], F[i], omega_rich, jac);
        s->set_min_iter(nu_pre);
        s->set_max_iter(nu_pre);
        cs = s;
      }
      else if(coarse == 2)
        cs = Solver::new_jacobi_precond(A[i], F[i], coarse_omega);
      hierarchy->push_level(A[i], F[i], cs);
    }
  }
  auto mgv = Solver::new_multigrid(hierarchy, Solver::MultiGridCycle::V);
  auto pcg = std::make_shared<Recorder<Solver::PCG<Csr, UFilter>>>(A[0], F[0], mgv);
  pcg->set_tol_rel(tol_rel);
  pcg->set_max_iter(max_iter);
  hierarchy->init();
  pcg->init();
  Vec vx = make_vec(x, L[0].A.rows);
  Vec vb = make_vec(b, L[0].A.rows);
  const auto t0 = std::chrono::steady_clock::now();
  Solver::Status st = Solver::solve(*pcg, vx, vb, A[0], F[0]);
  const auto t1 = std::chrono::steady_clock::now();
  if(solve_seconds) *solve_seconds = std::chrono::duration<double>(t1 - t0).count();
  std::memcpy(x, vx.elements(), sizeof(double) * L[0].A.rows);
  copy_hist(pcg->hist, hist, hist_cap);
  *iters = pcg->get_num_iter();
  *def_final = pcg->get_def_final();
  pcg->done();
  hierarchy->done();
  return int(st);
}

// kind: 0 = PCG, 1 = Richardson;  precond: 0 none, 1 Jacobi(omega_jac)
// use_apply: 0 -> Solver::solve (correct), 1 -> solver.apply (x formatted)
int ref_single_solve(int kind, const ref_csr_t* m, const std::uint64_t* filt, std::uint64_t nfilt,
                     int precond, double omega_jac, double omega_rich, double tol_rel,
                     std::uint64_t max_iter, int use_apply, double* x, const double* b,
                     double* hist, std::uint64_t hist_cap, std::uint64_t* iters,
                     double* def_final)
{
  ensure_runtime();
  Csr A = make_csr(*m);
  UFilter F = make_filter(m->rows, filt, nfilt);
  std::shared_ptr<Solver::SolverBase<Vec>> pre;
  if(precond) pre = Solver::new_jacobi_precond(A, F, omega_jac);
  std::shared_ptr<PISolver> s;
  std::vector<double>* h = nullptr;
  if(kind == 0)
  {
    auto p = std::make_shared<Recorder<Solver::PCG<Csr, UFilter>>>(A, F, pre);
    h = &p->hist;
    s = p;
  }
  else
  {
    auto p = std::make_shared<Recorder<Solver::Richardson<Csr, UFilter>>>(A, F, omega_rich, pre);
    h = &p->hist;
    s = p;
  }
  s->set_tol_rel(tol_rel);
  s->set_max_iter(max_iter);
  s->init();
  Vec vx = make_vec(x, m->rows);
  Vec vb = make_vec(b, m->rows);
  Solver::Status st;
  if(use_apply)
    st = s->apply(vx, vb);
  else
    st = Solver::solve(*s, vx, vb, A, F);
  std::memcpy(x, vx.elements(), sizeof(double) * m->rows);
  copy_hist(*h, hist, hist_cap);
  *iters = s->get_num_iter();
  *def_final = s->get_def_final();
  s->done();
  return int(st);
}

} // extern "C"

// ---------------- FP32 kernels and the mixed-precision V-cycle ----------------
typedef LAFEM::DenseVector<float, IT> Vec32;
typedef LAFEM::SparseMatrixCSR<float, IT> Csr32;
typedef LAFEM::UnitFilter<float, IT> UFilter32;
typedef LAFEM::Transfer<Csr32> Trans32;

namespace
{
  // A coarse solver of the float hierarchy that applies the FP64 inverse of the
  // filtered coarse matrix to the float right-hand side (one FP64 dot per row in
  // column order, rounded once to float; filtered rows get 0).  The reference
  // plugs coarse solvers in through SolverBase (multigrid.hpp:1549-1557); this
  // one is the driver's own: the product keeps its coarse solve in FP64.
  class DenseInverseCoarse : public Solver::SolverBase<Vec32>
  {
    Index _n;
    std::vector<double> _inv;
    std::vector<char> _mask;
  public:
    DenseInverseCoarse(const Csr& A, const std::uint64_t* filt, std::uint64_t nf) :
      _n(A.rows()), _inv(std::size_t(A.rows() * A.rows()), 0.0), _mask(std::size_t(A.rows()), 0)
    {
      for(std::uint64_t k = 0; k < nf; ++k) _mask[filt[k]] = 1;
      const IT* rp = A.row_ptr();
      const IT* ci = A.col_ind();
      const double* va = A.val();
      for(Index r = 0; r < _n; ++r)
      {
        if(_mask[r]) { _inv[r * _n + r] = 1.0; continue; }
        for(IT i = rp[r]; i < rp[r + 1]; ++i)
          if(!_mask[ci[i]]) _inv[r * _n + ci[i]] = va[i];
      }
      std::vector<int> piv(_n, 0);
      Math::invert_matrix(int(_n), int(_n), _inv.data(), piv.data());
    }
    virtual String name() const override { return "DenseInverseCoarse"; }
    virtual Solver::Status apply(Vec32& cor, const Vec32& def) override
    {
      const float* d = def.elements();
      float* c = cor.elements();
      for(Index r = 0; r < _n; ++r)
      {
        double s = 0.0;
        for(Index k = 0; k < _n; ++k) s = std::fma(_inv[r * _n + k], double(d[k]), s);
        c[r] = _mask[r] ? 0.0f : float(s);
      }
      return Solver::Status::success;
    }
  };

  Csr32 make_csr32(const ref_csr_t& m)
  {
    Csr32 a;
    a.convert(make_csr(m));
    return a;
  }

  UFilter32 make_filter32(Index n, const std::uint64_t* idx, std::uint64_t nf)
  {
    UFilter32 f(n);
    for(std::uint64_t i = 0; i < nf; ++i)
      f.add(IT(idx[i]), 0.0f);
    return f;
  }
}

extern "C" {

void ref_csr_apply_f32(float* r, float a, const float* x, float b, const float* y,
                       const float* val, const std::uint64_t* col, const std::uint64_t* rowptr,
                       std::uint64_t rows, std::uint64_t cols, std::uint64_t nnz, int transposed)
{
  ensure_runtime();
  LAFEM::Arch::Apply::csr_generic(r, a, x, b, y, val, (const IT*)col, (const IT*)rowptr,
                                  Index(rows), Index(cols), Index(nnz), transposed != 0);
}

void ref_component_invert_f32(float* r, const float* x, float s, std::uint64_t n)
{
  LAFEM::Arch::ComponentInvert::value_generic(r, x, s, Index(n));
}

// PCG<double> preconditioned by ConvertPrecond<double -> float> around a float
// MultiGrid V-cycle (convert_precond.hpp:100-119) whose levels hold
// SparseMatrixCSR<float> copies of the given matrices (LAFEM convert: each value
// rounded once), float JacobiPrecond / Richardson smoothers and float Transfer
// operators.  coarse: 1 float Richardson-Jacobi smoother, 2 float Jacobi
// (coarse_omega), 3 FP64 dense inverse (DenseInverseCoarse above).
int ref_mg_pcg_solve_mixed(const ref_level_t* L, int nl, int nu_pre, int nu_post, double omega_jac,
                           double omega_rich, int coarse, double coarse_omega, double tol_rel,
                           std::uint64_t max_iter, double* x, const double* b, double* hist,
                           std::uint64_t hist_cap, std::uint64_t* iters, double* def_final)
{
  ensure_runtime();
  Csr A0 = make_csr(L[0].A);
  UFilter F0 = make_filter(L[0].A.rows, L[0].filt, L[0].nfilt);
  std::vector<Csr32> A(nl);
  std::vector<UFilter32> F;
  std::vector<Trans32> T(nl);
  F.reserve(nl);
  for(int i = 0; i < nl; ++i)
  {
    A[i] = make_csr32(L[i].A);
    F.push_back(make_filter32(L[i].A.rows, L[i].filt, L[i].nfilt));
    if(i + 1 < nl)
      T[i] = Trans32(make_csr32(L[i].P), make_csr32(L[i].R));
  }
  auto hierarchy = std::make_shared<Solver::MultiGridHierarchy<Csr32, UFilter32, Trans32>>(nl);
  for(int i = 0; i < nl; ++i)
  {
    auto jac = Solver::new_jacobi_precond(A[i], F[i], float(omega_jac));
    if(i + 1 < nl)
    {
      auto pre = Solver::new_richardson(A[i], F[i], float(omega_rich), jac);
      pre->set_min_iter(nu_pre);
      pre->set_max_iter(nu_pre);
      auto post = Solver::new_richardson(A[i], F[i], float(omega_rich), jac);
      post->set_min_iter(nu_post);
      post->set_max_iter(nu_post);
      hierarchy->push_level(A[i], F[i], T[i], pre, post, pre);
    }
    else
    {
      std::shared_ptr<Solver::SolverBase<Vec32>> cs;
      if(coarse == 1)
      {
        auto s = Solver::new_richardson(A[i], F[i], float(omega_rich), jac);
        s->set_min_iter(nu_pre);
        s->set_max_iter(nu_pre);
        cs = s;
      }
      else if(coarse == 2)
        cs = Solver::new_jacobi_precond(A[i], F[i], float(coarse_omega));
      else if(coarse == 3)
        cs = std::make_shared<DenseInverseCoarse>(make_csr(L[i].A), L[i].filt, L[i].nfilt);
      hierarchy->push_level(A[i], F[i], cs);
    }
  }
  auto mgv = Solver::new_multigrid(hierarchy, Solver::MultiGridCycle::V);
  auto conv = Solver::new_convert_precond<Vec, Vec32>(mgv);
  auto pcg = std::make_shared<Recorder<Solver::PCG<Csr, UFilter>>>(A0, F0, conv);
  pcg->set_tol_rel(tol_rel);
  pcg->set_max_iter(max_iter);
  hierarchy->init();
  pcg->init();
  Vec vx = make_vec(x, L[0].A.rows);
  Vec vb = make_vec(b, L[0].A.rows);
  Solver::Status st = Solver::solve(*pcg, vx, vb, A0, F0);
  std::memcpy(x, vx.elements(), sizeof(double) * L[0].A.rows);
  copy_hist(pcg->hist, hist, hist_cap);
  *iters = pcg->get_num_iter();
  *def_final = pcg->get_def_final();
  pcg->done();
  hierarchy->done();
  return int(st);
}

} // extern "C"

// ---------------- Q1 Poisson assembly on the refined unit square/cube ----------------
struct ref_q1_t
{
  int dim;
  Index n, nnz, nfilt, pn, pnnz;
  std::vector<double> val, coords, rhs, pval;
  std::vector<std::uint64_t> col, rowptr, filt, pcol, prowptr;
};

template<int dim_>
static ref_q1_t* q1_build(int level, int with_prol, const char* cubature)
{
  typedef Shape::Hypercube<dim_> ShapeType;
  typedef Geometry::ConformalMesh<ShapeType> MeshType;
  typedef Geometry::MeshPart<MeshType> MeshPartType;
  typedef Trafo::Standard::Mapping<MeshType> TrafoType;
  typedef Space::Lagrange1::Element<TrafoType> SpaceType;

  // coarse mesh on level-1 (or level 0), refined once with the StandardRefinery
  const int lc = (with_prol && level > 0) ? level - 1 : level;
  const Index lci = Index(lc);
  Geometry::RefinedUnitCubeFactory<MeshType> fac_c(lci);
  MeshType mesh_c(fac_c);
  std::unique_ptr<MeshType> mesh_f;
  if(lc != level)
  {
    Geometry::StandardRefinery<MeshType> ref(mesh_c);
    mesh_f.reset(new MeshType(ref));
  }
  MeshType& mesh = mesh_f ? *mesh_f : mesh_c;
  TrafoType trafo(mesh);
  SpaceType space(trafo);
  Assembly::DomainAssembler<TrafoType> dasm(trafo);
  dasm.compile_all_elements();

  Csr mat;
  Assembly::SymbolicAssembler::assemble_matrix_std1(mat, space);
  mat.format();
  Assembly::Common::LaplaceOperator lap;
  Assembly::assemble_bilinear_operator_matrix_1(dasm, mat, lap, space, cubature);

  Geometry::BoundaryFactory<MeshType> bfac(mesh);
  MeshPartType boundary(bfac);
  Assembly::UnitFilterAssembler<MeshType> uasm;
  uasm.add_mesh_part(boundary);
  UFilter filter;
  uasm.assemble(filter, space);

  Vec rhs = mat.create_vector_r();
  rhs.format();
  Analytic::Common::SineBubbleFunction<dim_> sol_fn;
  Assembly::Common::LaplaceFunctional<decltype(sol_fn)> force(sol_fn);
  Assembly::assemble_linear_functional_vector(dasm, rhs, force, space, cubature);
  filter.filter_rhs(rhs);

  ref_q1_t* q = new ref_q1_t();
  q->dim = dim_;
  q->n = mat.rows();
  q->nnz = mat.used_elements();
  q->val.assign(mat.val(), mat.val() + q->nnz);
  q->col.assign(mat.col_ind(), mat.col_ind() + q->nnz);
  q->rowptr.assign(mat.row_ptr(), mat.row_ptr() + q->n + 1);
  q->rhs.assign(rhs.elements(), rhs.elements() + q->n);
  q->nfilt = filter.used_elements();
  q->filt.assign(filter.get_indices(), filter.get_indices() + q->nfilt);
  const auto& vtx = mesh.get_vertex_set();
  q->coords.resize(std::size_t(q->n) * dim_);
  for(Index i = 0; i < q->n; ++i)
    for(int k = 0; k < dim_; ++k)
      q->coords[std::size_t(i) * dim_ + k] = double(vtx[i][k]);
  q->pn = 0;
  q->pnnz = 0;
  if(lc != level)
  {
    // prolongation as in control/scalar_basic.hpp:282-331
    TrafoType trafo_c(mesh_c);
    SpaceType space_c(trafo_c);
    Csr prol;
    Assembly::SymbolicAssembler::assemble_matrix_2lvl(prol, space, space_c);
    Vec weight = prol.create_vector_l();
    prol.format();
    weight.format();
    Assembly::GridTransfer::assemble_prolongation(prol, weight, space, space_c, cubature);
    weight.component_invert(weight);
    prol.scale_rows(prol, weight);
    q->pn = prol.columns();
    q->pnnz = prol.used_elements();
    q->pval.assign(prol.val(), prol.val() + q->pnnz);
    q->pcol.assign(prol.col_ind(), prol.col_ind() + q->pnnz);
    q->prowptr.assign(prol.row_ptr(), prol.row_ptr() + q->n + 1);
  }
  return q;
}

extern "C" {

void* ref_q1_create(int dim, int level, int with_prol, const char* cubature)
{
  ensure_runtime();
  if(dim == 2) return q1_build<2>(level, with_prol, cubature);
  if(dim == 3) return q1_build<3>(level, with_prol, cubature);
  return nullptr;
}

void ref_q1_sizes(void* h, std::uint64_t* out)
{
  ref_q1_t* q = (ref_q1_t*)h;
  out[0] = q->n;
  out[1] = q->nnz;
  out[2] = q->nfilt;
  out[3] = q->pn;
  out[4] = q->pnnz;
}

void ref_q1_copy(void* h, double* val, std::uint64_t* col, std::uint64_t* rowptr,
                 std::uint64_t* filt, double* coords, double* rhs, double* pval,
                 std::uint64_t* pcol, std::uint64_t* prowptr)
{
  ref_q1_t* q = (ref_q1_t*)h;
  std::memcpy(val, q->val.data(), sizeof(double) * q->nnz);
  std::memcpy(col, q->col.data(), sizeof(std::uint64_t) * q->nnz);
  std::memcpy(rowptr, q->rowptr.data(), sizeof(std::uint64_t) * (q->n + 1));
  std::memcpy(filt, q->filt.data(), sizeof(std::uint64_t) * q->nfilt);
  std::memcpy(coords, q->coords.data(), sizeof(double) * q->coords.size());
  std::memcpy(rhs, q->rhs.data(), sizeof(double) * q->n);
  if(q->pnnz && pval)
  {
    std::memcpy(pval, q->pval.data(), sizeof(double) * q->pnnz);
    std::memcpy(pcol, q->pcol.data(), sizeof(std::uint64_t) * q->pnnz);
    std::memcpy(prowptr, q->prowptr.data(), sizeof(std::uint64_t) * (q->n + 1));
  }
}

void ref_q1_free(void* h) { delete (ref_q1_t*)h; }

// ---------------- Stokes: SaddlePointMatrix, AmaVanka, multigrid ----------------
// Mirrors orc_bcsr_t / orc_stokes_level_t of oracle/feat_oracle.h.
struct ref_bcsr_t
{
  std::uint64_t rows, cols;
  int bh, bw;
  const double* val;
  const std::uint64_t* col;
  const std::uint64_t* rowptr;
};

struct ref_stokes_level_t
{
  std::uint64_t nv, np;
  ref_bcsr_t A, B, D;
  const std::uint64_t* filt;
  const double* filt_val;
  std::uint64_t nfilt;
  const double* prim;
  const double* dual;
  ref_csr_t Pv, Rv, Pp, Rp;
};

} // extern "C"

namespace
{
  typedef LAFEM::SparseMatrixBCSR<DT, IT, 2, 2> MatA;
  typedef LAFEM::SparseMatrixBCSR<DT, IT, 2, 1> MatB;
  typedef LAFEM::SparseMatrixBCSR<DT, IT, 1, 2> MatD;
  typedef LAFEM::SaddlePointMatrix<MatA, MatB, MatD> SysMat;
  typedef LAFEM::DenseVectorBlocked<DT, IT, 2> VecV;
  typedef LAFEM::TupleVector<VecV, Vec> SysVec;
  typedef LAFEM::UnitFilterBlocked<DT, IT, 2> FiltV;
  typedef LAFEM::MeanFilter<DT, IT> FiltP;
  typedef LAFEM::TupleFilter<FiltV, FiltP> SysFilt;
  typedef LAFEM::SparseMatrixBWrappedCSR<DT, IT, 2> MatTV;
  typedef LAFEM::TupleDiagMatrix<MatTV, Csr> SysTransMat;
  typedef LAFEM::Transfer<SysTransMat> SysTrans;

  template<typename M_>
  M_ make_bcsr(const ref_bcsr_t& m)
  {
    const Index nnz = m.rowptr[m.rows];
    IVec ci = make_ivec(m.col, nnz);
    Vec va = make_vec(m.val, nnz * Index(m.bh * m.bw));
    IVec rp = make_ivec(m.rowptr, m.rows + 1);
    return M_(m.rows, m.cols, ci, va, rp);
  }

  SysMat make_saddle(const ref_stokes_level_t& L)
  {
    return SysMat(make_bcsr<MatA>(L.A), make_bcsr<MatB>(L.B), make_bcsr<MatD>(L.D));
  }

  SysFilt make_sys_filter(const ref_stokes_level_t& L)
  {
    FiltV fv(L.nv);
    for(std::uint64_t i = 0; i < L.nfilt; ++i)
    {
      Tiny::Vector<DT, 2> v;
      v[0] = L.filt_val[2 * i];
      v[1] = L.filt_val[2 * i + 1];
      fv.add(IT(L.filt[i]), v);
    }
    FiltP fp(make_vec(L.prim, L.np), make_vec(L.dual, L.np));
    return SysFilt(std::move(fv), std::move(fp));
  }

  SysVec make_sysvec(const double* p, const ref_stokes_level_t& L)
  {
    VecV u(L.nv);
    Vec q(L.np);
    if(p)
    {
      std::memcpy(u.template elements<LAFEM::Perspective::pod>(), p, sizeof(double) * 2 * L.nv);
      std::memcpy(q.elements(), p + 2 * L.nv, sizeof(double) * L.np);
    }
    else
    {
      u.format();
      q.format();
    }
    return SysVec(std::move(u), std::move(q));
  }

  void copy_sysvec(double* p, const SysVec& v, const ref_stokes_level_t& L)
  {
    std::memcpy(p, v.first().template elements<LAFEM::Perspective::pod>(), sizeof(double) * 2 * L.nv);
    std::memcpy(p + 2 * L.nv, v.rest().first().elements(), sizeof(double) * L.np);
  }

  // type-1 matrix of the benchmark (stokes_multigrid-bench.cpp:652-662)
  SysMat make_type1(const ref_stokes_level_t& L, const SysFilt& F)
  {
    SysMat m = make_saddle(L);
    F.first().filter_mat(m.block_a());
    F.first().filter_offdiag_row_mat(m.block_b());
    return m;
  }

  // exposes the assembled Vanka matrix of the reference AmaVanka
  class VankaX : public Solver::AmaVanka<SysMat, SysFilt>
  {
  public:
    using Solver::AmaVanka<SysMat, SysFilt>::AmaVanka;
    const typename Solver::Intern::AmaVankaMatrixHelper<SysMat>::VankaMatrix& vmat() const { return this->_vanka; }
    const std::vector<int>& mask() const { return this->_macro_mask; }
  };

  // SchwarzPrecond on one process (schwarz_precond.hpp:125-160): local solve, then filter_cor
  class LocalSchwarz : public Solver::SolverBase<SysVec>
  {
    std::shared_ptr<Solver::SolverBase<SysVec>> _local;
    const SysFilt& _filter;

  public:
    LocalSchwarz(std::shared_ptr<Solver::SolverBase<SysVec>> local, const SysFilt& f) : _local(local), _filter(f) {}
    String name() const override { return "Schwarz"; }
    void init_symbolic() override { _local->init_symbolic(); }
    void init_numeric() override { _local->init_numeric(); }
    void done_numeric() override { _local->done_numeric(); }
    void done_symbolic() override { _local->done_symbolic(); }
    Solver::Status apply(SysVec& c, const SysVec& d) override
    {
      Solver::Status st = _local->apply(c, d);
      if(st == Solver::Status::success) _filter.filter_cor(c);
      return st;
    }
  };

  template<typename M_>
  void copy_bcsr(const M_& m, std::uint64_t* ptr, std::uint64_t* col, double* val)
  {
    const Index nr = m.rows(), nnz = m.used_elements();
    for(Index i = 0; i <= nr; ++i) ptr[i] = m.row_ptr()[i];
    for(Index i = 0; i < nnz; ++i) col[i] = m.col_ind()[i];
    std::memcpy(val, m.template val<LAFEM::Perspective::pod>(), sizeof(double) * nnz * Index(M_::BlockHeight * M_::BlockWidth));
  }
}

extern "C" {

// AmaVanka on the type-1 matrix of level L; sizes {macros, nnz uu, up, pu, pp}
void* ref_vanka_create(const ref_stokes_level_t* L, double omega, int skip_singular, std::uint64_t* sizes)
{
  ensure_runtime();
  struct H
  {
    SysFilt F;
    SysMat M;
    std::shared_ptr<VankaX> V;
  };
  H* h = new H{make_sys_filter(*L), SysMat(), nullptr};
  h->M = make_type1(*L, h->F);
  h->V = std::make_shared<VankaX>(h->M, h->F, omega);
  h->V->set_skip_singular(skip_singular != 0);
  h->V->init();
  const auto& vm = h->V->vmat();
  sizes[0] = skip_singular ? h->V->mask().size() : 0;
  sizes[1] = vm.first().first().used_elements();
  sizes[2] = vm.first().rest().first().used_elements();
  sizes[3] = vm.rest().first().first().used_elements();
  sizes[4] = vm.rest().first().rest().first().used_elements();
  return h;
}

void ref_vanka_copy(void* hv, const ref_stokes_level_t* L, std::uint64_t** ptr, std::uint64_t** col, double** val,
                    int* mask)
{
  struct H
  {
    SysFilt F;
    SysMat M;
    std::shared_ptr<VankaX> V;
  };
  H* h = (H*)hv;
  const auto& vm = h->V->vmat();
  copy_bcsr(vm.first().first(), ptr[0], col[0], val[0]);
  copy_bcsr(vm.first().rest().first(), ptr[1], col[1], val[1]);
  copy_bcsr(vm.rest().first().first(), ptr[2], col[2], val[2]);
  copy_bcsr(vm.rest().first().rest().first(), ptr[3], col[3], val[3]);
  if(mask)
    for(std::size_t i = 0; i < h->V->mask().size(); ++i) mask[i] = h->V->mask()[i];
  (void)L;
}

// c = AmaVanka(d) (Vanka matrix apply + filter_cor)
void ref_vanka_apply(void* hv, const ref_stokes_level_t* L, double* c, const double* d)
{
  struct H
  {
    SysFilt F;
    SysMat M;
    std::shared_ptr<VankaX> V;
  };
  H* h = (H*)hv;
  SysVec vc = make_sysvec(nullptr, *L), vd = make_sysvec(d, *L);
  h->V->apply(vc, vd);
  copy_sysvec(c, vc, *L);
}

void ref_vanka_free(void* hv)
{
  struct H
  {
    SysFilt F;
    SysMat M;
    std::shared_ptr<VankaX> V;
  };
  H* h = (H*)hv;
  h->V->done();
  delete h;
}

// r = y + alpha*K*x (y != NULL) or r = K*x
void ref_saddle_apply(const ref_stokes_level_t* L, double* r, const double* x, const double* y, double alpha)
{
  ensure_runtime();
  SysMat K = make_saddle(*L);
  SysVec vr = make_sysvec(nullptr, *L), vx = make_sysvec(x, *L);
  if(y)
  {
    SysVec vy = make_sysvec(y, *L);
    K.apply(vr, vx, vy, alpha);
  }
  else
    K.apply(vr, vx);
  copy_sysvec(r, vr, *L);
}

// kind: 0 rhs, 1 sol, 2 def, 3 cor
void ref_stokes_filter(const ref_stokes_level_t* L, double* v, int kind)
{
  ensure_runtime();
  SysFilt F = make_sys_filter(*L);
  SysVec vv = make_sysvec(v, *L);
  if(kind == 0) F.filter_rhs(vv);
  else if(kind == 1) F.filter_sol(vv);
  else if(kind == 2) F.filter_def(vv);
  else F.filter_cor(vv);
  copy_sysvec(v, vv, *L);
}

void ref_stokes_transfer(const ref_stokes_level_t* fine, const ref_stokes_level_t* coarse, int prol, double* r,
                         const double* x)
{
  ensure_runtime();
  SysTransMat P(MatTV(make_csr(fine->Pv)), make_csr(fine->Pp));
  SysTransMat R(MatTV(make_csr(fine->Rv)), make_csr(fine->Rp));
  SysTrans T(std::move(P), std::move(R));
  if(prol)
  {
    SysVec vf = make_sysvec(nullptr, *fine), vc = make_sysvec(x, *coarse);
    T.prol(vf, vc);
    copy_sysvec(r, vf, *fine);
  }
  else
  {
    SysVec vf = make_sysvec(x, *fine), vc = make_sysvec(nullptr, *coarse);
    T.rest(vf, vc);
    copy_sysvec(r, vc, *coarse);
  }
}

void ref_stokes_smooth(const ref_stokes_level_t* L, int steps, double omega, double vanka_omega, double* sol,
                       const double* rhs)
{
  ensure_runtime();
  SysFilt F = make_sys_filter(*L);
  SysMat K = make_saddle(*L);
  SysMat K1 = make_type1(*L, F);
  auto vanka = Solver::new_amavanka(K1, F, vanka_omega);
  auto schwarz = std::make_shared<LocalSchwarz>(vanka, F);
  auto smoother = Solver::new_richardson(K, F, omega, schwarz);
  smoother->set_min_iter(steps);
  smoother->set_max_iter(steps);
  smoother->init();
  SysVec vs = make_sysvec(nullptr, *L), vr = make_sysvec(rhs, *L);
  smoother->apply(vs, vr);
  copy_sysvec(sol, vs, *L);
  smoother->done();
}

// Richardson(omega_outer, MG-V(Richardson(omega_rich, Schwarz(AmaVanka))^steps, coarse
// Schwarz(AmaVanka, skip_singular))) -- stokes_multigrid-bench.cpp:714-810, one process
int ref_stokes_mg_solve(const ref_stokes_level_t* L, int nl, int steps, double omega_rich, double vanka_omega,
                        int coarse_skip_singular, double omega_outer, double tol_rel, std::uint64_t max_iter,
                        double* x, const double* b, double* hist, std::uint64_t hist_cap, std::uint64_t* iters,
                        double* def_final, double* solve_seconds)
{
  ensure_runtime();
  std::deque<SysMat> K, K1;
  std::deque<SysFilt> F;
  std::deque<SysTrans> T;
  for(int i = 0; i < nl; ++i)
  {
    F.push_back(make_sys_filter(L[i]));
    K.push_back(make_saddle(L[i]));
    K1.push_back(make_type1(L[i], F.back()));
    if(i + 1 < nl)
    {
      SysTransMat P(MatTV(make_csr(L[i].Pv)), make_csr(L[i].Pp));
      SysTransMat R(MatTV(make_csr(L[i].Rv)), make_csr(L[i].Rp));
      T.emplace_back(std::move(P), std::move(R));
    }
    else
      T.emplace_back();
  }
  auto hierarchy = std::make_shared<Solver::MultiGridHierarchy<SysMat, SysFilt, SysTrans>>(nl);
  for(int i = 0; i < nl; ++i)
  {
    auto vanka = Solver::new_amavanka(K1[i], F[i], vanka_omega);
    if(i + 1 < nl)
    {
      auto schwarz = std::make_shared<LocalSchwarz>(vanka, F[i]);
      auto smoother = Solver::new_richardson(K[i], F[i], omega_rich, schwarz);
      smoother->set_min_iter(steps);
      smoother->set_max_iter(steps);
      hierarchy->push_level(K[i], F[i], T[i], smoother, smoother, smoother);
    }
    else
    {
      vanka->set_skip_singular(coarse_skip_singular != 0);
      auto schwarz = std::make_shared<LocalSchwarz>(vanka, F[i]);
      hierarchy->push_level(K[i], F[i], schwarz);
    }
  }
  auto mgv = Solver::new_multigrid(hierarchy, Solver::MultiGridCycle::V);
  auto solver = std::make_shared<Recorder<Solver::Richardson<SysMat, SysFilt>>>(K[0], F[0], omega_outer, mgv);
  solver->set_tol_rel(tol_rel);
  solver->set_max_iter(max_iter);
  hierarchy->init();
  solver->init();
  SysVec vx = make_sysvec(x, L[0]);
  SysVec vb = make_sysvec(b, L[0]);
  const auto t0 = std::chrono::steady_clock::now();
  Solver::Status st = solver->correct(vx, vb);
  const auto t1 = std::chrono::steady_clock::now();
  if(solve_seconds) *solve_seconds = std::chrono::duration<double>(t1 - t0).count();
  copy_sysvec(x, vx, L[0]);
  copy_hist(solver->hist, hist, hist_cap);
  *iters = solver->get_num_iter();
  *def_final = solver->get_def_final();
  solver->done();
  hierarchy->done();
  return int(st);
}

} // extern "C"

// ---------------- LAFEM file I/O (SparseMatrixCSR / DenseVector read_from, write_out) ----------------
// mode is the reference's FileMode value (lafem/base.hpp:31-47).  Valid inputs only: the
// reference aborts the process on a malformed file.
extern "C" {

int ref_csr_write(int mode, const char* path, const ref_csr_t* m, int symmetric)
{
  ensure_runtime();
  Csr a = make_csr(*m);
  a.write_out(LAFEM::FileMode(mode), String(path), symmetric != 0);
  return 0;
}

// First call with cap 0 for the sizes; a second with cap >= nnz (and rows + 1) copies the arrays.
int ref_csr_read(int mode, const char* path, std::uint64_t* rows, std::uint64_t* cols, std::uint64_t* nnz,
                 double* val, std::uint64_t* col, std::uint64_t* rowptr, std::uint64_t cap)
{
  ensure_runtime();
  Csr a;
  a.read_from(LAFEM::FileMode(mode), String(path));
  *rows = a.rows();
  *cols = a.columns();
  *nnz = a.used_elements();
  if(cap >= a.used_elements() && cap >= a.rows() + 1)
  {
    for(Index i = 0; i < a.used_elements(); ++i)
    {
      val[i] = a.val()[i];
      col[i] = a.col_ind()[i];
    }
    for(Index i = 0; i <= a.rows(); ++i) rowptr[i] = a.row_ptr()[i];
  }
  return 0;
}

int ref_dv_write(int mode, const char* path, const double* v, std::uint64_t n)
{
  ensure_runtime();
  Vec x = make_vec(v, n);
  x.write_out(LAFEM::FileMode(mode), String(path));
  return 0;
}

int ref_dv_read(int mode, const char* path, std::uint64_t* n, double* out, std::uint64_t cap)
{
  ensure_runtime();
  Vec x;
  x.read_from(LAFEM::FileMode(mode), String(path));
  *n = x.size();
  if(cap >= x.size())
    for(Index i = 0; i < x.size(); ++i) out[i] = x.elements()[i];
  return 0;
}

} // extern "C"

// ---------------- generic SOR / SSOR / ILU(p) preconditioners (CPU backend) ----------------
// kind 0: SORPrecond, 1: SSORPrecond (omega), 2: ILUPrecond (level p = int(omega));
// y = M^{-1} x with the empty unit filter.
extern "C" {

int ref_precond_apply(int kind, const ref_csr_t* m, double omega, const double* x, double* y)
{
  ensure_runtime();
  Csr A = make_csr(*m);
  UFilter F(m->rows);
  std::shared_ptr<Solver::SolverBase<Vec>> pre;
  if(kind == 0)
    pre = Solver::new_sor_precond(PreferredBackend::generic, A, F, omega);
  else if(kind == 1)
    pre = Solver::new_ssor_precond(PreferredBackend::generic, A, F, omega);
  else
    pre = Solver::new_ilu_precond(PreferredBackend::generic, A, F, int(omega));
  pre->init();
  Vec vx = make_vec(x, m->rows);
  Vec vy(m->rows, 0.0);
  pre->apply(vy, vx);
  std::memcpy(y, vy.elements(), sizeof(double) * m->rows);
  pre->done();
  return 0;
}

} // extern "C"

// ---------------- Q2/P1dc Stokes assembly by the reference (amavanka-test.cpp:86-117) ----------------
#include <kernel/space/lagrange2/element.hpp>
#include <kernel/space/discontinuous/element.hpp>
#include <kernel/assembly/bilinear_operator_assembler.hpp>
#include <kernel/assembly/gpdv_assembler.hpp>

struct ref_stokes_asm_t
{
  Index nv = 0, np = 0, nvert = 0, nedge = 0, ncell = 0;
  std::vector<double> a_val, b_val, d_val, coords;
  std::vector<std::uint64_t> a_col, a_ptr, b_col, b_ptr, d_col, d_ptr, edges, cells;
};

extern "C" {

// A = DuDvOperatorBlocked<2> (BCSR 2x2), B / D = GradPresDivVeloAssembler (BCSR 2x1 / 1x2) on the
// unit square refined `level` times, Lagrange2 velocity and StdPolyP<1> discontinuous pressure,
// cubature auto-degree:5; the mesh's vertex coordinates, edge->vertex and cell->vertex lists
// let the caller place every dof.
void* ref_stokes_asm_create(int level)
{
  ensure_runtime();
  typedef Shape::Hypercube<2> ShapeType;
  typedef Geometry::ConformalMesh<ShapeType> MeshType;
  typedef Trafo::Standard::Mapping<MeshType> TrafoType;
  typedef Space::Lagrange2::Element<TrafoType> SpaceV;
  typedef Space::Discontinuous::Element<TrafoType, Space::Discontinuous::Variant::StdPolyP<1>> SpaceP;
  const Index lvl = Index(level);
  Geometry::RefinedUnitCubeFactory<MeshType> fac(lvl);
  MeshType mesh(fac);
  TrafoType trafo(mesh);
  SpaceV space_v(trafo);
  SpaceP space_p(trafo);
  Cubature::DynamicFactory cubature("auto-degree:5");
  MatA mat_a;
  MatB mat_b;
  MatD mat_d;
  Assembly::SymbolicAssembler::assemble_matrix_std1(mat_a, space_v);
  mat_a.format();
  Assembly::Common::DuDvOperatorBlocked<2> dudv_op;
  Assembly::BilinearOperatorAssembler::assemble_matrix1(mat_a, dudv_op, space_v, cubature);
  Assembly::GradPresDivVeloAssembler::assemble(mat_b, mat_d, space_v, space_p, cubature);
  auto* s = new ref_stokes_asm_t();
  s->nv = mat_a.rows();
  s->np = mat_d.rows();
  auto grab = [](auto& m, int bs, std::vector<double>& val, std::vector<std::uint64_t>& col,
                 std::vector<std::uint64_t>& ptr) {
    const Index nnz = m.used_elements();
    val.assign(m.template val<LAFEM::Perspective::pod>(), m.template val<LAFEM::Perspective::pod>() + nnz * Index(bs));
    col.assign(m.col_ind(), m.col_ind() + nnz);
    ptr.assign(m.row_ptr(), m.row_ptr() + m.rows() + 1);
  };
  grab(mat_a, 4, s->a_val, s->a_col, s->a_ptr);
  grab(mat_b, 2, s->b_val, s->b_col, s->b_ptr);
  grab(mat_d, 2, s->d_val, s->d_col, s->d_ptr);
  const auto& vtx = mesh.get_vertex_set();
  s->nvert = vtx.get_num_vertices();
  s->coords.resize(2 * s->nvert);
  for(Index i = 0; i < s->nvert; ++i)
    for(int k = 0; k < 2; ++k) s->coords[2 * i + k] = double(vtx[i][k]);
  const auto& iev = mesh.template get_index_set<1, 0>();
  s->nedge = iev.get_num_entities();
  for(Index e = 0; e < s->nedge; ++e)
    for(int k = 0; k < 2; ++k) s->edges.push_back(iev[e][k]);
  const auto& icv = mesh.template get_index_set<2, 0>();
  s->ncell = icv.get_num_entities();
  for(Index c = 0; c < s->ncell; ++c)
    for(int k = 0; k < 4; ++k) s->cells.push_back(icv[c][k]);
  return s;
}

// out: nv, np, nnz(A), nnz(B), nnz(D) (blocks), vertices, edges, cells
void ref_stokes_asm_sizes(void* h, std::uint64_t* out)
{
  auto* s = (ref_stokes_asm_t*)h;
  out[0] = s->nv;
  out[1] = s->np;
  out[2] = s->a_col.size();
  out[3] = s->b_col.size();
  out[4] = s->d_col.size();
  out[5] = s->nvert;
  out[6] = s->nedge;
  out[7] = s->ncell;
}

void ref_stokes_asm_copy(void* h, double* a_val, std::uint64_t* a_col, std::uint64_t* a_ptr, double* b_val,
                         std::uint64_t* b_col, std::uint64_t* b_ptr, double* d_val, std::uint64_t* d_col,
                         std::uint64_t* d_ptr, double* coords, std::uint64_t* edges, std::uint64_t* cells)
{
  auto* s = (ref_stokes_asm_t*)h;
  std::copy(s->a_val.begin(), s->a_val.end(), a_val);
  std::copy(s->a_col.begin(), s->a_col.end(), a_col);
  std::copy(s->a_ptr.begin(), s->a_ptr.end(), a_ptr);
  std::copy(s->b_val.begin(), s->b_val.end(), b_val);
  std::copy(s->b_col.begin(), s->b_col.end(), b_col);
  std::copy(s->b_ptr.begin(), s->b_ptr.end(), b_ptr);
  std::copy(s->d_val.begin(), s->d_val.end(), d_val);
  std::copy(s->d_col.begin(), s->d_col.end(), d_col);
  std::copy(s->d_ptr.begin(), s->d_ptr.end(), d_ptr);
  std::copy(s->coords.begin(), s->coords.end(), coords);
  std::copy(s->edges.begin(), s->edges.end(), edges);
  std::copy(s->cells.begin(), s->cells.end(), cells);
}

void ref_stokes_asm_free(void* h) { delete (ref_stokes_asm_t*)h; }

} // extern "C"
