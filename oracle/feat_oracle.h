/*
 * feat_oracle.h -- CPU oracle for the LAFEM + Solver hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * generic (CPU) backend, used exclusively by tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py as the checker.  The product library
 * never links, loads or calls it.
 *
 * Arithmetic convention: the reference's optimised build compiles with
 * -O3 -march=<cpu> (build_system/configure_gcc.py:109-114), so GCC contracts
 * every `s += a*b` and `a*x + y` into one fused multiply-add on FMA hardware.
 * The oracle states each such contraction explicitly with fma() and is
 * itself compiled with -ffp-contract=off, so its rounding is defined by the
 * source and not by compiler flags.  Sums are sequential in index order, as
 * in the generic loops.
 *
 * Index type is uint64_t (reference `Index`, kernel/base_header.hpp).
 */
#ifndef FEAT_ORACLE_H
#define FEAT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded RNG: xorshift64* (kernel/util/random.hpp:84-110, 259-280) ---- */
uint64_t orc_rng_next(uint64_t* state);
/* fills v[0..n) with uniform values in [a,b] exactly as
 * DenseVector(Random&, n, a, b) (lafem/dense_vector.hpp:322-332, 860-869) */
void orc_rng_fill_f64(uint64_t* state, double* v, uint64_t n, double a, double b);

/* ---- CSR (lafem/arch/apply_generic.hpp:25-69) ---- */
void orc_csr_apply_f64(double* r, double a, const double* x, double b, const double* y,
                       const double* val, const uint64_t* col, const uint64_t* rowptr,
                       uint64_t rows, uint64_t cols, int transposed);
void orc_csr_apply_f32(float* r, float a, const float* x, float b, const float* y,
                       const float* val, const uint64_t* col, const uint64_t* rowptr,
                       uint64_t rows, uint64_t cols, int transposed);
/* ---- BCSR bh x bw blocks, row-major (lafem/arch/apply_generic.hpp:121-154) ---- */
void orc_bcsr_apply_f64(int bh, int bw, double* r, double a, const double* x, double b,
                        const double* y, const double* val, const uint64_t* col,
                        const uint64_t* rowptr, uint64_t rows);

/* ---- BLAS-1 (lafem/arch/<op>_generic.hpp) ---- */
void   orc_axpy_f64(double* r, double a, const double* x, const double* y, uint64_t n);
double orc_dot_f64(const double* x, const double* y, uint64_t n);
double orc_triple_dot_f64(const double* x, const double* y, const double* z, uint64_t n);
double orc_norm2_f64(const double* x, uint64_t n);
void   orc_scale_f64(double* r, const double* x, double s, uint64_t n);
void   orc_component_product_f64(double* r, const double* x, const double* y, uint64_t n);
void   orc_component_invert_f64(double* r, const double* x, double s, uint64_t n);
void   orc_diag_index(uint64_t* diag, const uint64_t* col, const uint64_t* rowptr, uint64_t rows);
void   orc_extract_diag_f64(double* diag, const double* val, const uint64_t* col,
                            const uint64_t* rowptr, uint64_t rows);
void   orc_unit_filter_rhs_f64(double* v, const double* sv_val, const uint64_t* sv_idx, uint64_t ue);
void   orc_unit_filter_def_f64(double* v, const uint64_t* sv_idx, uint64_t ue);
/* UnitFilterBlocked filter_rhs (rhs = 1) / filter_def (rhs = 0), block size bs */
void   orc_unit_filter_blocked_f64(int rhs, int bs, double* v, const double* sv_el, const uint64_t* sv_idx,
                                   uint64_t ue, int ign_nans);
void   orc_mean_filter_f64(double* v, const double* vec_prim, const double* vec_dual,
                           double volume, uint64_t n);

/* ---- dense Gauss-Jordan inverse with diagonal pivoting (util/math.hpp:1169-1274) ---- */
double orc_invert_matrix_f64(int n, int stride, double* a, int* p);

/* ---- solver configuration (solver/iterative.hpp:213-236 defaults) ---- */
typedef struct {
  uint64_t rows, cols, nnz;
  const double* val;
  const uint64_t* col;
  const uint64_t* rowptr;
} orc_csr_t;

typedef struct {
  double tol_rel, tol_abs, tol_abs_low, div_rel, div_abs, stag_rate;
  uint64_t min_iter, max_iter, min_stag_iter;
} orc_iter_cfg_t;

/* Status codes of solver/base.hpp:48-64 */
enum { ORC_UNDEFINED = 0, ORC_PROGRESS, ORC_SUCCESS, ORC_ABORTED, ORC_DIVERGED,
       ORC_MAX_ITER, ORC_STAGNATED };

typedef struct {
  int status;
  uint64_t num_iter;
  double def_init, def_final;
} orc_result_t;

void orc_iter_cfg_default(orc_iter_cfg_t* c);

/* coarse solver variants for the multigrid coarsest level */
enum {
  ORC_COARSE_NONE = 0,      /* copy rhs + filter_cor (multigrid.hpp:1560-1566)            */
  ORC_COARSE_SMOOTHER = 1,  /* Richardson-Jacobi with the level's smoother settings        */
  ORC_COARSE_JACOBI = 2,    /* one Jacobi application, omega = coarse_omega                */
  ORC_COARSE_DIRECT = 3     /* dense inverse of the filtered coarse matrix, applied as GEMV */
};

typedef struct {
  orc_csr_t A;              /* system matrix of this level                                  */
  const uint64_t* filt;     /* homogeneous unit-filter indices                              */
  uint64_t nfilt;
  orc_csr_t P;              /* prolongation from the next coarser level (rows = A.rows)     */
  orc_csr_t R;              /* restriction to the next coarser level                        */
} orc_mg_level_t;

typedef struct {
  int nu_pre, nu_post;
  double omega_jac;         /* JacobiPrecond omega (inv_diag = omega / a_ii)                */
  double omega_rich;        /* Richardson damping                                           */
  int coarse;               /* ORC_COARSE_*                                                  */
  double coarse_omega;      /* for ORC_COARSE_JACOBI                                         */
} orc_mg_cfg_t;

/* PCG (solver/pcg.hpp:159-288) preconditioned by a V-cycle multigrid
 * (solver/multigrid.hpp:1270-2000) with Richardson(Jacobi) smoothers
 * (solver/richardson.hpp:160-226, solver/jacobi_precond.hpp:136-164).
 * x: initial guess on input, solution on output.  def_hist (may be NULL) gets
 * def_init and the defect norm after every iteration (max_iter+1 entries). */
int orc_mg_pcg_solve(const orc_mg_level_t* levels, int nlevels, const orc_mg_cfg_t* mg,
                     const orc_iter_cfg_t* it, double* x, const double* b,
                     double* def_hist, orc_result_t* res);

/* The same PCG with an FP32 V-cycle as preconditioner (the ConvertPrecond
 * pattern, solver/convert_precond.hpp:100-119): FP64 outer residual, update and
 * convergence test; FP32 matrices (flags & ORC_MIXED_F16: FP16 values with a
 * power-of-two scale, FP32 arithmetic), FP32 Jacobi diagonals (float(omega) /
 * float(a_ii), as JacobiPrecond<float>), smoothing and transfers; the coarse
 * solve is FP64 for ORC_COARSE_DIRECT and FP32 otherwise.
 * ORC_MIXED_SLIP_INV64 plants a rounding slip (omega / a_ii in FP64, then
 * rounded): it exists only so that a test can show the pin catches it. */
enum { ORC_MIXED_F16 = 1, ORC_MIXED_SLIP_INV64 = 2 };
int orc_mg_pcg_solve_mixed(const orc_mg_level_t* levels, int nlevels, const orc_mg_cfg_t* mg,
                           const orc_iter_cfg_t* it, int flags, double* x, const double* b,
                           double* def_hist, orc_result_t* res);
/* FP16 values of a matrix as the FP16-valued form stores them (see feat_oracle.c) */
void orc_f16_values(const double* val, uint64_t nnz, float* out, float* scale);

/* Stand-alone iterative solvers on one CSR matrix with a unit filter.
 * precond: 0 = none, 1 = Jacobi(omega).  Variant of Solver::solve
 * (solver/iterative.hpp:930-938 -> correct()). */
int orc_pcg_solve(const orc_csr_t* A, const uint64_t* filt, uint64_t nfilt, int precond,
                  double omega, const orc_iter_cfg_t* it, double* x, const double* b,
                  double* def_hist, orc_result_t* res);
int orc_richardson_solve(const orc_csr_t* A, const uint64_t* filt, uint64_t nfilt, int precond,
                         double omega_jac, double omega_rich, const orc_iter_cfg_t* it,
                         double* x, const double* b, double* def_hist, orc_result_t* res);

/* Solver::apply() variants (x formatted to zero, b is the defect). */
int orc_pcg_apply(const orc_csr_t* A, const uint64_t* filt, uint64_t nfilt, int precond,
                  double omega, const orc_iter_cfg_t* it, double* x, const double* b,
                  double* def_hist, orc_result_t* res);
int orc_richardson_apply(const orc_csr_t* A, const uint64_t* filt, uint64_t nfilt, int precond,
                         double omega_jac, double omega_rich, const orc_iter_cfg_t* it,
                         double* x, const double* b, double* def_hist, orc_result_t* res);

/* ---- Stokes: SaddlePointMatrix<BCSR 2x2, 2x1, 1x2>, UnitFilterBlocked +
 * MeanFilter, AmaVanka, MG (feat_oracle_stokes.c) ---- */
typedef struct {
  uint64_t rows, cols;      /* in blocks */
  int bh, bw;
  const double* val;        /* row-major blocks */
  const uint64_t* col;
  const uint64_t* rowptr;
} orc_bcsr_t;
typedef struct {
  uint64_t nv, np;          /* velocity nodes (2 dofs each), pressure dofs */
  orc_bcsr_t A, B, D;       /* system matrix blocks (unfiltered) */
  const uint64_t* filt;     /* velocity unit filter: node indices ... */
  const double* filt_val;   /* ... and 2 values per node */
  uint64_t nfilt;
  const double* prim;       /* pressure mean filter primal / dual vectors */
  const double* dual;
  orc_csr_t Pv, Rv, Pp, Rp; /* scalar Q2 / P1dc transfers to the next coarser level */
} orc_stokes_level_t;
typedef struct {
  int steps;                /* Richardson smoothing steps (min = max) */
  double omega_rich;        /* smoother damping */
  double vanka_omega;       /* AmaVanka omega */
  int coarse_skip_singular; /* AmaVanka::set_skip_singular on the coarsest level */
  double omega_outer;       /* outer Richardson damping */
} orc_stokes_cfg_t;
enum { ORC_FILTER_RHS = 0, ORC_FILTER_SOL = 1, ORC_FILTER_DEF = 2, ORC_FILTER_COR = 3 };
typedef struct orc_vanka_s orc_vanka_t;
/* r = y + alpha*K*x, or r = K*x when y == NULL (saddle_point_matrix.hpp:446-491) */
void orc_saddle_apply(const orc_stokes_level_t* L, double* r, const double* x, const double* y,
                      double alpha);
void orc_stokes_filter(const orc_stokes_level_t* L, double* v, int kind);
/* AmaVanka init on the benchmark's type-1 (filtered) matrix; sizes:
 * {macros, nnz uu, nnz up, nnz pu, nnz pp, stride} */
orc_vanka_t* orc_vanka_create(const orc_stokes_level_t* L, double omega, int skip_singular);
void orc_vanka_sizes(const orc_vanka_t* V, uint64_t* out);
void orc_vanka_copy(const orc_vanka_t* V, uint64_t* ptr[4], uint64_t* col[4], double* val[4], int* mask);
void orc_vanka_apply(const orc_vanka_t* V, const orc_stokes_level_t* L, double* c, const double* d);
void orc_vanka_free(orc_vanka_t* V);
void orc_stokes_transfer(const orc_stokes_level_t* fine, const orc_stokes_level_t* coarse, int prol,
                         double* r, const double* x);
void orc_stokes_smooth(const orc_stokes_level_t* L, int steps, double omega, double vanka_omega,
                       double* sol, const double* rhs);
/* Richardson(omega_outer, MG-V(Richardson(omega_rich, Schwarz(AmaVanka)), coarse
 * Schwarz(AmaVanka))) -- benchmarks/stokes_multigrid-bench.cpp:714-810 */
int orc_stokes_mg_solve(const orc_stokes_level_t* levels, int nl, const orc_stokes_cfg_t* cfg,
                        const orc_iter_cfg_t* it, double* x, const double* b, double* hist,
                        orc_result_t* res);
/* ---- learned smoother network and features (feat_oracle_mlp.c) ---- */
void orc_mlp_forward(const float* p, const uint16_t* X, uint64_t n, int round_bf16, float* y);
void orc_mlp_grad(const float* p, const uint16_t* X, const float* t, uint64_t n, int round_bf16, float* grad);
void orc_mlp_features(uint64_t m, const double* r, const double* e, const double* dinv, const float* cfeat,
                      const uint64_t* node, uint64_t nsamp, uint16_t* X, float* t);
void orc_mlp_smooth(uint64_t m, const float* p, const double* r, const double* dinv, const float* cfeat,
                    double omega, double jacobi, double* x);
/* ---- multicolour SOR / SSOR (solver/sor_precond.cu, ssor_precond.cu semantics) ---- */
uint32_t orc_color_ranks(uint64_t n, const uint64_t* rowptr, const uint64_t* col, int row_colors, uint32_t* rank);
void orc_mcsor_apply(uint64_t n, const uint64_t* rowptr, const uint64_t* col, const double* val,
                     const uint32_t* rank, uint32_t ncolors, double omega, int ssor, const double* x, double* y);
/* ILU(0) (solver/ilu_precond.hpp generic backend): factors, then y = (D+U)^{-1}(I+L)^{-1} x */
void orc_ilu0_factor(uint64_t n, const uint64_t* rowptr, const uint64_t* col, const double* val, double* l,
                     double* u, double* d, uint64_t* rptr_l, uint64_t* rptr_u);
void orc_ilu0_apply(uint64_t n, const uint64_t* rowptr, const uint64_t* col, const double* l, const double* u,
                    const double* d, const uint64_t* rptr_l, const uint64_t* rptr_u, const double* x, double* y);
/* ILU(p) (solver/ilu_precond.hpp generic backend): the level-p split pattern (0, or -1 when
 * cap_l / cap_u are too small), the factors on a pattern, and y = (D+U)^{-1}(I+L)^{-1} x */
int orc_ilup_pattern(uint64_t n, const uint64_t* rowptr, const uint64_t* col, int p, uint64_t cap_l, uint64_t cap_u,
                     uint64_t* rl, uint64_t* cl, uint64_t* ru, uint64_t* cu, uint64_t* nl, uint64_t* nu);
void orc_ilu_factor_pattern(uint64_t n, const uint64_t* rowptr, const uint64_t* col, const double* val,
                            const uint64_t* rl, const uint64_t* cl, const uint64_t* ru, const uint64_t* cu, double* l,
                            double* u, double* d);
void orc_ilu_apply_pattern(uint64_t n, const uint64_t* rl, const uint64_t* cl, const uint64_t* ru, const uint64_t* cu,
                           const double* l, const double* u, const double* d, const double* x, double* y);
/* the generic sequential SOR / SSOR (solver/sor_precond.hpp:201-222, ssor_precond.hpp:177-231) */
void orc_sor_generic_apply(uint64_t n, const uint64_t* rowptr, const uint64_t* col, const double* val, double omega,
                           int ssor, const double* x, double* y);

#ifdef __cplusplus
}
#endif
#endif
