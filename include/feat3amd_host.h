/*
 * feat3amd_host.h -- C entry points of the C++ host layer (libfeat3amd_host.so).
 *
 * The host layer (include/feat3amd/{lafem,solver,assembly}.hpp) keeps the
 * reference's container and Solver::SolverBase API in C++; these functions
 * expose a solver tree and the structured problem setup to callers that are
 * not C++ (tests, benchmarks, FFI).  Device pointers as in feat3amd.h.
 */
#ifndef FEAT3AMD_HOST_H
#define FEAT3AMD_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- structured Q1 Poisson setup (host arrays; SymbolicAssembler + LaplaceOperator,
 * GridTransfer::assemble_prolongation, UnitFilterAssembler; see assembly.hpp) ----
 * sizes[0..4] = n, nnz(A), #boundary, nnz(P from level-1), cols(P) */
int feat_q1_sizes(int dim, int level, uint64_t* sizes);
int feat_q1_assemble(int dim, int level, double* val, uint64_t* col, uint64_t* rowptr,
                     uint64_t* bnd, double* pval, uint64_t* pcol, uint64_t* prowptr);
int feat_q1_rhs_sine(int dim, int level, double* b);
int feat_rhs_random(uint64_t seed, uint64_t n, double a, double b, double* out);

/* ---- LAFEM file I/O (SparseMatrixCSR / DenseVector read_from, write_out) ----
 * mode = FileMode (lafem/base.hpp:31-47): 0 fm_exp, 1 fm_dv, 2 fm_mtx, 4 fm_csr, 13 fm_binary.
 * Matrices: fm_csr / fm_binary (container image, exact), fm_mtx (MatrixMarket coordinate,
 * symmetric != 0 writes the lower triangle as "symmetric"); sparse_matrix_csr.hpp:877-1087.
 * Vectors: fm_dv / fm_binary, fm_mtx (array n x 1), fm_exp; dense_vector.hpp:549-723.
 * Host arrays.  A read returns an object holding the arrays: query the sizes, fetch, free. */
typedef struct feat_io_s* feat_io_t;
int feat_csr_write_file(int mode, const char* path, uint64_t rows, uint64_t cols, const double* val,
                        const uint64_t* col, const uint64_t* rowptr, int symmetric);
int feat_csr_read_file(int mode, const char* path, feat_io_t* out, uint64_t* rows, uint64_t* cols,
                       uint64_t* nnz);
int feat_io_fetch_csr(feat_io_t f, double* val, uint64_t* col, uint64_t* rowptr);
int feat_dv_write_file(int mode, const char* path, const double* v, uint64_t n);
int feat_dv_read_file(int mode, const char* path, feat_io_t* out, uint64_t* n);
int feat_io_fetch_dv(feat_io_t f, double* v);
int feat_io_free(feat_io_t f);

/* ---- Q2/P1dc Stokes levels on the unit square (lid-driven cavity) ----
 * Problem setup of benchmarks/stokes_multigrid-bench.cpp:576-697 (velocity
 * blocks A = DuDv, B = -grad, D = B^T, UnitFilterBlocked velocity boundary,
 * pressure MeanFilter, Q2 / P1dc prolongations), lexicographic numbering.
 * sizes: {nv, np, nnz A, nnz B, nnz D, nfilt, nnz Pv, nnz Pp, nv coarse, np coarse} */
int feat_stokes_sizes(int level, uint64_t* sizes);
typedef struct {
  double* a_val; uint64_t* a_col; uint64_t* a_ptr;   /* 2x2 blocks */
  double* b_val; uint64_t* b_col; uint64_t* b_ptr;   /* 2x1 blocks */
  double* d_val; uint64_t* d_col; uint64_t* d_ptr;   /* 1x2 blocks */
  uint64_t* filt; double* filt_val;                  /* 2 values per node */
  double* prim; double* dual;
  double* pv_val; uint64_t* pv_col; uint64_t* pv_ptr; /* may be NULL on level 0 */
  double* pp_val; uint64_t* pp_col; uint64_t* pp_ptr;
} feat_stokes_arrays_t;
int feat_stokes_assemble(int level, const feat_stokes_arrays_t* out);

/* ---- PCG(MultiGrid-V(Richardson(Jacobi))) solver tree ----
 * solver/pcg.hpp + solver/multigrid.hpp + solver/richardson.hpp + solver/jacobi_precond.hpp */
typedef struct feat_mgpcg_s* feat_mgpcg_t;

#define FEAT_COARSE_NONE 0      /* copy + filter_cor                       */
#define FEAT_COARSE_SMOOTHER 1  /* Richardson-Jacobi, nu_pre steps          */
#define FEAT_COARSE_JACOBI 2    /* one Jacobi application (coarse_omega)    */
#define FEAT_COARSE_DIRECT 3    /* dense inverse of the filtered matrix     */

typedef struct {
  int nu_pre, nu_post;
  double omega_jac, omega_rich;
  int coarse;
  double coarse_omega;
  double tol_rel;
  uint64_t max_iter;
  int force_generic;      /* 1: reference op-by-op path instead of the fused one */
  int csr_only;           /* 1: fused kernels on the CSR arrays instead of SELL-64 plans */
  int precision;          /* 0: FP64 V-cycle; 1: FP32 V-cycle inside the FP64 PCG
                             (ConvertPrecond, solver/convert_precond.hpp:100-119);
                             2: the FP32 cycle on FP16-valued matrices */
  int no_graphs;          /* 1: launch every kernel eagerly instead of replaying HIP graphs */
  int no_autotune;        /* 1: default SELL-64 launch shapes (no autotuning at init) */
  int no_coarse_fusion;   /* 1: launch every coarse-level operation separately (feat_mg_tail_f64 off) */
} feat_mgpcg_cfg_t;

typedef struct {          /* one level, host arrays; P maps the next coarser level onto this one */
  uint64_t rows;
  const double* val;
  const uint64_t* col;
  const uint64_t* rowptr;
  const uint64_t* filt;
  uint64_t nfilt;
  uint64_t pcols;
  const double* pval;
  const uint64_t* pcol;
  const uint64_t* prowptr;
} feat_level_host_t;

typedef struct {
  int status;             /* Solver::Status (solver/base.hpp:48-64) */
  uint64_t iters;
  double def_init, def_final;
} feat_solve_result_t;

int feat_mgpcg_create(feat_mgpcg_t* s, const feat_level_host_t* levels, int nlevels,
                      const feat_mgpcg_cfg_t* cfg);
/* builds the structured Q1 hierarchy levels lf (finest) .. lc (coarsest) itself: levels
 * above lc on the device (feat_q1_assemble_device), the coarsest on the host (its direct
 * solver factorises the host copy) */
int feat_mgpcg_create_q1(feat_mgpcg_t* s, int dim, int lf, int lc, const feat_mgpcg_cfg_t* cfg);
/* flags: FEAT_Q1_HOST_ASSEMBLY assembles every level on the host and uploads it (the
 * arrays are bit-identical either way) */
#define FEAT_Q1_HOST_ASSEMBLY 1
int feat_mgpcg_create_q1_ex(feat_mgpcg_t* s, int dim, int lf, int lc, const feat_mgpcg_cfg_t* cfg, int flags);
/* Solver::solve(pcg, x, b, A, F) -> PCG::correct: x is the initial guess (device), b device */
int feat_mgpcg_solve(feat_mgpcg_t s, double* x, const double* b, feat_solve_result_t* res,
                     double* hist, uint64_t hist_cap);
/* SparseMatrixCSR::write_out of level lvl's (device-resident) system matrix */
int feat_mgpcg_write_matrix(feat_mgpcg_t s, int lvl, int mode, const char* path, int symmetric);
/* device arrays of level `lvl`'s system matrix (0 = finest) */
int feat_mgpcg_matrix(feat_mgpcg_t s, int lvl, uint64_t* rows, uint64_t* nnz, const double** val,
                      const uint64_t** col, const uint64_t** rowptr);
/* New values (host array, same pattern) for level lvl's system matrix, then the
 * reference's numeric re-initialisation order: hierarchy init_numeric, then the
 * solver tree's init_numeric (multigrid.hpp:1225-1240) */
int feat_mgpcg_set_values(feat_mgpcg_t s, int lvl, const double* val_host);
/* first level of the coarse levels that run fused as one workgroup (feat_mg_tail_f64 /
 * feat_mg_tail_f32) in s's V-cycle; 0 = none.  No counterpart in the reference. */
int feat_mgpcg_coarse_fusion_level(feat_mgpcg_t s, int* level);
/* the number of fused levels that run in the grid form (feat_mg_tail_level_t sten + dlinear) */
int feat_mgpcg_coarse_fusion_grid(feat_mgpcg_t s, int* levels);
int feat_mgpcg_destroy(feat_mgpcg_t s);

/* ---- single Krylov solver with a device preconditioner ----
 * PCG (solver = 0, pcg.hpp) or Richardson (solver = 1, richardson.hpp, damping omega_rich)
 * around precond 0 none, 1 JacobiPrecond, 2 SORPrecond, 3 SSORPrecond (omega_pre =
 * damping; level-set order, the generic sweeps exactly), 4 ILUPrecond (ILU(p), the fill-in
 * level p = omega_pre, a whole number >= 0), 5 / 6 SOR /
 * SSOR in the multicolour order of the CUDA backend; homogeneous unit filter on filt[0..nfilt).  The matrix
 * is host CSR (copied to the device); x (initial guess) and b are device vectors. */
#define FEAT_PRECOND_NONE 0
#define FEAT_PRECOND_JACOBI 1
#define FEAT_PRECOND_SOR 2
#define FEAT_PRECOND_SSOR 3
#define FEAT_PRECOND_ILU 4
#define FEAT_PRECOND_SOR_MULTICOLOR 5
#define FEAT_PRECOND_SSOR_MULTICOLOR 6
typedef struct feat_krylov_s* feat_krylov_t;
int feat_krylov_create(feat_krylov_t* s, int solver, uint64_t n, const double* val, const uint64_t* col,
                       const uint64_t* rowptr, const uint64_t* filt, uint64_t nfilt, int precond,
                       double omega_pre, double omega_rich, double tol_rel, uint64_t max_iter);
int feat_krylov_solve(feat_krylov_t s, double* x, const double* b, feat_solve_result_t* res, double* hist,
                      uint64_t hist_cap);
/* cor = M^{-1} def of the preconditioner alone (device vectors) */
int feat_krylov_precond_apply(feat_krylov_t s, double* cor, const double* def);
int feat_krylov_destroy(feat_krylov_t s);
const char* feat_host_last_error(void);
/* Binds the host layer's runtime (Runtime::initialize, kernel/runtime.cpp:37-80)
 * to `device` and `stream` (NULL: the null stream).  Call before any other
 * host-layer function; later calls are ignored once initialised. */
int feat_runtime_init(int device, void* stream);
/* Flags of the SELL-64 plans the host layer builds from now on (feat3amd.h:
 * FEAT_SELL_NO_DICT keeps full values and column offsets even where the pair
 * dictionary applies).  No counterpart in the reference (format selection). */
int feat_runtime_sell_flags(int flags);
/* multigrid solvers initialised from now on use the grid forms where the levels qualify
 * (1, default): d-linear transfers as FEAT_FMT_DLINEAR descriptors, the fused coarse levels'
 * stencils and transfers formed in the kernel; 0: the plans and CSR transfers throughout */
int feat_runtime_grid_forms(int on);

/* ---- Stokes multigrid: Richardson(MG-V(Richardson(Schwarz(AmaVanka)))) ----
 * benchmarks/stokes_multigrid-bench.cpp:714-810 on levels level_fine..level_coarse
 * of the lid-driven cavity (one process). */
typedef struct {
  int steps;               /* smoother steps, min = max (default 5)         */
  double omega_rich;       /* smoother damping (0.7)                        */
  double vanka_omega;      /* AmaVanka omega (1.0)                          */
  int coarse_skip_singular;/* AmaVanka::set_skip_singular on the coarse level */
  double omega_outer;      /* outer Richardson damping (1.0)                */
  double tol_rel;
  uint64_t max_iter;
  int no_graphs;           /* 1: launch every kernel eagerly (no recorded iteration) */
  int no_coarse_fusion;    /* 1: the coarsest levels as separate launches (feat_stokes_tail_f64 off) */
  uint64_t tail_max_level_dofs;  /* largest level run in that workgroup (0: 512) */
} feat_stokes_cfg_t;
typedef struct feat_stokes_s* feat_stokes_t;
/* builds the levels and runs init_symbolic + init_numeric (AmaVanka on device) */
int feat_stokes_create(feat_stokes_t* s, int level_fine, int level_coarse, const feat_stokes_cfg_t* cfg);
/* x (initial guess, overwritten), b: device arrays of 2*nv + np doubles */
int feat_stokes_solve(feat_stokes_t s, double* x, const double* b, feat_solve_result_t* res,
                      double* hist, uint64_t hist_cap);
/* re-runs every level's AmaVanka::init_numeric; *seconds = wall time incl. sync */
int feat_stokes_init_numeric(feat_stokes_t s, double* seconds);
/* sizes of level lvl (0 = finest): {nv, np, macros, stride, nnz uu, up, pu, pp} */
int feat_stokes_sizes_of(feat_stokes_t s, int lvl, uint64_t* out);
/* host copies of the Vanka blocks (arrays of 4 pointers each; any may be NULL) and mask */
int feat_stokes_vanka_copy(feat_stokes_t s, int lvl, uint64_t** ptr, uint64_t** col, double** val, int* mask);
/* device-pointer operations on level lvl, for tests:
 * r = K x (y == NULL) or y + alpha K x;  c = Vanka(d) with `passes` filter_cor calls;
 * one smoother application;  filter kind 0 rhs, 1 sol, 2 def, 3 cor;
 * transfer between lvl and lvl+1: prol (r on lvl) or rest (r on lvl+1) */
int feat_stokes_apply(feat_stokes_t s, int lvl, double* r, const double* x, const double* y, double alpha);
int feat_stokes_vanka_apply(feat_stokes_t s, int lvl, double* c, const double* d, int passes);
int feat_stokes_smooth(feat_stokes_t s, int lvl, double* sol, const double* rhs);
int feat_stokes_filter(feat_stokes_t s, int lvl, double* v, int kind);
int feat_stokes_transfer(feat_stokes_t s, int lvl, int prol, double* r, const double* x);
int feat_stokes_destroy(feat_stokes_t s);

/* ---- batch of independent 2D Q1 diffusion problems (uneven sizes) ----
 * Instance i: level levels[i] (uniform in [lmin, lmax] from `seed`), log-normal
 * cell coefficient exp(sigma z) (seed + 1000 + i), PCG(MG-V(4,4) Richardson(Jacobi
 * 0.5), coarse level 1 Jacobi) to tol_rel, all instances advanced together. */
typedef struct {
  uint64_t seed;
  int ninst, lmin, lmax;
  double sigma;
  double tol_rel;
  uint64_t max_iter;
  int autotune;   /* 1: autotune every batch matrix's launch shape at setup */
  int graphs;     /* 1: replay one recorded PCG iteration per step */
  int tail_levels;  /* levels 1..tail_levels of every instance in one launch per V-cycle
                       (feat_batch_tail_f64, bit-identical; lowered to what one workgroup
                       holds); <= 0: every level as block-diagonal launches */
  int no_tiled;     /* 1: the smoother calls above the tail as SELL-64 launches instead of
                       the tiled structured-grid kernels (feat_batch_pre/post_tiled_f64) */
  int tiled_prol_arrays;  /* 1: the tiled post-smoothing reads the prolongation's ELL arrays
                             even where it is bilinear injection (formed from the node position) */
  int tail_ell;           /* 1: the coarse tail in ELL form even where its levels are structured grids */
  int no_tiled_restriction;  /* 1: the restriction of a tiled level as a launch of its own instead of
                                inside its tiled pre-smoothing */
} feat_batch_cfg_t;
typedef struct feat_batchsolver_s* feat_batchsolver_t;
int feat_batchsolver_create(feat_batchsolver_t* s, const feat_batch_cfg_t* cfg);
/* levels (ninst ints) and PCG-layout offsets (ninst + 1) */
int feat_batchsolver_layout(feat_batchsolver_t s, int* levels, uint64_t* off);
/* x, b: device arrays in the PCG layout; per-instance outputs are host arrays
 * (ninst entries each, hist ninst * (max_iter + 1), any may be NULL) */
int feat_batchsolver_solve(feat_batchsolver_t s, double* x, const double* b, int* status, uint64_t* iters,
                           double* def_init, double* def_final, double* hist);
/* chosen launch shapes and their SpMV times (us) of the batch matrices:
 * per level k = 1..lmax: A_k (and P_k, R_k for k > 1), then the PCG-layout matrix */
int feat_batchsolver_tuning(feat_batchsolver_t s, int* tunes, double* us, int cap, int* count);
/* levels run by the fused coarse tail (0: none); *grid (may be NULL): 1 when it runs in its
 * grid form (feat_batch_tail_t.grid) */
int feat_batchsolver_tail_levels(feat_batchsolver_t s, int* levels, int* grid);
/* the levels whose smoother calls run tiled: up to cap of them into levels, their number in *count */
int feat_batchsolver_tiled_levels(feat_batchsolver_t s, int* levels, int cap, int* count);
int feat_batchsolver_destroy(feat_batchsolver_t s);
/* setup pieces of one instance, for tests */
int feat_batch_levels(uint64_t seed, int ninst, int lmin, int lmax, int* out);
int feat_lognormal_cells(uint64_t seed, uint64_t n, double sigma, double* out);
/* Q1 diffusion matrix on `level` for a per-cell coefficient of `level_c` (>= level);
 * the pattern (rowptr, col) is the Q1 pattern of feat_q1_assemble(2, level) */
int feat_q1_diffusion_assemble(int level, int level_c, const double* coef, double* val, uint64_t* col,
                               uint64_t* rowptr);
/* 3D Q1 diffusion matrix of the tensor diag(1, 1, kz) on `level` for a per-cell coefficient of
 * `level_c` (>= level; cell (i, j, k) of level_c at i + 2^level_c (j + 2^level_c k)); the pattern
 * is the Q1 pattern of feat_q1_assemble(3, level).  2x2x2 Gauss points per cell. */
int feat_q1_diffusion3d_assemble(int level, int level_c, const double* coef, double kz, double* val,
                                 uint64_t* col, uint64_t* rowptr);

/* ---- multigrid with a learned finest-level smoother (3D Q1 anisotropic diffusion) ----
 * No counterpart in the reference (see include/feat3amd/learned.hpp). */
typedef struct {
  int level_fine, level_coarse;
  double sigma, kz;          /* log-normal coefficient, anisotropy diag(1,1,kz) */
  uint64_t seed;
  int nu;                    /* pre/post smoothing steps */
  double omega_jac, omega_learned, tol_rel;
  uint64_t max_iter;
  int n_init, n_snap;        /* training data: random starts x outer iterations */
  uint64_t batch;            /* training batch */
  float lr;
  int fcg_m;                 /* flexible CG: directions kept for A-orthogonalisation (>= 1) */
  int learned_steps;         /* learned steps per finest smoother call, the rest damped Jacobi
                                (0 or >= nu: all nu steps learned) */
} feat_learned_cfg_t;
typedef struct {
  int status;
  uint64_t iters;
  int fell_back;
  uint64_t fallback_iter;
  double def_init, def_final, seconds;
} feat_learned_result_t;
typedef struct feat_learned_s* feat_learned_t;
int feat_learned_create(feat_learned_t* s, const feat_learned_cfg_t* cfg);
/* residual/error snapshots of every finest smoothing step of the FP64 multigrid on the
 * error equation: Jacobi smoother (on_policy = 0) or the current learned one (on-policy
 * refresh from new random starts); *rho_jacobi: the Jacobi multigrid's contraction */
int feat_learned_generate(feat_learned_t s, int on_policy, double* rho_jacobi);
/* `steps` Adam steps; loss (steps floats, may be NULL): relative energy ep^T A ep / e^T A e
 * one learned step leaves on the step's snapshot */
int feat_learned_train(feat_learned_t s, int steps, uint64_t seed, float* loss);
/* training with model selection: `rounds` x `steps` Adam steps (on-policy data refresh
 * before rounds 2.. when on_policy), a validation solve on a second right-hand side after
 * each round; keeps the best round's parameters if they beat the Jacobi multigrid there.
 * round_iters (rounds entries, may be NULL): validation iterations, 0 = no convergence */
int feat_learned_fit(feat_learned_t s, int rounds, int steps, int on_policy, uint64_t seed,
                     int* best_round, uint64_t* jacobi_iters, uint64_t* best_iters, uint64_t* round_iters);
/* flexible PCG(MG-V) on the built-in right-hand side from x (device, n doubles;
 * NULL: from zero); learned = 0: Jacobi smoother on every level */
int feat_learned_solve(feat_learned_t s, int learned, int allow_fallback, double* x,
                       feat_learned_result_t* res, double* hist, uint64_t hist_cap);
int feat_learned_size(feat_learned_t s, uint64_t* n, uint64_t* m);
/* mean relative energy ep^T A ep / e^T A e one learned step leaves on the snapshots */
int feat_learned_energy(feat_learned_t s, double* energy);
/* parameters (FEAT_MLP_NPARAM floats): get (may be NULL), then set (may be NULL) */
int feat_learned_params(feat_learned_t s, float* get, const float* set);
/* device pointers of the FP32 parameters and their BF16 copies */
int feat_learned_param_ptrs(feat_learned_t s, float** params, uint16_t** wbf);
int feat_learned_destroy(feat_learned_t s);

#ifdef __cplusplus
}
#endif
#endif
