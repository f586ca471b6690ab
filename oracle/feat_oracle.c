/*
 * feat_oracle.c -- CPU oracle for the LAFEM + Solver hot path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference's generic
 * backend and solver loops.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it, and only as the checker.
 *
 * Every `s += a*b` / `a*x + y` of the reference is written as fma(), matching
 * the reference's -O3 -march=<cpu> build on FMA hardware; the file is built
 * with -ffp-contract=off so nothing else is fused.  Each function names the
 * reference lines it restates.
 */
#include "feat_oracle.h"
#include "orc_internal.h"

#include <float.h>
#include <immintrin.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EPS_F64 DBL_EPSILON

/* ------------------------------------------------------------------------- */
/* RNG: kernel/util/random.hpp:84-110 (next), :259-280 (floating point gen)   */
/* ------------------------------------------------------------------------- */
uint64_t orc_rng_next(uint64_t* s)
{
  uint64_t x = *s;
  x ^= x >> 12;
  x ^= x << 25;
  x ^= x >> 27;
  *s = x;
  return x * UINT64_C(2685821657736338717);
}

static double rng_unit(uint64_t* s)
{
  const uint64_t mask = UINT64_C(0xFFFFFFFF);
  const double d = 1.0 / (double)mask;
  return (double)(orc_rng_next(s) & mask) * d;
}

void orc_rng_fill_f64(uint64_t* s, double* v, uint64_t n, double a, double b)
{
  /* Random::operator()(a,b) orders the bounds (random.hpp:150-154);
   * gen_ranged: a + x*(b-a) if a+eps < b, else a (random.hpp:270-279). */
  double lo = a < b ? a : b, hi = a < b ? b : a;
  for(uint64_t i = 0; i < n; ++i)
  {
    double x = rng_unit(s);
    v[i] = (lo + EPS_F64 < hi) ? fma(x, hi - lo, lo) : lo;
  }
}

/* ------------------------------------------------------------------------- */
/* CSR apply: lafem/arch/apply_generic.hpp:25-69                              */
/*   r = a*A*x + b*y   (transposed: r = a*A^T*x + b*y)                        */
/* ------------------------------------------------------------------------- */
#define DEFINE_CSR_APPLY(NAME, DT, FMA, EPSV)                                        \
void NAME(DT* r, DT a, const DT* x, DT b, const DT* y, const DT* val,                 \
          const uint64_t* col, const uint64_t* rowptr, uint64_t rows, uint64_t cols, \
          int transposed)                                                            \
{                                                                                    \
  const uint64_t nr = transposed ? cols : rows;                                      \
  if((b < 0 ? -b : b) < EPSV)                                                        \
    memset(r, 0, sizeof(DT) * nr);                                                   \
  else if(r != y)                                                                    \
    memcpy(r, y, sizeof(DT) * nr);                                                   \
  if(transposed)                                                                     \
  {                                                                                  \
    DT ba = b / a;                                                                   \
    for(uint64_t c = 0; c < cols; ++c) r[c] = ba * r[c];                             \
    for(uint64_t row = 0; row < rows; ++row)                                         \
      for(uint64_t i = rowptr[row]; i < rowptr[row + 1]; ++i)                        \
        r[col[i]] = FMA(val[i], x[row], r[col[i]]);                                  \
    for(uint64_t c = 0; c < cols; ++c) r[c] = a * r[c];                              \
  }                                                                                  \
  else                                                                               \
  {                                                                                  \
    for(uint64_t row = 0; row < rows; ++row)                                         \
    {                                                                                \
      DT sum = 0;                                                                    \
      const uint64_t end = rowptr[row + 1];                                          \
      for(uint64_t i = rowptr[row]; i < end; ++i)                                    \
        sum = FMA(val[i], x[col[i]], sum);                                           \
      r[row] = FMA(sum, a, b * r[row]);                                              \
    }                                                                                \
  }                                                                                  \
}

DEFINE_CSR_APPLY(orc_csr_apply_f64, double, fma, DBL_EPSILON)
DEFINE_CSR_APPLY(orc_csr_apply_f32, float, fmaf, FLT_EPSILON)

/* BCSR: lafem/arch/apply_generic.hpp:121-154.  Blocks are bh x bw row-major;
 * the block loop runs i (block), h (row), w (column) in that order. */
void orc_bcsr_apply_f64(int bh, int bw, double* r, double a, const double* x, double b,
                        const double* y, const double* val, const uint64_t* col,
                        const uint64_t* rowptr, uint64_t rows)
{
  const uint64_t nr = rows * (uint64_t)bh;
  if(fabs(b) < EPS_F64)
    memset(r, 0, sizeof(double) * nr);
  else if(r != y)
    memcpy(r, y, sizeof(double) * nr);
  double bsum[16];
  for(uint64_t row = 0; row < rows; ++row)
  {
    for(int h = 0; h < bh; ++h) bsum[h] = 0.0;
    for(uint64_t i = rowptr[row]; i < rowptr[row + 1]; ++i)
    {
      const double* blk = val + i * (uint64_t)(bh * bw);
      const double* bx = x + col[i] * (uint64_t)bw;
      for(int h = 0; h < bh; ++h)
        for(int w = 0; w < bw; ++w)
          bsum[h] = fma(blk[h * bw + w], bx[w], bsum[h]);
    }
    for(int h = 0; h < bh; ++h)
      r[row * bh + h] = fma(bsum[h], a, b * r[row * bh + h]);
  }
}

/* ------------------------------------------------------------------------- */
/* BLAS-1                                                                     */
/* ------------------------------------------------------------------------- */
/* arch/axpy_generic.hpp:25-50 -- r = a*x + y.  The r == y and distinct
 * branches contract to one fma per element; the r == x branch is two
 * statements and rounds twice (checked bitwise against oracle/_ref). */
void orc_axpy_f64(double* r, double a, const double* x, const double* y, uint64_t n)
{
  if(r == x && r != y)
  {
    /* `r[i] *= a; r[i] += y[i];` -- two statements, two roundings */
    for(uint64_t i = 0; i < n; ++i)
      r[i] = r[i] * a + y[i];
    return;
  }
  for(uint64_t i = 0; i < n; ++i)
    r[i] = fma(a, x[i], y[i]);
}

/* arch/dot_product_generic.hpp:21-40 */
double orc_dot_f64(const double* x, const double* y, uint64_t n)
{
  double r = 0.0;
  if(x == y)
  {
    for(uint64_t i = 0; i < n; ++i)
      r = fma(x[i], x[i], r);
    return r;
  }
  /* the reference build leaves the x != y loop uncontracted (checked
   * bitwise against oracle/_ref) */
  for(uint64_t i = 0; i < n; ++i)
    r = r + x[i] * y[i];
  return r;
}

/* arch/dot_product_generic.hpp:44-70 (distinct-vector branch) */
double orc_triple_dot_f64(const double* x, const double* y, const double* z, uint64_t n)
{
  double r = 0.0;
  for(uint64_t i = 0; i < n; ++i)
    r = fma(x[i] * y[i], z[i], r);
  return r;
}

/* arch/norm_generic.hpp:24-33 */
double orc_norm2_f64(const double* x, uint64_t n)
{
  double r = 0.0;
  for(uint64_t i = 0; i < n; ++i)
    r = fma(x[i], x[i], r);
  return sqrt(r);
}

/* arch/scale_generic.hpp:22-38 */
void orc_scale_f64(double* r, const double* x, double s, uint64_t n)
{
  for(uint64_t i = 0; i < n; ++i)
    r[i] = x[i] * s;
}

/* arch/component_product_generic.hpp:22-52 */
void orc_component_product_f64(double* r, const double* x, const double* y, uint64_t n)
{
  for(uint64_t i = 0; i < n; ++i)
    r[i] = x[i] * y[i];
}

/* arch/component_invert_generic.hpp:21-36 -- r = s / x */
void orc_component_invert_f64(double* r, const double* x, double s, uint64_t n)
{
  for(uint64_t i = 0; i < n; ++i)
    r[i] = s / x[i];
}

/* arch/diagonal_generic.hpp:21-38 -- index of the diagonal entry, nnz if absent */
void orc_diag_index(uint64_t* diag, const uint64_t* col, const uint64_t* rowptr, uint64_t rows)
{
  for(uint64_t row = 0; row < rows; ++row)
  {
    diag[row] = rowptr[rows];
    for(uint64_t i = rowptr[row]; i < rowptr[row + 1]; ++i)
    {
      if(col[i] == row)
      {
        diag[row] = i;
        break;
      }
    }
  }
}

/* lafem/sparse_matrix_csr.hpp:2277-2290 -- diagonal values, 0 where absent */
void orc_extract_diag_f64(double* diag, const double* val, const uint64_t* col,
                          const uint64_t* rowptr, uint64_t rows)
{
  uint64_t* idx = (uint64_t*)malloc(sizeof(uint64_t) * (rows ? rows : 1));
  orc_diag_index(idx, col, rowptr, rows);
  for(uint64_t row = 0; row < rows; ++row)
    diag[row] = (idx[row] != rowptr[rows]) ? val[idx[row]] : 0.0;
  free(idx);
}

/* arch/unit_filter_generic.hpp:22-38 */
void orc_unit_filter_rhs_f64(double* v, const double* sv_val, const uint64_t* sv_idx, uint64_t ue)
{
  for(uint64_t i = 0; i < ue; ++i)
    v[sv_idx[i]] = sv_val[i];
}

void orc_unit_filter_def_f64(double* v, const uint64_t* sv_idx, uint64_t ue)
{
  for(uint64_t i = 0; i < ue; ++i)
    v[sv_idx[i]] = 0.0;
}

/* arch/unit_filter_blocked_generic.hpp:24-60: blocks of bs components; with ign_nans a
 * NaN filter value leaves its component unchanged (rhs and def alike) */
void orc_unit_filter_blocked_f64(int rhs, int bs, double* v, const double* sv_el, const uint64_t* sv_idx,
                                 uint64_t ue, int ign_nans)
{
  for(uint64_t i = 0; i < ue; ++i)
    for(int j = 0; j < bs; ++j)
    {
      const double e = sv_el ? sv_el[(uint64_t)bs * i + (uint64_t)j] : 0.0;
      if(ign_nans && isnan(e)) continue;
      v[(uint64_t)bs * sv_idx[i] + (uint64_t)j] = rhs ? e : 0.0;
    }
}

/* lafem/mean_filter.hpp:214-264.  filter_def/rhs: v += (-(v.prim)/vol) * dual;
 * filter_cor: v += (-(v.dual)/vol) * prim.  The caller passes the vector that
 * is dotted first and the one that is added second. */
void orc_mean_filter_f64(double* v, const double* w_dot, const double* w_add, double volume,
                         uint64_t n)
{
  const double alpha = -orc_dot_f64(v, w_dot, n) / volume;
  orc_axpy_f64(v, alpha, w_add, v, n);
}

/* ------------------------------------------------------------------------- */
/* util/math.hpp:1169-1274 -- Gauss-Jordan with diagonal pivoting             */
/* ------------------------------------------------------------------------- */
double orc_invert_matrix_f64(int n, int stride, double* a, int* p)
{
  if(n <= 0 || stride < n || !a || !p)
    return 0.0;
  if(n == 1)
  {
    double det = a[0];
    a[0] = 1.0 / det;
    return det;
  }
  for(int i = 0; i < n; ++i) p[i] = i;
  double det = 1.0;
  for(int k = 0; k < n; ++k)
  {
    double piv = fabs(a[p[k] * stride + p[k]]);
    int i = k;
    for(int j = k + 1; j < n; ++j)
    {
      double av = fabs(a[p[j] * stride + p[j]]);
      if(av > piv) { piv = av; i = j; }
    }
    if(i > k) { int t = p[k]; p[k] = p[i]; p[i] = t; }
    const int pk = p[k] * stride;
    det *= a[pk + p[k]];
    const double inv = 1.0 / a[pk + p[k]];
    a[pk + p[k]] = 1.0;
    for(int j = 0; j < n; ++j) a[pk + j] *= inv;
    for(int r = 0; r < n; ++r)
    {
      if(r == p[k]) continue;
      const int ro = r * stride;
      const double f = a[ro + p[k]];
      a[ro + p[k]] = 0.0;
      for(int j = 0; j < n; ++j)
        a[ro + j] = fma(-a[pk + j], f, a[ro + j]);
    }
  }
  return det;
}

/* ------------------------------------------------------------------------- */
/* Iterative solver bookkeeping: solver/iterative.hpp                         */
/* ------------------------------------------------------------------------- */
void orc_iter_cfg_default(orc_iter_cfg_t* c)
{
  /* iterative.hpp:213-236 */
  c->tol_rel = sqrt(EPS_F64);
  c->tol_abs = 1.0 / (EPS_F64 * EPS_F64);
  c->tol_abs_low = 0.0;
  c->div_rel = 1.0 / EPS_F64;
  c->div_abs = 1.0 / (EPS_F64 * EPS_F64);
  c->stag_rate = 0.95;
  c->min_iter = 0;
  c->max_iter = 100;
  c->min_stag_iter = 0;
}


/* iterative.hpp:748-770 */
int set_initial_defect(itstate_t* s, double def)
{
  s->def_init = s->def_cur = s->def_prev = def;
  s->num_iter = 0;
  s->num_stag_iter = 0;
  if(s->hist) s->hist[0] = def;
  if(!isfinite(def)) return ORC_ABORTED;
  if(def <= EPS_F64 * EPS_F64) return ORC_SUCCESS;
  return ORC_PROGRESS;
}

/* iterative.hpp:492-507, 795-836 */
int analyse_defect(itstate_t* s, uint64_t it, double cur, double prev, int check_stag)
{
  const orc_iter_cfg_t* c = s->c;
  if(!isfinite(cur)) return ORC_ABORTED;
  if(cur > c->div_abs || cur > c->div_rel * s->def_init) return ORC_DIVERGED;
  if(it < c->min_iter) return ORC_PROGRESS;
  if(cur <= c->tol_abs && (cur <= c->tol_rel * s->def_init || cur <= c->tol_abs_low))
    return ORC_SUCCESS;
  if(it >= c->max_iter) return ORC_MAX_ITER;
  if(check_stag && c->min_stag_iter > 0)
  {
    if(cur >= c->stag_rate * prev)
    {
      if(++s->num_stag_iter >= c->min_stag_iter) return ORC_STAGNATED;
    }
    else
      s->num_stag_iter = 0;
  }
  return ORC_PROGRESS;
}

/* iterative.hpp:853-885: the norm is skipped when min_iter == max_iter
 * (no plot, no stagnation check); the stale defect is analysed then. */
int set_new_defect(itstate_t* s, const double* def, uint64_t n)
{
  ++s->num_iter;
  s->def_prev = s->def_cur;
  int calc = (s->c->min_iter < s->c->max_iter) || (s->c->min_stag_iter > 0);
  if(calc)
    s->def_cur = orc_norm2_f64(def, n);
  if(s->hist && calc) s->hist[s->num_iter] = s->def_cur;
  return analyse_defect(s, s->num_iter, s->def_cur, s->def_prev, 1);
}

int status_success(int st)
{
  return st == ORC_SUCCESS || st == ORC_MAX_ITER || st == ORC_STAGNATED;
}

/* ------------------------------------------------------------------------- */
/* Smoother / preconditioner objects                                          */
/* ------------------------------------------------------------------------- */
typedef struct {
  const orc_csr_t* A;
  const uint64_t* filt;
  uint64_t nfilt;
  double* inv_diag;  /* NULL: no preconditioner */
} jac_t;

/* jacobi_precond.hpp:136-143: inv_diag = omega / diag(A) */
static double* jacobi_init(const orc_csr_t* A, double omega)
{
  double* d = (double*)malloc(sizeof(double) * (A->rows ? A->rows : 1));
  orc_extract_diag_f64(d, A->val, A->col, A->rowptr, A->rows);
  orc_component_invert_f64(d, d, omega, A->rows);
  return d;
}

/* jacobi_precond.hpp:159-164 or iterative.hpp:1067-1080 (no precond) */
static void precond_apply(const jac_t* j, double* cor, const double* def)
{
  const uint64_t n = j->A->rows;
  if(j->inv_diag)
    orc_component_product_f64(cor, j->inv_diag, def, n);
  else
    memcpy(cor, def, sizeof(double) * n);
  orc_unit_filter_def_f64(cor, j->filt, j->nfilt);
}

static void csr_apply(const orc_csr_t* A, double* r, double a, const double* x, double b,
                      const double* y)
{
  orc_csr_apply_f64(r, a, x, b, y, A->val, A->col, A->rowptr, A->rows, A->cols, 0);
}

/* richardson.hpp:160-226: apply(cor, def) with scratch vectors rdef, rcor. */
static int richardson_apply(const jac_t* j, double omega, const orc_iter_cfg_t* it, double* sol,
                            const double* rhs, double* rdef, double* rcor, double* hist,
                            uint64_t* iters)
{
  const uint64_t n = j->A->rows;
  itstate_t s = {it, 0, 0, 0, 0, 0, hist};
  memcpy(rdef, rhs, sizeof(double) * n);
  memset(sol, 0, sizeof(double) * n);
  int st = set_initial_defect(&s, orc_norm2_f64(rdef, n));
  while(st == ORC_PROGRESS)
  {
    precond_apply(j, rcor, rdef);
    orc_axpy_f64(sol, omega, rcor, sol, n);
    csr_apply(j->A, rdef, -1.0, sol, 1.0, rhs);
    orc_unit_filter_def_f64(rdef, j->filt, j->nfilt);
    st = set_new_defect(&s, rdef, n);
  }
  if(iters) *iters = s.num_iter;
  return st;
}

/* ------------------------------------------------------------------------- */
/* Multigrid V-cycle: multigrid.hpp:1270-1350, 1526-1589, 1713-2000           */
/* ------------------------------------------------------------------------- */
typedef struct {
  jac_t jac;
  double *rhs, *sol, *def, *cor, *rdef, *rcor;
  double* coarse_inv;   /* dense inverse for ORC_COARSE_DIRECT */
} mglvl_t;

typedef struct {
  const orc_mg_level_t* L;
  int nl;
  const orc_mg_cfg_t* cfg;
  mglvl_t* lv;
  orc_iter_cfg_t smooth_cfg;
} mg_t;

static double* dense_filtered_inverse(const orc_csr_t* A, const uint64_t* filt, uint64_t nfilt)
{
  const uint64_t n = A->rows;
  double* a = (double*)calloc(n * n, sizeof(double));
  int* p = (int*)malloc(sizeof(int) * n);
  char* f = (char*)calloc(n, 1);
  for(uint64_t k = 0; k < nfilt; ++k) f[filt[k]] = 1;
  for(uint64_t r = 0; r < n; ++r)
  {
    if(f[r]) { a[r * n + r] = 1.0; continue; }
    for(uint64_t i = A->rowptr[r]; i < A->rowptr[r + 1]; ++i)
      if(!f[A->col[i]]) a[r * n + A->col[i]] = A->val[i];
  }
  orc_invert_matrix_f64((int)n, (int)n, a, p);
  free(p);
  free(f);
  return a;
}

static void mg_init(mg_t* m, const orc_mg_level_t* L, int nl, const orc_mg_cfg_t* cfg)
{
  m->L = L;
  m->nl = nl;
  m->cfg = cfg;
  m->lv = (mglvl_t*)calloc((size_t)nl, sizeof(mglvl_t));
  orc_iter_cfg_default(&m->smooth_cfg);
  for(int i = 0; i < nl; ++i)
  {
    mglvl_t* v = &m->lv[i];
    const uint64_t n = L[i].A.rows;
    v->jac.A = &L[i].A;
    v->jac.filt = L[i].filt;
    v->jac.nfilt = L[i].nfilt;
    v->jac.inv_diag = jacobi_init(&L[i].A, cfg->omega_jac);
    double** vecs[6] = {&v->rhs, &v->sol, &v->def, &v->cor, &v->rdef, &v->rcor};
    for(int k = 0; k < 6; ++k) *vecs[k] = (double*)calloc(n ? n : 1, sizeof(double));
    if(i == nl - 1 && cfg->coarse == ORC_COARSE_DIRECT)
      v->coarse_inv = dense_filtered_inverse(&L[i].A, L[i].filt, L[i].nfilt);
  }
}

static void mg_free(mg_t* m)
{
  for(int i = 0; i < m->nl; ++i)
  {
    mglvl_t* v = &m->lv[i];
    free(v->jac.inv_diag);
    free(v->rhs); free(v->sol); free(v->def); free(v->cor); free(v->rdef); free(v->rcor);
    free(v->coarse_inv);
  }
  free(m->lv);
}

static void mg_smooth(mg_t* m, int i, int steps, double* x, const double* b)
{
  orc_iter_cfg_t c = m->smooth_cfg;
  c.min_iter = c.max_iter = (uint64_t)steps;
  mglvl_t* v = &m->lv[i];
  richardson_apply(&v->jac, m->cfg->omega_rich, &c, x, b, v->rdef, v->rcor, NULL, NULL);
}

static void mg_coarse(mg_t* m)
{
  const int c = m->nl - 1;
  mglvl_t* v = &m->lv[c];
  const orc_mg_level_t* L = &m->L[c];
  const uint64_t n = L->A.rows;
  switch(m->cfg->coarse)
  {
  case ORC_COARSE_SMOOTHER:
    mg_smooth(m, c, m->cfg->nu_pre, v->sol, v->rhs);
    break;
  case ORC_COARSE_JACOBI:
  {
    double* d = jacobi_init(&L->A, m->cfg->coarse_omega);
    orc_component_product_f64(v->sol, d, v->rhs, n);
    orc_unit_filter_def_f64(v->sol, L->filt, L->nfilt);
    free(d);
    break;
  }
  case ORC_COARSE_DIRECT:
    for(uint64_t r = 0; r < n; ++r)
    {
      double s = 0.0;
      for(uint64_t k = 0; k < n; ++k) s = fma(v->coarse_inv[r * n + k], v->rhs[k], s);
      v->sol[r] = s;
    }
    orc_unit_filter_def_f64(v->sol, L->filt, L->nfilt);
    break;
  default:
    memcpy(v->sol, v->rhs, sizeof(double) * n);
    orc_unit_filter_def_f64(v->sol, L->filt, L->nfilt);
    break;
  }
}

/* MultiGrid::apply (multigrid.hpp:1270-1330) with _apply_cycle_v (:1335-1350) */
static void mg_apply(mg_t* m, double* cor_out, const double* def_in)
{
  const int nl = m->nl;
  memcpy(m->lv[0].rhs, def_in, sizeof(double) * m->L[0].A.rows);
  /* _apply_rest (:1713-1840) */
  for(int i = 0; i < nl - 1; ++i)
  {
    mglvl_t* v = &m->lv[i];
    const orc_mg_level_t* L = &m->L[i];
    mg_smooth(m, i, m->cfg->nu_pre, v->sol, v->rhs);
    csr_apply(&L->A, v->def, -1.0, v->sol, 1.0, v->rhs);
    orc_unit_filter_def_f64(v->def, L->filt, L->nfilt);
    csr_apply(&L->R, m->lv[i + 1].rhs, 1.0, v->def, 0.0, m->lv[i + 1].rhs);
    orc_unit_filter_def_f64(m->lv[i + 1].rhs, m->L[i + 1].filt, m->L[i + 1].nfilt);
  }
  /* _apply_coarse (:1526-1589) */
  mg_coarse(m);
  /* _apply_prol (:1845-2000), fixed coarse grid correction omega = 1 */
  for(int i = nl - 2; i >= 0; --i)
  {
    mglvl_t* v = &m->lv[i];
    const orc_mg_level_t* L = &m->L[i];
    const uint64_t n = L->A.rows;
    csr_apply(&L->P, v->cor, 1.0, m->lv[i + 1].sol, 0.0, v->cor);
    orc_unit_filter_def_f64(v->cor, L->filt, L->nfilt);
    orc_axpy_f64(v->sol, 1.0, v->cor, v->sol, n);
    csr_apply(&L->A, v->def, -1.0, v->sol, 1.0, v->rhs);
    orc_unit_filter_def_f64(v->def, L->filt, L->nfilt);
    mg_smooth(m, i, m->cfg->nu_post, v->cor, v->def);
    orc_axpy_f64(v->sol, 1.0, v->cor, v->sol, n);
  }
  memcpy(cor_out, m->lv[0].sol, sizeof(double) * m->L[0].A.rows);
}

/* ------------------------------------------------------------------------- */
/* PCG: solver/pcg.hpp:159-288                                                */
/* ------------------------------------------------------------------------- */
typedef void (*precond_fn)(void* ctx, double* cor, const double* def);

static int pcg_intern(const orc_csr_t* A, const uint64_t* filt, uint64_t nfilt,
                      precond_fn M, void* mctx, const orc_iter_cfg_t* it, double* x, double* r,
                      double* hist, orc_result_t* res)
{
  const uint64_t n = A->rows;
  double* p = (double*)calloc(n ? n : 1, sizeof(double));
  double* t = (double*)calloc(n ? n : 1, sizeof(double)); /* q and z share storage */
  itstate_t s = {it, 0, 0, 0, 0, 0, hist};
  int st = set_initial_defect(&s, orc_norm2_f64(r, n));
  if(st == ORC_PROGRESS)
  {
    M(mctx, p, r);
    double gamma = orc_dot_f64(r, p, n);
    for(;;)
    {
      csr_apply(A, t, 1.0, p, 0.0, t);
      orc_unit_filter_def_f64(t, filt, nfilt);
      const double alpha = gamma / orc_dot_f64(t, p, n);
      orc_axpy_f64(x, alpha, p, x, n);
      orc_axpy_f64(r, -alpha, t, r, n);
      st = set_new_defect(&s, r, n);
      if(st != ORC_PROGRESS) break;
      M(mctx, t, r);
      const double gamma2 = gamma;
      gamma = orc_dot_f64(r, t, n);
      const double beta = gamma / gamma2;
      orc_axpy_f64(p, beta, p, t, n);
    }
  }
  if(res)
  {
    res->status = st;
    res->num_iter = s.num_iter;
    res->def_init = s.def_init;
    res->def_final = s.def_cur;
  }
  free(p);
  free(t);
  return st;
}

static void mg_precond(void* ctx, double* cor, const double* def)
{
  mg_apply((mg_t*)ctx, cor, def);
}

static void jac_precond(void* ctx, double* cor, const double* def)
{
  precond_apply((const jac_t*)ctx, cor, def);
}

/* PCG::correct (pcg.hpp:159-173): r = b - A*x, filter_def(r), iterate. */
int orc_mg_pcg_solve(const orc_mg_level_t* levels, int nlevels, const orc_mg_cfg_t* cfg,
                     const orc_iter_cfg_t* it, double* x, const double* b, double* hist,
                     orc_result_t* res)
{
  const orc_csr_t* A = &levels[0].A;
  const uint64_t n = A->rows;
  mg_t m;
  mg_init(&m, levels, nlevels, cfg);
  double* r = (double*)malloc(sizeof(double) * (n ? n : 1));
  csr_apply(A, r, -1.0, x, 1.0, b);
  orc_unit_filter_def_f64(r, levels[0].filt, levels[0].nfilt);
  int st = pcg_intern(A, levels[0].filt, levels[0].nfilt, mg_precond, &m, it, x, r, hist, res);
  free(r);
  mg_free(&m);
  return st;
}

static int single_level(const orc_csr_t* A, const uint64_t* filt, uint64_t nfilt, int precond,
                        double omega, jac_t* j)
{
  j->A = A;
  j->filt = filt;
  j->nfilt = nfilt;
  j->inv_diag = precond ? jacobi_init(A, omega) : NULL;
  return 0;
}

int orc_pcg_solve(const orc_csr_t* A, const uint64_t* filt, uint64_t nfilt, int precond,
                  double omega, const orc_iter_cfg_t* it, double* x, const double* b,
                  double* hist, orc_result_t* res)
{
  jac_t j;
  single_level(A, filt, nfilt, precond, omega, &j);
  const uint64_t n = A->rows;
  double* r = (double*)malloc(sizeof(double) * (n ? n : 1));
  csr_apply(A, r, -1.0, x, 1.0, b);
  orc_unit_filter_def_f64(r, filt, nfilt);
  int st = pcg_intern(A, filt, nfilt, jac_precond, &j, it, x, r, hist, res);
  free(r);
  free(j.inv_diag);
  return st;
}

/* PCG::apply (pcg.hpp:140-156): r = copy(def), x = 0, iterate. */
int orc_pcg_apply(const orc_csr_t* A, const uint64_t* filt, uint64_t nfilt, int precond,
                  double omega, const orc_iter_cfg_t* it, double* x, const double* b,
                  double* hist, orc_result_t* res)
{
  jac_t j;
  single_level(A, filt, nfilt, precond, omega, &j);
  const uint64_t n = A->rows;
  double* r = (double*)malloc(sizeof(double) * (n ? n : 1));
  memcpy(r, b, sizeof(double) * n);
  memset(x, 0, sizeof(double) * n);
  int st = pcg_intern(A, filt, nfilt, jac_precond, &j, it, x, r, hist, res);
  free(r);
  free(j.inv_diag);
  return st;
}

static int richardson_generic(const orc_csr_t* A, const uint64_t* filt, uint64_t nfilt,
                              int precond, double omega_jac, double omega_rich,
                              const orc_iter_cfg_t* it, double* x, const double* b, double* hist,
                              orc_result_t* res, int correct)
{
  jac_t j;
  single_level(A, filt, nfilt, precond, omega_jac, &j);
  const uint64_t n = A->rows;
  double* rdef = (double*)malloc(sizeof(double) * (n ? n : 1));
  double* rcor = (double*)malloc(sizeof(double) * (n ? n : 1));
  itstate_t s = {it, 0, 0, 0, 0, 0, hist};
  int st;
  if(correct)
  {
    /* richardson.hpp:174-184 (correct): defect of the given x */
    csr_apply(A, rdef, -1.0, x, 1.0, b);
    orc_unit_filter_def_f64(rdef, filt, nfilt);
  }
  else
  {
    memcpy(rdef, b, sizeof(double) * n);
    memset(x, 0, sizeof(double) * n);
  }
  st = set_initial_defect(&s, orc_norm2_f64(rdef, n));
  while(st == ORC_PROGRESS)
  {
    precond_apply(&j, rcor, rdef);
    orc_axpy_f64(x, omega_rich, rcor, x, n);
    csr_apply(A, rdef, -1.0, x, 1.0, b);
    orc_unit_filter_def_f64(rdef, filt, nfilt);
    st = set_new_defect(&s, rdef, n);
  }
  if(res)
  {
    res->status = st;
    res->num_iter = s.num_iter;
    res->def_init = s.def_init;
    res->def_final = s.def_cur;
  }
  free(rdef);
  free(rcor);
  free(j.inv_diag);
  return st;
}

int orc_richardson_solve(const orc_csr_t* A, const uint64_t* filt, uint64_t nfilt, int precond,
                         double omega_jac, double omega_rich, const orc_iter_cfg_t* it,
                         double* x, const double* b, double* hist, orc_result_t* res)
{
  return richardson_generic(A, filt, nfilt, precond, omega_jac, omega_rich, it, x, b, hist, res, 1);
}

int orc_richardson_apply(const orc_csr_t* A, const uint64_t* filt, uint64_t nfilt, int precond,
                         double omega_jac, double omega_rich, const orc_iter_cfg_t* it,
                         double* x, const double* b, double* hist, orc_result_t* res)
{
  return richardson_generic(A, filt, nfilt, precond, omega_jac, omega_rich, it, x, b, hist, res, 0);
}

/* ------------------------------------------------------------------------- */
/* Mixed precision: FP64 PCG preconditioned by an FP32 V-cycle                */
/* (the ConvertPrecond pattern, solver/convert_precond.hpp:100-119).  The     */
/* FP32 cycle keeps the op order of the FP64 one above; matrices and the      */
/* Jacobi diagonals (omega/a_ii in FP64) are rounded to FP32 once; the coarse */
/* solve applies the FP64 dense inverse to the FP32 right-hand side.          */
/* ------------------------------------------------------------------------- */
typedef struct {
  uint64_t rows, cols;
  float* val;
  const uint64_t* col;
  const uint64_t* rowptr;
  float scale;
} csr32_t;

typedef struct {
  csr32_t A, P, R;
  const uint64_t* filt;
  uint64_t nfilt;
  char* mask;
  float *inv, *rhs, *sol, *def, *cor, *xa, *xb;
  double* cinv;   /* coarse: FP64 dense inverse */
} lvl32_t;

/* FP16 values of a matrix: half(float(a)*2^-e) with 2^(e-1) <= max|float(a)| < 2^e
 * (no scaling when the maximum is 0 or not finite), returned as the FP32 value of
 * each FP16 number; *scale = 2^e multiplies every row sum.  F16C conversions,
 * round to nearest even. */
void orc_f16_values(const double* val, uint64_t nnz, float* out, float* scale)
{
  float inv = 1.0f;
  *scale = 1.0f;
  float mx = 0.0f;
  for(uint64_t i = 0; i < nnz; ++i)
  {
    const float a = fabsf((float)val[i]);
    if(a > mx) mx = a;
  }
  if(mx > 0.0f && isfinite(mx))
  {
    int e = 0;
    (void)frexpf(mx, &e);
    *scale = ldexpf(1.0f, e);
    inv = ldexpf(1.0f, -e);
  }
  for(uint64_t i = 0; i < nnz; ++i)
    out[i] = _cvtsh_ss(_cvtss_sh((float)val[i] * inv, 0));
}

/* half = 0: values rounded to FP32 (LAFEM convert, one static_cast per value).
 * half = 1: FP16 values (orc_f16_values). */
static csr32_t csr_to32(const orc_csr_t* m, int half)
{
  csr32_t r;
  r.rows = m->rows;
  r.cols = m->cols;
  r.col = m->col;
  r.rowptr = m->rowptr;
  r.scale = 1.0f;
  r.val = (float*)malloc(sizeof(float) * (m->nnz ? m->nnz : 1));
  if(half)
    orc_f16_values(m->val, m->nnz, r.val, &r.scale);
  else
    for(uint64_t i = 0; i < m->nnz; ++i) r.val[i] = (float)m->val[i];
  return r;
}

static float row_sum32(const csr32_t* A, uint64_t row, const float* x)
{
  float s = 0.0f;
  for(uint64_t i = A->rowptr[row]; i < A->rowptr[row + 1]; ++i)
    s = fmaf(A->val[i], x[A->col[i]], s);
  return s * A->scale;
}

static double nrm2sq32(const float* v, uint64_t n)
{
  double s = 0.0;
  for(uint64_t i = 0; i < n; ++i) s = fma((double)v[i], (double)v[i], s);
  return s;
}

static int skips32(double nrm2)
{
  const double d = sqrt(nrm2);
  return !isfinite(d) || d <= (double)FLT_EPSILON * (double)FLT_EPSILON;
}

/* Richardson(Jacobi) in FP32 from x = 0 (the fused smoother's semantics):
 * mode 0: x = result; mode 1: also d_out = last filtered defect; mode 2: acc += result */
static void smooth32(lvl32_t* L, int steps, float omega, const float* rhs, double rhs_nrm2, int use_skip,
                     float* x, float* d_out, float* acc)
{
  const uint64_t n = L->A.rows;
  const int skip = use_skip ? skips32(rhs_nrm2) : 0;
  float* cur = L->xa;
  float* nxt = L->xb;
  for(uint64_t i = 0; i < n; ++i)
  {
    float c = L->inv[i] * rhs[i];
    if(L->mask[i]) c = 0.0f;
    cur[i] = skip ? 0.0f : fmaf(omega, c, 0.0f);
  }
  for(int k = 1; k < steps; ++k)
  {
    for(uint64_t i = 0; i < n; ++i)
    {
      float d = fmaf(row_sum32(&L->A, i, cur), -1.0f, rhs[i]);
      if(L->mask[i]) d = 0.0f;
      float c = L->inv[i] * d;
      if(L->mask[i]) c = 0.0f;
      nxt[i] = skip ? 0.0f : fmaf(omega, c, cur[i]);
    }
    float* t = cur; cur = nxt; nxt = t;
  }
  if(d_out)
    for(uint64_t i = 0; i < n; ++i)
    {
      float d = fmaf(row_sum32(&L->A, i, cur), -1.0f, rhs[i]);
      d_out[i] = L->mask[i] ? 0.0f : d;
    }
  for(uint64_t i = 0; i < n; ++i)
  {
    if(acc) acc[i] = fmaf(1.0f, cur[i], acc[i]);
    else x[i] = cur[i];
  }
}

typedef struct {
  lvl32_t* L;
  int nl;
  const orc_mg_cfg_t* cfg;
} mg32_t;

static void mg32_init(mg32_t* m, const orc_mg_level_t* lv, int nl, const orc_mg_cfg_t* cfg, int flags)
{
  const int half = flags & ORC_MIXED_F16;
  m->nl = nl;
  m->cfg = cfg;
  m->L = (lvl32_t*)calloc((size_t)nl, sizeof(lvl32_t));
  for(int i = 0; i < nl; ++i)
  {
    lvl32_t* L = &m->L[i];
    const uint64_t n = lv[i].A.rows;
    L->A = csr_to32(&lv[i].A, half);
    if(i + 1 < nl)
    {
      L->P = csr_to32(&lv[i].P, half);
      L->R = csr_to32(&lv[i].R, half);
    }
    L->filt = lv[i].filt;
    L->nfilt = lv[i].nfilt;
    L->mask = (char*)calloc(n ? n : 1, 1);
    for(uint64_t k = 0; k < lv[i].nfilt; ++k) L->mask[lv[i].filt[k]] = 1;
    const double om = (i + 1 == nl && cfg->coarse == ORC_COARSE_JACOBI) ? cfg->coarse_omega : cfg->omega_jac;
    /* JacobiPrecond<float>::init_numeric (jacobi_precond.hpp:136-143): the diagonal
     * of the FP32 matrix (each a_ii rounded once), then float(omega) / d_ii in FP32
     * (arch/component_invert_generic.hpp:21-37).  The diagonal is a vector and keeps
     * FP32 precision in the FP16-valued form as well.  ORC_MIXED_SLIP_INV64 is a
     * planted slip for the oracle's own pin test: omega / a_ii in FP64, rounded. */
    double* d64 = (double*)malloc(sizeof(double) * (n ? n : 1));
    orc_extract_diag_f64(d64, lv[i].A.val, lv[i].A.col, lv[i].A.rowptr, n);
    L->inv = (float*)malloc(sizeof(float) * (n ? n : 1));
    for(uint64_t k = 0; k < n; ++k)
      L->inv[k] = (flags & ORC_MIXED_SLIP_INV64) ? (float)(om / d64[k]) : (float)om / (float)d64[k];
    free(d64);
    float** vecs[6] = {&L->rhs, &L->sol, &L->def, &L->cor, &L->xa, &L->xb};
    for(int k = 0; k < 6; ++k) *vecs[k] = (float*)calloc(n ? n : 1, sizeof(float));
    if(i + 1 == nl && cfg->coarse == ORC_COARSE_DIRECT)
      L->cinv = dense_filtered_inverse(&lv[i].A, lv[i].filt, lv[i].nfilt);
  }
}

static void mg32_free(mg32_t* m)
{
  for(int i = 0; i < m->nl; ++i)
  {
    lvl32_t* L = &m->L[i];
    free(L->A.val);
    if(i + 1 < m->nl) { free(L->P.val); free(L->R.val); }
    free(L->mask); free(L->inv); free(L->rhs); free(L->sol); free(L->def); free(L->cor);
    free(L->xa); free(L->xb); free(L->cinv);
  }
  free(m->L);
}

/* cor = MG32(def): the FP64 defect is rounded to FP32, the FP32 cycle runs,
 * the correction is widened back to FP64.  def_nrm2 is |def|^2 (FP64). */
static void mg32_apply(mg32_t* m, double* cor, const double* def, double def_nrm2)
{
  const int nl = m->nl, c = nl - 1;
  const orc_mg_cfg_t* cfg = m->cfg;
  const float om = (float)cfg->omega_rich;
  lvl32_t* L0 = &m->L[0];
  for(uint64_t i = 0; i < L0->A.rows; ++i) L0->rhs[i] = (float)def[i];
  double nrm = def_nrm2;
  for(int i = 0; i < c; ++i)
  {
    lvl32_t* L = &m->L[i];
    lvl32_t* Lc = &m->L[i + 1];
    smooth32(L, cfg->nu_pre, om, L->rhs, nrm, 1, L->sol, L->def, NULL);
    for(uint64_t r = 0; r < L->R.rows; ++r)
    {
      float v = fmaf(row_sum32(&L->R, r, L->def), 1.0f, 0.0f);
      Lc->rhs[r] = Lc->mask[r] ? 0.0f : v;
    }
    nrm = nrm2sq32(Lc->rhs, Lc->A.rows);
  }
  lvl32_t* Lc = &m->L[c];
  const uint64_t nc = Lc->A.rows;
  if(cfg->coarse == ORC_COARSE_DIRECT)
  {
    for(uint64_t r = 0; r < nc; ++r)
    {
      double s = 0.0;
      for(uint64_t k = 0; k < nc; ++k) s = fma(Lc->cinv[r * nc + k], (double)Lc->rhs[k], s);
      Lc->sol[r] = Lc->mask[r] ? 0.0f : (float)s;
    }
  }
  else if(cfg->coarse == ORC_COARSE_JACOBI)
    smooth32(Lc, 1, 1.0f, Lc->rhs, 0.0, 0, Lc->sol, NULL, NULL);
  else
    smooth32(Lc, cfg->nu_pre, om, Lc->rhs, nrm, 1, Lc->sol, NULL, NULL);
  for(int k = c - 1; k >= 0; --k)
  {
    lvl32_t* L = &m->L[k];
    const uint64_t n = L->A.rows;
    for(uint64_t r = 0; r < n; ++r)
    {
      float v = fmaf(row_sum32(&L->P, r, m->L[k + 1].sol), 1.0f, 0.0f);
      if(L->mask[r]) v = 0.0f;
      L->sol[r] = fmaf(1.0f, v, L->sol[r]);
    }
    for(uint64_t r = 0; r < n; ++r)
    {
      float d = fmaf(row_sum32(&L->A, r, L->sol), -1.0f, L->rhs[r]);
      L->def[r] = L->mask[r] ? 0.0f : d;
    }
    smooth32(L, cfg->nu_post, om, L->def, nrm2sq32(L->def, n), 1, NULL, NULL, L->sol);
  }
  for(uint64_t i = 0; i < L0->A.rows; ++i) cor[i] = (double)L0->sol[i];
}

typedef struct {
  mg32_t mg;
} mg32ctx_t;

static void mg32_precond(void* ctx, double* cor, const double* def)
{
  mg32_t* m = &((mg32ctx_t*)ctx)->mg;
  const uint64_t n = m->L[0].A.rows;
  mg32_apply(m, cor, def, orc_dot_f64(def, def, n));
}

int orc_mg_pcg_solve_mixed(const orc_mg_level_t* levels, int nlevels, const orc_mg_cfg_t* cfg,
                           const orc_iter_cfg_t* it, int flags, double* x, const double* b,
                           double* hist, orc_result_t* res)
{
  const orc_csr_t* A = &levels[0].A;
  const uint64_t n = A->rows;
  mg32ctx_t ctx;
  mg32_init(&ctx.mg, levels, nlevels, cfg, flags);
  double* r = (double*)malloc(sizeof(double) * (n ? n : 1));
  csr_apply(A, r, -1.0, x, 1.0, b);
  orc_unit_filter_def_f64(r, levels[0].filt, levels[0].nfilt);
  int st = pcg_intern(A, levels[0].filt, levels[0].nfilt, mg32_precond, &ctx, it, x, r, hist, res);
  free(r);
  mg32_free(&ctx.mg);
  return st;
}
