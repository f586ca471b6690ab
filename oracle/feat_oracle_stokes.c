/*
 * feat_oracle_stokes.c -- CPU oracle for the Stokes part of the hot path:
 * SaddlePointMatrix apply, the UnitFilterBlocked + MeanFilter system filter,
 * the AmaVanka smoother (macro deduction, local gather, Gauss-Jordan
 * inverse, scatter-add, row scaling, apply) and the multigrid
 * Richardson(MG-V(Richardson(Schwarz(AmaVanka)))) solver of the Stokes
 * benchmark (benchmarks/stokes_multigrid-bench.cpp:714-810), on one process.
 *
 * TEST INFRASTRUCTURE ONLY (see feat_oracle.h).  Same arithmetic convention:
 * explicit fma() where the reference's -O3 -march build contracts.
 *
 * Vector layout of a Stokes vector: [u (2 values per velocity node, node
 * major) | p (one value per pressure dof)].
 */
#include "feat_oracle.h"
#include "orc_internal.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EPS_F64 DBL_EPSILON

/* ------------------------------------------------------------------------- */
/* SaddlePointMatrix apply (lafem/saddle_point_matrix.hpp:446-491)            */
/* ------------------------------------------------------------------------- */
static void bcsr_apply(const orc_bcsr_t* M, double* r, double a, const double* x, double b,
                       const double* y)
{
  orc_bcsr_apply_f64(M->bh, M->bw, r, a, x, b, y, M->val, M->col, M->rowptr, M->rows);
}

void orc_saddle_apply(const orc_stokes_level_t* L, double* r, const double* x, const double* y,
                      double alpha)
{
  const uint64_t nu = 2 * L->nv;
  if(!y)
  {
    /* apply(r, x): A x, then r += B p; r_p = D u */
    bcsr_apply(&L->A, r, 1.0, x, 0.0, r);
    bcsr_apply(&L->B, r, 1.0, x + nu, 1.0, r);
    bcsr_apply(&L->D, r + nu, 1.0, x, 0.0, r + nu);
    return;
  }
  bcsr_apply(&L->A, r, alpha, x, 1.0, y);
  bcsr_apply(&L->B, r, alpha, x + nu, 1.0, r);
  bcsr_apply(&L->D, r + nu, alpha, x, 1.0, y + nu);
}

/* ------------------------------------------------------------------------- */
/* System filter: TupleFilter<UnitFilterBlocked<2>, MeanFilter>               */
/* (lafem/unit_filter_blocked.hpp:449-499, lafem/mean_filter.hpp:200-262,     */
/*  arch/unit_filter_blocked_generic.hpp:24-73)                               */
/* ------------------------------------------------------------------------- */
/* DenseVector::axpy (lafem/dense_vector.hpp:807-829): |alpha| < eps copies y
 * instead of calling the Arch kernel */
static void vec_axpy(double* r, double alpha, const double* x, const double* y, uint64_t n)
{
  if(fabs(alpha) < EPS_F64)
  {
    if(r != y) memcpy(r, y, sizeof(double) * n);
    return;
  }
  orc_axpy_f64(r, alpha, x, y, n);
}

/* MeanFilter::filter_rhs/def (w_dot = prim, w_add = dual) and filter_cor
 * (w_dot = dual, w_add = prim): v = v - (v.w_dot / vol) * w_add */
static void mean_filter(double* v, const double* w_dot, const double* w_add, double vol, uint64_t n)
{
  vec_axpy(v, -orc_dot_f64(v, w_dot, n) / vol, w_add, v, n);
}

static double mean_volume(const orc_stokes_level_t* L)
{
  return orc_dot_f64(L->prim, L->dual, L->np); /* mean_filter.hpp:73 */
}

void orc_stokes_filter(const orc_stokes_level_t* L, double* v, int kind)
{
  double* p = v + 2 * L->nv;
  const double vol = mean_volume(L);
  if(kind == ORC_FILTER_RHS || kind == ORC_FILTER_SOL)
  {
    for(uint64_t i = 0; i < L->nfilt; ++i)
    {
      v[2 * L->filt[i]] = L->filt_val[2 * i];
      v[2 * L->filt[i] + 1] = L->filt_val[2 * i + 1];
    }
  }
  else
  {
    for(uint64_t i = 0; i < L->nfilt; ++i)
      v[2 * L->filt[i]] = v[2 * L->filt[i] + 1] = 0.0;
  }
  if(kind == ORC_FILTER_SOL)
  {
    /* axpy(prim, v, sol_mean - v.dual/vol), sol_mean = 0 */
    const double alpha = 0.0 - orc_dot_f64(p, L->dual, L->np) / vol;
    vec_axpy(p, alpha, L->prim, p, L->np);
  }
  else if(kind == ORC_FILTER_COR)
    mean_filter(p, L->dual, L->prim, vol, L->np);
  else /* rhs, def */
    mean_filter(p, L->prim, L->dual, vol, L->np);
}

/* TupleVector::norm2 (lafem/tuple_vector.hpp:362-370): sqrt(sqr(|u|) + sqr(|p|)) */
static double stokes_norm2(const orc_stokes_level_t* L, const double* v)
{
  const double nu = orc_norm2_f64(v, 2 * L->nv);
  const double np = orc_norm2_f64(v + 2 * L->nv, L->np);
  return sqrt(fma(nu, nu, np * np));
}

static uint64_t stokes_n(const orc_stokes_level_t* L) { return 2 * L->nv + L->np; }

/* ------------------------------------------------------------------------- */
/* AmaVanka (solver/amavanka.hpp)                                             */
/* ------------------------------------------------------------------------- */
struct orc_vanka_s {
  uint64_t nv, np, nm, stride;
  uint64_t *vptr, *vidx, *pptr, *pidx;     /* macro -> sorted dofs */
  uint64_t *dvptr, *dvidx, *dpptr, *dpidx; /* dof -> macros (transposed graphs) */
  orc_bcsr_t uu, up, pu, pp;               /* Vanka matrix blocks */
  double *uu_val, *up_val, *pu_val, *pp_val;
  uint64_t *uu_ptr, *uu_col, *up_ptr, *up_col, *pu_ptr, *pu_col, *pp_ptr, *pp_col;
  int* mask;                               /* skip_singular: 1 = regular macro */
};

static int cmp_u64(const void* a, const void* b)
{
  const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* amavanka.hpp:224-350 deduct_macro_dofs for SaddlePointMatrix<BCSR> */
static void deduct_macros(orc_vanka_t* V, const orc_stokes_level_t* L)
{
  const uint64_t np = L->np;
  const uint64_t *rb = L->B.rowptr, *cb = L->B.col, *rd = L->D.rowptr, *cd = L->D.col;
  int* mask = (int*)calloc(np ? np : 1, sizeof(int));
  uint64_t* cnt = (uint64_t*)calloc(np ? np : 1, sizeof(uint64_t));
  uint64_t* touched = (uint64_t*)malloc(sizeof(uint64_t) * (np ? np : 1));
  V->pptr = (uint64_t*)malloc(sizeof(uint64_t) * (np + 1));
  V->pidx = (uint64_t*)malloc(sizeof(uint64_t) * (np ? np : 1));
  uint64_t nm = 0, np_idx = 0;
  V->pptr[0] = 0;
  for(uint64_t i = 0; i < np; ++i)
  {
    if(mask[i]) continue;
    uint64_t nt = 0;
    for(uint64_t kd = rd[i]; kd < rd[i + 1]; ++kd)
      for(uint64_t kb = rb[cd[kd]]; kb < rb[cd[kd] + 1]; ++kb)
      {
        const uint64_t c = cb[kb];
        if(cnt[c]++ == 0) touched[nt++] = c;
      }
    uint64_t deg = 0;
    for(uint64_t t = 0; t < nt; ++t)
      if(cnt[touched[t]] > deg) deg = cnt[touched[t]];
    /* map order = ascending pressure dof */
    qsort(touched, nt, sizeof(uint64_t), cmp_u64);
    for(uint64_t t = 0; t < nt; ++t)
    {
      if(cnt[touched[t]] == deg)
      {
        V->pidx[np_idx++] = touched[t];
        mask[touched[t]] = 1;
      }
      cnt[touched[t]] = 0;
    }
    V->pptr[++nm] = np_idx;
  }
  V->nm = nm;
  /* velocity macros: union of D's columns over the macro's pressure dofs, sorted */
  V->vptr = (uint64_t*)malloc(sizeof(uint64_t) * (nm + 1));
  uint64_t cap = 16 * (nm ? nm : 1), nv_idx = 0;
  V->vidx = (uint64_t*)malloc(sizeof(uint64_t) * cap);
  char* seen = (char*)calloc(L->nv ? L->nv : 1, 1);
  uint64_t* buf = (uint64_t*)malloc(sizeof(uint64_t) * (L->nv ? L->nv : 1));
  V->vptr[0] = 0;
  for(uint64_t m = 0; m < nm; ++m)
  {
    uint64_t nb = 0;
    for(uint64_t j = V->pptr[m]; j < V->pptr[m + 1]; ++j)
      for(uint64_t k = rd[V->pidx[j]]; k < rd[V->pidx[j] + 1]; ++k)
        if(!seen[cd[k]]) { seen[cd[k]] = 1; buf[nb++] = cd[k]; }
    qsort(buf, nb, sizeof(uint64_t), cmp_u64);
    if(nv_idx + nb > cap)
    {
      while(nv_idx + nb > cap) cap *= 2;
      V->vidx = (uint64_t*)realloc(V->vidx, sizeof(uint64_t) * cap);
    }
    for(uint64_t t = 0; t < nb; ++t) { V->vidx[nv_idx++] = buf[t]; seen[buf[t]] = 0; }
    V->vptr[m + 1] = nv_idx;
  }
  free(seen); free(buf); free(mask); free(cnt); free(touched);
}

/* transpose of a macro->dof graph (Adjacency::RenderType::transpose: sorted) */
static void transpose_graph(uint64_t nm, const uint64_t* ptr, const uint64_t* idx, uint64_t ndof,
                            uint64_t** tptr, uint64_t** tidx)
{
  uint64_t* p = (uint64_t*)calloc(ndof + 1, sizeof(uint64_t));
  for(uint64_t k = 0; k < ptr[nm]; ++k) ++p[idx[k] + 1];
  for(uint64_t i = 0; i < ndof; ++i) p[i + 1] += p[i];
  uint64_t* t = (uint64_t*)malloc(sizeof(uint64_t) * (ptr[nm] ? ptr[nm] : 1));
  uint64_t* fill = (uint64_t*)malloc(sizeof(uint64_t) * (ndof ? ndof : 1));
  memcpy(fill, p, sizeof(uint64_t) * ndof);
  for(uint64_t m = 0; m < nm; ++m)
    for(uint64_t k = ptr[m]; k < ptr[m + 1]; ++k) t[fill[idx[k]]++] = m;
  free(fill);
  *tptr = p;
  *tidx = t;
}

/* amavanka.hpp:663-673 alloc: injectify_sorted(dof_macros[row], macro_dofs[col]) */
static void alloc_block(uint64_t nrows, const uint64_t* dptr, const uint64_t* didx,
                        const uint64_t* mptr, const uint64_t* midx, uint64_t ncols, int bh, int bw,
                        orc_bcsr_t* M, uint64_t** mp, uint64_t** mc, double** mv)
{
  char* seen = (char*)calloc(ncols ? ncols : 1, 1);
  uint64_t* buf = (uint64_t*)malloc(sizeof(uint64_t) * (ncols ? ncols : 1));
  uint64_t* p = (uint64_t*)malloc(sizeof(uint64_t) * (nrows + 1));
  uint64_t cap = 32 * (nrows ? nrows : 1), nnz = 0;
  uint64_t* c = (uint64_t*)malloc(sizeof(uint64_t) * cap);
  p[0] = 0;
  for(uint64_t r = 0; r < nrows; ++r)
  {
    uint64_t nb = 0;
    for(uint64_t k = dptr[r]; k < dptr[r + 1]; ++k)
      for(uint64_t j = mptr[didx[k]]; j < mptr[didx[k] + 1]; ++j)
        if(!seen[midx[j]]) { seen[midx[j]] = 1; buf[nb++] = midx[j]; }
    qsort(buf, nb, sizeof(uint64_t), cmp_u64);
    if(nnz + nb > cap)
    {
      while(nnz + nb > cap) cap *= 2;
      c = (uint64_t*)realloc(c, sizeof(uint64_t) * cap);
    }
    for(uint64_t t = 0; t < nb; ++t) { c[nnz++] = buf[t]; seen[buf[t]] = 0; }
    p[r + 1] = nnz;
  }
  free(seen); free(buf);
  double* v = (double*)calloc(nnz ? nnz * (uint64_t)(bh * bw) : 1, sizeof(double));
  M->rows = nrows; M->cols = ncols; M->bh = bh; M->bw = bw;
  M->val = v; M->col = c; M->rowptr = p;
  *mp = p; *mc = c; *mv = v;
}

/* amavanka.hpp:477-547 gather of one BCSR block into the local matrix */
static void gather_block(const orc_bcsr_t* M, const double* vals, double* local, uint64_t stride,
                         const uint64_t* ridx, uint64_t nr, const uint64_t* cidx, uint64_t nc,
                         uint64_t roff, uint64_t coff)
{
  const int bh = M->bh, bw = M->bw;
  for(uint64_t i = 0; i < nr; ++i)
  {
    const uint64_t irow = ridx[i];
    uint64_t j = 0;
    for(uint64_t ai = M->rowptr[irow]; ai < M->rowptr[irow + 1]; ++ai)
    {
      const uint64_t icol = M->col[ai];
      while(j < nc && cidx[j] < icol) ++j;
      if(j < nc && cidx[j] == icol)
        for(int ii = 0; ii < bh; ++ii)
          for(int jj = 0; jj < bw; ++jj)
            local[(roff + i * bh + ii) * stride + coff + j * bw + jj] = vals[ai * (uint64_t)(bh * bw) + ii * bw + jj];
    }
  }
}

/* amavanka.hpp:836-908 scatter-add of one local block into a Vanka block */
static void scatter_block(const orc_bcsr_t* M, double* vals, const double* local, uint64_t stride,
                          const uint64_t* ridx, uint64_t nr, const uint64_t* cidx, uint64_t nc,
                          uint64_t roff, uint64_t coff)
{
  const int bh = M->bh, bw = M->bw;
  for(uint64_t i = 0; i < nr; ++i)
  {
    const uint64_t irow = ridx[i];
    uint64_t j = 0;
    for(uint64_t ai = M->rowptr[irow]; ai < M->rowptr[irow + 1]; ++ai)
    {
      const uint64_t icol = M->col[ai];
      while(j < nc && cidx[j] < icol) ++j;
      if(j < nc && cidx[j] == icol)
        for(int ii = 0; ii < bh; ++ii)
          for(int jj = 0; jj < bw; ++jj)
            vals[ai * (uint64_t)(bh * bw) + ii * bw + jj] += local[(roff + i * bh + ii) * stride + coff + j * bw + jj];
    }
  }
}

/* amavanka.hpp:1032-1066 scale_rows */
static void scale_block(orc_bcsr_t* M, double* vals, double omega, const uint64_t* dptr,
                        const uint64_t* didx, const int* mask)
{
  const uint64_t bs = (uint64_t)(M->bh * M->bw);
  for(uint64_t i = 0; i < M->rows; ++i)
  {
    uint64_t n = 0;
    if(!mask)
      n = dptr[i + 1] - dptr[i];
    else
      for(uint64_t j = dptr[i]; j < dptr[i + 1]; ++j) n += (uint64_t)mask[didx[j]];
    const double sc = omega / (double)(n > 1 ? n : 1);
    for(uint64_t j = M->rowptr[i]; j < M->rowptr[i + 1]; ++j)
      for(uint64_t q = 0; q < bs; ++q) vals[j * bs + q] *= sc;
  }
}

orc_vanka_t* orc_vanka_create(const orc_stokes_level_t* L, double omega, int skip_singular)
{
  orc_vanka_t* V = (orc_vanka_t*)calloc(1, sizeof(orc_vanka_t));
  V->nv = L->nv;
  V->np = L->np;
  /* type-1 matrices of the benchmark (stokes_multigrid-bench.cpp:654-662):
   * filter_mat(A) -> unit rows, filter_offdiag_row_mat(B) -> zero rows
   * (unit_filter_blocked.hpp:313-372) */
  double* a1 = (double*)malloc(sizeof(double) * 4 * (L->A.rowptr[L->nv] ? L->A.rowptr[L->nv] : 1));
  double* b1 = (double*)malloc(sizeof(double) * 2 * (L->B.rowptr[L->nv] ? L->B.rowptr[L->nv] : 1));
  memcpy(a1, L->A.val, sizeof(double) * 4 * L->A.rowptr[L->nv]);
  memcpy(b1, L->B.val, sizeof(double) * 2 * L->B.rowptr[L->nv]);
  for(uint64_t f = 0; f < L->nfilt; ++f)
  {
    const uint64_t ix = L->filt[f];
    for(uint64_t j = L->A.rowptr[ix]; j < L->A.rowptr[ix + 1]; ++j)
    {
      double* blk = a1 + 4 * j;
      blk[0] = blk[1] = blk[2] = blk[3] = 0.0;
      if(L->A.col[j] == ix) blk[0] = blk[3] = 1.0;
    }
    for(uint64_t j = L->B.rowptr[ix]; j < L->B.rowptr[ix + 1]; ++j)
      b1[2 * j] = b1[2 * j + 1] = 0.0;
  }
  deduct_macros(V, L);
  transpose_graph(V->nm, V->vptr, V->vidx, L->nv, &V->dvptr, &V->dvidx);
  transpose_graph(V->nm, V->pptr, V->pidx, L->np, &V->dpptr, &V->dpidx);
  alloc_block(L->nv, V->dvptr, V->dvidx, V->vptr, V->vidx, L->nv, 2, 2, &V->uu, &V->uu_ptr, &V->uu_col, &V->uu_val);
  alloc_block(L->nv, V->dvptr, V->dvidx, V->pptr, V->pidx, L->np, 2, 1, &V->up, &V->up_ptr, &V->up_col, &V->up_val);
  alloc_block(L->np, V->dpptr, V->dpidx, V->vptr, V->vidx, L->nv, 1, 2, &V->pu, &V->pu_ptr, &V->pu_col, &V->pu_val);
  alloc_block(L->np, V->dpptr, V->dpidx, V->pptr, V->pidx, L->np, 1, 1, &V->pp, &V->pp_ptr, &V->pp_col, &V->pp_val);
  if(skip_singular) V->mask = (int*)calloc(V->nm ? V->nm : 1, sizeof(int));
  /* calc_stride (amavanka.hpp:1118-1146) */
  uint64_t stride = 0;
  for(uint64_t m = 0; m < V->nm; ++m)
  {
    const uint64_t c = 2 * (V->vptr[m + 1] - V->vptr[m]) + (V->pptr[m + 1] - V->pptr[m]);
    if(c > stride) stride = c;
  }
  V->stride = stride;
  double* local = (double*)calloc(stride * stride + 1, sizeof(double));
  double* local_t = (double*)calloc(stride * stride + 1, sizeof(double));
  int* piv = (int*)malloc(sizeof(int) * (stride + 1));
  orc_bcsr_t A1 = L->A, B1 = L->B;
  A1.val = a1;
  B1.val = b1;
  /* init_numeric (amavanka.hpp:1478-1559) */
  for(uint64_t m = 0; m < V->nm; ++m)
  {
    const uint64_t* vi = V->vidx + V->vptr[m];
    const uint64_t* pi = V->pidx + V->pptr[m];
    const uint64_t nvm = V->vptr[m + 1] - V->vptr[m], npm = V->pptr[m + 1] - V->pptr[m];
    const uint64_t nu = 2 * nvm, n = nu + npm;
    /* gather (amavanka.hpp:605-625): A at (0,0), B at (0,nu), D at (nu,0) */
    gather_block(&A1, A1.val, local, stride, vi, nvm, vi, nvm, 0, 0);
    gather_block(&B1, B1.val, local, stride, vi, nvm, pi, npm, 0, nu);
    gather_block(&L->D, L->D.val, local, stride, pi, npm, vi, nvm, nu, 0);
    int regular = 1;
    if(skip_singular)
    {
      for(uint64_t i = 0; i < n; ++i)
        for(uint64_t j = 0; j < n; ++j) local_t[i * stride + j] = local[i * stride + j];
      orc_invert_matrix_f64((int)n, (int)stride, local, piv);
      double norm = 0.0;
      for(uint64_t i = 0; i < n; ++i)
        for(uint64_t j = 0; j < n; ++j)
        {
          double xij = (i == j) ? 1.0 : 0.0;
          for(uint64_t k = 0; k < n; ++k) xij = fma(-local_t[i * stride + k], local[k * stride + j], xij);
          norm = fma(xij, xij, norm);
        }
      regular = (norm < EPS_F64) ? 1 : 0;
      V->mask[m] = regular;
    }
    else
      orc_invert_matrix_f64((int)n, (int)stride, local, piv);
    if(regular)
    {
      scatter_block(&V->uu, V->uu_val, local, stride, vi, nvm, vi, nvm, 0, 0);
      scatter_block(&V->up, V->up_val, local, stride, vi, nvm, pi, npm, 0, nu);
      scatter_block(&V->pu, V->pu_val, local, stride, pi, npm, vi, nvm, nu, 0);
      scatter_block(&V->pp, V->pp_val, local, stride, pi, npm, pi, npm, nu, nu);
    }
    for(uint64_t i = 0; i < n; ++i)
      for(uint64_t j = 0; j < n; ++j) local[i * stride + j] = 0.0;
  }
  scale_block(&V->uu, V->uu_val, omega, V->dvptr, V->dvidx, V->mask);
  scale_block(&V->up, V->up_val, omega, V->dvptr, V->dvidx, V->mask);
  scale_block(&V->pu, V->pu_val, omega, V->dpptr, V->dpidx, V->mask);
  scale_block(&V->pp, V->pp_val, omega, V->dpptr, V->dpidx, V->mask);
  free(local); free(local_t); free(piv); free(a1); free(b1);
  return V;
}

void orc_vanka_free(orc_vanka_t* V)
{
  if(!V) return;
  free(V->vptr); free(V->vidx); free(V->pptr); free(V->pidx);
  free(V->dvptr); free(V->dvidx); free(V->dpptr); free(V->dpidx);
  free(V->uu_ptr); free(V->uu_col); free(V->uu_val);
  free(V->up_ptr); free(V->up_col); free(V->up_val);
  free(V->pu_ptr); free(V->pu_col); free(V->pu_val);
  free(V->pp_ptr); free(V->pp_col); free(V->pp_val);
  free(V->mask);
  free(V);
}

void orc_vanka_sizes(const orc_vanka_t* V, uint64_t* out)
{
  out[0] = V->nm;
  out[1] = V->uu_ptr[V->nv];
  out[2] = V->up_ptr[V->nv];
  out[3] = V->pu_ptr[V->np];
  out[4] = V->pp_ptr[V->np];
  out[5] = V->stride;
}

void orc_vanka_copy(const orc_vanka_t* V, uint64_t* ptr[4], uint64_t* col[4], double* val[4], int* mask)
{
  const orc_bcsr_t* B[4] = {&V->uu, &V->up, &V->pu, &V->pp};
  for(int b = 0; b < 4; ++b)
  {
    const uint64_t nr = B[b]->rows, nnz = B[b]->rowptr[nr], bs = (uint64_t)(B[b]->bh * B[b]->bw);
    if(ptr[b]) memcpy(ptr[b], B[b]->rowptr, sizeof(uint64_t) * (nr + 1));
    if(col[b]) memcpy(col[b], B[b]->col, sizeof(uint64_t) * nnz);
    if(val[b]) memcpy(val[b], B[b]->val, sizeof(double) * nnz * bs);
  }
  if(mask && V->mask) memcpy(mask, V->mask, sizeof(int) * V->nm);
}

/* Vanka matrix apply (TupleMatrix, lafem/tuple_matrix.hpp:305-315) + the
 * smoother's filter_cor (amavanka.hpp:1562-1567, num_steps = 1) */
void orc_vanka_apply(const orc_vanka_t* V, const orc_stokes_level_t* L, double* c, const double* d)
{
  const uint64_t nu = 2 * V->nv;
  bcsr_apply(&V->uu, c, 1.0, d, 0.0, c);
  bcsr_apply(&V->up, c, 1.0, d + nu, 1.0, c);
  bcsr_apply(&V->pu, c + nu, 1.0, d, 0.0, c + nu);
  bcsr_apply(&V->pp, c + nu, 1.0, d + nu, 1.0, c + nu);
  orc_stokes_filter(L, c, ORC_FILTER_COR);
}

/* ------------------------------------------------------------------------- */
/* Richardson(Schwarz(AmaVanka)) smoother, multigrid V-cycle, outer Richardson */
/* ------------------------------------------------------------------------- */
typedef struct {
  const orc_stokes_level_t* L;
  orc_vanka_t* V;
  double *rhs, *sol, *def, *cor, *rdef, *rcor;
} slvl_t;

/* SchwarzPrecond::apply on one process (solver/schwarz_precond.hpp:125-160):
 * local AmaVanka apply, then the global filter_cor */
static void schwarz_vanka(const slvl_t* v, double* c, const double* d)
{
  orc_vanka_apply(v->V, v->L, c, d);
  orc_stokes_filter(v->L, c, ORC_FILTER_COR);
}

/* Richardson::apply (richardson.hpp:160-226) with min_iter = max_iter = steps */
static void stokes_smooth(const slvl_t* v, int steps, double omega, double* sol, const double* rhs)
{
  const orc_stokes_level_t* L = v->L;
  const uint64_t n = stokes_n(L);
  orc_iter_cfg_t c;
  orc_iter_cfg_default(&c);
  c.min_iter = c.max_iter = (uint64_t)steps;
  itstate_t s = {&c, 0, 0, 0, 0, 0, NULL};
  memcpy(v->rdef, rhs, sizeof(double) * n);
  memset(sol, 0, sizeof(double) * n);
  int st = set_initial_defect(&s, stokes_norm2(L, v->rdef));
  while(st == ORC_PROGRESS)
  {
    schwarz_vanka(v, v->rcor, v->rdef);
    vec_axpy(sol, omega, v->rcor, sol, n);
    orc_saddle_apply(L, v->rdef, sol, rhs, -1.0);
    orc_stokes_filter(L, v->rdef, ORC_FILTER_DEF);
    ++s.num_iter; /* set_new_defect without the norm (min_iter == max_iter) */
    s.def_prev = s.def_cur;
    st = analyse_defect(&s, s.num_iter, s.def_cur, s.def_prev, 1);
  }
}

static void transfer_apply(const orc_csr_t* Mv, const orc_csr_t* Mp, double* r, uint64_t r_nv,
                           const double* x, uint64_t x_nv)
{
  /* TupleDiagMatrix(BWrappedCSR<2>, CSR) apply(r, x): per velocity row
   * bsum += val * x_block (arch/apply_generic.hpp:157-181), r = bsum */
  for(uint64_t row = 0; row < Mv->rows; ++row)
  {
    double s0 = 0.0, s1 = 0.0;
    for(uint64_t i = Mv->rowptr[row]; i < Mv->rowptr[row + 1]; ++i)
    {
      s0 = fma(Mv->val[i], x[2 * Mv->col[i]], s0);
      s1 = fma(Mv->val[i], x[2 * Mv->col[i] + 1], s1);
    }
    r[2 * row] = s0;
    r[2 * row + 1] = s1;
  }
  orc_csr_apply_f64(r + 2 * r_nv, 1.0, x + 2 * x_nv, 0.0, r + 2 * r_nv, Mp->val, Mp->col, Mp->rowptr,
                    Mp->rows, Mp->cols, 0);
}

static void stokes_mg_apply(slvl_t* lv, int nl, const orc_stokes_cfg_t* cfg, double* cor_out,
                            const double* def_in)
{
  memcpy(lv[0].rhs, def_in, sizeof(double) * stokes_n(lv[0].L));
  /* _apply_rest (multigrid.hpp:1713-1815) */
  for(int i = 0; i < nl - 1; ++i)
  {
    slvl_t* v = &lv[i];
    stokes_smooth(v, cfg->steps, cfg->omega_rich, v->sol, v->rhs);
    orc_saddle_apply(v->L, v->def, v->sol, v->rhs, -1.0);
    orc_stokes_filter(v->L, v->def, ORC_FILTER_DEF);
    transfer_apply(&v->L->Rv, &v->L->Rp, lv[i + 1].rhs, lv[i + 1].L->nv, v->def, v->L->nv);
    orc_stokes_filter(lv[i + 1].L, lv[i + 1].rhs, ORC_FILTER_DEF);
  }
  /* _apply_coarse (:1526-1570): coarse solver = Schwarz(AmaVanka) */
  schwarz_vanka(&lv[nl - 1], lv[nl - 1].sol, lv[nl - 1].rhs);
  /* _apply_prol (:1845-1980) */
  for(int i = nl - 2; i >= 0; --i)
  {
    slvl_t* v = &lv[i];
    const uint64_t n = stokes_n(v->L);
    transfer_apply(&v->L->Pv, &v->L->Pp, v->cor, v->L->nv, lv[i + 1].sol, lv[i + 1].L->nv);
    orc_stokes_filter(v->L, v->cor, ORC_FILTER_COR);
    vec_axpy(v->sol, 1.0, v->cor, v->sol, n);
    orc_saddle_apply(v->L, v->def, v->sol, v->rhs, -1.0);
    orc_stokes_filter(v->L, v->def, ORC_FILTER_DEF);
    stokes_smooth(v, cfg->steps, cfg->omega_rich, v->cor, v->def);
    vec_axpy(v->sol, 1.0, v->cor, v->sol, n);
  }
  memcpy(cor_out, lv[0].sol, sizeof(double) * stokes_n(lv[0].L));
}

/* prolongation (prol != 0: fine <- coarse with Pv/Pp of `fine`) or
 * restriction (coarse <- fine with Rv/Rp of `fine`) of a Stokes vector */
void orc_stokes_transfer(const orc_stokes_level_t* fine, const orc_stokes_level_t* coarse, int prol,
                         double* r, const double* x)
{
  if(prol)
    transfer_apply(&fine->Pv, &fine->Pp, r, fine->nv, x, coarse->nv);
  else
    transfer_apply(&fine->Rv, &fine->Rp, r, coarse->nv, x, fine->nv);
}

/* Richardson(omega, Schwarz(AmaVanka(vanka_omega)))::apply with `steps` steps */
void orc_stokes_smooth(const orc_stokes_level_t* L, int steps, double omega, double vanka_omega,
                       double* sol, const double* rhs)
{
  const uint64_t n = stokes_n(L);
  slvl_t v;
  memset(&v, 0, sizeof(v));
  v.L = L;
  v.V = orc_vanka_create(L, vanka_omega, 0);
  v.rdef = (double*)calloc(n, sizeof(double));
  v.rcor = (double*)calloc(n, sizeof(double));
  stokes_smooth(&v, steps, omega, sol, rhs);
  orc_vanka_free(v.V);
  free(v.rdef);
  free(v.rcor);
}

int orc_stokes_mg_solve(const orc_stokes_level_t* levels, int nl, const orc_stokes_cfg_t* cfg,
                        const orc_iter_cfg_t* it, double* x, const double* b, double* hist,
                        orc_result_t* res)
{
  slvl_t* lv = (slvl_t*)calloc((size_t)nl, sizeof(slvl_t));
  for(int i = 0; i < nl; ++i)
  {
    const uint64_t n = stokes_n(&levels[i]);
    lv[i].L = &levels[i];
    lv[i].V = orc_vanka_create(&levels[i], cfg->vanka_omega, i == nl - 1 ? cfg->coarse_skip_singular : 0);
    double** vecs[6] = {&lv[i].rhs, &lv[i].sol, &lv[i].def, &lv[i].cor, &lv[i].rdef, &lv[i].rcor};
    for(int k = 0; k < 6; ++k) *vecs[k] = (double*)calloc(n ? n : 1, sizeof(double));
  }
  const orc_stokes_level_t* L = &levels[0];
  const uint64_t n = stokes_n(L);
  double* d = (double*)calloc(n, sizeof(double));
  double* c = (double*)calloc(n, sizeof(double));
  /* Richardson::correct (richardson.hpp:174-185) with omega = 1 */
  orc_saddle_apply(L, d, x, b, -1.0);
  orc_stokes_filter(L, d, ORC_FILTER_DEF);
  itstate_t s = {it, 0, 0, 0, 0, 0, hist};
  int st = set_initial_defect(&s, stokes_norm2(L, d));
  while(st == ORC_PROGRESS)
  {
    stokes_mg_apply(lv, nl, cfg, c, d);
    vec_axpy(x, cfg->omega_outer, c, x, n);
    orc_saddle_apply(L, d, x, b, -1.0);
    orc_stokes_filter(L, d, ORC_FILTER_DEF);
    ++s.num_iter;
    s.def_prev = s.def_cur;
    const int calc = (it->min_iter < it->max_iter) || (it->min_stag_iter > 0);
    if(calc) s.def_cur = stokes_norm2(L, d);
    if(hist && calc) hist[s.num_iter] = s.def_cur;
    st = analyse_defect(&s, s.num_iter, s.def_cur, s.def_prev, 1);
  }
  if(res)
  {
    res->status = st;
    res->num_iter = s.num_iter;
    res->def_init = s.def_init;
    res->def_final = s.def_cur;
  }
  for(int i = 0; i < nl; ++i)
  {
    orc_vanka_free(lv[i].V);
    free(lv[i].rhs); free(lv[i].sol); free(lv[i].def); free(lv[i].cor); free(lv[i].rdef); free(lv[i].rcor);
  }
  free(lv); free(d); free(c);
  return st;
}
