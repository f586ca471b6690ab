/*
 * feat_oracle_smoothers.c -- CPU oracle for the multicolour SOR / SSOR sweeps and
 * ILU(0).  TEST INFRASTRUCTURE ONLY (see feat_oracle.h).
 *
 * Multicolour SOR / SSOR: the semantics of the reference's CUDA backend
 * (solver/sor_precond.cu:20-33 and 280-300, solver/ssor_precond.cu:25-57 and
 * 182-235), restated for a given colouring: colours processed in ascending order of
 * their row count (ties by colour number, the std::sort of (count, colour) pairs at
 * sor_precond.cu:207-218), y cleared first, every lower entry read (a row of a colour
 * not yet processed holds y = 0), the SSOR backward sweep over the colours in the
 * same order.  The colouring restates the product's rule (greedy first fit in row
 * order over the symmetrised pattern) so that tests can check it; the reference
 * takes cusparseDcsrcolor's, which is closed.
 */
#include "feat_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* rank[i] = processing rank of row i's colour; returns the number of colours */
uint32_t orc_color_ranks(uint64_t n, const uint64_t* rowptr, const uint64_t* col, int row_colors, uint32_t* rank)
{
  uint32_t* color = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t ncol = 0;
  if(row_colors)
  {
    for(uint64_t i = 0; i < n; ++i) color[i] = (uint32_t)i;
    ncol = (uint32_t)n;
  }
  else
  {
    /* neighbour j < i of row i: a_ij != 0 or a_ji != 0 (scan of all rows j < i, O(nnz * n)
       is too slow; use the explicit transposed lists) */
    uint64_t* tcnt = (uint64_t*)calloc(n + 1, sizeof(uint64_t));
    for(uint64_t k = 0; k < n; ++k)
      for(uint64_t e = rowptr[k]; e < rowptr[k + 1]; ++e)
        if(col[e] > k) ++tcnt[col[e] + 1];
    for(uint64_t i = 0; i < n; ++i) tcnt[i + 1] += tcnt[i];
    uint64_t* tk = (uint64_t*)malloc(sizeof(uint64_t) * (tcnt[n] ? tcnt[n] : 1));
    uint64_t* fill = (uint64_t*)malloc(sizeof(uint64_t) * (n ? n : 1));
    memcpy(fill, tcnt, sizeof(uint64_t) * n);
    for(uint64_t k = 0; k < n; ++k)
      for(uint64_t e = rowptr[k]; e < rowptr[k + 1]; ++e)
        if(col[e] > k) tk[fill[col[e]]++] = k;
    uint8_t* used = (uint8_t*)calloc(n + 1, 1);
    for(uint64_t i = 0; i < n; ++i)
    {
      for(uint64_t e = rowptr[i]; e < rowptr[i + 1]; ++e)
        if(col[e] < i) used[color[col[e]]] = 1;
      for(uint64_t e = tcnt[i]; e < tcnt[i + 1]; ++e) used[color[tk[e]]] = 1;
      uint32_t c = 0;
      while(used[c]) ++c;
      color[i] = c;
      if(c + 1 > ncol) ncol = c + 1;
      for(uint64_t e = rowptr[i]; e < rowptr[i + 1]; ++e)
        if(col[e] < i) used[color[col[e]]] = 0;
      for(uint64_t e = tcnt[i]; e < tcnt[i + 1]; ++e) used[color[tk[e]]] = 0;
    }
    free(used);
    free(fill);
    free(tk);
    free(tcnt);
  }
  /* order: ascending (count, colour) -- selection by repeated minimum */
  uint64_t* cnt = (uint64_t*)calloc(ncol ? ncol : 1, sizeof(uint64_t));
  for(uint64_t i = 0; i < n; ++i) ++cnt[color[i]];
  uint32_t* crank = (uint32_t*)malloc(sizeof(uint32_t) * (ncol ? ncol : 1));
  uint8_t* done = (uint8_t*)calloc(ncol ? ncol : 1, 1);
  uint32_t nr = 0;
  for(uint32_t r = 0; r < ncol; ++r)
  {
    int64_t best = -1;
    for(uint32_t c = 0; c < ncol; ++c)
      if(!done[c] && cnt[c] && (best < 0 || cnt[c] < cnt[best])) best = c;
    if(best < 0) break;
    done[best] = 1;
    crank[best] = nr++;
  }
  for(uint64_t i = 0; i < n; ++i) rank[i] = crank[color[i]];
  free(done);
  free(crank);
  free(cnt);
  free(color);
  return nr;
}

/* y = M^{-1} x: ssor == 0: SOR (omega (x - L y) / a), ssor != 0: forward, backward, scale */
void orc_mcsor_apply(uint64_t n, const uint64_t* rowptr, const uint64_t* col, const double* val,
                     const uint32_t* rank, uint32_t ncolors, double omega, int ssor, const double* x, double* y)
{
  for(uint64_t i = 0; i < n; ++i) y[i] = 0.0;
  for(uint32_t c = 0; c < ncolors; ++c)
    for(uint64_t i = 0; i < n; ++i)
    {
      if(rank[i] != c) continue;
      double d = 0.0;
      uint64_t k = rowptr[i];
      for(; col[k] < i; ++k) d = fma(val[k], y[col[k]], d);
      /* x - omega d: one fused multiply-add in the reference's build */
      y[i] = ssor ? fma(-omega, d, x[i]) / val[k] : omega * (x[i] - d) / val[k];
    }
  if(!ssor) return;
  for(uint32_t c = 0; c < ncolors; ++c)
    for(uint64_t i = 0; i < n; ++i)
    {
      if(rank[i] != c) continue;
      double d = 0.0;
      uint64_t k = rowptr[i + 1] - 1;
      for(; col[k] > i; --k) d = fma(val[k], y[col[k]], d);
      y[i] -= omega * d / val[k];
    }
  const double s = omega * (2.0 - omega);
  for(uint64_t i = 0; i < n; ++i) y[i] = y[i] * s;
}

/* The reference's generic sequential sweeps: SORPrecondWithBackend<generic>::_apply_intern
 * (solver/sor_precond.hpp:201-222) and SSORPrecondWithBackend<generic>::_apply_intern plus
 * the omega (2 - omega) scaling (solver/ssor_precond.hpp:177-231). */
void orc_sor_generic_apply(uint64_t n, const uint64_t* rowptr, const uint64_t* col, const double* val, double omega,
                           int ssor, const double* x, double* y)
{
  for(uint64_t i = 0; i < n; ++i) y[i] = x[i];
  for(uint64_t i = 0; i < n; ++i)
  {
    double d = 0.0;
    uint64_t k = rowptr[i];
    for(; col[k] < i; ++k) d = fma(val[k], y[col[k]], d);
    /* x - omega d: one fused multiply-add in the reference's build */
      y[i] = ssor ? fma(-omega, d, x[i]) / val[k] : omega * (x[i] - d) / val[k];
  }
  if(!ssor) return;
  for(uint64_t i = n; i > 0;)
  {
    --i;
    double d = 0.0;
    uint64_t k = rowptr[i + 1] - 1;
    for(; col[k] > i; --k) d = fma(val[k], y[col[k]], d);
    y[i] -= omega * d / val[k];
  }
  const double s = omega * (2.0 - omega);
  for(uint64_t i = 0; i < n; ++i) y[i] = y[i] * s;
}

/* ILU(0) of the reference's generic backend: ILUCoreSymbolic::set_struct_csr
 * (solver/ilu_precond.hpp:104-147: L = entries left of the diagonal, U = right of it),
 * copy_data_csr (:404-432), factorize_numeric_il_du (:453-527, the "(I+L)(D+U)" form
 * with the inverted diagonal stored), solve_il / solve_du (:540-582); the reference's
 * build contracts each `a -= b * c` into one fused multiply-add.
 * l, u: n_l / n_u values in the CSR order of A's lower / upper entries; d: n values. */
void orc_ilu0_factor(uint64_t n, const uint64_t* rowptr, const uint64_t* col, const double* val, double* l,
                     double* u, double* d, uint64_t* rptr_l, uint64_t* rptr_u)
{
  /* split structure */
  uint64_t nl = 0, nu = 0;
  rptr_l[0] = rptr_u[0] = 0;
  for(uint64_t i = 0; i < n; ++i)
  {
    for(uint64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
    {
      if(col[k] < i)
        l[nl++] = val[k];
      else if(col[k] == i)
        d[i] = val[k];
      else
        u[nu++] = val[k];
    }
    rptr_l[i + 1] = nl;
    rptr_u[i + 1] = nu;
  }
  /* column indices of L and U in the same order */
  uint64_t* cl = (uint64_t*)malloc(sizeof(uint64_t) * (nl ? nl : 1));
  uint64_t* cu = (uint64_t*)malloc(sizeof(uint64_t) * (nu ? nu : 1));
  nl = nu = 0;
  for(uint64_t i = 0; i < n; ++i)
    for(uint64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
    {
      if(col[k] < i) cl[nl++] = col[k];
      if(col[k] > i) cu[nu++] = col[k];
    }
  for(uint64_t i = 0; i < n; ++i)
  {
    const uint64_t ql = rptr_l[i + 1], qu = rptr_u[i + 1];
    for(uint64_t j = rptr_l[i]; j < ql; ++j)
    {
      const uint64_t cj = cl[j];
      uint64_t pl = j, pu = rptr_u[i];
      l[j] *= d[cj];
      uint64_t k = rptr_u[cj];
      for(; k < rptr_u[cj + 1]; ++k)
      {
        const uint64_t ck = cu[k];
        if(ck >= i) break;
        for(; (pl < ql) && (cl[pl] <= ck); ++pl)
          if(cl[pl] == ck) l[pl] = fma(-l[j], u[k], l[pl]);
      }
      if((k < rptr_u[cj + 1]) && (cu[k] == i))
      {
        d[i] = fma(-l[j], u[k], d[i]);
        ++k;
      }
      for(; k < rptr_u[cj + 1]; ++k)
      {
        const uint64_t ck = cu[k];
        for(; (pu < qu) && (cu[pu] <= ck); ++pu)
          if(cu[pu] == ck) u[pu] = fma(-l[j], u[k], u[pu]);
      }
    }
    d[i] = 1.0 / d[i];
  }
  free(cl);
  free(cu);
}

/* y = (D+U)^{-1} (I+L)^{-1} x with the factors of orc_ilu0_factor */
void orc_ilu0_apply(uint64_t n, const uint64_t* rowptr, const uint64_t* col, const double* l, const double* u,
                    const double* d, const uint64_t* rptr_l, const uint64_t* rptr_u, const double* x, double* y)
{
  for(uint64_t i = 0; i < n; ++i)
  {
    double r = x[i];
    uint64_t j = rptr_l[i];
    for(uint64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
      if(col[k] < i) r = fma(-l[j++], y[col[k]], r);
    y[i] = r;
  }
  for(uint64_t i = n; i > 0;)
  {
    --i;
    double r = y[i];
    uint64_t j = rptr_u[i];
    for(uint64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
      if(col[k] > i) r = fma(-u[j++], y[col[k]], r);
    y[i] = d[i] * r;
  }
}

/* ---- ILU(p) (solver/ilu_precond.hpp, generic backend) ----
 * The level-p pattern of ILUCoreSymbolic::factorize_symbolic (ilu_precond.hpp:191-262) with
 * its linear insertion _insert (:287-333): for every row i the level-0 entries of A split into
 * L (col < i) and U (col > i); then, walking the (growing) row of L, each entry (i, cj) of
 * level lj combines with every entry (cj, ck) of the already final row cj of U, level lk,
 * into (i, ck) of level lj + lk + 1 <= p, inserted into L (ck < i) or U (ck > i) at its
 * sorted place, or lowering the level of an existing entry.  Returns 0, or -1 when the
 * capacities cap_l / cap_u (entries) are too small. */
static uint64_t ilu_insert(uint64_t* idx, int* lvl, uint64_t* n, uint64_t i, uint64_t j, int l)
{
  for(; i < *n; ++i)
  {
    if(idx[i] == j)
    {
      if(l < lvl[i]) lvl[i] = l;
      return i + 1;
    }
    if(j < idx[i]) break;
  }
  for(uint64_t k = *n; k > i; --k)
  {
    idx[k] = idx[k - 1];
    lvl[k] = lvl[k - 1];
  }
  idx[i] = j;
  lvl[i] = l;
  ++*n;
  return i + 1;
}

int orc_ilup_pattern(uint64_t n, const uint64_t* rowptr, const uint64_t* col, int p, uint64_t cap_l, uint64_t cap_u,
                     uint64_t* rl, uint64_t* cl, uint64_t* ru, uint64_t* cu, uint64_t* nl_out, uint64_t* nu_out)
{
  int* ll = (int*)malloc(sizeof(int) * (cap_l ? cap_l : 1));
  int* lu = (int*)malloc(sizeof(int) * (cap_u ? cap_u : 1));
  uint64_t nl = 0, nu = 0;
  int rc = 0;
  rl[0] = ru[0] = 0;
  for(uint64_t i = 0; i < n && !rc; ++i)
  {
    for(uint64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
    {
      if(col[k] < i)
      {
        if(nl >= cap_l) { rc = -1; break; }
        cl[nl] = col[k];
        ll[nl++] = 0;
      }
      else if(col[k] > i)
      {
        if(nu >= cap_u) { rc = -1; break; }
        cu[nu] = col[k];
        lu[nu++] = 0;
      }
    }
    if(rc) break;
    if(p >= 1)
      for(uint64_t j = rl[i]; j < nl; ++j)
      {
        const uint64_t cj = cl[j];
        const int lj = ll[j];
        uint64_t olj = j, ouj = ru[i];
        for(uint64_t k = ru[cj]; k < ru[cj + 1]; ++k)
        {
          const uint64_t ck = cu[k];
          const int lev = lj + lu[k] + 1;
          if(lev > p) continue;
          if(ck < i)
          {
            if(nl + 1 > cap_l) { rc = -1; break; }
            olj = ilu_insert(cl, ll, &nl, olj, ck, lev);
          }
          else if(ck > i)
          {
            if(nu + 1 > cap_u) { rc = -1; break; }
            ouj = ilu_insert(cu, lu, &nu, ouj, ck, lev);
          }
        }
        if(rc) break;
      }
    rl[i + 1] = nl;
    ru[i + 1] = nu;
  }
  free(ll);
  free(lu);
  *nl_out = nl;
  *nu_out = nu;
  return rc;
}

/* ILUCoreScalar::copy_data_csr (ilu_precond.hpp:404-432) and factorize_numeric_il_du
 * (:453-527) on a split pattern (rl, cl, ru, cu): values of A where the pattern has A's
 * entries, zero at fill-in positions, then the IKJ elimination; d is stored inverted. */
void orc_ilu_factor_pattern(uint64_t n, const uint64_t* rowptr, const uint64_t* col, const double* val,
                            const uint64_t* rl, const uint64_t* cl, const uint64_t* ru, const uint64_t* cu, double* l,
                            double* u, double* d)
{
  for(uint64_t i = 0; i < n; ++i)
  {
    uint64_t ra = rowptr[i];
    const uint64_t xa = rowptr[i + 1];
    for(uint64_t j = rl[i]; j < rl[i + 1]; ++j)
    {
      if(ra < xa && cl[j] == col[ra])
        l[j] = val[ra++];
      else
        l[j] = 0.0;
    }
    d[i] = val[ra++];
    for(uint64_t j = ru[i]; j < ru[i + 1]; ++j)
    {
      if(ra < xa && cu[j] == col[ra])
        u[j] = val[ra++];
      else
        u[j] = 0.0;
    }
  }
  for(uint64_t i = 0; i < n; ++i)
  {
    const uint64_t ql = rl[i + 1], qu = ru[i + 1];
    for(uint64_t j = rl[i]; j < ql; ++j)
    {
      const uint64_t cj = cl[j];
      uint64_t pl = j, pu = ru[i];
      l[j] *= d[cj];
      uint64_t k = ru[cj];
      for(; k < ru[cj + 1]; ++k)
      {
        const uint64_t ck = cu[k];
        if(ck >= i) break;
        for(; (pl < ql) && (cl[pl] <= ck); ++pl)
          if(cl[pl] == ck) l[pl] = fma(-l[j], u[k], l[pl]);
      }
      if((k < ru[cj + 1]) && (cu[k] == i))
      {
        d[i] = fma(-l[j], u[k], d[i]);
        ++k;
      }
      for(; k < ru[cj + 1]; ++k)
      {
        const uint64_t ck = cu[k];
        for(; (pu < qu) && (cu[pu] <= ck); ++pu)
          if(cu[pu] == ck) u[pu] = fma(-l[j], u[k], u[pu]);
      }
    }
    d[i] = 1.0 / d[i];
  }
}

/* solve_il then solve_du (ilu_precond.hpp:540-582) on a split pattern */
void orc_ilu_apply_pattern(uint64_t n, const uint64_t* rl, const uint64_t* cl, const uint64_t* ru, const uint64_t* cu,
                           const double* l, const double* u, const double* d, const double* x, double* y)
{
  for(uint64_t i = 0; i < n; ++i)
  {
    double r = x[i];
    for(uint64_t j = rl[i]; j < rl[i + 1]; ++j) r = fma(-l[j], y[cl[j]], r);
    y[i] = r;
  }
  for(uint64_t i = n; i > 0;)
  {
    --i;
    double r = y[i];
    for(uint64_t j = ru[i]; j < ru[i + 1]; ++j) r = fma(-u[j], y[cu[j]], r);
    y[i] = d[i] * r;
  }
}
