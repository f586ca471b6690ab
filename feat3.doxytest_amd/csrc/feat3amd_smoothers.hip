// feat3amd_smoothers.hip -- SOR / SSOR sweeps for CSR matrices, rows of a "colour" in
// parallel, one launch per colour.
//
// Two orderings share the kernels:
//  * level sets (default): the colours are the level sets of the triangular dependency
//    graph (forward: lev(i) = 1 + max lev(j) over the lower entries j of row i; backward:
//    the same over the upper entries), processed in level order.  Every row sees the
//    final values of all rows it depends on, so the sweeps are exactly the reference's
//    generic sequential SORPrecond / SSORPrecond (solver/sor_precond.hpp:201-222,
//    ssor_precond.hpp:193-231), bit for bit (same sums in the same order).
//  * multicolour (FEAT_SOR_MULTICOLOR): what the reference's CUDA backend computes
//    (solver/sor_precond.cu:20-33, 280-300; ssor_precond.cu:25-57, 182-235): a proper
//    colouring (here greedy first fit in row order on the symmetrised pattern; the
//    reference takes cusparseDcsrcolor's), colours in ascending order of their row count
//    (ties by colour number), y cleared first so that lower entries of colours not yet
//    processed read 0 (left out here), the SSOR backward sweep over the colours in the
//    same order as the forward sweep.  Few launches, but the SSOR so defined is not
//    symmetric, and PCG around it stalls, as it would around the reference's.
// Update formulas (both orderings):
//   SOR          y_i = omega (x_i - sum_{j<i} a_ij y_j) / a_ii
//   SSOR forward y_i = (x_i - omega sum_{j<i} a_ij y_j) / a_ii
//   SSOR backward y_i -= omega (sum_{j>i} a_ij y_j) / a_ii
// (sums in the CSR's column order: ascending for j<i, descending for j>i).
//
// Layout: per sweep direction the rows are permuted colour by colour (ascending row
// number inside a colour) and each colour is cut into slices of 64 consecutive positions
// stored column-major (SELL-64): values, u32 columns, u16 row lengths.  One lane per row,
// coalesced value/column loads, gathers of y.
#include "feat3amd_common.hpp"

#include <algorithm>
#include <numeric>
#include <utility>
#include <vector>

namespace {

constexpr uint64_t kNoEntry = ~uint64_t(0);

struct SliceSet
{
  uint64_t* soff = nullptr;   // per slice: first slot; one extra at the end
  uint16_t* len = nullptr;    // per position
  double* val = nullptr;      // slots
  uint32_t* col = nullptr;
  uint64_t* map = nullptr;    // slot -> CSR entry (kNoEntry for padding)
  uint64_t slots = 0;
};

// MODE 0: SOR forward, 1: SSOR forward, 2: SSOR backward
template<int MODE>
__global__ __launch_bounds__(256) void k_sor_color(uint64_t off, uint64_t cnt, uint64_t sb,
                                                   const uint32_t* __restrict__ perm, const uint64_t* __restrict__ soff,
                                                   const uint16_t* __restrict__ len, const double* __restrict__ val,
                                                   const uint32_t* __restrict__ col, const double* __restrict__ diag,
                                                   double* __restrict__ y, const double* __restrict__ x, double omega)
{
  const uint64_t q = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if(q >= cnt) return;
  const uint64_t p = off + q;
  const uint32_t row = perm[p];
  const uint64_t base = soff[sb + q / 64] + (q % 64);
  const uint32_t L = len[p];
  double d = 0.0;
  // 8 entries in flight: loads first, then the fma chain in entry order
  for(uint32_t k = 0; k < L; k += 8)
  {
    const uint32_t rem = L - k < 8 ? L - k : 8;
    double v[8], yv[8];
#pragma unroll
    for(uint32_t t = 0; t < 8; ++t)
    {
      const uint64_t e = base + 64 * uint64_t(k + t);
      v[t] = t < rem ? val[e] : 0.0;
      yv[t] = t < rem ? y[col[e]] : 0.0;
    }
#pragma unroll
    for(uint32_t t = 0; t < 8; ++t)
      if(t < rem) d = dfma(v[t], yv[t], d);
  }
  if(MODE == 0)
    y[row] = omega * (x[row] - d) / diag[row];
  else if(MODE == 1)
    y[row] = dfma(-omega, d, x[row]) / diag[row];   // x - omega d contracted, as the reference's build does
  else
    y[row] = y[row] - omega * d / diag[row];
}

__global__ void k_gather_values(uint64_t n, const uint64_t* __restrict__ map, const double* __restrict__ src,
                                double* __restrict__ dst)
{
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    dst[i] = map[i] == kNoEntry ? 0.0 : src[map[i]];
}

template<typename T>
int upload(T*& dst, const std::vector<T>& src)
{
  int rc = check_hip(hipMalloc(&dst, sizeof(T) * (src.empty() ? 1 : src.size())), "feat_sor_init_symbolic: hipMalloc");
  if(rc) return rc;
  if(!src.empty()) rc = check_hip(hipMemcpy(dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice),
                                   "feat_sor_init_symbolic: upload");
  return rc;
}

} // namespace

// the rows of one sweep direction, colour by colour
struct Part
{
  std::vector<uint64_t> color_off;     // positions of each colour, in processing order (ncolors + 1)
  std::vector<uint64_t> slice_base;    // first slice of each colour (ncolors + 1)
  std::vector<uint32_t> rank_of_row;   // processing rank of each row's colour
  std::vector<uint32_t> perm_h;        // position -> row
  uint32_t* perm = nullptr;
  SliceSet ss;
};

struct feat_sor_s
{
  uint64_t n = 0, nnz = 0;
  int flags = 0;
  uint64_t* diag_idx = nullptr;
  double* diag = nullptr;
  Part fw, bw;
};

namespace {

void free_slices(SliceSet& s)
{
  (void)hipFree(s.soff);
  (void)hipFree(s.len);
  (void)hipFree(s.val);
  (void)hipFree(s.col);
  (void)hipFree(s.map);
  s = SliceSet();
}

void free_sor(feat_sor_s* s)
{
  if(!s) return;
  (void)hipFree(s->fw.perm);
  (void)hipFree(s->bw.perm);
  (void)hipFree(s->diag_idx);
  (void)hipFree(s->diag);
  free_slices(s->fw.ss);
  free_slices(s->bw.ss);
  delete s;
}

// positions, colour offsets and slice bases of a partition given each row's rank
void make_part(Part& P, uint64_t n, uint32_t nranks)
{
  P.color_off.assign(nranks + 1, 0);
  for(uint64_t i = 0; i < n; ++i) ++P.color_off[P.rank_of_row[i] + 1];
  for(uint32_t r = 0; r < nranks; ++r) P.color_off[r + 1] += P.color_off[r];
  P.perm_h.resize(n);
  std::vector<uint64_t> pos(P.color_off.begin(), P.color_off.end() - 1);
  for(uint64_t i = 0; i < n; ++i) P.perm_h[pos[P.rank_of_row[i]]++] = uint32_t(i);
  P.slice_base.assign(nranks + 1, 0);
  for(uint32_t r = 0; r < nranks; ++r)
    P.slice_base[r + 1] = P.slice_base[r] + (P.color_off[r + 1] - P.color_off[r] + 63) / 64;
}

// entries of each position (CSR entry numbers, in sweep order) -> SELL-64 slices per colour
int build_slices(const Part& s, const std::vector<std::vector<uint64_t>>& ent, const std::vector<uint64_t>& colv,
                 SliceSet& out)
{
  const size_t nc = s.color_off.size() - 1;
  std::vector<uint64_t> soff{0};
  std::vector<uint16_t> len(s.perm_h.size());
  std::vector<uint64_t> map;
  for(size_t c = 0; c < nc; ++c)
    for(uint64_t p0 = s.color_off[c]; p0 < s.color_off[c + 1]; p0 += 64)
    {
      const uint64_t p1 = std::min<uint64_t>(p0 + 64, s.color_off[c + 1]);
      size_t w = 0;
      for(uint64_t p = p0; p < p1; ++p)
      {
        if(ent[p].size() > 65535) return fail(FEAT_ERR_INVALID, "feat_sor_init_symbolic: row longer than 65535 entries");
        len[p] = uint16_t(ent[p].size());
        w = std::max(w, ent[p].size());
      }
      const uint64_t b = map.size();
      map.resize(b + 64 * w, kNoEntry);
      for(uint64_t p = p0; p < p1; ++p)
        for(size_t k = 0; k < ent[p].size(); ++k) map[b + 64 * k + (p - p0)] = ent[p][k];
      soff.push_back(map.size());
    }
  std::vector<uint32_t> col(map.size(), 0);
  for(size_t k = 0; k < map.size(); ++k)
    if(map[k] != kNoEntry) col[k] = uint32_t(colv[map[k]]);
  out.slots = map.size();
  int rc = upload(out.soff, soff);
  if(!rc) rc = upload(out.len, len);
  if(!rc) rc = upload(out.col, col);
  if(!rc) rc = upload(out.map, map);
  if(!rc) rc = check_hip(hipMalloc(&out.val, sizeof(double) * (map.empty() ? 1 : map.size())), "feat_sor_init_symbolic: hipMalloc");
  return rc;
}

template<int MODE>
int sweep(feat_handle_t h, const feat_sor_s* s, const Part& P, double* y, const double* x, double omega)
{
  const size_t nc = P.color_off.size() - 1;
  for(size_t c = 0; c < nc; ++c)
  {
    const uint64_t off = P.color_off[c], cnt = P.color_off[c + 1] - off;
    if(!cnt) continue;
    hipLaunchKernelGGL(k_sor_color<MODE>, dim3(unsigned((cnt + 255) / 256)), dim3(256), 0, h->stream, off, cnt,
                       P.slice_base[c], P.perm, P.ss.soff, P.ss.len, P.ss.val, P.ss.col, s->diag, y, x, omega);
  }
  return check_launch("feat_sor sweep");
}

} // namespace

extern "C" {

int feat_sor_init_symbolic(feat_handle_t h, const feat_csr_f64_t* A, int flags, feat_sor_t* out)
{
  if(!h || !A || !out) return fail(FEAT_ERR_INVALID, "feat_sor_init_symbolic: null argument");
  if(A->rows != A->cols) return fail(FEAT_ERR_INVALID, "feat_sor_init_symbolic: matrix is not square");
  if(A->rows >= (uint64_t(1) << 32)) return fail(FEAT_ERR_INVALID, "feat_sor_init_symbolic: more than 2^32-1 rows");
  const uint64_t n = A->rows, nnz = A->nnz;
  // the pattern on the host (set-up)
  std::vector<uint64_t> rp(n + 1), colv(nnz);
  int rc = check_hip(hipMemcpyAsync(rp.data(), A->rowptr, sizeof(uint64_t) * (n + 1), hipMemcpyDeviceToHost, h->stream),
                     "feat_sor_init_symbolic: download");
  if(!rc && nnz)
    rc = check_hip(hipMemcpyAsync(colv.data(), A->col, sizeof(uint64_t) * nnz, hipMemcpyDeviceToHost, h->stream),
                   "feat_sor_init_symbolic: download");
  if(!rc) rc = check_hip(hipStreamSynchronize(h->stream), "feat_sor_init_symbolic: sync");
  if(rc) return rc;
  if(rp[0] != 0 || rp[n] != nnz) return fail(FEAT_ERR_INVALID, "feat_sor_init_symbolic: bad row pointers");
  std::vector<uint64_t> didx(n, kNoEntry);
  for(uint64_t i = 0; i < n; ++i)
    for(uint64_t k = rp[i]; k < rp[i + 1]; ++k)
    {
      if(colv[k] >= n) return fail(FEAT_ERR_INVALID, "feat_sor_init_symbolic: column index out of range");
      if(k > rp[i] && colv[k] <= colv[k - 1]) return fail(FEAT_ERR_INVALID, "feat_sor_init_symbolic: unsorted row");
      if(colv[k] == i) didx[i] = k;
    }
  for(uint64_t i = 0; i < n; ++i)
    if(didx[i] == kNoEntry) return fail(FEAT_ERR_INVALID, "feat_sor_init_symbolic: missing main diagonal entry");

  auto s = new feat_sor_s();
  s->n = n;
  s->nnz = nnz;
  s->flags = flags;
  // forward entries: lower (ascending column); backward entries: upper (descending column)
  if(flags & FEAT_SOR_MULTICOLOR)
  {
    // greedy first fit in row order over the symmetrised pattern
    std::vector<uint32_t> color(n);
    uint32_t ncol = 0;
    std::vector<uint64_t> tp(n + 1, 0), tk;   // rows k < i that hold column i
    for(uint64_t k = 0; k < n; ++k)
      for(uint64_t e = rp[k]; e < rp[k + 1]; ++e)
        if(colv[e] > k) ++tp[colv[e] + 1];
    for(uint64_t i = 0; i < n; ++i) tp[i + 1] += tp[i];
    tk.resize(tp[n]);
    {
      std::vector<uint64_t> fill(tp.begin(), tp.end() - 1);
      for(uint64_t k = 0; k < n; ++k)
        for(uint64_t e = rp[k]; e < rp[k + 1]; ++e)
          if(colv[e] > k) tk[fill[colv[e]]++] = k;
    }
    std::vector<uint64_t> seen;   // seen[c] == i + 1: colour c used by a neighbour of row i
    for(uint64_t i = 0; i < n; ++i)
    {
      auto mark = [&](uint64_t j) {
        if(j < i)
        {
          if(seen.size() <= color[j]) seen.resize(color[j] + 1, 0);
          seen[color[j]] = i + 1;
        }
      };
      for(uint64_t e = rp[i]; e < rp[i + 1]; ++e) mark(colv[e]);
      for(uint64_t e = tp[i]; e < tp[i + 1]; ++e) mark(tk[e]);
      uint32_t c = 0;
      while(c < seen.size() && seen[c] == i + 1) ++c;
      color[i] = c;
      ncol = std::max(ncol, c + 1);
    }
    // colours in ascending order of their row count, ties by colour number
    std::vector<uint64_t> cnt(ncol, 0);
    for(uint64_t i = 0; i < n; ++i) ++cnt[color[i]];
    std::vector<std::pair<uint64_t, uint32_t>> ord;
    for(uint32_t c = 0; c < ncol; ++c)
      if(cnt[c]) ord.push_back({cnt[c], c});
    std::sort(ord.begin(), ord.end());
    std::vector<uint32_t> rank(ncol, 0);
    for(uint32_t r = 0; r < ord.size(); ++r) rank[ord[r].second] = r;
    s->fw.rank_of_row.resize(n);
    for(uint64_t i = 0; i < n; ++i) s->fw.rank_of_row[i] = rank[color[i]];
    make_part(s->fw, n, uint32_t(ord.size()));
    s->bw.rank_of_row = s->fw.rank_of_row;   // the backward sweep keeps the forward colour order
    make_part(s->bw, n, uint32_t(ord.size()));
  }
  else
  {
    // level sets of the lower and of the upper triangle
    s->fw.rank_of_row.assign(n, 0);
    uint32_t nf = n ? 1 : 0;
    for(uint64_t i = 0; i < n; ++i)
    {
      uint32_t l = 0;
      for(uint64_t e = rp[i]; e < rp[i + 1] && colv[e] < i; ++e) l = std::max(l, s->fw.rank_of_row[colv[e]] + 1);
      s->fw.rank_of_row[i] = l;
      nf = std::max(nf, l + 1);
    }
    s->bw.rank_of_row.assign(n, 0);
    uint32_t nb = n ? 1 : 0;
    for(uint64_t i = n; i-- > 0;)
    {
      uint32_t l = 0;
      for(uint64_t e = rp[i + 1]; e > rp[i] && colv[e - 1] > i; --e) l = std::max(l, s->bw.rank_of_row[colv[e - 1]] + 1);
      s->bw.rank_of_row[i] = l;
      nb = std::max(nb, l + 1);
    }
    make_part(s->fw, n, nf);
    make_part(s->bw, n, nb);
  }
  std::vector<std::vector<uint64_t>> fe(n), be(n);
  for(uint64_t p = 0; p < n; ++p)
  {
    const uint64_t i = s->fw.perm_h[p];
    for(uint64_t e = rp[i]; e < rp[i + 1] && colv[e] < i; ++e)
      if(s->fw.rank_of_row[colv[e]] < s->fw.rank_of_row[i]) fe[p].push_back(e);
  }
  for(uint64_t p = 0; p < n; ++p)
  {
    const uint64_t i = s->bw.perm_h[p];
    for(uint64_t e = rp[i + 1]; e > rp[i] && colv[e - 1] > i; --e) be[p].push_back(e - 1);
  }
  int rc2 = upload(s->fw.perm, s->fw.perm_h);
  if(!rc2) rc2 = upload(s->bw.perm, s->bw.perm_h);
  if(!rc2) rc2 = upload(s->diag_idx, didx);
  if(!rc2) rc2 = check_hip(hipMalloc(&s->diag, sizeof(double) * (n ? n : 1)), "feat_sor_init_symbolic: hipMalloc");
  if(!rc2) rc2 = build_slices(s->fw, fe, colv, s->fw.ss);
  if(!rc2) rc2 = build_slices(s->bw, be, colv, s->bw.ss);
  if(!rc2) rc2 = feat_sor_init_numeric(h, s, A);
  if(rc2)
  {
    free_sor(s);
    return rc2;
  }
  *out = s;
  return FEAT_OK;
}

int feat_sor_init_numeric(feat_handle_t h, feat_sor_t s, const feat_csr_f64_t* A)
{
  if(!h || !s || !A) return fail(FEAT_ERR_INVALID, "feat_sor_init_numeric: null argument");
  if(A->rows != s->n || A->nnz != s->nnz) return fail(FEAT_ERR_INVALID, "feat_sor_init_numeric: matrix does not match the plan");
  hipLaunchKernelGGL(k_gather_values, dim3(grid_for(s->n)), dim3(kThreads), 0, h->stream, s->n, s->diag_idx, A->val, s->diag);
  for(const SliceSet* ss : {&s->fw.ss, &s->bw.ss})
    if(ss->slots)
      hipLaunchKernelGGL(k_gather_values, dim3(grid_for(ss->slots)), dim3(kThreads), 0, h->stream, ss->slots, ss->map,
                         A->val, ss->val);
  return check_launch("feat_sor_init_numeric");
}

int feat_sor_colors(feat_sor_t s, int backward, int* ncolors, uint32_t* rank_of_row, uint64_t* rows_per_color)
{
  if(!s) return fail(FEAT_ERR_INVALID, "feat_sor_colors: null plan");
  const Part& P = backward ? s->bw : s->fw;
  const size_t nc = P.color_off.size() - 1;
  if(ncolors) *ncolors = int(nc);
  if(rank_of_row) std::copy(P.rank_of_row.begin(), P.rank_of_row.end(), rank_of_row);
  if(rows_per_color)
    for(size_t c = 0; c < nc; ++c) rows_per_color[c] = P.color_off[c + 1] - P.color_off[c];
  return FEAT_OK;
}

int feat_sor_apply(feat_handle_t h, feat_sor_t s, double* y, const double* x, double omega)
{
  if(!h || !s || !y || !x) return fail(FEAT_ERR_INVALID, "feat_sor_apply: null argument");
  if(y == x) return fail(FEAT_ERR_ALIAS, "feat_sor_apply: y and x alias");
  const int rc = feat_set_memory_f64(h, y, 0.0, s->n);   // a fill kernel: safe under graph capture
  return rc ? rc : sweep<0>(h, s, s->fw, y, x, omega);
}

int feat_ssor_forward_apply(feat_handle_t h, feat_sor_t s, double* y, const double* x, double omega)
{
  if(!h || !s || !y || !x) return fail(FEAT_ERR_INVALID, "feat_ssor_forward_apply: null argument");
  if(y == x) return fail(FEAT_ERR_ALIAS, "feat_ssor_forward_apply: y and x alias");
  const int rc = feat_set_memory_f64(h, y, 0.0, s->n);
  return rc ? rc : sweep<1>(h, s, s->fw, y, x, omega);
}

int feat_ssor_backward_apply(feat_handle_t h, feat_sor_t s, double* y, double omega)
{
  if(!h || !s || !y) return fail(FEAT_ERR_INVALID, "feat_ssor_backward_apply: null argument");
  return sweep<2>(h, s, s->bw, y, nullptr, omega);
}

int feat_sor_done_symbolic(feat_sor_t s)
{
  free_sor(s);
  return FEAT_OK;
}

} // extern "C"

// ============================== ILU(0) ==============================
// ILUPrecond's generic backend (solver/ilu_precond.hpp:104-147 structure, 404-432 data,
// 453-527 numeric factorisation, 540-582 solves) run level by level: row i of the
// factorisation reads only the finished rows of its L pattern (their inverted diagonal
// and U rows) and writes only its own L, D and U entries, so the rows of one level set of
// the lower triangle are factorised in parallel with the sequential operation order of
// each row kept -- the factors, and the solves, are the generic ones bit for bit (each
// `a -= b * c` as one fused multiply-add, as the reference's build contracts it).
namespace {

__global__ __launch_bounds__(256) void k_ilu_factor_level(uint64_t off, uint64_t cnt, const uint32_t* __restrict__ perm,
                                                          const uint64_t* __restrict__ rptr_l, const uint32_t* __restrict__ cl,
                                                          const uint64_t* __restrict__ rptr_u, const uint32_t* __restrict__ cu,
                                                          double* __restrict__ l, double* __restrict__ u,
                                                          double* __restrict__ d)
{
  const uint64_t q = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if(q >= cnt) return;
  const uint32_t i = perm[off + q];
  const uint64_t ql = rptr_l[i + 1], qu = rptr_u[i + 1];
  double di = d[i];
  for(uint64_t j = rptr_l[i]; j < ql; ++j)
  {
    const uint32_t cj = cl[j];
    uint64_t pl = j, pu = rptr_u[i];
    const double lj = l[j] * d[cj];
    l[j] = lj;
    const uint64_t ke = rptr_u[cj + 1];
    uint64_t k = rptr_u[cj];
    for(; k < ke; ++k)
    {
      const uint32_t ck = cu[k];
      if(ck >= i) break;
      for(; pl < ql && cl[pl] <= ck; ++pl)
        if(cl[pl] == ck) l[pl] = dfma(-lj, u[k], l[pl]);
    }
    if(k < ke && cu[k] == i)
    {
      di = dfma(-lj, u[k], di);
      ++k;
    }
    for(; k < ke; ++k)
    {
      const uint32_t ck = cu[k];
      for(; pu < qu && cu[pu] <= ck; ++pu)
        if(cu[pu] == ck) u[pu] = dfma(-lj, u[k], u[pu]);
    }
  }
  d[i] = 1.0 / di;
}

// (I+L) y = x over one level set of the lower triangle
__global__ __launch_bounds__(256) void k_ilu_solve_l(uint64_t off, uint64_t cnt, const uint32_t* __restrict__ perm,
                                                     const uint64_t* __restrict__ rptr_l, const uint32_t* __restrict__ cl,
                                                     const double* __restrict__ l, double* y, const double* x)
{
  const uint64_t q = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if(q >= cnt) return;
  const uint32_t i = perm[off + q];
  double r = x[i];
  const uint64_t j1 = rptr_l[i + 1];
  for(uint64_t j = rptr_l[i]; j < j1; j += 8)
  {
    const uint64_t rem = j1 - j < 8 ? j1 - j : 8;
    double v[8], yv[8];
#pragma unroll
    for(uint32_t t = 0; t < 8; ++t)
    {
      v[t] = t < rem ? l[j + t] : 0.0;
      yv[t] = t < rem ? y[cl[j + t]] : 0.0;
    }
#pragma unroll
    for(uint32_t t = 0; t < 8; ++t)
      if(t < rem) r = dfma(-v[t], yv[t], r);
  }
  y[i] = r;
}

// (D+U) y = y over one level set of the upper triangle (D holds the inverted diagonal)
__global__ __launch_bounds__(256) void k_ilu_solve_u(uint64_t off, uint64_t cnt, const uint32_t* __restrict__ perm,
                                                     const uint64_t* __restrict__ rptr_u, const uint32_t* __restrict__ cu,
                                                     const double* __restrict__ u, const double* __restrict__ d, double* y)
{
  const uint64_t q = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if(q >= cnt) return;
  const uint32_t i = perm[off + q];
  double r = y[i];
  const uint64_t j1 = rptr_u[i + 1];
  for(uint64_t j = rptr_u[i]; j < j1; j += 8)
  {
    const uint64_t rem = j1 - j < 8 ? j1 - j : 8;
    double v[8], yv[8];
#pragma unroll
    for(uint32_t t = 0; t < 8; ++t)
    {
      v[t] = t < rem ? u[j + t] : 0.0;
      yv[t] = t < rem ? y[cu[j + t]] : 0.0;
    }
#pragma unroll
    for(uint32_t t = 0; t < 8; ++t)
      if(t < rem) r = dfma(-v[t], yv[t], r);
  }
  y[i] = d[i] * r;
}

} // namespace

struct feat_ilu_s
{
  uint64_t n = 0, nnz = 0, nl = 0, nu = 0;
  std::vector<uint64_t> fw_off, bw_off;   // level offsets into the permutations
  uint32_t *fw_perm = nullptr, *bw_perm = nullptr;
  uint64_t *rptr_l = nullptr, *rptr_u = nullptr;
  uint32_t *cl = nullptr, *cu = nullptr;
  uint64_t *map_l = nullptr, *map_u = nullptr, *map_d = nullptr;
  double *l = nullptr, *u = nullptr, *d = nullptr;
};

namespace {

void free_ilu(feat_ilu_s* s)
{
  if(!s) return;
  for(void* p : {(void*)s->fw_perm, (void*)s->bw_perm, (void*)s->rptr_l, (void*)s->rptr_u, (void*)s->cl, (void*)s->cu,
                 (void*)s->map_l, (void*)s->map_u, (void*)s->map_d, (void*)s->l, (void*)s->u, (void*)s->d})
    (void)hipFree(p);
  delete s;
}

// level sets: rank per row -> permutation (ascending rows inside a level) and offsets
void level_perm(const std::vector<uint32_t>& lev, uint32_t nlev, std::vector<uint32_t>& perm, std::vector<uint64_t>& off)
{
  const uint64_t n = lev.size();
  off.assign(nlev + 1, 0);
  for(uint64_t i = 0; i < n; ++i) ++off[lev[i] + 1];
  for(uint32_t r = 0; r < nlev; ++r) off[r + 1] += off[r];
  perm.resize(n);
  std::vector<uint64_t> pos(off.begin(), off.end() - 1);
  for(uint64_t i = 0; i < n; ++i) perm[pos[lev[i]]++] = uint32_t(i);
}

// ILUCoreSymbolic::factorize_symbolic (ilu_precond.hpp:191-262): the level-p pattern from
// the level-0 split (rl, cl, ru, cu).  Row by row, every entry (i, cj) of the growing row of
// L combines with the entries (cj, ck) of the final row cj of U into (i, ck) of level
// lev(i,cj) + lev(cj,ck) + 1; entries of level <= p are inserted at their sorted place (or
// lower an existing entry's level), as the reference's linear insertion does.  The value
// maps ml / mu are rebuilt: A's entry index where the pattern has it, kNoEntry at fill-in.
void ilu_level_p(uint64_t n, int p, const std::vector<uint64_t>& rp, const std::vector<uint64_t>& colv,
                 std::vector<uint64_t>& rl, std::vector<uint32_t>& cl, std::vector<uint64_t>& ml,
                 std::vector<uint64_t>& ru, std::vector<uint32_t>& cu, std::vector<uint64_t>& mu)
{
  std::vector<uint64_t> nrl{0}, nru{0};
  std::vector<uint32_t> ncl, ncu;
  std::vector<int> lvl_l, lvl_u;
  auto insert = [](std::vector<uint32_t>& idx, std::vector<int>& lvl, uint64_t i, uint32_t j, int l) -> uint64_t {
    const uint64_t m = idx.size();
    for(; i < m; ++i)
    {
      if(idx[i] == j)
      {
        if(l < lvl[i]) lvl[i] = l;
        return i + 1;
      }
      if(j < idx[i]) break;
    }
    idx.insert(idx.begin() + std::ptrdiff_t(i), j);
    lvl.insert(lvl.begin() + std::ptrdiff_t(i), l);
    return i + 1;
  };
  for(uint64_t i = 0; i < n; ++i)
  {
    for(uint64_t j = rl[i]; j < rl[i + 1]; ++j)
    {
      ncl.push_back(cl[j]);
      lvl_l.push_back(0);
    }
    for(uint64_t j = ru[i]; j < ru[i + 1]; ++j)
    {
      ncu.push_back(cu[j]);
      lvl_u.push_back(0);
    }
    for(uint64_t j = nrl[i]; j < ncl.size(); ++j)
    {
      const uint32_t cj = ncl[j];
      const int lj = lvl_l[j];
      uint64_t olj = j, ouj = nru[i];
      for(uint64_t k = nru[cj]; k < nru[cj + 1]; ++k)
      {
        const uint32_t ck = ncu[k];
        const int lev = lj + lvl_u[k] + 1;
        if(lev > p) continue;
        if(ck < i)
          olj = insert(ncl, lvl_l, olj, ck, lev);
        else if(ck > i)
          ouj = insert(ncu, lvl_u, ouj, ck, lev);
      }
    }
    nrl.push_back(ncl.size());
    nru.push_back(ncu.size());
  }
  // value maps: merge each pattern row with A's (sorted) row
  std::vector<uint64_t> nml(ncl.size(), kNoEntry), nmu(ncu.size(), kNoEntry);
  for(uint64_t i = 0; i < n; ++i)
  {
    uint64_t ka = rp[i];
    for(uint64_t j = nrl[i]; j < nrl[i + 1]; ++j)
    {
      while(ka < rp[i + 1] && colv[ka] < ncl[j]) ++ka;
      if(ka < rp[i + 1] && colv[ka] == ncl[j]) nml[j] = ka;
    }
    ka = rp[i];
    for(uint64_t j = nru[i]; j < nru[i + 1]; ++j)
    {
      while(ka < rp[i + 1] && colv[ka] < ncu[j]) ++ka;
      if(ka < rp[i + 1] && colv[ka] == ncu[j]) nmu[j] = ka;
    }
  }
  rl.swap(nrl);
  cl.swap(ncl);
  ml.swap(nml);
  ru.swap(nru);
  cu.swap(ncu);
  mu.swap(nmu);
}

} // namespace

extern "C" {

int feat_ilu_init_symbolic(feat_handle_t h, const feat_csr_f64_t* A, feat_ilu_t* out)
{
  return feat_ilu_init_symbolic_p(h, A, 0, out);
}

int feat_ilu_init_symbolic_p(feat_handle_t h, const feat_csr_f64_t* A, int p, feat_ilu_t* out)
{
  if(!h || !A || !out || p < 0) return fail(FEAT_ERR_INVALID, "feat_ilu_init_symbolic: null argument or p < 0");
  if(A->rows != A->cols) return fail(FEAT_ERR_INVALID, "feat_ilu_init_symbolic: matrix is not square");
  if(A->rows >= (uint64_t(1) << 32)) return fail(FEAT_ERR_INVALID, "feat_ilu_init_symbolic: more than 2^32-1 rows");
  const uint64_t n = A->rows, nnz = A->nnz;
  std::vector<uint64_t> rp(n + 1), colv(nnz);
  int rc = check_hip(hipMemcpyAsync(rp.data(), A->rowptr, sizeof(uint64_t) * (n + 1), hipMemcpyDeviceToHost, h->stream),
                     "feat_ilu_init_symbolic: download");
  if(!rc && nnz)
    rc = check_hip(hipMemcpyAsync(colv.data(), A->col, sizeof(uint64_t) * nnz, hipMemcpyDeviceToHost, h->stream),
                   "feat_ilu_init_symbolic: download");
  if(!rc) rc = check_hip(hipStreamSynchronize(h->stream), "feat_ilu_init_symbolic: sync");
  if(rc) return rc;
  if(rp[0] != 0 || rp[n] != nnz) return fail(FEAT_ERR_INVALID, "feat_ilu_init_symbolic: bad row pointers");
  // L / D / U split (set_struct_csr: every row needs its diagonal entry)
  std::vector<uint64_t> rl{0}, ru{0}, ml, mu, md(n);
  std::vector<uint32_t> cl, cu;
  for(uint64_t i = 0; i < n; ++i)
  {
    bool diag = false;
    for(uint64_t k = rp[i]; k < rp[i + 1]; ++k)
    {
      if(colv[k] >= n) return fail(FEAT_ERR_INVALID, "feat_ilu_init_symbolic: column index out of range");
      if(k > rp[i] && colv[k] <= colv[k - 1]) return fail(FEAT_ERR_INVALID, "feat_ilu_init_symbolic: unsorted row");
      if(colv[k] < i)
      {
        cl.push_back(uint32_t(colv[k]));
        ml.push_back(k);
      }
      else if(colv[k] == i)
      {
        md[i] = k;
        diag = true;
      }
      else
      {
        cu.push_back(uint32_t(colv[k]));
        mu.push_back(k);
      }
    }
    if(!diag) return fail(FEAT_ERR_INVALID, "feat_ilu_init_symbolic: missing main diagonal entry");
    rl.push_back(cl.size());
    ru.push_back(cu.size());
  }
  if(p > 0)
    ilu_level_p(n, p, rp, colv, rl, cl, ml, ru, cu, mu);
  // level sets of L (factorisation and forward solve) and of U (backward solve)
  std::vector<uint32_t> lf(n, 0), lb(n, 0);
  uint32_t nf = n ? 1 : 0, nb = n ? 1 : 0;
  for(uint64_t i = 0; i < n; ++i)
  {
    uint32_t v = 0;
    for(uint64_t j = rl[i]; j < rl[i + 1]; ++j) v = std::max(v, lf[cl[j]] + 1);
    lf[i] = v;
    nf = std::max(nf, v + 1);
  }
  for(uint64_t i = n; i-- > 0;)
  {
    uint32_t v = 0;
    for(uint64_t j = ru[i]; j < ru[i + 1]; ++j) v = std::max(v, lb[cu[j]] + 1);
    lb[i] = v;
    nb = std::max(nb, v + 1);
  }
  auto s = new feat_ilu_s();
  s->n = n;
  s->nnz = nnz;
  s->nl = cl.size();
  s->nu = cu.size();
  std::vector<uint32_t> pf, pb;
  level_perm(lf, nf, pf, s->fw_off);
  level_perm(lb, nb, pb, s->bw_off);
  rc = upload(s->fw_perm, pf);
  if(!rc) rc = upload(s->bw_perm, pb);
  if(!rc) rc = upload(s->rptr_l, rl);
  if(!rc) rc = upload(s->rptr_u, ru);
  if(!rc) rc = upload(s->cl, cl);
  if(!rc) rc = upload(s->cu, cu);
  if(!rc) rc = upload(s->map_l, ml);
  if(!rc) rc = upload(s->map_u, mu);
  if(!rc) rc = upload(s->map_d, md);
  if(!rc) rc = check_hip(hipMalloc(&s->l, sizeof(double) * (s->nl ? s->nl : 1)), "feat_ilu_init_symbolic: hipMalloc");
  if(!rc) rc = check_hip(hipMalloc(&s->u, sizeof(double) * (s->nu ? s->nu : 1)), "feat_ilu_init_symbolic: hipMalloc");
  if(!rc) rc = check_hip(hipMalloc(&s->d, sizeof(double) * (n ? n : 1)), "feat_ilu_init_symbolic: hipMalloc");
  if(rc)
  {
    free_ilu(s);
    return rc;
  }
  *out = s;
  return FEAT_OK;
}

int feat_ilu_init_numeric(feat_handle_t h, feat_ilu_t s, const feat_csr_f64_t* A)
{
  if(!h || !s || !A) return fail(FEAT_ERR_INVALID, "feat_ilu_init_numeric: null argument");
  if(A->rows != s->n || A->nnz != s->nnz) return fail(FEAT_ERR_INVALID, "feat_ilu_init_numeric: matrix does not match the plan");
  if(s->nl)
    hipLaunchKernelGGL(k_gather_values, dim3(grid_for(s->nl)), dim3(kThreads), 0, h->stream, s->nl, s->map_l, A->val, s->l);
  if(s->nu)
    hipLaunchKernelGGL(k_gather_values, dim3(grid_for(s->nu)), dim3(kThreads), 0, h->stream, s->nu, s->map_u, A->val, s->u);
  hipLaunchKernelGGL(k_gather_values, dim3(grid_for(s->n)), dim3(kThreads), 0, h->stream, s->n, s->map_d, A->val, s->d);
  for(size_t v = 0; v + 1 < s->fw_off.size(); ++v)
  {
    const uint64_t off = s->fw_off[v], cnt = s->fw_off[v + 1] - off;
    hipLaunchKernelGGL(k_ilu_factor_level, dim3(unsigned((cnt + 255) / 256)), dim3(256), 0, h->stream, off, cnt,
                       s->fw_perm, s->rptr_l, s->cl, s->rptr_u, s->cu, s->l, s->u, s->d);
  }
  return check_launch("feat_ilu_init_numeric");
}

int feat_ilu_apply(feat_handle_t h, feat_ilu_t s, double* y, const double* x)
{
  if(!h || !s || !y || !x) return fail(FEAT_ERR_INVALID, "feat_ilu_apply: null argument");
  for(size_t v = 0; v + 1 < s->fw_off.size(); ++v)
  {
    const uint64_t off = s->fw_off[v], cnt = s->fw_off[v + 1] - off;
    hipLaunchKernelGGL(k_ilu_solve_l, dim3(unsigned((cnt + 255) / 256)), dim3(256), 0, h->stream, off, cnt, s->fw_perm,
                       s->rptr_l, s->cl, s->l, y, x);
  }
  for(size_t v = 0; v + 1 < s->bw_off.size(); ++v)
  {
    const uint64_t off = s->bw_off[v], cnt = s->bw_off[v + 1] - off;
    hipLaunchKernelGGL(k_ilu_solve_u, dim3(unsigned((cnt + 255) / 256)), dim3(256), 0, h->stream, off, cnt, s->bw_perm,
                       s->rptr_u, s->cu, s->u, s->d, y);
  }
  return check_launch("feat_ilu_apply");
}

int feat_ilu_info(feat_ilu_t s, uint64_t* nnz_l, uint64_t* nnz_u, int* levels_l, int* levels_u)
{
  if(!s) return fail(FEAT_ERR_INVALID, "feat_ilu_info: null plan");
  if(nnz_l) *nnz_l = s->nl;
  if(nnz_u) *nnz_u = s->nu;
  if(levels_l) *levels_l = int(s->fw_off.size() - 1);
  if(levels_u) *levels_u = int(s->bw_off.size() - 1);
  return FEAT_OK;
}

int feat_ilu_factors(feat_handle_t h, feat_ilu_t s, double* l, double* u, double* d)
{
  if(!h || !s) return fail(FEAT_ERR_INVALID, "feat_ilu_factors: null argument");
  int rc = FEAT_OK;
  if(l && s->nl) rc = check_hip(hipMemcpyAsync(l, s->l, sizeof(double) * s->nl, hipMemcpyDeviceToHost, h->stream), "feat_ilu_factors");
  if(!rc && u && s->nu) rc = check_hip(hipMemcpyAsync(u, s->u, sizeof(double) * s->nu, hipMemcpyDeviceToHost, h->stream), "feat_ilu_factors");
  if(!rc && d) rc = check_hip(hipMemcpyAsync(d, s->d, sizeof(double) * s->n, hipMemcpyDeviceToHost, h->stream), "feat_ilu_factors");
  if(!rc) rc = check_hip(hipStreamSynchronize(h->stream), "feat_ilu_factors");
  return rc;
}

int feat_ilu_done_symbolic(feat_ilu_t s)
{
  free_ilu(s);
  return FEAT_OK;
}

} // extern "C"
