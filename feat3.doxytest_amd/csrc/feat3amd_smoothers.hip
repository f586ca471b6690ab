// feat3amd_smoothers.hip -- multicolour SOR / SSOR sweeps for CSR matrices.
//
// What the reference's CUDA backend computes for SORPrecond / SSORPrecond
// (solver/sor_precond.cu:20-33, 280-300; solver/ssor_precond.cu:25-57, 182-235):
// the rows are coloured so that no two rows of one colour are coupled, the colours
// are processed in ascending order of their row count (ties: colour number), and
// within a colour every row is independent:
//   SOR          y = 0; per colour: y_i = omega (x_i - sum_{j<i} a_ij y_j) / a_ii
//   SSOR forward y = 0; per colour: y_i = (x_i - omega sum_{j<i} a_ij y_j) / a_ii
//   SSOR backward, the same colour order: y_i -= omega (sum_{j>i} a_ij y_j) / a_ii
// (the sums run in the CSR's column order, ascending for j<i, descending for j>i; a
// lower entry of a colour processed later reads y_j = 0 and is left out here).
// The colouring is this library's own (greedy first fit in row order on the
// symmetrised pattern; the reference takes cusparseDcsrcolor's), so the colour
// classes, and with them the preconditioner, can differ from a CUDA run; the
// update formulas and the colour order rule are the reference's.
//
// Layout: the rows are permuted colour by colour (ascending row number inside a
// colour); each colour is cut into slices of 64 consecutive positions stored
// column-major (SELL-64), one slice set for the forward entries (lower entries of
// earlier colours) and one for the backward entries (all upper entries, descending
// column order).  One launch per colour and sweep, one lane per row, coalesced
// value/column loads, gathers of y.
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
  for(uint32_t k = 0; k < L; ++k) d = dfma(val[base + 64 * uint64_t(k)], y[col[base + 64 * uint64_t(k)]], d);
  if(MODE == 0)
    y[row] = omega * (x[row] - d) / diag[row];
  else if(MODE == 1)
    y[row] = (x[row] - omega * d) / diag[row];
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

struct feat_sor_s
{
  uint64_t n = 0, nnz = 0;
  std::vector<uint64_t> color_off;     // positions of each colour, in processing order (ncolors + 1)
  std::vector<uint64_t> slice_base;    // first slice of each colour (ncolors + 1)
  std::vector<uint32_t> rank_of_row;   // processing rank of each row's colour
  uint32_t* perm = nullptr;            // position -> row
  uint64_t* diag_idx = nullptr;
  double* diag = nullptr;
  SliceSet fw, bw;
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
  (void)hipFree(s->perm);
  (void)hipFree(s->diag_idx);
  (void)hipFree(s->diag);
  free_slices(s->fw);
  free_slices(s->bw);
  delete s;
}

// entries of each position (CSR entry numbers, in sweep order) -> SELL-64 slices per colour
int build_slices(const feat_sor_s& s, const std::vector<std::vector<uint64_t>>& ent, const std::vector<uint64_t>& colv,
                 SliceSet& out)
{
  const size_t nc = s.color_off.size() - 1;
  std::vector<uint64_t> soff{0};
  std::vector<uint16_t> len(s.n);
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
int sweep(feat_handle_t h, const feat_sor_s* s, const SliceSet& ss, double* y, const double* x, double omega)
{
  const size_t nc = s->color_off.size() - 1;
  for(size_t c = 0; c < nc; ++c)
  {
    const uint64_t off = s->color_off[c], cnt = s->color_off[c + 1] - off;
    if(!cnt) continue;
    hipLaunchKernelGGL(k_sor_color<MODE>, dim3(unsigned((cnt + 255) / 256)), dim3(256), 0, h->stream, off, cnt,
                       s->slice_base[c], s->perm, ss.soff, ss.len, ss.val, ss.col, s->diag, y, x, omega);
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

  // colouring: greedy first fit in row order over the symmetrised pattern
  std::vector<uint32_t> color(n);
  uint32_t ncol = 0;
  if(flags & FEAT_SOR_ROW_COLORS)
  {
    std::iota(color.begin(), color.end(), 0u);
    ncol = uint32_t(n);
  }
  else
  {
    // rows k < i that hold column i (transposed strictly upper pattern)
    std::vector<uint64_t> tp(n + 1, 0), tk;
    for(uint64_t k = 0; k < n; ++k)
      for(uint64_t e = rp[k]; e < rp[k + 1]; ++e)
        if(colv[e] > k) ++tp[colv[e] + 1];
    for(uint64_t i = 0; i < n; ++i) tp[i + 1] += tp[i];
    tk.resize(tp[n]);
    std::vector<uint64_t> fill(tp.begin(), tp.end() - 1);
    for(uint64_t k = 0; k < n; ++k)
      for(uint64_t e = rp[k]; e < rp[k + 1]; ++e)
        if(colv[e] > k) tk[fill[colv[e]]++] = k;
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

  auto s = new feat_sor_s();
  s->n = n;
  s->nnz = nnz;
  s->rank_of_row.resize(n);
  for(uint64_t i = 0; i < n; ++i) s->rank_of_row[i] = rank[color[i]];
  s->color_off.assign(ord.size() + 1, 0);
  for(size_t r = 0; r < ord.size(); ++r) s->color_off[r + 1] = s->color_off[r] + ord[r].first;
  std::vector<uint32_t> perm(n);
  {
    std::vector<uint64_t> pos(s->color_off.begin(), s->color_off.end() - 1);
    for(uint64_t i = 0; i < n; ++i) perm[pos[s->rank_of_row[i]]++] = uint32_t(i);
  }
  s->slice_base.assign(ord.size() + 1, 0);
  for(size_t r = 0; r < ord.size(); ++r)
    s->slice_base[r + 1] = s->slice_base[r] + (s->color_off[r + 1] - s->color_off[r] + 63) / 64;
  // forward: lower entries of earlier colours (ascending column); backward: upper entries (descending)
  std::vector<std::vector<uint64_t>> fe(n), be(n);
  for(uint64_t p = 0; p < n; ++p)
  {
    const uint64_t i = perm[p];
    for(uint64_t e = rp[i]; e < rp[i + 1] && colv[e] < i; ++e)
      if(s->rank_of_row[colv[e]] < s->rank_of_row[i]) fe[p].push_back(e);
    for(uint64_t e = rp[i + 1]; e > rp[i] && colv[e - 1] > i; --e) be[p].push_back(e - 1);
  }
  rc = upload(s->perm, perm);
  if(!rc) rc = upload(s->diag_idx, didx);
  if(!rc) rc = check_hip(hipMalloc(&s->diag, sizeof(double) * (n ? n : 1)), "feat_sor_init_symbolic: hipMalloc");
  if(!rc) rc = build_slices(*s, fe, colv, s->fw);
  if(!rc) rc = build_slices(*s, be, colv, s->bw);
  if(!rc) rc = feat_sor_init_numeric(h, s, A);
  if(rc)
  {
    free_sor(s);
    return rc;
  }
  *out = s;
  return FEAT_OK;
}

int feat_sor_init_numeric(feat_handle_t h, feat_sor_t s, const feat_csr_f64_t* A)
{
  if(!h || !s || !A) return fail(FEAT_ERR_INVALID, "feat_sor_init_numeric: null argument");
  if(A->rows != s->n || A->nnz != s->nnz) return fail(FEAT_ERR_INVALID, "feat_sor_init_numeric: matrix does not match the plan");
  hipLaunchKernelGGL(k_gather_values, dim3(grid_for(s->n)), dim3(kThreads), 0, h->stream, s->n, s->diag_idx, A->val, s->diag);
  if(s->fw.slots)
    hipLaunchKernelGGL(k_gather_values, dim3(grid_for(s->fw.slots)), dim3(kThreads), 0, h->stream, s->fw.slots, s->fw.map,
                       A->val, s->fw.val);
  if(s->bw.slots)
    hipLaunchKernelGGL(k_gather_values, dim3(grid_for(s->bw.slots)), dim3(kThreads), 0, h->stream, s->bw.slots, s->bw.map,
                       A->val, s->bw.val);
  return check_launch("feat_sor_init_numeric");
}

int feat_sor_colors(feat_sor_t s, int* ncolors, uint32_t* rank_of_row, uint64_t* rows_per_color)
{
  if(!s) return fail(FEAT_ERR_INVALID, "feat_sor_colors: null plan");
  const size_t nc = s->color_off.size() - 1;
  if(ncolors) *ncolors = int(nc);
  if(rank_of_row) std::copy(s->rank_of_row.begin(), s->rank_of_row.end(), rank_of_row);
  if(rows_per_color)
    for(size_t c = 0; c < nc; ++c) rows_per_color[c] = s->color_off[c + 1] - s->color_off[c];
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
