// feat3amd_stokes.hip -- gfx950 kernels of the Stokes path: SaddlePointMatrix /
// AmaVanka-matrix apply, the pressure mean filter, BWrappedCSR transfers and the
// AmaVanka numeric factorisation (batched 21x21 Gauss-Jordan, one workgroup per
// macro, local matrices in LDS).
//
// Every per-row sum runs in the reference's storage order with one fma per
// product (bit-identical to the generic BCSR loops); the local inverses follow
// Math::invert_matrix operation by operation, and every Vanka entry is the sum
// of its macros' contributions in ascending macro order, as the sequential
// scatter-add forms it -- so the Vanka matrix is bit-identical too.  The C3
// sizes (33k velocity + 12k pressure dofs at level 6) live in the MALL/L2, so
// these kernels are latency-bound; one thread per block row keeps the
// reference's arithmetic exactly.
#include "feat3amd_common.hpp"

namespace {

constexpr int kMaxStride = 32;   // local Vanka matrices up to 32x32 in LDS
constexpr int kVankaThreads = 64;

// ---------------------------------------------------------------------------
// tuple / saddle-point apply
// ---------------------------------------------------------------------------
struct TupleDev
{
  uint64_t nv, np;
  const double* val[4];
  const uint64_t* col[4];
  const uint64_t* ptr[4];
};

constexpr int kUnr = 8;   // blocks per row fetched ahead in the tuple SpMV
__global__ __launch_bounds__(kThreads) void k_tuple_apply(TupleDev M, double* __restrict__ r,
                                                          const double* __restrict__ x,
                                                          const double* __restrict__ y, double alpha,
                                                          const uint32_t* __restrict__ vmask)
{
  const uint64_t nrows = M.nv + M.np;
  const double* xp = x + 2 * M.nv;
  for(uint64_t row = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; row < nrows;
      row += uint64_t(gridDim.x) * blockDim.x)
  {
    if(row < M.nv)
    {
      double s0 = 0.0, s1 = 0.0;
      uint64_t k = M.ptr[0][row];
      const uint64_t ke = M.ptr[0][row + 1];
      // blocks are fetched kUnr at a time before any x is gathered (same fma order)
      for(; k + kUnr <= ke; k += kUnr)
      {
        double4 b[kUnr];
        uint64_t c[kUnr];
        double2 xv[kUnr];
#pragma unroll
        for(int u = 0; u < kUnr; ++u)
        {
          b[u] = reinterpret_cast<const double4*>(M.val[0])[k + u];
          c[u] = M.col[0][k + u];
        }
#pragma unroll
        for(int u = 0; u < kUnr; ++u) xv[u] = reinterpret_cast<const double2*>(x)[c[u]];
#pragma unroll
        for(int u = 0; u < kUnr; ++u)
        {
          s0 = dfma(b[u].x, xv[u].x, s0);
          s0 = dfma(b[u].y, xv[u].y, s0);
          s1 = dfma(b[u].z, xv[u].x, s1);
          s1 = dfma(b[u].w, xv[u].y, s1);
        }
      }
      for(; k < ke; ++k)
      {
        const double* b = M.val[0] + 4 * k;
        const uint64_t c = M.col[0][k];
        const double x0 = x[2 * c], x1 = x[2 * c + 1];
        s0 = dfma(b[0], x0, s0);
        s0 = dfma(b[1], x1, s0);
        s1 = dfma(b[2], x0, s1);
        s1 = dfma(b[3], x1, s1);
      }
      // A.apply(r, x_u, y_u, alpha): r = sum*alpha + beta*y (beta = 0 without y)
      double r0 = y ? dfma(s0, alpha, y[2 * row]) : dfma(s0, 1.0, 0.0);
      double r1 = y ? dfma(s1, alpha, y[2 * row + 1]) : dfma(s1, 1.0, 0.0);
      const double a1 = y ? alpha : 1.0;
      if(M.val[1])
      {
        double t0 = 0.0, t1 = 0.0;
        uint64_t kb = M.ptr[1][row];
        const uint64_t kbe = M.ptr[1][row + 1];
        for(; kb + kUnr <= kbe; kb += kUnr)
        {
          double2 b[kUnr];
          uint64_t c[kUnr];
          double xv[kUnr];
#pragma unroll
          for(int u = 0; u < kUnr; ++u)
          {
            b[u] = reinterpret_cast<const double2*>(M.val[1])[kb + u];
            c[u] = M.col[1][kb + u];
          }
#pragma unroll
          for(int u = 0; u < kUnr; ++u) xv[u] = xp[c[u]];
#pragma unroll
          for(int u = 0; u < kUnr; ++u)
          {
            t0 = dfma(b[u].x, xv[u], t0);
            t1 = dfma(b[u].y, xv[u], t1);
          }
        }
        for(; kb < kbe; ++kb)
        {
          const double* b = M.val[1] + 2 * kb;
          const double xv = xp[M.col[1][kb]];
          t0 = dfma(b[0], xv, t0);
          t1 = dfma(b[1], xv, t1);
        }
        // B.apply(r, x_p, r, alpha)
        r0 = dfma(t0, a1, r0);
        r1 = dfma(t1, a1, r1);
      }
      if(is_masked(vmask, row)) r0 = r1 = 0.0;
      r[2 * row] = r0;
      r[2 * row + 1] = r1;
    }
    else
    {
      const uint64_t pr = row - M.nv;
      double s = 0.0;
      uint64_t k = M.ptr[2][pr];
      const uint64_t ke = M.ptr[2][pr + 1];
      for(; k + kUnr <= ke; k += kUnr)
      {
        double2 b[kUnr], xv[kUnr];
        uint64_t c[kUnr];
#pragma unroll
        for(int u = 0; u < kUnr; ++u)
        {
          b[u] = reinterpret_cast<const double2*>(M.val[2])[k + u];
          c[u] = M.col[2][k + u];
        }
#pragma unroll
        for(int u = 0; u < kUnr; ++u) xv[u] = reinterpret_cast<const double2*>(x)[c[u]];
#pragma unroll
        for(int u = 0; u < kUnr; ++u)
        {
          s = dfma(b[u].x, xv[u].x, s);
          s = dfma(b[u].y, xv[u].y, s);
        }
      }
      for(; k < ke; ++k)
      {
        const double* b = M.val[2] + 2 * k;
        const uint64_t c = M.col[2][k];
        s = dfma(b[0], x[2 * c], s);
        s = dfma(b[1], x[2 * c + 1], s);
      }
      double rp = y ? dfma(s, alpha, y[2 * M.nv + pr]) : dfma(s, 1.0, 0.0);
      if(M.val[3])
      {
        double t = 0.0;
        for(uint64_t k = M.ptr[3][pr]; k < M.ptr[3][pr + 1]; ++k)
          t = dfma(M.val[3][k], xp[M.col[3][k]], t);
        rp = dfma(t, y ? alpha : 1.0, rp);
      }
      r[2 * M.nv + pr] = rp;
    }
  }
}

// ---------------------------------------------------------------------------
// mean filter passes (one workgroup) and small vector kernels
// ---------------------------------------------------------------------------
constexpr int kMeanThreads = 1024;
__global__ __launch_bounds__(kMeanThreads) void k_mean_passes(double* __restrict__ v, const double* __restrict__ w_dot,
                                                              const double* __restrict__ w_add, double volume,
                                                              uint64_t n, int passes, int sol_form)
{
  __shared__ double red[kMeanThreads / 64];
  __shared__ double s_alpha;
  for(int pass = 0; pass < passes; ++pass)
  {
    // thread t adds v_i w_i for i = t, t + 1024, ... in order, four products fetched ahead
    double part = 0.0;
    uint64_t i = threadIdx.x;
    for(; i + 3 * kMeanThreads < n; i += 4 * kMeanThreads)
    {
      double a[4], b[4];
#pragma unroll
      for(int u = 0; u < 4; ++u)
      {
        a[u] = v[i + u * kMeanThreads];
        b[u] = w_dot[i + u * kMeanThreads];
      }
#pragma unroll
      for(int u = 0; u < 4; ++u) part += a[u] * b[u];
    }
    for(; i < n; i += kMeanThreads) part += v[i] * w_dot[i];
    const double dot = block_sum(part, red);
    if(threadIdx.x == 0) s_alpha = sol_form ? 0.0 - dot / volume : -dot / volume;
    __syncthreads();
    const double alpha = s_alpha;
    if(fabs(alpha) >= DBL_EPSILON)
      for(uint64_t i = threadIdx.x; i < n; i += blockDim.x) v[i] = dfma(alpha, w_add[i], v[i]);
    __syncthreads();
  }
}

__global__ __launch_bounds__(kThreads) void k_sumsq2(const double* __restrict__ v, uint64_t nu, uint64_t np,
                                                     GridReduce gr)
{
  double su = 0.0, sp = 0.0;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for(uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nu + np; i += stride)
  {
    const double t = v[i];
    if(i < nu)
      su = dfma(t, t, su);
    else
      sp = dfma(t, t, sp);
  }
  grid_reduce_finish(gr, su, sp);
}

__global__ void k_tuple_norm_finish(const double* __restrict__ sums, double* __restrict__ out)
{
  const double nu = sqrt(sums[0]), np = sqrt(sums[1]);
  *out = sqrt(dfma(nu, nu, np * np));
}

__global__ __launch_bounds__(kThreads) void k_axpy_gated(double* __restrict__ x, double omega,
                                                         const double* __restrict__ c, uint64_t n,
                                                         const double* __restrict__ gate)
{
  if(gate)
  {
    const double g = *gate;
    if(!isfinite(g) || g <= DBL_EPSILON * DBL_EPSILON) return;
  }
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for(uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    x[i] = dfma(omega, c[i], x[i]);
}

__global__ __launch_bounds__(kThreads) void k_csrsb2(double* __restrict__ r, const double* __restrict__ x,
                                                     const double* __restrict__ val,
                                                     const uint64_t* __restrict__ col,
                                                     const uint64_t* __restrict__ rowptr, uint64_t rows)
{
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for(uint64_t row = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; row < rows; row += stride)
  {
    double s0 = 0.0, s1 = 0.0;
    for(uint64_t k = rowptr[row]; k < rowptr[row + 1]; ++k)
    {
      const double a = val[k];
      const uint64_t c = col[k];
      s0 = dfma(a, x[2 * c], s0);
      s1 = dfma(a, x[2 * c + 1], s1);
    }
    // (bsum * a) + (b * r) with a = 1, b = 0
    r[2 * row] = dfma(s0, 1.0, 0.0);
    r[2 * row + 1] = dfma(s1, 1.0, 0.0);
  }
}

// ---------------------------------------------------------------------------
// AmaVanka numeric factorisation
// ---------------------------------------------------------------------------
struct VankaGraphDev
{
  uint64_t nm, nv, np, stride;
  const uint64_t *vptr, *vidx, *pptr, *pidx, *dvptr, *dvidx, *dpptr, *dpidx;
};

// One workgroup (one wavefront) per macro: gather, Gauss-Jordan, singularity check.
__global__ __launch_bounds__(kVankaThreads) void k_vanka_local(VankaGraphDev g, TupleDev K,
                                                               const uint32_t* __restrict__ vmask,
                                                               int skip_singular, double* __restrict__ inv_out,
                                                               int* __restrict__ macro_mask)
{
  __shared__ double a[kMaxStride * kMaxStride];
  __shared__ double t[kMaxStride * kMaxStride];
  __shared__ double xs[kMaxStride * kMaxStride];
  __shared__ double fac[kMaxStride];
  __shared__ int p[kMaxStride];
  __shared__ int s_pk;
  const uint64_t m = blockIdx.x;
  const int lane = threadIdx.x;
  const int S = int(g.stride);
  const uint64_t* vl = g.vidx + g.vptr[m];
  const uint64_t* pl = g.pidx + g.pptr[m];
  const int nvm = int(g.vptr[m + 1] - g.vptr[m]);
  const int npm = int(g.pptr[m + 1] - g.pptr[m]);
  const int nu = 2 * nvm, n = nu + npm;
  for(int e = lane; e < S * S; e += kVankaThreads) a[e] = 0.0;
  __syncthreads();
  // gather (amavanka.hpp:477-547, 605-625): one local row per lane
  if(lane < n)
  {
    double* ar = a + lane * S;
    if(lane < nu)
    {
      const int ii = lane & 1;
      const uint64_t irow = vl[lane >> 1];
      const bool dir = is_masked(vmask, irow);
      int j = 0;
      for(uint64_t k = K.ptr[0][irow]; k < K.ptr[0][irow + 1]; ++k)
      {
        const uint64_t icol = K.col[0][k];
        while(j < nvm && vl[j] < icol) ++j;
        if(j < nvm && vl[j] == icol)
          for(int jj = 0; jj < 2; ++jj)
            ar[2 * j + jj] = dir ? ((icol == irow && jj == ii) ? 1.0 : 0.0) : K.val[0][4 * k + 2 * ii + jj];
      }
      j = 0;
      for(uint64_t k = K.ptr[1][irow]; k < K.ptr[1][irow + 1]; ++k)
      {
        const uint64_t icol = K.col[1][k];
        while(j < npm && pl[j] < icol) ++j;
        if(j < npm && pl[j] == icol) ar[nu + j] = dir ? 0.0 : K.val[1][2 * k + ii];
      }
    }
    else
    {
      const uint64_t irow = pl[lane - nu];
      int j = 0;
      for(uint64_t k = K.ptr[2][irow]; k < K.ptr[2][irow + 1]; ++k)
      {
        const uint64_t icol = K.col[2][k];
        while(j < nvm && vl[j] < icol) ++j;
        if(j < nvm && vl[j] == icol)
        {
          ar[2 * j] = K.val[2][2 * k];
          ar[2 * j + 1] = K.val[2][2 * k + 1];
        }
      }
    }
    p[lane] = lane;
  }
  __syncthreads();
  if(skip_singular)
    for(int e = lane; e < S * S; e += kVankaThreads) t[e] = a[e];
  // Gauss-Jordan with diagonal pivoting (util/math.hpp:1184-1268)
  if(n == 1)
  {
    if(lane == 0) a[0] = 1.0 / a[0];
    __syncthreads();
  }
  else
    for(int k = 0; k < n; ++k)
    {
      if(lane == 0)
      {
        double piv = fabs(a[p[k] * S + p[k]]);
        int i = k;
        for(int j = k + 1; j < n; ++j)
        {
          const double av = fabs(a[p[j] * S + p[j]]);
          if(av > piv)
          {
            piv = av;
            i = j;
          }
        }
        if(i > k)
        {
          const int tmp = p[k];
          p[k] = p[i];
          p[i] = tmp;
        }
        s_pk = p[k];
      }
      __syncthreads();
      const int pk = s_pk, off = pk * S;
      const double inv = 1.0 / a[off + pk];
      if(lane < n) fac[lane] = a[lane * S + pk];
      __syncthreads();
      if(lane < n) a[off + lane] = (lane == pk ? 1.0 : a[off + lane]) * inv;
      __syncthreads();
      for(int e = lane; e < n * n; e += kVankaThreads)
      {
        const int r = e / n, j = e - r * n;
        if(r == pk) continue;
        const double base = (j == pk) ? 0.0 : a[r * S + j];
        a[r * S + j] = dfma(-a[off + j], fac[r], base);
      }
      __syncthreads();
    }
  // singularity check |I - A*A^-1|_F^2 < eps (amavanka.hpp:1506-1530)
  if(skip_singular)
  {
    for(int e = lane; e < n * n; e += kVankaThreads)
    {
      const int i = e / n, j = e - i * n;
      double xij = (i == j) ? 1.0 : 0.0;
      for(int k = 0; k < n; ++k) xij = dfma(-t[i * S + k], a[k * S + j], xij);
      xs[e] = xij;
    }
    __syncthreads();
    if(lane == 0)
    {
      double norm = 0.0;
      for(int e = 0; e < n * n; ++e) norm = dfma(xs[e], xs[e], norm);
      macro_mask[m] = (norm < DBL_EPSILON) ? 1 : 0;
    }
  }
  double* out = inv_out + m * uint64_t(S) * uint64_t(S);
  for(int e = lane; e < S * S; e += kVankaThreads) out[e] = a[e];
}

__device__ __forceinline__ int find_pos(const uint64_t* list, int len, uint64_t v)
{
  for(int i = 0; i < len; ++i)
    if(list[i] == v) return i;
  return -1;
}

// One thread per Vanka block row: every entry is the ascending-macro sum of the
// local inverses (the scatter-add order), times omega / #macros (scale_rows).
__global__ __launch_bounds__(kThreads) void k_vanka_assemble(VankaGraphDev g, TupleDev V,
                                                             const double* __restrict__ inv,
                                                             const int* __restrict__ macro_mask, double omega)
{
  const uint64_t nrows = g.nv + g.np;
  const int S = int(g.stride);
  for(uint64_t row = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; row < nrows;
      row += uint64_t(gridDim.x) * blockDim.x)
  {
    const bool vel = row < g.nv;
    const uint64_t dof = vel ? row : row - g.nv;
    const uint64_t* mp = vel ? g.dvptr : g.dpptr;
    const uint64_t* mi = vel ? g.dvidx : g.dpidx;
    uint64_t cnt = 0;
    for(uint64_t q = mp[dof]; q < mp[dof + 1]; ++q) cnt += macro_mask ? uint64_t(macro_mask[mi[q]]) : 1u;
    const double sc = omega / double(cnt > 1 ? cnt : 1);
    const int bh = vel ? 2 : 1;
    for(int b = 0; b < 2; ++b)
    {
      const int blk = vel ? b : 2 + b;       // velocity rows: blocks 0, 1; pressure rows: 2, 3
      const bool colvel = (b == 0);
      const int bw = colvel ? 2 : 1;
      double* val = const_cast<double*>(V.val[blk]);
      for(uint64_t k = V.ptr[blk][dof]; k < V.ptr[blk][dof + 1]; ++k)
      {
        const uint64_t c = V.col[blk][k];
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for(uint64_t q = mp[dof]; q < mp[dof + 1]; ++q)
        {
          const uint64_t m = mi[q];
          if(macro_mask && !macro_mask[m]) continue;
          const uint64_t* vl = g.vidx + g.vptr[m];
          const uint64_t* pl = g.pidx + g.pptr[m];
          const int nvm = int(g.vptr[m + 1] - g.vptr[m]), npm = int(g.pptr[m + 1] - g.pptr[m]);
          const int nu = 2 * nvm;
          const int li = vel ? find_pos(vl, nvm, dof) : find_pos(pl, npm, dof);
          const int lj = colvel ? find_pos(vl, nvm, c) : find_pos(pl, npm, c);
          if(li < 0 || lj < 0) continue;
          const int r0 = vel ? 2 * li : nu + li;
          const int c0 = colvel ? 2 * lj : nu + lj;
          const double* L = inv + m * uint64_t(S) * uint64_t(S);
          for(int ii = 0; ii < bh; ++ii)
            for(int jj = 0; jj < bw; ++jj) acc[ii * bw + jj] += L[(r0 + ii) * S + c0 + jj];
        }
        for(int e = 0; e < bh * bw; ++e) val[k * uint64_t(bh * bw) + e] = acc[e] * sc;
      }
    }
  }
}

TupleDev to_dev(const feat_tuple_f64_t* M)
{
  TupleDev d;
  d.nv = M->nv;
  d.np = M->np;
  for(int b = 0; b < 4; ++b)
  {
    d.val[b] = M->val[b];
    d.col[b] = M->col[b];
    d.ptr[b] = M->ptr[b];
  }
  return d;
}

bool tuple_ok(const feat_tuple_f64_t* M, int first, int last)
{
  if(!M) return false;
  for(int b = first; b <= last; ++b)
    if(!M->val[b] || !M->col[b] || !M->ptr[b]) return false;
  return true;
}

int rows_grid(uint64_t rows) { return grid_for(rows, 1); }

} // namespace

extern "C" {

int feat_tuple_apply_f64(feat_handle_t h, const feat_tuple_f64_t* M, double* r, const double* x,
                         const double* y, double alpha, const uint32_t* vmask)
{
  if(!h || !tuple_ok(M, 0, 0) || !tuple_ok(M, 2, 2) || !r || !x)
    return fail(FEAT_ERR_INVALID, "feat_tuple_apply_f64: null argument");
  if(M->val[1] && (!M->col[1] || !M->ptr[1])) return fail(FEAT_ERR_INVALID, "feat_tuple_apply_f64: block 1");
  if(M->val[3] && (!M->col[3] || !M->ptr[3])) return fail(FEAT_ERR_INVALID, "feat_tuple_apply_f64: block 3");
  if(r == x) return fail(FEAT_ERR_ALIAS, "feat_tuple_apply_f64: r must not alias x");
  const uint64_t rows = M->nv + M->np;
  if(!rows) return FEAT_OK;
  hipLaunchKernelGGL(k_tuple_apply, dim3(rows_grid(rows)), dim3(kThreads), 0, h->stream, to_dev(M), r, x, y, alpha,
                     vmask);
  return check_launch("feat_tuple_apply_f64");
}

int feat_mean_filter_passes_f64(feat_handle_t h, double* v, const double* w_dot, const double* w_add,
                                double volume, uint64_t n, int passes, int sol_form)
{
  if(!h || (n && (!v || !w_dot || !w_add)) || passes < 0 || passes > 2)
    return fail(FEAT_ERR_INVALID, "feat_mean_filter_passes_f64: bad argument");
  if(!n || !passes) return FEAT_OK;
  hipLaunchKernelGGL(k_mean_passes, dim3(1), dim3(kMeanThreads), 0, h->stream, v, w_dot, w_add, volume, n, passes,
                     sol_form);
  return check_launch("feat_mean_filter_passes_f64");
}

int feat_tuple_norm2_f64(feat_handle_t h, const double* v, uint64_t nu, uint64_t np, double* out,
                         const feat_reduce_ws_t* ws)
{
  if(!h || !out || (nu + np && !v)) return fail(FEAT_ERR_INVALID, "feat_tuple_norm2_f64: null argument");
  if(need_ws(ws, out)) return fail(FEAT_ERR_INVALID, "feat_tuple_norm2_f64: workspace required");
  double* sums = h->scal + 8;   // two scratch scalars of the handle
  const int grid = grid_for(nu + np, 8);
  const GridReduce gd = reduce_for_grid(ws_reduce(ws, sums, sums + 1), uint64_t(grid));
  hipLaunchKernelGGL(k_sumsq2, dim3(grid), dim3(kThreads), 0, h->stream, v, nu, np, gd);
  finish_reduce(h->stream, gd, uint64_t(grid));
  hipLaunchKernelGGL(k_tuple_norm_finish, dim3(1), dim3(1), 0, h->stream, sums, out);
  return check_launch("feat_tuple_norm2_f64");
}

int feat_axpy_gated_f64(feat_handle_t h, double* x, double omega, const double* c, uint64_t n,
                        const double* gate_norm)
{
  if(!h || (n && (!x || !c))) return fail(FEAT_ERR_INVALID, "feat_axpy_gated_f64: null argument");
  // DenseVector::axpy: |alpha| < eps copies y (= x) -- nothing to do
  if(!n || fabs(omega) < DBL_EPSILON) return FEAT_OK;
  hipLaunchKernelGGL(k_axpy_gated, dim3(grid_for(n, 4)), dim3(kThreads), 0, h->stream, x, omega, c, n, gate_norm);
  return check_launch("feat_axpy_gated_f64");
}

int feat_csrsb2_apply_f64_u64(feat_handle_t h, double* r, const double* x, const double* val, const uint64_t* col,
                              const uint64_t* rowptr, uint64_t rows)
{
  if(!h || (rows && (!r || !x || !rowptr))) return fail(FEAT_ERR_INVALID, "feat_csrsb2_apply_f64_u64: null argument");
  if(r == x) return fail(FEAT_ERR_ALIAS, "feat_csrsb2_apply_f64_u64: r must not alias x");
  if(!rows) return FEAT_OK;
  hipLaunchKernelGGL(k_csrsb2, dim3(rows_grid(rows)), dim3(kThreads), 0, h->stream, r, x, val, col, rowptr, rows);
  return check_launch("feat_csrsb2_apply_f64_u64");
}

int feat_vanka_numeric_f64(feat_handle_t h, const feat_vanka_graph_t* g, const feat_tuple_f64_t* K,
                           const uint32_t* vmask, double omega, int skip_singular, double* local_inv,
                           int* macro_mask, const feat_tuple_f64_t* V)
{
  if(!h || !g || !tuple_ok(K, 0, 2) || !tuple_ok(V, 0, 3) || !local_inv || (skip_singular && !macro_mask))
    return fail(FEAT_ERR_INVALID, "feat_vanka_numeric_f64: null argument");
  if(g->stride < 1 || g->stride > uint64_t(kMaxStride))
    return fail(FEAT_ERR_UNSUPPORTED, "feat_vanka_numeric_f64: local matrix size must be 1..32");
  if(K->nv != g->nv || K->np != g->np || V->nv != g->nv || V->np != g->np)
    return fail(FEAT_ERR_INVALID, "feat_vanka_numeric_f64: size mismatch");
  VankaGraphDev gd{g->nm, g->nv, g->np, g->stride, g->vptr, g->vidx, g->pptr, g->pidx,
                   g->dvptr, g->dvidx, g->dpptr, g->dpidx};
  if(g->nm)
    hipLaunchKernelGGL(k_vanka_local, dim3(unsigned(g->nm)), dim3(kVankaThreads), 0, h->stream, gd, to_dev(K),
                       vmask, skip_singular, local_inv, macro_mask);
  const uint64_t rows = g->nv + g->np;
  if(rows)
    hipLaunchKernelGGL(k_vanka_assemble, dim3(rows_grid(rows)), dim3(kThreads), 0, h->stream, gd, to_dev(V),
                       local_inv, skip_singular ? macro_mask : nullptr, omega);
  return check_launch("feat_vanka_numeric_f64");
}

} // extern "C"
