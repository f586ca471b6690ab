// feat3amd_batch.hip -- kernels for batches of independent solves stored back
// to back (the batched MG-PCG of many uneven Poisson instances).
//
// Work is split into chunks of at most FEAT_BATCH_CHUNK elements that never
// cross an instance boundary; one workgroup per chunk reads its instance's
// scalars (alpha, beta, the active flag) once.  Per-instance sums are formed
// from the chunk partials in ascending chunk order by one thread per instance,
// so they are deterministic and independent of the launch shape.
#include "feat3amd_common.hpp"

namespace {

struct BatchDev
{
  uint32_t ninst, nchunks;
  const uint64_t* off;
  const uint32_t* chunk_inst;
  const uint64_t* chunk_lo;
  const uint64_t* chunk_hi;
  const uint32_t* inst_chunk;
  double* partials;
};

BatchDev to_dev(const feat_batch_t* b)
{
  return BatchDev{b->ninst, b->nchunks, b->off, b->chunk_inst, b->chunk_lo, b->chunk_hi, b->inst_chunk, b->partials};
}

// the chunk's partial sum into part (b.partials when null)
__device__ __forceinline__ void chunk_partial(const BatchDev& b, double v, double* part = nullptr)
{
  __shared__ double red[kThreads / 64];
  const double t = block_sum(v, red);
  if(threadIdx.x == 0) (part ? part : b.partials)[blockIdx.x] = t;
}

// instance i's sum of its chunk partials, in ascending chunk order (as k_batch_finish)
__device__ __forceinline__ double inst_sum(const BatchDev& b, const double* __restrict__ part, uint32_t i)
{
  double s = 0.0;
  for(uint32_t c = b.inst_chunk[i]; c < b.inst_chunk[i + 1]; ++c) s += part[c];
  return s;
}

__global__ __launch_bounds__(kThreads) void k_batch_dot(BatchDev b, const double* __restrict__ x,
                                                        const double* __restrict__ y, double* part)
{
  const uint64_t lo = b.chunk_lo[blockIdx.x], hi = b.chunk_hi[blockIdx.x];
  double s = 0.0;
  for(uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) s = dfma(x[i], y[i], s);
  chunk_partial(b, s, part);
}

// out[i] = sum of instance i's chunk partials (ascending); only_active: skip inactive
__global__ void k_batch_finish(BatchDev b, double* __restrict__ out, const int* __restrict__ active)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= b.ninst) return;
  if(active && !active[i]) return;
  double s = 0.0;
  for(uint32_t c = b.inst_chunk[i]; c < b.inst_chunk[i + 1]; ++c) s += b.partials[c];
  out[i] = s;
}

// g_part / tp_part (may be null): gamma / tp as the instance sums of those chunk partials
__global__ __launch_bounds__(kThreads) void k_batch_update(BatchDev b, double* __restrict__ x, double* __restrict__ r,
                                                           const double* __restrict__ p, const double* __restrict__ t,
                                                           const double* __restrict__ gamma,
                                                           const double* __restrict__ tp, const int* __restrict__ active,
                                                           const double* __restrict__ g_part,
                                                           const double* __restrict__ tp_part, double* rr_part)
{
  const uint32_t inst = b.chunk_inst[blockIdx.x];
  const uint64_t lo = b.chunk_lo[blockIdx.x], hi = b.chunk_hi[blockIdx.x];
  double s = 0.0;
  if(active[inst])
  {
    const double g = g_part ? inst_sum(b, g_part, inst) : gamma[inst];
    const double q = tp_part ? inst_sum(b, tp_part, inst) : tp[inst];
    const double alpha = g / q;
    // DenseVector::axpy: |alpha| < eps leaves x and r unchanged
    const bool upd = fabs(alpha) >= DBL_EPSILON;
    for(uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x)
    {
      double ri = r[i];
      if(upd)
      {
        x[i] = dfma(alpha, p[i], x[i]);
        ri = dfma(-alpha, t[i], ri);
        r[i] = ri;
      }
      s = dfma(ri, ri, s);
    }
  }
  chunk_partial(b, s, rr_part);
}

// g_part (may be null): gamma_new as the instance sum of those chunk partials
__global__ __launch_bounds__(kThreads) void k_batch_direction(BatchDev b, double* __restrict__ p,
                                                              const double* __restrict__ z,
                                                              const double* __restrict__ gnew,
                                                              const double* __restrict__ gold,
                                                              const int* __restrict__ active, int first,
                                                              const double* __restrict__ g_part)
{
  const uint32_t inst = b.chunk_inst[blockIdx.x];
  if(!active[inst]) return;
  const uint64_t lo = b.chunk_lo[blockIdx.x], hi = b.chunk_hi[blockIdx.x];
  const double beta = first ? 0.0 : (g_part ? inst_sum(b, g_part, inst) : gnew[inst]) / gold[inst];
  // DenseVector::axpy(p, z, beta): |beta| < eps copies z; else r == x branch: p *= beta; p += z
  const bool copy = first || fabs(beta) < DBL_EPSILON;
  for(uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x)
  {
    if(copy)
      p[i] = z[i];
    else
    {
      const double q = p[i] * beta;
      p[i] = q + z[i];
    }
  }
}

__global__ void k_batch_copy_scalars(uint32_t n, double* __restrict__ dst, const double* __restrict__ src,
                                     const int* __restrict__ active)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n && (!active || active[i])) dst[i] = src[i];
}

struct StopDev
{
  double tol_rel, tol_abs, div_rel, div_abs;
  uint64_t min_iter, max_iter;
};

// Status codes of solver/base.hpp:48-64
constexpr int kProgress = 1, kSuccess = 2, kAborted = 3, kDiverged = 4, kMaxIter = 5;

// rr_part (may be null): rr as the instance sums of those chunk partials; gamma (may be
// null): first, for the instances active on entry, gamma[i] = instance sum of g_part
__global__ void k_batch_status(uint32_t n, const double* __restrict__ rr, StopDev c, feat_batch_state_t st, int initial,
                               BatchDev b, const double* __restrict__ rr_part, double* __restrict__ gamma,
                               const double* __restrict__ g_part)
{
  __shared__ int cnt[kThreads / 64];
  int mine = 0;
  for(uint32_t i = threadIdx.x; i < n; i += blockDim.x)
  {
    if(gamma && st.active[i]) gamma[i] = inst_sum(b, g_part, i);
    const double def = sqrt(rr_part ? inst_sum(b, rr_part, i) : rr[i]);
    if(initial)
    {
      // iterative.hpp:748-770
      st.def_init[i] = def;
      st.def_cur[i] = def;
      st.iters[i] = 0;
      int s = kProgress;
      if(!isfinite(def))
        s = kAborted;
      else if(def <= DBL_EPSILON * DBL_EPSILON)
        s = kSuccess;
      st.status[i] = s;
      st.active[i] = (s == kProgress) ? 1 : 0;
      if(st.hist) st.hist[uint64_t(i) * st.hist_cap] = def;
    }
    else if(st.active[i])
    {
      // iterative.hpp:795-885 (the defect norm is computed: min_iter < max_iter)
      const uint64_t it = ++st.iters[i];
      st.def_cur[i] = def;
      if(st.hist && it < st.hist_cap) st.hist[uint64_t(i) * st.hist_cap + it] = def;
      int s = kProgress;
      if(!isfinite(def))
        s = kAborted;
      else if(def > c.div_abs || def > c.div_rel * st.def_init[i])
        s = kDiverged;
      else if(it < c.min_iter)
        s = kProgress;
      else if(def <= c.tol_abs && def <= c.tol_rel * st.def_init[i])
        s = kSuccess;
      else if(it >= c.max_iter)
        s = kMaxIter;
      st.status[i] = s;
      st.active[i] = (s == kProgress) ? 1 : 0;
    }
    mine += st.active[i];
  }
  for(int off = 32; off > 0; off >>= 1) mine += __shfl_down(mine, off, 64);
  if((threadIdx.x & 63) == 0) cnt[threadIdx.x >> 6] = mine;
  __syncthreads();
  if(threadIdx.x == 0)
  {
    int t = 0;
    for(int w = 0; w < int(blockDim.x >> 6); ++w) t += cnt[w];
    *st.nactive = t;
  }
}

__global__ __launch_bounds__(kThreads) void k_batch_gather(uint32_t nseg, const uint64_t* __restrict__ dst_lo,
                                                           const uint64_t* __restrict__ src_lo,
                                                           const uint64_t* __restrict__ len, double* __restrict__ dst,
                                                           const double* __restrict__ src)
{
  for(uint32_t sgm = blockIdx.y; sgm < nseg; sgm += gridDim.y)
  {
    const uint64_t n = len[sgm], d0 = dst_lo[sgm], s0 = src_lo[sgm];
    for(uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
      dst[d0 + i] = src[s0 + i];
  }
}

__global__ void k_copy_int(const int* __restrict__ src, int* __restrict__ dst) { *dst = *src; }

// ---------------------------------------------------------------------------
// Coarse tail of the batched V-cycle: one workgroup per instance runs its levels
// min(finest, T) .. 1 with every vector in LDS.  A thread owns rows tid and tid + kBt of
// a level (a level has at most FEAT_BATCH_TAIL_MAX_ROWS < 2 kBt rows) and holds their
// entries of A in registers for a whole smoother call, so the sweeps between barriers
// touch LDS only.  The matrices are in ELL form (entry u of row j at [u*n + j]): a row's
// entries load in one coalesced round trip with no row-pointer read in front of them, and
// the loads of R (next to A's at the start of the pre-smoothing) and of P and A (on the
// way up) are issued together.  Each row is formed with the operations of the batched
// launches it stands for: k_smooth_first (first iterate), EpiSmooth modes 0 / 1 / 2,
// EpiMasked (restriction), EpiProlAdd and EpiDefect, with the row sum as the sequential
// fma over the row's entries in CSR order (the order of every SELL-64 plan of those
// launches): bit-identical results.
constexpr int kBt = 1024;
constexpr int kBtRows = 2;  // rows per thread
static_assert(FEAT_BATCH_TAIL_MAX_ROWS <= kBt * kBtRows, "rows per thread");

template<int W>
struct EllRow
{
  double v[W];
  uint32_t c[W];  // column, local to the instance's block of the source vector
  int len;
};

// row j's entries (len = 0 when off); every entry below the level's width is loaded,
// the row length only selects the terms of the sum
template<int W>
__device__ __forceinline__ void ell_load(EllRow<W>& R, const double* __restrict__ val,
                                         const uint16_t* __restrict__ col, const uint8_t* __restrict__ len,
                                         uint32_t n, uint32_t w, uint32_t j, bool on)
{
  R.len = 0;
  if(!on) return;
  R.len = len[j];
#pragma unroll
  for(int u = 0; u < W; ++u)
    if(uint32_t(u) < w)
    {
      R.v[u] = val[uint64_t(u) * n + j];
      R.c[u] = col[uint64_t(u) * n + j];
    }
}

template<int W>
__device__ __forceinline__ double ell_dot(const EllRow<W>& R, const double* x)
{
  double s = 0.0;
#pragma unroll
  for(int u = 0; u < W; ++u)
    if(u < R.len) s = dfma(R.v[u], x[R.c[u]], s);
  return s;
}

__global__ __launch_bounds__(kBt) void k_batch_tail(feat_batch_tail_t Targ, const double* __restrict__ r,
                                                    double* __restrict__ z)
{
  constexpr int WA = FEAT_BATCH_TAIL_MAX_ROW_LEN, WP = FEAT_BATCH_TAIL_MAX_PROL_LEN, WR = FEAT_BATCH_TAIL_MAX_ROW_LEN;
  // the level descriptors are indexed by level at run time: one copy in LDS
  __shared__ feat_batch_tail_t T;
  __shared__ uint64_t s_loff[FEAT_BATCH_TAIL_MAX_LEVELS + 1];
  __shared__ uint32_t s_base[FEAT_BATCH_TAIL_MAX_LEVELS + 1];
  // per level (at s_base[k]): right-hand side, solution, Jacobi inverse diagonal, filter flag
  __shared__ double s_rhs[FEAT_BATCH_TAIL_MAX_TOTAL], s_sol[FEAT_BATCH_TAIL_MAX_TOTAL];
  __shared__ double s_inv[FEAT_BATCH_TAIL_MAX_TOTAL];
  __shared__ uint8_t s_m[FEAT_BATCH_TAIL_MAX_TOTAL];
  // the two smoother buffers and the defect of the level being worked on
  __shared__ double s_b0[FEAT_BATCH_TAIL_MAX_ROWS], s_b1[FEAT_BATCH_TAIL_MAX_ROWS], s_d[FEAT_BATCH_TAIL_MAX_ROWS];
  const uint32_t inst = blockIdx.x;
  const int fin = Targ.finest[inst];
  const int top = fin < Targ.nlevels ? fin : Targ.nlevels;
  const bool deep = fin > Targ.nlevels;
  if(threadIdx.x == 0)
  {
    T = Targ;
    uint32_t o = 0;
    for(int k = 1; k <= top; ++k)
    {
      s_loff[k] = Targ.loff[uint64_t(k - 1) * Targ.ninst + inst];
      s_base[k] = o;
      o += Targ.lv[k - 1].n;
    }
  }
  __syncthreads();
  const int tid = threadIdx.x;
  const double omega = T.omega;
  for(int k = 1; k <= top; ++k)
  {
    const feat_batch_tail_level_t& L = T.lv[k - 1];
    const uint64_t g0 = s_loff[k];
    const uint32_t b = s_base[k];
    for(uint32_t j = tid; j < L.n; j += kBt)
    {
      s_inv[b + j] = L.inv[g0 + j];
      s_m[b + j] = is_masked(L.mask, g0 + j) ? 1 : 0;
    }
  }
  {
    const uint32_t n = T.lv[top - 1].n, b = s_base[top];
    for(uint32_t j = tid; j < n; j += kBt)
      s_rhs[b + j] = deep ? T.rhs_deep[s_loff[top] + j] : r[T.pcg_off[inst] + j];
  }
  __syncthreads();

  // A's rows of this thread on level k, in registers
  EllRow<WA> A[kBtRows];
  auto load_a = [&](int k) {
    const feat_batch_tail_level_t& L = T.lv[k - 1];
    const double* val = L.a_val + s_loff[k] * L.wa;
#pragma unroll
    for(int q = 0; q < kBtRows; ++q)
    {
      const uint32_t j = tid + q * kBt;
      ell_load(A[q], val, L.a_col, L.a_len, L.n, L.wa, j, j < L.n);
    }
  };

  // Richardson(Jacobi) pre-smoothing from x = 0 (feat_rj_smooth_f64 with x = sol, tmp and
  // d_out): iterates alternate between sol and s_b1 so that the last one lands in sol;
  // the defect of the last iterate goes to s_d; then the rhs of level k-1 = filter(R s_d)
  auto pre_restrict = [&](int k) {
    const feat_batch_tail_level_t& L = T.lv[k - 1];
    const uint32_t n = L.n, b = s_base[k];
    const uint32_t nc = T.lv[k - 2].n, bc = s_base[k - 1];
    const double* rhs = s_rhs + b;
    const double* inv = s_inv + b;
    const uint8_t* m = s_m + b;
    double* x = s_sol + b;
    const int nu = T.nu_pre;
    auto buf = [&](int s) { return ((nu - s) & 1) ? s_b1 : x; };
    load_a(k);
    EllRow<WR> R;  // level k-1 has at most kBt rows
    ell_load(R, L.r_val, L.r_col, L.r_len, nc, L.wr, tid, uint32_t(tid) < nc);
    {
      double* x1 = buf(1);
#pragma unroll
      for(int q = 0; q < kBtRows; ++q)
      {
        const uint32_t j = tid + q * kBt;
        if(j < n)
        {
          double c = inv[j] * rhs[j];
          if(m[j]) c = 0.0;
          x1[j] = dfma(omega, c, 0.0);
        }
      }
    }
    for(int s = 1; s <= nu; ++s)
    {
      __syncthreads();
      const double* xk = buf(s);
      double* xo = buf(s + 1);
#pragma unroll
      for(int q = 0; q < kBtRows; ++q)
      {
        const uint32_t j = tid + q * kBt;
        if(j < n)
        {
          const double sum = ell_dot(A[q], xk);
          double d = dfma(sum, -1.0, rhs[j]);
          if(m[j]) d = 0.0;
          if(s == nu)
            s_d[j] = d;
          else
          {
            double c = inv[j] * d;
            if(m[j]) c = 0.0;
            xo[j] = dfma(omega, c, xk[j]);
          }
        }
      }
    }
    __syncthreads();
    if(uint32_t(tid) < nc)
    {
      double v = dfma(ell_dot(R, s_d), 1.0, 0.0);
      if(s_m[bc + tid]) v = 0.0;
      s_rhs[bc + tid] = v;
    }
    __syncthreads();
  };

  // sol of level k += filter(P sol of level k-1), then s_d = filter(rhs - A sol)
  auto prol_defect = [&](int k) {
    const feat_batch_tail_level_t& L = T.lv[k - 1];
    const uint32_t n = L.n, b = s_base[k];
    const double* xc = s_sol + s_base[k - 1];
    double* x = s_sol + b;
    const uint8_t* m = s_m + b;
    EllRow<WP> P[kBtRows];
#pragma unroll
    for(int q = 0; q < kBtRows; ++q)
    {
      const uint32_t j = tid + q * kBt;
      ell_load(P[q], L.p_val, L.p_col, L.p_len, n, L.wp, j, j < n);
    }
    load_a(k);
#pragma unroll
    for(int q = 0; q < kBtRows; ++q)
    {
      const uint32_t j = tid + q * kBt;
      if(j < n)
      {
        double c = dfma(ell_dot(P[q], xc), 1.0, 0.0);
        if(m[j]) c = 0.0;
        x[j] = dfma(1.0, c, x[j]);
      }
    }
    __syncthreads();
#pragma unroll
    for(int q = 0; q < kBtRows; ++q)
    {
      const uint32_t j = tid + q * kBt;
      if(j < n)
      {
        double d = dfma(ell_dot(A[q], x), -1.0, s_rhs[b + j]);
        if(m[j]) d = 0.0;
        s_d[j] = d;
      }
    }
    __syncthreads();
  };

  // Richardson(Jacobi) post-smoothing on s_d from 0, its last iterate added to sol
  // (feat_rj_smooth_f64 with x = cor, tmp and acc = sol); A's rows are still loaded
  auto post = [&](int k) {
    const uint32_t n = T.lv[k - 1].n, b = s_base[k];
    const double* inv = s_inv + b;
    const uint8_t* m = s_m + b;
    double* x = s_sol + b;
    const int nu = T.nu_post;
    auto buf = [&](int s) { return ((nu - s) & 1) ? s_b1 : s_b0; };
    {
      double* x1 = buf(1);
#pragma unroll
      for(int q = 0; q < kBtRows; ++q)
      {
        const uint32_t j = tid + q * kBt;
        if(j < n)
        {
          double c = inv[j] * s_d[j];
          if(m[j]) c = 0.0;
          const double xn = dfma(omega, c, 0.0);
          if(nu == 1)
            x[j] = dfma(1.0, xn, x[j]);
          else
            x1[j] = xn;
        }
      }
    }
    for(int s = 1; s <= nu - 1; ++s)
    {
      __syncthreads();
      const double* xk = buf(s);
      double* xo = buf(s + 1);
#pragma unroll
      for(int q = 0; q < kBtRows; ++q)
      {
        const uint32_t j = tid + q * kBt;
        if(j < n)
        {
          double d = dfma(ell_dot(A[q], xk), -1.0, s_d[j]);
          if(m[j]) d = 0.0;
          double c = inv[j] * d;
          if(m[j]) c = 0.0;
          const double xn = dfma(omega, c, xk[j]);
          if(s == nu - 1)
            x[j] = dfma(1.0, xn, x[j]);
          else
            xo[j] = xn;
        }
      }
    }
    __syncthreads();
  };

  for(int k = top; k >= 2; --k) pre_restrict(k);
  {
    // level 1: one Jacobi step with omega 1 from x = 0
    const uint32_t n = T.lv[0].n, b = s_base[1];
    for(uint32_t j = tid; j < n; j += kBt)
    {
      double c = s_inv[b + j] * s_rhs[b + j];
      if(s_m[b + j]) c = 0.0;
      s_sol[b + j] = dfma(1.0, c, 0.0);
    }
    __syncthreads();
  }
  for(int k = 2; k <= top; ++k)
  {
    prol_defect(k);
    post(k);
  }
  {
    const uint32_t n = T.lv[top - 1].n, b = s_base[top];
    for(uint32_t j = tid; j < n; j += kBt)
    {
      if(deep)
        T.sol_deep[s_loff[top] + j] = s_sol[b + j];
      else
        z[T.pcg_off[inst] + j] = s_sol[b + j];
    }
  }
}

// ---------------------------------------------------------------------------
// Grid form of the coarse tail: every level a structured n1 x n1 grid per instance, A as
// the nine stencil values of the free rows, P bilinear injection and R its transpose
// (formed from the node positions).  Every matrix value is read once at kernel start:
// the top level's rows into registers, the levels below into LDS, so the V-cycle itself
// touches LDS only.  A filtered row never uses its row sum (its defect and correction
// are zeroed by every kernel it stands for), so only free (interior) rows are summed;
// each sum is the sequential fma over the row's entries in CSR order, as everywhere.
__device__ __forceinline__ double grid9(const double* v, size_t stride, const double* x, int j, int n1)
{
  double s = 0.0;
  s = dfma(v[0 * stride], x[j - n1 - 1], s);
  s = dfma(v[1 * stride], x[j - n1], s);
  s = dfma(v[2 * stride], x[j - n1 + 1], s);
  s = dfma(v[3 * stride], x[j - 1], s);
  s = dfma(v[4 * stride], x[j], s);
  s = dfma(v[5 * stride], x[j + 1], s);
  s = dfma(v[6 * stride], x[j + n1 - 1], s);
  s = dfma(v[7 * stride], x[j + n1], s);
  s = dfma(v[8 * stride], x[j + n1 + 1], s);
  return s;
}

// P row (x, y) of the fine grid (n1 per side) over the coarse grid xc: bilinear injection
__device__ __forceinline__ double grid_prol(const double* xc, int j, int n1)
{
  const int nc1 = (n1 + 1) / 2, gx = j % n1, gy = j / n1;
  const double* xr = xc + (gy >> 1) * nc1 + (gx >> 1);
  const bool ox = gx & 1, oy = gy & 1;
  if(!ox && !oy) return dfma(1.0, xr[0], 0.0);
  if(!oy) return dfma(0.5, xr[1], dfma(0.5, xr[0], 0.0));
  if(!ox) return dfma(0.5, xr[nc1], dfma(0.5, xr[0], 0.0));
  return dfma(0.25, xr[nc1 + 1], dfma(0.25, xr[nc1], dfma(0.25, xr[1], dfma(0.25, xr[0], 0.0))));
}

// R row (X, Y) of the coarse grid (nc1 per side) over the fine grid d (n1 = 2 nc1 - 1):
// the fine nodes (2X + dx, 2Y + dy) inside the grid, ascending, weights w(dx) w(dy)
__device__ __forceinline__ double grid_rest(const double* d, int jc, int n1)
{
  const int nc1 = (n1 + 1) / 2, X = jc % nc1, Y = jc / nc1;
  double s = 0.0;
#pragma unroll
  for(int dy = -1; dy <= 1; ++dy)
#pragma unroll
    for(int dx = -1; dx <= 1; ++dx)
    {
      const int fx = 2 * X + dx, fy = 2 * Y + dy;
      if(fx >= 0 && fx < n1 && fy >= 0 && fy < n1)
        s = dfma((dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0), d[fy * n1 + fx], s);
    }
  return s;
}

__global__ __launch_bounds__(kBt) void k_batch_tail_grid(feat_batch_tail_t Targ, const double* __restrict__ r,
                                                         double* __restrict__ z)
{
  __shared__ feat_batch_tail_t T;
  __shared__ uint64_t s_loff[FEAT_BATCH_TAIL_MAX_LEVELS + 1];
  __shared__ uint32_t s_base[FEAT_BATCH_TAIL_MAX_LEVELS + 1];
  __shared__ double s_rhs[FEAT_BATCH_TAIL_MAX_TOTAL], s_sol[FEAT_BATCH_TAIL_MAX_TOTAL];
  __shared__ double s_inv[FEAT_BATCH_TAIL_MAX_TOTAL];
  __shared__ uint8_t s_m[FEAT_BATCH_TAIL_MAX_TOTAL];
  __shared__ double s_b0[FEAT_BATCH_TAIL_MAX_ROWS], s_b1[FEAT_BATCH_TAIL_MAX_ROWS], s_d[FEAT_BATCH_TAIL_MAX_ROWS];
  // stencil values of the levels below the top: level k's value u of row j at 9 s_base[k] + u n + j
  __shared__ double s_a[9 * FEAT_BATCH_TAIL_MAX_BELOW];
  const uint32_t inst = blockIdx.x;
  const int fin = Targ.finest[inst];
  const int top = fin < Targ.nlevels ? fin : Targ.nlevels;
  const bool deep = fin > Targ.nlevels;
  if(threadIdx.x == 0)
  {
    T = Targ;
    uint32_t o = 0;
    for(int k = 1; k <= top; ++k)
    {
      s_loff[k] = Targ.loff[uint64_t(k - 1) * Targ.ninst + inst];
      s_base[k] = o;
      o += Targ.lv[k - 1].n;
    }
  }
  __syncthreads();
  const int tid = threadIdx.x;
  const double omega = T.omega;
  // every read of the matrices, the inverse diagonals, the filters and the rhs, issued together
  double at[kBtRows][9];
  {
    const feat_batch_tail_level_t& L = T.lv[top - 1];
    const double* sv = L.sval + s_loff[top] * 9;
#pragma unroll
    for(int q = 0; q < kBtRows; ++q)
    {
      const uint32_t j = tid + q * kBt;
#pragma unroll
      for(int u = 0; u < 9; ++u) at[q][u] = j < L.n ? sv[uint64_t(u) * L.n + j] : 0.0;
    }
  }
  for(int k = 1; k <= top; ++k)
  {
    const feat_batch_tail_level_t& L = T.lv[k - 1];
    const uint64_t g0 = s_loff[k];
    const uint32_t b = s_base[k];
    for(uint32_t j = tid; j < L.n; j += kBt)
    {
      s_inv[b + j] = L.inv[g0 + j];
      s_m[b + j] = is_masked(L.mask, g0 + j) ? 1 : 0;
    }
    if(k < top)
    {
      const double* sv = L.sval + g0 * 9;
      for(uint32_t e = tid; e < 9 * L.n; e += kBt) s_a[9 * b + e] = sv[e];
    }
  }
  {
    const uint32_t n = T.lv[top - 1].n, b = s_base[top];
    for(uint32_t j = tid; j < n; j += kBt)
      s_rhs[b + j] = deep ? T.rhs_deep[s_loff[top] + j] : r[T.pcg_off[inst] + j];
  }
  __syncthreads();

  // row sum of A on level k (row j = tid + q kBt), free rows only
  auto asum = [&](int k, int q, const double* x) -> double {
    const feat_batch_tail_level_t& L = T.lv[k - 1];
    const int j = tid + q * kBt, n1 = int(L.n1);
    if(k == top) return grid9(at[q], 1, x, j, n1);
    return grid9(s_a + 9 * s_base[k] + j, L.n, x, j, n1);
  };

  // pre-smoothing from 0 with the defect of the last iterate into s_d, then the restriction
  auto pre_restrict = [&](int k) {
    const feat_batch_tail_level_t& L = T.lv[k - 1];
    const uint32_t n = L.n, b = s_base[k];
    const uint32_t nc = T.lv[k - 2].n, bc = s_base[k - 1];
    const double* rhs = s_rhs + b;
    const double* inv = s_inv + b;
    const uint8_t* m = s_m + b;
    double* x = s_sol + b;
    const int nu = T.nu_pre;
    auto buf = [&](int s) { return ((nu - s) & 1) ? s_b1 : x; };
    {
      double* x1 = buf(1);
#pragma unroll
      for(int q = 0; q < kBtRows; ++q)
      {
        const uint32_t j = tid + q * kBt;
        if(j < n)
        {
          double c = inv[j] * rhs[j];
          if(m[j]) c = 0.0;
          x1[j] = dfma(omega, c, 0.0);
        }
      }
    }
    for(int s = 1; s <= nu; ++s)
    {
      __syncthreads();
      const double* xk = buf(s);
      double* xo = buf(s + 1);
#pragma unroll
      for(int q = 0; q < kBtRows; ++q)
      {
        const uint32_t j = tid + q * kBt;
        if(j < n)
        {
          double d = 0.0;
          if(!m[j])
            d = dfma(asum(k, q, xk), -1.0, rhs[j]);
          if(s == nu)
            s_d[j] = d;
          else
          {
            double c = inv[j] * d;
            if(m[j]) c = 0.0;
            xo[j] = dfma(omega, c, xk[j]);
          }
        }
      }
    }
    __syncthreads();
    if(uint32_t(tid) < nc)
    {
      double v = 0.0;
      if(!s_m[bc + tid]) v = dfma(grid_rest(s_d, tid, int(L.n1)), 1.0, 0.0);
      s_rhs[bc + tid] = v;
    }
    __syncthreads();
  };

  // sol += filter(P sol of level k-1), then s_d = filter(rhs - A sol)
  auto prol_defect = [&](int k) {
    const feat_batch_tail_level_t& L = T.lv[k - 1];
    const uint32_t n = L.n, b = s_base[k];
    const double* xc = s_sol + s_base[k - 1];
    double* x = s_sol + b;
    const uint8_t* m = s_m + b;
#pragma unroll
    for(int q = 0; q < kBtRows; ++q)
    {
      const uint32_t j = tid + q * kBt;
      if(j < n)
      {
        double c = 0.0;
        if(!m[j]) c = dfma(grid_prol(xc, int(j), int(L.n1)), 1.0, 0.0);
        x[j] = dfma(1.0, c, x[j]);
      }
    }
    __syncthreads();
#pragma unroll
    for(int q = 0; q < kBtRows; ++q)
    {
      const uint32_t j = tid + q * kBt;
      if(j < n)
      {
        double d = 0.0;
        if(!m[j]) d = dfma(asum(k, q, x), -1.0, s_rhs[b + j]);
        s_d[j] = d;
      }
    }
    __syncthreads();
  };

  // post-smoothing on s_d from 0, its last iterate added to sol
  auto post = [&](int k) {
    const uint32_t n = T.lv[k - 1].n, b = s_base[k];
    const double* inv = s_inv + b;
    const uint8_t* m = s_m + b;
    double* x = s_sol + b;
    const int nu = T.nu_post;
    auto buf = [&](int s) { return ((nu - s) & 1) ? s_b1 : s_b0; };
    {
      double* x1 = buf(1);
#pragma unroll
      for(int q = 0; q < kBtRows; ++q)
      {
        const uint32_t j = tid + q * kBt;
        if(j < n)
        {
          double c = inv[j] * s_d[j];
          if(m[j]) c = 0.0;
          const double xn = dfma(omega, c, 0.0);
          if(nu == 1)
            x[j] = dfma(1.0, xn, x[j]);
          else
            x1[j] = xn;
        }
      }
    }
    for(int s = 1; s <= nu - 1; ++s)
    {
      __syncthreads();
      const double* xk = buf(s);
      double* xo = buf(s + 1);
#pragma unroll
      for(int q = 0; q < kBtRows; ++q)
      {
        const uint32_t j = tid + q * kBt;
        if(j < n)
        {
          double c = 0.0;
          if(!m[j]) c = inv[j] * dfma(asum(k, q, xk), -1.0, s_d[j]);
          const double xn = dfma(omega, c, xk[j]);
          if(s == nu - 1)
            x[j] = dfma(1.0, xn, x[j]);
          else
            xo[j] = xn;
        }
      }
    }
    __syncthreads();
  };

  for(int k = top; k >= 2; --k) pre_restrict(k);
  {
    const uint32_t n = T.lv[0].n, b = s_base[1];
    for(uint32_t j = tid; j < n; j += kBt)
    {
      double c = s_inv[b + j] * s_rhs[b + j];
      if(s_m[b + j]) c = 0.0;
      s_sol[b + j] = dfma(1.0, c, 0.0);
    }
    __syncthreads();
  }
  for(int k = 2; k <= top; ++k)
  {
    prol_defect(k);
    post(k);
  }
  {
    const uint32_t n = T.lv[top - 1].n, b = s_base[top];
    for(uint32_t j = tid; j < n; j += kBt)
    {
      if(deep)
        T.sol_deep[s_loff[top] + j] = s_sol[b + j];
      else
        z[T.pcg_off[inst] + j] = s_sol[b + j];
    }
  }
}

// ---------------------------------------------------------------------------
// Tiled smoother calls of a batched level on structured grids.  A workgroup owns a
// kTlTile x kTlTile tile of one instance's n1 x n1 grid and works on the tile plus a halo of
// kTlHalo nodes (clipped to the grid), kTlPts region nodes per thread with their nine
// stencil values, right-hand side, inverse diagonal and filter flag in registers and the
// iterates in LDS.  A sweep over the region is exact at the nodes at least s nodes away
// from a cut (non-grid) edge of the region after s stencil applications, so with at most
// kTlHalo applications the tile's nodes are exact; nodes nearer a cut edge are computed
// with a zero row sum and never reach the tile.  Filtered rows never use their row sum
// (every kernel they stand for sets their defect / correction to zero), and a free row's
// sum is the sequential fma over its nine neighbours in CSR (lexicographic) order, so the
// results are bit-identical to the SELL-64 launches of the same calls.
constexpr int kTl = 512;
constexpr int kTlTile = 32;
constexpr int kTlPts = 4;   // region nodes per thread (42 x 42 <= 4 * 512)

// REST (pre only): the restriction of the defect to the next coarser level as well, for the
// coarse nodes whose fine centre lies in the tile; the defect of the tile's neighbours is then
// needed exact, so the halo is one node wider
template<bool POST, bool REST = false>
__global__ __launch_bounds__(kTl) void k_batch_tiled(feat_batch_tiled_t t)
{
  constexpr int kTlHalo = FEAT_BATCH_TILED_MAX_STEPS + (REST ? 1 : 0);
  constexpr int kTlReg = kTlTile + 2 * kTlHalo;
  static_assert(kTlReg * kTlReg <= kTlPts * kTl, "region nodes per thread");
  __shared__ double s_b[POST ? 3 : 2][kTlReg * kTlReg];
  const int n1 = int(t.n1);
  const int tiles = (n1 + kTlTile - 1) / kTlTile;
  const uint32_t ntl = uint32_t(tiles) * uint32_t(tiles);
  const uint32_t p = blockIdx.x / ntl, tl = blockIdx.x % ntl;
  const int tx0 = int(tl % uint32_t(tiles)) * kTlTile, ty0 = int(tl / uint32_t(tiles)) * kTlTile;
  const int rx0 = max(tx0 - kTlHalo, 0), ry0 = max(ty0 - kTlHalo, 0);
  const int rw = min(tx0 + kTlTile + kTlHalo, n1) - rx0, rh = min(ty0 + kTlTile + kTlHalo, n1) - ry0;
  const uint64_t nk = uint64_t(n1) * uint64_t(n1), base = uint64_t(p) * nk;
  const int tid = threadIdx.x;
  const double omega = t.omega;
  const int nu = t.nu;

  // an instance entering the V-cycle here reads its right-hand side from r (and post
  // leaves its solution in z too)
  const uint64_t top = t.top ? t.top[p] : ~uint64_t(0);
  auto rhs_at = [&](uint32_t rw_) { return top != ~uint64_t(0) ? t.r[top + rw_] : t.rhs[base + rw_]; };
  double a[kTlPts][9], rv[kTlPts], iv[kTlPts], dq[kTlPts], xs[kTlPts];
  int li[kTlPts];
  uint32_t row[kTlPts];
  bool on[kTlPts], m[kTlPts], in[kTlPts], out[kTlPts];
#pragma unroll
  for(int q = 0; q < kTlPts; ++q)
  {
    const int idx = tid + q * kTl;
    on[q] = idx < rw * rh;
    const int lx = on[q] ? idx % rw : 0, ly = on[q] ? idx / rw : 0;
    const int gx = rx0 + lx, gy = ry0 + ly;
    li[q] = idx;
    row[q] = uint32_t(gy) * uint32_t(n1) + uint32_t(gx);
    in[q] = lx > 0 && lx < rw - 1 && ly > 0 && ly < rh - 1;
    out[q] = on[q] && gx >= tx0 && gx < tx0 + kTlTile && gy >= ty0 && gy < ty0 + kTlTile;
    m[q] = true;
    rv[q] = iv[q] = dq[q] = xs[q] = 0.0;
    if(on[q])
    {
      const uint64_t g = base + row[q];
      m[q] = is_masked(t.mask, g);
      if(!POST) rv[q] = rhs_at(row[q]);
      iv[q] = t.inv[g];
      if(POST) xs[q] = t.sol[g];
    }
  }
  // the stencil values (after the prolongation's loads when there is one: fewer live registers)
  auto load_stencil = [&]() {
#pragma unroll
    for(int q = 0; q < kTlPts; ++q)
#pragma unroll
      for(int u = 0; u < 9; ++u)
        a[q][u] = (on[q] && !m[q] && in[q]) ? t.sval[(uint64_t(p) * 9 + u) * nk + row[q]] : 0.0;
  };
  auto stencil = [&](int q, const double* x) -> double {
    double s = 0.0;
    if(in[q] && !m[q])
    {
      const int c = li[q];
      s = dfma(a[q][0], x[c - rw - 1], s);
      s = dfma(a[q][1], x[c - rw], s);
      s = dfma(a[q][2], x[c - rw + 1], s);
      s = dfma(a[q][3], x[c - 1], s);
      s = dfma(a[q][4], x[c], s);
      s = dfma(a[q][5], x[c + 1], s);
      s = dfma(a[q][6], x[c + rw - 1], s);
      s = dfma(a[q][7], x[c + rw], s);
      s = dfma(a[q][8], x[c + rw + 1], s);
    }
    return s;
  };
  auto buf = [&](int s) { return ((nu - s) & 1) ? s_b[1] : s_b[0]; };

  if constexpr(!POST)
  {
    load_stencil();
    // first iterate from x = 0 (k_smooth_first)
    double* x1 = buf(1);
#pragma unroll
    for(int q = 0; q < kTlPts; ++q)
      if(on[q])
      {
        double c = iv[q] * rv[q];
        if(m[q]) c = 0.0;
        x1[li[q]] = dfma(omega, c, 0.0);
      }
    for(int s = 1; s <= nu; ++s)
    {
      __syncthreads();
      const double* xk = buf(s);
      double* xo = buf(s + 1);
#pragma unroll
      for(int q = 0; q < kTlPts; ++q)
        if(on[q])
        {
          double d = dfma(stencil(q, xk), -1.0, rv[q]);
          if(m[q]) d = 0.0;
          if(s == nu)
            dq[q] = d;
          else
          {
            double c = iv[q] * d;
            if(m[q]) c = 0.0;
            xo[li[q]] = dfma(omega, c, xk[li[q]]);
          }
        }
    }
    const double* xl = buf(nu);
#pragma unroll
    for(int q = 0; q < kTlPts; ++q)
      if(out[q])
      {
        t.sol[base + row[q]] = xl[li[q]];
        if(!REST) t.def[base + row[q]] = dq[q];
      }
    if constexpr(REST)
    {
      // the defect in the iterate buffer no longer read, then rhs_c = filter(R def) for the
      // coarse nodes (X, Y) with (2X, 2Y) in the tile: the fine nodes around (2X, 2Y) inside
      // the grid in ascending order, weights 1/2 per nonzero offset (EpiMasked)
      double* dl = buf(nu + 1);
#pragma unroll
      for(int q = 0; q < kTlPts; ++q)
        if(on[q]) dl[li[q]] = dq[q];
      __syncthreads();
      const int nc1 = (n1 + 1) / 2;
      const int X0 = tx0 / 2, Y0 = ty0 / 2;
      const int cw = (min(tx0 + kTlTile, n1) - 1) / 2 - X0 + 1, chh = (min(ty0 + kTlTile, n1) - 1) / 2 - Y0 + 1;
      for(int e = tid; e < cw * chh; e += kTl)
      {
        const int X = X0 + e % cw, Y = Y0 + e / cw;
        double sum = 0.0;
#pragma unroll
        for(int dy = -1; dy <= 1; ++dy)
#pragma unroll
          for(int dx = -1; dx <= 1; ++dx)
          {
            const int gx = 2 * X + dx, gy = 2 * Y + dy;
            if(gx >= 0 && gx < n1 && gy >= 0 && gy < n1)
              sum = dfma((dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0), dl[(gy - ry0) * rw + (gx - rx0)], sum);
          }
        const uint64_t gc = t.coff[p] + uint64_t(Y) * nc1 + X;
        double v = dfma(sum, 1.0, 0.0);
        if(is_masked(t.cmask, gc)) v = 0.0;
        t.crhs[gc] = v;
      }
    }
  }
  else
  {
    // sol += filter(P xc)
    const double* xc = t.xc + t.coff[p];
    if(t.p_struct)
    {
      // bilinear injection: the row's entries in CSR order, from the node position
      const int nc1 = (n1 + 1) / 2;
      load_stencil();
#pragma unroll
      for(int q = 0; q < kTlPts; ++q)
        if(on[q])
        {
          const int gx = int(row[q] % uint32_t(n1)), gy = int(row[q] / uint32_t(n1));
          const double* xr = xc + (gy >> 1) * nc1 + (gx >> 1);
          const bool ox = gx & 1, oy = gy & 1;
          double sum;
          if(!ox && !oy)
            sum = dfma(1.0, xr[0], 0.0);
          else if(!oy)
            sum = dfma(0.5, xr[1], dfma(0.5, xr[0], 0.0));
          else if(!ox)
            sum = dfma(0.5, xr[nc1], dfma(0.5, xr[0], 0.0));
          else
            sum = dfma(0.25, xr[nc1 + 1], dfma(0.25, xr[nc1], dfma(0.25, xr[1], dfma(0.25, xr[0], 0.0))));
          double c = dfma(sum, 1.0, 0.0);
          if(m[q]) c = 0.0;
          xs[q] = dfma(1.0, c, xs[q]);
          s_b[2][li[q]] = xs[q];
        }
    }
    else
    {
#pragma unroll
    for(int q = 0; q < kTlPts; ++q)
      if(on[q])
      {
        const uint32_t j = row[q];
        const int len = t.p_len[j];
        double pv[FEAT_BATCH_TAIL_MAX_PROL_LEN];
        uint32_t pc[FEAT_BATCH_TAIL_MAX_PROL_LEN];
#pragma unroll
        for(int u = 0; u < FEAT_BATCH_TAIL_MAX_PROL_LEN; ++u)
          if(uint32_t(u) < t.wp)
          {
            pv[u] = t.p_val[uint64_t(u) * nk + j];
            pc[u] = t.p_col[uint64_t(u) * nk + j];
          }
        double sum = 0.0;
#pragma unroll
        for(int u = 0; u < FEAT_BATCH_TAIL_MAX_PROL_LEN; ++u)
          if(u < len) sum = dfma(pv[u], xc[pc[u]], sum);
        double c = dfma(sum, 1.0, 0.0);
        if(m[q]) c = 0.0;
        xs[q] = dfma(1.0, c, xs[q]);
        s_b[2][li[q]] = xs[q];
      }
    load_stencil();
    }
    __syncthreads();
    // defect, then the first iterate on it
    double* y1 = buf(1);
#pragma unroll
    for(int q = 0; q < kTlPts; ++q)
      if(on[q])
      {
        double d = dfma(stencil(q, s_b[2]), -1.0, rhs_at(row[q]));
        if(m[q]) d = 0.0;
        dq[q] = d;
        double c = iv[q] * d;
        if(m[q]) c = 0.0;
        const double xn = dfma(omega, c, 0.0);
        if(nu == 1)
          xs[q] = dfma(1.0, xn, s_b[2][li[q]]);
        else
          y1[li[q]] = xn;
      }
    for(int s = 1; s <= nu - 1; ++s)
    {
      __syncthreads();
      const double* yk = buf(s);
      double* yo = buf(s + 1);
#pragma unroll
      for(int q = 0; q < kTlPts; ++q)
        if(on[q])
        {
          double d = dfma(stencil(q, yk), -1.0, dq[q]);
          if(m[q]) d = 0.0;
          double c = iv[q] * d;
          if(m[q]) c = 0.0;
          const double xn = dfma(omega, c, yk[li[q]]);
          if(s == nu - 1)
            xs[q] = dfma(1.0, xn, s_b[2][li[q]]);
          else
            yo[li[q]] = xn;
        }
    }
#pragma unroll
    for(int q = 0; q < kTlPts; ++q)
      if(out[q])
      {
        t.sol[base + row[q]] = xs[q];
        if(top != ~uint64_t(0)) t.z[top + row[q]] = xs[q];
      }
  }
}

int tiled_launch(feat_handle_t h, const feat_batch_tiled_t* t, bool post, const char* name)
{
  if(!h || !t) return fail(FEAT_ERR_INVALID, name);
  if(t->n1 < 3 || t->n1 > 65535 || t->nu < 1 || t->nu > FEAT_BATCH_TILED_MAX_STEPS)
    return fail(FEAT_ERR_INVALID, "batch tiled smoother: grid size or steps out of range");
  if(!t->ninst) return FEAT_OK;
  if(!t->sval || !t->inv || !t->mask || !t->rhs || !t->sol || (!post && !t->rest && !t->def))
    return fail(FEAT_ERR_INVALID, "batch tiled smoother: null array");
  if(!post && t->rest && (!t->coff || !t->crhs || !t->cmask))
    return fail(FEAT_ERR_INVALID, "batch tiled smoother: restriction target missing");
  if(!post && t->rest && (static_cast<const void*>(t->crhs) == t->sol || t->crhs == t->rhs))
    return fail(FEAT_ERR_ALIAS, "batch tiled smoother: restriction target aliases a level vector");
  if(post && (!t->xc || !t->coff ||
              (!t->p_struct && (!t->p_val || !t->p_col || !t->p_len || t->wp > FEAT_BATCH_TAIL_MAX_PROL_LEN))))
    return fail(FEAT_ERR_INVALID, "batch tiled smoother: prolongation incomplete");
  if(t->top && (!t->r || (post && !t->z)))
    return fail(FEAT_ERR_INVALID, "batch tiled smoother: r / z missing for the entering instances");
  if(t->top && (static_cast<const void*>(t->r) == t->sol || (post && static_cast<const void*>(t->r) == t->z)))
    return fail(FEAT_ERR_ALIAS, "batch tiled smoother: r aliases an output");
  if(static_cast<const void*>(t->rhs) == t->sol || (!post && (t->def == t->sol || static_cast<const void*>(t->rhs) == t->def)) ||
     (post && static_cast<const void*>(t->xc) == t->sol))
    return fail(FEAT_ERR_ALIAS, "batch tiled smoother: vectors alias");
  const uint64_t tiles = (t->n1 + kTlTile - 1) / kTlTile;
  const uint64_t grid = tiles * tiles * t->ninst;
  if(grid > 0x7fffffffull) return fail(FEAT_ERR_INVALID, "batch tiled smoother: grid too large");
  if(post)
    hipLaunchKernelGGL((k_batch_tiled<true>), dim3(uint32_t(grid)), dim3(kTl), 0, h->stream, *t);
  else if(t->rest)
    hipLaunchKernelGGL((k_batch_tiled<false, true>), dim3(uint32_t(grid)), dim3(kTl), 0, h->stream, *t);
  else
    hipLaunchKernelGGL((k_batch_tiled<false>), dim3(uint32_t(grid)), dim3(kTl), 0, h->stream, *t);
  return check_launch(name);
}

bool batch_ok(const feat_batch_t* b)
{
  return b && b->off && b->chunk_inst && b->chunk_lo && b->chunk_hi && b->inst_chunk && b->partials;
}

int finish(feat_handle_t h, const feat_batch_t* b, double* out, const int* active, const char* name)
{
  hipLaunchKernelGGL(k_batch_finish, dim3((b->ninst + 63) / 64), dim3(64), 0, h->stream, to_dev(b), out, active);
  return check_launch(name);
}

} // namespace

extern "C" {

int feat_batch_dot_f64(feat_handle_t h, const feat_batch_t* b, const double* x, const double* y, double* out)
{
  if(!h || !batch_ok(b) || !x || !y || !out) return fail(FEAT_ERR_INVALID, "feat_batch_dot_f64: null argument");
  if(!b->nchunks) return FEAT_OK;
  hipLaunchKernelGGL(k_batch_dot, dim3(b->nchunks), dim3(kThreads), 0, h->stream, to_dev(b), x, y, (double*)nullptr);
  int rc = check_launch("feat_batch_dot_f64");
  return rc ? rc : finish(h, b, out, nullptr, "feat_batch_dot_f64");
}

int feat_batch_pcg_update_f64(feat_handle_t h, const feat_batch_t* b, double* x, double* r, const double* p,
                              const double* t, const double* gamma, const double* tp, const int* active, double* rr)
{
  if(!h || !batch_ok(b) || !x || !r || !p || !t || !gamma || !tp || !active || !rr)
    return fail(FEAT_ERR_INVALID, "feat_batch_pcg_update_f64: null argument");
  if(!b->nchunks) return FEAT_OK;
  hipLaunchKernelGGL(k_batch_update, dim3(b->nchunks), dim3(kThreads), 0, h->stream, to_dev(b), x, r, p, t, gamma, tp,
                     active, (const double*)nullptr, (const double*)nullptr, (double*)nullptr);
  int rc = check_launch("feat_batch_pcg_update_f64");
  return rc ? rc : finish(h, b, rr, active, "feat_batch_pcg_update_f64");
}

int feat_batch_pcg_direction_f64(feat_handle_t h, const feat_batch_t* b, double* p, const double* z,
                                 const double* gamma_new, const double* gamma_old, const int* active, int first)
{
  if(!h || !batch_ok(b) || !p || !z || !active || (!first && (!gamma_new || !gamma_old)))
    return fail(FEAT_ERR_INVALID, "feat_batch_pcg_direction_f64: null argument");
  if(!b->nchunks) return FEAT_OK;
  hipLaunchKernelGGL(k_batch_direction, dim3(b->nchunks), dim3(kThreads), 0, h->stream, to_dev(b), p, z, gamma_new,
                     gamma_old, active, first, (const double*)nullptr);
  return check_launch("feat_batch_pcg_direction_f64");
}

int feat_batch_copy_scalars_f64(feat_handle_t h, uint32_t ninst, double* dst, const double* src, const int* active)
{
  if(!h || (ninst && (!dst || !src))) return fail(FEAT_ERR_INVALID, "feat_batch_copy_scalars_f64: null argument");
  if(!ninst) return FEAT_OK;
  hipLaunchKernelGGL(k_batch_copy_scalars, dim3((ninst + 63) / 64), dim3(64), 0, h->stream, ninst, dst, src, active);
  return check_launch("feat_batch_copy_scalars_f64");
}

int feat_batch_status_f64(feat_handle_t h, uint32_t ninst, const double* rr, const feat_batch_stop_t* stop,
                          const feat_batch_state_t* st, int initial)
{
  if(!h || !rr || !stop || !st || !st->def_init || !st->def_cur || !st->iters || !st->status || !st->active ||
     !st->nactive)
    return fail(FEAT_ERR_INVALID, "feat_batch_status_f64: null argument");
  StopDev c{stop->tol_rel, stop->tol_abs, stop->div_rel, stop->div_abs, stop->min_iter, stop->max_iter};
  hipLaunchKernelGGL(k_batch_status, dim3(1), dim3(kThreads), 0, h->stream, ninst, rr, c, *st, initial, BatchDev{},
                     (const double*)nullptr, (double*)nullptr, (const double*)nullptr);
  return check_launch("feat_batch_status_f64");
}

int feat_batch_dot_partials_f64(feat_handle_t h, const feat_batch_t* b, const double* x, const double* y, double* part)
{
  if(!h || !batch_ok(b) || !x || !y || !part) return fail(FEAT_ERR_INVALID, "feat_batch_dot_partials_f64: null argument");
  if(!b->nchunks) return FEAT_OK;
  hipLaunchKernelGGL(k_batch_dot, dim3(b->nchunks), dim3(kThreads), 0, h->stream, to_dev(b), x, y, part);
  return check_launch("feat_batch_dot_partials_f64");
}

int feat_batch_pcg_update_partials_f64(feat_handle_t h, const feat_batch_t* b, double* x, double* r, const double* p,
                                       const double* t, const double* g_part, const double* tp_part,
                                       const int* active, double* rr_part)
{
  if(!h || !batch_ok(b) || !x || !r || !p || !t || !g_part || !tp_part || !active || !rr_part)
    return fail(FEAT_ERR_INVALID, "feat_batch_pcg_update_partials_f64: null argument");
  if(rr_part == g_part || rr_part == tp_part)
    return fail(FEAT_ERR_ALIAS, "feat_batch_pcg_update_partials_f64: rr_part aliases an input");
  if(!b->nchunks) return FEAT_OK;
  hipLaunchKernelGGL(k_batch_update, dim3(b->nchunks), dim3(kThreads), 0, h->stream, to_dev(b), x, r, p, t,
                     (const double*)nullptr, (const double*)nullptr, active, g_part, tp_part, rr_part);
  return check_launch("feat_batch_pcg_update_partials_f64");
}

int feat_batch_status_partials_f64(feat_handle_t h, const feat_batch_t* b, const double* rr_part,
                                   const feat_batch_stop_t* stop, const feat_batch_state_t* st, double* gamma,
                                   const double* g_part)
{
  if(!h || !batch_ok(b) || !rr_part || !stop || !st || !st->def_init || !st->def_cur || !st->iters || !st->status ||
     !st->active || !st->nactive || (gamma && !g_part))
    return fail(FEAT_ERR_INVALID, "feat_batch_status_partials_f64: null argument");
  StopDev c{stop->tol_rel, stop->tol_abs, stop->div_rel, stop->div_abs, stop->min_iter, stop->max_iter};
  hipLaunchKernelGGL(k_batch_status, dim3(1), dim3(kThreads), 0, h->stream, b->ninst, (const double*)nullptr, c, *st,
                     0, to_dev(b), rr_part, gamma, g_part);
  return check_launch("feat_batch_status_partials_f64");
}

int feat_batch_pcg_direction_partials_f64(feat_handle_t h, const feat_batch_t* b, double* p, const double* z,
                                          const double* g_part, const double* gamma_old, const int* active)
{
  if(!h || !batch_ok(b) || !p || !z || !g_part || !gamma_old || !active)
    return fail(FEAT_ERR_INVALID, "feat_batch_pcg_direction_partials_f64: null argument");
  if(!b->nchunks) return FEAT_OK;
  hipLaunchKernelGGL(k_batch_direction, dim3(b->nchunks), dim3(kThreads), 0, h->stream, to_dev(b), p, z,
                     (const double*)nullptr, gamma_old, active, 0, g_part);
  return check_launch("feat_batch_pcg_direction_partials_f64");
}

int feat_batch_gather_f64(feat_handle_t h, uint32_t nseg, const uint64_t* dst_lo, const uint64_t* src_lo,
                          const uint64_t* len, double* dst, const double* src)
{
  if(!h || (nseg && (!dst_lo || !src_lo || !len || !dst || !src)))
    return fail(FEAT_ERR_INVALID, "feat_batch_gather_f64: null argument");
  if(!nseg) return FEAT_OK;
  hipLaunchKernelGGL(k_batch_gather, dim3(64, nseg < 4096 ? nseg : 4096), dim3(kThreads), 0, h->stream, nseg, dst_lo,
                     src_lo, len, dst, src);
  return check_launch("feat_batch_gather_f64");
}

int feat_batch_tail_f64(feat_handle_t h, const feat_batch_tail_t* t, const double* r, double* z)
{
  const char* name = "feat_batch_tail_f64";
  if(!h || !t || !t->finest || !t->loff || !t->pcg_off) return fail(FEAT_ERR_INVALID, name);
  const int nl = t->nlevels;
  if(nl < 1 || nl > FEAT_BATCH_TAIL_MAX_LEVELS || t->nu_pre < 1 || t->nu_post < 1)
    return fail(FEAT_ERR_INVALID, "feat_batch_tail_f64: nlevels or smoothing steps out of range");
  uint64_t total = 0;
  for(int k = 1; k <= nl; ++k)
  {
    const feat_batch_tail_level_t& L = t->lv[k - 1];
    const bool transfer = k > 1;
    if(t->grid)
    {
      if(L.n < 1 || !L.inv || !L.mask || !L.sval || L.n1 < 3 || !(L.n1 & 1) || uint64_t(L.n1) * L.n1 != L.n ||
         (transfer && t->lv[k - 2].n1 != (L.n1 + 1) / 2))
        return fail(FEAT_ERR_INVALID, "feat_batch_tail_f64: grid level descriptor inconsistent");
      if(k == nl && total > FEAT_BATCH_TAIL_MAX_BELOW)
        return fail(FEAT_ERR_INVALID, "feat_batch_tail_f64: levels below the top exceed the LDS stencil store");
    }
    else
    {
      if(L.n < 1 || !L.inv || !L.mask || !L.a_val || !L.a_col || !L.a_len ||
         (transfer && (!L.p_val || !L.p_col || !L.p_len || !L.r_val || !L.r_col || !L.r_len)))
        return fail(FEAT_ERR_INVALID, "feat_batch_tail_f64: level descriptor incomplete");
      if(L.wa > FEAT_BATCH_TAIL_MAX_ROW_LEN ||
         (transfer && (L.wr > FEAT_BATCH_TAIL_MAX_ROW_LEN || L.wp > FEAT_BATCH_TAIL_MAX_PROL_LEN)))
        return fail(FEAT_ERR_INVALID, "feat_batch_tail_f64: ELL width above the limit");
    }
    total += L.n;
    // the restriction to a level gives one row to each thread
    if(L.n > FEAT_BATCH_TAIL_MAX_ROWS || (k < nl && L.n > uint64_t(kBt))) total = FEAT_BATCH_TAIL_MAX_TOTAL + 1;
  }
  if(total > FEAT_BATCH_TAIL_MAX_TOTAL)
    return fail(FEAT_ERR_INVALID, "feat_batch_tail_f64: levels exceed the workgroup's LDS vectors");
  if(!t->ninst) return FEAT_OK;
  if(!r || !z || !t->rhs_deep || !t->sol_deep) return fail(FEAT_ERR_INVALID, "feat_batch_tail_f64: null vector");
  if(static_cast<const void*>(r) == z || t->rhs_deep == t->sol_deep)
    return fail(FEAT_ERR_ALIAS, "feat_batch_tail_f64: input and output alias");
  if(t->grid)
    hipLaunchKernelGGL(k_batch_tail_grid, dim3(t->ninst), dim3(kBt), 0, h->stream, *t, r, z);
  else
    hipLaunchKernelGGL(k_batch_tail, dim3(t->ninst), dim3(kBt), 0, h->stream, *t, r, z);
  return check_launch(name);
}

int feat_batch_pre_tiled_f64(feat_handle_t h, const feat_batch_tiled_t* t)
{
  return tiled_launch(h, t, false, "feat_batch_pre_tiled_f64");
}

int feat_batch_post_tiled_f64(feat_handle_t h, const feat_batch_tiled_t* t)
{
  return tiled_launch(h, t, true, "feat_batch_post_tiled_f64");
}

int feat_copy_int_to_host(feat_handle_t h, const int* dev, int* host_pinned)
{
  if(!h || !dev || !host_pinned) return fail(FEAT_ERR_INVALID, "feat_copy_int_to_host: null argument");
  return check_hip(hipMemcpyAsync(host_pinned, dev, sizeof(int), hipMemcpyDeviceToHost, h->stream), "hipMemcpyAsync");
}

} // extern "C"
