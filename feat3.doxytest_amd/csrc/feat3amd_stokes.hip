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

// a CSR matrix with u64 indices (transfers)
struct StCsr
{
  const double* val;
  const uint64_t* col;
  const uint64_t* ptr;
  uint64_t rows;
};

// Tuple apply staged through LDS.  A workgroup covers kTupRows block rows of one kind:
// velocity rows (blocks 0 = A, 2x2, and 1 = B, 2x1) or pressure rows (2 = D, 1x2, and
// 3 = C, 1x1).  All 256 threads load the workgroup's blocks and column indices and gather
// the x they multiply into LDS -- every load of a pass issued before the first gather,
// every gather before the first LDS store -- then one lane per row adds its blocks from LDS
// in storage order, with the fmas of the row-serial loop (bit-identical to it).  A
// workgroup whose rows hold more blocks than one pass stages takes several passes (48 Q2
// rows hold at most 24 x 25 + 24 x 15 = 960 A blocks: one pass).
constexpr int kTupRows = 48;
constexpr int kTupE0 = 1024;   // blocks of the first matrix (A or D) per pass
constexpr int kTupE1 = 512;    // blocks of the second (B or C) per pass
constexpr int kTupU0 = kTupE0 / kThreads, kTupU1 = kTupE1 / kThreads;

struct TupleRowSums
{
  double s0, s1, t0, t1;   // first-matrix sums (rows 0/1 of a block), second-matrix sums
};

__device__ __forceinline__ TupleRowSums tuple_row_sums(const TupleDev& M, const double* __restrict__ x, bool vel,
                                                      uint64_t rb, int nr)
{
  __shared__ double4 s_v0[kTupE0];   // 32 KB: A blocks (D: .x .y)
  __shared__ double2 s_x0[kTupE0];   // 16 KB
  __shared__ double2 s_v1[kTupE1];   //  8 KB: B blocks (C: .x)
  __shared__ double s_x1[kTupE1];    //  4 KB
  __shared__ uint64_t s_p0[kTupRows + 1], s_p1[kTupRows + 1];
  const int tid = threadIdx.x;
  const int m0 = vel ? 0 : 2, m1 = vel ? 1 : 3;
  const bool has1 = M.val[m1] != nullptr;
  const double* xp = x + 2 * M.nv;
  if(tid <= nr)
  {
    s_p0[tid] = M.ptr[m0][rb + tid];
    s_p1[tid] = has1 ? M.ptr[m1][rb + tid] : 0;
  }
  __syncthreads();
  const uint64_t b0 = s_p0[0], e0 = s_p0[nr], b1 = s_p1[0], e1 = s_p1[nr];
  const uint64_t n0 = (e0 - b0 + kTupE0 - 1) / kTupE0, n1 = (e1 - b1 + kTupE1 - 1) / kTupE1;
  const uint64_t npass = n0 > n1 ? n0 : n1;
  const bool rl = tid < nr;
  uint64_t ra0 = 0, re0 = 0, ra1 = 0, re1 = 0;
  if(rl)
  {
    ra0 = s_p0[tid];
    re0 = s_p0[tid + 1];
    ra1 = s_p1[tid];
    re1 = s_p1[tid + 1];
  }
  TupleRowSums o{0.0, 0.0, 0.0, 0.0};
  const double* v0 = M.val[m0];
  const uint64_t* c0 = M.col[m0];
  const double* v1 = M.val[m1];
  const uint64_t* c1 = M.col[m1];
  for(uint64_t q = 0; q < npass; ++q)
  {
    const uint64_t w0 = b0 + q * kTupE0, w1 = b1 + q * kTupE1;
    double4 a[kTupU0];
    double2 xa[kTupU0], b[kTupU1];
    uint64_t ca[kTupU0], cb[kTupU1];
    double xb[kTupU1];
#pragma unroll
    for(int u = 0; u < kTupU0; ++u)
    {
      const uint64_t e = w0 + uint64_t(tid + u * kThreads);
      ca[u] = 0;
      a[u] = double4{0.0, 0.0, 0.0, 0.0};
      if(e < e0)
      {
        if(vel)
          a[u] = reinterpret_cast<const double4*>(v0)[e];
        else
        {
          const double2 d = reinterpret_cast<const double2*>(v0)[e];
          a[u].x = d.x;
          a[u].y = d.y;
        }
        ca[u] = c0[e];
      }
    }
#pragma unroll
    for(int u = 0; u < kTupU1; ++u)
    {
      const uint64_t e = w1 + uint64_t(tid + u * kThreads);
      cb[u] = 0;
      b[u] = double2{0.0, 0.0};
      if(e < e1)
      {
        if(vel)
          b[u] = reinterpret_cast<const double2*>(v1)[e];
        else
          b[u].x = v1[e];
        cb[u] = c1[e];
      }
    }
#pragma unroll
    for(int u = 0; u < kTupU0; ++u)
    {
      xa[u] = double2{0.0, 0.0};
      if(w0 + uint64_t(tid + u * kThreads) < e0) xa[u] = reinterpret_cast<const double2*>(x)[ca[u]];
    }
#pragma unroll
    for(int u = 0; u < kTupU1; ++u)
    {
      xb[u] = 0.0;
      if(w1 + uint64_t(tid + u * kThreads) < e1) xb[u] = xp[cb[u]];
    }
#pragma unroll
    for(int u = 0; u < kTupU0; ++u)
      if(w0 + uint64_t(tid + u * kThreads) < e0)
      {
        s_v0[tid + u * kThreads] = a[u];
        s_x0[tid + u * kThreads] = xa[u];
      }
#pragma unroll
    for(int u = 0; u < kTupU1; ++u)
      if(w1 + uint64_t(tid + u * kThreads) < e1)
      {
        s_v1[tid + u * kThreads] = b[u];
        s_x1[tid + u * kThreads] = xb[u];
      }
    __syncthreads();
    if(rl)
    {
      // four entries' LDS loads before their fmas (same fma order)
      const uint64_t lo0 = ra0 > w0 ? ra0 : w0, h0 = w0 + kTupE0, hi0 = re0 < h0 ? re0 : h0;
      uint64_t k = lo0;
      for(; k + 4 <= hi0; k += 4)
      {
        double4 v[4];
        double2 xv[4];
#pragma unroll
        for(int u = 0; u < 4; ++u)
        {
          v[u] = s_v0[k + u - w0];
          xv[u] = s_x0[k + u - w0];
        }
#pragma unroll
        for(int u = 0; u < 4; ++u)
        {
          o.s0 = dfma(v[u].x, xv[u].x, o.s0);
          o.s0 = dfma(v[u].y, xv[u].y, o.s0);
          if(vel)
          {
            o.s1 = dfma(v[u].z, xv[u].x, o.s1);
            o.s1 = dfma(v[u].w, xv[u].y, o.s1);
          }
        }
      }
      for(; k < hi0; ++k)
      {
        const double4 v = s_v0[k - w0];
        const double2 xv = s_x0[k - w0];
        o.s0 = dfma(v.x, xv.x, o.s0);
        o.s0 = dfma(v.y, xv.y, o.s0);
        if(vel)
        {
          o.s1 = dfma(v.z, xv.x, o.s1);
          o.s1 = dfma(v.w, xv.y, o.s1);
        }
      }
      const uint64_t lo1 = ra1 > w1 ? ra1 : w1, h1 = w1 + kTupE1, hi1 = re1 < h1 ? re1 : h1;
      k = lo1;
      for(; k + 4 <= hi1; k += 4)
      {
        double2 v[4];
        double xv[4];
#pragma unroll
        for(int u = 0; u < 4; ++u)
        {
          v[u] = s_v1[k + u - w1];
          xv[u] = s_x1[k + u - w1];
        }
#pragma unroll
        for(int u = 0; u < 4; ++u)
        {
          o.t0 = dfma(v[u].x, xv[u], o.t0);
          if(vel) o.t1 = dfma(v[u].y, xv[u], o.t1);
        }
      }
      for(; k < hi1; ++k)
      {
        const double2 v = s_v1[k - w1];
        const double xv = s_x1[k - w1];
        o.t0 = dfma(v.x, xv, o.t0);
        if(vel) o.t1 = dfma(v.y, xv, o.t1);
      }
    }
    __syncthreads();
  }
  return o;
}

// The pressure mean filter of MeanFilter::filter_def / filter_cor / filter_sol
// (mean_filter.hpp:200-262) on v[0, n), `passes` times: dot = v . w_dot, alpha = -dot/volume
// (sol_form: 0 - dot/volume), v += alpha w_add unless |alpha| < eps (DenseVector::axpy's copy,
// dense_vector.hpp:815-820).  The dot is summed as by 1,024 threads: virtual thread t adds
// v_i w_i for i = t, t + 1,024, ... in order, each wave of 64 virtual threads is summed by
// wave_sum, the 16 wave sums in order -- so a workgroup of NT = 256 or 1,024 threads gives
// the same bits.  Every thread of the workgroup calls it.
constexpr int kMeanThreads = 1024;
constexpr int kMeanBatch = 4;    // elements per virtual thread whose loads are issued together

// v[i] = a * w[i] + v[i] over [0, n) by the NT threads of the workgroup, kUpdBatch loads of
// each thread in flight (16 for 256 threads, 4 for 512 or more)
template<int NT>
__device__ __forceinline__ void axpy_block(double* __restrict__ v, double a, const double* __restrict__ w, uint64_t n)
{
  constexpr int kUpdBatch = NT >= 512 ? 4 : 16;
  for(uint64_t i0 = threadIdx.x; i0 < n; i0 += uint64_t(NT) * kUpdBatch)
  {
    double vv[kUpdBatch], ww[kUpdBatch];
#pragma unroll
    for(int u = 0; u < kUpdBatch; ++u)
    {
      const uint64_t i = i0 + uint64_t(u) * NT;
      vv[u] = i < n ? v[i] : 0.0;
      ww[u] = i < n ? w[i] : 0.0;
    }
#pragma unroll
    for(int u = 0; u < kUpdBatch; ++u)
    {
      const uint64_t i = i0 + uint64_t(u) * NT;
      if(i < n) v[i] = dfma(a, ww[u], vv[u]);
    }
  }
}

template<int NT>
__device__ __forceinline__ void mean_passes_block(double* __restrict__ v, const double* __restrict__ w_dot,
                                  const double* __restrict__ w_add, double volume, uint64_t n, int passes,
                                  int sol_form, double* red /* kMeanThreads / 64 */, double* s_alpha)
{
  static_assert(NT == 256 || NT == 512 || NT == 1024, "workgroup size");
  constexpr int kV = kMeanThreads / NT;   // virtual threads per thread: T + k NT, k < kV
  const int T = threadIdx.x, w = T >> 6;
  for(int pass = 0; pass < passes; ++pass)
  {
    double part[kV];
#pragma unroll
    for(int k = 0; k < kV; ++k) part[k] = 0.0;
    // element j of virtual thread t is i = t + j * 1024; kMeanBatch of them per round
    for(uint64_t j0 = 0; uint64_t(T) + j0 * kMeanThreads < n; j0 += kMeanBatch)
    {
      double a[kV][kMeanBatch], b[kV][kMeanBatch];
#pragma unroll
      for(int k = 0; k < kV; ++k)
#pragma unroll
        for(int u = 0; u < kMeanBatch; ++u)
        {
          const uint64_t i = uint64_t(T + k * NT) + (j0 + u) * kMeanThreads;
          a[k][u] = i < n ? v[i] : 0.0;
          b[k][u] = i < n ? w_dot[i] : 0.0;
        }
#pragma unroll
      for(int k = 0; k < kV; ++k)
#pragma unroll
        for(int u = 0; u < kMeanBatch; ++u)
          if(uint64_t(T + k * NT) + (j0 + u) * kMeanThreads < n) part[k] += a[k][u] * b[k][u];
    }
#pragma unroll
    for(int k = 0; k < kV; ++k)
    {
      const double ws = wave_sum(part[k]);
      if((T & 63) == 0) red[w + k * (NT / 64)] = ws;
    }
    __syncthreads();
    if(T == 0)
    {
      double rr[kMeanThreads / 64];
#pragma unroll
      for(int i = 0; i < kMeanThreads / 64; ++i) rr[i] = red[i];
      double dot = rr[0];
#pragma unroll
      for(int i = 1; i < kMeanThreads / 64; ++i) dot += rr[i];
      *s_alpha = sol_form ? 0.0 - dot / volume : -dot / volume;
    }
    __syncthreads();
    const double alpha = *s_alpha;
    if(fabs(alpha) >= DBL_EPSILON) axpy_block<NT>(v, alpha, w_add, n);
    __syncthreads();
  }
}

// gate of the smoother update (k_axpy_gated): skipped when the smoother's initial defect norm
// is not finite or below eps^2
__device__ __forceinline__ bool gate_open(const double* gate)
{
  if(!gate) return true;
  const double g = *gate;
  return isfinite(g) && g > DBL_EPSILON * DBL_EPSILON;
}

// What follows a tuple apply in the same launch: the pressure mean filter on r (passes 0 =
// none) and the smoother update x_upd += omega r (gated; x_upd null = none).
struct TuplePost
{
  const double* w_dot;
  const double* w_add;
  double volume;
  int passes, sol_form;
  double* x_upd;
  double omega;
  const double* gate;
  uint32_t* counter;   // zero between launches; counts the pressure workgroups in
  int tail;            // 1: filter + pressure update in the last workgroup; 0: left to k_mean_passes
  int x_zero;          // x_upd holds zeros and is not read: the update writes omega r, or zeros
                       // where it is skipped (the smoother's first step after sol.format())
};

// the update's value for one entry: x + omega r (x_zero: 0 + omega r), or x unchanged (x_zero:
// zero) when the update is skipped
__device__ __forceinline__ double post_update(bool open, bool x_zero, double omega, double r, double x)
{
  return open ? dfma(omega, r, x_zero ? 0.0 : x) : (x_zero ? 0.0 : x);
}

// The pressure part of a launch's TuplePost, run by every pressure workgroup after it stored
// its rows of rp: each publishes its rows and counts in; the last one applies the mean filter
// passes to rp[0, np) and the update to x_upd_p[0, np).
__device__ __forceinline__ void post_pressure_tail(const TuplePost& post, double* rp, double* x_upd_p, uint64_t np,
                                                   uint32_t npb)
{
  __shared__ double red[kMeanThreads / 64];
  __shared__ double s_alpha;
  __shared__ int last;
  // agent-scope release by every writer, then count in
  __threadfence();
  __syncthreads();
  if(threadIdx.x == 0)
  {
    const uint32_t prev = __hip_atomic_fetch_add(post.counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    last = (prev == npb - 1) ? 1 : 0;
  }
  __syncthreads();
  if(!last) return;
  __threadfence();   // acquire: the other workgroups' rows
  mean_passes_block<kThreads>(rp, post.w_dot, post.w_add, post.volume, np, post.passes, post.sol_form, red, &s_alpha);
  if(x_upd_p)
  {
    const bool open = fabs(post.omega) >= DBL_EPSILON && gate_open(post.gate);
    if(open && !post.x_zero)
      axpy_block<kThreads>(x_upd_p, post.omega, rp, np);
    else if(open || post.x_zero)
      for(uint64_t i = threadIdx.x; i < np; i += kThreads) x_upd_p[i] = post_update(open, true, post.omega, rp[i], 0.0);
  }
  if(threadIdx.x == 0) __hip_atomic_store(post.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Pressure parts above this many rows are filtered by k_mean_passes (1,024 threads) in a launch
// of its own: one 256-thread workgroup needs ~15 dependent memory round trips for 12,288 rows
// (Q2/P1dc level 6: 36 us fused against 21 us for the separate launches; 14 against 19 us at
// level 5 with 3,072).
constexpr uint64_t kTupTailMaxP = 4096;

// r = K x (y null) or r = y + alpha K x, velocity rows zeroed where vmask is set
// (SaddlePointMatrix::apply, saddle_point_matrix.hpp:446-491: A.apply then B.apply into the
// velocity part, D.apply then C.apply into the pressure part).  Workgroups 0 .. nvb-1 take
// the velocity rows, the rest the pressure rows.  POST: then the TuplePost operations --
// velocity workgroups update their own rows of x_upd (the filter leaves velocity rows alone);
// pressure workgroups publish their rows and count in, and the last one filters the whole
// pressure part and updates x_upd's.
template<bool POST>
__global__ __launch_bounds__(kThreads) void k_tuple_apply(TupleDev M, double* __restrict__ r,
                                                          const double* __restrict__ x,
                                                          const double* __restrict__ y, double alpha,
                                                          const uint32_t* __restrict__ vmask, uint32_t nvb,
                                                          TuplePost post)
{
  const bool vel = blockIdx.x < nvb;
  const uint64_t nrows = vel ? M.nv : M.np;
  const uint64_t rb = uint64_t(vel ? blockIdx.x : blockIdx.x - nvb) * kTupRows;
  const uint64_t left = nrows - rb;
  const int nr = int(left < uint64_t(kTupRows) ? left : uint64_t(kTupRows));
  const bool rl = int(threadIdx.x) < nr;
  const uint64_t row = rb + threadIdx.x;
  // the row's own operands, loaded before the staging
  double y0 = 0.0, y1 = 0.0, u0 = 0.0, u1 = 0.0;
  bool masked = false;
  const bool open = POST && fabs(post.omega) >= DBL_EPSILON && gate_open(post.gate);
  const bool upd = POST && post.x_upd != nullptr && vel && (open || post.x_zero);
  const bool upd_read = upd && open && !post.x_zero;
  if(rl)
  {
    if(vel)
    {
      if(y)
      {
        y0 = y[2 * row];
        y1 = y[2 * row + 1];
      }
      masked = is_masked(vmask, row);
      if(upd_read)
      {
        u0 = post.x_upd[2 * row];
        u1 = post.x_upd[2 * row + 1];
      }
    }
    else if(y)
      y0 = y[2 * M.nv + row];
  }
  const TupleRowSums o = tuple_row_sums(M, x, vel, rb, nr);
  const double a1 = y ? alpha : 1.0;
  if(rl)
  {
    if(vel)
    {
      // A.apply(r, x_u, y_u, alpha): r = sum*alpha + beta*y (beta = 0 without y)
      double r0 = y ? dfma(o.s0, alpha, y0) : dfma(o.s0, 1.0, 0.0);
      double r1 = y ? dfma(o.s1, alpha, y1) : dfma(o.s1, 1.0, 0.0);
      if(M.val[1])
      {
        // B.apply(r, x_p, r, alpha)
        r0 = dfma(o.t0, a1, r0);
        r1 = dfma(o.t1, a1, r1);
      }
      if(masked) r0 = r1 = 0.0;
      r[2 * row] = r0;
      r[2 * row + 1] = r1;
      if(upd)
      {
        post.x_upd[2 * row] = post_update(open, post.x_zero, post.omega, r0, u0);
        post.x_upd[2 * row + 1] = post_update(open, post.x_zero, post.omega, r1, u1);
      }
    }
    else
    {
      double rp = y ? dfma(o.s0, alpha, y0) : dfma(o.s0, 1.0, 0.0);
      if(M.val[3]) rp = dfma(o.t0, a1, rp);
      r[2 * M.nv + row] = rp;
    }
  }
  if constexpr(POST)
  {
    if(vel || !post.tail) return;
    post_pressure_tail(post, r + 2 * M.nv, post.x_upd ? post.x_upd + 2 * M.nv : nullptr, M.np, gridDim.x - nvb);
  }
}

// Stokes transfer with what follows it (TupleDiagMatrix<BWrappedCSR<2>, CSR>::apply, then
// filter_def or filter_cor and an optional axpy): r = [V x_u ; P x_p] (velocity rows: the scalar
// CSR V on both components, k_csrsb2; pressure rows: P, k_csr_rows with a = 1, b = 0), velocity
// rows zeroed by vmask, `post.passes` MeanFilter passes on the pressure part, then x_upd +=
// post.omega r.  Rows are staged through LDS like the tuple apply; workgroups 0 .. nvb-1 take
// the velocity rows.  x: [u (2 per node, nvx nodes) | p].
constexpr int kTrRows = 32;     // rows per workgroup (restriction rows hold up to 25 entries)
constexpr int kTrE = 1024;      // entries staged per pass
template<bool POST>
__global__ __launch_bounds__(kThreads) void k_stokes_transfer(StCsr V, StCsr P, const double* __restrict__ x, uint64_t nvx,
                                                              double* __restrict__ r, const uint32_t* __restrict__ vmask,
                                                              uint32_t nvb, TuplePost post)
{
  __shared__ double s_a[kTrE];
  __shared__ double2 s_x[kTrE];
  const int t = threadIdx.x;
  const bool vel = blockIdx.x < nvb;
  const StCsr& M = vel ? V : P;
  const uint64_t rb = uint64_t(vel ? blockIdx.x : blockIdx.x - nvb) * kTrRows;
  const uint64_t left = M.rows - rb;
  const int nr = int(left < uint64_t(kTrRows) ? left : uint64_t(kTrRows));
  const bool rl = t < nr;
  const uint64_t row = rb + t;
  const uint64_t b = M.ptr[rb], e = M.ptr[rb + nr];
  uint64_t ra = 0, re = 0;
  double u0 = 0.0, u1 = 0.0;
  const bool open = POST && fabs(post.omega) >= DBL_EPSILON && gate_open(post.gate);
  const bool upd = POST && post.x_upd != nullptr && vel && (open || post.x_zero);
  const bool upd_read = upd && open && !post.x_zero;
  if(rl)
  {
    ra = M.ptr[row];
    re = M.ptr[row + 1];
    if(upd_read)
    {
      u0 = post.x_upd[2 * row];
      u1 = post.x_upd[2 * row + 1];
    }
  }
  const double* xp = x + 2 * nvx;
  double s0 = 0.0, s1 = 0.0;
  for(uint64_t w = b; w < e; w += kTrE)
  {
    double a[kTrE / kThreads];
    uint64_t c[kTrE / kThreads];
    double2 xv[kTrE / kThreads];
#pragma unroll
    for(int u = 0; u < kTrE / kThreads; ++u)
    {
      const uint64_t k = w + uint64_t(t + u * kThreads);
      a[u] = k < e ? M.val[k] : 0.0;
      c[u] = k < e ? M.col[k] : 0;
    }
#pragma unroll
    for(int u = 0; u < kTrE / kThreads; ++u)
    {
      xv[u] = double2{0.0, 0.0};
      if(w + uint64_t(t + u * kThreads) < e)
      {
        if(vel)
          xv[u] = double2{x[2 * c[u]], x[2 * c[u] + 1]};
        else
          xv[u].x = xp[c[u]];
      }
    }
#pragma unroll
    for(int u = 0; u < kTrE / kThreads; ++u)
    {
      s_a[t + u * kThreads] = a[u];
      s_x[t + u * kThreads] = xv[u];
    }
    __syncthreads();
    if(rl)
    {
      const uint64_t lo = ra > w ? ra : w, hw = w + kTrE, hi = re < hw ? re : hw;
      for(uint64_t k = lo; k < hi; ++k)
      {
        const double aa = s_a[k - w];
        const double2 xx = s_x[k - w];
        s0 = dfma(aa, xx.x, s0);
        if(vel) s1 = dfma(aa, xx.y, s1);
      }
    }
    __syncthreads();
  }
  if(rl)
  {
    if(vel)
    {
      double r0 = dfma(s0, 1.0, 0.0), r1 = dfma(s1, 1.0, 0.0);
      if(POST && is_masked(vmask, row)) r0 = r1 = 0.0;
      r[2 * row] = r0;
      r[2 * row + 1] = r1;
      if(upd)
      {
        post.x_upd[2 * row] = post_update(open, post.x_zero, post.omega, r0, u0);
        post.x_upd[2 * row + 1] = post_update(open, post.x_zero, post.omega, r1, u1);
      }
    }
    else
      r[2 * V.rows + row] = dfma(s0, 1.0, 0.0);
  }
  if constexpr(POST)
  {
    if(vel || !post.tail) return;
    post_pressure_tail(post, r + 2 * V.rows, post.x_upd ? post.x_upd + 2 * V.rows : nullptr, P.rows,
                       gridDim.x - nvb);
  }
}

// ---------------------------------------------------------------------------
// mean filter passes (one workgroup) and small vector kernels
// ---------------------------------------------------------------------------
// the passes, then (x_upd != null) x_upd += omega v gated as k_axpy_gated
__global__ __launch_bounds__(kMeanThreads) void k_mean_passes(double* __restrict__ v, const double* __restrict__ w_dot,
                                                              const double* __restrict__ w_add, double volume,
                                                              uint64_t n, int passes, int sol_form,
                                                              double* __restrict__ x_upd, double omega,
                                                              const double* __restrict__ gate, int x_zero)
{
  __shared__ double red[kMeanThreads / 64];
  __shared__ double s_alpha;
  mean_passes_block<kMeanThreads>(v, w_dot, w_add, volume, n, passes, sol_form, red, &s_alpha);
  if(!x_upd) return;
  const bool open = fabs(omega) >= DBL_EPSILON && gate_open(gate);
  if(open && !x_zero)
    axpy_block<kMeanThreads>(x_upd, omega, v, n);
  else if(open || x_zero)
    for(uint64_t i = threadIdx.x; i < n; i += kMeanThreads) x_upd[i] = post_update(open, true, omega, v[i], 0.0);
}

// sums of squares of v[0, nu) and v[nu, nu + np); with norm != null (and the sum finished in
// this launch) the workgroup that completes it also writes sqrt(|u|^2 + |p|^2) as
// k_tuple_norm_finish does
__global__ __launch_bounds__(kThreads) void k_sumsq2(const double* __restrict__ v, uint64_t nu, uint64_t np,
                                                     GridReduce gr, double* __restrict__ norm)
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
  if(grid_reduce_finish(gr, su, sp) && norm && threadIdx.x == 0)
  {
    const double a = sqrt(*gr.out0), b = sqrt(*gr.out1);
    *norm = sqrt(dfma(a, a, b * b));
  }
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
// the coarse levels of the Stokes V-cycle in one workgroup
// ---------------------------------------------------------------------------
// Every operation of StokesMultiGrid::mg_apply on the tail's levels -- smoother
// (format, norm gate, AmaVanka apply + filter_cor + gated update, defect + filter_def),
// defect, restriction + filter_def, coarse AmaVanka apply, prolongation + filter_cor,
// the two axpys -- with the arithmetic of the kernels the separate launches run, so
// the tail is bit-identical to them.  The level vectors stay in LDS; matrix blocks and
// column indices are staged through LDS in passes (every load of a pass issued before
// the first LDS store), each row then summed by its thread in storage order with x read
// from the LDS vectors.
constexpr int kSt = 512;                     // threads of the tail workgroup (256 registers each)
constexpr int kStE0 = 2048, kStE1 = 1024;    // entries per pass: A / D / transfer, B / C
constexpr int kStMaxLevels = FEAT_STOKES_TAIL_MAX_LEVELS;

struct StStage
{
  double4 v0[kStE0];    // 64 KB (transfers: 8,192 doubles)
  uint32_t c0[kStE0];   //  8 KB
  double2 v1[kStE1];    // 16 KB
  uint32_t c1[kStE1];   //  4 KB
};

struct StLevel
{
  TupleDev K, V;
  const uint32_t* vmask;
  const double* prim;
  const double* dual;
  double vol;
  StCsr rv, rp, pv, pp;   // restriction to / prolongation from the next coarser level
};

struct StTail
{
  int nl, steps;
  double omega;
  StLevel lv[kStMaxLevels];
};

// Row sums over staged entries [lo, hi) of a pass starting at w, in storage order: the
// entries' LDS loads and the x they address are fetched four at a time before their fmas.
// First matrix (A: 2x2 blocks, vel; D: 1x2 blocks), x with two components per column:
__device__ __forceinline__ void st_sum0(const StStage& S, uint64_t w, uint64_t lo, uint64_t hi, const double* x,
                                        bool vel, double& s0, double& s1)
{
  uint64_t k = lo;
  for(; k + 4 <= hi; k += 4)
  {
    double4 v[4];
    uint32_t c[4];
    double x0[4], x1[4];
#pragma unroll
    for(int u = 0; u < 4; ++u)
    {
      v[u] = S.v0[k + u - w];
      c[u] = S.c0[k + u - w];
    }
#pragma unroll
    for(int u = 0; u < 4; ++u)
    {
      x0[u] = x[2 * c[u]];
      x1[u] = x[2 * c[u] + 1];
    }
#pragma unroll
    for(int u = 0; u < 4; ++u)
    {
      s0 = dfma(v[u].x, x0[u], s0);
      s0 = dfma(v[u].y, x1[u], s0);
      if(vel)
      {
        s1 = dfma(v[u].z, x0[u], s1);
        s1 = dfma(v[u].w, x1[u], s1);
      }
    }
  }
  for(; k < hi; ++k)
  {
    const double4 v = S.v0[k - w];
    const uint32_t c = S.c0[k - w];
    const double x0 = x[2 * c], x1 = x[2 * c + 1];
    s0 = dfma(v.x, x0, s0);
    s0 = dfma(v.y, x1, s0);
    if(vel)
    {
      s1 = dfma(v.z, x0, s1);
      s1 = dfma(v.w, x1, s1);
    }
  }
}
// second matrix (B: 2x1 blocks, vel; C: 1x1), x with one component per column
__device__ __forceinline__ void st_sum1(const StStage& S, uint64_t w, uint64_t lo, uint64_t hi, const double* xp,
                                        bool vel, double& t0, double& t1)
{
  uint64_t k = lo;
  for(; k + 4 <= hi; k += 4)
  {
    double2 v[4];
    double xv[4];
#pragma unroll
    for(int u = 0; u < 4; ++u) v[u] = S.v1[k + u - w];
#pragma unroll
    for(int u = 0; u < 4; ++u) xv[u] = xp[S.c1[k + u - w]];
#pragma unroll
    for(int u = 0; u < 4; ++u)
    {
      t0 = dfma(v[u].x, xv[u], t0);
      if(vel) t1 = dfma(v[u].y, xv[u], t1);
    }
  }
  for(; k < hi; ++k)
  {
    const double2 v = S.v1[k - w];
    const double xv = xp[S.c1[k - w]];
    t0 = dfma(v.x, xv, t0);
    if(vel) t1 = dfma(v.y, xv, t1);
  }
}

// r = y + alpha M x (y null: r = M x) over rows [rb, rb + nr) of one kind, vectors in LDS
// (k_tuple_apply's arithmetic)
__device__ __forceinline__ void st_tuple_group(const TupleDev& M, bool vel, uint64_t rb, int nr, double* r, const double* x,
                               const double* y, double alpha, const uint32_t* vmask, StStage& S)
{
  const int t = threadIdx.x;
  const int m0 = vel ? 0 : 2, m1 = vel ? 1 : 3;
  const bool has1 = M.val[m1] != nullptr;
  const uint64_t* p0 = M.ptr[m0];
  const uint64_t* p1 = M.ptr[m1];
  const uint64_t b0 = p0[rb], e0 = p0[rb + nr];
  const uint64_t b1 = has1 ? p1[rb] : 0, e1 = has1 ? p1[rb + nr] : 0;
  const bool rl = t < nr;
  uint64_t ra0 = 0, re0 = 0, ra1 = 0, re1 = 0;
  if(rl)
  {
    ra0 = p0[rb + t];
    re0 = p0[rb + t + 1];
    if(has1)
    {
      ra1 = p1[rb + t];
      re1 = p1[rb + t + 1];
    }
  }
  const uint64_t n0 = (e0 - b0 + kStE0 - 1) / kStE0, n1 = (e1 - b1 + kStE1 - 1) / kStE1;
  const uint64_t npass = n0 > n1 ? n0 : n1;
  const double* v0 = M.val[m0];
  const uint64_t* c0 = M.col[m0];
  const double* v1 = M.val[m1];
  const uint64_t* c1 = M.col[m1];
  const double* xp = x + 2 * M.nv;
  double s0 = 0.0, s1 = 0.0, t0 = 0.0, t1 = 0.0;
  for(uint64_t q = 0; q < npass; ++q)
  {
    const uint64_t w0 = b0 + q * kStE0, w1 = b1 + q * kStE1;
    double4 a[kStE0 / kSt];
    uint32_t ca[kStE0 / kSt];
    double2 b[kStE1 / kSt];
    uint32_t cb[kStE1 / kSt];
#pragma unroll
    for(int u = 0; u < kStE0 / kSt; ++u)
    {
      const uint64_t e = w0 + uint64_t(t + u * kSt);
      a[u] = double4{0.0, 0.0, 0.0, 0.0};
      ca[u] = 0;
      if(e < e0)
      {
        if(vel)
          a[u] = reinterpret_cast<const double4*>(v0)[e];
        else
        {
          const double2 d = reinterpret_cast<const double2*>(v0)[e];
          a[u].x = d.x;
          a[u].y = d.y;
        }
        ca[u] = uint32_t(c0[e]);
      }
    }
#pragma unroll
    for(int u = 0; u < kStE1 / kSt; ++u)
    {
      const uint64_t e = w1 + uint64_t(t + u * kSt);
      b[u] = double2{0.0, 0.0};
      cb[u] = 0;
      if(e < e1)
      {
        if(vel)
          b[u] = reinterpret_cast<const double2*>(v1)[e];
        else
          b[u].x = v1[e];
        cb[u] = uint32_t(c1[e]);
      }
    }
#pragma unroll
    for(int u = 0; u < kStE0 / kSt; ++u)
    {
      S.v0[t + u * kSt] = a[u];
      S.c0[t + u * kSt] = ca[u];
    }
#pragma unroll
    for(int u = 0; u < kStE1 / kSt; ++u)
    {
      S.v1[t + u * kSt] = b[u];
      S.c1[t + u * kSt] = cb[u];
    }
    __syncthreads();
    if(rl)
    {
      const uint64_t lo0 = ra0 > w0 ? ra0 : w0, h0 = w0 + kStE0, hi0 = re0 < h0 ? re0 : h0;
      st_sum0(S, w0, lo0, hi0, x, vel, s0, s1);
      const uint64_t lo1 = ra1 > w1 ? ra1 : w1, h1 = w1 + kStE1, hi1 = re1 < h1 ? re1 : h1;
      st_sum1(S, w1, lo1, hi1, xp, vel, t0, t1);
    }
    __syncthreads();
  }
  if(rl)
  {
    const uint64_t row = rb + t;
    const double a1 = y ? alpha : 1.0;
    if(vel)
    {
      double r0 = y ? dfma(s0, alpha, y[2 * row]) : dfma(s0, 1.0, 0.0);
      double r1 = y ? dfma(s1, alpha, y[2 * row + 1]) : dfma(s1, 1.0, 0.0);
      if(has1)
      {
        r0 = dfma(t0, a1, r0);
        r1 = dfma(t1, a1, r1);
      }
      if(is_masked(vmask, row)) r0 = r1 = 0.0;
      r[2 * row] = r0;
      r[2 * row + 1] = r1;
    }
    else
    {
      double rp = y ? dfma(s0, alpha, y[2 * M.nv + row]) : dfma(s0, 1.0, 0.0);
      if(has1) rp = dfma(t0, a1, rp);
      r[2 * M.nv + row] = rp;
    }
  }
  __syncthreads();
}

// The same for all rows at once when nv + np <= kSt: the velocity rows' A blocks and the
// pressure rows' D blocks form one staged sequence (B and C blocks another), so a pass serves
// both kinds and the row pointers of both are fetched in one round trip.
__device__ __forceinline__ void st_tuple_all(const TupleDev& M, double* r, const double* x, const double* y,
                                             double alpha, const uint32_t* vmask, StStage& S)
{
  const int t = threadIdx.x;
  const uint64_t nv = M.nv, np = M.np;
  const bool hasB = M.val[1] != nullptr, hasC = M.val[3] != nullptr;
  const uint64_t bA = M.ptr[0][0], eA = M.ptr[0][nv], bD = M.ptr[2][0], eD = M.ptr[2][np];
  const uint64_t bB = hasB ? M.ptr[1][0] : 0, eB = hasB ? M.ptr[1][nv] : 0;
  const uint64_t bC = hasC ? M.ptr[3][0] : 0, eC = hasC ? M.ptr[3][np] : 0;
  const uint64_t nA = eA - bA, nD = eD - bD, nB = eB - bB, nC = eC - bC;
  const bool vel = uint64_t(t) < nv, pre = !vel && uint64_t(t) < nv + np;
  const uint64_t row = vel ? uint64_t(t) : uint64_t(t) - nv;
  // this row's range in the two staged sequences
  uint64_t ra0 = 0, re0 = 0, ra1 = 0, re1 = 0;
  double y0 = 0.0, y1 = 0.0;
  if(vel)
  {
    ra0 = M.ptr[0][row] - bA;
    re0 = M.ptr[0][row + 1] - bA;
    if(hasB)
    {
      ra1 = M.ptr[1][row] - bB;
      re1 = M.ptr[1][row + 1] - bB;
    }
    if(y)
    {
      y0 = y[2 * row];
      y1 = y[2 * row + 1];
    }
  }
  else if(pre)
  {
    ra0 = nA + (M.ptr[2][row] - bD);
    re0 = nA + (M.ptr[2][row + 1] - bD);
    if(hasC)
    {
      ra1 = nB + (M.ptr[3][row] - bC);
      re1 = nB + (M.ptr[3][row + 1] - bC);
    }
    if(y) y0 = y[2 * nv + row];
  }
  const uint64_t n0 = nA + nD, n1 = nB + nC;
  const uint64_t p0 = (n0 + kStE0 - 1) / kStE0, p1 = (n1 + kStE1 - 1) / kStE1;
  const uint64_t npass = p0 > p1 ? p0 : p1;
  const double* xp = x + 2 * nv;
  double s0 = 0.0, s1 = 0.0, t0 = 0.0, t1 = 0.0;
  for(uint64_t q = 0; q < npass; ++q)
  {
    const uint64_t w0 = q * kStE0, w1 = q * kStE1;
    double4 a[kStE0 / kSt];
    uint32_t ca[kStE0 / kSt];
    double2 b[kStE1 / kSt];
    uint32_t cb[kStE1 / kSt];
#pragma unroll
    for(int u = 0; u < kStE0 / kSt; ++u)
    {
      const uint64_t j = w0 + uint64_t(t + u * kSt);
      a[u] = double4{0.0, 0.0, 0.0, 0.0};
      ca[u] = 0;
      if(j < nA)
      {
        a[u] = reinterpret_cast<const double4*>(M.val[0])[bA + j];
        ca[u] = uint32_t(M.col[0][bA + j]);
      }
      else if(j < n0)
      {
        const double2 d = reinterpret_cast<const double2*>(M.val[2])[bD + (j - nA)];
        a[u].x = d.x;
        a[u].y = d.y;
        ca[u] = uint32_t(M.col[2][bD + (j - nA)]);
      }
    }
#pragma unroll
    for(int u = 0; u < kStE1 / kSt; ++u)
    {
      const uint64_t j = w1 + uint64_t(t + u * kSt);
      b[u] = double2{0.0, 0.0};
      cb[u] = 0;
      if(j < nB)
      {
        b[u] = reinterpret_cast<const double2*>(M.val[1])[bB + j];
        cb[u] = uint32_t(M.col[1][bB + j]);
      }
      else if(j < n1)
      {
        b[u].x = M.val[3][bC + (j - nB)];
        cb[u] = uint32_t(M.col[3][bC + (j - nB)]);
      }
    }
#pragma unroll
    for(int u = 0; u < kStE0 / kSt; ++u)
    {
      S.v0[t + u * kSt] = a[u];
      S.c0[t + u * kSt] = ca[u];
    }
#pragma unroll
    for(int u = 0; u < kStE1 / kSt; ++u)
    {
      S.v1[t + u * kSt] = b[u];
      S.c1[t + u * kSt] = cb[u];
    }
    __syncthreads();
    if(vel || pre)
    {
      const uint64_t lo0 = ra0 > w0 ? ra0 : w0, h0 = w0 + kStE0, hi0 = re0 < h0 ? re0 : h0;
      st_sum0(S, w0, lo0, hi0, x, vel, s0, s1);
      const uint64_t lo1 = ra1 > w1 ? ra1 : w1, h1 = w1 + kStE1, hi1 = re1 < h1 ? re1 : h1;
      st_sum1(S, w1, lo1, hi1, xp, vel, t0, t1);
    }
    __syncthreads();
  }
  const double a1 = y ? alpha : 1.0;
  if(vel)
  {
    double r0 = y ? dfma(s0, alpha, y0) : dfma(s0, 1.0, 0.0);
    double r1 = y ? dfma(s1, alpha, y1) : dfma(s1, 1.0, 0.0);
    if(hasB)
    {
      r0 = dfma(t0, a1, r0);
      r1 = dfma(t1, a1, r1);
    }
    if(is_masked(vmask, row)) r0 = r1 = 0.0;
    r[2 * row] = r0;
    r[2 * row + 1] = r1;
  }
  else if(pre)
  {
    double rp = y ? dfma(s0, alpha, y0) : dfma(s0, 1.0, 0.0);
    if(hasC) rp = dfma(t0, a1, rp);
    r[2 * nv + row] = rp;
  }
  __syncthreads();
}

__device__ __forceinline__ void st_tuple(const TupleDev& M, double* r, const double* x, const double* y, double alpha,
                         const uint32_t* vmask, StStage& S)
{
  if(M.nv + M.np <= uint64_t(kSt))
  {
    st_tuple_all(M, r, x, y, alpha, vmask, S);
    return;
  }
  for(uint64_t rb = 0; rb < M.nv; rb += kSt)
    st_tuple_group(M, true, rb, int(M.nv - rb < uint64_t(kSt) ? M.nv - rb : uint64_t(kSt)), r, x, y, alpha, vmask, S);
  for(uint64_t rb = 0; rb < M.np; rb += kSt)
    st_tuple_group(M, false, rb, int(M.np - rb < uint64_t(kSt) ? M.np - rb : uint64_t(kSt)), r, x, y, alpha, vmask, S);
}

// r = R x with NC components per row (NC = 2: k_csrsb2, BWrappedCSR; NC = 1: k_csr_rows with
// a = 1, b = 0): sums from zero in storage order, then fma(sum, 1, 0)
template<int NC>
__device__ __forceinline__ void st_csr(const StCsr& R, const double* x, double* r, StStage& S)
{
  double* sv = reinterpret_cast<double*>(S.v0);
  constexpr int kE = kStE0;
  const int t = threadIdx.x;
  for(uint64_t rb = 0; rb < R.rows; rb += kSt)
  {
    const int nr = int(R.rows - rb < uint64_t(kSt) ? R.rows - rb : uint64_t(kSt));
    const uint64_t b = R.ptr[rb], e = R.ptr[rb + nr];
    const bool rl = t < nr;
    const uint64_t ra = rl ? R.ptr[rb + t] : 0, re = rl ? R.ptr[rb + t + 1] : 0;
    double s0 = 0.0, s1 = 0.0;
    for(uint64_t w = b; w < e; w += kE)
    {
      double v[kE / kSt];
      uint32_t c[kE / kSt];
#pragma unroll
      for(int u = 0; u < kE / kSt; ++u)
      {
        const uint64_t k = w + uint64_t(t + u * kSt);
        v[u] = k < e ? R.val[k] : 0.0;
        c[u] = k < e ? uint32_t(R.col[k]) : 0u;
      }
#pragma unroll
      for(int u = 0; u < kE / kSt; ++u)
      {
        sv[t + u * kSt] = v[u];
        S.c0[t + u * kSt] = c[u];
      }
      __syncthreads();
      if(rl)
      {
        const uint64_t lo = ra > w ? ra : w, hw = w + kE, hi = re < hw ? re : hw;
        for(uint64_t k = lo; k < hi; ++k)
        {
          const double a = sv[k - w];
          const uint32_t cc = S.c0[k - w];
          s0 = dfma(a, x[NC * cc], s0);
          if(NC == 2) s1 = dfma(a, x[NC * cc + 1], s1);
        }
      }
      __syncthreads();
    }
    if(rl)
    {
      r[NC * (rb + t)] = dfma(s0, 1.0, 0.0);
      if(NC == 2) r[NC * (rb + t) + 1] = dfma(s1, 1.0, 0.0);
    }
  }
  __syncthreads();
}

// r = T x for a Stokes transfer (velocity rows by V with two components, pressure rows by P),
// both staged in one sequence when rows(V) + rows(P) <= kSt (else st_csr twice);
// x: [u (2 per node, nvx nodes) | p], r: [u (2 per node, V.rows nodes) | p]
__device__ __forceinline__ void st_transfer(const StCsr& V, const StCsr& P, const double* x, uint64_t nvx, double* r,
                                            StStage& S)
{
  if(V.rows + P.rows > uint64_t(kSt))
  {
    st_csr<2>(V, x, r, S);
    st_csr<1>(P, x + 2 * nvx, r + 2 * V.rows, S);
    return;
  }
  double* sv = reinterpret_cast<double*>(S.v0);
  constexpr int kE = kStE0;
  const int t = threadIdx.x;
  const uint64_t bV = V.ptr[0], eV = V.ptr[V.rows], bP = P.ptr[0], eP = P.ptr[P.rows];
  const uint64_t nV = eV - bV, n = nV + (eP - bP);
  const bool vel = uint64_t(t) < V.rows, pre = !vel && uint64_t(t) < V.rows + P.rows;
  const uint64_t row = vel ? uint64_t(t) : uint64_t(t) - V.rows;
  uint64_t ra = 0, re = 0;
  if(vel)
  {
    ra = V.ptr[row] - bV;
    re = V.ptr[row + 1] - bV;
  }
  else if(pre)
  {
    ra = nV + (P.ptr[row] - bP);
    re = nV + (P.ptr[row + 1] - bP);
  }
  const double* xp = x + 2 * nvx;
  double s0 = 0.0, s1 = 0.0;
  for(uint64_t w = 0; w < n; w += kE)
  {
    double v[kE / kSt];
    uint32_t c[kE / kSt];
#pragma unroll
    for(int u = 0; u < kE / kSt; ++u)
    {
      const uint64_t j = w + uint64_t(t + u * kSt);
      v[u] = 0.0;
      c[u] = 0u;
      if(j < nV)
      {
        v[u] = V.val[bV + j];
        c[u] = uint32_t(V.col[bV + j]);
      }
      else if(j < n)
      {
        v[u] = P.val[bP + (j - nV)];
        c[u] = uint32_t(P.col[bP + (j - nV)]);
      }
    }
#pragma unroll
    for(int u = 0; u < kE / kSt; ++u)
    {
      sv[t + u * kSt] = v[u];
      S.c0[t + u * kSt] = c[u];
    }
    __syncthreads();
    const uint64_t lo = ra > w ? ra : w, hw = w + kE, hi = re < hw ? re : hw;
    if(vel || pre)
    {
      const double* xx = vel ? x : xp;
      const int nc = vel ? 2 : 1;
      uint64_t k = lo;
      for(; k + 4 <= hi; k += 4)
      {
        double a[4], x0[4], x1[4];
        uint32_t cc[4];
#pragma unroll
        for(int u = 0; u < 4; ++u)
        {
          a[u] = sv[k + u - w];
          cc[u] = S.c0[k + u - w];
        }
#pragma unroll
        for(int u = 0; u < 4; ++u)
        {
          x0[u] = xx[nc * cc[u]];
          x1[u] = vel ? xx[nc * cc[u] + 1] : 0.0;
        }
#pragma unroll
        for(int u = 0; u < 4; ++u)
        {
          s0 = dfma(a[u], x0[u], s0);
          if(vel) s1 = dfma(a[u], x1[u], s1);
        }
      }
      for(; k < hi; ++k)
      {
        const double a = sv[k - w];
        const uint32_t cc = S.c0[k - w];
        s0 = dfma(a, xx[nc * cc], s0);
        if(vel) s1 = dfma(a, xx[nc * cc + 1], s1);
      }
    }
    __syncthreads();
  }
  if(vel)
  {
    r[2 * row] = dfma(s0, 1.0, 0.0);
    r[2 * row + 1] = dfma(s1, 1.0, 0.0);
  }
  else if(pre)
    r[2 * V.rows + row] = dfma(s0, 1.0, 0.0);
  __syncthreads();
}

// the norm of feat_tuple_norm2_f64 (k_sumsq2 on grid_for(n, 8) workgroups, the in-kernel grid
// finish, k_tuple_norm_finish) for up to kSt / 256 of those workgroups (n <= 1,024 kSt / 256);
// result in *out
__device__ __forceinline__ void st_norm(const double* v, uint64_t nu, uint64_t np, double* redu, double* redp, double* out)
{
  const uint64_t n = nu + np;
  const uint64_t G = (n + 2047) / 2048;
  const int T = threadIdx.x, b = T >> 8, t = T & 255, w = T >> 6;
  double su = 0.0, sp = 0.0;
  if(uint64_t(b) < G)
    for(uint64_t i = uint64_t(b) * 256 + t; i < n; i += G * 256)
    {
      const double q = v[i];
      if(i < nu)
        su = dfma(q, q, su);
      else
        sp = dfma(q, q, sp);
    }
  su = wave_sum(su);
  sp = wave_sum(sp);
  if((T & 63) == 0)
  {
    redu[w] = su;
    redp[w] = sp;
  }
  __syncthreads();
  if(w == 0)
  {
    // partial of virtual workgroup `lane` (its block sum), then the grid finish's block sum
    double pu = 0.0, pp = 0.0;
    if(uint64_t(T) < G)
    {
      pu = redu[4 * T];
      pp = redp[4 * T];
      for(int i = 1; i < 4; ++i)
      {
        pu += redu[4 * T + i];
        pp += redp[4 * T + i];
      }
      pu = 0.0 + pu;
      pp = 0.0 + pp;
    }
    pu = wave_sum(pu);
    pp = wave_sum(pp);
    if(T == 0)
    {
      for(int i = 1; i < 4; ++i)
      {
        pu += 0.0;
        pp += 0.0;
      }
      const double a = sqrt(pu), c = sqrt(pp);
      *out = sqrt(dfma(a, a, c * c));
    }
  }
  __syncthreads();
}

// one AmaVanka multigrid smoother application (StokesMultiGrid::smooth)
__device__ __forceinline__ void st_smooth(const StLevel& L, int steps, double omega, double* sol, const double* rhs, double* rdef,
                          double* rcor, StStage& S, double* red, double* red2, double* scal)
{
  const uint64_t n = 2 * L.K.nv + L.K.np;
  for(uint64_t i = threadIdx.x; i < n; i += kSt) sol[i] = 0.0;
  __syncthreads();
  st_norm(rhs, 2 * L.K.nv, L.K.np, red, red2, scal);
  const bool open = fabs(omega) >= DBL_EPSILON && gate_open(scal);
  for(int s = 0; s < steps; ++s)
  {
    const double* d = (s == 0) ? rhs : rdef;
    st_tuple(L.V, rcor, d, nullptr, 1.0, L.vmask, S);
    mean_passes_block<kSt>(rcor + 2 * L.K.nv, L.dual, L.prim, L.vol, L.K.np, 2, 0, red, scal + 1);
    if(open)
    {
      for(uint64_t i = threadIdx.x; i < n; i += kSt) sol[i] = dfma(omega, rcor[i], sol[i]);
      __syncthreads();
    }
    if(s + 1 < steps)
    {
      st_tuple(L.K, rdef, sol, rhs, -1.0, L.vmask, S);
      mean_passes_block<kSt>(rdef + 2 * L.K.nv, L.prim, L.dual, L.vol, L.K.np, 1, 0, red, scal + 1);
    }
  }
}

__device__ __forceinline__ void st_defect(const StLevel& L, double* d, const double* x, const double* b, StStage& S, double* red,
                          double* scal)
{
  st_tuple(L.K, d, x, b, -1.0, L.vmask, S);
  mean_passes_block<kSt>(d + 2 * L.K.nv, L.prim, L.dual, L.vol, L.K.np, 1, 0, red, scal + 1);
}

// UnitFilterBlocked::filter_def on the velocity part
__device__ __forceinline__ void st_unit_def(const StLevel& L, double* v)
{
  if(L.vmask)
    for(uint64_t i = threadIdx.x; i < L.K.nv; i += kSt)
      if(is_masked(L.vmask, i))
      {
        v[2 * i] = 0.0;
        v[2 * i + 1] = 0.0;
      }
  __syncthreads();
}

__device__ __forceinline__ void st_axpy1(double* sol, const double* cor, uint64_t n)
{
  for(uint64_t i = threadIdx.x; i < n; i += kSt) sol[i] = dfma(1.0, cor[i], sol[i]);
  __syncthreads();
}

// The V-cycle over the tail's levels (0 = its finest): rhs of level 0 in, its solution out.
__global__ __launch_bounds__(kSt) void k_stokes_tail(StTail Targ, const double* __restrict__ rhs_in,
                                                     double* __restrict__ sol_out)
{
  // the level descriptors are indexed by level at run time: one copy in LDS
  __shared__ StTail T;
  if(threadIdx.x == 0) T = Targ;
  __syncthreads();
  __shared__ StStage S;
  __shared__ double s_vec[6 * FEAT_STOKES_TAIL_MAX_DOFS];
  __shared__ double red[kMeanThreads / 64], red2[kMeanThreads / 64], scal[4];
  // every level's MeanFilter vectors and velocity mask, read by each filter pass
  __shared__ double s_pd[2 * FEAT_STOKES_TAIL_MAX_PRESSURE];
  __shared__ uint32_t s_mask[FEAT_STOKES_TAIL_MAX_MASK_WORDS];
  {
    uint64_t op = 0, om = 0;
    for(int l = 0; l < T.nl; ++l)
    {
      StLevel& L = T.lv[l];
      const uint64_t np = L.K.np, nw = L.vmask ? (L.K.nv + 31) / 32 : 0;
      for(uint64_t i = threadIdx.x; i < np; i += kSt)
      {
        s_pd[op + i] = L.prim[i];
        s_pd[op + np + i] = L.dual[i];
      }
      for(uint64_t i = threadIdx.x; i < nw; i += kSt) s_mask[om + i] = L.vmask[i];
      __syncthreads();
      if(threadIdx.x == 0)
      {
        L.prim = s_pd + op;
        L.dual = s_pd + op + np;
        if(L.vmask) L.vmask = s_mask + om;
      }
      op += 2 * np;
      om += nw;
    }
  }
  // per level six vectors: rhs, sol, def, cor, rdef, rcor
  __syncthreads();
  enum { RHS = 0, SOL, DEF, COR, RDEF, RCOR };
  auto vec = [&](int l, int k) -> double* {
    uint64_t o = 0;
#pragma unroll
    for(int i = 0; i < kStMaxLevels; ++i)
      if(i < l) o += 6 * (2 * T.lv[i].K.nv + T.lv[i].K.np);
    return s_vec + o + uint64_t(k) * (2 * T.lv[l].K.nv + T.lv[l].K.np);
  };
  const int nl = T.nl;
  {
    const uint64_t n0 = 2 * T.lv[0].K.nv + T.lv[0].K.np;
    for(uint64_t i = threadIdx.x; i < n0; i += kSt) vec(0, RHS)[i] = rhs_in[i];
    __syncthreads();
  }
  for(int l = 0; l + 1 < nl; ++l)
  {
    const StLevel& L = T.lv[l];
    const StLevel& C = T.lv[l + 1];
    st_smooth(L, T.steps, T.omega, vec(l, SOL), vec(l, RHS), vec(l, RDEF), vec(l, RCOR), S, red, red2, scal);
    st_defect(L, vec(l, DEF), vec(l, SOL), vec(l, RHS), S, red, scal);
    st_transfer(L.rv, L.rp, vec(l, DEF), L.K.nv, vec(l + 1, RHS), S);
    st_unit_def(C, vec(l + 1, RHS));
    mean_passes_block<kSt>(vec(l + 1, RHS) + 2 * C.K.nv, C.prim, C.dual, C.vol, C.K.np, 1, 0, red, scal + 1);
  }
  {
    const StLevel& Lc = T.lv[nl - 1];
    st_tuple(Lc.V, vec(nl - 1, SOL), vec(nl - 1, RHS), nullptr, 1.0, Lc.vmask, S);
    mean_passes_block<kSt>(vec(nl - 1, SOL) + 2 * Lc.K.nv, Lc.dual, Lc.prim, Lc.vol, Lc.K.np, 2, 0, red, scal + 1);
  }
  for(int l = nl - 2; l >= 0; --l)
  {
    const StLevel& L = T.lv[l];
    const StLevel& C = T.lv[l + 1];
    const uint64_t n = 2 * L.K.nv + L.K.np;
    st_transfer(L.pv, L.pp, vec(l + 1, SOL), C.K.nv, vec(l, COR), S);
    st_unit_def(L, vec(l, COR));
    mean_passes_block<kSt>(vec(l, COR) + 2 * L.K.nv, L.dual, L.prim, L.vol, L.K.np, 1, 0, red, scal + 1);
    st_axpy1(vec(l, SOL), vec(l, COR), n);
    st_defect(L, vec(l, DEF), vec(l, SOL), vec(l, RHS), S, red, scal);
    st_smooth(L, T.steps, T.omega, vec(l, COR), vec(l, DEF), vec(l, RDEF), vec(l, RCOR), S, red, red2, scal);
    st_axpy1(vec(l, SOL), vec(l, COR), n);
  }
  const uint64_t n0 = 2 * T.lv[0].K.nv + T.lv[0].K.np;
  for(uint64_t i = threadIdx.x; i < n0; i += kSt) sol_out[i] = vec(0, SOL)[i];
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
  const uint64_t nvb = (M->nv + kTupRows - 1) / kTupRows, npb = (M->np + kTupRows - 1) / kTupRows;
  if(nvb + npb > uint64_t(INT32_MAX)) return fail(FEAT_ERR_UNSUPPORTED, "feat_tuple_apply_f64: too many rows");
  hipLaunchKernelGGL(k_tuple_apply<false>, dim3(unsigned(nvb + npb)), dim3(kThreads), 0, h->stream, to_dev(M), r, x, y,
                     alpha, vmask, uint32_t(nvb), TuplePost{});
  return check_launch("feat_tuple_apply_f64");
}

int feat_tuple_apply_post_f64(feat_handle_t h, const feat_tuple_f64_t* M, double* r, const double* x, const double* y,
                              double alpha, const uint32_t* vmask, const double* w_dot, const double* w_add,
                              double volume, int passes, int sol_form, double* x_upd, double omega,
                              const double* gate_norm, int x_upd_zero, const feat_reduce_ws_t* ws)
{
  if(!h || !tuple_ok(M, 0, 0) || !tuple_ok(M, 2, 2) || !r || !x)
    return fail(FEAT_ERR_INVALID, "feat_tuple_apply_post_f64: null argument");
  if(M->val[1] && (!M->col[1] || !M->ptr[1])) return fail(FEAT_ERR_INVALID, "feat_tuple_apply_post_f64: block 1");
  if(M->val[3] && (!M->col[3] || !M->ptr[3])) return fail(FEAT_ERR_INVALID, "feat_tuple_apply_post_f64: block 3");
  if(passes < 0 || passes > 2 || (passes && M->np && (!w_dot || !w_add)))
    return fail(FEAT_ERR_INVALID, "feat_tuple_apply_post_f64: mean filter arguments");
  if(!ws || !ws->counter) return fail(FEAT_ERR_INVALID, "feat_tuple_apply_post_f64: workspace required");
  if(r == x) return fail(FEAT_ERR_ALIAS, "feat_tuple_apply_post_f64: r must not alias x");
  if(x_upd && (x_upd == x || x_upd == r || x_upd == y))
    return fail(FEAT_ERR_ALIAS, "feat_tuple_apply_post_f64: x_upd must not alias x, r or y");
  const uint64_t rows = M->nv + M->np;
  if(!rows) return FEAT_OK;
  const uint64_t nvb = (M->nv + kTupRows - 1) / kTupRows, npb = (M->np + kTupRows - 1) / kTupRows;
  if(nvb + npb > uint64_t(INT32_MAX)) return fail(FEAT_ERR_UNSUPPORTED, "feat_tuple_apply_post_f64: too many rows");
  const int tail = (npb && M->np <= kTupTailMaxP) ? 1 : 0;
  const int xz = (x_upd && x_upd_zero) ? 1 : 0;
  const TuplePost post{w_dot, w_add, volume, passes, sol_form, x_upd, omega, gate_norm, ws->counter, tail, xz};
  hipLaunchKernelGGL(k_tuple_apply<true>, dim3(unsigned(nvb + npb)), dim3(kThreads), 0, h->stream, to_dev(M), r, x, y,
                     alpha, vmask, uint32_t(nvb), post);
  if(npb && !tail && (passes || x_upd))
    hipLaunchKernelGGL(k_mean_passes, dim3(1), dim3(kMeanThreads), 0, h->stream, r + 2 * M->nv, w_dot, w_add, volume,
                       M->np, passes, sol_form, x_upd ? x_upd + 2 * M->nv : nullptr, omega, gate_norm, xz);
  return check_launch("feat_tuple_apply_post_f64");
}

int feat_stokes_transfer_post_f64(feat_handle_t h, const feat_csr_f64_t* V, const feat_csr_f64_t* P, const double* x,
                                  uint64_t nvx, double* r, const uint32_t* vmask, const double* w_dot,
                                  const double* w_add, double volume, int passes, double* x_upd,
                                  const feat_reduce_ws_t* ws)
{
  if(!h || !V || !P || !x || !r || (V->rows && (!V->rowptr || (V->nnz && (!V->val || !V->col)))) ||
     (P->rows && (!P->rowptr || (P->nnz && (!P->val || !P->col)))))
    return fail(FEAT_ERR_INVALID, "feat_stokes_transfer_post_f64: null argument");
  if(V->cols != nvx) return fail(FEAT_ERR_INVALID, "feat_stokes_transfer_post_f64: V columns != nvx");
  if(passes < 0 || passes > 2 || (passes && P->rows && (!w_dot || !w_add)))
    return fail(FEAT_ERR_INVALID, "feat_stokes_transfer_post_f64: mean filter arguments");
  if(!ws || !ws->counter) return fail(FEAT_ERR_INVALID, "feat_stokes_transfer_post_f64: workspace required");
  if(r == x || (x_upd && (x_upd == x || x_upd == r)))
    return fail(FEAT_ERR_ALIAS, "feat_stokes_transfer_post_f64: r, x and x_upd must be distinct");
  const uint64_t nvb = (V->rows + kTrRows - 1) / kTrRows, npb = (P->rows + kTrRows - 1) / kTrRows;
  if(!nvb && !npb) return FEAT_OK;
  if(nvb + npb > uint64_t(INT32_MAX)) return fail(FEAT_ERR_UNSUPPORTED, "feat_stokes_transfer_post_f64: too many rows");
  const int tail = (npb && P->rows <= kTupTailMaxP) ? 1 : 0;
  const TuplePost post{w_dot, w_add, volume, passes, 0, x_upd, 1.0, nullptr, ws->counter, tail, 0};
  const StCsr v{V->val, V->col, V->rowptr, V->rows}, p{P->val, P->col, P->rowptr, P->rows};
  hipLaunchKernelGGL(k_stokes_transfer<true>, dim3(unsigned(nvb + npb)), dim3(kThreads), 0, h->stream, v, p, x, nvx, r,
                     vmask, uint32_t(nvb), post);
  if(npb && !tail && (passes || x_upd))
    hipLaunchKernelGGL(k_mean_passes, dim3(1), dim3(kMeanThreads), 0, h->stream, r + 2 * V->rows, w_dot, w_add, volume,
                       P->rows, passes, 0, x_upd ? x_upd + 2 * V->rows : nullptr, 1.0, nullptr, 0);
  return check_launch("feat_stokes_transfer_post_f64");
}

int feat_mean_filter_passes_f64(feat_handle_t h, double* v, const double* w_dot, const double* w_add,
                                double volume, uint64_t n, int passes, int sol_form)
{
  if(!h || (n && (!v || !w_dot || !w_add)) || passes < 0 || passes > 2)
    return fail(FEAT_ERR_INVALID, "feat_mean_filter_passes_f64: bad argument");
  if(!n || !passes) return FEAT_OK;
  hipLaunchKernelGGL(k_mean_passes, dim3(1), dim3(kMeanThreads), 0, h->stream, v, w_dot, w_add, volume, n, passes,
                     sol_form, nullptr, 0.0, nullptr, 0);
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
  // the norm is formed by the workgroup that completes the sums, unless they are deferred
  hipLaunchKernelGGL(k_sumsq2, dim3(grid), dim3(kThreads), 0, h->stream, v, nu, np, gd, gd.defer ? nullptr : out);
  if(gd.defer)
  {
    finish_reduce(h->stream, gd, uint64_t(grid));
    hipLaunchKernelGGL(k_tuple_norm_finish, dim3(1), dim3(1), 0, h->stream, sums, out);
  }
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

int feat_stokes_tail_f64(feat_handle_t h, const feat_stokes_tail_t* t, const double* rhs, double* sol)
{
  if(!h || !t || !rhs || !sol) return fail(FEAT_ERR_INVALID, "feat_stokes_tail_f64: null argument");
  if(t->nlevels < 1 || t->nlevels > FEAT_STOKES_TAIL_MAX_LEVELS || t->steps < 1)
    return fail(FEAT_ERR_INVALID, "feat_stokes_tail_f64: 1..FEAT_STOKES_TAIL_MAX_LEVELS levels, steps >= 1");
  StTail T{};
  T.nl = t->nlevels;
  T.steps = t->steps;
  T.omega = t->omega;
  uint64_t dofs = 0;
  for(int l = 0; l < t->nlevels; ++l)
  {
    const feat_stokes_tail_level_t& L = t->lv[l];
    if(!tuple_ok(&L.K, 0, 2) || !tuple_ok(&L.V, 0, 3) || (L.K.val[3] && (!L.K.col[3] || !L.K.ptr[3])))
      return fail(FEAT_ERR_INVALID, "feat_stokes_tail_f64: matrix blocks");
    if(L.V.nv != L.K.nv || L.V.np != L.K.np || (L.K.np && (!L.prim || !L.dual)))
      return fail(FEAT_ERR_INVALID, "feat_stokes_tail_f64: level sizes");
    dofs += 2 * L.K.nv + L.K.np;
    if(l + 1 < t->nlevels)
    {
      const feat_stokes_tail_level_t& C = t->lv[l + 1];
      const feat_csr_f64_t* m[4] = {&L.rv, &L.rp, &L.pv, &L.pp};
      for(const feat_csr_f64_t* c : m)
        if(!c->val || !c->col || !c->rowptr) return fail(FEAT_ERR_INVALID, "feat_stokes_tail_f64: transfer");
      if(L.rv.rows != C.K.nv || L.rv.cols != L.K.nv || L.rp.rows != C.K.np || L.rp.cols != L.K.np ||
         L.pv.rows != L.K.nv || L.pv.cols != C.K.nv || L.pp.rows != L.K.np || L.pp.cols != C.K.np)
        return fail(FEAT_ERR_INVALID, "feat_stokes_tail_f64: transfer shapes");
    }
    StLevel& D = T.lv[l];
    D.K = to_dev(&L.K);
    D.V = to_dev(&L.V);
    D.vmask = L.vmask;
    D.prim = L.prim;
    D.dual = L.dual;
    D.vol = L.volume;
    D.rv = StCsr{L.rv.val, L.rv.col, L.rv.rowptr, L.rv.rows};
    D.rp = StCsr{L.rp.val, L.rp.col, L.rp.rowptr, L.rp.rows};
    D.pv = StCsr{L.pv.val, L.pv.col, L.pv.rowptr, L.pv.rows};
    D.pp = StCsr{L.pp.val, L.pp.col, L.pp.rowptr, L.pp.rows};
  }
  static_assert(FEAT_STOKES_TAIL_MAX_DOFS <= 2048 * (kSt / 256), "the norm's workgroups");
  if(dofs > FEAT_STOKES_TAIL_MAX_DOFS)
    return fail(FEAT_ERR_UNSUPPORTED, "feat_stokes_tail_f64: more than FEAT_STOKES_TAIL_MAX_DOFS unknowns");
  uint64_t pres = 0, words = 0;
  for(int l = 0; l < t->nlevels; ++l)
  {
    pres += t->lv[l].K.np;
    words += t->lv[l].vmask ? (t->lv[l].K.nv + 31) / 32 : 0;
  }
  if(pres > FEAT_STOKES_TAIL_MAX_PRESSURE || words > FEAT_STOKES_TAIL_MAX_MASK_WORDS)
    return fail(FEAT_ERR_UNSUPPORTED, "feat_stokes_tail_f64: pressure / filter sizes above the LDS budget");
  hipLaunchKernelGGL(k_stokes_tail, dim3(1), dim3(kSt), 0, h->stream, T, rhs, sol);
  return check_launch("feat_stokes_tail_f64");
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
