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
      const uint64_t lo0 = ra0 > w0 ? ra0 : w0, h0 = w0 + kTupE0, hi0 = re0 < h0 ? re0 : h0;
      for(uint64_t k = lo0; k < hi0; ++k)
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
      for(uint64_t k = lo1; k < hi1; ++k)
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
constexpr int kUpdBatch = 16;    // elements per thread per batch of an update loop

// v[i] = a * w[i] + v[i] over [0, n) by the NT threads of the workgroup, kUpdBatch loads of
// each thread in flight
template<int NT>
__device__ __forceinline__ void axpy_block(double* __restrict__ v, double a, const double* __restrict__ w, uint64_t n)
{
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
__device__ void mean_passes_block(double* __restrict__ v, const double* __restrict__ w_dot,
                                  const double* __restrict__ w_add, double volume, uint64_t n, int passes,
                                  int sol_form, double* red /* kMeanThreads / 64 */, double* s_alpha)
{
  static_assert(NT == 256 || NT == 1024, "workgroup size");
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
      double dot = red[0];
      for(int i = 1; i < kMeanThreads / 64; ++i) dot += red[i];
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
};

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
  const bool upd = POST && post.x_upd != nullptr && vel && fabs(post.omega) >= DBL_EPSILON && gate_open(post.gate);
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
      if(upd)
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
        post.x_upd[2 * row] = dfma(post.omega, r0, u0);
        post.x_upd[2 * row + 1] = dfma(post.omega, r1, u1);
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
    __shared__ double red[kMeanThreads / 64];
    __shared__ double s_alpha;
    __shared__ int last;
    // publish this workgroup's pressure rows (agent-scope release by every writer), then count in
    __threadfence();
    __syncthreads();
    if(threadIdx.x == 0)
    {
      const uint32_t npb = gridDim.x - nvb;
      const uint32_t prev = __hip_atomic_fetch_add(post.counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      last = (prev == npb - 1) ? 1 : 0;
    }
    __syncthreads();
    if(!last) return;
    __threadfence();   // acquire: the other workgroups' rows
    double* rp = r + 2 * M.nv;
    mean_passes_block<kThreads>(rp, post.w_dot, post.w_add, post.volume, M.np, post.passes, post.sol_form, red,
                                &s_alpha);
    if(post.x_upd && fabs(post.omega) >= DBL_EPSILON && gate_open(post.gate))
      axpy_block<kThreads>(post.x_upd + 2 * M.nv, post.omega, rp, M.np);
    if(threadIdx.x == 0) __hip_atomic_store(post.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
                                                              const double* __restrict__ gate)
{
  __shared__ double red[kMeanThreads / 64];
  __shared__ double s_alpha;
  mean_passes_block<kMeanThreads>(v, w_dot, w_add, volume, n, passes, sol_form, red, &s_alpha);
  if(x_upd && fabs(omega) >= DBL_EPSILON && gate_open(gate)) axpy_block<kMeanThreads>(x_upd, omega, v, n);
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
  const uint64_t nvb = (M->nv + kTupRows - 1) / kTupRows, npb = (M->np + kTupRows - 1) / kTupRows;
  if(nvb + npb > uint64_t(INT32_MAX)) return fail(FEAT_ERR_UNSUPPORTED, "feat_tuple_apply_f64: too many rows");
  hipLaunchKernelGGL(k_tuple_apply<false>, dim3(unsigned(nvb + npb)), dim3(kThreads), 0, h->stream, to_dev(M), r, x, y,
                     alpha, vmask, uint32_t(nvb), TuplePost{});
  return check_launch("feat_tuple_apply_f64");
}

int feat_tuple_apply_post_f64(feat_handle_t h, const feat_tuple_f64_t* M, double* r, const double* x, const double* y,
                              double alpha, const uint32_t* vmask, const double* w_dot, const double* w_add,
                              double volume, int passes, int sol_form, double* x_upd, double omega,
                              const double* gate_norm, const feat_reduce_ws_t* ws)
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
  const int tail = M->np <= kTupTailMaxP ? 1 : 0;
  const TuplePost post{w_dot, w_add, volume, passes, sol_form, x_upd, omega, gate_norm, ws->counter, tail};
  if(npb && tail)
    hipLaunchKernelGGL(k_tuple_apply<true>, dim3(unsigned(nvb + npb)), dim3(kThreads), 0, h->stream, to_dev(M), r, x,
                       y, alpha, vmask, uint32_t(nvb), post);
  else
    hipLaunchKernelGGL(k_tuple_apply<true>, dim3(unsigned(nvb + npb)), dim3(kThreads), 0, h->stream, to_dev(M), r, x,
                       y, alpha, vmask, uint32_t(nvb), TuplePost{w_dot, w_add, volume, passes, sol_form, x_upd, omega,
                                                                  gate_norm, ws->counter, 0});
  if(npb && !tail && (passes || x_upd))
    hipLaunchKernelGGL(k_mean_passes, dim3(1), dim3(kMeanThreads), 0, h->stream, r + 2 * M->nv, w_dot, w_add, volume,
                       M->np, passes, sol_form, x_upd ? x_upd + 2 * M->nv : nullptr, omega, gate_norm);
  return check_launch("feat_tuple_apply_post_f64");
}

int feat_mean_filter_passes_f64(feat_handle_t h, double* v, const double* w_dot, const double* w_add,
                                double volume, uint64_t n, int passes, int sol_form)
{
  if(!h || (n && (!v || !w_dot || !w_add)) || passes < 0 || passes > 2)
    return fail(FEAT_ERR_INVALID, "feat_mean_filter_passes_f64: bad argument");
  if(!n || !passes) return FEAT_OK;
  hipLaunchKernelGGL(k_mean_passes, dim3(1), dim3(kMeanThreads), 0, h->stream, v, w_dot, w_add, volume, n, passes,
                     sol_form, nullptr, 0.0, nullptr);
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
