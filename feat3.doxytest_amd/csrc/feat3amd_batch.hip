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

__device__ __forceinline__ void chunk_partial(const BatchDev& b, double v)
{
  __shared__ double red[kThreads / 64];
  const double t = block_sum(v, red);
  if(threadIdx.x == 0) b.partials[blockIdx.x] = t;
}

__global__ __launch_bounds__(kThreads) void k_batch_dot(BatchDev b, const double* __restrict__ x,
                                                        const double* __restrict__ y)
{
  const uint64_t lo = b.chunk_lo[blockIdx.x], hi = b.chunk_hi[blockIdx.x];
  double s = 0.0;
  for(uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) s = dfma(x[i], y[i], s);
  chunk_partial(b, s);
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

__global__ __launch_bounds__(kThreads) void k_batch_update(BatchDev b, double* __restrict__ x, double* __restrict__ r,
                                                           const double* __restrict__ p, const double* __restrict__ t,
                                                           const double* __restrict__ gamma,
                                                           const double* __restrict__ tp, const int* __restrict__ active)
{
  const uint32_t inst = b.chunk_inst[blockIdx.x];
  const uint64_t lo = b.chunk_lo[blockIdx.x], hi = b.chunk_hi[blockIdx.x];
  double s = 0.0;
  if(active[inst])
  {
    const double alpha = gamma[inst] / tp[inst];
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
  chunk_partial(b, s);
}

__global__ __launch_bounds__(kThreads) void k_batch_direction(BatchDev b, double* __restrict__ p,
                                                              const double* __restrict__ z,
                                                              const double* __restrict__ gnew,
                                                              const double* __restrict__ gold,
                                                              const int* __restrict__ active, int first)
{
  const uint32_t inst = b.chunk_inst[blockIdx.x];
  if(!active[inst]) return;
  const uint64_t lo = b.chunk_lo[blockIdx.x], hi = b.chunk_hi[blockIdx.x];
  const double beta = first ? 0.0 : gnew[inst] / gold[inst];
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

__global__ void k_batch_status(uint32_t n, const double* __restrict__ rr, StopDev c, feat_batch_state_t st, int initial)
{
  __shared__ int cnt[kThreads / 64];
  int mine = 0;
  for(uint32_t i = threadIdx.x; i < n; i += blockDim.x)
  {
    const double def = sqrt(rr[i]);
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
  hipLaunchKernelGGL(k_batch_dot, dim3(b->nchunks), dim3(kThreads), 0, h->stream, to_dev(b), x, y);
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
                     active);
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
                     gamma_old, active, first);
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
  hipLaunchKernelGGL(k_batch_status, dim3(1), dim3(kThreads), 0, h->stream, ninst, rr, c, *st, initial);
  return check_launch("feat_batch_status_f64");
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

int feat_copy_int_to_host(feat_handle_t h, const int* dev, int* host_pinned)
{
  if(!h || !dev || !host_pinned) return fail(FEAT_ERR_INVALID, "feat_copy_int_to_host: null argument");
  return check_hip(hipMemcpyAsync(host_pinned, dev, sizeof(int), hipMemcpyDeviceToHost, h->stream), "hipMemcpyAsync");
}

} // extern "C"
