// feat3amd_common.hpp -- definitions shared by the HIP translation units of
// libfeat3amd.so: launch constants, error helpers, the handle, and the
// wavefront/workgroup/grid reduction helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <string>
#include <type_traits>

#include "feat3amd.h"

namespace feat3amd_detail {
extern thread_local std::string g_err;
int fail(int code, const char* what);
int check_hip(hipError_t e, const char* where);
int check_launch(const char* where);
// the transposed CSR apply in the reference's sequential order (feat3amd_transpose.hip)
template<typename DT, typename IT>
int csr_apply_transposed(feat_handle_t h, DT* r, DT a, const DT* x, DT b, bool bzero, const DT* y, const DT* val,
                         const IT* col, const IT* rowptr, uint64_t rows, uint64_t cols, uint64_t nnz, const char* name);
// the same with every buffer in the per-device scratch array (feat3amd_transpose.hip)
template<typename DT, typename IT>
int csr_apply_transposed_noalloc(feat_handle_t h, DT* r, DT a, const DT* x, DT b, bool bzero, const DT* y,
                                 const DT* val, const IT* col, const IT* rowptr, uint64_t rows, uint64_t cols,
                                 uint64_t nnz, const char* name);
// looks up this device's copy of that scratch array (once per device; the device is current)
int noalloc_scratch_attach(int device);
} // namespace feat3amd_detail
using feat3amd_detail::check_hip;
using feat3amd_detail::check_launch;
using feat3amd_detail::fail;

namespace {

constexpr int kThreads = 256;
constexpr int kXcd = 8;
constexpr int kMaxGrid = FEAT_REDUCE_MAX_BLOCKS;
constexpr int kRowsPerBlock = 256;  // rows per CSR row-block (one lane per row)
constexpr int kCsrUnroll = 4;       // CSR entries each thread loads per pass
constexpr int kChunk = kCsrUnroll * kThreads;   // nonzeros staged in LDS per pass
constexpr int kSellUnroll = 8;      // SELL entries per row in flight


inline int grid_for(uint64_t work, int per_thread = 1)
{
  uint64_t b = (work + uint64_t(kThreads) * per_thread - 1) / (uint64_t(kThreads) * per_thread);
  if(b < 1) b = 1;
  if(b > uint64_t(kMaxGrid)) b = kMaxGrid;
  return int(b);
}

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool is_masked(const uint32_t* mask, uint64_t i)
{
  return mask != nullptr && ((mask[i >> 5] >> (i & 31u)) & 1u);
}

__device__ __forceinline__ double dfma(double a, double b, double c) { return __fma_rn(a, b, c); }
__device__ __forceinline__ float dfma(float a, float b, float c) { return __fmaf_rn(a, b, c); }

// Skip rule of IterativeSolver::_set_initial_defect (iterative.hpp:761-766)
// for a smoother whose initial defect norm is sqrt(nrm2).
template<typename DT>
__device__ __forceinline__ bool smoother_skips(const double* nrm2)
{
  if(nrm2 == nullptr) return false;
  const double d = sqrt(*nrm2);
  const double eps = std::is_same<DT, float>::value ? double(FLT_EPSILON) : DBL_EPSILON;
  return !isfinite(d) || d <= eps * eps;
}

__device__ __forceinline__ double wave_sum(double v)
{
  for(int off = 32; off > 0; off >>= 1)
    v += __shfl_down(v, off, 64);
  return v;
}

// Sum over the workgroup; the result is valid in thread 0.
__device__ double block_sum(double v, double* red /* >= kThreads/64 */)
{
  v = wave_sum(v);
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  if(lane == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  if(threadIdx.x == 0)
  {
    t = red[0];
    for(int i = 1; i < int(blockDim.x >> 6); ++i) t += red[i];
  }
  __syncthreads();
  return t;
}

// Deterministic grid reduction of up to two values: partial per workgroup, summed
// in index order -- by the last workgroup to arrive (small grids), or, with `defer`,
// by a one-workgroup kernel launched right after (large grids: the workgroups then
// store their partials and end without any atomic or wait).  Both sum the same
// way: thread t adds partials t, t + kThreads, ... in order, then a block sum.
struct GridReduce
{
  double* partials;   // 2 * kMaxGrid
  uint32_t* counter;
  double* out0;       // may be null
  double* out1;       // may be null
  int defer;          // partials only; k_grid_finish completes the sum
};

// the in-order sum of nblocks partials (kThreads threads; result to out0/out1)
template<bool ATOMIC>
__device__ void grid_reduce_tail(const GridReduce& g, uint32_t nblocks, double* red)
{
  const bool r0 = g.out0 != nullptr, r1 = g.out1 != nullptr;
  // eight loads are issued before any is added, so a large grid costs few round trips
  double s0 = 0.0, s1 = 0.0;
  for(uint32_t i0 = threadIdx.x; i0 < nblocks; i0 += 8 * blockDim.x)
  {
    double v0[8], v1[8];
#pragma unroll
    for(int u = 0; u < 8; ++u)
    {
      const uint32_t i = i0 + uint32_t(u) * blockDim.x;
      const bool in = i < nblocks;
      if(ATOMIC)
      {
        v0[u] = (r0 && in) ? __hip_atomic_load(&g.partials[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
        v1[u] = (r1 && in) ? __hip_atomic_load(&g.partials[kMaxGrid + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                           : 0.0;
      }
      else
      {
        v0[u] = (r0 && in) ? g.partials[i] : 0.0;
        v1[u] = (r1 && in) ? g.partials[kMaxGrid + i] : 0.0;
      }
    }
#pragma unroll
    for(int u = 0; u < 8; ++u)
      if(i0 + uint32_t(u) * blockDim.x < nblocks)
      {
        s0 += v0[u];
        s1 += v1[u];
      }
  }
  if(r0) s0 = block_sum(s0, red);
  if(r1) s1 = block_sum(s1, red);
  if(threadIdx.x == 0)
  {
    if(r0) *g.out0 = s0;
    if(r1) *g.out1 = s1;
  }
}

// Returns true in the workgroup that completed the sum (its thread 0 wrote out0/out1); false
// everywhere else and when the sum is deferred or nothing is reduced.
__device__ bool grid_reduce_finish(const GridReduce& g, double v0, double v1)
{
  __shared__ double red[kThreads / 64];
  __shared__ int last;
  const bool r0 = g.out0 != nullptr, r1 = g.out1 != nullptr;
  if(!r0 && !r1) return false;
  double b0 = r0 ? block_sum(v0, red) : 0.0;
  double b1 = r1 ? block_sum(v1, red) : 0.0;
  if(g.defer)
  {
    if(threadIdx.x == 0)
    {
      if(r0) g.partials[blockIdx.x] = b0;
      if(r1) g.partials[kMaxGrid + blockIdx.x] = b1;
    }
    return false;
  }
  if(threadIdx.x == 0)
  {
    // Publish the partials with agent-scope atomic exchanges: they are performed
    // at the device's coherence point, so no cache writeback (an agent-scope
    // release fence flushes the XCD's L2 and costs ~50 ns per workgroup).  The
    // returned values are kept live so the exchanges are issued with return,
    // and s_waitcnt(0) waits for them before this workgroup is counted in.
    double o0 = 0.0, o1 = 0.0;
    if(r0) o0 = __hip_atomic_exchange(&g.partials[blockIdx.x], b0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if(r1) o1 = __hip_atomic_exchange(&g.partials[kMaxGrid + blockIdx.x], b1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);
    asm volatile("" ::"v"(o0), "v"(o1));
    const uint32_t prev = __hip_atomic_fetch_add(g.counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    last = (prev == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if(!last) return false;
  grid_reduce_tail<true>(g, gridDim.x, red);
  if(threadIdx.x == 0) __hip_atomic_store(g.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// completes a deferred grid reduction of nblocks partials (one workgroup of kThreads)
__global__ __launch_bounds__(kThreads) void k_grid_finish(GridReduce g, uint32_t nblocks)
{
  __shared__ double red[kThreads / 64];
  grid_reduce_tail<false>(g, nblocks, red);
}

// grids above this many workgroups defer their reduction to k_grid_finish (below it the
// in-kernel finish costs less than the extra launch)
constexpr uint64_t kDeferGrid = 512;

__host__ inline GridReduce reduce_for_grid(GridReduce g, uint64_t grid)
{
  g.defer = (grid > kDeferGrid && (g.out0 != nullptr || g.out1 != nullptr)) ? 1 : 0;
  return g;
}
__host__ inline void finish_reduce(hipStream_t s, const GridReduce& g, uint64_t grid)
{
  if(g.defer) hipLaunchKernelGGL(k_grid_finish, dim3(1), dim3(kThreads), 0, s, g, uint32_t(grid));
}

// Contiguous range of row-blocks for this workgroup's XCD; the workgroups of
// one XCD stride through it.
struct XcdRange
{
  uint64_t first, end, stride;
};

__device__ __forceinline__ XcdRange xcd_range(uint64_t nitems)
{
  const uint32_t g = gridDim.x, b = blockIdx.x;
  const uint32_t x = b % kXcd, j = b / kXcd;
  const uint32_t q = g / kXcd, rem = g % kXcd;
  const uint32_t gx = q + (x < rem ? 1u : 0u);
  const uint64_t pre = uint64_t(x) * q + (x < rem ? x : rem);
  XcdRange r;
  const uint64_t begin = nitems * pre / g;
  r.end = nitems * (pre + gx) / g;
  r.first = begin + j;
  r.stride = gx;
  return r;
}

} // namespace

// ===========================================================================
// handle
// ===========================================================================
struct feat_graph_s
{
  hipGraph_t graph;
  hipGraphExec_t exec;
};

struct feat_handle_s
{
  int device;
  hipStream_t stream;
  bool own_stream;
  hipEvent_t event;   // for feat_handle_wait
  double* partials;   // 2 * kMaxGrid
  uint32_t* counter;
  double* scal;       // scratch device scalars
  double* host;       // pinned host scalars
};

namespace {
inline GridReduce handle_reduce(feat_handle_t h, double* out0, double* out1)
{
  GridReduce g;
  g.partials = h->partials;
  g.counter = h->counter;
  g.out0 = out0;
  g.out1 = out1;
  g.defer = 0;
  return g;
}

inline GridReduce ws_reduce(const feat_reduce_ws_t* ws, double* out0, double* out1)
{
  GridReduce g;
  g.partials = ws ? ws->partials : nullptr;
  g.counter = ws ? ws->counter : nullptr;
  g.out0 = out0;
  g.out1 = out1;
  g.defer = 0;
  return g;
}

inline bool need_ws(const feat_reduce_ws_t* ws, const void* a, const void* b = nullptr)
{
  return (a != nullptr || b != nullptr) && (ws == nullptr || ws->partials == nullptr || ws->counter == nullptr);
}
} // namespace

