// feat3amd_kernels.hip -- gfx950 kernels and the C-ABI of include/feat3amd.h.
//
// Design (MI355X / CDNA4):
//  * CSR SpMV: rows are processed in row-blocks of kRowsPerBlock (256, one lane
//    per row).  A 256-thread workgroup streams the contiguous nonzero range of its
//    row-block through LDS in passes of kChunk entries (val and the gathered
//    x[col] side by side, fully coalesced 64-wide loads, all of a pass in flight
//    before the first store), then each lane sums its row's entries from LDS in
//    storage order with one fma per entry.  The per-row arithmetic is therefore exactly the
//    reference's sequential loop (arch/apply_generic.hpp:57-66), bit for bit,
//    while HBM sees only coalesced streams.  Fused epilogues apply the unit
//    filter (bit mask), the defect/Richardson/Jacobi updates and reductions.
//  * Grids are capped (kMaxGrid) and row-blocks are assigned per XCD in
//    contiguous ranges: a workgroup with blockIdx b runs on XCD b % 8, so the
//    x-gathers of neighbouring row-blocks share that XCD's L2.
//  * Reductions are deterministic: per-workgroup partials (fixed mapping) and a
//    last-workgroup-done finish that sums the partials in index order.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <utility>
#include <vector>
#include <string>
#include <type_traits>

#include "feat3amd_common.hpp"

namespace feat3amd_detail {

thread_local std::string g_err;

int fail(int code, const char* what)
{
  g_err = what;
  return code;
}

int check_hip(hipError_t e, const char* where)
{
  if(e == hipSuccess) return FEAT_OK;
  g_err = std::string(where) + ": " + hipGetErrorString(e);
  return FEAT_ERR_HIP;
}

int check_launch(const char* where)
{
  return check_hip(hipGetLastError(), where);
}

} // namespace feat3amd_detail

namespace {

// EpiPre<E>::on: E loads its row operands through E::pre, ahead of the row sum
template<class E, class = void>
struct EpiPre
{
  struct type
  {
  };
  static constexpr bool on = false;
};
template<class E>
struct EpiPre<E, std::void_t<typename E::Pre>>
{
  typedef typename E::Pre type;
  static constexpr bool on = true;
};

// ---------------------------------------------------------------------------
// CSR row-block kernel with a fused epilogue.
//   Epi::row(row, sum, acc0, acc1) is called once per row by its lane.
// ---------------------------------------------------------------------------
template<typename DT, typename IT, class Epi>
__global__ __launch_bounds__(kThreads) void k_csr_rows(const DT* __restrict__ val,
                                                       const IT* __restrict__ col,
                                                       const IT* __restrict__ rowptr,
                                                       const DT* __restrict__ x, uint64_t rows,
                                                       Epi epi, GridReduce gr)
{
  static_assert(kChunk == kCsrUnroll * kThreads, "one pass = kCsrUnroll entries per thread");
  __shared__ DT s_val[kChunk];
  __shared__ DT s_x[kChunk];
  const uint64_t nrb = (rows + kRowsPerBlock - 1) / kRowsPerBlock;
  const XcdRange xr = xcd_range(nrb);
  const int t = threadIdx.x;
  double acc0 = 0.0, acc1 = 0.0;
  epi.begin();
  for(uint64_t rb = xr.first; rb < xr.end; rb += xr.stride)
  {
    const uint64_t r0 = rb * kRowsPerBlock;
    const uint64_t nr = (rows - r0) < uint64_t(kRowsPerBlock) ? (rows - r0) : uint64_t(kRowsPerBlock);
    const uint64_t p0 = uint64_t(rowptr[r0]);
    const uint64_t p1 = uint64_t(rowptr[r0 + nr]);
    const bool mine = uint64_t(t) < nr;
    uint64_t rs = 0, re = 0;
    typename EpiPre<Epi>::type pf{};
    if(mine)
    {
      rs = uint64_t(rowptr[r0 + t]);
      re = uint64_t(rowptr[r0 + t + 1]);
      if constexpr(EpiPre<Epi>::on) pf = epi.pre(r0 + t);
    }
    DT sum = DT(0);
    for(uint64_t c0 = p0; c0 < p1; c0 += kChunk)
    {
      const uint64_t cn = (p1 - c0) < uint64_t(kChunk) ? (p1 - c0) : uint64_t(kChunk);
      // every value and column of the pass is requested before the first gather of x,
      // and every gather before the first LDS store (the matrix is streamed once:
      // nontemporal loads keep x resident in L2)
      DT v[kCsrUnroll], xv[kCsrUnroll];
      uint64_t cc[kCsrUnroll];
#pragma unroll
      for(int u = 0; u < kCsrUnroll; ++u)
      {
        const uint64_t j = uint64_t(t) + uint64_t(u) * kThreads;
        v[u] = j < cn ? __builtin_nontemporal_load(val + c0 + j) : DT(0);
        cc[u] = j < cn ? uint64_t(__builtin_nontemporal_load(col + c0 + j)) : 0;
      }
#pragma unroll
      for(int u = 0; u < kCsrUnroll; ++u)
        xv[u] = uint64_t(t) + uint64_t(u) * kThreads < cn ? x[cc[u]] : DT(0);
#pragma unroll
      for(int u = 0; u < kCsrUnroll; ++u)
      {
        s_val[t + u * kThreads] = v[u];
        s_x[t + u * kThreads] = xv[u];
      }
      __syncthreads();
      if(mine)
      {
        const uint64_t kb = rs > c0 ? rs : c0;
        const uint64_t ke = re < c0 + cn ? re : c0 + cn;
        for(uint64_t k = kb; k < ke; ++k)
          sum = dfma(s_val[k - c0], s_x[k - c0], sum);
      }
      __syncthreads();
    }
    if(mine)
    {
      if constexpr(EpiPre<Epi>::on)
        epi.row(r0 + t, sum, pf, acc0, acc1);
      else
        epi.row(r0 + t, sum, acc0, acc1);
    }
  }
  grid_reduce_finish(gr, acc0, acc1);
}

// r = a*sum + b*y  (y not read when |b| < eps)          [arch/apply_generic.hpp:30-66]
template<typename DT>
struct EpiApply
{
  DT* r;
  const DT* y;
  DT a, b;
  bool bzero;
  __device__ void begin() {}
  // the row's operands, loadable before its sum is formed
  struct Pre
  {
    DT yv;
  };
  __device__ __forceinline__ Pre pre(uint64_t i) const { return Pre{bzero ? DT(0) : y[i]}; }
  __device__ __forceinline__ void row(uint64_t i, DT sum, const Pre& p, double&, double&) const
  {
    r[i] = dfma(sum, a, b * p.yv);
  }
  __device__ __forceinline__ void row(uint64_t i, DT sum, double& a0, double& a1) const { row(i, sum, pre(i), a0, a1); }
};

// The first iterate of the Richardson(Jacobi) smoother that takes the row's value v as its
// right-hand side, formed where v is produced: x1 = filter(omega * (inv_diag * v)), the
// operations of k_smooth_first without its skip rule (the smoother's next step, which has
// the norm, applies it).  Unused when x1 is null.
template<typename DT>
struct FirstIter
{
  const DT* inv;
  DT* x1;
  DT omega;
  __device__ __forceinline__ DT load(uint64_t i) const { return x1 ? inv[i] : DT(0); }
  __device__ __forceinline__ void store(uint64_t i, DT v, DT inv_i, bool m) const
  {
    if(!x1) return;
    DT c = inv_i * v;
    if(m) c = DT(0);
    x1[i] = dfma(omega, c, DT(0));
  }
};

// r = filter(A*x) with optional |r|^2 and r.w       [transfer.hpp:247-251 + unit_filter]
template<typename DT>
struct EpiMasked
{
  DT* r;
  const uint32_t* mask;
  const DT* dotw;
  bool nrm, dot;
  FirstIter<DT> first;   // the pre-smoother's first iterate on r (restriction)
  __device__ void begin() {}
  // the row's operands, loadable before its sum is formed
  struct Pre
  {
    DT w, inv;
    bool m;
  };
  __device__ __forceinline__ Pre pre(uint64_t i) const
  {
    return Pre{dot ? dotw[i] : DT(0), first.load(i), is_masked(mask, i)};
  }
  __device__ __forceinline__ void row(uint64_t i, DT sum, const Pre& p, double& a0, double& a1) const
  {
    DT v = dfma(sum, DT(1), DT(0));
    if(p.m) v = DT(0);
    r[i] = v;
    first.store(i, v, p.inv, p.m);
    if(nrm) a0 = fma(double(v), double(v), a0);
    if(dot) a1 = fma(double(v), double(p.w), a1);
  }
  __device__ __forceinline__ void row(uint64_t i, DT sum, double& a0, double& a1) const { row(i, sum, pre(i), a0, a1); }
};

// d = filter(rhs - A*x) with optional |d|^2          [sparse_matrix_csr.hpp:1807-1843]
template<typename DT>
struct EpiDefect
{
  DT* d;
  const DT* rhs;
  const uint32_t* mask;
  bool nrm;
  FirstIter<DT> first;   // the post-smoother's first iterate on d
  __device__ void begin() {}
  // the row's operands, loadable before its sum is formed
  struct Pre
  {
    DT rhs, inv;
    bool m;
  };
  __device__ __forceinline__ Pre pre(uint64_t i) const { return Pre{rhs[i], first.load(i), is_masked(mask, i)}; }
  __device__ __forceinline__ void row(uint64_t i, DT sum, const Pre& p, double& a0, double&) const
  {
    DT v = dfma(sum, DT(-1), p.rhs);
    if(p.m) v = DT(0);
    d[i] = v;
    first.store(i, v, p.inv, p.m);
    if(nrm) a0 = fma(double(v), double(v), a0);
  }
  __device__ __forceinline__ void row(uint64_t i, DT sum, double& a0, double& a1) const { row(i, sum, pre(i), a0, a1); }
};

// sol += filter(P*xc)                                  [multigrid.hpp:1885-1930]
template<typename DT>
struct EpiProlAdd
{
  DT* sol;
  const uint32_t* mask;
  __device__ void begin() {}
  // the row's operands, loadable before its sum is formed (sol[i] is read only by its row)
  struct Pre
  {
    DT s;
    bool m;
  };
  __device__ __forceinline__ Pre pre(uint64_t i) const { return Pre{sol[i], is_masked(mask, i)}; }
  __device__ __forceinline__ void row(uint64_t i, DT sum, const Pre& p, double&, double&) const
  {
    DT c = dfma(sum, DT(1), DT(0));
    if(p.m) c = DT(0);
    sol[i] = dfma(DT(1), c, p.s);
  }
  __device__ __forceinline__ void row(uint64_t i, DT sum, double& a0, double& a1) const { row(i, sum, pre(i), a0, a1); }
};

// One Richardson(Jacobi) step, fused (richardson.hpp:205-221, jacobi_precond.hpp:159-164):
//   d = filter(rhs - A*x_k)
//   mode 0 (next):  x_{k+1} = x_k + omega * filter(inv∘d)
//   mode 1 (defect): dout = d
//   mode 2 (acc):    acc += x_k + omega * filter(inv∘d);   optional acc.w
template<typename DT>
struct EpiSmooth
{
  const DT* rhs;
  const DT* inv;
  const uint32_t* mask;
  const DT* xk;
  DT* out;          // x_{k+1} (mode 0), d (mode 1), acc (mode 2)
  const DT* dotw;
  const double* rhs_nrm2;
  DT omega;
  int mode;
  bool skip;
  __device__ void begin() { skip = smoother_skips<DT>(rhs_nrm2); }
  // the row's operands, loadable before its sum is formed (out[i] is read only by its row)
  struct Pre
  {
    DT rhs, inv, xk, acc, w;
    bool m;
  };
  __device__ __forceinline__ Pre pre(uint64_t i) const
  {
    Pre p{rhs[i], DT(0), DT(0), DT(0), DT(0), is_masked(mask, i)};
    if(mode != 1)
    {
      p.inv = inv[i];
      p.xk = xk[i];
    }
    if(mode == 2)
    {
      p.acc = out[i];
      if(dotw) p.w = dotw[i];
    }
    return p;
  }
  __device__ __forceinline__ void row(uint64_t i, DT sum, const Pre& p, double&, double& a1) const
  {
    DT d = dfma(sum, DT(-1), p.rhs);
    if(p.m) d = DT(0);
    if(mode == 1)
    {
      out[i] = d;
      return;
    }
    DT c = p.inv * d;
    if(p.m) c = DT(0);
    DT xn = dfma(omega, c, p.xk);
    if(skip) xn = DT(0);
    if(mode == 0)
      out[i] = xn;
    else
    {
      const DT a = dfma(DT(1), xn, p.acc);
      out[i] = a;
      if(dotw) a1 = a1 + double(a) * double(p.w);
    }
  }
  __device__ __forceinline__ void row(uint64_t i, DT sum, double& a0, double& a1) const { row(i, sum, pre(i), a0, a1); }
};

template<typename DT, typename IT, class Epi>
int launch_rows(hipStream_t s, const DT* val, const IT* col, const IT* rowptr, const DT* x,
                uint64_t rows, const Epi& epi, const GridReduce& gr, const char* name)
{
  if(rows == 0) return FEAT_OK;
  const uint64_t nrb = (rows + kRowsPerBlock - 1) / kRowsPerBlock;
  const int grid = int(nrb < uint64_t(kMaxGrid) ? nrb : uint64_t(kMaxGrid));
  const GridReduce gd = reduce_for_grid(gr, uint64_t(grid));
  hipLaunchKernelGGL((k_csr_rows<DT, IT, Epi>), dim3(grid), dim3(kThreads), 0, s, val, col, rowptr,
                     x, rows, epi, gd);
  finish_reduce(s, gd, uint64_t(grid));
  return check_launch(name);
}

// ---------------------------------------------------------------------------
// SELL-64 engine: one lane per row, coalesced 64-wide loads of val/col, the
// row's entries summed in CSR order (4 entries per step in flight).
// ---------------------------------------------------------------------------
template<bool DELTA, typename CT>
__device__ __forceinline__ uint64_t sell_col(CT c, uint64_t row)
{
  if(DELTA)
    return uint64_t(std::int64_t(row) + std::int64_t(c));
  return uint64_t(c);
}

// stored value -> arithmetic type (uint16_t carries FP16 bits)
template<typename DT, typename VT>
__device__ __forceinline__ DT ld_val(const VT* p)
{
  return DT(__builtin_nontemporal_load(p));
}
template<>
__device__ __forceinline__ float ld_val<float, uint16_t>(const uint16_t* p)
{
  return float(__builtin_bit_cast(_Float16, __builtin_nontemporal_load(p)));
}

// a stored value already loaded -> arithmetic type
template<typename DT, typename VT>
__device__ __forceinline__ DT cvt_val(VT v)
{
  if constexpr(std::is_same<VT, uint16_t>::value)
    return DT(__builtin_bit_cast(_Float16, v));
  else
    return DT(v);
}

// VCODE: val holds 1-byte codes into vtab (256 values, staged in LDS), else the values.
template<typename DT, typename VT, typename CT, bool DELTA, class Epi, int UNR = kSellUnroll, bool VCODE = false>
__global__ __launch_bounds__(kThreads) void k_sell_rows(const VT* __restrict__ val,
                                                        const CT* __restrict__ col,
                                                        const uint64_t* __restrict__ soff,
                                                        const uint16_t* __restrict__ rlen,
                                                        const DT* __restrict__ x, uint64_t rows,
                                                        DT scale, Epi epi, GridReduce gr,
                                                        const DT* __restrict__ vtab = nullptr)
{
  __shared__ DT s_tab[VCODE ? kThreads : 1];
  if constexpr(VCODE)
  {
    static_assert(kThreads == 256, "one table entry per thread");
    s_tab[threadIdx.x] = vtab[threadIdx.x];
    __syncthreads();
  }
  auto value = [&](const VT* p) -> DT {
    if constexpr(VCODE)
      return s_tab[__builtin_nontemporal_load(p)];
    else
      return ld_val<DT, VT>(p);
  };
  const uint64_t nch = (rows + kThreads - 1) / kThreads;
  const XcdRange xr = xcd_range(nch);
  const int lane = threadIdx.x & 63;
  double acc0 = 0.0, acc1 = 0.0;
  epi.begin();
  for(uint64_t ch = xr.first; ch < xr.end; ch += xr.stride)
  {
    const uint64_t row = ch * kThreads + threadIdx.x;
    if(row < rows)
    {
      const uint64_t off = soff[row >> 6];
      const int len = rlen[row];
      const VT* v = val + off + lane;
      const CT* c = col + off + lane;
      typename EpiPre<Epi>::type pf{};
      if constexpr(EpiPre<Epi>::on) pf = epi.pre(row);
      DT sum = DT(0);
      int k = 0;
      // matrix arrays are streamed once: nontemporal loads keep x resident in L2
      for(; k + UNR <= len; k += UNR)
      {
        DT vv[UNR];
        uint64_t cc[UNR];
        DT xv[UNR];
#pragma unroll
        for(int u = 0; u < UNR; ++u)
        {
          vv[u] = value(v + (k + u) * 64);
          cc[u] = sell_col<DELTA>(__builtin_nontemporal_load(c + (k + u) * 64), row);
        }
#pragma unroll
        for(int u = 0; u < UNR; ++u) xv[u] = x[cc[u]];
#pragma unroll
        for(int u = 0; u < UNR; ++u) sum = dfma(vv[u], xv[u], sum);
      }
      for(; k < len; ++k)
        sum = dfma(value(v + k * 64), x[sell_col<DELTA>(__builtin_nontemporal_load(c + k * 64), row)], sum);
      if(!VCODE && !std::is_same<VT, DT>::value) sum = sum * scale;  // exact: a power of two
      if constexpr(EpiPre<Epi>::on)
        epi.row(row, sum, pf, acc0, acc1);
      else
        epi.row(row, sum, acc0, acc1);
    }
  }
  grid_reduce_finish(gr, acc0, acc1);
}

// ---------------------------------------------------------------------------
// Pair-dictionary SELL-64: each stored entry is one byte, the code of its
// (column - row, value) pair in a table of at most 256 pairs that the
// workgroup stages in LDS.  Entry k of a row sits in byte k % 4 of word k / 4
// of its lane (word w of lane l of slice s at codes[soff[s] + 64 w + l]), so a
// lane fetches four entries per load and a wavefront reads 256 contiguous
// bytes.  Row sums run in CSR order with one fma per entry, bit-identical to
// the CSR loop.
// ---------------------------------------------------------------------------
template<typename DT, class Epi, int G>
__global__ __launch_bounds__(kThreads) void k_sell_dict_rows(const DT* __restrict__ tab,
                                                             const uint32_t* __restrict__ codes,
                                                             const uint64_t* __restrict__ soff, uint32_t sw,
                                                             const uint16_t* __restrict__ rlen,
                                                             const DT* __restrict__ x, uint64_t rows, Epi epi,
                                                             GridReduce gr)
{
  static_assert(kThreads == 256, "one table entry per thread");
  static_assert(G == 4 || G == 8 || G == 16, "gather group");
  __shared__ DT s_val[256];
  __shared__ int s_del[256];
  s_val[threadIdx.x] = tab[threadIdx.x];
  s_del[threadIdx.x] = reinterpret_cast<const int*>(tab + 256)[threadIdx.x];
  __syncthreads();
  const uint64_t nch = (rows + kThreads - 1) / kThreads;
  const XcdRange xr = xcd_range(nch);
  const int lane = threadIdx.x & 63;
  double acc0 = 0.0, acc1 = 0.0;
  epi.begin();
  for(uint64_t ch = xr.first; ch < xr.end; ch += xr.stride)
  {
    const uint64_t row = ch * kThreads + threadIdx.x;
    if(row < rows)
    {
      // uniform slices: the code address needs no offset load
      const uint32_t* cw = codes + (sw ? (row >> 6) * 64 * uint64_t(sw) : soff[row >> 6]) + lane;
      const int len = rlen[row];
      const int nw = (len + 3) >> 2;
      DT sum = DT(0);
      // up to 8 code words (32 entries) are fetched before any column is gathered;
      // the gathers then run G entries at a time
      for(int w0 = 0; w0 < nw; w0 += 8)
      {
        uint32_t wd[8];
#pragma unroll
        for(int u = 0; u < 8; ++u) wd[u] = w0 + u < nw ? __builtin_nontemporal_load(cw + (w0 + u) * 64) : 0u;
#pragma unroll
        for(int g0 = 0; g0 < 32; g0 += G)
        {
          const int kb = 4 * w0 + g0;
          if(kb >= len) break;
          DT vv[G], xv[G];
#pragma unroll
          for(int j = 0; j < G; ++j)
          {
            const int e = g0 + j;
            const uint32_t c = (wd[e >> 2] >> (8 * (e & 3))) & 255u;
            vv[j] = s_val[c];
            xv[j] = x[kb + j < len ? uint64_t(std::int64_t(row) + s_del[c]) : uint64_t(0)];
          }
#pragma unroll
          for(int j = 0; j < G; ++j)
            if(kb + j < len) sum = dfma(vv[j], xv[j], sum);
        }
      }
      epi.row(row, sum, acc0, acc1);
    }
  }
  grid_reduce_finish(gr, acc0, acc1);
}

// DPP wave shifts (wave64): lane l receives lane l-1's value (shr) or lane l+1's (shl); the
// edge lane gets 0 and is overwritten by the caller
constexpr int kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138;
template<int CTRL>
__device__ __forceinline__ double wave_shift(double v)
{
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const int lo = __builtin_amdgcn_update_dpp(0, int(uint32_t(b)), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(uint32_t(b >> 32)), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, uint64_t(uint32_t(lo)) | (uint64_t(uint32_t(hi)) << 32));
}
template<int CTRL>
__device__ __forceinline__ float wave_shift(float v)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

// Diagonal plan (FEAT_IDX_DIAG): the distinct column offsets are covered by triples
// (c-1, c, c+1) around kMaxTrip or fewer centres c_g; byte 3g+j of a row's code words is the
// pair code of its entry at offset c_g + j - 1 (kDiagNone where the row has none).  The rows
// of a wave are consecutive, so x[row + c - 1] is the left lane's x[row' + c] and
// x[row + c + 1] the right lane's: one load per triple and two wave shifts.
constexpr int kMaxTrip = 10;
constexpr uint32_t kDiagNone = 255u;   // slot without an entry

// SpMV-type row loop over the diagonal plan: the slots run in ascending offset order (=
// ascending columns, the CSR order), each lane adds its entry of a slot if it has one.
// Metadata after the pair table: the number of triples, their centres.  A wave covers
// kHaloRows = 62 rows with its lanes 1..62; lanes 0 and 63 only load x for the rows just
// outside (base - 1 and base + 62), so the left / right value of every triple comes from
// a neighbouring lane by a wave shift and no wave needs loads for its edges (3 % idle
// lanes against 9 masked loads and 20 registers per wave: 36.7 -> 31.1 us on 3D Q1 L7).
constexpr int kHaloRows = 62;
constexpr int kHaloChunk = kHaloRows * (kThreads / 64);   // rows per workgroup chunk

// NT > 0: the number of triples and the code words per row (NW) are fixed at compile time
// (the 27-point stencil: 9 triples in 7 words), so the offsets sit in scalar registers and
// every load of a chunk is issued straight-line (3D Q1 level 7 Jacobi step: 30 -> 27 us);
// NT == 0 reads both from the plan.  The code words go through the caches like the
// vectors: a level's codes are re-read by every smoothing step and stay resident with them
// in the last-level cache (nontemporal code loads: 30 against 27 us).
template<typename DT, class Epi, int NT = 0, int NW = 0>
__global__ __launch_bounds__(kThreads) void k_sell_diag_rows(const DT* __restrict__ tab,
                                                             const uint32_t* __restrict__ codes, uint32_t nw,
                                                             const DT* __restrict__ x, uint64_t rows, uint64_t cols,
                                                             Epi epi, GridReduce gr)
{
  static_assert(kThreads == 256, "one table entry per thread");
  static_assert(NT >= 0 && NT <= kMaxTrip && NW >= 0 && NW <= 8 && (NT == 0) == (NW == 0), "diagonal plan shape");
  constexpr int kT = NT ? NT : kMaxTrip, kW = NW ? NW : 8;
  __shared__ DT s_val[256];
  const DT tv = tab[threadIdx.x];
  const int* meta = reinterpret_cast<const int*>(tab + 256) + 256;
  const int nt = NT ? NT : meta[0];
  const int nwr = NW ? NW : int(nw);
  int off[kT];
#pragma unroll
  for(int g = 0; g < kT; ++g) off[g] = (NT || g < nt) ? meta[1 + g] : 0;
  const uint64_t nch = (rows + kHaloChunk - 1) / kHaloChunk;
  const XcdRange xr = xcd_range(nch);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const std::int64_t last = std::int64_t(cols) - 1;
  double acc0 = 0.0, acc1 = 0.0;
  epi.begin();
  for(uint64_t ch = xr.first; ch < xr.end; ch += xr.stride)
  {
    const std::int64_t base = std::int64_t(ch) * kHaloChunk + std::int64_t(wave) * kHaloRows;
    const std::int64_t row = base + lane - 1;   // lanes 0 and 63: halo rows
    const bool live = base < std::int64_t(rows);
    const bool mine = lane >= 1 && lane <= kHaloRows && row < std::int64_t(rows);
    uint32_t wd[kW];
    DT cen[kT];
    typename EpiPre<Epi>::type pf{};
#pragma unroll
    for(int w = 0; w < kW; ++w) wd[w] = 0xFFFFFFFFu;
    if(mine)
    {
      const uint32_t* cw = codes + uint64_t(row >> 6) * 64 * uint64_t(nwr) + uint64_t(row & 63);
#pragma unroll
      for(int w = 0; w < kW; ++w)
        if(NW || w < nwr) wd[w] = cw[w * 64];
      if constexpr(EpiPre<Epi>::on) pf = epi.pre(uint64_t(row));
    }
    if(live)
    {
#pragma unroll
      for(int g = 0; g < kT; ++g)
        if(NT || g < nt)
        {
          std::int64_t i = row + off[g];
          i = i < 0 ? 0 : (i > last ? last : i);
          cen[g] = x[i];
        }
    }
    if(ch == xr.first)   // block-uniform: every wave of the block passes here once
    {
      s_val[threadIdx.x] = tv;
      __syncthreads();
    }
    if(!live) continue;
    DT sum = DT(0);
#pragma unroll
    for(int g = 0; g < kT; ++g)
      if(NT || g < nt)
      {
        const DT v3[3] = {wave_shift<kDppWaveShr1>(cen[g]), cen[g], wave_shift<kDppWaveShl1>(cen[g])};
#pragma unroll
        for(int j = 0; j < 3; ++j)
        {
          const int t = 3 * g + j;
          const uint32_t c = (wd[t >> 2] >> (8 * (t & 3))) & 255u;
          const DT p = dfma(s_val[c], v3[j], sum);   // code 255 (no entry): discarded
          sum = c != kDiagNone ? p : sum;
        }
      }
    if(mine)
    {
      if constexpr(EpiPre<Epi>::on)
        epi.row(uint64_t(row), sum, pf, acc0, acc1);
      else
        epi.row(uint64_t(row), sum, acc0, acc1);
    }
  }
  grid_reduce_finish(gr, acc0, acc1);
}

// d-linear injection between structured grids (FEAT_FMT_DLINEAR): row sums formed from the
// node positions, in the ascending column order of the assembled transfer
struct TailGeom
{
  int dim, n1;   // dim 0: not a grid
};

// prolongation row i of the fine grid (n1 per side): the coarse nodes around the fine node in
// ascending order, each weighted with 1/2 per odd axis.  The (up to 8) loads are issued
// together, then summed in that order.
// I: the index type (int32_t where the grid has fewer than 2^31 nodes: cheaper divisions)
template<typename DT, typename I = int64_t>
__device__ __forceinline__ DT dlinear_prol_row(TailGeom g, I i, const DT* x)
{
  const I n1 = g.n1, nc = (n1 + 1) / 2;
  const I fx = i % n1, fy = (i / n1) % n1, fz = g.dim == 3 ? i / (n1 * n1) : 0;
  const int ox = int(fx & 1), oy = int(fy & 1), oz = int(fz & 1);
  const DT w = DT((ox ? 0.5 : 1.0) * (oy ? 0.5 : 1.0) * (oz ? 0.5 : 1.0));
  const I c0 = (fz >> 1) * nc * nc + (fy >> 1) * nc + (fx >> 1);
  DT v[8];
#pragma unroll
  for(int t = 0; t < 8; ++t)
  {
    const int ix = t & 1, iy = (t >> 1) & 1, iz = t >> 2;
    const bool on = ix <= ox && iy <= oy && iz <= oz;
    v[t] = on ? x[c0 + iz * nc * nc + iy * nc + ix] : DT(0);
  }
  DT s = DT(0);
#pragma unroll
  for(int t = 0; t < 8; ++t)
  {
    const int ix = t & 1, iy = (t >> 1) & 1, iz = t >> 2;
    if(ix <= ox && iy <= oy && iz <= oz) s = dfma(w, v[t], s);
  }
  return s;
}

// restriction row jc of the coarse grid: the fine nodes (2X + dx, 2Y + dy, 2Z + dz) inside
// the grid in ascending order, weights 1/2 per nonzero offset; the (up to 27) loads are
// issued together, then summed in that order
template<typename DT, typename I = int64_t>
__device__ __forceinline__ DT dlinear_rest_row(TailGeom g, I jc, const DT* x)
{
  const I n1 = g.n1, nc = (n1 + 1) / 2;
  const I X = jc % nc, Y = (jc / nc) % nc, Z = g.dim == 3 ? jc / (nc * nc) : 0;
  DT v[27];
#pragma unroll
  for(int t = 0; t < 27; ++t)
  {
    const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
    const I fx = 2 * X + dx, fy = 2 * Y + dy, fz = 2 * Z + dz;
    const bool on = fx >= 0 && fx < n1 && fy >= 0 && fy < n1 && (g.dim == 3 ? (fz >= 0 && fz < n1) : dz == 0);
    v[t] = on ? x[(fz * n1 + fy) * n1 + fx] : DT(0);
  }
  DT s = DT(0);
#pragma unroll
  for(int t = 0; t < 27; ++t)
  {
    const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
    const I fx = 2 * X + dx, fy = 2 * Y + dy, fz = 2 * Z + dz;
    const bool on = fx >= 0 && fx < n1 && fy >= 0 && fy < n1 && (g.dim == 3 ? (fz >= 0 && fz < n1) : dz == 0);
    if(on) s = dfma(DT((dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0) * (dz ? 0.5 : 1.0)), v[t], s);
  }
  return s;
}

template<typename DT, class Epi, bool REST, typename I>
__global__ __launch_bounds__(kThreads) void k_dlinear_rows(int dim, int n1, const DT* __restrict__ x, uint64_t rows,
                                                           Epi epi, GridReduce gr)
{
  double acc0 = 0.0, acc1 = 0.0;
  epi.begin();
  const TailGeom g{dim, n1};
  for(uint64_t i = blockIdx.x * uint64_t(kThreads) + threadIdx.x; i < rows; i += uint64_t(gridDim.x) * kThreads)
  {
    typename EpiPre<Epi>::type pf{};
    if constexpr(EpiPre<Epi>::on) pf = epi.pre(i);
    const DT sum = REST ? dlinear_rest_row<DT, I>(g, I(i), x) : dlinear_prol_row<DT, I>(g, I(i), x);
    if constexpr(EpiPre<Epi>::on)
      epi.row(i, sum, pf, acc0, acc1);
    else
      epi.row(i, sum, acc0, acc1);
  }
  grid_reduce_finish(gr, acc0, acc1);
}

// Banded plan (FEAT_FMT_BANDED): row i of an n1^DIM grid takes its neighbours (i+dx, j+dy,
// k+dz) inside the grid, columns formed from the node position, values slot-major.  All
// loads of a row are issued together, then summed in ascending column order (slot order),
// which is the CSR order the plan was checked against at set-up.  The value arrays are
// streamed once (nontemporal loads); x stays in L2.  Indices and byte offsets are 32-bit:
// the plan requires 27 * rows * 8 B < 4 GiB (fewer than 19.8 M rows).
template<typename DT, typename VT, class Epi, int DIM>
__global__ __launch_bounds__(kThreads) void k_band_rows(const VT* __restrict__ val, int n1,
                                                        const DT* __restrict__ x, uint64_t rows, DT scale,
                                                        Epi epi, GridReduce gr)
{
  constexpr int NS = DIM == 3 ? 27 : 9;
  double acc0 = 0.0, acc1 = 0.0;
  epi.begin();
  const int32_t m = n1, mm = n1 * n1;
  // the chunk-to-workgroup mapping of the SELL-64 launch, so grid reductions (norms, dots)
  // take the same partial sums as on the plain plan
  const uint64_t nch = (rows + kThreads - 1) / kThreads;
  const XcdRange xr = xcd_range(nch);
  for(uint64_t ch = xr.first; ch < xr.end; ch += xr.stride)
  {
    const uint64_t r = ch * kThreads + threadIdx.x;
    if(r >= rows) continue;
    typename EpiPre<Epi>::type pf{};
    if constexpr(EpiPre<Epi>::on) pf = epi.pre(r);
    const int32_t i = int32_t(r);
    const int32_t fx = i % m, fy = DIM == 3 ? (i / m) % m : i / m, fz = DIM == 3 ? i / mm : 0;
    const bool lx = fx > 0, hx = fx + 1 < m, ly = fy > 0, hy = fy + 1 < m, lz = fz > 0, hz = fz + 1 < m;
    const uint32_t ro = uint32_t(i) * uint32_t(sizeof(VT)), st = uint32_t(rows) * uint32_t(sizeof(VT));
    // Stored values stay raw until the sum.  16-bit values: every load is issued
    // unconditionally (each slot array spans all rows, and an outside neighbour reads x[i];
    // both are left out of the sum), since a predicated 16-bit load is compiled into a
    // branch that waits for it before converting (FP16 plan 1.8x slower).  Wider values
    // keep predicated loads, which measured faster for them (FP64 solve 70.4 against 78.9 ms).
    VT v[NS];
    DT xv[NS];
#pragma unroll
    for(int t = 0; t < NS; ++t)
    {
      const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = DIM == 3 ? t / 9 - 1 : 0;
      const bool on = (dx < 0 ? lx : dx > 0 ? hx : true) && (dy < 0 ? ly : dy > 0 ? hy : true) &&
                      (dz < 0 ? lz : dz > 0 ? hz : true);
      // 32-bit byte offsets from one base per array (the plan keeps its values below 4 GiB):
      // one offset register per load instead of a 64-bit address
      const VT* vp = (const VT*)((const char*)val + (ro + uint32_t(t) * st));
      if constexpr(sizeof(VT) == 2)
      {
        v[t] = __builtin_nontemporal_load(vp);
        xv[t] = *(const DT*)((const char*)x + uint32_t(on ? i + dx + dy * m + dz * mm : i) * uint32_t(sizeof(DT)));
      }
      else
      {
        v[t] = on ? __builtin_nontemporal_load(vp) : VT(0);
        xv[t] = on ? *(const DT*)((const char*)x + uint32_t(i + dx + dy * m + dz * mm) * uint32_t(sizeof(DT))) : DT(0);
      }
    }
    DT sum = DT(0);
#pragma unroll
    for(int t = 0; t < NS; ++t)
    {
      const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = DIM == 3 ? t / 9 - 1 : 0;
      const bool on = (dx < 0 ? lx : dx > 0 ? hx : true) && (dy < 0 ? ly : dy > 0 ? hy : true) &&
                      (dz < 0 ? lz : dz > 0 ? hz : true);
      if(on) sum = dfma(cvt_val<DT, VT>(v[t]), xv[t], sum);
    }
    if(!std::is_same<VT, DT>::value) sum = sum * scale;  // exact: a power of two
    if constexpr(EpiPre<Epi>::on)
      epi.row(r, sum, pf, acc0, acc1);
    else
      epi.row(r, sum, acc0, acc1);
  }
  grid_reduce_finish(gr, acc0, acc1);
}

// dimension of a banded plan: 3 where n1^3 = rows, else 2
__host__ __device__ inline int band_dim(uint64_t rows, uint64_t n1) { return n1 * n1 * n1 == rows ? 3 : 2; }

// Runs Epi over the rows of a descriptor (CSR or SELL-64).
template<typename DT, class Epi, class M>
int launch_mat(hipStream_t s, const M* A, const DT* x, const Epi& epi, const GridReduce& gr, const char* name)
{
  if(A->rows == 0) return FEAT_OK;
  if(A->format == FEAT_FMT_DLINEAR)
  {
    const uint64_t nb = (A->rows + kThreads - 1) / kThreads;
    const int grid = int(nb < uint64_t(kMaxGrid) ? nb : uint64_t(kMaxGrid));
    const GridReduce gd = reduce_for_grid(gr, uint64_t(grid));
    // 32-bit index arithmetic while the fine grid has fewer than 2^31 nodes
    const bool small = A->rows < (uint64_t(1) << 31) && A->cols < (uint64_t(1) << 31);
    if(A->value_kind && small)
      hipLaunchKernelGGL((k_dlinear_rows<DT, Epi, true, int32_t>), dim3(grid), dim3(kThreads), 0, s, A->index,
                         int(A->slice_words), x, A->rows, epi, gd);
    else if(A->value_kind)
      hipLaunchKernelGGL((k_dlinear_rows<DT, Epi, true, int64_t>), dim3(grid), dim3(kThreads), 0, s, A->index,
                         int(A->slice_words), x, A->rows, epi, gd);
    else if(small)
      hipLaunchKernelGGL((k_dlinear_rows<DT, Epi, false, int32_t>), dim3(grid), dim3(kThreads), 0, s, A->index,
                         int(A->slice_words), x, A->rows, epi, gd);
    else
      hipLaunchKernelGGL((k_dlinear_rows<DT, Epi, false, int64_t>), dim3(grid), dim3(kThreads), 0, s, A->index,
                         int(A->slice_words), x, A->rows, epi, gd);
    finish_reduce(s, gd, uint64_t(grid));
    return check_launch(name);
  }
  if(A->format == FEAT_FMT_BANDED)
  {
    const uint64_t nb = (A->rows + kThreads - 1) / kThreads;   // the SELL-64 default grid
    const int grid = int(nb < uint64_t(kMaxGrid) ? nb : uint64_t(kMaxGrid));
    const GridReduce gd = reduce_for_grid(gr, uint64_t(grid));
    const int n1 = int(A->slice_words);
    const bool d3 = band_dim(A->rows, A->slice_words) == 3;
    const DT sc = DT(A->scale);
    if(A->value_kind == FEAT_VAL_F16)
    {
      if constexpr(!std::is_same<DT, float>::value)
        return fail(FEAT_ERR_UNSUPPORTED, "FP16 values need FP32 arithmetic");
      else
      {
        const uint16_t* vh = (const uint16_t*)(const void*)A->val;
        if(d3)
          hipLaunchKernelGGL((k_band_rows<DT, uint16_t, Epi, 3>), dim3(grid), dim3(kThreads), 0, s, vh, n1, x, A->rows, sc, epi, gd);
        else
          hipLaunchKernelGGL((k_band_rows<DT, uint16_t, Epi, 2>), dim3(grid), dim3(kThreads), 0, s, vh, n1, x, A->rows, sc, epi, gd);
      }
    }
    else if(d3)
      hipLaunchKernelGGL((k_band_rows<DT, DT, Epi, 3>), dim3(grid), dim3(kThreads), 0, s, A->val, n1, x, A->rows, sc, epi, gd);
    else
      hipLaunchKernelGGL((k_band_rows<DT, DT, Epi, 2>), dim3(grid), dim3(kThreads), 0, s, A->val, n1, x, A->rows, sc, epi, gd);
    finish_reduce(s, gd, uint64_t(grid));
    return check_launch(name);
  }
  if(A->format == FEAT_FMT_CSR)
  {
    if(A->index == FEAT_IDX_U64)
      return launch_rows<DT, uint64_t>(s, A->val, (const uint64_t*)A->col, (const uint64_t*)A->ptr, x, A->rows, epi, gr, name);
    if(A->index == FEAT_IDX_U32)
      return launch_rows<DT, uint32_t>(s, A->val, (const uint32_t*)A->col, (const uint32_t*)A->ptr, x, A->rows, epi, gr, name);
  }
  else if(A->format == FEAT_FMT_SELL)
  {
    // one workgroup per 256-row chunk up to the reduction-partials limit, so the
    // hardware dispatcher balances the load; larger matrices stride.  A tuned
    // descriptor caps the grid lower (workgroups stride over the chunks) -- except
    // for launches that reduce: their partial sums follow the grid, and the default
    // grid keeps every dot and norm independent of the tuning.
    const uint64_t nch = (A->rows + kThreads - 1) / kThreads;
    const bool reduces = gr.out0 != nullptr || gr.out1 != nullptr;
    const int cap_code = reduces ? 0 : (A->tune >> 2) & 7;
    const uint64_t cap = cap_code ? (uint64_t(kMaxGrid) >> cap_code) : uint64_t(kMaxGrid);
    const int grid = int(nch < cap ? nch : cap);
    const GridReduce gd = reduce_for_grid(gr, uint64_t(grid));
    const int unr_code = A->tune & 3;
    const uint64_t* soff = (const uint64_t*)A->ptr;
    const DT sc = DT(A->scale);
    if(A->value_kind == FEAT_VAL_F16)
    {
      if(!std::is_same<DT, float>::value) return fail(FEAT_ERR_UNSUPPORTED, "FP16 values need FP32 arithmetic");
      const uint16_t* vh = (const uint16_t*)(const void*)A->val;
      if(A->index == FEAT_IDX_U32)
        hipLaunchKernelGGL((k_sell_rows<DT, uint16_t, uint32_t, false, Epi>), dim3(grid), dim3(kThreads), 0, s, vh,
                           (const uint32_t*)A->col, soff, A->rowlen, x, A->rows, sc, epi, gd);
      else if(A->index == FEAT_IDX_D16)
        hipLaunchKernelGGL((k_sell_rows<DT, uint16_t, int16_t, true, Epi>), dim3(grid), dim3(kThreads), 0, s, vh,
                           (const int16_t*)A->col, soff, A->rowlen, x, A->rows, sc, epi, gd);
      else
        return fail(FEAT_ERR_UNSUPPORTED, "SELL index kind");
      finish_reduce(s, gd, uint64_t(grid));
      return check_launch(name);
    }
    if(A->index == FEAT_IDX_DIAG)
    {
      // chunks of kHaloChunk rows (62 per wave); a dot or norm is summed per chunk, so it
      // rounds differently from the same sum over the 256-row chunks of the other plans
      const uint64_t nh = (A->rows + kHaloChunk - 1) / kHaloChunk;
      const int gh = int(nh < cap ? nh : cap);
      const GridReduce gdh = reduce_for_grid(gr, uint64_t(gh));
      // 7 code words per row means 9 triples (27 slots; 8 or 10 triples need 6 or 8 words)
      if(A->slice_words == 7)
        hipLaunchKernelGGL((k_sell_diag_rows<DT, Epi, 9, 7>), dim3(gh), dim3(kThreads), 0, s,
                           reinterpret_cast<const DT*>(A->val), static_cast<const uint32_t*>(A->col), A->slice_words, x,
                           A->rows, A->cols, epi, gdh);
      else
        hipLaunchKernelGGL((k_sell_diag_rows<DT, Epi>), dim3(gh), dim3(kThreads), 0, s, reinterpret_cast<const DT*>(A->val),
                           static_cast<const uint32_t*>(A->col), A->slice_words, x, A->rows, A->cols, epi, gdh);
      finish_reduce(s, gdh, uint64_t(gh));
      return check_launch(name);
    }
    if(A->index == FEAT_IDX_DICT)
    {
      const DT* tab = reinterpret_cast<const DT*>(A->val);
      const uint32_t* cw = static_cast<const uint32_t*>(A->col);
      if(unr_code == 1)
        hipLaunchKernelGGL((k_sell_dict_rows<DT, Epi, 4>), dim3(grid), dim3(kThreads), 0, s, tab, cw, soff, A->slice_words,
                           A->rowlen, x, A->rows, epi, gd);
      else if(unr_code == 3)
        hipLaunchKernelGGL((k_sell_dict_rows<DT, Epi, 16>), dim3(grid), dim3(kThreads), 0, s, tab, cw, soff, A->slice_words,
                           A->rowlen, x, A->rows, epi, gd);
      else
        hipLaunchKernelGGL((k_sell_dict_rows<DT, Epi, 8>), dim3(grid), dim3(kThreads), 0, s, tab, cw, soff, A->slice_words,
                           A->rowlen, x, A->rows, epi, gd);
      finish_reduce(s, gd, uint64_t(grid));
      return check_launch(name);
    }
    if(A->value_kind == FEAT_VAL_CODE8)
    {
      const DT* vt = reinterpret_cast<const DT*>(A->val);
      const uint8_t* vc = reinterpret_cast<const uint8_t*>(vt + 256);
      if(A->index == FEAT_IDX_U32)
      {
        if(std::is_same<DT, double>::value && unr_code == 1)
          hipLaunchKernelGGL((k_sell_rows<DT, uint8_t, uint32_t, false, Epi, 4, true>), dim3(grid), dim3(kThreads), 0, s,
                             vc, (const uint32_t*)A->col, soff, A->rowlen, x, A->rows, DT(1), epi, gd, vt);
        else if(std::is_same<DT, double>::value && unr_code == 3)
          hipLaunchKernelGGL((k_sell_rows<DT, uint8_t, uint32_t, false, Epi, 16, true>), dim3(grid), dim3(kThreads), 0, s,
                             vc, (const uint32_t*)A->col, soff, A->rowlen, x, A->rows, DT(1), epi, gd, vt);
        else
          hipLaunchKernelGGL((k_sell_rows<DT, uint8_t, uint32_t, false, Epi, kSellUnroll, true>), dim3(grid),
                             dim3(kThreads), 0, s, vc, (const uint32_t*)A->col, soff, A->rowlen, x, A->rows, DT(1), epi,
                             gd, vt);
      }
      else if(A->index == FEAT_IDX_D16)
      {
        if(std::is_same<DT, double>::value && unr_code == 1)
          hipLaunchKernelGGL((k_sell_rows<DT, uint8_t, int16_t, true, Epi, 4, true>), dim3(grid), dim3(kThreads), 0, s,
                             vc, (const int16_t*)A->col, soff, A->rowlen, x, A->rows, DT(1), epi, gd, vt);
        else if(std::is_same<DT, double>::value && unr_code == 3)
          hipLaunchKernelGGL((k_sell_rows<DT, uint8_t, int16_t, true, Epi, 16, true>), dim3(grid), dim3(kThreads), 0, s,
                             vc, (const int16_t*)A->col, soff, A->rowlen, x, A->rows, DT(1), epi, gd, vt);
        else
          hipLaunchKernelGGL((k_sell_rows<DT, uint8_t, int16_t, true, Epi, kSellUnroll, true>), dim3(grid),
                             dim3(kThreads), 0, s, vc, (const int16_t*)A->col, soff, A->rowlen, x, A->rows, DT(1), epi,
                             gd, vt);
      }
      else
        return fail(FEAT_ERR_UNSUPPORTED, "SELL index kind");
      finish_reduce(s, gd, uint64_t(grid));
      return check_launch(name);
    }
    if(A->index == FEAT_IDX_U32)
    {
      if(std::is_same<DT, double>::value && unr_code == 1)
        hipLaunchKernelGGL((k_sell_rows<DT, DT, uint32_t, false, Epi, 4>), dim3(grid), dim3(kThreads), 0, s, A->val,
                           (const uint32_t*)A->col, soff, A->rowlen, x, A->rows, sc, epi, gd);
      else if(std::is_same<DT, double>::value && unr_code == 3)
        hipLaunchKernelGGL((k_sell_rows<DT, DT, uint32_t, false, Epi, 16>), dim3(grid), dim3(kThreads), 0, s, A->val,
                           (const uint32_t*)A->col, soff, A->rowlen, x, A->rows, sc, epi, gd);
      else
        hipLaunchKernelGGL((k_sell_rows<DT, DT, uint32_t, false, Epi>), dim3(grid), dim3(kThreads), 0, s, A->val,
                           (const uint32_t*)A->col, soff, A->rowlen, x, A->rows, sc, epi, gd);
      finish_reduce(s, gd, uint64_t(grid));
      return check_launch(name);
    }
    if(A->index == FEAT_IDX_D16)
    {
      if(std::is_same<DT, double>::value && unr_code == 1)
        hipLaunchKernelGGL((k_sell_rows<DT, DT, int16_t, true, Epi, 4>), dim3(grid), dim3(kThreads), 0, s, A->val,
                           (const int16_t*)A->col, soff, A->rowlen, x, A->rows, sc, epi, gd);
      else if(std::is_same<DT, double>::value && unr_code == 3)
        hipLaunchKernelGGL((k_sell_rows<DT, DT, int16_t, true, Epi, 16>), dim3(grid), dim3(kThreads), 0, s, A->val,
                           (const int16_t*)A->col, soff, A->rowlen, x, A->rows, sc, epi, gd);
      else
        hipLaunchKernelGGL((k_sell_rows<DT, DT, int16_t, true, Epi>), dim3(grid), dim3(kThreads), 0, s, A->val,
                           (const int16_t*)A->col, soff, A->rowlen, x, A->rows, sc, epi, gd);
      finish_reduce(s, gd, uint64_t(grid));
      return check_launch(name);
    }
  }
  return fail(FEAT_ERR_UNSUPPORTED, "matrix descriptor: unknown format/index combination");
}

// SELL-64 plan construction from a device CSR (u64).
__global__ void k_sell_scan_rows(const uint64_t* rp, const uint64_t* col, uint64_t rows, uint16_t* rlen,
                                 uint32_t* slmax, uint32_t* stats /* [0] max |col-row| (saturated), [1] long-row flag */)
{
  for(uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < rows; r += uint64_t(gridDim.x) * blockDim.x)
  {
    const uint64_t b = rp[r], e = rp[r + 1];
    const uint64_t len = e - b;
    uint64_t md = 0;
    for(uint64_t i = b; i < e; ++i)
    {
      const uint64_t d = col[i] > r ? col[i] - r : r - col[i];
      md = d > md ? d : md;
    }
    if(len > 65535) atomicMax(&stats[1], 1u);
    rlen[r] = uint16_t(len > 65535 ? 65535 : len);
    atomicMax(&slmax[r >> 6], uint32_t(len > 0xFFFFFFFFull ? 0xFFFFFFFFull : len));
    atomicMax(&stats[0], uint32_t(md > 0xFFFFFFFFull ? 0xFFFFFFFFull : md));
  }
}

template<typename VT>
__device__ __forceinline__ VT st_val(double a, float inv_scale)
{
  return VT(a);
}
template<>
__device__ __forceinline__ uint16_t st_val<uint16_t>(double a, float inv_scale)
{
  // FP16 bits of half(float(a) * 2^-e): round to FP32, scale exactly, round to FP16
  return __builtin_bit_cast(uint16_t, _Float16(float(a) * inv_scale));
}

__global__ void k_absmax(const double* val, uint64_t n, uint32_t* out)
{
  uint32_t m = 0;
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
  {
    const float a = std::fabs(float(val[i]));
    const uint32_t b = __builtin_bit_cast(uint32_t, a);
    m = b > m ? b : m;
  }
  atomicMax(out, m);
}

template<typename DT, typename CT, bool DELTA>
__global__ void k_sell_fill(const double* val, const uint64_t* col, const uint64_t* rp, uint64_t rows,
                            const uint64_t* soff, DT* sval, CT* scol, float inv_scale)
{
  for(uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < rows; r += uint64_t(gridDim.x) * blockDim.x)
  {
    const uint64_t s = r >> 6, lane = r & 63;
    const uint64_t off = soff[s], smax = (soff[s + 1] - soff[s]) / 64;
    const uint64_t b = rp[r], len = rp[r + 1] - b;
    for(uint64_t k = 0; k < smax; ++k)
    {
      const uint64_t p = off + k * 64 + lane;
      if(k < len)
      {
        sval[p] = st_val<DT>(val[b + k], inv_scale);
        scol[p] = DELTA ? CT(std::int64_t(col[b + k]) - std::int64_t(r)) : CT(col[b + k]);
      }
      else
      {
        sval[p] = DT(0);
        scol[p] = DELTA ? CT(0) : CT(r);
      }
    }
  }
}

// device-to-device copy, 16 bytes per lane (the runtime's copy kernel moves vectors at
// about 1.4 TB/s)
__global__ __launch_bounds__(kThreads) void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n)
{
  for(uint64_t i = blockIdx.x * uint64_t(kThreads) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kThreads)
    dst[i] = src[i];
}

// Banded-plan check: 1 in *bad where a row's CSR columns are not exactly its grid
// neighbours inside the n1^dim grid in ascending order (set-up only)
__global__ void k_band_check(const uint64_t* col, const uint64_t* rp, uint64_t rows, int dim, int64_t n1, int* bad)
{
  for(uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < rows; r += uint64_t(gridDim.x) * blockDim.x)
  {
    const int64_t i = int64_t(r);
    const int64_t fx = i % n1, fy = (i / n1) % n1, fz = dim == 3 ? i / (n1 * n1) : 0;
    uint64_t p = rp[r];
    const uint64_t e = rp[r + 1];
    bool ok = true;
    for(int t = 0; t < 27 && ok; ++t)
    {
      const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
      if(dim == 2 && t >= 9) break;
      const int64_t gz = dim == 3 ? fz + dz : 0;
      if(fx + dx < 0 || fx + dx >= n1 || fy + dy < 0 || fy + dy >= n1 || gz < 0 || gz >= n1) continue;
      const int64_t c = i + dx + dy * n1 + (dim == 3 ? dz * n1 * n1 : 0);
      if(p >= e || col[p] != uint64_t(c))
        ok = false;
      else
        ++p;
    }
    if(!ok || p != e) atomicOr(bad, 1);
  }
}

// Banded-plan fill: slot t of row r at bval[t * rows + r], 0 where the neighbour is outside
template<typename DT>
__global__ void k_band_fill(const double* val, const uint64_t* rp, uint64_t rows, int dim, int64_t n1, DT* bval,
                            float inv_scale)
{
  for(uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < rows; r += uint64_t(gridDim.x) * blockDim.x)
  {
    const int64_t i = int64_t(r);
    const int64_t fx = i % n1, fy = (i / n1) % n1, fz = dim == 3 ? i / (n1 * n1) : 0;
    uint64_t p = rp[r];
    const int ns = dim == 3 ? 27 : 9;
    for(int t = 0; t < ns; ++t)
    {
      const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = dim == 3 ? t / 9 - 1 : 0;
      const bool on = fx + dx >= 0 && fx + dx < n1 && fy + dy >= 0 && fy + dy < n1 && fz + dz >= 0 && fz + dz < n1;
      bval[uint64_t(t) * rows + r] = on ? st_val<DT>(val[p++], inv_scale) : DT(0);
    }
  }
}

// bit pattern of a stored value (the key of the value-code table) and back
template<typename DT>
__host__ __device__ inline uint64_t value_key(DT v)
{
  if constexpr(sizeof(DT) == 8)
    return __builtin_bit_cast(uint64_t, v);
  else if constexpr(sizeof(DT) == 4)
    return __builtin_bit_cast(uint32_t, v);
  else
    return __builtin_bit_cast(uint16_t, v);
}
template<typename DT>
__host__ __device__ inline DT key_value(uint64_t k)
{
  if constexpr(sizeof(DT) == 8)
    return __builtin_bit_cast(DT, k);
  else if constexpr(sizeof(DT) == 4)
    return __builtin_bit_cast(DT, uint32_t(k));
  else
    return __builtin_bit_cast(DT, uint16_t(k));
}

// Value-code fill: the SELL layout of k_sell_fill with each value replaced by its index
// in keys[0..nk) (the ascending bit patterns of the distinct values in the arithmetic type,
// found on the host).  A value whose key is not in the table (a device conversion that
// rounds otherwise than the host's) sets *miss; the plan then keeps full values.
template<typename DT, typename CT, bool DELTA>
__global__ void k_sell_fill_code(const double* val, const uint64_t* col, const uint64_t* rp, uint64_t rows,
                                 const uint64_t* soff, const uint64_t* keys, int nk, uint8_t* codes, CT* scol,
                                 int* miss)
{
  for(uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < rows; r += uint64_t(gridDim.x) * blockDim.x)
  {
    const uint64_t s = r >> 6, lane = r & 63;
    const uint64_t off = soff[s], smax = (soff[s + 1] - soff[s]) / 64;
    const uint64_t b = rp[r], len = rp[r + 1] - b;
    for(uint64_t k = 0; k < smax; ++k)
    {
      const uint64_t p = off + k * 64 + lane;
      if(k < len)
      {
        const uint64_t key = value_key<DT>(DT(val[b + k]));
        int lo = 0, hi = nk - 1;   // binary search for the key
        while(lo < hi)
        {
          const int mid = (lo + hi) >> 1;
          if(keys[mid] < key)
            lo = mid + 1;
          else
            hi = mid;
        }
        if(keys[lo] != key) *miss = 1;
        codes[p] = uint8_t(lo);
        scol[p] = DELTA ? CT(std::int64_t(col[b + k]) - std::int64_t(r)) : CT(col[b + k]);
      }
      else
      {
        codes[p] = 0;
        scol[p] = DELTA ? CT(0) : CT(r);
      }
    }
  }
}

// Pair-dictionary construction.  Distinct value bit patterns and distinct
// column - row offsets are collected in two open-addressing tables of
// kDictSlots keys (relaxed reads first, compare-and-swap only to insert); the
// host numbers the keys in ascending order, marks the (value, offset) code
// pairs present, numbers those, and the fill kernel writes every entry's pair
// code.  Collection stops early once either table holds more than 256 keys.
constexpr uint32_t kDictSlots = 1024;
constexpr uint64_t kDictEmpty = ~uint64_t(0);

__device__ __forceinline__ uint32_t dict_hash(uint64_t key)
{
  return uint32_t((key * 0x9E3779B97F4A7C15ull) >> 54) & (kDictSlots - 1);
}
__device__ __forceinline__ uint64_t dict_delta_key(uint64_t col, uint64_t row)
{
  return uint64_t(std::int64_t(col) - std::int64_t(row)) ^ (uint64_t(1) << 63);
}

// slot of `key`, inserting it if absent; -1 after kDictProbes occupied slots (at most 256
// keys in 1024 slots: longer probe runs mean the matrix has too many pairs anyway)
constexpr uint32_t kDictProbes = 64;
__device__ int dict_insert(uint64_t* tab, uint32_t* count, uint64_t key)
{
  uint32_t h = dict_hash(key);
  for(uint32_t p = 0; p < kDictProbes; ++p)
  {
    const uint64_t cur = __hip_atomic_load(tab + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if(cur == key) return int(h);
    if(cur == kDictEmpty)
    {
      const uint64_t prev = atomicCAS(reinterpret_cast<unsigned long long*>(tab + h), (unsigned long long)kDictEmpty,
                                      (unsigned long long)key);
      if(prev == kDictEmpty)
      {
        atomicAdd(count, 1u);
        return int(h);
      }
      if(prev == key) return int(h);
    }
    h = (h + 1) & (kDictSlots - 1);
  }
  return -1;
}

// slot of a key known to be present
__device__ __forceinline__ uint32_t dict_find(const uint64_t* tab, uint64_t key)
{
  uint32_t h = dict_hash(key);
  for(uint32_t p = 0; p < kDictSlots && tab[h] != key; ++p) h = (h + 1) & (kDictSlots - 1);
  return h;
}

// state[0], state[1]: distinct values / offsets inserted; state[2]: failure flag
__global__ void k_dict_collect(const double* val, const uint64_t* col, const uint64_t* rp, uint64_t rows,
                               uint64_t* vtab, uint64_t* dtab, uint32_t* state)
{
  for(uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < rows; r += uint64_t(gridDim.x) * blockDim.x)
  {
    if(__hip_atomic_load(state, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > 256u ||
       __hip_atomic_load(state + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > 256u ||
       __hip_atomic_load(state + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
      return;
    for(uint64_t i = rp[r]; i < rp[r + 1]; ++i)
    {
      const uint64_t vb = __builtin_bit_cast(uint64_t, val[i]);
      if(vb == kDictEmpty || dict_insert(vtab, state, vb) < 0 || dict_insert(dtab, state + 1, dict_delta_key(col[i], r)) < 0 ||
         __hip_atomic_load(state, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > 256u ||
         __hip_atomic_load(state + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > 256u)
      {
        atomicOr(state + 2, 1u);
        return;
      }
    }
  }
}

// present[cv * 256 + cd] = 1 for every entry's (value code, offset code)
__global__ void k_dict_mark(const double* val, const uint64_t* col, const uint64_t* rp, uint64_t rows,
                            const uint64_t* vtab, const uint64_t* dtab, const uint8_t* vcode, const uint8_t* dcode,
                            uint32_t* present)
{
  for(uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < rows; r += uint64_t(gridDim.x) * blockDim.x)
    for(uint64_t i = rp[r]; i < rp[r + 1]; ++i)
    {
      const uint32_t cv = vcode[dict_find(vtab, __builtin_bit_cast(uint64_t, val[i]))];
      const uint32_t cd = dcode[dict_find(dtab, dict_delta_key(col[i], r))];
      present[cv * 256u + cd] = 1u;
    }
}

// the 4-entries-per-word code layout of the plan (padding bytes 0, never read)
__global__ void k_dict_fill(const double* val, const uint64_t* col, const uint64_t* rp, uint64_t rows,
                            const uint64_t* soff, const uint64_t* vtab, const uint64_t* dtab, const uint8_t* vcode,
                            const uint8_t* dcode, const uint8_t* pcode, uint32_t* codes)
{
  for(uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < rows; r += uint64_t(gridDim.x) * blockDim.x)
  {
    const uint64_t s = r >> 6, lane = r & 63;
    const uint64_t off = soff[s], nw = (soff[s + 1] - soff[s]) / 64;
    const uint64_t b = rp[r], len = rp[r + 1] - b;
    for(uint64_t w = 0; w < nw; ++w)
    {
      uint32_t word = 0;
      for(uint64_t j = 0; j < 4; ++j)
      {
        const uint64_t k = 4 * w + j;
        if(k < len)
        {
          const uint32_t cv = vcode[dict_find(vtab, __builtin_bit_cast(uint64_t, val[b + k]))];
          const uint32_t cd = dcode[dict_find(dtab, dict_delta_key(col[b + k], r))];
          word |= uint32_t(pcode[cv * 256u + cd]) << (8 * j);
        }
      }
      codes[off + w * 64 + lane] = word;
    }
  }
}

// the slot layout of the diagonal plan: byte slot_of[d] of a row's code words holds the pair
// code of its entry at the d-th smallest column offset, 255 where the row has none; rows past
// the end of the matrix (the last slice) hold 255 only.  state: set when a row is not
// strictly ascending (two entries would share a slot or leave CSR order)
__global__ void k_dict_fill_diag(const double* val, const uint64_t* col, const uint64_t* rp, uint64_t rows,
                                 uint64_t padded, uint32_t nw, const uint64_t* vtab, const uint64_t* dtab,
                                 const uint8_t* vcode, const uint8_t* dcode, const uint8_t* pcode, const uint8_t* slot_of,
                                 uint32_t* codes, uint32_t* state)
{
  for(uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < padded; r += uint64_t(gridDim.x) * blockDim.x)
  {
    uint32_t word[8];
#pragma unroll
    for(int w = 0; w < 8; ++w) word[w] = 0xFFFFFFFFu;
    if(r < rows)
      for(uint64_t i = rp[r]; i < rp[r + 1]; ++i)
      {
        if(i > rp[r] && col[i] <= col[i - 1]) atomicOr(state, 1u);
        const uint32_t cv = vcode[dict_find(vtab, __builtin_bit_cast(uint64_t, val[i]))];
        const uint32_t cd = dcode[dict_find(dtab, dict_delta_key(col[i], r))];
        const uint32_t sl = slot_of[cd], code = pcode[cv * 256u + cd], sh = 8u * (sl & 3u);
#pragma unroll
        for(uint32_t w = 0; w < 8; ++w)
          if(w == (sl >> 2)) word[w] = (word[w] & ~(0xFFu << sh)) | (code << sh);
      }
    const uint64_t s = r >> 6, lane = r & 63;
#pragma unroll
    for(uint32_t w = 0; w < 8; ++w)
      if(w < nw) codes[(s * nw + w) * 64 + lane] = word[w];
  }
}

// ---------------------------------------------------------------------------
// elementwise kernels
// ---------------------------------------------------------------------------
// bandwidth probe: every thread streams 16-byte nontemporal loads grid-stride and sums
// them; the sum is stored only if it equals a value no real data produces, so the loads
// cannot be dropped and nothing is written
typedef double feat_d2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(kThreads) void k_stream_read(const feat_d2* __restrict__ p, uint64_t n2, double* sink)
{
  double s = 0.0;
  for(uint64_t i = blockIdx.x * uint64_t(kThreads) + threadIdx.x; i < n2; i += uint64_t(gridDim.x) * kThreads)
  {
    const feat_d2 v = __builtin_nontemporal_load(p + i);
    s += v.x + v.y;
  }
  if(s == -1.2345e300) *sink = s;
}

template<typename DT>
__global__ void k_fill(DT* p, DT v, uint64_t n)
{
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

template<typename DO, typename DI>
__global__ void k_convert(DO* o, const DI* in, uint64_t n)
{
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    o[i] = DO(in[i]);
}

__global__ void k_convert_f64_f16(uint16_t* o, const double* in, uint64_t n)
{
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
  {
    _Float16 h = _Float16(in[i]);
    o[i] = __builtin_bit_cast(uint16_t, h);
  }
}

__global__ void k_convert_f16_f64(double* o, const uint16_t* in, uint64_t n)
{
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    o[i] = double(__builtin_bit_cast(_Float16, in[i]));
}

// axpy: r == x (and not y) rounds a*x first, like `r[i] *= a; r[i] += y[i];`
template<typename DT>
__global__ void k_axpy(DT* r, DT a, const DT* x, const DT* y, uint64_t n, int two_round)
{
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
  {
    if(two_round)
    {
      const DT t = x[i] * a;
      r[i] = t + y[i];
    }
    else
      r[i] = dfma(a, x[i], y[i]);
  }
}

__global__ void k_scale(double* r, const double* x, double s, uint64_t n)
{
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    r[i] = x[i] * s;
}

__global__ void k_cprod(double* r, const double* x, const double* y, uint64_t n)
{
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    r[i] = x[i] * y[i];
}

template<typename DT>
__global__ void k_cinv(DT* r, const DT* x, DT s, uint64_t n)
{
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    r[i] = s / x[i];
}

// kind 0: sum x*y; 1: sum x*x (fma); 2: sum x*y*z
__global__ __launch_bounds__(kThreads) void k_reduce(const double* x, const double* y, const double* z,
                                                     uint64_t n, int kind, GridReduce gr)
{
  double acc = 0.0;
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
  {
    if(kind == 0)
      acc = acc + x[i] * y[i];
    else if(kind == 1)
      acc = fma(x[i], x[i], acc);
    else
      acc = fma(x[i] * y[i], z[i], acc);
  }
  grid_reduce_finish(gr, acc, 0.0);
}

__global__ void k_sqrt_inplace(double* p)
{
  if(threadIdx.x == 0 && blockIdx.x == 0) *p = sqrt(*p);
}

template<typename IT>
__global__ void k_diag_index(IT* diag, const IT* col, const IT* rowptr, uint64_t rows)
{
  for(uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < rows; r += uint64_t(gridDim.x) * blockDim.x)
  {
    IT d = rowptr[rows];
    for(IT i = rowptr[r]; i < rowptr[r + 1]; ++i)
      if(uint64_t(col[i]) == r)
      {
        d = i;
        break;
      }
    diag[r] = d;
  }
}

__global__ void k_extract_diag(double* diag, const double* val, const uint64_t* col,
                               const uint64_t* rowptr, uint64_t rows)
{
  for(uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < rows; r += uint64_t(gridDim.x) * blockDim.x)
  {
    double d = 0.0;
    for(uint64_t i = rowptr[r]; i < rowptr[r + 1]; ++i)
      if(col[i] == r)
      {
        d = val[i];
        break;
      }
    diag[r] = d;
  }
}

__global__ void k_filter_rhs(double* v, const double* sv, const uint64_t* idx, uint64_t ue)
{
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < ue; i += uint64_t(gridDim.x) * blockDim.x)
    v[idx[i]] = sv[i];
}

__global__ void k_filter_def(double* v, const uint64_t* idx, uint64_t ue)
{
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < ue; i += uint64_t(gridDim.x) * blockDim.x)
    v[idx[i]] = 0.0;
}

// UnitFilterBlocked: one thread per filtered component (entry i / bs, component i % bs),
// so the value reads are coalesced; RHS writes the value, DEF a zero; with ign_nans a
// NaN value leaves its component alone (unit_filter_blocked_generic.hpp:24-60)
template<bool RHS>
__global__ void k_filter_blocked(double* v, int bs, const double* sv, const uint64_t* idx, uint64_t ue, int ign_nans)
{
  const uint64_t total = ue * uint64_t(bs);
  for(uint64_t t = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; t < total; t += uint64_t(gridDim.x) * blockDim.x)
  {
    const uint64_t e = t / uint64_t(bs), k = t - e * uint64_t(bs);
    if(ign_nans && sv[t] != sv[t]) continue;
    v[idx[e] * uint64_t(bs) + k] = RHS ? sv[t] : 0.0;
  }
}

__global__ void k_mask_zero(uint32_t* m, uint64_t nw)
{
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < nw; i += uint64_t(gridDim.x) * blockDim.x)
    m[i] = 0u;
}

__global__ void k_mask_set(uint32_t* m, const uint64_t* idx, uint64_t ue)
{
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < ue; i += uint64_t(gridDim.x) * blockDim.x)
    atomicOr(&m[idx[i] >> 5], 1u << (idx[i] & 31u));
}

// mean filter: v += (-(v.w_dot)/vol) * w_add; alpha computed from the device scalar
__global__ void k_mean_apply(double* v, const double* w_add, const double* dotv, double vol, uint64_t n)
{
  const double alpha = -(*dotv) / vol;
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    v[i] = fma(alpha, w_add[i], v[i]);
}

// BCSR: one lane per block row, sequential over blocks, h, w (arch/apply_generic.hpp:121-154)
template<int BH, int BW>
__global__ void k_bcsr(double* r, double a, const double* x, double b, const double* y,
                       const double* val, const uint64_t* col, const uint64_t* rowptr,
                       uint64_t rows, int bzero)
{
  for(uint64_t row = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; row < rows; row += uint64_t(gridDim.x) * blockDim.x)
  {
    double bs[BH];
#pragma unroll
    for(int h = 0; h < BH; ++h) bs[h] = 0.0;
    for(uint64_t i = rowptr[row]; i < rowptr[row + 1]; ++i)
    {
      const double* blk = val + i * (BH * BW);
      const double* bx = x + col[i] * BW;
#pragma unroll
      for(int h = 0; h < BH; ++h)
#pragma unroll
        for(int w = 0; w < BW; ++w)
          bs[h] = __fma_rn(blk[h * BW + w], bx[w], bs[h]);
    }
#pragma unroll
    for(int h = 0; h < BH; ++h)
    {
      const double yv = bzero ? 0.0 : y[row * BH + h];
      r[row * BH + h] = __fma_rn(bs[h], a, b * yv);
    }
  }
}

// Richardson(Jacobi) first step from x = 0: x_1 = omega*filter(inv∘rhs); optional acc += x_1
template<typename DT>
__global__ void k_smooth_first(DT* x1, const DT* inv, const DT* rhs, const uint32_t* mask, DT omega,
                               const double* rhs_nrm2, DT* acc, const DT* dotw, uint64_t n,
                               GridReduce gr)
{
  const bool skip = smoother_skips<DT>(rhs_nrm2);
  double a1 = 0.0;
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
  {
    DT c = inv[i] * rhs[i];
    if(is_masked(mask, i)) c = DT(0);
    DT xn = dfma(omega, c, DT(0));
    if(skip) xn = DT(0);
    if(acc)
    {
      const DT a = dfma(DT(1), xn, acc[i]);
      acc[i] = a;
      if(dotw) a1 = a1 + double(a) * double(dotw[i]);
    }
    else
      x1[i] = xn;
  }
  grid_reduce_finish(gr, 0.0, a1);
}

template<typename DT>
__global__ void k_dense_gemv(int n, const DT* ainv, const DT* b, DT* x, const uint32_t* mask)
{
  for(int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
  {
    DT s = DT(0);
    for(int k = 0; k < n; ++k) s = dfma(ainv[size_t(r) * n + k], b[k], s);
    if(is_masked(mask, uint64_t(r))) s = DT(0);
    x[r] = s;
  }
}

__global__ __launch_bounds__(kThreads) void k_pcg_update(double* x, double* r, const double* p, const double* q,
                                                         const double* gamma, const double* qp, uint64_t n,
                                                         GridReduce gr, FirstIter<double> first,
                                                         const uint32_t* first_mask)
{
  const double alpha = (*gamma) / (*qp);
  double acc = 0.0;
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
  {
    const double inv = first.load(i);
    x[i] = fma(alpha, p[i], x[i]);
    const double rn = fma(-alpha, q[i], r[i]);
    r[i] = rn;
    first.store(i, rn, inv, is_masked(first_mask, i));
    acc = fma(rn, rn, acc);
  }
  grid_reduce_finish(gr, acc, 0.0);
}

__global__ void k_pcg_direction(double* p, const double* z, const double* gamma_new, double* gamma_old, uint64_t n)
{
  const double beta = (*gamma_new) / (*gamma_old);
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
  {
    const double t = p[i] * beta;
    p[i] = t + z[i];
  }
}

__global__ void k_copy_scalar(double* dst, const double* src)
{
  if(threadIdx.x == 0 && blockIdx.x == 0) *dst = *src;
}

} // namespace

extern "C" {

const char* feat_last_error(void) { return feat3amd_detail::g_err.c_str(); }

int feat_handle_create(feat_handle_t* out, int device, feat_stream_t stream)
{
  if(!out) return fail(FEAT_ERR_INVALID, "feat_handle_create: null handle pointer");
  int ndev = 0;
  if(hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0)
    return fail(FEAT_ERR_NO_DEVICE, "feat_handle_create: no such HIP device");
  int rc = check_hip(hipSetDevice(device), "hipSetDevice");
  if(rc) return rc;
  feat_handle_s* h = new feat_handle_s();
  h->device = device;
  // NULL is the device's null stream (ordered with every blocking stream,
  // torch's default stream included); the handle never creates a stream.
  h->own_stream = false;
  h->stream = hipStream_t(stream);
  rc = check_hip(hipMalloc(&h->partials, sizeof(double) * 2 * kMaxGrid), "hipMalloc(partials)");
  if(!rc) rc = check_hip(hipMalloc(&h->counter, 64), "hipMalloc(counter)");
  if(!rc) rc = check_hip(hipMalloc(&h->scal, sizeof(double) * 16), "hipMalloc(scal)");
  if(!rc) rc = check_hip(hipHostMalloc(&h->host, sizeof(double) * 16, 0), "hipHostMalloc");
  if(!rc) rc = check_hip(hipEventCreateWithFlags(&h->event, hipEventDisableTiming), "hipEventCreate");
  if(!rc) rc = feat3amd_detail::noalloc_scratch_attach(device);
  if(!rc) rc = check_hip(hipMemsetAsync(h->counter, 0, 64, h->stream), "hipMemset");
  if(!rc) rc = check_hip(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
  if(rc)
  {
    feat_handle_destroy(h);
    return rc;
  }
  *out = h;
  return FEAT_OK;
}

int feat_handle_destroy(feat_handle_t h)
{
  if(!h) return FEAT_OK;
  (void)hipSetDevice(h->device);
  (void)hipStreamSynchronize(h->stream);
  if(h->partials) (void)hipFree(h->partials);
  if(h->counter) (void)hipFree(h->counter);
  if(h->scal) (void)hipFree(h->scal);
  if(h->host) (void)hipHostFree(h->host);
  if(h->event) (void)hipEventDestroy(h->event);
  if(h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return FEAT_OK;
}

int feat_stream_create(feat_stream_t* s)
{
  if(!s) return fail(FEAT_ERR_INVALID, "feat_stream_create: null pointer");
  hipStream_t st = nullptr;
  int rc = check_hip(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreateWithFlags");
  if(!rc) *s = st;
  return rc;
}

int feat_stream_destroy(feat_stream_t s)
{
  return s ? check_hip(hipStreamDestroy(hipStream_t(s)), "hipStreamDestroy") : FEAT_OK;
}

int feat_handle_wait(feat_handle_t waiter, feat_handle_t signaler)
{
  if(!waiter || !signaler) return fail(FEAT_ERR_INVALID, "feat_handle_wait: null handle");
  if(waiter->stream == signaler->stream) return FEAT_OK;
  int rc = check_hip(hipEventRecord(signaler->event, signaler->stream), "hipEventRecord");
  if(!rc) rc = check_hip(hipStreamWaitEvent(waiter->stream, signaler->event, 0), "hipStreamWaitEvent");
  return rc;
}

int feat_graph_begin(feat_handle_t h)
{
  if(!h || !h->stream) return fail(FEAT_ERR_INVALID, "feat_graph_begin: needs a handle on a non-null stream");
  return check_hip(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
}

int feat_graph_end(feat_handle_t h, feat_graph_t* g)
{
  if(!h || !g) return fail(FEAT_ERR_INVALID, "feat_graph_end: null argument");
  *g = nullptr;
  hipGraph_t graph = nullptr;
  int rc = check_hip(hipStreamEndCapture(h->stream, &graph), "hipStreamEndCapture");
  if(rc) return rc;
  hipGraphExec_t exec = nullptr;
  rc = check_hip(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  if(rc)
  {
    (void)hipGraphDestroy(graph);
    return rc;
  }
  *g = new feat_graph_s{graph, exec};
  return FEAT_OK;
}

int feat_graph_launch(feat_handle_t h, feat_graph_t g)
{
  if(!h || !g) return fail(FEAT_ERR_INVALID, "feat_graph_launch: null argument");
  return check_hip(hipGraphLaunch(g->exec, h->stream), "hipGraphLaunch");
}

int feat_graph_destroy(feat_graph_t g)
{
  if(!g) return FEAT_OK;
  (void)hipGraphExecDestroy(g->exec);
  (void)hipGraphDestroy(g->graph);
  delete g;
  return FEAT_OK;
}

int feat_copy_scalar_to_host(feat_handle_t h, const double* dev, double* host_pinned)
{
  if(!h || !dev || !host_pinned) return fail(FEAT_ERR_INVALID, "feat_copy_scalar_to_host: null argument");
  return check_hip(hipMemcpyAsync(host_pinned, dev, sizeof(double), hipMemcpyDeviceToHost, h->stream), "hipMemcpyAsync");
}

int feat_handle_stream(feat_handle_t h, feat_stream_t* s)
{
  if(!h || !s) return fail(FEAT_ERR_INVALID, "feat_handle_stream: null argument");
  *s = h->stream;
  return FEAT_OK;
}

int feat_sync(feat_handle_t h)
{
  if(!h) return fail(FEAT_ERR_INVALID, "feat_sync: null handle");
  return check_hip(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
}

int feat_malloc(void** p, size_t bytes)
{
  if(!p) return fail(FEAT_ERR_INVALID, "feat_malloc: null pointer");
  if(bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if(e != hipSuccess) return fail(FEAT_ERR_NOMEM, "feat_malloc: hipMalloc failed");
  return FEAT_OK;
}

int feat_free(void* p) { return p ? check_hip(hipFree(p), "hipFree") : FEAT_OK; }

int feat_host_malloc(void** p, size_t bytes)
{
  if(!p) return fail(FEAT_ERR_INVALID, "feat_host_malloc: null pointer");
  hipError_t e = hipHostMalloc(p, bytes ? bytes : 8, 0);
  if(e != hipSuccess) return fail(FEAT_ERR_NOMEM, "feat_host_malloc: hipHostMalloc failed");
  return FEAT_OK;
}

int feat_host_free(void* p) { return p ? check_hip(hipHostFree(p), "hipHostFree") : FEAT_OK; }

int feat_copy_device(feat_handle_t h, void* dst, const void* src, size_t bytes)
{
  if(!h || (bytes && (!dst || !src))) return fail(FEAT_ERR_INVALID, "feat_copy_device: null argument");
  if(!bytes || dst == src) return FEAT_OK;
  if(((uintptr_t(dst) | uintptr_t(src) | bytes) & 15) != 0)
    return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, h->stream), "hipMemcpyAsync");
  const uint64_t n16 = bytes / 16;
  const uint64_t nb = (n16 + kThreads - 1) / kThreads;
  const int grid = int(nb < uint64_t(kMaxGrid) ? nb : uint64_t(kMaxGrid));
  hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(kThreads), 0, h->stream, (const uint4*)src, (uint4*)dst, n16);
  return check_launch("feat_copy_device");
}

int feat_copy(feat_handle_t h, void* dst, const void* src, size_t bytes)
{
  if(!h || (bytes && (!dst || !src))) return fail(FEAT_ERR_INVALID, "feat_copy: null argument");
  if(!bytes || dst == src) return FEAT_OK;
  return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, h->stream), "hipMemcpyAsync");
}

int feat_stream_read_f64(feat_handle_t h, const double* x, uint64_t n, double* sink)
{
  if(!h || (n && !x) || !sink) return fail(FEAT_ERR_INVALID, "feat_stream_read_f64: null argument");
  if(n & 1) return fail(FEAT_ERR_INVALID, "feat_stream_read_f64: n must be even");
  if((reinterpret_cast<uintptr_t>(x) & 15) != 0) return fail(FEAT_ERR_INVALID, "feat_stream_read_f64: x not 16-byte aligned");
  if(!n) return FEAT_OK;
  // 4096 workgroups (16 per CU) measured as the fastest read shape on 128 MiB - 1 GiB
  const uint64_t n2 = n / 2;
  const uint64_t blocks = (n2 + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(k_stream_read, dim3(blocks < 4096 ? blocks : 4096), dim3(kThreads), 0, h->stream,
                     reinterpret_cast<const feat_d2*>(x), n2, sink);
  return check_launch("feat_stream_read_f64");
}

int feat_set_memory_f64(feat_handle_t h, double* p, double v, uint64_t n)
{
  if(!h || (n && !p)) return fail(FEAT_ERR_INVALID, "feat_set_memory_f64: null argument");
  if(!n) return FEAT_OK;
  hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(n, 4)), dim3(kThreads), 0, h->stream, p, v, n);
  return check_launch("k_fill");
}

int feat_set_memory_f32(feat_handle_t h, float* p, float v, uint64_t n)
{
  if(!h || (n && !p)) return fail(FEAT_ERR_INVALID, "feat_set_memory_f32: null argument");
  if(!n) return FEAT_OK;
  hipLaunchKernelGGL(k_fill<float>, dim3(grid_for(n, 4)), dim3(kThreads), 0, h->stream, p, v, n);
  return check_launch("k_fill");
}

#define CONVERT_FN(NAME, DO, DI)                                                                  \
  int NAME(feat_handle_t h, DO* dst, const DI* src, uint64_t n)                                  \
  {                                                                                               \
    if(!h || (n && (!dst || !src))) return fail(FEAT_ERR_INVALID, #NAME ": null argument");     \
    if(!n) return FEAT_OK;                                                                        \
    hipLaunchKernelGGL((k_convert<DO, DI>), dim3(grid_for(n, 4)), dim3(kThreads), 0, h->stream, \
                       dst, src, n);                                                              \
    return check_launch(#NAME);                                                                   \
  }
CONVERT_FN(feat_convert_f64_f32, float, double)
CONVERT_FN(feat_convert_f32_f64, double, float)
CONVERT_FN(feat_convert_u64_u32, uint32_t, uint64_t)
#undef CONVERT_FN

int feat_convert_f64_f16(feat_handle_t h, uint16_t* dst, const double* src, uint64_t n)
{
  if(!h || (n && (!dst || !src))) return fail(FEAT_ERR_INVALID, "feat_convert_f64_f16: null argument");
  if(!n) return FEAT_OK;
  hipLaunchKernelGGL(k_convert_f64_f16, dim3(grid_for(n, 4)), dim3(kThreads), 0, h->stream, dst, src, n);
  return check_launch("k_convert_f64_f16");
}

int feat_convert_f16_f64(feat_handle_t h, double* dst, const uint16_t* src, uint64_t n)
{
  if(!h || (n && (!dst || !src))) return fail(FEAT_ERR_INVALID, "feat_convert_f16_f64: null argument");
  if(!n) return FEAT_OK;
  hipLaunchKernelGGL(k_convert_f16_f64, dim3(grid_for(n, 4)), dim3(kThreads), 0, h->stream, dst, src, n);
  return check_launch("k_convert_f16_f64");
}

// ---------------------------------------------------------------------------
// SpMV
// ---------------------------------------------------------------------------
} // extern "C"

namespace {
template<typename DT, typename IT>
int csr_apply_impl(feat_handle_t h, DT* r, DT a, const DT* x, DT b, const DT* y, const DT* val,
                   const IT* col, const IT* rowptr, uint64_t rows, uint64_t cols, uint64_t nnz,
                   int transposed, const char* name)
{
  if(!h) return fail(FEAT_ERR_INVALID, "csr_apply: null handle");
  const DT eps = std::is_same<DT, float>::value ? DT(FLT_EPSILON) : DT(DBL_EPSILON);
  const bool bzero = (b < 0 ? -b : b) < eps;
  const uint64_t nr = transposed ? cols : rows;
  const uint64_t nx = transposed ? rows : cols;
  if((nr && !r) || (nx && !x) || (nnz && (!val || !col)) || !rowptr || (!bzero && nr && !y))
    return fail(FEAT_ERR_INVALID, "csr_apply: null argument");
  if(r == x && nr) return fail(FEAT_ERR_ALIAS, "csr_apply: r and x must not alias");
  if(!transposed)
  {
    EpiApply<DT> e{r, y, a, b, bzero};
    GridReduce g{nullptr, nullptr, nullptr, nullptr};
    return launch_rows<DT, IT>(h->stream, val, col, rowptr, x, rows, e, g, name);
  }
  // r = (b/a)*y ; r[col] += val*x[row] ; r = a*r    (arch/apply_generic.hpp:37-53)
  return feat3amd_detail::csr_apply_transposed<DT, IT>(h, r, a, x, b, bzero, y, val, col, rowptr, rows, cols, nnz, name);
}
} // namespace

extern "C" {

int feat_csr_apply_transposed_noalloc_f64_u64(feat_handle_t h, double* r, double a, const double* x, double b,
                                              const double* y, const double* val, const uint64_t* col,
                                              const uint64_t* rowptr, uint64_t rows, uint64_t cols, uint64_t nnz)
{
  const char* name = "feat_csr_apply_transposed_noalloc_f64_u64";
  if(!h) return fail(FEAT_ERR_INVALID, "csr_apply_transposed_noalloc: null handle");
  const bool bzero = (b < 0 ? -b : b) < DBL_EPSILON;
  if((cols && !r) || (rows && !x) || (nnz && (!val || !col)) || !rowptr || (!bzero && cols && !y))
    return fail(FEAT_ERR_INVALID, "csr_apply_transposed_noalloc: null argument");
  if(r == x && cols) return fail(FEAT_ERR_ALIAS, "csr_apply_transposed_noalloc: r and x must not alias");
  return feat3amd_detail::csr_apply_transposed_noalloc<double, uint64_t>(h, r, a, x, b, bzero, y, val, col, rowptr, rows,
                                                                         cols, nnz, name);
}

int feat_csr_apply_f64_u64(feat_handle_t h, double* r, double a, const double* x, double b,
                           const double* y, const double* val, const uint64_t* col,
                           const uint64_t* rowptr, uint64_t rows, uint64_t cols, uint64_t nnz,
                           int transposed)
{
  return csr_apply_impl<double, uint64_t>(h, r, a, x, b, y, val, col, rowptr, rows, cols, nnz,
                                          transposed, "feat_csr_apply_f64_u64");
}

int feat_csr_apply_f64_u32(feat_handle_t h, double* r, double a, const double* x, double b,
                           const double* y, const double* val, const uint32_t* col,
                           const uint32_t* rowptr, uint64_t rows, uint64_t cols, uint64_t nnz,
                           int transposed)
{
  return csr_apply_impl<double, uint32_t>(h, r, a, x, b, y, val, col, rowptr, rows, cols, nnz,
                                          transposed, "feat_csr_apply_f64_u32");
}

int feat_csr_apply_f32_u32(feat_handle_t h, float* r, float a, const float* x, float b,
                           const float* y, const float* val, const uint32_t* col,
                           const uint32_t* rowptr, uint64_t rows, uint64_t cols, uint64_t nnz,
                           int transposed)
{
  return csr_apply_impl<float, uint32_t>(h, r, a, x, b, y, val, col, rowptr, rows, cols, nnz,
                                         transposed, "feat_csr_apply_f32_u32");
}

int feat_csr_apply_f32_u64(feat_handle_t h, float* r, float a, const float* x, float b,
                           const float* y, const float* val, const uint64_t* col,
                           const uint64_t* rowptr, uint64_t rows, uint64_t cols, uint64_t nnz,
                           int transposed)
{
  return csr_apply_impl<float, uint64_t>(h, r, a, x, b, y, val, col, rowptr, rows, cols, nnz,
                                         transposed, "feat_csr_apply_f32_u64");
}

int feat_bcsr_apply_f64_u64(feat_handle_t h, int bh, int bw, double* r, double a,
                            const double* x, double b, const double* y, const double* val,
                            const uint64_t* col, const uint64_t* rowptr, uint64_t rows,
                            uint64_t cols, uint64_t nnz)
{
  if(!h) return fail(FEAT_ERR_INVALID, "feat_bcsr_apply: null handle");
  const bool bzero = std::fabs(b) < DBL_EPSILON;
  if((rows && !r) || (cols && !x) || (nnz && (!val || !col)) || !rowptr || (!bzero && rows && !y))
    return fail(FEAT_ERR_INVALID, "feat_bcsr_apply: null argument");
  if(r == x && rows) return fail(FEAT_ERR_ALIAS, "feat_bcsr_apply: r and x must not alias");
  if(!rows) return FEAT_OK;
  dim3 g(grid_for(rows)), t(kThreads);
  if(bh == 2 && bw == 2)
    hipLaunchKernelGGL((k_bcsr<2, 2>), g, t, 0, h->stream, r, a, x, b, y, val, col, rowptr, rows, int(bzero));
  else if(bh == 2 && bw == 1)
    hipLaunchKernelGGL((k_bcsr<2, 1>), g, t, 0, h->stream, r, a, x, b, y, val, col, rowptr, rows, int(bzero));
  else if(bh == 1 && bw == 2)
    hipLaunchKernelGGL((k_bcsr<1, 2>), g, t, 0, h->stream, r, a, x, b, y, val, col, rowptr, rows, int(bzero));
  else if(bh == 3 && bw == 3)
    hipLaunchKernelGGL((k_bcsr<3, 3>), g, t, 0, h->stream, r, a, x, b, y, val, col, rowptr, rows, int(bzero));
  else
    return fail(FEAT_ERR_UNSUPPORTED, "feat_bcsr_apply: block shape not provided");
  return check_launch("feat_bcsr_apply_f64_u64");
}

// ---------------------------------------------------------------------------
// BLAS-1
// ---------------------------------------------------------------------------
int feat_axpy_f64(feat_handle_t h, double* r, double a, const double* x, const double* y, uint64_t n)
{
  if(!h || (n && (!r || !x || !y))) return fail(FEAT_ERR_INVALID, "feat_axpy_f64: null argument");
  if(!n) return FEAT_OK;
  const int two = (r == x && r != y) ? 1 : 0;
  hipLaunchKernelGGL(k_axpy<double>, dim3(grid_for(n, 4)), dim3(kThreads), 0, h->stream, r, a, x, y, n, two);
  return check_launch("feat_axpy_f64");
}

int feat_axpy_f32(feat_handle_t h, float* r, float a, const float* x, const float* y, uint64_t n)
{
  if(!h || (n && (!r || !x || !y))) return fail(FEAT_ERR_INVALID, "feat_axpy_f32: null argument");
  if(!n) return FEAT_OK;
  const int two = (r == x && r != y) ? 1 : 0;
  hipLaunchKernelGGL(k_axpy<float>, dim3(grid_for(n, 4)), dim3(kThreads), 0, h->stream, r, a, x, y, n, two);
  return check_launch("feat_axpy_f32");
}

} // extern "C"

namespace {
int reduce_impl(feat_handle_t h, double* res, const double* x, const double* y, const double* z,
                uint64_t n, int kind, bool take_sqrt, const char* name)
{
  if(!h || !res || (n && (!x || !y || (kind == 2 && !z)))) return fail(FEAT_ERR_INVALID, name);
  if(!n)
    return feat_set_memory_f64(h, res, 0.0, 1);
  const int grid = grid_for(n, 8);
  const GridReduce gd = reduce_for_grid(handle_reduce(h, res, nullptr), uint64_t(grid));
  hipLaunchKernelGGL(k_reduce, dim3(grid), dim3(kThreads), 0, h->stream, x, y, z, n, kind, gd);
  finish_reduce(h->stream, gd, uint64_t(grid));
  if(take_sqrt)
    hipLaunchKernelGGL(k_sqrt_inplace, dim3(1), dim3(1), 0, h->stream, res);
  return check_launch(name);
}

int to_host(feat_handle_t h, const double* dev, double* host)
{
  if(!host) return fail(FEAT_ERR_INVALID, "null host result");
  int rc = check_hip(hipMemcpyAsync(h->host, dev, sizeof(double), hipMemcpyDeviceToHost, h->stream), "hipMemcpyAsync");
  if(!rc) rc = check_hip(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
  if(!rc) *host = h->host[0];
  return rc;
}
} // namespace

extern "C" {

int feat_dot_f64(feat_handle_t h, double* res, const double* x, const double* y, uint64_t n)
{
  const int kind = (x == y) ? 1 : 0;
  return reduce_impl(h, res, x, y, nullptr, n, kind, false, "feat_dot_f64");
}

int feat_dot_f64_host(feat_handle_t h, double* res_host, const double* x, const double* y, uint64_t n)
{
  if(!h) return fail(FEAT_ERR_INVALID, "feat_dot_f64_host: null handle");
  int rc = feat_dot_f64(h, h->scal, x, y, n);
  return rc ? rc : to_host(h, h->scal, res_host);
}

int feat_triple_dot_f64_host(feat_handle_t h, double* res_host, const double* x, const double* y,
                             const double* z, uint64_t n)
{
  if(!h) return fail(FEAT_ERR_INVALID, "feat_triple_dot_f64_host: null handle");
  int rc = reduce_impl(h, h->scal, x, y, z, n, 2, false, "feat_triple_dot_f64_host");
  return rc ? rc : to_host(h, h->scal, res_host);
}

int feat_norm2_f64(feat_handle_t h, double* res, const double* x, uint64_t n)
{
  return reduce_impl(h, res, x, x, nullptr, n, 1, true, "feat_norm2_f64");
}

int feat_norm2_f64_host(feat_handle_t h, double* res_host, const double* x, uint64_t n)
{
  if(!h) return fail(FEAT_ERR_INVALID, "feat_norm2_f64_host: null handle");
  int rc = feat_norm2_f64(h, h->scal, x, n);
  return rc ? rc : to_host(h, h->scal, res_host);
}

int feat_scale_f64(feat_handle_t h, double* r, const double* x, double s, uint64_t n)
{
  if(!h || (n && (!r || !x))) return fail(FEAT_ERR_INVALID, "feat_scale_f64: null argument");
  if(!n) return FEAT_OK;
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n, 4)), dim3(kThreads), 0, h->stream, r, x, s, n);
  return check_launch("feat_scale_f64");
}

int feat_component_product_f64(feat_handle_t h, double* r, const double* x, const double* y, uint64_t n)
{
  if(!h || (n && (!r || !x || !y))) return fail(FEAT_ERR_INVALID, "feat_component_product_f64: null argument");
  if(!n) return FEAT_OK;
  hipLaunchKernelGGL(k_cprod, dim3(grid_for(n, 4)), dim3(kThreads), 0, h->stream, r, x, y, n);
  return check_launch("feat_component_product_f64");
}

int feat_component_invert_f64(feat_handle_t h, double* r, const double* x, double s, uint64_t n)
{
  if(!h || (n && (!r || !x))) return fail(FEAT_ERR_INVALID, "feat_component_invert_f64: null argument");
  if(!n) return FEAT_OK;
  hipLaunchKernelGGL(k_cinv<double>, dim3(grid_for(n, 4)), dim3(kThreads), 0, h->stream, r, x, s, n);
  return check_launch("feat_component_invert_f64");
}

int feat_component_invert_f32(feat_handle_t h, float* r, const float* x, float s, uint64_t n)
{
  if(!h || (n && (!r || !x))) return fail(FEAT_ERR_INVALID, "feat_component_invert_f32: null argument");
  if(!n) return FEAT_OK;
  hipLaunchKernelGGL(k_cinv<float>, dim3(grid_for(n, 4)), dim3(kThreads), 0, h->stream, r, x, s, n);
  return check_launch("feat_component_invert_f32");
}

int feat_csr_diag_index_u64(feat_handle_t h, uint64_t* diag, const uint64_t* col, const uint64_t* rowptr,
                            uint64_t rows)
{
  if(!h || !rowptr || (rows && (!diag || !col))) return fail(FEAT_ERR_INVALID, "feat_csr_diag_index_u64: null argument");
  if(!rows) return FEAT_OK;
  hipLaunchKernelGGL(k_diag_index<uint64_t>, dim3(grid_for(rows)), dim3(kThreads), 0, h->stream, diag, col, rowptr, rows);
  return check_launch("feat_csr_diag_index_u64");
}

int feat_csr_extract_diag_f64_u64(feat_handle_t h, double* diag, const double* val, const uint64_t* col,
                                  const uint64_t* rowptr, uint64_t rows)
{
  if(!h || !rowptr || (rows && (!diag || !col || !val))) return fail(FEAT_ERR_INVALID, "feat_csr_extract_diag: null argument");
  if(!rows) return FEAT_OK;
  hipLaunchKernelGGL(k_extract_diag, dim3(grid_for(rows)), dim3(kThreads), 0, h->stream, diag, val, col, rowptr, rows);
  return check_launch("feat_csr_extract_diag_f64_u64");
}

// ---------------------------------------------------------------------------
// filters
// ---------------------------------------------------------------------------
int feat_unit_filter_rhs_f64_u64(feat_handle_t h, double* v, const double* sv_val, const uint64_t* sv_idx, uint64_t ue)
{
  if(!h || (ue && (!v || !sv_val || !sv_idx))) return fail(FEAT_ERR_INVALID, "feat_unit_filter_rhs: null argument");
  if(!ue) return FEAT_OK;
  hipLaunchKernelGGL(k_filter_rhs, dim3(grid_for(ue)), dim3(kThreads), 0, h->stream, v, sv_val, sv_idx, ue);
  return check_launch("feat_unit_filter_rhs_f64_u64");
}

int feat_unit_filter_def_f64_u64(feat_handle_t h, double* v, const uint64_t* sv_idx, uint64_t ue)
{
  if(!h || (ue && (!v || !sv_idx))) return fail(FEAT_ERR_INVALID, "feat_unit_filter_def: null argument");
  if(!ue) return FEAT_OK;
  hipLaunchKernelGGL(k_filter_def, dim3(grid_for(ue)), dim3(kThreads), 0, h->stream, v, sv_idx, ue);
  return check_launch("feat_unit_filter_def_f64_u64");
}

int feat_unit_filter_mask_u64(feat_handle_t h, uint32_t* mask, uint64_t n, const uint64_t* sv_idx, uint64_t ue)
{
  const uint64_t nw = (n + 31) / 32;
  if(!h || (nw && !mask) || (ue && !sv_idx)) return fail(FEAT_ERR_INVALID, "feat_unit_filter_mask: null argument");
  if(nw)
    hipLaunchKernelGGL(k_mask_zero, dim3(grid_for(nw)), dim3(kThreads), 0, h->stream, mask, nw);
  if(ue)
    hipLaunchKernelGGL(k_mask_set, dim3(grid_for(ue)), dim3(kThreads), 0, h->stream, mask, sv_idx, ue);
  return check_launch("feat_unit_filter_mask_u64");
}

int feat_unit_filter_blocked_rhs_f64_u64(feat_handle_t h, int bs, double* v, const double* sv_elements,
                                         const uint64_t* sv_indices, uint64_t ue, int ign_nans)
{
  if(!h || bs <= 0 || (ue && (!v || !sv_elements || !sv_indices)))
    return fail(FEAT_ERR_INVALID, "feat_unit_filter_blocked_rhs: bad argument");
  if(!ue) return FEAT_OK;
  hipLaunchKernelGGL(k_filter_blocked<true>, dim3(grid_for(ue * uint64_t(bs))), dim3(kThreads), 0, h->stream, v, bs,
                     sv_elements, sv_indices, ue, ign_nans);
  return check_launch("feat_unit_filter_blocked_rhs_f64_u64");
}

int feat_unit_filter_blocked_def_f64_u64(feat_handle_t h, int bs, double* v, const double* sv_elements,
                                         const uint64_t* sv_indices, uint64_t ue, int ign_nans)
{
  if(!h || bs <= 0 || (ue && (!v || !sv_indices || (ign_nans && !sv_elements))))
    return fail(FEAT_ERR_INVALID, "feat_unit_filter_blocked_def: bad argument");
  if(!ue) return FEAT_OK;
  hipLaunchKernelGGL(k_filter_blocked<false>, dim3(grid_for(ue * uint64_t(bs))), dim3(kThreads), 0, h->stream, v, bs,
                     sv_elements, sv_indices, ue, ign_nans);
  return check_launch("feat_unit_filter_blocked_def_f64_u64");
}

int feat_mean_filter_f64(feat_handle_t h, double* v, const double* w_dot, const double* w_add, double volume, uint64_t n)
{
  if(!h || (n && (!v || !w_dot || !w_add))) return fail(FEAT_ERR_INVALID, "feat_mean_filter_f64: null argument");
  if(!n) return FEAT_OK;
  int rc = reduce_impl(h, h->scal + 1, v, w_dot, nullptr, n, (v == w_dot) ? 1 : 0, false, "feat_mean_filter_f64");
  if(rc) return rc;
  hipLaunchKernelGGL(k_mean_apply, dim3(grid_for(n, 4)), dim3(kThreads), 0, h->stream, v, w_add, h->scal + 1, volume, n);
  return check_launch("feat_mean_filter_f64");
}

// ---------------------------------------------------------------------------
// matrix descriptors and SELL-64 plans
// ---------------------------------------------------------------------------
int feat_matrix_from_csr_f64(const feat_csr_f64_t* A, feat_matrix_f64_t* out)
{
  if(!A || !out) return fail(FEAT_ERR_INVALID, "feat_matrix_from_csr_f64: null argument");
  *out = feat_matrix_f64_t{FEAT_FMT_CSR, FEAT_IDX_U64, A->rows, A->cols, A->nnz, A->val, A->col, A->rowptr, nullptr,
                           FEAT_VAL_NATIVE, 1.0f};
  return FEAT_OK;
}

} // extern "C"

namespace {
template<typename M>
int dlinear_desc(int dim, uint32_t n1, int transpose, M* out, const char* name)
{
  if(!out || (dim != 2 && dim != 3) || n1 < 3 || !(n1 & 1) || (transpose != 0 && transpose != 1))
    return fail(FEAT_ERR_INVALID, name);
  const uint64_t nc = (uint64_t(n1) + 1) / 2;
  const uint64_t nf = dim == 3 ? uint64_t(n1) * n1 * n1 : uint64_t(n1) * n1, ncn = dim == 3 ? nc * nc * nc : nc * nc;
  *out = M{};
  out->format = FEAT_FMT_DLINEAR;
  out->index = dim;
  out->rows = transpose ? ncn : nf;
  out->cols = transpose ? nf : ncn;
  out->value_kind = transpose;
  out->scale = 1.0f;
  out->slice_words = n1;
  return FEAT_OK;
}
} // namespace

extern "C" {

int feat_matrix_dlinear_f64(int dim, uint32_t n1, int transpose, feat_matrix_f64_t* out)
{
  return dlinear_desc(dim, n1, transpose, out, "feat_matrix_dlinear_f64");
}

int feat_matrix_dlinear_f32(int dim, uint32_t n1, int transpose, feat_matrix_f32_t* out)
{
  return dlinear_desc(dim, n1, transpose, out, "feat_matrix_dlinear_f32");
}

} // extern "C"

namespace {
// Pair-dictionary plan of a scanned CSR (slice maxima hmax).  Returns FEAT_OK with *out
// filled, or FEAT_ERR_UNSUPPORTED when the matrix has more than 256 distinct
// (column - row, value) pairs (the caller then builds the plain plan).
template<typename DT, typename M>
int dict_create(feat_handle_t h, const feat_csr_f64_t* A, const std::vector<uint32_t>& hmax, uint16_t* rlen, M* out,
                const char* name, int flags)
{
  const uint64_t n = A->rows, ns = (n + 63) / 64;
  uint64_t *vtab = nullptr, *dtab = nullptr, *soff = nullptr;
  uint32_t *state = nullptr, *present = nullptr, *codes = nullptr;
  uint8_t* maps = nullptr;  // vcode[kDictSlots], dcode[kDictSlots], pcode[65536]
  DT* table = nullptr;
  auto cleanup = [&]() {
    for(void* p : {(void*)vtab, (void*)dtab, (void*)soff, (void*)state, (void*)present, (void*)codes, (void*)maps,
                   (void*)table})
      if(p) (void)hipFree(p);
  };
  const size_t tb = sizeof(uint64_t) * kDictSlots;
  int rc = check_hip(hipMalloc(&vtab, tb), "dict table");
  if(!rc) rc = check_hip(hipMalloc(&dtab, tb), "dict table");
  if(!rc) rc = check_hip(hipMalloc(&state, 3 * sizeof(uint32_t)), "dict state");
  if(!rc) rc = check_hip(hipMemsetAsync(vtab, 0xFF, tb, h->stream), "memset");
  if(!rc) rc = check_hip(hipMemsetAsync(dtab, 0xFF, tb, h->stream), "memset");
  if(!rc) rc = check_hip(hipMemsetAsync(state, 0, 3 * sizeof(uint32_t), h->stream), "memset");
  if(rc) { cleanup(); return rc; }
  hipLaunchKernelGGL(k_dict_collect, dim3(grid_for(n)), dim3(kThreads), 0, h->stream, A->val, A->col, A->rowptr, n,
                     vtab, dtab, state);
  std::vector<uint64_t> hv(kDictSlots), hd(kDictSlots);
  uint32_t hs[3] = {0, 0, 0};
  rc = check_launch(name);
  if(!rc) rc = check_hip(hipMemcpyAsync(hv.data(), vtab, tb, hipMemcpyDeviceToHost, h->stream), "copy");
  if(!rc) rc = check_hip(hipMemcpyAsync(hd.data(), dtab, tb, hipMemcpyDeviceToHost, h->stream), "copy");
  if(!rc) rc = check_hip(hipMemcpyAsync(hs, state, sizeof(hs), hipMemcpyDeviceToHost, h->stream), "copy");
  if(!rc) rc = check_hip(hipStreamSynchronize(h->stream), "sync");
  if(rc) { cleanup(); return rc; }
  if(hs[2] || hs[0] > 256 || hs[1] > 256)
  {
    cleanup();
    return FEAT_ERR_UNSUPPORTED;
  }
  // codes of the distinct keys in ascending key order (deterministic)
  std::vector<uint8_t> hmap(2 * kDictSlots + 65536 + 256, 0);
  std::vector<std::pair<uint64_t, uint32_t>> kv, kd;
  for(uint32_t i = 0; i < kDictSlots; ++i)
  {
    if(hv[i] != kDictEmpty) kv.push_back({hv[i], i});
    if(hd[i] != kDictEmpty) kd.push_back({hd[i], i});
  }
  std::sort(kv.begin(), kv.end());
  std::sort(kd.begin(), kd.end());
  for(size_t c = 0; c < kv.size(); ++c) hmap[kv[c].second] = uint8_t(c);
  for(size_t c = 0; c < kd.size(); ++c) hmap[kDictSlots + kd[c].second] = uint8_t(c);
  rc = check_hip(hipMalloc(&maps, hmap.size()), "dict maps");
  if(!rc) rc = check_hip(hipMalloc(&present, sizeof(uint32_t) * 65536), "dict pairs");
  if(!rc) rc = check_hip(hipMemcpyAsync(maps, hmap.data(), hmap.size(), hipMemcpyHostToDevice, h->stream), "copy");
  if(!rc) rc = check_hip(hipMemsetAsync(present, 0, sizeof(uint32_t) * 65536, h->stream), "memset");
  if(rc) { cleanup(); return rc; }
  hipLaunchKernelGGL(k_dict_mark, dim3(grid_for(n)), dim3(kThreads), 0, h->stream, A->val, A->col, A->rowptr, n, vtab,
                     dtab, maps, maps + kDictSlots, present);
  std::vector<uint32_t> hp(65536);
  rc = check_launch(name);
  if(!rc) rc = check_hip(hipMemcpyAsync(hp.data(), present, sizeof(uint32_t) * 65536, hipMemcpyDeviceToHost, h->stream), "copy");
  if(!rc) rc = check_hip(hipStreamSynchronize(h->stream), "sync");
  if(rc) { cleanup(); return rc; }
  // pair table: values in the plan's arithmetic type, then the offsets as int32, then
  // the diagonal plan's metadata (count, offsets, kinds)
  std::vector<unsigned char> htab(256 * sizeof(DT) + (256 + 1 + kMaxTrip) * sizeof(int), 0);
  DT* tv = reinterpret_cast<DT*>(htab.data());
  int* td = reinterpret_cast<int*>(htab.data() + 256 * sizeof(DT));
  uint32_t np = 0;
  for(uint32_t cv = 0; cv < kv.size(); ++cv)
    for(uint32_t cd = 0; cd < kd.size(); ++cd)
      if(hp[cv * 256u + cd])
      {
        if(np == 256)
        {
          cleanup();
          return FEAT_ERR_UNSUPPORTED;
        }
        tv[np] = DT(__builtin_bit_cast(double, kv[cv].first));
        td[np] = int(std::int64_t(kd[cd].first ^ (uint64_t(1) << 63)));
        hmap[2 * kDictSlots + cv * 256u + cd] = uint8_t(np);
        ++np;
      }
  // diagonal plan: the distinct offsets covered by at most kMaxTrip triples (c-1, c, c+1),
  // greedily from the smallest; a free code for "no entry"; slot padding within a quarter
  const uint64_t nd = kd.size();
  std::vector<int> centres;
  std::vector<uint8_t> slot_of(256, 0);
  {
    std::vector<std::int64_t> o(nd);
    for(uint64_t t = 0; t < nd; ++t) o[t] = std::int64_t(kd[t].first ^ (uint64_t(1) << 63));
    for(uint64_t t = 0; t < nd && centres.size() <= uint64_t(kMaxTrip);)
    {
      const std::int64_t c = o[t] + 1;
      const uint32_t g = uint32_t(centres.size());
      centres.push_back(int(c));
      for(; t < nd && o[t] <= c + 1; ++t) slot_of[t] = uint8_t(3 * g + uint32_t(o[t] - (c - 1)));
    }
  }
  const uint64_t nslot = 3 * centres.size();
  const bool diag = !(flags & FEAT_SELL_NO_DIAG) && nd && centres.size() <= uint64_t(kMaxTrip) && np < 256 &&
                    ns * 64 * nslot <= A->nnz + A->nnz / 4 + 64 * nslot;
  if(diag)
  {
    int* meta = td + 256;
    meta[0] = int(centres.size());
    for(size_t g = 0; g < centres.size(); ++g) meta[1 + g] = centres[g];
  }
  // word offsets of the slices: ceil(slice max / 4) words per lane; when padding every
  // slice to the longest costs at most a quarter more, all slices get that stride
  std::vector<uint64_t> hoff(ns + 1, 0);
  uint64_t maxw = 0;
  for(uint64_t i = 0; i < ns; ++i)
  {
    const uint64_t w = (uint64_t(hmax[i]) + 3) / 4;
    hoff[i + 1] = hoff[i] + w * 64;
    maxw = w > maxw ? w : maxw;
  }
  const bool uniform = ns * 64 * maxw * 4 <= hoff[ns] * 5 && maxw < (uint64_t(1) << 16);
  if(uniform)
    for(uint64_t i = 0; i <= ns; ++i) hoff[i] = i * 64 * maxw;
  const uint32_t dw = uint32_t((nslot + 3) / 4);
  const uint64_t words = diag ? ns * 64 * dw : hoff[ns];
  std::copy(slot_of.begin(), slot_of.end(), hmap.begin() + 2 * kDictSlots + 65536);
  rc = check_hip(hipMalloc(&table, htab.size()), "dict pair table");
  if(!rc) rc = check_hip(hipMalloc(&soff, sizeof(uint64_t) * (ns + 1)), "sell soff");
  if(!rc) rc = check_hip(hipMalloc(&codes, sizeof(uint32_t) * (words ? words : 1)), "dict codes");
  if(!rc) rc = check_hip(hipMemcpyAsync(table, htab.data(), htab.size(), hipMemcpyHostToDevice, h->stream), "copy");
  if(!rc) rc = check_hip(hipMemcpyAsync(soff, hoff.data(), sizeof(uint64_t) * (ns + 1), hipMemcpyHostToDevice, h->stream), "copy");
  if(!rc) rc = check_hip(hipMemcpyAsync(maps, hmap.data(), hmap.size(), hipMemcpyHostToDevice, h->stream), "copy");
  if(rc) { cleanup(); return rc; }
  uint32_t unsorted = 0;
  if(diag)
  {
    rc = check_hip(hipMemsetAsync(state, 0, sizeof(uint32_t), h->stream), "memset");
    if(rc) { cleanup(); return rc; }
    hipLaunchKernelGGL(k_dict_fill_diag, dim3(grid_for(ns * 64)), dim3(kThreads), 0, h->stream, A->val, A->col,
                       A->rowptr, n, ns * 64, dw, vtab, dtab, maps, maps + kDictSlots, maps + 2 * kDictSlots,
                       maps + 2 * kDictSlots + 65536, codes, state);
    rc = check_launch(name);
    if(!rc) rc = check_hip(hipMemcpyAsync(&unsorted, state, sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream), "copy");
  }
  else
  {
    hipLaunchKernelGGL(k_dict_fill, dim3(grid_for(n)), dim3(kThreads), 0, h->stream, A->val, A->col, A->rowptr, n, soff,
                       vtab, dtab, maps, maps + kDictSlots, maps + 2 * kDictSlots, codes);
    rc = check_launch(name);
  }
  if(!rc) rc = check_hip(hipStreamSynchronize(h->stream), "sync");
  if(!rc && unsorted)
  {
    // rows out of column order keep the entry-order plan
    cleanup();
    return dict_create<DT>(h, A, hmax, rlen, out, name, flags | FEAT_SELL_NO_DIAG);
  }
  for(void** p : {(void**)&vtab, (void**)&dtab, (void**)&state, (void**)&present, (void**)&maps})
  {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if(rc) { cleanup(); return rc; }
  out->format = FEAT_FMT_SELL;
  out->index = diag ? FEAT_IDX_DIAG : FEAT_IDX_DICT;
  out->slice_words = diag ? dw : (uniform ? uint32_t(maxw) : 0u);
  out->rows = n;
  out->cols = A->cols;
  out->nnz = 4 * words;
  out->val = (decltype(out->val))(void*)table;
  out->value_kind = FEAT_VAL_NATIVE;
  out->scale = 1.0f;
  out->tune = 0;
  out->col = codes;
  out->ptr = soff;
  out->rowlen = rlen;
  return FEAT_OK;
}

// Banded plan of a square CSR whose rows are the clipped 3^dim-point stencils of an n1^dim
// grid (n1 >= 3; the 3D grid is tried first, then the 2D one).  Returns
// FEAT_ERR_UNSUPPORTED, without an error record, where the pattern is anything else.
template<typename DT, typename M>
int band_create(feat_handle_t h, const feat_csr_f64_t* A, float scale, float inv_scale, M* out)
{
  const uint64_t n = A->rows;
  uint64_t cand[2][2];
  int nc = 0;
  const uint64_t r3 = uint64_t(std::llround(std::cbrt(double(n)))), r2 = uint64_t(std::llround(std::sqrt(double(n))));
  for(uint64_t q = r3 > 0 ? r3 - 1 : 0; q <= r3 + 1; ++q)
    if(q >= 3 && q * q * q == n && A->nnz == (3 * q - 2) * (3 * q - 2) * (3 * q - 2))
    {
      cand[nc][0] = 3;
      cand[nc++][1] = q;
      break;
    }
  for(uint64_t q = r2 > 0 ? r2 - 1 : 0; q <= r2 + 1; ++q)
    if(q >= 3 && q * q == n && A->nnz == (3 * q - 2) * (3 * q - 2))
    {
      cand[nc][0] = 2;
      cand[nc++][1] = q;
      break;
    }
  if(!nc) return FEAT_ERR_UNSUPPORTED;
  int* bad = nullptr;
  int rc = check_hip(hipMalloc(&bad, sizeof(int)), "banded plan flag");
  int dim = 0;
  uint64_t n1 = 0;
  for(int c = 0; c < nc && !rc && !dim; ++c)
  {
    int hb = 1;
    rc = check_hip(hipMemsetAsync(bad, 0, sizeof(int), h->stream), "memset");
    if(!rc)
    {
      hipLaunchKernelGGL(k_band_check, dim3(grid_for(n)), dim3(kThreads), 0, h->stream, A->col, A->rowptr, n,
                         int(cand[c][0]), int64_t(cand[c][1]), bad);
      rc = check_launch("banded plan check");
    }
    if(!rc) rc = check_hip(hipMemcpyAsync(&hb, bad, sizeof(int), hipMemcpyDeviceToHost, h->stream), "copy");
    if(!rc) rc = check_hip(hipStreamSynchronize(h->stream), "sync");
    if(!rc && !hb)
    {
      dim = int(cand[c][0]);
      n1 = cand[c][1];
    }
  }
  if(bad) (void)hipFree(bad);
  if(rc) return rc;
  if(!dim) return FEAT_ERR_UNSUPPORTED;
  const uint64_t ns = dim == 3 ? 27 : 9;
  DT* bval = nullptr;
  rc = check_hip(hipMalloc(&bval, sizeof(DT) * ns * n), "banded plan values");
  if(!rc)
  {
    hipLaunchKernelGGL((k_band_fill<DT>), dim3(grid_for(n)), dim3(kThreads), 0, h->stream, A->val, A->rowptr, n, dim,
                       int64_t(n1), bval, inv_scale);
    rc = check_launch("banded plan fill");
  }
  if(!rc) rc = check_hip(hipStreamSynchronize(h->stream), "sync");
  if(rc)
  {
    if(bval) (void)hipFree(bval);
    return rc;
  }
  out->format = FEAT_FMT_BANDED;
  out->index = 0;
  out->rows = n;
  out->cols = n;
  out->nnz = ns * n;
  out->val = (decltype(out->val))(void*)bval;
  out->col = nullptr;
  out->ptr = nullptr;
  out->rowlen = nullptr;
  out->value_kind = std::is_same<DT, uint16_t>::value ? FEAT_VAL_F16 : FEAT_VAL_NATIVE;
  out->scale = scale;
  out->tune = 0;
  out->slice_words = uint32_t(n1);
  return FEAT_OK;
}

template<typename DT, typename M>
int sell_create_impl(feat_handle_t h, const feat_csr_f64_t* A, M* out, const char* name, int flags = 0)
{
  // DT is the stored value type: double, float, or uint16_t (FP16 bits)
  if(!h || !A || !out || !A->rowptr || (A->nnz && (!A->val || !A->col))) return fail(FEAT_ERR_INVALID, name);
  std::memset(out, 0, sizeof(M));
  const uint64_t n = A->rows, ns = (n + 63) / 64;
  uint16_t* rlen = nullptr;
  uint32_t* slmax = nullptr;
  uint32_t* stats = nullptr;
  uint64_t* soff = nullptr;
  DT* sval = nullptr;
  void* scol = nullptr;
  auto cleanup = [&]() {
    if(rlen) (void)hipFree(rlen);
    if(slmax) (void)hipFree(slmax);
    if(stats) (void)hipFree(stats);
    if(soff) (void)hipFree(soff);
    if(sval) (void)hipFree(sval);
    if(scol) (void)hipFree(scol);
  };
  int rc = check_hip(hipMalloc(&rlen, sizeof(uint16_t) * (n ? n : 1)), "sell rlen");
  if(!rc) rc = check_hip(hipMalloc(&slmax, sizeof(uint32_t) * (ns ? ns : 1)), "sell slmax");
  if(!rc) rc = check_hip(hipMalloc(&stats, sizeof(uint32_t) * 2), "sell stats");
  if(!rc) rc = check_hip(hipMalloc(&soff, sizeof(uint64_t) * (ns + 1)), "sell soff");
  if(!rc) rc = check_hip(hipMemsetAsync(slmax, 0, sizeof(uint32_t) * (ns ? ns : 1), h->stream), "memset");
  if(!rc) rc = check_hip(hipMemsetAsync(stats, 0, sizeof(uint32_t) * 2, h->stream), "memset");
  if(rc) { cleanup(); return rc; }
  if(n)
    hipLaunchKernelGGL(k_sell_scan_rows, dim3(grid_for(n)), dim3(kThreads), 0, h->stream, A->rowptr, A->col, n, rlen, slmax, stats);
  std::vector<uint32_t> hmax(ns);
  uint32_t hst[2] = {0, 0};
  rc = check_launch(name);
  if(!rc && ns) rc = check_hip(hipMemcpyAsync(hmax.data(), slmax, sizeof(uint32_t) * ns, hipMemcpyDeviceToHost, h->stream), "copy");
  if(!rc) rc = check_hip(hipMemcpyAsync(hst, stats, sizeof(hst), hipMemcpyDeviceToHost, h->stream), "copy");
  if(!rc) rc = check_hip(hipStreamSynchronize(h->stream), "sync");
  if(rc) { cleanup(); return rc; }
  std::vector<uint64_t> hoff(ns + 1, 0);
  for(uint64_t i = 0; i < ns; ++i) hoff[i + 1] = hoff[i] + uint64_t(hmax[i]) * 64;
  const uint64_t total = hoff[ns];
  // reject skewed row lengths (padding beyond twice the nonzeros) on matrices
  // where it matters; small matrices (partial slices) are always accepted
  if(hst[1] || (total > 2 * A->nnz + 64 * ns && total > (uint64_t(1) << 16)))
  {
    cleanup();
    return fail(FEAT_ERR_UNSUPPORTED, "SELL-64 plan: row too long or padding too large; keep CSR");
  }
  if(!(flags & FEAT_SELL_NO_DICT) && !std::is_same<DT, uint16_t>::value && n && A->nnz && hst[0] <= 0x7FFFFFFFu)
  {
    (void)hipFree(slmax);
    (void)hipFree(stats);
    (void)hipFree(soff);
    slmax = nullptr;
    stats = nullptr;
    soff = nullptr;
    rc = dict_create<DT>(h, A, hmax, rlen, out, name, flags);
    if(rc == FEAT_OK) return FEAT_OK;
    if(rc != FEAT_ERR_UNSUPPORTED)
    {
      cleanup();
      return rc;
    }
    rc = check_hip(hipMalloc(&soff, sizeof(uint64_t) * (ns + 1)), "sell soff");
    if(rc) { cleanup(); return rc; }
  }
  const bool d16 = hst[0] <= 32767u;
  if(!d16 && A->cols > 0xFFFFFFFFull)
  {
    cleanup();
    return fail(FEAT_ERR_UNSUPPORTED, "SELL-64 plan: columns exceed 32 bits");
  }
  const size_t cb = d16 ? sizeof(int16_t) : sizeof(uint32_t);
  // FP16: scale 2^e with 2^(e-1) <= max|a| < 2^e, so stored values lie in (-1, 1)
  float scale = 1.0f, inv_scale = 1.0f;
  if(std::is_same<DT, uint16_t>::value && A->nnz)
  {
    uint32_t* am = nullptr;
    uint32_t hm = 0;
    rc = check_hip(hipMalloc(&am, sizeof(uint32_t)), "absmax");
    if(!rc) rc = check_hip(hipMemsetAsync(am, 0, sizeof(uint32_t), h->stream), "memset");
    if(!rc) hipLaunchKernelGGL(k_absmax, dim3(grid_for(A->nnz, 8)), dim3(kThreads), 0, h->stream, A->val, A->nnz, am);
    if(!rc) rc = check_hip(hipMemcpyAsync(&hm, am, sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream), "copy");
    if(!rc) rc = check_hip(hipStreamSynchronize(h->stream), "sync");
    if(am) (void)hipFree(am);
    if(rc) { cleanup(); return rc; }
    const float mx = __builtin_bit_cast(float, hm);
    if(mx > 0.0f && std::isfinite(mx))
    {
      int e = 0;
      (void)std::frexp(mx, &e);
      scale = std::ldexp(1.0f, e);
      inv_scale = std::ldexp(1.0f, -e);
    }
  }
  // value codes: at most 256 distinct values (bit patterns in the arithmetic type).  The
  // values are scanned in chunks, and the scan stops at the 257th distinct value.
  std::vector<uint64_t> vkeys;
  if(!(flags & FEAT_SELL_NO_DICT) && !std::is_same<DT, uint16_t>::value && n && A->nnz)
  {
    const uint64_t chunk = std::min<uint64_t>(A->nnz, uint64_t(1) << 20);
    std::vector<double> hv(chunk);
    uint64_t last = ~uint64_t(0);
    bool stop = false;
    for(uint64_t c0 = 0; c0 < A->nnz && !stop; c0 += chunk)
    {
      const uint64_t cn = std::min<uint64_t>(chunk, A->nnz - c0);
      rc = check_hip(hipMemcpyAsync(hv.data(), A->val + c0, sizeof(double) * cn, hipMemcpyDeviceToHost, h->stream), "copy");
      if(!rc) rc = check_hip(hipStreamSynchronize(h->stream), "sync");
      if(rc) { cleanup(); return rc; }
      for(uint64_t i = 0; i < cn; ++i)
      {
        const double d = hv[i];
        if(std::isnan(d))   // NaN payloads may convert differently on host and device
        {
          stop = true;
          break;
        }
        const uint64_t key = value_key<DT>(DT(d));
        if(key == last) continue;
        last = key;
        auto it = std::lower_bound(vkeys.begin(), vkeys.end(), key);
        if(it != vkeys.end() && *it == key) continue;
        if(vkeys.size() == 256)
        {
          stop = true;
          break;
        }
        vkeys.insert(it, key);
      }
    }
    if(stop) vkeys.clear();
  }
  // no pair dictionary and no value codes: the banded plan where the rows are grid stencils
  if(vkeys.empty() && !(flags & (FEAT_SELL_NO_DICT | FEAT_SELL_NO_BANDED)) && n >= 27 && n == A->cols &&
     27 * n * sizeof(double) < (uint64_t(1) << 32) && A->nnz)
  {
    rc = band_create<DT>(h, A, scale, inv_scale, out);
    if(rc == FEAT_OK)
    {
      cleanup();
      return FEAT_OK;
    }
    if(rc != FEAT_ERR_UNSUPPORTED)
    {
      cleanup();
      return rc;
    }
    rc = FEAT_OK;
  }
  bool vcode = !vkeys.empty();
  uint64_t* dkeys = nullptr;
  const size_t vbytes = vcode ? sizeof(DT) * 256 + (total ? total : 1) : sizeof(DT) * (total ? total : 1);
  rc = check_hip(hipMalloc(&sval, vbytes), "sell val");
  if(!rc) rc = check_hip(hipMalloc(&scol, cb * (total ? total : 1)), "sell col");
  if(!rc) rc = check_hip(hipMemcpyAsync(soff, hoff.data(), sizeof(uint64_t) * (ns + 1), hipMemcpyHostToDevice, h->stream), "copy");
  if(!rc && vcode)
  {
    std::vector<DT> tab(256, DT(0));
    for(size_t i = 0; i < vkeys.size(); ++i) tab[i] = key_value<DT>(vkeys[i]);
    rc = check_hip(hipMalloc(&dkeys, sizeof(uint64_t) * vkeys.size()), "value keys");
    if(!rc) rc = check_hip(hipMemcpyAsync(dkeys, vkeys.data(), sizeof(uint64_t) * vkeys.size(), hipMemcpyHostToDevice, h->stream), "copy");
    if(!rc) rc = check_hip(hipMemcpyAsync(sval, tab.data(), sizeof(DT) * 256, hipMemcpyHostToDevice, h->stream), "copy");
    int* dmiss = nullptr;
    int hmiss = 0;
    if constexpr(!std::is_same<DT, uint16_t>::value)   // FP16 plans keep their values
    {
      if(!rc) rc = check_hip(hipMalloc(&dmiss, sizeof(int)), "value-code miss flag");
      if(!rc) rc = check_hip(hipMemsetAsync(dmiss, 0, sizeof(int), h->stream), "memset");
      if(!rc)
      {
        uint8_t* codes = reinterpret_cast<uint8_t*>(sval + 256);
        if(d16)
          hipLaunchKernelGGL((k_sell_fill_code<DT, int16_t, true>), dim3(grid_for(n)), dim3(kThreads), 0, h->stream,
                             A->val, A->col, A->rowptr, n, soff, dkeys, int(vkeys.size()), codes, (int16_t*)scol, dmiss);
        else
          hipLaunchKernelGGL((k_sell_fill_code<DT, uint32_t, false>), dim3(grid_for(n)), dim3(kThreads), 0, h->stream,
                             A->val, A->col, A->rowptr, n, soff, dkeys, int(vkeys.size()), codes, (uint32_t*)scol, dmiss);
        rc = check_launch(name);
        if(!rc) rc = check_hip(hipMemcpyAsync(&hmiss, dmiss, sizeof(int), hipMemcpyDeviceToHost, h->stream), "copy");
        if(!rc) rc = check_hip(hipStreamSynchronize(h->stream), "sync");
      }
    }
    if(dmiss) (void)hipFree(dmiss);
    if(dkeys) (void)hipFree(dkeys);
    if(rc) { cleanup(); return rc; }
    if(hmiss)   // a value missed the host's table: full values instead of codes
    {
      vcode = false;
      (void)hipFree(sval);
      sval = nullptr;
      rc = check_hip(hipMalloc(&sval, sizeof(DT) * (total ? total : 1)), "sell val");
      if(rc) { cleanup(); return rc; }
    }
  }
  if(rc) { cleanup(); return rc; }
  if(n && !vcode)
  {
    if(d16)
      hipLaunchKernelGGL((k_sell_fill<DT, int16_t, true>), dim3(grid_for(n)), dim3(kThreads), 0, h->stream, A->val, A->col,
                         A->rowptr, n, soff, sval, (int16_t*)scol, inv_scale);
    else
      hipLaunchKernelGGL((k_sell_fill<DT, uint32_t, false>), dim3(grid_for(n)), dim3(kThreads), 0, h->stream, A->val, A->col,
                         A->rowptr, n, soff, sval, (uint32_t*)scol, inv_scale);
  }
  rc = check_launch(name);
  if(!rc) rc = check_hip(hipStreamSynchronize(h->stream), "sync");
  (void)hipFree(slmax);
  (void)hipFree(stats);
  slmax = nullptr;
  stats = nullptr;
  if(rc) { cleanup(); return rc; }
  out->format = FEAT_FMT_SELL;
  out->index = d16 ? FEAT_IDX_D16 : FEAT_IDX_U32;
  out->rows = n;
  out->cols = A->cols;
  out->nnz = total;
  out->val = (decltype(out->val))(void*)sval;
  out->value_kind = std::is_same<DT, uint16_t>::value ? FEAT_VAL_F16 : (vcode ? FEAT_VAL_CODE8 : FEAT_VAL_NATIVE);
  out->scale = scale;
  out->tune = 0;
  out->col = scol;
  out->ptr = soff;
  out->rowlen = rlen;
  return FEAT_OK;
}

template<typename M>
int sell_destroy_impl(M* m)
{
  if(!m) return FEAT_OK;
  if(m->format != FEAT_FMT_SELL && m->format != FEAT_FMT_BANDED)
    return fail(FEAT_ERR_INVALID, "feat_sell_destroy: not a SELL-64 or banded plan");
  if(m->val) (void)hipFree(const_cast<void*>(static_cast<const void*>(m->val)));
  if(m->col) (void)hipFree(const_cast<void*>(m->col));
  if(m->ptr) (void)hipFree(const_cast<void*>(m->ptr));
  if(m->rowlen) (void)hipFree(const_cast<uint16_t*>(m->rowlen));
  std::memset(m, 0, sizeof(M));
  return FEAT_OK;
}

template<typename M>
bool mat_ok(const M* A)
{
  if(A && A->format == FEAT_FMT_DLINEAR)
  {
    const uint64_t n1 = A->slice_words, nc = (n1 + 1) / 2;
    const uint64_t nf = A->index == 3 ? n1 * n1 * n1 : n1 * n1, ncn = A->index == 3 ? nc * nc * nc : nc * nc;
    return (A->index == 2 || A->index == 3) && n1 >= 3 && (n1 & 1) && (A->value_kind == 0 || A->value_kind == 1) &&
           A->rows == (A->value_kind ? ncn : nf) && A->cols == (A->value_kind ? nf : ncn) && nf <= (uint64_t(1) << 62);
  }
  if(A && A->format == FEAT_FMT_BANDED)
  {
    const uint64_t q = A->slice_words;
    return A->val && q >= 3 && A->rows == A->cols && 27 * A->rows * sizeof(double) < (uint64_t(1) << 32) &&
           (q * q * q == A->rows || q * q == A->rows) &&
           (A->value_kind == FEAT_VAL_NATIVE || A->value_kind == FEAT_VAL_F16);
  }
  return A && A->ptr && (A->nnz == 0 || (A->val && A->col)) && (A->format != FEAT_FMT_SELL || A->rowlen);
}

template<typename DT, typename M>
int defect_impl(feat_handle_t h, const M* A, DT* d, const DT* rhs, const DT* x, const uint32_t* mask,
                double* nrm2, const feat_reduce_ws_t* ws, const char* name, FirstIter<DT> first = {nullptr, nullptr, DT(0)})
{
  if(!h || !mat_ok(A) || (A->rows && (!d || !rhs || !x))) return fail(FEAT_ERR_INVALID, name);
  if(d == x) return fail(FEAT_ERR_ALIAS, name);
  if(first.x1 && ((A->rows && !first.inv) || first.x1 == d || first.x1 == x || first.x1 == rhs))
    return fail(FEAT_ERR_INVALID, name);
  if(need_ws(ws, nrm2)) return fail(FEAT_ERR_INVALID, "reduction workspace missing");
  EpiDefect<DT> e{d, rhs, mask, nrm2 != nullptr, first};
  return launch_mat<DT>(h->stream, A, x, e, ws_reduce(ws, nrm2, nullptr), name);
}

template<typename DT, typename M>
int masked_impl(feat_handle_t h, const M* A, DT* r, const DT* x, const uint32_t* mask, double* nrm2,
                const DT* dotw, double* dot, const feat_reduce_ws_t* ws, const char* name,
                FirstIter<DT> first = {nullptr, nullptr, DT(0)})
{
  if(!h || !mat_ok(A) || (A->rows && !r) || (A->cols && !x) || (dot && !dotw)) return fail(FEAT_ERR_INVALID, name);
  if(r == x) return fail(FEAT_ERR_ALIAS, name);
  if(first.x1 && ((A->rows && !first.inv) || first.x1 == r || first.x1 == x)) return fail(FEAT_ERR_INVALID, name);
  if(need_ws(ws, nrm2, dot)) return fail(FEAT_ERR_INVALID, "reduction workspace missing");
  EpiMasked<DT> e{r, mask, dotw, nrm2 != nullptr, dot != nullptr, first};
  return launch_mat<DT>(h->stream, A, x, e, ws_reduce(ws, nrm2, dot), name);
}

template<typename DT, typename M>
int prol_impl(feat_handle_t h, const M* P, DT* sol, const DT* xc, const uint32_t* mask, const char* name)
{
  if(!h || !mat_ok(P) || (P->rows && !sol) || (P->cols && !xc)) return fail(FEAT_ERR_INVALID, name);
  if(sol == xc) return fail(FEAT_ERR_ALIAS, name);
  EpiProlAdd<DT> e{sol, mask};
  GridReduce g{nullptr, nullptr, nullptr, nullptr};
  return launch_mat<DT>(h->stream, P, xc, e, g, name);
}

// iterate k (1-based) of the smoother lives in buf[(steps - k) & 1], so the last
// stored iterate lands in x
// x1_ready: the first iterate (without the skip rule) is already in B(1), written by the
// kernel that produced rhs (FirstIter); valid for steps >= 2, where step 1 applies the rule
template<typename DT, typename M>
int smooth_impl(feat_handle_t h, const M* A, const DT* inv, const uint32_t* mask, int steps, DT omega,
                const DT* rhs, const double* rhs_nrm2, DT* x, DT* tmp, DT* d_out, DT* acc, const DT* dotw,
                double* dot, const feat_reduce_ws_t* ws, const char* name, bool x1_ready = false)
{
  if(x1_ready && steps < 2) return fail(FEAT_ERR_INVALID, "rj_smooth: a first iterate from the producer needs steps >= 2");
  if(!h || !mat_ok(A) || A->rows != A->cols || steps < 1) return fail(FEAT_ERR_INVALID, name);
  const uint64_t n = A->rows;
  if(n && (!inv || !rhs || !tmp || !x)) return fail(FEAT_ERR_INVALID, name);
  if(x == tmp || (acc && (acc == x || acc == tmp)) || (d_out && (d_out == x || d_out == tmp)))
    return fail(FEAT_ERR_ALIAS, name);
  if(need_ws(ws, dot)) return fail(FEAT_ERR_INVALID, "reduction workspace missing");
  if(d_out && acc) return fail(FEAT_ERR_INVALID, "rj_smooth: d_out and acc are exclusive");
  if(!n) return FEAT_OK;
  DT* buf[2] = {x, tmp};
  auto B = [&](int k) { return buf[(steps - k) & 1]; };
  const int nsp = d_out ? steps : steps - 1;
  GridReduce none{nullptr, nullptr, nullptr, nullptr};
  const bool acc_first = (acc != nullptr && steps == 1);
  if(!x1_ready)
  {
    const int grid = grid_for(n, 4);
    GridReduce g = reduce_for_grid(acc_first ? ws_reduce(ws, nullptr, dot) : none, uint64_t(grid));
    hipLaunchKernelGGL((k_smooth_first<DT>), dim3(grid), dim3(kThreads), 0, h->stream, B(1), inv, rhs, mask,
                       omega, rhs_nrm2, acc_first ? acc : (DT*)nullptr, acc_first ? dotw : (const DT*)nullptr, n, g);
    finish_reduce(h->stream, g, uint64_t(grid));
    int rc = check_launch(name);
    if(rc) return rc;
  }
  for(int k = 1; k <= nsp; ++k)
  {
    EpiSmooth<DT> e;
    e.rhs = rhs;
    e.inv = inv;
    e.mask = mask;
    e.xk = B(k);
    e.dotw = nullptr;
    e.rhs_nrm2 = rhs_nrm2;
    e.omega = omega;
    e.skip = false;
    GridReduce g = none;
    if(d_out && k == steps)
    {
      e.mode = 1;
      e.out = d_out;
    }
    else if(acc && k == steps - 1)
    {
      e.mode = 2;
      e.out = acc;
      e.dotw = dotw;
      g = ws_reduce(ws, nullptr, dot);
    }
    else
    {
      e.mode = 0;
      e.out = B(k + 1);
    }
    int rc = launch_mat<DT>(h->stream, A, e.xk, e, g, name);
    if(rc) return rc;
  }
  return FEAT_OK;
}

__global__ __launch_bounds__(kThreads) void k_convert_dot(double* z, const float* z32, const double* w, uint64_t n, GridReduce gr)
{
  double acc = 0.0;
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
  {
    const double v = double(z32[i]);
    z[i] = v;
    if(w) acc = acc + v * w[i];
  }
  grid_reduce_finish(gr, acc, 0.0);
}

__global__ void k_dense_gemv_mixed(int n, const double* ainv, const float* b, float* x, const uint32_t* mask)
{
  for(int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
  {
    double s = 0.0;
    for(int k = 0; k < n; ++k) s = __fma_rn(ainv[size_t(r) * n + k], double(b[k]), s);
    if(is_masked(mask, uint64_t(r))) s = 0.0;
    x[r] = float(s);
  }
}
} // namespace

extern "C" {

int feat_sell_create_f64(feat_handle_t h, const feat_csr_f64_t* A, feat_matrix_f64_t* out)
{
  return sell_create_impl<double>(h, A, out, "feat_sell_create_f64");
}

int feat_sell_create_f32(feat_handle_t h, const feat_csr_f64_t* A, feat_matrix_f32_t* out)
{
  return sell_create_impl<float>(h, A, out, "feat_sell_create_f32");
}

int feat_sell_create_ex_f64(feat_handle_t h, const feat_csr_f64_t* A, int flags, feat_matrix_f64_t* out)
{
  return sell_create_impl<double>(h, A, out, "feat_sell_create_ex_f64", flags);
}

int feat_sell_create_ex_f32(feat_handle_t h, const feat_csr_f64_t* A, int flags, feat_matrix_f32_t* out)
{
  return sell_create_impl<float>(h, A, out, "feat_sell_create_ex_f32", flags);
}

int feat_sell_create_f16(feat_handle_t h, const feat_csr_f64_t* A, feat_matrix_f32_t* out)
{
  return sell_create_impl<uint16_t>(h, A, out, "feat_sell_create_f16");
}

int feat_sell_create_ex_f16(feat_handle_t h, const feat_csr_f64_t* A, int flags, feat_matrix_f32_t* out)
{
  return sell_create_impl<uint16_t>(h, A, out, "feat_sell_create_ex_f16", flags);
}

int feat_sell_destroy_f64(feat_matrix_f64_t* m) { return sell_destroy_impl(m); }
int feat_sell_destroy_f32(feat_matrix_f32_t* m) { return sell_destroy_impl(m); }

int feat_matrix_apply_f64(feat_handle_t h, const feat_matrix_f64_t* A, double* r, double a, const double* x, double b,
                          const double* y)
{
  if(!h || !mat_ok(A)) return fail(FEAT_ERR_INVALID, "feat_matrix_apply_f64: bad argument");
  const bool bzero = std::fabs(b) < DBL_EPSILON;
  if((A->rows && !r) || (A->cols && !x) || (!bzero && A->rows && !y)) return fail(FEAT_ERR_INVALID, "feat_matrix_apply_f64: null vector");
  if(r == x && A->rows) return fail(FEAT_ERR_ALIAS, "feat_matrix_apply_f64: r and x must not alias");
  EpiApply<double> e{r, y, a, b, bzero};
  GridReduce g{nullptr, nullptr, nullptr, nullptr};
  return launch_mat<double>(h->stream, A, x, e, g, "feat_matrix_apply_f64");
}

int feat_autotune_matrix_f64(feat_handle_t h, feat_matrix_f64_t* A, const double* x, double* r, int reps,
                             double* best_us)
{
  if(!h || !mat_ok(A) || reps < 1 || (A->rows && (!x || !r))) return fail(FEAT_ERR_INVALID, "feat_autotune_matrix_f64: bad argument");
  if(r == x && A->rows) return fail(FEAT_ERR_ALIAS, "feat_autotune_matrix_f64: r and x must not alias");
  A->tune = 0;
  if(best_us) *best_us = -1.0;
  if(A->format != FEAT_FMT_SELL || A->rows == 0) return FEAT_OK;  // only SELL launch shapes are tunable
  // The diagonal kernel keeps its full grid: it has no unroll choice, a capped grid times
  // the same for r = A*x (within noise), and the smoothing epilogue, whose row operands
  // lengthen each chunk's chain, loses up to half its speed on a capped grid.
  if(A->index == FEAT_IDX_DIAG) return FEAT_OK;
  const uint64_t nch = (A->rows + kThreads - 1) / kThreads;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = check_hip(hipEventCreate(&e0), "hipEventCreate");
  if(!rc) rc = check_hip(hipEventCreate(&e1), "hipEventCreate");
  int best = 0;
  float best_ms = 0.0f;
  for(int unr = 1; unr <= 3 && !rc; ++unr)
    for(int cap = 0; cap <= 5 && !rc; ++cap)
    {
      if(cap && (uint64_t(kMaxGrid) >> cap) >= nch) continue;  // same launch as the uncapped grid
      A->tune = unr | (cap << 2);
      for(int w = 0; w < 2 && !rc; ++w) rc = feat_matrix_apply_f64(h, A, r, 1.0, x, 0.0, nullptr);
      if(!rc) rc = check_hip(hipEventRecord(e0, h->stream), "hipEventRecord");
      for(int k = 0; k < reps && !rc; ++k) rc = feat_matrix_apply_f64(h, A, r, 1.0, x, 0.0, nullptr);
      if(!rc) rc = check_hip(hipEventRecord(e1, h->stream), "hipEventRecord");
      if(!rc) rc = check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
      float ms = 0.0f;
      if(!rc) rc = check_hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
      if(!rc && (best == 0 || ms < best_ms))
      {
        best = A->tune;
        best_ms = ms;
      }
    }
  if(e0) (void)hipEventDestroy(e0);
  if(e1) (void)hipEventDestroy(e1);
  A->tune = rc ? 0 : best;
  if(!rc && best_us) *best_us = 1e3 * double(best_ms) / reps;
  return rc;
}

int feat_defect_f64(feat_handle_t h, const feat_matrix_f64_t* A, double* d, const double* rhs, const double* x,
                    const uint32_t* mask, double* nrm2, const feat_reduce_ws_t* ws)
{
  return defect_impl<double>(h, A, d, rhs, x, mask, nrm2, ws, "feat_defect_f64");
}

int feat_apply_masked_f64(feat_handle_t h, const feat_matrix_f64_t* A, double* r, const double* x,
                          const uint32_t* mask, double* nrm2, const double* dotw, double* dot,
                          const feat_reduce_ws_t* ws)
{
  return masked_impl<double>(h, A, r, x, mask, nrm2, dotw, dot, ws, "feat_apply_masked_f64");
}

int feat_defect_first_f64(feat_handle_t h, const feat_matrix_f64_t* A, double* d, const double* rhs, const double* x,
                          const uint32_t* mask, double* nrm2, const double* inv_diag, double omega, double* x1,
                          const feat_reduce_ws_t* ws)
{
  if(!x1) return fail(FEAT_ERR_INVALID, "feat_defect_first_f64: null x1");
  return defect_impl<double>(h, A, d, rhs, x, mask, nrm2, ws, "feat_defect_first_f64",
                             FirstIter<double>{inv_diag, x1, omega});
}

int feat_apply_masked_first_f64(feat_handle_t h, const feat_matrix_f64_t* A, double* r, const double* x,
                                const uint32_t* mask, double* nrm2, const double* inv_diag, double omega, double* x1,
                                const feat_reduce_ws_t* ws)
{
  if(!x1) return fail(FEAT_ERR_INVALID, "feat_apply_masked_first_f64: null x1");
  return masked_impl<double>(h, A, r, x, mask, nrm2, (const double*)nullptr, nullptr, ws, "feat_apply_masked_first_f64",
                             FirstIter<double>{inv_diag, x1, omega});
}

int feat_prol_add_f64(feat_handle_t h, const feat_matrix_f64_t* P, double* sol, const double* xc, const uint32_t* mask)
{
  return prol_impl<double>(h, P, sol, xc, mask, "feat_prol_add_f64");
}

int feat_rj_smooth_f64(feat_handle_t h, const feat_matrix_f64_t* A, const double* inv_diag, const uint32_t* mask,
                       int steps, double omega, const double* rhs, const double* rhs_nrm2, double* x, double* tmp,
                       double* d_out, double* acc, const double* dotw, double* dot, const feat_reduce_ws_t* ws)
{
  return smooth_impl<double>(h, A, inv_diag, mask, steps, omega, rhs, rhs_nrm2, x, tmp, d_out, acc, dotw, dot, ws,
                             "feat_rj_smooth_f64");
}

int feat_rj_smooth_from1_f64(feat_handle_t h, const feat_matrix_f64_t* A, const double* inv_diag,
                             const uint32_t* mask, int steps, double omega, const double* rhs, const double* rhs_nrm2,
                             double* x, double* tmp, double* d_out, double* acc, const double* dotw, double* dot,
                             const feat_reduce_ws_t* ws)
{
  return smooth_impl<double>(h, A, inv_diag, mask, steps, omega, rhs, rhs_nrm2, x, tmp, d_out, acc, dotw, dot, ws,
                             "feat_rj_smooth_from1_f64", true);
}

int feat_dense_gemv_f64(feat_handle_t h, int n, const double* ainv, const double* b, double* x, const uint32_t* mask)
{
  if(!h || n < 0 || (n && (!ainv || !b || !x))) return fail(FEAT_ERR_INVALID, "feat_dense_gemv_f64: bad argument");
  if(b == x && n) return fail(FEAT_ERR_ALIAS, "feat_dense_gemv_f64: b and x must not alias");
  if(!n) return FEAT_OK;
  hipLaunchKernelGGL(k_dense_gemv<double>, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, h->stream, n, ainv, b, x, mask);
  return check_launch("feat_dense_gemv_f64");
}

int feat_pcg_update_f64(feat_handle_t h, double* x, double* r, const double* p, const double* q,
                        const double* gamma, const double* qp, double* rnrm2, uint64_t n,
                        const feat_reduce_ws_t* ws)
{
  if(!h || !gamma || !qp || !rnrm2 || (n && (!x || !r || !p || !q))) return fail(FEAT_ERR_INVALID, "feat_pcg_update_f64: null argument");
  if(need_ws(ws, rnrm2)) return fail(FEAT_ERR_INVALID, "reduction workspace missing");
  const int grid = grid_for(n, 4);
  const GridReduce gd = reduce_for_grid(ws_reduce(ws, rnrm2, nullptr), uint64_t(grid));
  hipLaunchKernelGGL(k_pcg_update, dim3(grid), dim3(kThreads), 0, h->stream, x, r, p, q, gamma, qp, n, gd,
                     FirstIter<double>{nullptr, nullptr, 0.0}, (const uint32_t*)nullptr);
  finish_reduce(h->stream, gd, uint64_t(grid));
  return check_launch("feat_pcg_update_f64");
}

int feat_pcg_update_first_f64(feat_handle_t h, double* x, double* r, const double* p, const double* q,
                              const double* gamma, const double* qp, double* rnrm2, uint64_t n,
                              const double* inv_diag, const uint32_t* mask, double omega, double* x1,
                              const feat_reduce_ws_t* ws)
{
  if(!h || !gamma || !qp || !rnrm2 || (n && (!x || !r || !p || !q || !inv_diag || !x1)))
    return fail(FEAT_ERR_INVALID, "feat_pcg_update_first_f64: null argument");
  // x1 may take q's place (each element of q is read before the element of x1 is written)
  if(x1 == x || x1 == r || x1 == p) return fail(FEAT_ERR_ALIAS, "feat_pcg_update_first_f64: x1 aliases x, r or p");
  if(need_ws(ws, rnrm2)) return fail(FEAT_ERR_INVALID, "reduction workspace missing");
  const int grid = grid_for(n, 4);
  const GridReduce gd = reduce_for_grid(ws_reduce(ws, rnrm2, nullptr), uint64_t(grid));
  hipLaunchKernelGGL(k_pcg_update, dim3(grid), dim3(kThreads), 0, h->stream, x, r, p, q, gamma, qp, n, gd,
                     FirstIter<double>{inv_diag, x1, omega}, mask);
  finish_reduce(h->stream, gd, uint64_t(grid));
  return check_launch("feat_pcg_update_first_f64");
}

int feat_pcg_direction_f64(feat_handle_t h, double* p, const double* z, const double* gamma_new, double* gamma_old, uint64_t n)
{
  if(!h || !gamma_new || !gamma_old || (n && (!p || !z))) return fail(FEAT_ERR_INVALID, "feat_pcg_direction_f64: null argument");
  if(n)
    hipLaunchKernelGGL(k_pcg_direction, dim3(grid_for(n, 4)), dim3(kThreads), 0, h->stream, p, z, gamma_new, gamma_old, n);
  hipLaunchKernelGGL(k_copy_scalar, dim3(1), dim3(1), 0, h->stream, gamma_old, gamma_new);
  return check_launch("feat_pcg_direction_f64");
}

int feat_read_scalar_f64(feat_handle_t h, const double* dev, double* host)
{
  if(!h || !dev || !host) return fail(FEAT_ERR_INVALID, "feat_read_scalar_f64: null argument");
  return to_host(h, dev, host);
}

// ---------------- FP32 inner solver kernels ----------------
int feat_apply_masked_f32(feat_handle_t h, const feat_matrix_f32_t* A, float* r, const float* x,
                          const uint32_t* mask, double* nrm2, const feat_reduce_ws_t* ws)
{
  return masked_impl<float>(h, A, r, x, mask, nrm2, nullptr, nullptr, ws, "feat_apply_masked_f32");
}

int feat_defect_f32(feat_handle_t h, const feat_matrix_f32_t* A, float* d, const float* rhs, const float* x,
                    const uint32_t* mask, double* nrm2, const feat_reduce_ws_t* ws)
{
  return defect_impl<float>(h, A, d, rhs, x, mask, nrm2, ws, "feat_defect_f32");
}

int feat_prol_add_f32(feat_handle_t h, const feat_matrix_f32_t* P, float* sol, const float* xc, const uint32_t* mask)
{
  return prol_impl<float>(h, P, sol, xc, mask, "feat_prol_add_f32");
}

int feat_rj_smooth_f32(feat_handle_t h, const feat_matrix_f32_t* A, const float* inv_diag, const uint32_t* mask,
                       int steps, float omega, const float* rhs, const double* rhs_nrm2, float* x, float* tmp,
                       float* d_out, float* acc, const feat_reduce_ws_t* ws)
{
  return smooth_impl<float>(h, A, inv_diag, mask, steps, omega, rhs, rhs_nrm2, x, tmp, d_out, acc, nullptr, nullptr,
                            ws, "feat_rj_smooth_f32");
}

int feat_dense_gemv_f32(feat_handle_t h, int n, const float* ainv, const float* b, float* x, const uint32_t* mask)
{
  if(!h || n < 0 || (n && (!ainv || !b || !x))) return fail(FEAT_ERR_INVALID, "feat_dense_gemv_f32: bad argument");
  if(b == x && n) return fail(FEAT_ERR_ALIAS, "feat_dense_gemv_f32: b and x must not alias");
  if(!n) return FEAT_OK;
  hipLaunchKernelGGL(k_dense_gemv<float>, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, h->stream, n, ainv, b, x, mask);
  return check_launch("feat_dense_gemv_f32");
}

int feat_convert_dot_f32_f64(feat_handle_t h, double* z, const float* z32, const double* w, uint64_t n, double* dot,
                             const feat_reduce_ws_t* ws)
{
  if(!h || (n && (!z || !z32)) || (dot && !w)) return fail(FEAT_ERR_INVALID, "feat_convert_dot_f32_f64: null argument");
  if(need_ws(ws, dot)) return fail(FEAT_ERR_INVALID, "reduction workspace missing");
  const int grid = grid_for(n, 4);
  const GridReduce gd = reduce_for_grid(ws_reduce(ws, dot, nullptr), uint64_t(grid));
  hipLaunchKernelGGL(k_convert_dot, dim3(grid), dim3(kThreads), 0, h->stream, z, z32, dot ? w : nullptr, n, gd);
  finish_reduce(h->stream, gd, uint64_t(grid));
  return check_launch("feat_convert_dot_f32_f64");
}

int feat_dense_gemv_mixed(feat_handle_t h, int n, const double* ainv, const float* b, float* x, const uint32_t* mask)
{
  if(!h || n < 0 || (n && (!ainv || !b || !x))) return fail(FEAT_ERR_INVALID, "feat_dense_gemv_mixed: bad argument");
  if(!n) return FEAT_OK;
  hipLaunchKernelGGL(k_dense_gemv_mixed, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, h->stream, n, ainv, b, x, mask);
  return check_launch("feat_dense_gemv_mixed");
}

} // extern "C"

// ---------------------------------------------------------------------------
// fused coarse levels: one 1024-thread workgroup, vectors in LDS
// ---------------------------------------------------------------------------
namespace {

constexpr int kTailThreads = 1024;

template<typename DT>
struct TailBufs
{
  DT *rhs, *sol, *tmp, *def, *cor;
};

// One thread per row on every level of the tail (at most kTailThreads rows), so each
// thread keeps its row's code words, inverted diagonal entry and filter bit in registers
// for the whole V-cycle: the smoothing and defect sweeps then touch only LDS.
constexpr int kTailSmoothed = FEAT_MG_TAIL_MAX_LEVELS - 1;

template<typename DT>
struct TailRow
{
  uint32_t wd[8];
  DT inv;
  bool masked;
};

// sum of the thread's DIAG-plan row (slot order = CSR order) with x in LDS
template<typename DT>
__device__ __forceinline__ DT tail_row_a(const TailRow<DT> R, const int* meta, const DT* tab, uint64_t i, const DT* x)
{
  const int nt = meta[0];
  DT sum = DT(0);
#pragma unroll
  for(int g = 0; g < kMaxTrip; ++g)
    if(g < nt)
    {
      const std::int64_t c = std::int64_t(i) + meta[1 + g];
#pragma unroll
      for(int j = 0; j < 3; ++j)
      {
        const int t = 3 * g + j;
        const uint32_t code = (R.wd[t >> 2] >> (8 * (t & 3))) & 255u;
        if(code != kDiagNone) sum = dfma(tab[code], x[c + j - 1], sum);
      }
    }
  return sum;
}


// sum of a free row of an n1^dim grid level whose rows all hold the stencil sten (3^dim values
// at the offsets (dz, dy, dx) in lexicographic order = CSR order), x in LDS
template<typename DT>
__device__ __forceinline__ DT tail_row_grid(const DT* sten, TailGeom g, int i, const DT* x)
{
  DT s = DT(0);
  const int zl = g.dim == 3 ? -1 : 0, zh = g.dim == 3 ? 1 : 0, nn = g.n1 * g.n1;
  int u = 0;
  for(int dz = zl; dz <= zh; ++dz)
#pragma unroll
    for(int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for(int dx = -1; dx <= 1; ++dx, ++u) s = dfma(sten[u], x[i + dz * nn + dy * g.n1 + dx], s);
  return s;
}

// the level's row: grid form or plan
// GRID: every fused level in the grid form (the plan code is left out of the kernel)
template<bool GRID, typename DT>
__device__ __forceinline__ DT tail_row(const TailRow<DT>& R, const int* meta, const DT* tab, TailGeom g, uint64_t i,
                                       const DT* x)
{
  if constexpr(GRID)
    return tail_row_grid(tab, g, int(i), x);
  else
    return g.dim ? tail_row_grid(tab, g, int(i), x) : tail_row_a(R, meta, tab, i, x);
}

template<typename DT>
__device__ __forceinline__ DT tail_prol_grid(TailGeom g, int i, const DT* x)
{
  return dlinear_prol_row<DT, int>(g, i, x);
}

template<typename DT>
__device__ __forceinline__ DT tail_rest_grid(TailGeom g, int jc, const DT* x)
{
  return dlinear_rest_row<DT, int>(g, jc, x);
}

// sum of a transfer row from its FP64 CSR arrays; an FP32 cycle rounds each value to FP32
// as its FP32 plan does
template<typename DT>
__device__ __forceinline__ DT tail_row_csr(const double* val, const uint64_t* col, const uint64_t* rp, uint64_t r,
                                           const DT* x)
{
  // 8 entries in flight, then the fmas in entry order
  const uint64_t k1 = rp[r + 1];
  DT sum = DT(0);
  for(uint64_t k = rp[r]; k < k1; k += 8)
  {
    DT v[8];
    uint64_t c[8];
#pragma unroll
    for(int u = 0; u < 8; ++u)
    {
      v[u] = k + u < k1 ? DT(val[k + u]) : DT(0);
      c[u] = k + u < k1 ? col[k + u] : 0;
    }
#pragma unroll
    for(int u = 0; u < 8; ++u)
      if(k + u < k1) sum = dfma(v[u], x[c[u]], sum);
  }
  return sum;
}

// block-wide sum of one value per thread (all threads call it)
__device__ double tail_sum(double v, double* red)
{
  v = wave_sum(v);
  if((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for(int w = 0; w < kTailThreads / 64; ++w) t += red[w];
  __syncthreads();
  return t;
}

// the skip rule of smoother_skips with the norm already summed
template<typename DT>
__device__ __forceinline__ bool tail_skips(double nrm2)
{
  const double d = sqrt(nrm2);
  const double eps = std::is_same<DT, float>::value ? double(FLT_EPSILON) : DBL_EPSILON;
  return !isfinite(d) || d <= eps * eps;
}

// Richardson(Jacobi) with `steps` steps from x = 0 on rhs (smooth_impl): with def the last
// step writes the defect of x_steps into def and x_steps ends in b0; else the last update is
// accumulated into acc.  b0/b1 are the ping-pong pair.  Thread = row `live ? threadIdx.x : -`.
template<bool GRID, typename DT>
__device__ __forceinline__ void tail_smooth(const TailRow<DT> R, bool live, const int* meta, const DT* tab, TailGeom geo,
                                            int steps, DT omega, bool skip, const DT* rhs, DT* b0, DT* b1, DT* def, DT* acc)
{
  const uint64_t i = threadIdx.x;
  auto B = [&](int k) { return ((steps - k) & 1) ? b1 : b0; };
  const bool acc_first = acc != nullptr && steps == 1;
  if(live)
  {
    DT c = R.inv * rhs[i];
    if(R.masked) c = DT(0);
    DT xn = dfma(omega, c, DT(0));
    if(skip) xn = DT(0);
    if(acc_first)
      acc[i] = dfma(DT(1), xn, acc[i]);
    else
      B(1)[i] = xn;
  }
  __syncthreads();
  const int nsp = def ? steps : steps - 1;
  for(int k = 1; k <= nsp; ++k)
  {
    const DT* xk = B(k);
    const int mode = (def && k == steps) ? 1 : ((acc && k == steps - 1) ? 2 : 0);
    DT* out = mode == 1 ? def : (mode == 2 ? acc : B(k + 1));
    if(live)
    {
      // a filtered row's defect is zero whatever its sum: the sum is skipped
      DT d = DT(0);
      if(!R.masked) d = dfma(tail_row<GRID>(R, meta, tab, geo, i, xk), DT(-1), rhs[i]);
      if(mode == 1)
        out[i] = d;
      else
      {
        DT c = R.inv * d;
        if(R.masked) c = DT(0);
        DT xn = dfma(omega, c, xk[i]);
        if(skip) xn = DT(0);
        out[i] = mode == 0 ? xn : dfma(DT(1), xn, out[i]);
      }
    }
    __syncthreads();
  }
}

// bytes of the tail's LDS vectors and pair tables, rounded up to a double
template<typename DT, typename TT>
__host__ __device__ inline uint64_t tail_words_bytes(const TT& t)
{
  uint64_t words = 0;
  for(int l = 0; l < t.nlevels && l < FEAT_MG_TAIL_MAX_LEVELS; ++l)
    words += (l + 1 < t.nlevels ? 5 * t.lv[l].n + 256 : 2 * t.lv[l].n);
  return (words * sizeof(DT) + 7) / 8 * 8;
}

// TT: feat_mg_tail_t (FP64 cycle) or feat_mg_tail_f32_t (FP32 cycle of the mixed multigrid)
template<typename DT, typename TT, bool GRID>
__global__ __launch_bounds__(kTailThreads) void k_mg_tail(TT t, const DT* __restrict__ rhs_top,
                                                          const double* __restrict__ nrm_top, DT* __restrict__ sol_top)
{
  extern __shared__ double lds_words[];
  __shared__ double red[kTailThreads / 64];
  DT* lds = reinterpret_cast<DT*>(lds_words);
  const int nl = t.nlevels, c = nl - 1;
  const uint64_t tid = threadIdx.x;
  TailBufs<DT> b[FEAT_MG_TAIL_MAX_LEVELS];
  TailRow<DT> row[kTailSmoothed];
  const int* meta[kTailSmoothed];
  TailGeom geo[kTailSmoothed];
  DT* tabs = lds;   // the smoothed levels' pair-table values, 256 each
  // the coarsest level's vectors by name: indexing b with the run-time c would put b in scratch
  DT* crhs = nullptr;
  DT* csol = nullptr;
  {
    DT* p = lds + 256 * c;
#pragma unroll
    for(int l = 0; l < FEAT_MG_TAIL_MAX_LEVELS; ++l)
      if(l < nl)
      {
        const uint64_t n = t.lv[l].n;
        b[l].rhs = p;
        b[l].sol = p + n;
        if(l == c)
        {
          crhs = p;
          csol = p + n;
        }
        p += 2 * n;
        if(l < c)
        {
          b[l].tmp = p;
          b[l].def = p + n;
          b[l].cor = p + 2 * n;
          p += 3 * n;
        }
      }
  }
#pragma unroll
  for(int l = 0; l < kTailSmoothed; ++l)
    if(l < c)
    {
      const auto& L = t.lv[l];
      // grid form: the stencil in the pair table's place, no code words
      geo[l] = TailGeom{L.sten ? L.dim : 0, int(L.n1)};
      meta[l] = L.sten ? nullptr : reinterpret_cast<const int*>(L.table + 256) + 256;
      if(GRID || L.sten)
        for(int e = threadIdx.x; e < (L.dim == 3 ? 27 : 9); e += kTailThreads) tabs[256 * l + e] = L.sten[e];
      else
        for(int e = threadIdx.x; e < 256; e += kTailThreads) tabs[256 * l + e] = L.table[e];
      if(tid < L.n)
      {
        if(!GRID && !L.sten)
        {
          const uint32_t* cw = L.codes + (tid >> 6) * 64 * uint64_t(L.words) + (tid & 63);
#pragma unroll
          for(int w = 0; w < 8; ++w) row[l].wd[w] = w < int(L.words) ? cw[w * 64] : 0xFFFFFFFFu;
        }
        row[l].inv = L.inv_diag[tid];
        row[l].masked = is_masked(L.mask, tid);
      }
    }
  if(tid < t.lv[0].n) b[0].rhs[tid] = rhs_top[tid];
  // the coarse inverse in LDS behind the vectors when it fits (as the launch sized it)
  const double* cinv = t.coarse_inverse;
  {
    const uint64_t nc = t.lv[c].n;
    const uint64_t off = tail_words_bytes<DT>(t);
    if(off + nc * nc * sizeof(double) <= FEAT_MG_TAIL_MAX_LDS)
    {
      double* p = reinterpret_cast<double*>(reinterpret_cast<char*>(lds_words) + off);
      for(uint64_t e = tid; e < nc * nc; e += kTailThreads) p[e] = t.coarse_inverse[e];
      cinv = p;
    }
  }
  __syncthreads();
  double nrm[FEAT_MG_TAIL_MAX_LEVELS];
  nrm[0] = *nrm_top;
  // restriction leg
#pragma unroll
  for(int l = 0; l < kTailSmoothed; ++l)
    if(l < c)
    {
      const auto& L = t.lv[l];
      const bool live = tid < L.n;
      tail_smooth<GRID>(row[l], live, meta[l], tabs + 256 * l, geo[l], L.pre_steps, L.pre_omega, tail_skips<DT>(nrm[l]),
                  b[l].rhs, b[l].sol, b[l].tmp, b[l].def, (DT*)nullptr);
      const auto& C = t.lv[l + 1];
      double part = 0.0;
      if(tid < C.n)
      {
        // a filtered row's value is zero whatever its sum: the sum is skipped
        DT v = DT(0);
        if(!is_masked(C.mask, tid))
        {
          if constexpr(GRID)
            v = dfma(tail_rest_grid(TailGeom{L.dim, int(L.n1)}, int(tid), b[l].def), DT(1), DT(0));
          else
            v = dfma(L.dlinear ? tail_rest_grid(TailGeom{L.dim, int(L.n1)}, int(tid), b[l].def)
                               : tail_row_csr(L.r_val, L.r_col, L.r_ptr, tid, b[l].def),
                     DT(1), DT(0));
        }
        b[l + 1].rhs[tid] = v;
        part = double(v) * double(v);
      }
      nrm[l + 1] = tail_sum(part, red);   // also orders the rhs stores before their use
    }
  // coarsest level: dense inverse, summed in FP64 (feat_dense_gemv_f64 / _mixed)
  {
    const int n = int(t.lv[c].n);
    if(int(tid) < n)
    {
      const double* ar = cinv + size_t(tid) * n;
      double sv = 0.0;
      for(int k = 0; k < n; k += 8)
      {
        double a[8];
#pragma unroll
        for(int u = 0; u < 8; ++u) a[u] = k + u < n ? ar[k + u] : 0.0;
#pragma unroll
        for(int u = 0; u < 8; ++u)
          if(k + u < n) sv = dfma(a[u], double(crhs[k + u]), sv);
      }
      if(is_masked(t.lv[c].mask, tid)) sv = 0.0;
      csol[tid] = DT(sv);
    }
    __syncthreads();
  }
  // prolongation leg
#pragma unroll
  for(int l = kTailSmoothed - 1; l >= 0; --l)
    if(l < c)
    {
      const auto& L = t.lv[l];
      const bool live = tid < L.n;
      if(live)
      {
        DT cc = DT(0);
        if(!row[l].masked)
        {
          if constexpr(GRID)
            cc = dfma(tail_prol_grid(TailGeom{L.dim, int(L.n1)}, int(tid), b[l + 1].sol), DT(1), DT(0));
          else
            cc = dfma(L.dlinear ? tail_prol_grid(TailGeom{L.dim, int(L.n1)}, int(tid), b[l + 1].sol)
                                : tail_row_csr(L.p_val, L.p_col, L.p_ptr, tid, b[l + 1].sol),
                      DT(1), DT(0));
        }
        b[l].sol[tid] = dfma(DT(1), cc, b[l].sol[tid]);
      }
      __syncthreads();
      double part = 0.0;
      if(live)
      {
        DT v = DT(0);
        if(!row[l].masked) v = dfma(tail_row<GRID>(row[l], meta[l], tabs + 256 * l, geo[l], tid, b[l].sol), DT(-1), b[l].rhs[tid]);
        b[l].def[tid] = v;
        part = double(v) * double(v);
      }
      const double dn = tail_sum(part, red);
      tail_smooth<GRID>(row[l], live, meta[l], tabs + 256 * l, geo[l], L.post_steps, L.post_omega, tail_skips<DT>(dn),
                  b[l].def, b[l].cor, b[l].tmp, (DT*)nullptr, b[l].sol);
    }
  if(tid < t.lv[0].n) sol_top[tid] = b[0].sol[tid];
}

template<typename DT, typename TT>
int mg_tail_launch(feat_handle_t h, const TT* t, const DT* rhs, const double* rhs_nrm2, DT* sol, const char* name)
{
  if(t)
    for(int l = 0; l + 1 < t->nlevels && l < FEAT_MG_TAIL_MAX_LEVELS; ++l)
    {
      const auto& L = t->lv[l];
      const uint64_t n1 = L.n1, nn = L.dim == 3 ? n1 * n1 * n1 : n1 * n1, nc = (n1 + 1) / 2;
      const uint64_t ncn = L.dim == 3 ? nc * nc * nc : nc * nc;
      if((L.sten || L.dlinear) && (L.dim < 2 || L.dim > 3 || n1 < 3 || !(n1 & 1) || nn != L.n || ncn != t->lv[l + 1].n))
        return fail(FEAT_ERR_INVALID, "fused coarse levels: grid form inconsistent with the level sizes");
    }
  if(!h || !t || !rhs || !rhs_nrm2 || !sol || t->nlevels < 2 || t->nlevels > FEAT_MG_TAIL_MAX_LEVELS ||
     !t->coarse_inverse)
    return fail(FEAT_ERR_INVALID, name);
  uint64_t words = 0;
  for(int l = 0; l < t->nlevels; ++l)
  {
    const auto& L = t->lv[l];
    words += (l + 1 < t->nlevels ? 5 * L.n + 256 : 2 * L.n);
    if(L.n > uint64_t(kTailThreads)) return fail(FEAT_ERR_INVALID, "fused coarse levels: a level has more than 1024 rows");
    if(l + 1 < t->nlevels &&
       (!L.table || !L.codes || L.words > 8 || !L.inv_diag || !L.r_val || !L.r_col || !L.r_ptr || !L.p_val ||
        !L.p_col || !L.p_ptr || L.pre_steps < 1 || L.post_steps < 1))
      return fail(FEAT_ERR_INVALID, "fused coarse levels: incomplete level");
  }
  if(words * sizeof(DT) > FEAT_MG_TAIL_MAX_LDS) return fail(FEAT_ERR_INVALID, "fused coarse levels: levels exceed the LDS budget");
  // the coarse inverse is staged behind the vectors when it fits
  const uint64_t off = tail_words_bytes<DT>(*t), nc = t->lv[t->nlevels - 1].n;
  const uint64_t lds = off + nc * nc * sizeof(double) <= FEAT_MG_TAIL_MAX_LDS ? off + nc * nc * sizeof(double)
                                                                               : words * sizeof(DT);
  // every smoothed level in the grid form: the kernel without the plan code
  bool grid = true;
  for(int l = 0; l + 1 < t->nlevels; ++l) grid = grid && t->lv[l].sten && t->lv[l].dlinear;
  if(grid)
    hipLaunchKernelGGL((k_mg_tail<DT, TT, true>), dim3(1), dim3(kTailThreads), lds, h->stream, *t, rhs, rhs_nrm2, sol);
  else
    hipLaunchKernelGGL((k_mg_tail<DT, TT, false>), dim3(1), dim3(kTailThreads), lds, h->stream, *t, rhs, rhs_nrm2, sol);
  return check_launch(name);
}

} // namespace

extern "C" int feat_mg_tail_f64(feat_handle_t h, const feat_mg_tail_t* t, const double* rhs, const double* rhs_nrm2,
                                double* sol)
{
  return mg_tail_launch(h, t, rhs, rhs_nrm2, sol, "feat_mg_tail_f64");
}

extern "C" int feat_mg_tail_f32(feat_handle_t h, const feat_mg_tail_f32_t* t, const float* rhs, const double* rhs_nrm2,
                                float* sol)
{
  return mg_tail_launch(h, t, rhs, rhs_nrm2, sol, "feat_mg_tail_f32");
}
