// feat3amd_transpose.hip -- the transposed CSR apply of Arch::Apply::csr in a fixed order.
//
// The reference's generic transposed apply (arch/apply_generic.hpp:37-53) is
//     r = (b/a) * (|b| < eps ? 0 : y);   r[col[i]] += val[i] * x[row];   r = a * r
// with the scatter running over the rows in order, so every r[c] accumulates its
// column's entries in ascending entry order.  Here each call builds the column
// order of the entries (a stable radix sort of the column indices, so entries of
// one column keep their CSR order) and then one lane per column gathers its
// entries in that order: the same additions in the same order, with no atomics,
// so the result is deterministic and bit-identical to the sequential loop.
// The call allocates its sort buffers stream-ordered (this path is set-up work in
// the reference's solvers, which form R = P^T explicitly); the allocation-free variant
// keeps them in a static device array instead.
#include "feat3amd_common.hpp"

#include <mutex>

#include <hipcub/hipcub.hpp>

// Scratch of the allocation-free variant (feat_csr_apply_transposed_noalloc_f64_u64): the
// sort buffers live in this file-scope device array, one copy per GPU, held until the
// process ends.  Host code never passes `scratch` itself to a kernel or copy (on the host
// the name is only a stand-in for the symbol): it uses the per-device address that
// hipGetSymbolAddress returned at handle creation.  Every call on one GPU shares that
// GPU's copy, so calls that overlap on one GPU race; nothing in it needs to start at zero.
static __device__ __align__(256) unsigned char scratch[FEAT_TR_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr size_t kTrScratchBytes = size_t(FEAT_TR_NOALLOC_SCRATCH_BYTES);
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

void* scratch_of(int device)
{
  if(device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  return g_scratch_ptr[device];
}

template<typename IT>
__global__ void k_tr_keys(const IT* col, uint64_t nnz, uint32_t* keys, uint32_t* idx)
{
  for(uint64_t e = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; e < nnz; e += uint64_t(gridDim.x) * blockDim.x)
  {
    keys[e] = uint32_t(col[e]);
    idx[e] = uint32_t(e);
  }
}

template<typename IT>
__global__ void k_tr_row_of(const IT* rowptr, uint64_t rows, uint32_t* row_of)
{
  for(uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < rows; r += uint64_t(gridDim.x) * blockDim.x)
    for(uint64_t e = uint64_t(rowptr[r]); e < uint64_t(rowptr[r + 1]); ++e) row_of[e] = uint32_t(r);
}

// colptr[c] = first position of column c in the sorted keys (empty columns included)
__global__ void k_tr_colptr(const uint32_t* keys, uint64_t nnz, uint64_t cols, uint32_t* colptr)
{
  for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i <= nnz; i += uint64_t(gridDim.x) * blockDim.x)
  {
    const int64_t kp = i == 0 ? -1 : int64_t(keys[i - 1]);
    const int64_t kc = i == nnz ? int64_t(cols) : int64_t(keys[i]);
    for(int64_t c = kp + 1; c <= kc; ++c) colptr[c] = uint32_t(i);
  }
}

template<typename DT>
__global__ void k_tr_gather(DT* r, DT a, DT ba, const DT* x, const DT* y, int bzero, const DT* val,
                            const uint32_t* colptr, const uint32_t* order, const uint32_t* row_of, uint64_t cols)
{
  for(uint64_t c = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; c < cols; c += uint64_t(gridDim.x) * blockDim.x)
  {
    DT acc = ba * (bzero ? DT(0) : y[c]);
    for(uint32_t k = colptr[c]; k < colptr[c + 1]; ++k)
    {
      const uint32_t e = order[k];
      acc = dfma(val[e], x[row_of[e]], acc);
    }
    r[c] = a * acc;
  }
}

template<typename T>
int dev_alloc(T** p, uint64_t n, hipStream_t s)
{
  return check_hip(hipMallocAsync(reinterpret_cast<void**>(p), sizeof(T) * (n ? n : 1), s), "transposed apply: buffer");
}

// 256-byte aligned bump layout inside the scratch array
struct ScratchLayout
{
  size_t off = 0;
  template<typename T>
  size_t take(uint64_t n)
  {
    const size_t o = off;
    off = (off + sizeof(T) * size_t(n ? n : 1) + 255) & ~size_t(255);
    return o;
  }
};

} // namespace

namespace feat3amd_detail {

int noalloc_scratch_attach(int device)
{
  if(device < 0 || device >= kMaxDevices) return fail(FEAT_ERR_UNSUPPORTED, "no-alloc scratch: device index above 63");
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  if(g_scratch_ptr[device]) return FEAT_OK;
  void* p = nullptr;
  const int rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  if(rc) return rc;
  g_scratch_ptr[device] = p;
  return FEAT_OK;
}

// The transposed apply of csr_apply_transposed with every buffer in the device's scratch
// array: no allocation, no free; sizes that do not fit are rejected before any launch.
template<typename DT, typename IT>
int csr_apply_transposed_noalloc(feat_handle_t h, DT* r, DT a, const DT* x, DT b, bool bzero, const DT* y,
                                 const DT* val, const IT* col, const IT* rowptr, uint64_t rows, uint64_t cols,
                                 uint64_t nnz, const char* name)
{
  if(nnz >= (uint64_t(1) << 31) || rows >= (uint64_t(1) << 32) || cols >= (uint64_t(1) << 32))
    return fail(FEAT_ERR_UNSUPPORTED, "csr_apply_transposed_noalloc: needs fewer than 2^31 entries and 2^32 rows, columns");
  unsigned char* base = static_cast<unsigned char*>(scratch_of(h->device));
  if(!base) return fail(FEAT_ERR_NO_DEVICE, "csr_apply_transposed_noalloc: scratch not attached on the handle's device");
  if(!cols) return FEAT_OK;
  hipStream_t s = h->stream;
  int end_bit = 1;
  while(end_bit < 32 && (uint64_t(1) << end_bit) < cols) ++end_bit;
  ScratchLayout lay;
  const size_t o_colptr = lay.take<uint32_t>(cols + 1);
  const size_t o_k0 = lay.take<uint32_t>(nnz), o_k1 = lay.take<uint32_t>(nnz);
  const size_t o_v0 = lay.take<uint32_t>(nnz), o_v1 = lay.take<uint32_t>(nnz);
  const size_t o_row = lay.take<uint32_t>(nnz);
  auto at = [&](size_t o) { return reinterpret_cast<uint32_t*>(base + o); };
  hipcub::DoubleBuffer<uint32_t> dk(at(o_k0), at(o_k1)), dv(at(o_v0), at(o_v1));
  size_t tmp_bytes = 0;
  if(nnz)
  {
    const int rc = check_hip(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, dk, dv, int(nnz), 0, end_bit, s),
                             "transposed apply: sort size");
    if(rc) return rc;
  }
  const size_t o_tmp = lay.take<unsigned char>(tmp_bytes);
  if(lay.off > kTrScratchBytes)
    return fail(FEAT_ERR_UNSUPPORTED, "csr_apply_transposed_noalloc: the matrix needs more than the 32 MiB scratch "
                                      "(about 20 bytes per entry and 4 per column)");
  uint32_t* colptr = at(o_colptr);
  if(nnz)
  {
    hipLaunchKernelGGL((k_tr_keys<IT>), dim3(grid_for(nnz, 4)), dim3(kThreads), 0, s, col, nnz, dk.Current(),
                       dv.Current());
    hipLaunchKernelGGL((k_tr_row_of<IT>), dim3(grid_for(rows)), dim3(kThreads), 0, s, rowptr, rows, at(o_row));
    int rc = check_launch(name);
    if(!rc)
      rc = check_hip(hipcub::DeviceRadixSort::SortPairs(static_cast<void*>(base + o_tmp), tmp_bytes, dk, dv, int(nnz), 0,
                                                        end_bit, s),
                     "transposed apply: sort");
    if(rc) return rc;
  }
  hipLaunchKernelGGL(k_tr_colptr, dim3(grid_for(nnz + 1, 4)), dim3(kThreads), 0, s, dk.Current(), nnz, cols, colptr);
  hipLaunchKernelGGL((k_tr_gather<DT>), dim3(grid_for(cols)), dim3(kThreads), 0, s, r, a, DT(b / a), x, y, int(bzero), val,
                     colptr, dv.Current(), at(o_row), cols);
  return check_launch(name);
}

template int csr_apply_transposed_noalloc<double, uint64_t>(feat_handle_t, double*, double, const double*, double, bool,
                                                            const double*, const double*, const uint64_t*,
                                                            const uint64_t*, uint64_t, uint64_t, uint64_t, const char*);

template<typename DT, typename IT>
int csr_apply_transposed(feat_handle_t h, DT* r, DT a, const DT* x, DT b, bool bzero, const DT* y, const DT* val,
                         const IT* col, const IT* rowptr, uint64_t rows, uint64_t cols, uint64_t nnz, const char* name)
{
  if(nnz >= (uint64_t(1) << 31) || rows >= (uint64_t(1) << 32) || cols >= (uint64_t(1) << 32))
    return fail(FEAT_ERR_UNSUPPORTED, "csr_apply: transposed apply needs fewer than 2^31 entries and 2^32 rows, columns");
  if(!cols) return FEAT_OK;
  hipStream_t s = h->stream;
  uint32_t *keys = nullptr, *keys_s = nullptr, *idx = nullptr, *order = nullptr, *row_of = nullptr, *colptr = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int end_bit = 1;
  while(end_bit < 32 && (uint64_t(1) << end_bit) < cols) ++end_bit;
  int rc = FEAT_OK;
  auto release = [&]() {
    for(void* p : {(void*)keys, (void*)keys_s, (void*)idx, (void*)order, (void*)row_of, (void*)colptr, tmp})
      if(p) (void)hipFreeAsync(p, s);
  };
  rc = dev_alloc(&colptr, cols + 1, s);
  if(!rc && nnz)
  {
    rc = dev_alloc(&keys, nnz, s);
    if(!rc) rc = dev_alloc(&keys_s, nnz, s);
    if(!rc) rc = dev_alloc(&idx, nnz, s);
    if(!rc) rc = dev_alloc(&order, nnz, s);
    if(!rc) rc = dev_alloc(&row_of, nnz, s);
    if(!rc)
      rc = check_hip(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys, keys_s, idx, order, int(nnz), 0, end_bit, s),
                     "transposed apply: sort size");
    if(!rc) rc = check_hip(hipMallocAsync(&tmp, tmp_bytes ? tmp_bytes : 1, s), "transposed apply: sort buffer");
    if(!rc)
    {
      hipLaunchKernelGGL((k_tr_keys<IT>), dim3(grid_for(nnz, 4)), dim3(kThreads), 0, s, col, nnz, keys, idx);
      hipLaunchKernelGGL((k_tr_row_of<IT>), dim3(grid_for(rows)), dim3(kThreads), 0, s, rowptr, rows, row_of);
      rc = check_launch(name);
    }
    if(!rc)
      rc = check_hip(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys_s, idx, order, int(nnz), 0, end_bit, s),
                     "transposed apply: sort");
  }
  if(!rc)
  {
    hipLaunchKernelGGL(k_tr_colptr, dim3(grid_for(nnz + 1, 4)), dim3(kThreads), 0, s, keys_s, nnz, cols, colptr);
    hipLaunchKernelGGL((k_tr_gather<DT>), dim3(grid_for(cols)), dim3(kThreads), 0, s, r, a, DT(b / a), x, y, int(bzero),
                       val, colptr, order, row_of, cols);
    rc = check_launch(name);
  }
  release();
  return rc;
}

template int csr_apply_transposed<double, uint64_t>(feat_handle_t, double*, double, const double*, double, bool,
                                                    const double*, const double*, const uint64_t*, const uint64_t*,
                                                    uint64_t, uint64_t, uint64_t, const char*);
template int csr_apply_transposed<double, uint32_t>(feat_handle_t, double*, double, const double*, double, bool,
                                                    const double*, const double*, const uint32_t*, const uint32_t*,
                                                    uint64_t, uint64_t, uint64_t, const char*);
template int csr_apply_transposed<float, uint64_t>(feat_handle_t, float*, float, const float*, float, bool,
                                                   const float*, const float*, const uint64_t*, const uint64_t*,
                                                   uint64_t, uint64_t, uint64_t, const char*);
template int csr_apply_transposed<float, uint32_t>(feat_handle_t, float*, float, const float*, float, bool,
                                                   const float*, const float*, const uint32_t*, const uint32_t*,
                                                   uint64_t, uint64_t, uint64_t, const char*);

} // namespace feat3amd_detail
