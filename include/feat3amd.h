/*
 * feat3amd.h -- C-ABI of the MI355X (gfx950) LAFEM + Solver hot path.
 *
 * This is the drop-in boundary.  Each entry point replaces one raw-pointer
 * kernel of the reference's LAFEM::Arch layer (kernel/lafem/arch/<op>.hpp,
 * dispatched by BACKEND_SKELETON_* in kernel/backend.hpp:55-120) or one fused
 * group of them used by the solver loops (kernel/solver/<solver>.hpp).  The citation
 * above each declaration names the reference interface it replaces.
 *
 * Conventions
 *  - All array arguments are DEVICE pointers (hipMalloc or feat_malloc) unless
 *    the name ends in _host.  The callee never keeps or frees them.
 *  - Every call is stream-ordered on the handle's stream and returns an int
 *    status (FEAT_OK = 0).  Reductions write a DEVICE scalar; the *_host
 *    variants synchronise the stream and return a host value, matching the
 *    reference's `DT_ value(...)` signatures.
 *  - Aliasing rules follow the reference: r == y allowed in SpMV
 *    (lafem/sparse_matrix_csr.hpp:1799), r == x forbidden (:1750,1837);
 *    axpy allows r == x and r == y (arch/axpy_generic.hpp:25-50).
 *  - Index types: u64 (reference default `Index`) and u32 (FEAT_INDEX_U32).
 *  - No entry point allocates device memory in the hot path: reductions use
 *    the handle's preallocated workspace; fused solver kernels take their
 *    workspace (feat_reduce_ws_t) from the caller.
 */
#ifndef FEAT3AMD_H
#define FEAT3AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- status codes ---------------- */
#define FEAT_OK 0
#define FEAT_ERR_INVALID 1     /* bad argument (NULL pointer, size mismatch)   */
#define FEAT_ERR_HIP 2         /* HIP runtime error                            */
#define FEAT_ERR_ALIAS 3       /* forbidden aliasing (e.g. SpMV r == x)        */
#define FEAT_ERR_NOMEM 4       /* allocation failed (setup calls only)         */
#define FEAT_ERR_UNSUPPORTED 5 /* variant not provided                         */
#define FEAT_ERR_NO_DEVICE 6   /* no gfx950 device / library not initialised    */

typedef struct feat_handle_s* feat_handle_t;
typedef void* feat_stream_t; /* a hipStream_t */

/* ---------------- runtime (kernel/runtime.cpp:37-148, util/cuda_util.cu:100-136) ---------------- */
/* Thread-local message of the last failing call. */
const char* feat_last_error(void);
/* Creates a handle on `device` that launches on `stream` (NULL: the device's
 * null stream).  Allocates the handle's reduction workspace. */
int feat_handle_create(feat_handle_t* h, int device, feat_stream_t stream);
int feat_handle_destroy(feat_handle_t h);
int feat_handle_stream(feat_handle_t h, feat_stream_t* stream);
int feat_sync(feat_handle_t h);
/* Streams, ordering and graphs (no reference counterpart: the reference's kernels
 * are synchronous).  A handle on a non-null stream can record a HIP graph of the
 * calls made on it between feat_graph_begin and feat_graph_end (setup time), and
 * replay it with feat_graph_launch.  Calls made while recording must not
 * synchronise or allocate. */
typedef struct feat_graph_s* feat_graph_t;
int feat_stream_create(feat_stream_t* s);    /* non-blocking, default priority */
int feat_stream_destroy(feat_stream_t s);
/* `waiter`'s stream waits for the work issued so far on `signaler`'s stream */
int feat_handle_wait(feat_handle_t waiter, feat_handle_t signaler);
int feat_graph_begin(feat_handle_t h);
int feat_graph_end(feat_handle_t h, feat_graph_t* g);
int feat_graph_launch(feat_handle_t h, feat_graph_t g);
int feat_graph_destroy(feat_graph_t g);
/* stream-ordered copy of one device scalar into pinned host memory (recordable) */
int feat_copy_scalar_to_host(feat_handle_t h, const double* dev, double* host_pinned);
/* device memory (MemoryPool::allocate_memory/release_memory, util/memory_pool.hpp:79-147) */
int feat_malloc(void** p, size_t bytes);
int feat_free(void* p);
int feat_host_malloc(void** p, size_t bytes); /* pinned host memory */
int feat_host_free(void* p);
/* MemoryPool::copy / set_memory (util/memory_pool.hpp:181-251) */
int feat_copy(feat_handle_t h, void* dst, const void* src, size_t bytes);      /* any direction */
/* Device-to-device copy by a kernel (16 bytes per lane where both pointers and the size
 * are 16-byte aligned, else hipMemcpyAsync); both buffers must be device memory.
 * DenseVector::copy (dense_vector.hpp) on device containers. */
int feat_copy_device(feat_handle_t h, void* dst, const void* src, size_t bytes);
/* Bandwidth probe (no reference counterpart): reads the n doubles of x once with 16-byte
 * nontemporal loads and writes nothing (sink is a device scalar that is never written in
 * practice).  n even, x 16-byte aligned.  The measured read ceiling for roofline fractions. */
int feat_stream_read_f64(feat_handle_t h, const double* x, uint64_t n, double* sink);
int feat_set_memory_f64(feat_handle_t h, double* p, double v, uint64_t n);
int feat_set_memory_f32(feat_handle_t h, float* p, float v, uint64_t n);
/* MemoryPool::convert (util/memory_pool.hpp:231-251, util/cuda_util.cu:226-244) */
int feat_convert_f64_f32(feat_handle_t h, float* dst, const double* src, uint64_t n);
int feat_convert_f32_f64(feat_handle_t h, double* dst, const float* src, uint64_t n);
int feat_convert_f64_f16(feat_handle_t h, uint16_t* dst, const double* src, uint64_t n);
int feat_convert_f16_f64(feat_handle_t h, double* dst, const uint16_t* src, uint64_t n);
int feat_convert_u64_u32(feat_handle_t h, uint32_t* dst, const uint64_t* src, uint64_t n);

/* ---------------- SpMV (arch/apply.hpp:51-79, generic arch/apply_generic.hpp:25-69) ----------------
 * r = a*A*x + b*y, or r = a*A^T*x + b*y when transposed != 0.  |b| < eps
 * means y is not read (r is treated as 0), as in the generic code.  The
 * non-transposed form sums each row in storage order with one fma per entry
 * (bit-identical to the sequential CPU loop); the transposed form computes
 * (b/a)*y, adds each column's entries in ascending entry order (the order of the
 * generic row-by-row scatter) and scales by a: bit-identical too.  The
 * transposed form allocates stream-ordered sort buffers (about 20 bytes per
 * entry) in each call and needs nnz < 2^31; it is not for recorded graphs. */
int feat_csr_apply_f64_u64(feat_handle_t h, double* r, double a, const double* x, double b,
                           const double* y, const double* val, const uint64_t* col,
                           const uint64_t* rowptr, uint64_t rows, uint64_t cols, uint64_t nnz,
                           int transposed);
int feat_csr_apply_f64_u32(feat_handle_t h, double* r, double a, const double* x, double b,
                           const double* y, const double* val, const uint32_t* col,
                           const uint32_t* rowptr, uint64_t rows, uint64_t cols, uint64_t nnz,
                           int transposed);
int feat_csr_apply_f32_u32(feat_handle_t h, float* r, float a, const float* x, float b,
                           const float* y, const float* val, const uint32_t* col,
                           const uint32_t* rowptr, uint64_t rows, uint64_t cols, uint64_t nnz,
                           int transposed);
int feat_csr_apply_f32_u64(feat_handle_t h, float* r, float a, const float* x, float b,
                           const float* y, const float* val, const uint64_t* col,
                           const uint64_t* rowptr, uint64_t rows, uint64_t cols, uint64_t nnz,
                           int transposed);

/* The transposed form of feat_csr_apply_f64_u64 (transposed != 0) without any allocation:
 * no counterpart in the reference.  It is for callers that cannot allocate device memory
 * in the call (HIP graph capture, serving loops, harnesses that forbid it).  Its sort
 * buffers live in a static device array of FEAT_TR_NOALLOC_SCRATCH_BYTES per GPU, held
 * until the process ends and looked up when a handle is created; a matrix needing more
 * (about 20 bytes per entry plus 4 per column and the sort's own few KB) is rejected with
 * FEAT_ERR_UNSUPPORTED before anything runs.  Same arithmetic, bit-identical results.
 * Calls that overlap on one GPU (any handle, any stream) share that array and race. */
#define FEAT_TR_NOALLOC_SCRATCH_BYTES (32u << 20)
int feat_csr_apply_transposed_noalloc_f64_u64(feat_handle_t h, double* r, double a, const double* x,
                                              double b, const double* y, const double* val,
                                              const uint64_t* col, const uint64_t* rowptr,
                                              uint64_t rows, uint64_t cols, uint64_t nnz);

/* BCSR with bh x bw row-major blocks, (bh,bw) in {(2,2),(2,1),(1,2),(3,3)}
 * (arch/apply.hpp:108-133, generic arch/apply_generic.hpp:121-154). */
int feat_bcsr_apply_f64_u64(feat_handle_t h, int bh, int bw, double* r, double a,
                            const double* x, double b, const double* y, const double* val,
                            const uint64_t* col, const uint64_t* rowptr, uint64_t rows,
                            uint64_t cols, uint64_t nnz);

/* ---------------- BLAS-1 (arch/{axpy,dot_product,norm,scale,component_product,component_invert}.hpp) ---------------- */
/* Axpy::value: r = a*x + y (arch/axpy.hpp; generic arch/axpy_generic.hpp:25-50).
 * r == x rounds a*x before adding y, like the generic two-statement branch. */
int feat_axpy_f64(feat_handle_t h, double* r, double a, const double* x, const double* y,
                  uint64_t n);
int feat_axpy_f32(feat_handle_t h, float* r, float a, const float* x, const float* y,
                  uint64_t n);
/* DotProduct::value (arch/dot_product.hpp; generic :21-40): deterministic
 * two-stage reduction; result written to the device scalar *res. */
int feat_dot_f64(feat_handle_t h, double* res, const double* x, const double* y, uint64_t n);
int feat_dot_f64_host(feat_handle_t h, double* res_host, const double* x, const double* y,
                      uint64_t n);
/* TripleDotProduct::value (arch/dot_product_generic.hpp:44-70): sum x*y*z */
int feat_triple_dot_f64_host(feat_handle_t h, double* res_host, const double* x,
                             const double* y, const double* z, uint64_t n);
/* Norm2::value (arch/norm.hpp; generic arch/norm_generic.hpp:24-33): sqrt(sum x^2) */
int feat_norm2_f64(feat_handle_t h, double* res, const double* x, uint64_t n);
int feat_norm2_f64_host(feat_handle_t h, double* res_host, const double* x, uint64_t n);
/* Scale::value (arch/scale_generic.hpp:22-38): r = x*s */
int feat_scale_f64(feat_handle_t h, double* r, const double* x, double s, uint64_t n);
/* ComponentProduct::value (arch/component_product_generic.hpp:22-52): r = x∘y (FP64
 * throughout; the reference's CUDA kernel rounds through float, component_product.cu:27) */
int feat_component_product_f64(feat_handle_t h, double* r, const double* x, const double* y,
                               uint64_t n);
/* ComponentInvert::value (arch/component_invert_generic.hpp:21-36): r = s/x */
int feat_component_invert_f64(feat_handle_t h, double* r, const double* x, double s,
                              uint64_t n);
/* the same for float (JacobiPrecond<float>::init_numeric, jacobi_precond.hpp:136-143,
 * inverts the FP32 diagonal with float(omega)) */
int feat_component_invert_f32(feat_handle_t h, float* r, const float* x, float s, uint64_t n);
/* Diagonal::csr (arch/diagonal_generic.hpp:21-38): index of the diagonal
 * entry of each row, nnz where absent */
int feat_csr_diag_index_u64(feat_handle_t h, uint64_t* diag, const uint64_t* col,
                            const uint64_t* rowptr, uint64_t rows);
/* SparseMatrixCSR::extract_diag (lafem/sparse_matrix_csr.hpp:2277-2303): diag values, 0 where absent */
int feat_csr_extract_diag_f64_u64(feat_handle_t h, double* diag, const double* val,
                                  const uint64_t* col, const uint64_t* rowptr, uint64_t rows);

/* ---------------- filters ---------------- */
/* UnitFilter::filter_rhs/filter_sol (lafem/unit_filter.hpp:376-395,
 * arch/unit_filter_generic.hpp:22-28): v[idx[i]] = sv[i] */
int feat_unit_filter_rhs_f64_u64(feat_handle_t h, double* v, const double* sv_val,
                                 const uint64_t* sv_idx, uint64_t ue);
/* UnitFilter::filter_def/filter_cor (unit_filter.hpp:402-420,
 * arch/unit_filter_generic.hpp:31-38): v[idx[i]] = 0 */
int feat_unit_filter_def_f64_u64(feat_handle_t h, double* v, const uint64_t* sv_idx,
                                 uint64_t ue);
/* Bit mask of the unit filter (1 bit per row, 32-bit words, ceil(n/32) words,
 * zeroed by the caller-visible call itself).  Used by the fused kernels to
 * apply filter_def/filter_cor in their epilogues. */
int feat_unit_filter_mask_u64(feat_handle_t h, uint32_t* mask, uint64_t n,
                              const uint64_t* sv_idx, uint64_t ue);
/* UnitFilterBlocked::filter_rhs / filter_def (lafem/unit_filter_blocked.hpp:449-495) ->
 * Arch::UnitFilterBlocked::filter_rhs<bs> / filter_def<bs>(v, sv_elements, sv_indices, ue,
 * ign_nans) (arch/unit_filter_blocked_generic.hpp:24-60): for entry i and component j,
 * v[bs*sv_indices[i] + j] = sv_elements[bs*i + j] (rhs) or 0 (def); with ign_nans a NaN
 * in sv_elements[bs*i + j] leaves that component unchanged.  filter_def reads
 * sv_elements only when ign_nans is set (may be NULL otherwise). */
int feat_unit_filter_blocked_rhs_f64_u64(feat_handle_t h, int bs, double* v, const double* sv_elements,
                                         const uint64_t* sv_indices, uint64_t ue, int ign_nans);
int feat_unit_filter_blocked_def_f64_u64(feat_handle_t h, int bs, double* v, const double* sv_elements,
                                         const uint64_t* sv_indices, uint64_t ue, int ign_nans);
/* MeanFilter::filter_def / filter_cor (lafem/mean_filter.hpp:214-264):
 * v += (-(v . w_dot) / volume) * w_add, all on device */
int feat_mean_filter_f64(feat_handle_t h, double* v, const double* w_dot, const double* w_add,
                         double volume, uint64_t n);

/* ---------------- fused solver kernels ----------------
 * These replace op sequences of solver/{pcg,richardson,jacobi_precond,multigrid}.hpp
 * with fewer HBM passes.  Each keeps the reference's per-element arithmetic
 * (same roundings, same per-row summation order); only the order of the
 * global sums inside dot/norm differs (deterministic tree).
 *
 * A reduction workspace: `partials` holds 2*FEAT_REDUCE_MAX_BLOCKS doubles,
 * `counter` one zero-initialised uint32 (left at zero after every call). */
#define FEAT_REDUCE_MAX_BLOCKS 16384
typedef struct {
  double* partials;
  uint32_t* counter;
} feat_reduce_ws_t;

typedef struct {            /* a CSR matrix on the device (u64 indices) */
  uint64_t rows, cols, nnz;
  const double* val;
  const uint64_t* col;
  const uint64_t* rowptr;
} feat_csr_f64_t;

/* Matrix descriptor of the fused kernels: the caller's CSR, or a SELL-64
 * execution plan of it.  SELL-64 stores each slice of 64 consecutive rows
 * column-major (entry k of row 64s+l at slice_off[s] + 64k + l), so one
 * wavefront reads 64 consecutive values per load; every row keeps its CSR
 * entry order, so each row sum is bit-identical to the CSR loop.  Column
 * indices are u32 (absolute) or i16 (col - row) when the bandwidth allows. */
#define FEAT_FMT_CSR 0
#define FEAT_FMT_SELL 1
#define FEAT_FMT_DLINEAR 2   /* no arrays: d-linear injection between structured grids,
                              * see feat_matrix_dlinear_f64 (index = dim, slice_words =
                              * n1, value_kind = transpose)                          */
#define FEAT_FMT_BANDED 3    /* the clipped 3^dim-point stencil rows of a structured grid
                              * (n1 nodes per side, lexicographic, x fastest; rows = cols =
                              * n1^dim): columns formed from the node position, values
                              * slot-major (val[t * rows + i], slot t = (dx+1) + 3 (dy+1) +
                              * 9 (dz+1), in ascending column order; 0 where the neighbour
                              * lies outside the grid); slice_words = n1, the dimension is 3
                              * where n1^3 = rows, else 2; col, ptr, rowlen unused        */
#define FEAT_IDX_U64 0   /* CSR: u64 col/rowptr                              */
#define FEAT_IDX_U32 1   /* CSR: u32 col/rowptr; SELL: u32 absolute columns  */
#define FEAT_IDX_D16 2   /* SELL: int16 column minus row                      */
#define FEAT_IDX_DICT 3  /* SELL: one-byte codes of (column - row, value) pairs;
                          * val -> pair table: 256 values of the arithmetic type
                          * followed by 256 int32 offsets; col -> codes, four
                          * entries per 32-bit lane word; ptr -> word offsets   */
#define FEAT_IDX_DIAG 4  /* SELL: the pair dictionary by diagonal: byte t of a
                          * row's code words is the pair code of its entry at the
                          * t-th smallest column offset (255: none); val -> the
                          * pair table followed by int32 count, offsets, and how
                          * each x value is obtained; col -> codes, slice_words
                          * words per row, 64-row slices lane-interleaved        */
#define FEAT_VAL_NATIVE 0 /* values in the descriptor's arithmetic type     */
#define FEAT_VAL_F16 1    /* FP32 arithmetic on FP16 values times `scale`   */
#define FEAT_VAL_CODE8 2  /* SELL (i16/u32 index): each stored value is a
                           * 1-byte code into a table of 256 values of the
                           * arithmetic type; val -> the table, followed by
                           * the codes in the layout of the columns          */
typedef struct {
  int format, index;
  uint64_t rows, cols;
  uint64_t nnz;              /* CSR nonzeros; SELL: stored (padded) entries   */
  const double* val;
  const void* col;
  const void* ptr;           /* CSR row pointer / SELL slice offsets (u64)    */
  const uint16_t* rowlen;    /* SELL row lengths                             */
  int value_kind;            /* FEAT_VAL_NATIVE or FEAT_VAL_CODE8 (SELL)      */
  float scale;               /* 1                                             */
  int tune;                  /* SELL launch shape, 0 = default (autotuner)    */
  uint32_t slice_words;      /* pair-dictionary plans: code words per lane in
                                every slice (slice s at s * 64 * slice_words),
                                0: per-slice offsets in ptr                   */
} feat_matrix_f64_t;

typedef struct {             /* FP32 arithmetic (mixed-precision inner solver) */
  int format, index;
  uint64_t rows, cols, nnz;
  const float* val;          /* FP32 values, or FP16 values (value_kind F16)  */
  const void* col;
  const void* ptr;
  const uint16_t* rowlen;
  int value_kind;            /* FEAT_VAL_NATIVE, _F16 or _CODE8 (SELL only)    */
  float scale;               /* row sums are multiplied by this power of two  */
  int tune;                  /* SELL launch shape, 0 = default                */
  uint32_t slice_words;      /* as feat_matrix_f64_t::slice_words             */
} feat_matrix_f32_t;

/* A CSR descriptor (no copy) of the caller's arrays. */
int feat_matrix_from_csr_f64(const feat_csr_f64_t* A, feat_matrix_f64_t* out);
/* A matrix-free descriptor of the d-linear injection from the grid with (n1 + 1) / 2 nodes
 * per side to the grid with n1 (odd, >= 3) per side, dim 2 or 3, nodes in lexicographic
 * order.  transpose = 0: rows = the fine nodes, each taking the coarse nodes around it in
 * ascending order with the weight 1/2 per odd axis (the prolongation); transpose = 1: its
 * transpose (the restriction), each coarse node taking the fine nodes (2X + dx, ...),
 * dx in {-1, 0, 1}, inside the grid in ascending order.  Row sums run in that order, which
 * is the CSR order of the assembled Q1 grid transfer of such grids, so where every entry of
 * an assembled transfer is this injection (the caller checks) the results are bit-identical
 * to its CSR or SELL-64 plan.  Every entry point that takes a matrix descriptor accepts it.
 * No counterpart in the reference (its transfers are assembled matrices). */
int feat_matrix_dlinear_f64(int dim, uint32_t n1, int transpose, feat_matrix_f64_t* out);
int feat_matrix_dlinear_f32(int dim, uint32_t n1, int transpose, feat_matrix_f32_t* out);
/* Builds the SELL-64 plan of a CSR matrix on the device (setup time; allocates).
 * Fails with FEAT_ERR_UNSUPPORTED when a row has more than 65535 entries or the
 * slice padding would exceed the CSR size; the caller then keeps the CSR.
 * A matrix with at most 256 distinct (column - row, value) pairs -- FE matrices
 * on uniformly refined meshes: 42 for 3D Q1 -- gets the lossless pair-dictionary
 * plan (FEAT_IDX_DICT, one byte per entry); others get i16/u32 columns and, with
 * at most 256 distinct values (transfer operators: 4), 1-byte value codes
 * (FEAT_VAL_CODE8), else full values. */
int feat_sell_create_f64(feat_handle_t h, const feat_csr_f64_t* A, feat_matrix_f64_t* out);
/* Same with flags: FEAT_SELL_NO_DICT builds the plain plan (full values) even
 * where the pair dictionary or the value codes apply. */
#define FEAT_SELL_NO_DICT 1
#define FEAT_SELL_NO_DIAG 2   /* keep the entry-order dictionary (FEAT_IDX_DICT) */
/* A square matrix without a pair dictionary or value codes whose every row holds exactly the
 * grid neighbours (i+dx, j+dy, k+dz), |d| <= 1, inside an n1^dim grid, in ascending order
 * (Q1 matrices with variable coefficients) gets the banded plan (FEAT_FMT_BANDED: 8, 4 or
 * 2 bytes per entry in FP64, FP32 or FP16 values against 10, 6 or 4 in the plain plan;
 * the same row sums, bit for bit; up to 19.8 M rows).  FEAT_SELL_NO_BANDED keeps the
 * plain plan. */
#define FEAT_SELL_NO_BANDED 4
int feat_sell_create_ex_f64(feat_handle_t h, const feat_csr_f64_t* A, int flags, feat_matrix_f64_t* out);
int feat_sell_create_ex_f32(feat_handle_t h, const feat_csr_f64_t* A, int flags, feat_matrix_f32_t* out);
int feat_sell_create_ex_f16(feat_handle_t h, const feat_csr_f64_t* A, int flags, feat_matrix_f32_t* out);
/* Same plan with values rounded to FP32 (ConvertPrecond, convert_precond.hpp:100-119). */
int feat_sell_create_f32(feat_handle_t h, const feat_csr_f64_t* A, feat_matrix_f32_t* out);
/* Same plan with FP16 values: v16 = half(float(a) * 2^-e) with 2^(e-1) <= max|a| < 2^e;
 * row sums sum(fmaf(float(v16), x, s)) are multiplied by scale = 2^e. */
int feat_sell_create_f16(feat_handle_t h, const feat_csr_f64_t* A, feat_matrix_f32_t* out);
/* Frees a plan made by feat_sell_create_* (SELL-64 or banded). */
int feat_sell_destroy_f64(feat_matrix_f64_t* m);
int feat_sell_destroy_f32(feat_matrix_f32_t* m);

/* r = a*A*x + b*y on a descriptor (same semantics as feat_csr_apply_f64_u64). */
int feat_matrix_apply_f64(feat_handle_t h, const feat_matrix_f64_t* A, double* r, double a,
                          const double* x, double b, const double* y);

/* Defect d = rhs - A*x (SparseMatrixCSR::apply(r,x,y,-1), sparse_matrix_csr.hpp:1807-1843)
 * followed by filter_def through `mask` (may be NULL) and, when nrm2 != NULL,
 * the squared norm of the filtered defect into the device scalar *nrm2. */
int feat_defect_f64(feat_handle_t h, const feat_matrix_f64_t* A, double* d, const double* rhs,
                    const double* x, const uint32_t* mask, double* nrm2,
                    const feat_reduce_ws_t* ws);
/* r = A*x (b = 0) with filter mask and optional squared norm (Transfer::rest + filter_def,
 * multigrid.hpp:1801-1805).  `dotw` != NULL adds the dot of the result with dotw into *dot. */
int feat_apply_masked_f64(feat_handle_t h, const feat_matrix_f64_t* A, double* r,
                          const double* x, const uint32_t* mask, double* nrm2,
                          const double* dotw, double* dot, const feat_reduce_ws_t* ws);
/* Prolongation + correction: sol += filter_cor(P*xc) (Transfer::prol, filter_cor, axpy;
 * multigrid.hpp:1885-1930). */
int feat_prol_add_f64(feat_handle_t h, const feat_matrix_f64_t* P, double* sol,
                      const double* xc, const uint32_t* mask);

/* Richardson(JacobiPrecond) smoother with min_iter == max_iter == steps
 * (richardson.hpp:160-226, jacobi_precond.hpp:159-164): starting from x = 0,
 * `steps` times { c = filter_cor(inv_diag∘d); x += omega*c; d = rhs - A*x; filter_def(d) }.
 * The smoother is skipped (x = 0, d = rhs) when sqrt(*rhs_nrm2) is not finite or
 * <= eps^2 (iterative.hpp:748-770); rhs_nrm2 is a device scalar (NULL: never skip).
 *   x        out: the smoother result (n); scratch in acc mode.   tmp: scratch (n)
 *   d_out    optional: the last filtered defect (n) -- equals the defect the multigrid
 *            recomputes after pre-smoothing (multigrid.hpp:1754-1774)
 *   acc      optional: if set, the result is added to it instead of stored in x
 *            (post-smoothing `sol += cor`, multigrid.hpp:1977); the last defect is
 *            then not formed (it is unused by the reference)
 *   dotw/dot optional: dot(acc_after, dotw) into the device scalar *dot. */
int feat_rj_smooth_f64(feat_handle_t h, const feat_matrix_f64_t* A, const double* inv_diag,
                       const uint32_t* mask, int steps, double omega, const double* rhs,
                       const double* rhs_nrm2, double* x, double* tmp, double* d_out,
                       double* acc, const double* dotw, double* dot,
                       const feat_reduce_ws_t* ws);

/* The smoother's first iterate formed where its right-hand side is produced (multigrid.hpp:
 * 1753 formats x, richardson.hpp:205-221 then takes the first step): the producer also
 * writes x1 = filter_cor(omega * (inv_diag ∘ v)) of its output v, without the skip rule,
 * and feat_rj_smooth_from1_f64 (same arguments as feat_rj_smooth_f64, steps >= 2) starts
 * from x1 in its first ping-pong buffer -- tmp for even steps, x for odd -- and applies the
 * skip rule in its next step.  Same values as the separate first step; no reference
 * counterpart (one vector pass less per smoother call).
 *   feat_defect_first_f64:       feat_defect_f64 + x1 of the post-smoother (rhs = d)
 *   feat_apply_masked_first_f64: feat_apply_masked_f64 (no dot) + x1 of the pre-smoother
 *   feat_pcg_update_first_f64:   feat_pcg_update_f64 + x1 of the preconditioner's
 *                                pre-smoother on r, filtered through `mask`; x1 may
 *                                be q (not x, r or p) */
int feat_defect_first_f64(feat_handle_t h, const feat_matrix_f64_t* A, double* d, const double* rhs,
                          const double* x, const uint32_t* mask, double* nrm2, const double* inv_diag,
                          double omega, double* x1, const feat_reduce_ws_t* ws);
int feat_apply_masked_first_f64(feat_handle_t h, const feat_matrix_f64_t* A, double* r,
                                const double* x, const uint32_t* mask, double* nrm2,
                                const double* inv_diag, double omega, double* x1,
                                const feat_reduce_ws_t* ws);
int feat_rj_smooth_from1_f64(feat_handle_t h, const feat_matrix_f64_t* A, const double* inv_diag,
                             const uint32_t* mask, int steps, double omega, const double* rhs,
                             const double* rhs_nrm2, double* x, double* tmp, double* d_out,
                             double* acc, const double* dotw, double* dot,
                             const feat_reduce_ws_t* ws);
int feat_pcg_update_first_f64(feat_handle_t h, double* x, double* r, const double* p, const double* q,
                              const double* gamma, const double* qp, double* rnrm2, uint64_t n,
                              const double* inv_diag, const uint32_t* mask, double omega, double* x1,
                              const feat_reduce_ws_t* ws);

/* Dense coarse solve x = filter_cor(Ainv * b) with a row-major n x n inverse. */
int feat_dense_gemv_f64(feat_handle_t h, int n, const double* ainv, const double* b, double* x,
                        const uint32_t* mask);

/* PCG updates (pcg.hpp:238-282) on device scalars:
 *   feat_pcg_update: alpha = (*gamma)/(*qp); x += alpha*p; r -= alpha*q; *rnrm2 = r.r
 *   feat_pcg_direction: beta = (*gamma_new)/(*gamma_old); p = beta*p + z (two roundings,
 *                       the axpy r == x branch); *gamma_old = *gamma_new */
int feat_pcg_update_f64(feat_handle_t h, double* x, double* r, const double* p, const double* q,
                        const double* gamma, const double* qp, double* rnrm2, uint64_t n,
                        const feat_reduce_ws_t* ws);
int feat_pcg_direction_f64(feat_handle_t h, double* p, const double* z, const double* gamma_new,
                           double* gamma_old, uint64_t n);

/* Reads a device scalar into host memory after synchronising the stream. */
int feat_read_scalar_f64(feat_handle_t h, const double* dev, double* host);

/* ---------------- mixed precision (solver/convert_precond.hpp:100-119) ----------------
 * The same fused kernels on FP32 values (row sums with fmaf in CSR order; norms and
 * dots accumulated in FP64). */
int feat_apply_masked_f32(feat_handle_t h, const feat_matrix_f32_t* A, float* r, const float* x,
                          const uint32_t* mask, double* nrm2, const feat_reduce_ws_t* ws);
int feat_defect_f32(feat_handle_t h, const feat_matrix_f32_t* A, float* d, const float* rhs,
                    const float* x, const uint32_t* mask, double* nrm2,
                    const feat_reduce_ws_t* ws);
int feat_prol_add_f32(feat_handle_t h, const feat_matrix_f32_t* P, float* sol, const float* xc,
                      const uint32_t* mask);
int feat_rj_smooth_f32(feat_handle_t h, const feat_matrix_f32_t* A, const float* inv_diag,
                       const uint32_t* mask, int steps, float omega, const float* rhs,
                       const double* rhs_nrm2, float* x, float* tmp, float* d_out, float* acc,
                       const feat_reduce_ws_t* ws);
int feat_dense_gemv_f32(feat_handle_t h, int n, const float* ainv, const float* b, float* x,
                        const uint32_t* mask);
/* ConvertPrecond's way back: z = double(z32) and *dot = z . w in FP64. */
int feat_convert_dot_f32_f64(feat_handle_t h, double* z, const float* z32, const double* w,
                             uint64_t n, double* dot, const feat_reduce_ws_t* ws);
/* FP64 coarse solve inside the FP32 cycle: x = float(filter_cor(Ainv * double(b))). */
int feat_dense_gemv_mixed(feat_handle_t h, int n, const double* ainv, const float* b, float* x,
                          const uint32_t* mask);

/* ---------------- Stokes: SaddlePointMatrix, system filter, AmaVanka ----------------
 * A Stokes vector is [u | p]: 2 values per velocity node (node major), then one per
 * pressure dof.  A feat_tuple_f64_t is a 2x2 tuple of BCSR blocks over that layout
 * (u64 col / rowptr, row-major blocks):
 *   block 0: nv x nv, 2x2 blocks   block 1: nv x np, 2x1 blocks
 *   block 2: np x nv, 1x2 blocks   block 3: np x np, 1x1 blocks (val == NULL: absent)
 * SaddlePointMatrix<A, B, D> (lafem/saddle_point_matrix.hpp:117-140) is {A, B, D, -};
 * the AmaVanka matrix (solver/amavanka.hpp:102-119, a TupleMatrix) uses all four. */
typedef struct {
  uint64_t nv, np;
  const double* val[4];
  const uint64_t* col[4];
  const uint64_t* ptr[4];
} feat_tuple_f64_t;

/* r = y + alpha*M*x (y != NULL) or r = M*x, with the reference's op order: row
 * block 0 is A.apply(r, x_u, y_u, alpha) then B.apply(r, x_p, r, alpha), row block 1
 * D.apply(r_p, x_u, y_p, alpha) then C (saddle_point_matrix.hpp:446-491,
 * tuple_matrix.hpp:305-315; each block as Arch::Apply::bcsr, apply_generic.hpp:121-154).
 * vmask (may be NULL): velocity nodes whose two rows are zeroed afterwards
 * (UnitFilterBlocked::filter_def, unit_filter_blocked.hpp:478-499).  r must not alias x. */
int feat_tuple_apply_f64(feat_handle_t h, const feat_tuple_f64_t* M, double* r, const double* x,
                         const double* y, double alpha, const uint32_t* vmask);
/* The same apply followed by what the Stokes smoother does next
 * (richardson.hpp:160-226 with SchwarzPrecond(AmaVanka), amavanka.hpp:1562-1567,
 * schwarz_precond.hpp:155-158): `passes` (0..2) MeanFilter passes on r's pressure part, as
 * feat_mean_filter_passes_f64(r + 2 nv, w_dot, w_add, volume, np, passes, sol_form), then,
 * when x_upd != NULL, x_upd += omega * r gated by *gate_norm as feat_axpy_gated_f64.  Results
 * are bit-identical to the separate calls.  With np <= 4096 all of it is one launch, the
 * filter running in the workgroup that finishes last (ws->counter: zero before the call and
 * after it); larger pressure parts take a second, one-workgroup launch for the filter and the
 * pressure rows of the update.  x_upd must not alias x, r or y. */
int feat_tuple_apply_post_f64(feat_handle_t h, const feat_tuple_f64_t* M, double* r, const double* x,
                              const double* y, double alpha, const uint32_t* vmask, const double* w_dot,
                              const double* w_add, double volume, int passes, int sol_form, double* x_upd,
                              double omega, const double* gate_norm, int x_upd_zero, const feat_reduce_ws_t* ws);
/* x_upd_zero != 0: x_upd is taken to hold zeros and is not read -- the update writes
 * omega * r, or zeros where it is skipped -- which is the smoother's first step after
 * vec_cor.format() (Richardson::apply, richardson.hpp:166) without the format. */

/* Stokes transfer and what follows it, in one launch (two above 4096 pressure rows):
 * r = [V x_u ; P x_p] (TupleDiagMatrix<BWrappedCSR<2>, CSR>::apply, control/stokes_blocked.hpp:
 * 193-202, transfer.hpp:241-259: velocity rows by the scalar CSR V on both components as
 * feat_csrsb2_apply_f64_u64, pressure rows as feat_csr_apply_f64_u64 with a = 1, b = 0),
 * velocity rows zeroed by vmask (UnitFilterBlocked::filter_def), `passes` MeanFilter passes on
 * the pressure part (w_dot, w_add, volume: filter_def or filter_cor), then, with x_upd != NULL,
 * x_upd += r (DenseVector::axpy with alpha 1, multigrid.hpp:1930).  Bit-identical to those
 * calls in a row.  x: [u (2 per node, nvx nodes) | p]; r and x_upd: [u (2 per node, V->rows
 * nodes) | p]; x, r and x_upd distinct; ws->counter as for feat_tuple_apply_post_f64. */
int feat_stokes_transfer_post_f64(feat_handle_t h, const feat_csr_f64_t* V, const feat_csr_f64_t* P,
                                  const double* x, uint64_t nvx, double* r, const uint32_t* vmask,
                                  const double* w_dot, const double* w_add, double volume, int passes,
                                  double* x_upd, const feat_reduce_ws_t* ws);

/* The coarse levels of the Stokes multigrid V-cycle in one workgroup (no counterpart in the
 * reference: a launch-count fusion of StokesMultiGrid's mg_apply over its coarsest levels,
 * multigrid.hpp:1270-1350, 1713-1980 with the smoother of richardson.hpp:160-226 and
 * SchwarzPrecond(AmaVanka)).  lv[0] is the finest level of the tail, lv[nlevels-1] the
 * coarsest (solved by one AmaVanka application + filter_cor twice); on every other level:
 * smoother (steps Richardson steps with damping omega, gated by the right-hand side's
 * norm), defect + filter_def, restriction (rv, rp) + filter_def, and on the way up
 * prolongation (pv, pp) + filter_cor, sol += cor, defect, smoother on the defect, sol +=
 * cor.  Every operation has the arithmetic of the separate entry points
 * (feat_tuple_apply_f64, feat_mean_filter_passes_f64, feat_axpy_gated_f64,
 * feat_tuple_norm2_f64, feat_csrsb2_apply_f64_u64, feat_csr_apply_f64_u64,
 * feat_unit_filter_blocked_def_f64_u64, feat_axpy_f64), so the result is bit-identical to
 * them.  rhs: level 0's right-hand side (read), sol: its solution (written); the levels'
 * other vectors live in the workgroup's LDS: sum over the levels of 2 nv + np at most
 * FEAT_STOKES_TAIL_MAX_DOFS (and of np, of the mask words at most the two limits below). */
#define FEAT_STOKES_TAIL_MAX_LEVELS 4
#define FEAT_STOKES_TAIL_MAX_DOFS 1100
#define FEAT_STOKES_TAIL_MAX_PRESSURE 384     /* sum over the levels of np */
#define FEAT_STOKES_TAIL_MAX_MASK_WORDS 32    /* sum over the levels of ceil(nv / 32) */
typedef struct {
  feat_tuple_f64_t K;             /* SaddlePointMatrix {A, B, D, -} */
  feat_tuple_f64_t V;             /* AmaVanka matrix (four blocks) */
  const uint32_t* vmask;          /* velocity nodes of the UnitFilterBlocked (may be NULL) */
  const double* prim;             /* MeanFilter vectors and volume (mean_filter.hpp:200-262) */
  const double* dual;
  double volume;
  feat_csr_f64_t rv, rp;          /* restriction to the next coarser level (velocity nodes, pressure) */
  feat_csr_f64_t pv, pp;          /* prolongation from it */
} feat_stokes_tail_level_t;
typedef struct {
  int nlevels, steps;
  double omega;
  feat_stokes_tail_level_t lv[FEAT_STOKES_TAIL_MAX_LEVELS];
} feat_stokes_tail_t;
int feat_stokes_tail_f64(feat_handle_t h, const feat_stokes_tail_t* t, const double* rhs, double* sol);

/* MeanFilter passes on one vector v of length n, in one workgroup
 * (lafem/mean_filter.hpp:213-262): per pass alpha = -(v . w_dot)/volume
 * (sol_form: 0 - (v . w_dot)/volume) and, unless |alpha| < eps (DenseVector::axpy,
 * dense_vector.hpp:815-820), v = alpha*w_add + v.  passes = 1 or 2 (two filter_cor
 * calls in a row, as AmaVanka followed by SchwarzPrecond). */
int feat_mean_filter_passes_f64(feat_handle_t h, double* v, const double* w_dot,
                                const double* w_add, double volume, uint64_t n, int passes,
                                int sol_form);

/* TupleVector::norm2 of [u | p] (tuple_vector.hpp:362-370):
 * sqrt(|u|^2 + |p|^2) with |.| = DenseVector::norm2, written to *out (device). */
int feat_tuple_norm2_f64(feat_handle_t h, const double* v, uint64_t nu, uint64_t np, double* out,
                         const feat_reduce_ws_t* ws);

/* x = omega*c + x (DenseVector::axpy, r == y branch), skipped entirely when the
 * device value *gate_norm (may be NULL) is non-finite or <= eps^2: the skip rule
 * of IterativeSolver::_set_initial_defect (iterative.hpp:761-766) for a smoother
 * whose iterate stays formatted. */
int feat_axpy_gated_f64(feat_handle_t h, double* x, double omega, const double* c, uint64_t n,
                        const double* gate_norm);

/* Velocity transfer: SparseMatrixBWrappedCSR<2> apply r = P*x on 2-blocked vectors
 * (Arch::Apply::csrsb_generic, apply_generic.hpp:157-181; a = 1, b = 0). */
int feat_csrsb2_apply_f64_u64(feat_handle_t h, double* r, const double* x, const double* val,
                              const uint64_t* col, const uint64_t* rowptr, uint64_t rows);

/* AmaVanka macro graphs (amavanka.hpp:224-350 deduct_macro_dofs and their
 * transposes), device arrays. */
typedef struct {
  uint64_t nm, nv, np, stride;
  const uint64_t* vptr; const uint64_t* vidx;   /* macro -> velocity nodes, ascending */
  const uint64_t* pptr; const uint64_t* pidx;   /* macro -> pressure dofs, ascending  */
  const uint64_t* dvptr; const uint64_t* dvidx; /* velocity node -> macros, ascending */
  const uint64_t* dpptr; const uint64_t* dpidx; /* pressure dof -> macros, ascending  */
} feat_vanka_graph_t;

/* AmaVanka::init_numeric (amavanka.hpp:1478-1559) on device for K = {A, B, D}:
 * per macro (one workgroup each) gather the local matrix of the benchmark's type-1
 * matrix (vmask rows: unit rows in A, zero rows in B; unit_filter_blocked.hpp:313-372),
 * invert it by Gauss-Jordan with diagonal pivoting (util/math.hpp:1169-1274), and with
 * skip_singular mark macros with |I - A*A^-1|_F^2 >= eps (macro_mask, nm ints); then
 * every Vanka matrix entry is the sum of the local inverses in ascending macro order
 * (the reference's sequential scatter-add) times omega / #macros of its row
 * (scale_rows, :1032-1066).  V holds the Vanka pattern; its val arrays are written.
 * local_inv: nm * stride * stride doubles of scratch.  stride <= 32. */
int feat_vanka_numeric_f64(feat_handle_t h, const feat_vanka_graph_t* g, const feat_tuple_f64_t* K,
                           const uint32_t* vmask, double omega, int skip_singular,
                           double* local_inv, int* macro_mask, const feat_tuple_f64_t* V);

/* ---------------- batches of independent solves ----------------
 * Instances stored back to back: instance i owns elements [off[i], off[i+1]).
 * Batch kernels run one workgroup per chunk of at most FEAT_BATCH_CHUNK elements
 * inside one instance (chunk_inst[c], range [chunk_lo[c], chunk_hi[c])); the chunks
 * of instance i are inst_chunk[i] .. inst_chunk[i+1]-1 in ascending order, so
 * per-instance sums are deterministic.  partials: nchunks doubles of scratch.
 * These entry points have no single reference function: they run, for every
 * instance at once, the scalar operations of PCG::_apply_intern (pcg.hpp:188-288)
 * with each instance's own scalars and stopping state. */
#define FEAT_BATCH_CHUNK 4096
typedef struct {
  uint32_t ninst, nchunks;
  const uint64_t* off;
  const uint32_t* chunk_inst;
  const uint64_t* chunk_lo;
  const uint64_t* chunk_hi;
  const uint32_t* inst_chunk;
  double* partials;
} feat_batch_t;

/* out[i] = x_i . y_i for every instance (fma accumulation) */
int feat_batch_dot_f64(feat_handle_t h, const feat_batch_t* b, const double* x, const double* y,
                       double* out);
/* per active instance (active[i] != 0): alpha = gamma[i] / tp[i];
 * x += alpha p, r -= alpha t (DenseVector::axpy r == y branch, skipped when
 * |alpha| < eps, dense_vector.hpp:815-820); rr[i] = r_i . r_i */
int feat_batch_pcg_update_f64(feat_handle_t h, const feat_batch_t* b, double* x, double* r,
                              const double* p, const double* t, const double* gamma,
                              const double* tp, const int* active, double* rr);
/* per active instance: p = z (first != 0) or p = (gamma_new[i]/gamma_old[i]) p + z
 * (pcg.hpp:276-282: p.axpy(p, z, beta), the r == x branch rounds twice) */
int feat_batch_pcg_direction_f64(feat_handle_t h, const feat_batch_t* b, double* p, const double* z,
                                 const double* gamma_new, const double* gamma_old,
                                 const int* active, int first);
/* device copy of every instance's gamma_new into gamma_old for the active ones */
int feat_batch_copy_scalars_f64(feat_handle_t h, uint32_t ninst, double* dst, const double* src,
                                const int* active);

/* Per-instance IterativeSolver state (iterative.hpp:748-885) on the device. */
typedef struct {
  double tol_rel, tol_abs, div_rel, div_abs;
  uint64_t min_iter, max_iter;
} feat_batch_stop_t;
typedef struct {
  double* def_init;   /* ninst */
  double* def_cur;    /* ninst */
  uint64_t* iters;    /* ninst */
  int* status;        /* ninst, solver/base.hpp:48-64 codes */
  int* active;        /* ninst: status == progress */
  double* hist;       /* ninst * hist_cap, or NULL */
  uint64_t hist_cap;
  int* nactive;       /* one int: number of active instances */
} feat_batch_state_t;
/* initial != 0: _set_initial_defect(sqrt(rr[i])) for every instance;
 * else _set_new_defect(sqrt(rr[i])) for the active ones */
int feat_batch_status_f64(feat_handle_t h, uint32_t ninst, const double* rr,
                          const feat_batch_stop_t* stop, const feat_batch_state_t* st, int initial);
/* The same steps with the per-instance sums left as chunk partials (nchunks doubles
 * each) and summed by the kernel that uses them, in ascending chunk order, so every
 * value equals the one of the entry points above (one launch fewer per sum):
 * dot: part[c] = chunk c's partial of x . y;
 * update: gamma = instance sum of g_part, tp = of tp_part, the r.r partials into rr_part;
 * status (not initial): rr = instance sum of rr_part; first, when gamma is not NULL, for
 * the instances active on entry gamma[i] = instance sum of g_part (the copy of
 * feat_batch_copy_scalars_f64 after the previous direction);
 * direction (not first): gamma_new = instance sum of g_part. */
int feat_batch_dot_partials_f64(feat_handle_t h, const feat_batch_t* b, const double* x, const double* y,
                                double* part);
int feat_batch_pcg_update_partials_f64(feat_handle_t h, const feat_batch_t* b, double* x, double* r,
                                       const double* p, const double* t, const double* g_part,
                                       const double* tp_part, const int* active, double* rr_part);
int feat_batch_status_partials_f64(feat_handle_t h, const feat_batch_t* b, const double* rr_part,
                                   const feat_batch_stop_t* stop, const feat_batch_state_t* st,
                                   double* gamma, const double* g_part);
int feat_batch_pcg_direction_partials_f64(feat_handle_t h, const feat_batch_t* b, double* p, const double* z,
                                          const double* g_part, const double* gamma_old, const int* active);
/* copies segments dst[dst_lo[c] + k] = src[src_lo[c] + k], k < len[c], c < nseg
 * (between the PCG layout and a multigrid level's layout) */
int feat_batch_gather_f64(feat_handle_t h, uint32_t nseg, const uint64_t* dst_lo,
                          const uint64_t* src_lo, const uint64_t* len, double* dst,
                          const double* src);
/* copies a device int to pinned host memory (stream-ordered) */
int feat_copy_int_to_host(feat_handle_t h, const int* dev, int* host_pinned);

/* The coarse levels 1..T of the batched V-cycle (solver/multigrid.hpp:1270-2000 per
 * instance, as the batched launches run it: Richardson(Jacobi) pre-smoothing with the
 * defect of the last iterate, restriction, one damped Jacobi step (omega 1) on level 1,
 * prolongation + defect + post-smoothing accumulated into the level's solution), one
 * workgroup per instance with the instance's vectors of every tail level in LDS.
 * Level k (index k-1 of lv) holds, back to back, the instances whose finest level is
 * >= k, n rows each; instance i's rows start at loff[(k-1)*ninst + i] (a multiple of n).
 * Its matrices in ELL form, entry u of local row j at [u*n + j] (R: [u*n_{k-1} + j]):
 *   A  values per instance at a_val + loff*wa, the column pattern shared;
 *   P  (rows of level k, columns of level k-1) and R = P^T: values and pattern shared
 *      by every instance (the geometric transfer of one mesh);
 * columns local to the instance's block, row lengths in *_len, entries past a row's
 * length unused (any column, any value).  An instance whose finest level f is <= T
 * enters at level f with r[pcg_off[i] + j] and leaves z[pcg_off[i] + j]; one with f > T
 * takes level T's right-hand side from rhs_deep and leaves level T's solution in
 * sol_deep (both in level T's layout).  Every row is formed with the operations of
 * feat_rj_smooth_f64, feat_apply_masked_f64, feat_prol_add_f64 and feat_defect_f64 (row
 * sums as the sequential fma over the row's entries in CSR order), so the results are
 * bit-identical to those calls on the same matrices in CSR form.  Limits: n of every
 * level <= FEAT_BATCH_TAIL_MAX_ROWS (below level T: <= 1024), the sum of n over the levels <=
 * FEAT_BATCH_TAIL_MAX_TOTAL, widths wa, wr <= FEAT_BATCH_TAIL_MAX_ROW_LEN, wp <=
 * FEAT_BATCH_TAIL_MAX_PROL_LEN.  No single reference function: the coarse end of
 * MultiGrid::apply for every instance at once. */
#define FEAT_BATCH_TAIL_MAX_LEVELS 8
#define FEAT_BATCH_TAIL_MAX_ROWS 1100
#define FEAT_BATCH_TAIL_MAX_TOTAL 1536
#define FEAT_BATCH_TAIL_MAX_ROW_LEN 9
#define FEAT_BATCH_TAIL_MAX_PROL_LEN 4
#define FEAT_BATCH_TAIL_MAX_BELOW 448   /* grid form: rows of one instance below level T, summed */
typedef struct {
  uint32_t n;                  /* rows of one instance on this level */
  uint32_t n1;                 /* grid form: nodes per side (n = n1 * n1) */
  const double* sval;          /* grid form: 9-point stencil values, as feat_batch_tiled_t */
  uint32_t wa, wp, wr;         /* ELL widths of A, P, R (P, R unused on level 1) */
  const double* a_val;
  const uint16_t* a_col;
  const uint8_t* a_len;
  const double* p_val;
  const uint16_t* p_col;
  const uint8_t* p_len;
  const double* r_val;
  const uint16_t* r_col;
  const uint8_t* r_len;
  const double* inv;           /* Jacobi: omega_jac / diag (level 1: 1 / diag), level layout */
  const uint32_t* mask;        /* Dirichlet rows (UnitFilter), bit per row of the level layout */
} feat_batch_tail_level_t;
typedef struct {
  int nlevels;                 /* T */
  uint32_t ninst;
  int nu_pre, nu_post;
  double omega;                /* Richardson damping of the smoothers */
  const int* finest;           /* per instance */
  const uint64_t* loff;        /* T * ninst */
  const uint64_t* pcg_off;     /* per instance */
  const double* rhs_deep;
  double* sol_deep;
  /* 1: grid form.  Every level is a structured n1 x n1 grid per instance whose free rows
   * hold the 9-point stencil (values in sval), P is bilinear injection from level k-1
   * and R its transpose, with the weights 1, 1/2, 1/4 in CSR order (checked by the
   * caller).  The ELL arrays are not read, every matrix is read once at kernel start,
   * and the levels below T hold at most FEAT_BATCH_TAIL_MAX_BELOW rows in all. */
  int grid;
  feat_batch_tail_level_t lv[FEAT_BATCH_TAIL_MAX_LEVELS];
} feat_batch_tail_t;
int feat_batch_tail_f64(feat_handle_t h, const feat_batch_tail_t* t, const double* r, double* z);

/* The smoother calls of one batched level on structured grids, tiled: instance p's rows
 * are p*n1*n1 .. of the level layout, lexicographic on an n1 x n1 grid; every row that the
 * unit filter (mask) leaves free is an interior node whose matrix row holds exactly the
 * nine neighbours (dy, dx) in {-1,0,1}^2 in CSR order, with values at
 * sval[(p*9 + u)*n1*n1 + row] (u = 3*(dy+1) + (dx+1); filtered rows: any values, unused).
 * One workgroup per 32x32 tile of one instance runs all sweeps of a smoother call in LDS
 * on the tile and a halo of FEAT_BATCH_TILED_MAX_STEPS nodes.
 *   pre:  feat_rj_smooth_f64(nu, omega, rhs, x = sol, d_out = def) from x = 0;
 *   post: feat_prol_add_f64(P, sol, xc), feat_defect_f64(rhs, sol) and
 *         feat_rj_smooth_f64(nu, omega, defect, acc = sol) from 0, the prolongation from
 *         level k-1 in ELL form shared by the instances (entry u of row j at [u*n1*n1 + j],
 *         columns local to the coarse block, which starts at coff[p] of xc's layout).
 * Each row is formed with the operations of those calls (row sums in CSR order), so the
 * results are bit-identical to them.  1 <= nu <= FEAT_BATCH_TILED_MAX_STEPS, wp <=
 * FEAT_BATCH_TAIL_MAX_PROL_LEN.  No single reference function: the smoother calls of
 * MultiGrid::apply (multigrid.hpp:1790-1930) for every instance of a level at once. */
#define FEAT_BATCH_TILED_MAX_STEPS 4
typedef struct {
  uint32_t n1, ninst;
  int nu;
  double omega;
  const double* sval;
  const double* inv;           /* omega_jac / diag */
  const uint32_t* mask;
  const double* rhs;
  double* sol;
  double* def;                 /* pre: the defect of the last iterate (out) */
  const double* xc;            /* post: level k-1 solution */
  const uint64_t* coff;        /* post: per instance */
  uint32_t wp;
  const double* p_val;
  const uint16_t* p_col;
  const uint8_t* p_len;
  /* may be NULL: per instance, the offset in r / z (the PCG layout) of an instance whose
   * finest level is this one (its right-hand side is read from r instead of rhs, and post
   * also writes its solution to z), UINT64_MAX for the others */
  const uint64_t* top;
  const double* r;
  double* z;
  /* 1: the prolongation is bilinear injection between the grids, row (x, y) of level k
   * taking the level k-1 nodes (x/2 .. (x+1)/2, y/2 .. (y+1)/2) in lexicographic order with
   * the weights 1, 1/2, 1/4 (checked by the caller): formed from the node position, the
   * p_* arrays are not read */
  int p_struct;
  /* pre only, 1: the restriction of the defect to level k-1 in the same launch instead of
   * def (not written): crhs[coff[p] + Y (n1+1)/2 + X] = filter(R def) with R the transpose
   * of bilinear injection (the fine nodes around (2X, 2Y) inside the grid in ascending
   * order, weights 1/2 per nonzero offset; checked by the caller), filtered by cmask, as
   * feat_apply_masked_f64 forms it */
  int rest;
  double* crhs;
  const uint32_t* cmask;
} feat_batch_tiled_t;
int feat_batch_pre_tiled_f64(feat_handle_t h, const feat_batch_tiled_t* t);
int feat_batch_post_tiled_f64(feat_handle_t h, const feat_batch_tiled_t* t);

/* ---------------- structured Q1 hierarchy on the device ----------------
 * The Q1 Laplace problem on level `level` of the unit square (dim 2) / cube (dim 3),
 * lexicographic node numbering: stiffness matrix (SymbolicAssembler + LaplaceOperator),
 * Dirichlet boundary list (UnitFilterAssembler), prolongation from level-1 and its
 * transpose, the restriction (GridTransfer::assemble_prolongation and transpose,
 * control/scalar_basic.hpp:282-331).  Bit-identical to the host assembly of
 * include/feat3amd/assembly.hpp.  sizes: [rows, stiffness nnz, boundary nodes,
 * prolongation nnz, coarse rows]; the restriction has the prolongation's nnz.
 * No counterpart in the reference (its assembly runs on the host). */
int feat_q1_device_sizes(int dim, int level, uint64_t* sizes);
/* Any of bnd / the prolongation arrays / the restriction arrays may be NULL (skipped). */
int feat_q1_assemble_device(feat_handle_t h, int dim, int level, double* val, uint64_t* col,
                            uint64_t* rowptr, uint64_t* bnd, double* pval, uint64_t* pcol,
                            uint64_t* prowptr, double* rval, uint64_t* rcol, uint64_t* rrowptr);

/* ---------------- Q2/P1dc Stokes level on the device ----------------
 * The lid-driven-cavity level of benchmarks/stokes_multigrid-bench.cpp:576-697 on the unit
 * square (2^level cells per axis, lexicographic velocity nodes, P1dc dofs 3e..3e+2): velocity
 * block A = DuDvOperatorBlocked<2> (2x2 blocks, assembly/common_operators.hpp:875-961),
 * B = -(div) (2x1, control/stokes_blocked.hpp:784-798), D = B^T (1x2), UnitFilterBlocked
 * velocity boundary (lid u = (1, 0)), MeanFilter vectors (mean_filter_assembler.hpp:67-85),
 * Q2 and P1dc prolongations from level-1 and their transposes (GridTransfer,
 * control/stokes_blocked.hpp).  Bit-identical to the host assembly of
 * include/feat3amd/stokes_assembly.hpp (and the host transpose for the restrictions).
 * sizes: {nv, np, nnz A, nnz B, nnz D, nfilt, nnz Pv (= nnz Rv), nnz Pp (= nnz Rp),
 * nv coarse, np coarse}.  Every group of three pointers (filt/filt_val and prim/dual as
 * pairs) is either complete or all NULL (skipped); level 0 has no transfer.
 * No counterpart in the reference (its assembly runs on the host). */
typedef struct {
  double* a_val; uint64_t* a_col; uint64_t* a_ptr;
  double* b_val; uint64_t* b_col; uint64_t* b_ptr;
  double* d_val; uint64_t* d_col; uint64_t* d_ptr;
  uint64_t* filt; double* filt_val;
  double* prim; double* dual;
  double* pv_val; uint64_t* pv_col; uint64_t* pv_ptr;
  double* pp_val; uint64_t* pp_col; uint64_t* pp_ptr;
  double* rv_val; uint64_t* rv_col; uint64_t* rv_ptr;
  double* rp_val; uint64_t* rp_col; uint64_t* rp_ptr;
} feat_stokes_level_dev_t;
int feat_stokes_device_sizes(int level, uint64_t* sizes);
int feat_stokes_assemble_device(feat_handle_t h, int level, const feat_stokes_level_dev_t* out);

/* ---------------- SOR / SSOR sweeps (CSR, f64) ----------------
 * SORPrecond / SSORPrecond (solver/sor_precond.hpp, solver/ssor_precond.hpp) with the rows of
 * one "colour" updated in parallel, one launch per colour:
 *   feat_sor_apply:           y_i = omega (x_i - sum_{j<i} a_ij y_j) / a_ii
 *   feat_ssor_forward_apply:  y_i = (x_i - omega sum_{j<i} a_ij y_j) / a_ii
 *   feat_ssor_backward_apply: y_i -= omega (sum_{j>i} a_ij y_j) / a_ii
 * (SSORPrecond::apply then scales by omega (2 - omega) and filters, ssor_precond.hpp:177-183).
 * Default ordering: the colours are the level sets of the lower (forward) and upper
 * (backward) triangle, so the result is exactly the generic sequential sweep
 * (sor_precond.hpp:201-222, ssor_precond.hpp:193-231); a 27-point 3D stencil in lexicographic
 * order has 7N+1 levels per sweep.  FEAT_SOR_MULTICOLOR: the CUDA backend's scheme
 * (sor_precond.hpp:21-28 Intern::cuda_sor_init_symbolic / cuda_sor_apply, ssor_precond.hpp:
 * 20-24 cuda_ssor_forward_apply / cuda_ssor_backward_apply): a proper colouring (greedy first
 * fit in row order on the symmetrised pattern; the reference takes cusparseDcsrcolor's),
 * colours in ascending order of their row count (ties by colour number), y cleared first,
 * backward sweep in the forward colour order -- 8 colours for a 27-point stencil, but the
 * resulting SSOR is not symmetric (as the reference's CUDA SSOR).
 * A: device CSR with sorted rows and a main diagonal entry in every row; the plan keeps a
 * copy of the values: call feat_sor_init_numeric after they change.  y must not alias x. */
typedef struct feat_sor_s* feat_sor_t;
#define FEAT_SOR_MULTICOLOR 1
int feat_sor_init_symbolic(feat_handle_t h, const feat_csr_f64_t* A, int flags, feat_sor_t* out);
int feat_sor_init_numeric(feat_handle_t h, feat_sor_t s, const feat_csr_f64_t* A);
/* the forward (backward = 0) or backward partition: number of colours, the processing rank of
 * each row's colour (n entries), rows per colour in processing order; host arrays, may be NULL */
int feat_sor_colors(feat_sor_t s, int backward, int* ncolors, uint32_t* rank_of_row, uint64_t* rows_per_color);
int feat_sor_apply(feat_handle_t h, feat_sor_t s, double* y, const double* x, double omega);
int feat_ssor_forward_apply(feat_handle_t h, feat_sor_t s, double* y, const double* x, double omega);
int feat_ssor_backward_apply(feat_handle_t h, feat_sor_t s, double* y, double omega);
int feat_sor_done_symbolic(feat_sor_t s);

/* ---------------- ILU(0) (CSR, f64) ----------------
 * ILUPrecond, generic backend with p = 0 (solver/ilu_precond.hpp: ILUCoreSymbolic::
 * set_struct_csr :104-147, copy_data_csr :404-432, factorize_numeric_il_du :453-527,
 * solve_il / solve_du :540-582; the CUDA backend's cusparse csrilu02 computes the same
 * factorisation): (I+L)(D+U) ~ A on A's pattern, D stored inverted.  Factorisation and
 * forward solve run over the level sets of L, the backward solve over those of U, one launch
 * per level, so every row keeps the sequential operation order: bit-identical to the generic
 * backend.  init_numeric copies the values of A and factorises; apply computes
 * y = (D+U)^{-1} (I+L)^{-1} x (y may alias x).  A: device CSR, sorted rows, every row with
 * its diagonal entry.
 * feat_ilu_init_symbolic_p: ILU(p), the level-p pattern of ILUCoreSymbolic::factorize_symbolic
 * (ilu_precond.hpp:191-262, the same linear insertion), values of A copied onto it with zeros
 * at fill-in (copy_data_csr), the same level-scheduled factorisation and solves on it;
 * p = 0 is feat_ilu_init_symbolic. */
typedef struct feat_ilu_s* feat_ilu_t;
int feat_ilu_init_symbolic(feat_handle_t h, const feat_csr_f64_t* A, feat_ilu_t* out);
int feat_ilu_init_symbolic_p(feat_handle_t h, const feat_csr_f64_t* A, int p, feat_ilu_t* out);
int feat_ilu_init_numeric(feat_handle_t h, feat_ilu_t s, const feat_csr_f64_t* A);
int feat_ilu_apply(feat_handle_t h, feat_ilu_t s, double* y, const double* x);
/* entries of L and U, number of level sets of L and of U */
int feat_ilu_info(feat_ilu_t s, uint64_t* nnz_l, uint64_t* nnz_u, int* levels_l, int* levels_u);
/* copies the factors to host arrays (L and U in the order of the pattern's rows: A's lower /
 * upper entries for ILU(0), the level-p pattern's for ILU(p); the inverted diagonal); any
 * pointer may be NULL */
int feat_ilu_factors(feat_handle_t h, feat_ilu_t s, double* l, double* u, double* d);
int feat_ilu_done_symbolic(feat_ilu_t s);

/* ---------------- fused coarse levels of the V-cycle ----------------
 * One workgroup (1024 threads) runs the coarse part of MultiGrid::apply (multigrid.hpp:
 * 1713-2000) from level t down to the coarsest and back, with every vector of those levels
 * in LDS: per smoothed level the fused Richardson(Jacobi) pre-smoothing with its final
 * defect (feat_rj_smooth_f64 semantics), the masked restriction with |rhs|^2, then on the
 * coarsest level the dense inverse (feat_dense_gemv_f64), and back up the prolongation-add,
 * the defect with |d|^2 and the post-smoothing accumulated into the solution.  Every row
 * value is computed with the same operations in the same order as the separate launches
 * (matrix sums in CSR / slot order), so the solution is bit-identical; the |.|^2 sums
 * feeding the smoothers' skip rule (exact zero or non-finite) are summed in another order.
 * lv[0 .. nlevels-2]: smoothed levels (A as a FEAT_IDX_DIAG plan, restriction to and
 * prolongation from the next level as CSR); lv[nlevels-1]: the coarsest level (n, mask).
 * rhs / rhs_nrm2: level t's right-hand side and |rhs|^2 (device); sol: level t's solution
 * (written).  The caller keeps 8 bytes x (5 n + 256 over the smoothed levels + 2 n coarse)
 * within FEAT_MG_TAIL_MAX_LDS bytes (when the coarse inverse, 8 n_c^2 bytes, also fits, it
 * is staged in LDS).  No counterpart in the reference (a launch-count fusion). */
#define FEAT_MG_TAIL_MAX_LEVELS 6
#define FEAT_MG_TAIL_MAX_ROWS 1024   /* rows per level (one thread per row) */
#define FEAT_MG_TAIL_MAX_LDS 61440
typedef struct {
  uint64_t n;
  const double* table; const uint32_t* codes; uint32_t words;   /* FEAT_IDX_DIAG plan of A */
  const double* inv_diag; const uint32_t* mask;
  const double* r_val; const uint64_t* r_col; const uint64_t* r_ptr;   /* restriction to the next level */
  const double* p_val; const uint64_t* p_col; const uint64_t* p_ptr;   /* prolongation from the next level */
  int pre_steps, post_steps;
  double pre_omega, post_omega;
  /* grid form (0 / NULL: unused).  dim (2 or 3) and n1: the level is an n1^dim grid in
   * lexicographic order.  sten: the 3^dim values that every free row holds at the offsets
   * (dz, dy, dx) in {-1,0,1}^dim in lexicographic order, read instead of the plan; rows the
   * mask filters skip their sum, which they never use.  dlinear: the prolongation from the
   * next level ((n1 + 1) / 2 per side) is d-linear injection (weights 1, 1/2, ... in
   * ascending columns) and the restriction its transpose, formed from the node positions
   * instead of read.  All of it checked by the caller against the matrices. */
  int dim;
  uint32_t n1;
  const double* sten;
  int dlinear;
} feat_mg_tail_level_t;
typedef struct {
  int nlevels;
  feat_mg_tail_level_t lv[FEAT_MG_TAIL_MAX_LEVELS];
  const double* coarse_inverse;   /* n x n row-major inverse of the filtered coarsest matrix */
} feat_mg_tail_t;
int feat_mg_tail_f64(feat_handle_t h, const feat_mg_tail_t* t, const double* rhs, const double* rhs_nrm2, double* sol);
/* The same for the FP32 V-cycle of the mixed-precision multigrid (feat_rj_smooth_f32,
 * feat_apply_masked_f32, feat_dense_gemv_mixed, feat_prol_add_f32, feat_defect_f32
 * semantics): FP32 pair tables (the FP32 FEAT_IDX_DIAG plans), FP32 inverted diagonals,
 * omegas and vectors; the transfers' FP64 CSR values are rounded to FP32 as the FP32
 * plans round them; the coarse inverse is applied in FP64 and rounded once.  The caller
 * keeps 4 bytes x (5 n + 256 over the smoothed levels + 2 n coarse) within
 * FEAT_MG_TAIL_MAX_LDS bytes. */
typedef struct {
  uint64_t n;
  const float* table; const uint32_t* codes; uint32_t words;   /* FP32 FEAT_IDX_DIAG plan of A */
  const float* inv_diag; const uint32_t* mask;
  const double* r_val; const uint64_t* r_col; const uint64_t* r_ptr;
  const double* p_val; const uint64_t* p_col; const uint64_t* p_ptr;
  int pre_steps, post_steps;
  float pre_omega, post_omega;
  int dim;                     /* grid form, as feat_mg_tail_level_t (sten: FP32 values) */
  uint32_t n1;
  const float* sten;
  int dlinear;
} feat_mg_tail_level_f32_t;
typedef struct {
  int nlevels;
  feat_mg_tail_level_f32_t lv[FEAT_MG_TAIL_MAX_LEVELS];
  const double* coarse_inverse;
} feat_mg_tail_f32_t;
int feat_mg_tail_f32(feat_handle_t h, const feat_mg_tail_f32_t* t, const float* rhs, const double* rhs_nrm2, float* sol);

/* ---------------- autotuner ----------------
 * Launch shapes of the SELL-64 row kernels that every fused entry point uses
 * (entries per row in flight: 4, 8 or 16 -- pair-dictionary plans: 4, 8 or 16
 * codes; workgroup cap: all row chunks, or 8192/4096/2048/1024/512 workgroups
 * striding over them; launches that also compute a dot or norm keep the full
 * grid).  Every shape gives bit-identical results (each row is summed in CSR
 * order, reductions do not depend on the shape).  The tuner times
 * r = A*x for every shape with HIP events (reps launches each) and stores the
 * fastest in A->tune; *best_us (may be NULL) receives its time per launch.
 * Diagonal dictionary plans (FEAT_IDX_DIAG) have one shape: tune stays 0. */
int feat_autotune_matrix_f64(feat_handle_t h, feat_matrix_f64_t* A, const double* x, double* r,
                             int reps, double* best_us);

/* ---------------- learned smoother network ----------------
 * No counterpart in the reference.  MLP 54 -> 64 -> 64 -> 1, GELU, on BF16 matrix
 * cores (FP32 accumulation).  Parameters (FP32 master, FEAT_MLP_NPARAM floats):
 *   W1[64][64] (row = hidden unit, inputs 54..63 are padding and stay 0) at 0,
 *   b1[64] at 4096, W2[64][64] at 4160, b2[64] at 8256, w3[64] at 8320, b3 at 8384.
 * wbf: BF16 copies of W1 then W2 (2 * 4096 values).  Network:
 *   z1 = W1 x + b1, h1 = bf16(gelu(z1)), z2 = W2 h1 + b2, y = w3 . gelu(z2) + b3,
 * x a row of 64 BF16 inputs.  Training loss: mean of (y - t)^2. */
#define FEAT_MLP_IN 54
#define FEAT_MLP_HID 64
#define FEAT_MLP_NPARAM 8385
#define FEAT_MLP_NPARAM_PAD 8448
int feat_mlp_to_bf16(feat_handle_t h, const float* params, uint16_t* wbf);
int feat_mlp_infer(feat_handle_t h, const float* params, const uint16_t* wbf, const uint16_t* X,
                   uint64_t nsamp, float* y);
/* grad (FEAT_MLP_NPARAM_PAD floats; grad[FEAT_MLP_NPARAM_PAD-1] = sum of squared
 * errors) of the mean squared error over nsamp samples; partials: scratch of
 * partials_cap * FEAT_MLP_NPARAM_PAD floats, partials_cap >= 4 w + ceil(4 w / 32) with
 * w = min(256, ceil(nsamp / 64)) (1056 always suffices) */
int feat_mlp_grad(feat_handle_t h, const float* params, const uint16_t* wbf, const uint16_t* X,
                  const float* t, uint64_t nsamp, float* partials, uint64_t partials_cap, float* grad);
/* one Adam step (bias-corrected, step >= 1); refreshes wbf */
int feat_mlp_adam(feat_handle_t h, float* params, uint16_t* wbf, const float* grad, float* m, float* v,
                  int step, float lr, float beta1, float beta2, float eps);
/* Inputs of the learned smoother on a structured m^3 node grid (lexicographic, x
 * fastest): for interior node i, rho_j = r_j / a_ii over its 3x3x3 patch (j = dz,dy,dx
 * lexicographic), s = max |rho_j|; x = [rho / s (27), cfeat over the patch (27), 0 (10)].
 * Training samples: node[k] of snapshot snap[k] of R and E (snapshots of n = m^3
 * doubles); target t = e_i / s.  Rows of boundary nodes or s == 0 are zero (t = 0). */
int feat_mlp_features(feat_handle_t h, uint64_t m, const double* R, const double* E,
                      const double* dinv, const float* cfeat, const uint64_t* node,
                      const uint32_t* snap, uint64_t nsamp, uint16_t* X, float* t);
/* learned smoothing step on every interior node: x_i += omega * (jacobi * r_i / a_ii
 * + s_i * y_i), y the network output for the patch of the defect r (dinv = 1/diag(A)):
 * damped Jacobi plus the learned correction */
int feat_mlp_smooth_f64(feat_handle_t h, uint64_t m, const float* params, const uint16_t* wbf,
                        const double* r, const double* dinv, const float* cfeat, double omega,
                        double jacobi, double* x);
/* the same step's direction: c_i = omega * (jacobi * r_i / a_ii + s_i * y_i) on every
 * interior node with s_i > 0, c_i = 0 elsewhere, and *dot = c . r (device scalar,
 * deterministic reduction through ws) */
int feat_mlp_direction_f64(feat_handle_t h, uint64_t m, const float* params, const uint16_t* wbf,
                           const double* r, const double* dinv, const float* cfeat, double omega,
                           double jacobi, double* c, double* dot, const feat_reduce_ws_t* ws);
/* inputs of every node (whole-field training): X (m^3 rows), s, rho_i = r_i / a_ii */
int feat_mlp_features_full(feat_handle_t h, uint64_t m, const double* r, const double* dinv,
                           const float* cfeat, uint16_t* X, double* s, double* rho);
/* energy-optimal step of the learned smoother along c (q = A c, *a = c . d, *b = c . q on
 * the device): alpha = *a / *b (0 unless *b > 0 and the ratio is finite), x += alpha c,
 * d -= alpha q -- the error energy never grows */
int feat_mlp_step_f64(feat_handle_t h, uint64_t n, double* x, double* d, const double* c,
                      const double* q, const double* a, const double* b);
/* energy-norm training of the learned correction on a field pair (r = A e, e):
 * stage 0 (coef = the Jacobi weight): ep = e - (coef rho + s y) (the error after one
 * learned step, 0 where s = 0; e may be NULL for e = 0);
 * stage 1 (coef = a loss weight w), given q = A ep: t = y + w s q / 2, so that the
 * squared-error gradient of feat_mlp_grad over all n nodes is the gradient of
 * w ep^T A ep / (2 n) */
int feat_mlp_energy(feat_handle_t h, uint64_t n, int stage, double coef, const float* y,
                    const double* s, const double* rho, const double* e, const double* q,
                    double* ep, float* t);

#ifdef __cplusplus
}
#endif
#endif
