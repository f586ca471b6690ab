// feat3amd_assembly.hip -- the structured Q1 hierarchy assembled on the device.
//
// Stiffness matrix, prolongation, restriction (its transpose) and Dirichlet
// boundary list of the Q1 Laplace problem on a 2^L (x 2^L (x 2^L)) grid with
// lexicographic node numbering: the same entries, in the same order and with the
// same floating-point operations as the host assembly, hence bit-identical arrays.
// What the reference assembles with SymbolicAssembler + LaplaceOperator,
// GridTransfer::assemble_prolongation and the transposed restriction
// (control/scalar_basic.hpp:282-331), UnitFilterAssembler.
//
// On a tensor grid every row pointer has a closed form (a product of per-axis
// prefix sums), so a thread computes its row without a scan.  The kernels are
// write-only (16 B per entry); a workgroup builds kStageRows consecutive rows in
// LDS and then stores their contiguous entry range with consecutive lanes on
// consecutive entries, instead of each lane writing its own row (a 27-entry stride
// between the lanes of one store).
#include "feat3amd_common.hpp"

namespace {

constexpr int kStageRows = 128;

struct Q1Grid
{
  int dim;
  uint64_t m, N;     // nodes per axis, cells per axis
};

__device__ __forceinline__ void coords(int dim, uint64_t m, uint64_t r, uint64_t c[3])
{
  c[0] = r % m;
  c[1] = (r / m) % m;
  c[2] = dim == 3 ? r / (m * m) : 0;
}

// start of row r from per-axis lengths len(c) and prefix sums pre(c) (T = pre of the whole axis)
template<class Axis>
__device__ __forceinline__ uint64_t tensor_start(int dim, uint64_t m, const Axis& ax, uint64_t r)
{
  uint64_t c[3];
  coords(dim, m, r, c);
  const uint64_t T = ax.total();
  return dim == 3 ? T * T * ax.pre(c[2]) + ax.len(c[2]) * (T * ax.pre(c[1]) + ax.len(c[1]) * ax.pre(c[0]))
                  : T * ax.pre(c[1]) + ax.len(c[1]) * ax.pre(c[0]);
}

// ---- stiffness matrix: per axis, a node has (c > 0) + 1 + (c < N) neighbours
struct LapAxis
{
  uint64_t N;
  __device__ uint64_t len(uint64_t c) const { return (c > 0) + 1 + (c < N); }
  __device__ uint64_t pre(uint64_t c) const { return c ? 3 * c - 1 : 0; }   // c <= N
  __device__ uint64_t total() const { return 3 * N + 1; }
};

struct LaplaceRows
{
  Q1Grid g;
  uint64_t n, nnz;
  double ev[4];   // entry value per number of differing coordinates, for one shared element

  __device__ uint64_t start(uint64_t r) const { return r >= n ? nnz : tensor_start(g.dim, g.m, LapAxis{g.N}, r); }
  template<class Emit>
  __device__ void emit(uint64_t r, Emit&& out) const
  {
    uint64_t c[3];
    coords(g.dim, g.m, r, c);
    const int64_t c0 = int64_t(c[0]), c1 = int64_t(c[1]), c2 = int64_t(c[2]), N = int64_t(g.N);
    const uint64_t sz = g.dim == 3 ? g.m : 1;
    const int64_t lo2 = g.dim == 3 ? -1 : 0, hi2 = g.dim == 3 ? 1 : 0;
    for(int64_t d2 = lo2; d2 <= hi2; ++d2)
      for(int64_t d1 = -1; d1 <= 1; ++d1)
        for(int64_t d0 = -1; d0 <= 1; ++d0)
        {
          const int64_t q0 = c0 + d0, q1 = c1 + d1, q2 = c2 + d2;
          if(q0 < 0 || q1 < 0 || q2 < 0 || q0 > N || q1 > N || (g.dim == 3 && q2 > N)) continue;
          // elements containing both nodes: a factor 2 per equal interior coordinate
          uint64_t ne = 1;
          int ndiff = 0;
          const int64_t cc[3] = {c0, c1, c2}, dd[3] = {d0, d1, d2};
          for(int d = 0; d < g.dim; ++d)
          {
            if(dd[d] != 0)
              ++ndiff;
            else
              ne *= (cc[d] > 0 && cc[d] < N) ? 2 : 1;
          }
          out(uint64_t(q0) + g.m * (uint64_t(q1) + sz * uint64_t(q2)), double(ne) * ev[ndiff]);
        }
  }
};

// ---- prolongation from level L-1: per axis an odd fine coordinate has 2 parents
struct ProAxis
{
  uint64_t m;
  __device__ uint64_t len(uint64_t c) const { return (c & 1) ? 2 : 1; }
  __device__ uint64_t pre(uint64_t c) const { return c + c / 2; }
  __device__ uint64_t total() const { return m + m / 2; }
};

struct ProlongationRows
{
  Q1Grid gf;
  uint64_t mc, n, nnz;

  __device__ uint64_t start(uint64_t r) const { return r >= n ? nnz : tensor_start(gf.dim, gf.m, ProAxis{gf.m}, r); }
  template<class Emit>
  __device__ void emit(uint64_t r, Emit&& out) const
  {
    uint64_t c[3];
    coords(gf.dim, gf.m, r, c);
    uint64_t ci[3][2] = {};
    double w[3][2] = {};
    int cnt[3] = {1, 1, 1};
    for(int d = 0; d < 3; ++d)
    {
      if(d >= gf.dim)
      {
        ci[d][0] = 0;
        w[d][0] = 1.0;
        continue;
      }
      if(c[d] & 1)
      {
        cnt[d] = 2;
        ci[d][0] = (c[d] - 1) / 2;
        ci[d][1] = (c[d] + 1) / 2;
        w[d][0] = w[d][1] = 0.5;
      }
      else
      {
        ci[d][0] = c[d] / 2;
        w[d][0] = 1.0;
      }
    }
    for(int a2 = 0; a2 < cnt[2]; ++a2)
      for(int a1 = 0; a1 < cnt[1]; ++a1)
        for(int a0 = 0; a0 < cnt[0]; ++a0)
          out(ci[0][a0] + mc * (ci[1][a1] + (gf.dim == 3 ? mc * ci[2][a2] : 0)), w[0][a0] * w[1][a1] * w[2][a2]);
  }
};

// ---- restriction = prolongation^T: coarse node c is a parent of the fine nodes
// 2c-1, 2c, 2c+1 (within the grid) of each axis, with weights 1/2, 1, 1/2
struct ResAxis
{
  uint64_t mc;
  __device__ uint64_t len(uint64_t c) const { return 3 - (c == 0) - (c + 1 == mc); }
  __device__ uint64_t pre(uint64_t c) const { return c ? 3 * c - 1 : 0; }   // c < mc
  __device__ uint64_t total() const { return 3 * mc - 2; }
};

struct RestrictionRows
{
  Q1Grid gc;
  uint64_t mf, n, nnz;

  __device__ uint64_t start(uint64_t r) const { return r >= n ? nnz : tensor_start(gc.dim, gc.m, ResAxis{gc.m}, r); }
  template<class Emit>
  __device__ void emit(uint64_t r, Emit&& out) const
  {
    const uint64_t mc = gc.m;
    uint64_t c[3];
    coords(gc.dim, mc, r, c);
    // per axis: fine coordinates (ascending) and the weights the prolongation gives them
    uint64_t fi[3][3] = {};
    double w[3][3] = {};
    int cnt[3] = {1, 1, 1};
    for(int d = 0; d < 3; ++d)
    {
      if(d >= gc.dim)
      {
        fi[d][0] = 0;
        w[d][0] = 1.0;
        continue;
      }
      int q = 0;
      if(c[d] > 0)
      {
        fi[d][q] = 2 * c[d] - 1;
        w[d][q++] = 0.5;
      }
      fi[d][q] = 2 * c[d];
      w[d][q++] = 1.0;
      if(c[d] + 1 < mc)
      {
        fi[d][q] = 2 * c[d] + 1;
        w[d][q++] = 0.5;
      }
      cnt[d] = q;
    }
    // the product in the prolongation's order: axis 0, then 1, then 2
    for(int a2 = 0; a2 < cnt[2]; ++a2)
      for(int a1 = 0; a1 < cnt[1]; ++a1)
        for(int a0 = 0; a0 < cnt[0]; ++a0)
          out(fi[0][a0] + mf * (fi[1][a1] + (gc.dim == 3 ? mf * fi[2][a2] : 0)), w[0][a0] * w[1][a1] * w[2][a2]);
  }
};

// One workgroup of kStageRows threads per chunk of kStageRows rows: thread t emits row
// r0 + t into LDS at its offset inside the chunk, then the chunk's entries (contiguous in
// val / col) are stored by consecutive lanes.  MAXLEN bounds the entries of one row.
template<int MAXLEN, class Rows>
__global__ __launch_bounds__(kStageRows) void k_assemble_rows(Rows rows, double* __restrict__ val,
                                                              uint64_t* __restrict__ col,
                                                              uint64_t* __restrict__ rowptr)
{
  __shared__ double sv[kStageRows * MAXLEN];
  __shared__ uint64_t sc[kStageRows * MAXLEN];
  const uint64_t n = rows.n;
  for(uint64_t r0 = uint64_t(blockIdx.x) * kStageRows; r0 < n; r0 += uint64_t(gridDim.x) * kStageRows)
  {
    const uint64_t r = r0 + threadIdx.x;
    const uint64_t k0 = rows.start(r0);
    const uint64_t k1 = rows.start(r0 + kStageRows < n ? r0 + kStageRows : n);
    if(r < n)
    {
      uint64_t k = rows.start(r);
      rowptr[r] = k;
      if(r + 1 == n) rowptr[n] = k1;
      k -= k0;
      rows.emit(r, [&](uint64_t c, double v) {
        sv[k] = v;
        sc[k] = c;
        ++k;
      });
    }
    __syncthreads();
    for(uint64_t i = threadIdx.x; i < k1 - k0; i += kStageRows)
    {
      val[k0 + i] = sv[i];
      col[k0 + i] = sc[i];
    }
    __syncthreads();
  }
}

// ---- boundary nodes in ascending order: entry j of the list
__global__ void k_q1_boundary(Q1Grid g, uint64_t nb, uint64_t* __restrict__ idx)
{
  const uint64_t m = g.m, N = g.N;
  const uint64_t ring = 4 * m - 4;   // boundary nodes of one inner z-plane (3D)
  for(uint64_t j = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; j < nb; j += uint64_t(gridDim.x) * blockDim.x)
  {
    // position inside one plane: row y = 0 (m nodes), rows 1..N-1 (x = 0, x = N), row N (m nodes)
    auto in_plane = [&](uint64_t t, uint64_t& x, uint64_t& y) {
      if(t < m)
      {
        y = 0;
        x = t;
      }
      else if(t < m + 2 * (N - 1))
      {
        const uint64_t u = t - m;
        y = 1 + u / 2;
        x = (u & 1) ? N : 0;
      }
      else
      {
        y = N;
        x = t - m - 2 * (N - 1);
      }
    };
    uint64_t x = 0, y = 0, z = 0;
    if(g.dim == 2)
      in_plane(j, x, y);
    else if(j < m * m)
    {
      y = j / m;
      x = j % m;
    }
    else if(j < m * m + (N - 1) * ring)
    {
      const uint64_t u = j - m * m;
      z = 1 + u / ring;
      in_plane(u % ring, x, y);
    }
    else
    {
      const uint64_t u = j - m * m - (N - 1) * ring;
      z = N;
      y = u / m;
      x = u % m;
    }
    idx[j] = x + m * (y + (g.dim == 3 ? m * z : 0));
  }
}

Q1Grid grid_of(int dim, int level)
{
  const uint64_t m = (uint64_t(1) << level) + 1;
  return Q1Grid{dim, m, m - 1};
}

uint64_t ipow_u(uint64_t b, int e)
{
  uint64_t r = 1;
  for(int i = 0; i < e; ++i) r *= b;
  return r;
}

int grid_rows(uint64_t n)
{
  const uint64_t b = (n + kStageRows - 1) / kStageRows;
  return int(b < 65536 ? (b ? b : 1) : 65536);
}

} // namespace

extern "C" {

int feat_q1_device_sizes(int dim, int level, uint64_t* sizes)
{
  if((dim != 2 && dim != 3) || level < 0 || level > 12 || !sizes) return fail(FEAT_ERR_INVALID, "feat_q1_device_sizes: bad argument");
  const Q1Grid g = grid_of(dim, level);
  sizes[0] = ipow_u(g.m, dim);                                                // rows
  sizes[1] = ipow_u(3 * g.N + 1, dim);                                        // stiffness nonzeros
  sizes[2] = sizes[0] - ipow_u(g.N - 1, dim);                                 // boundary nodes (N >= 1)
  sizes[3] = level > 0 ? ipow_u(g.m + g.m / 2, dim) : 0;                      // prolongation nonzeros
  sizes[4] = level > 0 ? ipow_u((uint64_t(1) << (level - 1)) + 1, dim) : 0;   // coarse rows
  return FEAT_OK;
}

int feat_q1_assemble_device(feat_handle_t h, int dim, int level, double* val, uint64_t* col, uint64_t* rowptr,
                            uint64_t* bnd, double* pval, uint64_t* pcol, uint64_t* prowptr, double* rval, uint64_t* rcol,
                            uint64_t* rrowptr)
{
  uint64_t sz[5];
  int rc = feat_q1_device_sizes(dim, level, sz);
  if(rc) return rc;
  if(!h || !val || !col || !rowptr) return fail(FEAT_ERR_INVALID, "feat_q1_assemble_device: null stiffness arrays");
  const bool pro = pval || pcol || prowptr, res = rval || rcol || rrowptr;
  if((pro && !(pval && pcol && prowptr)) || (res && !(rval && rcol && rrowptr)))
    return fail(FEAT_ERR_INVALID, "feat_q1_assemble_device: incomplete transfer arrays");
  if(level == 0 && (pro || res)) return fail(FEAT_ERR_INVALID, "feat_q1_assemble_device: level 0 has no transfer");
  const Q1Grid g = grid_of(dim, level);
  LaplaceRows lap{g, sz[0], sz[1], {}};
  if(dim == 2)
  {
    lap.ev[0] = 4.0 / 6.0;
    lap.ev[1] = -1.0 / 6.0;
    lap.ev[2] = -2.0 / 6.0;
    lap.ev[3] = 0.0;
  }
  else
  {
    const double hh = 1.0 / double(g.N);
    lap.ev[0] = hh / 3.0;
    lap.ev[1] = 0.0;
    lap.ev[2] = -hh / 12.0;
    lap.ev[3] = -hh / 12.0;
  }
  hipLaunchKernelGGL((k_assemble_rows<27, LaplaceRows>), dim3(grid_rows(sz[0])), dim3(kStageRows), 0, h->stream, lap,
                     val, col, rowptr);
  if(bnd) hipLaunchKernelGGL(k_q1_boundary, dim3(grid_for(sz[2])), dim3(kThreads), 0, h->stream, g, sz[2], bnd);
  if(pro)
    hipLaunchKernelGGL((k_assemble_rows<8, ProlongationRows>), dim3(grid_rows(sz[0])), dim3(kStageRows), 0, h->stream,
                       ProlongationRows{g, grid_of(dim, level - 1).m, sz[0], sz[3]}, pval, pcol, prowptr);
  if(res)
    hipLaunchKernelGGL((k_assemble_rows<27, RestrictionRows>), dim3(grid_rows(sz[4])), dim3(kStageRows), 0, h->stream,
                       RestrictionRows{grid_of(dim, level - 1), g.m, sz[4], sz[3]}, rval, rcol, rrowptr);
  return check_launch("feat_q1_assemble_device");
}

} // extern "C"
