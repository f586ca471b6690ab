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

#include <feat3amd/stokes_assembly.hpp>

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


// ===================== Q2/P1dc Stokes level (unit square, lid-driven cavity) =====================
// The arrays of the host assembly (include/feat3amd/stokes_assembly.hpp, what
// benchmarks/stokes_multigrid-bench.cpp:576-697 assembles): A = DuDv (2x2 blocks), B = -grad
// (2x1), D = B^T (1x2), velocity boundary list and values, pressure mean-filter vectors, Q2 and
// P1dc prolongations from level-1 and their transposes.  One thread per (row, slot): lanes of a
// wave write consecutive entries of consecutive rows.  The element tables are formed on the host
// by the host assembly's own code; the device only sums them in element order or copies them.

struct KeTable { double v[9][9][4]; };   // element block (test I, trial J), row-major 2x2
struct FgTable { double v[9][3][2]; };   // B element entries before the -(h/2) scale

struct StokesGrid { uint64_t N, m; double h; };   // cells per axis, Q2 nodes per axis, 1/N

// Q2 node adjacency per axis (also the restriction's row pattern on the coarse node line):
// an even coordinate reaches a-2..a+2 (clipped), an odd one a-1..a+1
struct Q2Axis
{
  uint64_t m;
  __device__ uint64_t lo(uint64_t a) const { return (a & 1) ? a - 1 : (a >= 2 ? a - 2 : 0); }
  __device__ uint64_t len(uint64_t a) const
  {
    if(a & 1) return 3;
    return (a + 2 <= m - 1 ? a + 2 : m - 1) - lo(a) + 1;
  }
  __device__ uint64_t pre(uint64_t a) const { return 8 * (a / 2) + ((a & 1) ? 5 : 0) - (a > 0 ? 2 : 0); }   // a < m
  __device__ uint64_t total() const { return 8 * ((m - 1) / 2) + 1; }
};

template<class Axis>
__device__ __forceinline__ uint64_t start2(uint64_t m, const Axis& ax, uint64_t r)
{
  return ax.total() * ax.pre(r / m) + ax.len(r / m) * ax.pre(r % m);
}

struct StokesARows
{
  StokesGrid g;
  uint64_t n, nnz;
  KeTable ke;
  __device__ uint64_t start(uint64_t r) const { return start2(g.m, Q2Axis{g.m}, r); }
  __device__ uint64_t len(uint64_t r) const { return Q2Axis{g.m}.len(r % g.m) * Q2Axis{g.m}.len(r / g.m); }
  __device__ void entry(uint64_t r, uint32_t s, double* v, uint64_t& c) const
  {
    const Q2Axis ax{g.m};
    const uint64_t a = r % g.m, b = r / g.m, lx = ax.len(a);
    const uint64_t ca = ax.lo(a) + s % lx, cb = ax.lo(b) + s / lx;
    c = cb * g.m + ca;
    // elements holding both nodes, in the host's element order (ey, then ex)
    auto erange = [&](uint64_t p, uint64_t q, uint64_t& lo, uint64_t& hi) {
      const uint64_t mx = p > q ? p : q, mn = p < q ? p : q;
      lo = mx <= 2 ? 0 : (mx - 1) / 2;
      hi = mn / 2 < g.N - 1 ? mn / 2 : g.N - 1;
    };
    uint64_t x0, x1, y0, y1;
    erange(a, ca, x0, x1);
    erange(b, cb, y0, y1);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for(uint64_t ey = y0; ey <= y1; ++ey)
      for(uint64_t ex = x0; ex <= x1; ++ex)
      {
        const int I = int((b - 2 * ey) * 3 + (a - 2 * ex)), J = int((cb - 2 * ey) * 3 + (ca - 2 * ex));
        s0 += ke.v[I][J][0];
        s1 += ke.v[I][J][1];
        s2 += ke.v[I][J][2];
        s3 += ke.v[I][J][3];
      }
    v[0] = s0;
    v[1] = s1;
    v[2] = s2;
    v[3] = s3;
  }
};

// elements (1D) holding Q2 node coordinate a
struct ElemAxis
{
  uint64_t N;
  __device__ uint64_t lo(uint64_t a) const { return (a & 1) ? (a - 1) / 2 : (a >= 2 ? a / 2 - 1 : 0); }
  __device__ uint64_t len(uint64_t a) const
  {
    if(a & 1) return 1;
    return (a / 2 < N ? a / 2 : N - 1) - lo(a) + 1;
  }
  __device__ uint64_t pre(uint64_t a) const { return 3 * (a / 2) + ((a & 1) ? 2 : 0) - (a > 0 ? 1 : 0); }
  __device__ uint64_t total() const { return 3 * N; }
};

struct StokesBRows
{
  StokesGrid g;
  uint64_t n, nnz;
  FgTable fg;
  __device__ uint64_t start(uint64_t r) const { return 3 * start2(g.m, ElemAxis{g.N}, r); }
  __device__ uint64_t len(uint64_t r) const { return 3 * ElemAxis{g.N}.len(r % g.m) * ElemAxis{g.N}.len(r / g.m); }
  __device__ void entry(uint64_t r, uint32_t s, double* v, uint64_t& c) const
  {
    const ElemAxis ax{g.N};
    const uint64_t a = r % g.m, b = r / g.m, lx = ax.len(a);
    const uint32_t p = s % 3, t = s / 3;
    const uint64_t ex = ax.lo(a) + t % lx, ey = ax.lo(b) + t / lx;
    const int L = int((b - 2 * ey) * 3 + (a - 2 * ex));
    c = 3 * (ey * g.N + ex) + p;
    const double sc = -(0.5 * g.h);
    v[0] = sc * fg.v[L][p][0];
    v[1] = sc * fg.v[L][p][1];
  }
};

struct StokesDRows
{
  StokesGrid g;
  uint64_t n, nnz;
  FgTable fg;
  __device__ uint64_t start(uint64_t r) const { return 9 * r; }
  __device__ uint64_t len(uint64_t) const { return 9; }
  __device__ void entry(uint64_t r, uint32_t L, double* v, uint64_t& c) const
  {
    const uint64_t e = r / 3, ex = e % g.N, ey = e / g.N;
    const uint32_t p = uint32_t(r % 3);
    c = (2 * ey + L / 3) * g.m + 2 * ex + L % 3;
    const double sc = -(0.5 * g.h);
    v[0] = sc * fg.v[L][p][0];
    v[1] = sc * fg.v[L][p][1];
  }
};

// Q2 prolongation: a fine node at an even coordinate copies coarse node a/2, an odd one
// interpolates the three nodes of coarse element a/4 at t = -1/2 or +1/2
__device__ __forceinline__ int q2_w1d(uint64_t a, uint64_t* ci, double* cw)
{
  if(!(a & 1))
  {
    ci[0] = a / 2;
    cw[0] = 1.0;
    return 1;
  }
  const uint64_t ec = a / 4;
  const bool left = (a % 4) == 1;
  ci[0] = 2 * ec;
  ci[1] = 2 * ec + 1;
  ci[2] = 2 * ec + 2;
  cw[0] = left ? 0.375 : -0.125;
  cw[1] = 0.75;
  cw[2] = left ? -0.125 : 0.375;
  return 3;
}

struct PvAxis
{
  __device__ uint64_t len(uint64_t a) const { return (a & 1) ? 3 : 1; }
  __device__ uint64_t pre(uint64_t a) const { return 4 * (a / 2) + (a & 1); }
  uint64_t T;
  __device__ uint64_t total() const { return T; }
};

struct StokesPvRows
{
  StokesGrid g;   // fine
  uint64_t mc, n, nnz;
  __device__ uint64_t start(uint64_t r) const { return start2(g.m, PvAxis{4 * g.N + 1}, r); }
  __device__ uint64_t len(uint64_t r) const { return PvAxis{0}.len(r % g.m) * PvAxis{0}.len(r / g.m); }
  __device__ void entry(uint64_t r, uint32_t s, double* v, uint64_t& c) const
  {
    uint64_t cx[3], cy[3];
    double wx[3], wy[3];
    const int nx = q2_w1d(r % g.m, cx, wx);
    q2_w1d(r / g.m, cy, wy);
    const int i = int(s % uint32_t(nx)), j = int(s / uint32_t(nx));
    c = cy[j] * mc + cx[i];
    v[0] = wx[i] * wy[j];
  }
};

// restriction Pv^T: fine nodes (ascending) holding coarse node A on one axis, with the weights
// the prolongation gives them
__device__ __forceinline__ int q2_r1d(uint64_t A, uint64_t mf, uint64_t* fi, double* fw)
{
  int q = 0;
  auto put = [&](int64_t a, double w) {
    if(a >= 0 && uint64_t(a) < mf)
    {
      fi[q] = uint64_t(a);
      fw[q++] = w;
    }
  };
  const int64_t k = int64_t(A / 2);
  if(A & 1)
  {
    put(4 * k + 1, 0.75);
    put(4 * k + 2, 1.0);
    put(4 * k + 3, 0.75);
  }
  else
  {
    put(4 * k - 3, -0.125);
    put(4 * k - 1, 0.375);
    put(4 * k, 1.0);
    put(4 * k + 1, 0.375);
    put(4 * k + 3, -0.125);
  }
  return q;
}

struct StokesRvRows
{
  uint64_t mc, mf, n, nnz;
  __device__ uint64_t start(uint64_t r) const { return start2(mc, Q2Axis{mc}, r); }
  __device__ uint64_t len(uint64_t r) const { return Q2Axis{mc}.len(r % mc) * Q2Axis{mc}.len(r / mc); }
  __device__ void entry(uint64_t r, uint32_t s, double* v, uint64_t& c) const
  {
    uint64_t fa[5], fb[5];
    double wa[5], wb[5];
    const int nx = q2_r1d(r % mc, mf, fa, wa);
    q2_r1d(r / mc, mf, fb, wb);
    const int i = int(s % uint32_t(nx)), j = int(s / uint32_t(nx));
    c = fb[j] * mf + fa[i];
    v[0] = wa[i] * wb[j];   // the prolongation's product: x weight times y weight
  }
};

// P1dc prolongation: fine element f of coarse element ec, rows 3f (constant), 3f+1 (x^), 3f+2 (y^)
struct StokesPpRows
{
  uint64_t N, n, nnz;   // fine cells per axis
  __device__ uint64_t start(uint64_t r) const
  {
    const uint64_t k = r % 3;
    return 5 * (r / 3) + (k >= 1 ? 3 : 0) + (k >= 2 ? 1 : 0);
  }
  __device__ uint64_t len(uint64_t r) const { return r % 3 == 0 ? 3 : 1; }
  __device__ void entry(uint64_t r, uint32_t s, double* v, uint64_t& c) const
  {
    const uint64_t f = r / 3, k = r % 3, ex = f % N, ey = f / N, Nc = N / 2;
    const uint64_t ec = (ey / 2) * Nc + ex / 2;
    if(k == 0)
    {
      c = 3 * ec + s;
      v[0] = s == 0 ? 1.0 : (s == 1 ? ((ex % 2) ? 0.5 : -0.5) : ((ey % 2) ? 0.5 : -0.5));
    }
    else
    {
      c = 3 * ec + k;
      v[0] = 0.5;
    }
  }
};

// restriction Pp^T: coarse rows 3ec (4 entries), 3ec+1 and 3ec+2 (8 entries each), over the four
// children f0 < f1 < f2 < f3 of the coarse element
struct StokesRpRows
{
  uint64_t N, n, nnz;   // fine cells per axis; n = coarse pressure dofs
  __device__ uint64_t start(uint64_t r) const
  {
    const uint64_t k = r % 3;
    return 20 * (r / 3) + (k >= 1 ? 4 : 0) + (k >= 2 ? 8 : 0);
  }
  __device__ uint64_t len(uint64_t r) const { return r % 3 == 0 ? 4 : 8; }
  __device__ void entry(uint64_t r, uint32_t s, double* v, uint64_t& c) const
  {
    const uint64_t Nc = N / 2, ec = r / 3, k = r % 3, cx = ec % Nc, cy = ec / Nc;
    const uint32_t ch = k == 0 ? s : s / 2;
    const uint64_t ex = 2 * cx + (ch & 1), ey = 2 * cy + (ch >> 1), f = ey * N + ex;
    if(k == 0)
    {
      c = 3 * f;
      v[0] = 1.0;
    }
    else if(s % 2 == 0)
    {
      c = 3 * f;
      v[0] = k == 1 ? ((ex % 2) ? 0.5 : -0.5) : ((ey % 2) ? 0.5 : -0.5);
    }
    else
    {
      c = 3 * f + k;
      v[0] = 0.5;
    }
  }
};

// thread t -> row t / MAXLEN, slot t % MAXLEN; BS values per entry
template<int MAXLEN, int BS, class Rows>
__global__ __launch_bounds__(256) void k_assemble_entries(Rows rows, double* __restrict__ val,
                                                          uint64_t* __restrict__ col, uint64_t* __restrict__ rowptr)
{
  const uint64_t total = rows.n * MAXLEN;
  for(uint64_t t = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; t < total; t += uint64_t(gridDim.x) * blockDim.x)
  {
    const uint64_t r = t / MAXLEN;
    const uint32_t s = uint32_t(t % MAXLEN);
    const uint64_t k = rows.start(r) + s;
    if(s == 0)
    {
      rowptr[r] = k;
      if(r + 1 == rows.n) rowptr[rows.n] = rows.nnz;
    }
    if(s < rows.len(r))
    {
      uint64_t c;
      rows.entry(r, s, val + k * BS, c);
      col[k] = c;
    }
  }
}

// velocity boundary nodes (ascending) with the lid values, pressure mean-filter vectors
__global__ void k_stokes_filters(StokesGrid g, uint64_t nb, uint64_t* __restrict__ filt, double* __restrict__ fval,
                                 uint64_t np, double* __restrict__ prim, double* __restrict__ dual)
{
  const uint64_t m = g.m, M = m - 1;
  for(uint64_t j = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; j < nb + np; j += uint64_t(gridDim.x) * blockDim.x)
  {
    if(j < nb)
    {
      if(!filt) continue;
      uint64_t x, y;
      if(j < m)
      {
        y = 0;
        x = j;
      }
      else if(j < m + 2 * (M - 1))
      {
        const uint64_t u = j - m;
        y = 1 + u / 2;
        x = (u & 1) ? M : 0;
      }
      else
      {
        y = M;
        x = j - m - 2 * (M - 1);
      }
      filt[j] = y * m + x;
      fval[2 * j] = (y == M && x > 0 && x < M) ? 1.0 : 0.0;
      fval[2 * j + 1] = 0.0;
    }
    else if(prim)
    {
      const uint64_t i = j - nb;
      prim[i] = i % 3 == 0 ? 1.0 : 0.0;
      dual[i] = i % 3 == 0 ? g.h * g.h : 0.0;
    }
  }
}

template<int MAXLEN, int BS, class Rows>
void launch_entries(hipStream_t st, const Rows& rows, double* val, uint64_t* col, uint64_t* rowptr)
{
  hipLaunchKernelGGL((k_assemble_entries<MAXLEN, BS, Rows>), dim3(grid_for(rows.n * MAXLEN)), dim3(kThreads), 0, st,
                     rows, val, col, rowptr);
}

// element tables by the host assembly's own formulas (stokes_assembly.hpp)
void stokes_tables(KeTable& ke, FgTable& fg)
{
  static const feat3amd::Assembly::detail::Moments mo;
  for(int I = 0; I < 9; ++I)
    for(int J = 0; J < 9; ++J)
    {
      const int i = I % 3, j = I / 3, ip = J % 3, jp = J / 3;
      const double gxx = mo.K[i][ip] * mo.M[j][jp];
      const double gyy = mo.M[i][ip] * mo.K[j][jp];
      const double gyx = mo.E[ip][i] * mo.E[j][jp];
      const double gxy = mo.E[i][ip] * mo.E[jp][j];
      ke.v[I][J][0] = gxx + gyy + gxx;
      ke.v[I][J][1] = gyx;
      ke.v[I][J][2] = gxy;
      ke.v[I][J][3] = gxx + gyy + gyy;
    }
  for(int I = 0; I < 9; ++I)
  {
    const int i = I % 3, j = I / 3;
    for(int p = 0; p < 3; ++p)
    {
      const int mx = p == 1 ? 1 : 0, my = p == 2 ? 1 : 0;
      fg.v[I][p][0] = mo.F[i][mx] * mo.G[j][my];
      fg.v[I][p][1] = mo.G[i][mx] * mo.F[j][my];
    }
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

int feat_stokes_device_sizes(int level, uint64_t* sizes)
{
  if(level < 0 || level > 12 || !sizes) return fail(FEAT_ERR_INVALID, "feat_stokes_device_sizes: bad argument");
  const uint64_t N = uint64_t(1) << level, m = 2 * N + 1;
  sizes[0] = m * m;                                         // velocity nodes
  sizes[1] = 3 * N * N;                                     // pressure dofs
  sizes[2] = (8 * N + 1) * (8 * N + 1);                     // A blocks
  sizes[3] = 27 * N * N;                                    // B blocks
  sizes[4] = 27 * N * N;                                    // D blocks
  sizes[5] = 4 * m - 4;                                     // boundary velocity nodes
  sizes[6] = level > 0 ? (4 * N + 1) * (4 * N + 1) : 0;     // Pv entries (= Rv entries)
  sizes[7] = level > 0 ? 5 * N * N : 0;                     // Pp entries (= Rp entries)
  sizes[8] = level > 0 ? (N + 1) * (N + 1) : 0;             // coarse velocity nodes
  sizes[9] = level > 0 ? 3 * (N / 2) * (N / 2) : 0;         // coarse pressure dofs
  return FEAT_OK;
}

int feat_stokes_assemble_device(feat_handle_t h, int level, const feat_stokes_level_dev_t* o)
{
  uint64_t sz[10];
  int rc = feat_stokes_device_sizes(level, sz);
  if(rc) return rc;
  if(!h || !o) return fail(FEAT_ERR_INVALID, "feat_stokes_assemble_device: null argument");
  auto group = [](const void* a, const void* b, const void* c) { return (a != nullptr) + (b != nullptr) + (c != nullptr); };
  const int ga = group(o->a_val, o->a_col, o->a_ptr), gb = group(o->b_val, o->b_col, o->b_ptr),
            gd = group(o->d_val, o->d_col, o->d_ptr), gf = group(o->filt, o->filt_val, o->filt),
            gm = group(o->prim, o->dual, o->prim), gpv = group(o->pv_val, o->pv_col, o->pv_ptr),
            gpp = group(o->pp_val, o->pp_col, o->pp_ptr), grv = group(o->rv_val, o->rv_col, o->rv_ptr),
            grp = group(o->rp_val, o->rp_col, o->rp_ptr);
  for(int gcount : {ga, gb, gd, gf, gm, gpv, gpp, grv, grp})
    if(gcount != 0 && gcount != 3) return fail(FEAT_ERR_INVALID, "feat_stokes_assemble_device: incomplete array group");
  if(level == 0 && (gpv || gpp || grv || grp)) return fail(FEAT_ERR_INVALID, "feat_stokes_assemble_device: level 0 has no transfer");
  const uint64_t N = uint64_t(1) << level;
  const StokesGrid g{N, 2 * N + 1, 1.0 / double(N)};
  KeTable ke;
  FgTable fg;
  stokes_tables(ke, fg);
  hipStream_t st = h->stream;
  if(ga) launch_entries<25, 4>(st, StokesARows{g, sz[0], sz[2], ke}, o->a_val, o->a_col, o->a_ptr);
  if(gb) launch_entries<12, 2>(st, StokesBRows{g, sz[0], sz[3], fg}, o->b_val, o->b_col, o->b_ptr);
  if(gd) launch_entries<9, 2>(st, StokesDRows{g, sz[1], sz[4], fg}, o->d_val, o->d_col, o->d_ptr);
  if(gf || gm)
    hipLaunchKernelGGL(k_stokes_filters, dim3(grid_for(sz[5] + sz[1])), dim3(kThreads), 0, st, g, sz[5], o->filt,
                       o->filt_val, sz[1], o->prim, o->dual);
  if(gpv) launch_entries<9, 1>(st, StokesPvRows{g, N + 1, sz[0], sz[6]}, o->pv_val, o->pv_col, o->pv_ptr);
  if(gpp) launch_entries<3, 1>(st, StokesPpRows{N, sz[1], sz[7]}, o->pp_val, o->pp_col, o->pp_ptr);
  if(grv) launch_entries<25, 1>(st, StokesRvRows{N + 1, 2 * N + 1, sz[8], sz[6]}, o->rv_val, o->rv_col, o->rv_ptr);
  if(grp) launch_entries<8, 1>(st, StokesRpRows{N, sz[9], sz[7]}, o->rp_val, o->rp_col, o->rp_ptr);
  return check_launch("feat_stokes_assemble_device");
}

} // extern "C"
