// feat3amd/assembly.hpp -- structured Q1 Poisson problems on the unit
// square/cube (setup time, host).
//
// Produces what the reference assembles with
//   SymbolicAssembler::assemble_matrix_std1 + LaplaceOperator   (Q1 stiffness),
//   GridTransfer::assemble_prolongation + row scaling           (control/scalar_basic.hpp:282-331),
//   UnitFilterAssembler on the boundary                          (homogeneous Dirichlet),
// on a 2^L x 2^L (x 2^L) grid of squares/cubes, but in lexicographic node
// numbering (x fastest) instead of the refinement numbering of
// geometry/intern/standard_vertex_refiner.hpp.  Matrix entries are the exact
// element sums: 2D element matrix (1/6)[4 -1 -2 -1], 3D h*[1/3, 0, -1/12, -1/12]
// by number of differing coordinates.  The stiffness pattern is the element
// adjacency pattern (3^d stencil incl. the stored zero face couplings in 3D);
// the prolongation stores only the interpolation weights {1, 1/2, 1/4, 1/8}
// (the reference's 2-level pattern additionally stores zeros, which change no
// result).
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

namespace feat3amd
{
  namespace Assembly
  {
    struct CsrHost
    {
      std::uint64_t rows = 0, cols = 0;
      std::vector<double> val;
      std::vector<std::uint64_t> col, rowptr;
    };

    inline std::uint64_t q1_nodes_1d(int level) { return (std::uint64_t(1) << level) + 1; }

    inline std::uint64_t ipow(std::uint64_t b, int e)
    {
      std::uint64_t r = 1;
      for(int i = 0; i < e; ++i) r *= b;
      return r;
    }

    /// Q1 Laplace stiffness matrix on level `level` of the unit square (dim 2) or cube (dim 3).
    inline CsrHost q1_laplace(int dim, int level)
    {
      const std::uint64_t m = q1_nodes_1d(level), N = m - 1;
      const std::uint64_t n = ipow(m, dim);
      const double h = 1.0 / double(N);
      // element value by number of differing coordinates (0..dim)
      double ev[4];
      if(dim == 2)
      {
        ev[0] = 4.0 / 6.0;
        ev[1] = -1.0 / 6.0;
        ev[2] = -2.0 / 6.0;
        ev[3] = 0.0;
      }
      else
      {
        ev[0] = h / 3.0;
        ev[1] = 0.0;
        ev[2] = -h / 12.0;
        ev[3] = -h / 12.0;
      }
      CsrHost A;
      A.rows = A.cols = n;
      A.rowptr.resize(n + 1);
      // row lengths
      auto len1 = [&](std::uint64_t c) -> std::uint64_t { return (c > 0) + 1 + (c < N); };
      A.rowptr[0] = 0;
      for(std::uint64_t r = 0; r < n; ++r)
      {
        std::uint64_t c[3] = {r % m, (r / m) % m, dim == 3 ? r / (m * m) : 0};
        std::uint64_t l = 1;
        for(int d = 0; d < dim; ++d) l *= len1(c[d]);
        A.rowptr[r + 1] = A.rowptr[r] + l;
      }
      A.val.resize(A.rowptr[n]);
      A.col.resize(A.rowptr[n]);
      const std::uint64_t sz = dim == 3 ? m : 1;
      for(std::uint64_t r = 0; r < n; ++r)
      {
        const std::int64_t c0 = std::int64_t(r % m), c1 = std::int64_t((r / m) % m), c2 = dim == 3 ? std::int64_t(r / (m * m)) : 0;
        std::uint64_t k = A.rowptr[r];
        const std::int64_t lo2 = dim == 3 ? -1 : 0, hi2 = dim == 3 ? 1 : 0;
        for(std::int64_t d2 = lo2; d2 <= hi2; ++d2)
          for(std::int64_t d1 = -1; d1 <= 1; ++d1)
            for(std::int64_t d0 = -1; d0 <= 1; ++d0)
            {
              const std::int64_t q0 = c0 + d0, q1 = c1 + d1, q2 = c2 + d2;
              if(q0 < 0 || q1 < 0 || q2 < 0 || q0 > std::int64_t(N) || q1 > std::int64_t(N) || (dim == 3 && q2 > std::int64_t(N)))
                continue;
              // number of elements containing both nodes: product over axes of
              // (2 if same coordinate and interior, 1 if same on the boundary, 1 if differing)
              std::uint64_t ne = 1;
              int ndiff = 0;
              const std::int64_t cc[3] = {c0, c1, c2}, dd[3] = {d0, d1, d2};
              for(int d = 0; d < dim; ++d)
              {
                if(dd[d] != 0)
                  ++ndiff;
                else
                  ne *= (cc[d] > 0 && cc[d] < std::int64_t(N)) ? 2 : 1;
              }
              A.col[k] = std::uint64_t(q0) + m * (std::uint64_t(q1) + sz * std::uint64_t(q2));
              A.val[k] = double(ne) * ev[ndiff];
              ++k;
            }
      }
      return A;
    }

    /// Prolongation from level-1 to level (interpolation weights only).
    inline CsrHost q1_prolongation(int dim, int level)
    {
      const std::uint64_t mf = q1_nodes_1d(level), mc = q1_nodes_1d(level - 1);
      const std::uint64_t nf = ipow(mf, dim), nc = ipow(mc, dim);
      CsrHost P;
      P.rows = nf;
      P.cols = nc;
      P.rowptr.resize(nf + 1);
      P.rowptr[0] = 0;
      for(std::uint64_t r = 0; r < nf; ++r)
      {
        std::uint64_t c[3] = {r % mf, (r / mf) % mf, dim == 3 ? r / (mf * mf) : 0};
        std::uint64_t l = 1;
        for(int d = 0; d < dim; ++d) l *= (c[d] & 1) ? 2 : 1;
        P.rowptr[r + 1] = P.rowptr[r] + l;
      }
      P.val.resize(P.rowptr[nf]);
      P.col.resize(P.rowptr[nf]);
      for(std::uint64_t r = 0; r < nf; ++r)
      {
        const std::uint64_t c[3] = {r % mf, (r / mf) % mf, dim == 3 ? r / (mf * mf) : 0};
        // per axis: list of (coarse index, weight), ascending
        std::uint64_t ci[3][2];
        double w[3][2];
        int cnt[3] = {1, 1, 1};
        for(int d = 0; d < 3; ++d)
        {
          if(d >= dim)
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
        std::uint64_t k = P.rowptr[r];
        for(int a2 = 0; a2 < cnt[2]; ++a2)
          for(int a1 = 0; a1 < cnt[1]; ++a1)
            for(int a0 = 0; a0 < cnt[0]; ++a0)
            {
              P.col[k] = ci[0][a0] + mc * (ci[1][a1] + (dim == 3 ? mc * ci[2][a2] : 0));
              P.val[k] = w[0][a0] * w[1][a1] * w[2][a2];
              ++k;
            }
      }
      return P;
    }

    /// Explicit CSR transpose with ascending column order in each row
    /// (SparseMatrixCSR::transpose, used for R = P^T in control/scalar_basic.hpp:327).
    inline CsrHost transpose(const CsrHost& A)
    {
      CsrHost T;
      T.rows = A.cols;
      T.cols = A.rows;
      T.rowptr.assign(A.cols + 1, 0);
      for(auto c : A.col) ++T.rowptr[c + 1];
      for(std::uint64_t i = 0; i < A.cols; ++i) T.rowptr[i + 1] += T.rowptr[i];
      T.val.resize(A.val.size());
      T.col.resize(A.col.size());
      std::vector<std::uint64_t> pos(T.rowptr.begin(), T.rowptr.end() - 1);
      for(std::uint64_t r = 0; r < A.rows; ++r)
        for(std::uint64_t i = A.rowptr[r]; i < A.rowptr[r + 1]; ++i)
        {
          const std::uint64_t k = pos[A.col[i]]++;
          T.col[k] = r;
          T.val[k] = A.val[i];
        }
      return T;
    }

    /// Boundary node indices (ascending) = homogeneous Dirichlet unit filter.
    inline std::vector<std::uint64_t> q1_boundary(int dim, int level)
    {
      const std::uint64_t m = q1_nodes_1d(level), N = m - 1, n = ipow(m, dim);
      std::vector<std::uint64_t> b;
      for(std::uint64_t r = 0; r < n; ++r)
      {
        const std::uint64_t c[3] = {r % m, (r / m) % m, dim == 3 ? r / (m * m) : 1};
        bool on = false;
        for(int d = 0; d < dim; ++d) on = on || c[d] == 0 || c[d] == N;
        if(on) b.push_back(r);
      }
      return b;
    }

    /// Q1 load vector of f = -Laplace(prod sin(pi x_k)) = dim*pi^2*prod sin(pi x_k)
    /// (Analytic::Common::SineBubbleFunction, analytic/common.hpp:206-232), integrated
    /// exactly: int sin(pi x) phi_i dx = sin(pi x_i) * 2(1 - cos(pi h)) / (pi^2 h).
    inline std::vector<double> q1_rhs_sine(int dim, int level)
    {
      const std::uint64_t m = q1_nodes_1d(level), N = m - 1, n = ipow(m, dim);
      const double pi = 3.14159265358979323846, h = 1.0 / double(N);
      const double f1 = 2.0 * (1.0 - std::cos(pi * h)) / (pi * pi * h);
      std::vector<double> s(m);
      for(std::uint64_t i = 0; i < m; ++i) s[i] = std::sin(pi * double(i) * h) * f1;
      std::vector<double> b(n);
      for(std::uint64_t r = 0; r < n; ++r)
      {
        double v = double(dim) * pi * pi;
        v *= s[r % m];
        v *= s[(r / m) % m];
        if(dim == 3) v *= s[r / (m * m)];
        b[r] = v;
      }
      return b;
    }

    /// xorshift64* uniform values in [a,b] (kernel/util/random.hpp:84-110, 259-280;
    /// DenseVector(Random&, n, a, b), lafem/dense_vector.hpp:322-332).
    inline void random_fill(std::uint64_t seed, double* v, std::uint64_t n, double a, double b)
    {
      std::uint64_t x = seed ? seed : 28054777172512ull;
      const double lo = a < b ? a : b, hi = a < b ? b : a;
      const double d = 1.0 / double(0xFFFFFFFFull);
      for(std::uint64_t i = 0; i < n; ++i)
      {
        x ^= x >> 12;
        x ^= x << 25;
        x ^= x >> 27;
        const std::uint64_t r = x * 2685821657736338717ull;
        const double u = double(r & 0xFFFFFFFFull) * d;
        v[i] = (lo + 2.220446049250313e-16 < hi) ? std::fma(u, hi - lo, lo) : lo;
      }
    }
  } // namespace Assembly
} // namespace feat3amd
