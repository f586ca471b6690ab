// feat3amd/stokes_assembly.hpp -- structured Q2/P1dc Stokes problems on the
// unit square (setup time, host).
//
// Produces, on a 2^L x 2^L grid of squares, what the reference assembles for
// the Stokes benchmark (benchmarks/stokes_multigrid-bench.cpp:576-697):
//   A  = DuDvOperatorBlocked<2>            (assembly/common_operators.hpp:875-961)
//        block(a,b) of (test I, trial J) = d_ab (grad phi_J . grad phi_I) + d_a phi_J d_b phi_I
//   B  = -(d_k phi_I, psi_p)               (control/stokes_blocked.hpp:784-798, scale -1)
//   D  = B^T
//   velocity UnitFilterBlocked on the whole boundary (lid-driven cavity:
//        u = (1,0) on the open top edge, 0 elsewhere),
//   pressure MeanFilter (assembly/mean_filter_assembler.hpp:67-85):
//        primal = interpolant of 1, dual = integrals of the basis functions,
//   Q2 and P1dc prolongations (GridTransfer, then shrink(1e-3) drops the
//        numerical zeros), restrictions = their transposes.
// Element integrals are formed from exact 1D moment tables (3-point Gauss
// is exact for every integrand here); the reference integrates the same
// polynomials with gauss-legendre:3, so values agree to rounding.
//
// Numbering (lexicographic, x fastest):
//   velocity node (a,b) of the (2N+1)^2 Q2 node grid -> b*(2N+1)+a,
//   P1dc dofs of element e = ey*N+ex -> 3e (constant), 3e+1 (x^), 3e+2 (y^),
//   where (x^, y^) are the element's reference coordinates in [-1,1]^2
//   (space/discontinuous/evaluator.hpp:285-352, the linearised inverse trafo
//   at the element centre).
#pragma once

#include <feat3amd/assembly.hpp>

#include <cstdint>
#include <vector>

namespace feat3amd
{
  namespace Assembly
  {
    /// block-CSR (row-major bh x bw blocks), as SparseMatrixBCSR stores them
    struct BcsrHost
    {
      std::uint64_t rows = 0, cols = 0; // in blocks
      int bh = 1, bw = 1;
      std::vector<double> val;
      std::vector<std::uint64_t> col, rowptr;
    };

    struct StokesLevelHost
    {
      int level = 0;
      std::uint64_t nv = 0, np = 0; // velocity nodes, pressure dofs
      BcsrHost A, B, D;
      std::vector<std::uint64_t> filt;  // Dirichlet velocity nodes (ascending)
      std::vector<double> filt_val;     // 2 values per node
      std::vector<double> prim, dual;   // mean filter vectors
      CsrHost Pv, Pp;                   // prolongation from level-1 (empty on level 0)
    };

    namespace detail
    {
      // 1D quadratic Lagrange basis on [-1,1] with nodes -1, 0, 1 and its derivative
      inline double lag(int i, double t)
      {
        return i == 0 ? 0.5 * t * (t - 1.0) : (i == 1 ? 1.0 - t * t : 0.5 * t * (t + 1.0));
      }
      inline double dlag(int i, double t) { return i == 0 ? t - 0.5 : (i == 1 ? -2.0 * t : t + 0.5); }

      struct Moments
      {
        double M[3][3], K[3][3], E[3][3]; // int L_a L_b, int L_a' L_b', int L_a' L_b
        double F[3][2], G[3][2];          // int L_a' t^m, int L_a t^m  (m = 0, 1)
        Moments()
        {
          const double q = std::sqrt(0.6), gp[3] = {-q, 0.0, q}, gw[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
          for(int a = 0; a < 3; ++a)
          {
            for(int b = 0; b < 3; ++b)
            {
              M[a][b] = K[a][b] = E[a][b] = 0.0;
              for(int k = 0; k < 3; ++k)
              {
                M[a][b] += gw[k] * lag(a, gp[k]) * lag(b, gp[k]);
                K[a][b] += gw[k] * dlag(a, gp[k]) * dlag(b, gp[k]);
                E[a][b] += gw[k] * dlag(a, gp[k]) * lag(b, gp[k]);
              }
            }
            for(int m = 0; m < 2; ++m)
            {
              F[a][m] = G[a][m] = 0.0;
              for(int k = 0; k < 3; ++k)
              {
                const double tm = m == 0 ? 1.0 : gp[k];
                F[a][m] += gw[k] * dlag(a, gp[k]) * tm;
                G[a][m] += gw[k] * lag(a, gp[k]) * tm;
              }
            }
          }
        }
      };

      // sorted columns of a 1D index range, for the tensor-product patterns
      inline void node_range(std::uint64_t a, std::uint64_t nmax, std::uint64_t& lo, std::uint64_t& hi)
      {
        // a Q2 node at an even coordinate belongs to up to two elements, at an odd one to one
        if(a % 2 == 0)
        {
          lo = a >= 2 ? a - 2 : 0;
          hi = a + 2 <= nmax ? a + 2 : nmax;
        }
        else
        {
          lo = a - 1;
          hi = a + 1;
        }
      }
    } // namespace detail

    /// Q2/P1dc Stokes operator, filters and mean filter on level `level` of the unit square.
    inline StokesLevelHost stokes_q2p1dc(int level)
    {
      StokesLevelHost S;
      S.level = level;
      const std::uint64_t N = std::uint64_t(1) << level, m = 2 * N + 1;
      const double h = 1.0 / double(N);
      S.nv = m * m;
      S.np = 3 * N * N;
      static const detail::Moments mo;

      // --- A: element-adjacency pattern of the Q2 nodes, 2x2 blocks
      BcsrHost& A = S.A;
      A.rows = A.cols = S.nv;
      A.bh = A.bw = 2;
      A.rowptr.assign(S.nv + 1, 0);
      std::vector<std::uint64_t> xlo(S.nv), xhi(S.nv), ylo(S.nv), yhi(S.nv);
      for(std::uint64_t b = 0; b < m; ++b)
        for(std::uint64_t a = 0; a < m; ++a)
        {
          const std::uint64_t r = b * m + a;
          detail::node_range(a, m - 1, xlo[r], xhi[r]);
          detail::node_range(b, m - 1, ylo[r], yhi[r]);
          A.rowptr[r + 1] = A.rowptr[r] + (xhi[r] - xlo[r] + 1) * (yhi[r] - ylo[r] + 1);
        }
      const std::uint64_t nnzA = A.rowptr[S.nv];
      A.col.resize(nnzA);
      A.val.assign(nnzA * 4, 0.0);
      for(std::uint64_t r = 0; r < S.nv; ++r)
      {
        std::uint64_t k = A.rowptr[r];
        for(std::uint64_t bb = ylo[r]; bb <= yhi[r]; ++bb)
          for(std::uint64_t aa = xlo[r]; aa <= xhi[r]; ++aa)
            A.col[k++] = bb * m + aa;
      }
      auto posA = [&](std::uint64_t r, std::uint64_t c) {
        const std::uint64_t ca = c % m, cb = c / m;
        return A.rowptr[r] + (cb - ylo[r]) * (xhi[r] - xlo[r] + 1) + (ca - xlo[r]);
      };
      // element block matrices (h-independent in 2D)
      double Ke[9][9][2][2];
      for(int I = 0; I < 9; ++I)
        for(int J = 0; J < 9; ++J)
        {
          const int i = I % 3, j = I / 3, ip = J % 3, jp = J / 3;
          const double gxx = mo.K[i][ip] * mo.M[j][jp];
          const double gyy = mo.M[i][ip] * mo.K[j][jp];
          const double gyx = mo.E[ip][i] * mo.E[j][jp]; // int d_y phi_I d_x phi_J
          const double gxy = mo.E[i][ip] * mo.E[jp][j]; // int d_x phi_I d_y phi_J
          Ke[I][J][0][0] = gxx + gyy + gxx;
          Ke[I][J][0][1] = gyx;
          Ke[I][J][1][0] = gxy;
          Ke[I][J][1][1] = gxx + gyy + gyy;
        }
      auto enode = [&](std::uint64_t ex, std::uint64_t ey, int L) { return (2 * ey + L / 3) * m + 2 * ex + L % 3; };
      for(std::uint64_t ey = 0; ey < N; ++ey)
        for(std::uint64_t ex = 0; ex < N; ++ex)
          for(int I = 0; I < 9; ++I)
          {
            const std::uint64_t r = enode(ex, ey, I);
            for(int J = 0; J < 9; ++J)
            {
              double* blk = &A.val[posA(r, enode(ex, ey, J)) * 4];
              for(int q = 0; q < 4; ++q) blk[q] += Ke[I][J][q / 2][q % 2];
            }
          }

      // --- B (2x1 blocks) and D = B^T (1x2 blocks)
      double Be[9][3][2];
      for(int I = 0; I < 9; ++I)
      {
        const int i = I % 3, j = I / 3;
        for(int p = 0; p < 3; ++p)
        {
          const int mx = p == 1 ? 1 : 0, my = p == 2 ? 1 : 0;
          Be[I][p][0] = -(0.5 * h) * (mo.F[i][mx] * mo.G[j][my]);
          Be[I][p][1] = -(0.5 * h) * (mo.G[i][mx] * mo.F[j][my]);
        }
      }
      BcsrHost& B = S.B;
      B.rows = S.nv;
      B.cols = S.np;
      B.bh = 2;
      B.bw = 1;
      B.rowptr.assign(S.nv + 1, 0);
      auto elems_of = [&](std::uint64_t a, std::uint64_t& lo, std::uint64_t& hi) {
        // elements (1D) that contain Q2 node coordinate a
        if(a % 2 == 0)
        {
          lo = a >= 2 ? a / 2 - 1 : 0;
          hi = a / 2 < N ? a / 2 : N - 1;
        }
        else
          lo = hi = (a - 1) / 2;
      };
      for(std::uint64_t r = 0; r < S.nv; ++r)
      {
        std::uint64_t e0x, e1x, e0y, e1y;
        elems_of(r % m, e0x, e1x);
        elems_of(r / m, e0y, e1y);
        B.rowptr[r + 1] = B.rowptr[r] + 3 * (e1x - e0x + 1) * (e1y - e0y + 1);
      }
      B.col.resize(B.rowptr[S.nv]);
      B.val.assign(B.rowptr[S.nv] * 2, 0.0);
      for(std::uint64_t r = 0; r < S.nv; ++r)
      {
        std::uint64_t e0x, e1x, e0y, e1y, k = B.rowptr[r];
        elems_of(r % m, e0x, e1x);
        elems_of(r / m, e0y, e1y);
        for(std::uint64_t ey = e0y; ey <= e1y; ++ey)
          for(std::uint64_t ex = e0x; ex <= e1x; ++ex)
          {
            const std::uint64_t e = ey * N + ex;
            // local node of r in element e
            const int L = int((r / m - 2 * ey) * 3 + (r % m - 2 * ex));
            for(int p = 0; p < 3; ++p)
            {
              B.col[k] = 3 * e + std::uint64_t(p);
              B.val[2 * k] = Be[L][p][0];
              B.val[2 * k + 1] = Be[L][p][1];
              ++k;
            }
          }
      }
      BcsrHost& D = S.D;
      D.rows = S.np;
      D.cols = S.nv;
      D.bh = 1;
      D.bw = 2;
      D.rowptr.assign(S.np + 1, 0);
      for(std::uint64_t r = 0; r < S.np; ++r) D.rowptr[r + 1] = D.rowptr[r] + 9;
      D.col.resize(S.np * 9);
      D.val.resize(S.np * 18);
      for(std::uint64_t ey = 0; ey < N; ++ey)
        for(std::uint64_t ex = 0; ex < N; ++ex)
          for(int p = 0; p < 3; ++p)
          {
            const std::uint64_t r = 3 * (ey * N + ex) + std::uint64_t(p);
            for(int L = 0; L < 9; ++L) // element nodes in ascending global order
            {
              const std::uint64_t k = D.rowptr[r] + std::uint64_t(L);
              D.col[k] = enode(ex, ey, L);
              D.val[2 * k] = Be[L][p][0];
              D.val[2 * k + 1] = Be[L][p][1];
            }
          }

      // --- velocity boundary (lid-driven cavity) and pressure mean filter
      for(std::uint64_t b = 0; b < m; ++b)
        for(std::uint64_t a = 0; a < m; ++a)
          if(a == 0 || b == 0 || a == m - 1 || b == m - 1)
          {
            S.filt.push_back(b * m + a);
            const bool lid = (b == m - 1) && a > 0 && a < m - 1;
            S.filt_val.push_back(lid ? 1.0 : 0.0);
            S.filt_val.push_back(0.0);
          }
      S.prim.assign(S.np, 0.0);
      S.dual.assign(S.np, 0.0);
      for(std::uint64_t e = 0; e < N * N; ++e)
      {
        S.prim[3 * e] = 1.0;
        S.dual[3 * e] = h * h; // int x^ = int y^ = 0 on the element
      }

      // --- prolongations from level-1
      if(level > 0)
      {
        const std::uint64_t Nc = N / 2, mc = 2 * Nc + 1;
        // 1D Q2 interpolation weights of fine node A from the coarse node line
        auto w1d = [&](std::uint64_t Af, std::uint64_t* ci, double* cw) -> int {
          if(Af % 2 == 0)
          {
            ci[0] = Af / 2;
            cw[0] = 1.0;
            return 1;
          }
          const std::uint64_t ec = Af / 4;
          const bool left = (Af % 4) == 1; // t = -1/2 in the coarse element
          ci[0] = 2 * ec;
          ci[1] = 2 * ec + 1;
          ci[2] = 2 * ec + 2;
          cw[0] = left ? 0.375 : -0.125;
          cw[1] = 0.75;
          cw[2] = left ? -0.125 : 0.375;
          return 3;
        };
        CsrHost& P = S.Pv;
        P.rows = S.nv;
        P.cols = mc * mc;
        P.rowptr.assign(S.nv + 1, 0);
        for(std::uint64_t b = 0; b < m; ++b)
          for(std::uint64_t a = 0; a < m; ++a)
          {
            std::uint64_t cx[3], cy[3];
            double wx[3], wy[3];
            const int nx = w1d(a, cx, wx), ny = w1d(b, cy, wy);
            for(int j = 0; j < ny; ++j)
              for(int i = 0; i < nx; ++i)
              {
                P.col.push_back(cy[j] * mc + cx[i]);
                P.val.push_back(wx[i] * wy[j]);
              }
            P.rowptr[b * m + a + 1] = P.col.size();
          }
        CsrHost& Q = S.Pp;
        Q.rows = S.np;
        Q.cols = 3 * Nc * Nc;
        Q.rowptr.assign(S.np + 1, 0);
        for(std::uint64_t ey = 0; ey < N; ++ey)
          for(std::uint64_t ex = 0; ex < N; ++ex)
          {
            const std::uint64_t ec = (ey / 2) * Nc + ex / 2, f = ey * N + ex;
            const double sx = (ex % 2) ? 0.5 : -0.5, sy = (ey % 2) ? 0.5 : -0.5;
            // constant: c0 + sx*c1 + sy*c2 ; x^: c1/2 ; y^: c2/2
            Q.col.insert(Q.col.end(), {3 * ec, 3 * ec + 1, 3 * ec + 2});
            Q.val.insert(Q.val.end(), {1.0, sx, sy});
            Q.rowptr[3 * f + 1] = Q.col.size();
            Q.col.push_back(3 * ec + 1);
            Q.val.push_back(0.5);
            Q.rowptr[3 * f + 2] = Q.col.size();
            Q.col.push_back(3 * ec + 2);
            Q.val.push_back(0.5);
            Q.rowptr[3 * f + 3] = Q.col.size();
          }
      }
      return S;
    }
  } // namespace Assembly
} // namespace feat3amd
