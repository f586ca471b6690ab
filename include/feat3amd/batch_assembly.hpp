// feat3amd/batch_assembly.hpp -- the batch of independent 2D Q1 diffusion
// problems (setup time, host).
//
// Each instance is -div(c grad u) = f on the unit square, homogeneous Dirichlet
// boundary, Q1 on a 2^L x 2^L grid (lexicographic numbering as in
// assembly.hpp), with a diffusion coefficient c that is constant on each cell
// of the instance's finest level: log-normal, c = exp(sigma * z), z standard
// normal (Box-Muller on xorshift64* uniforms).  The reference has no
// variable-coefficient operator (assembly/common_operators.hpp:36
// LaplaceOperator is constant-coefficient); with c == 1 this operator is
// exactly the Q1 Laplace matrix of assembly.hpp.  Element matrices use the
// 2x2 Gauss rule (exact for a cell-wise constant coefficient); on the coarser
// multigrid levels every Gauss point samples the finest cell containing it,
// which is what re-assembling the same coefficient function on each level
// does (geometric multigrid).
#pragma once

#include <feat3amd/assembly.hpp>

#include <cmath>
#include <cstdint>
#include <vector>

namespace feat3amd
{
  namespace Assembly
  {
    /// gradient products of the Q1 basis at the 4 Gauss points of the reference
    /// square: G[q][a][b] = (d_t phi_a d_t phi_b + d_s phi_a d_s phi_b)(x_q), local
    /// node a = ay*2 + ax (lexicographic); the map to a square of width h scales by 1
    struct Q1GaussTable
    {
      double G[4][4][4];
      double pt[4][2];
      Q1GaussTable()
      {
        const double g = 1.0 / std::sqrt(3.0);
        const double gp[2] = {-g, g};
        for(int q = 0; q < 4; ++q)
        {
          const double t = gp[q % 2], s = gp[q / 2];
          pt[q][0] = t;
          pt[q][1] = s;
          double dt[4], ds[4];
          for(int a = 0; a < 4; ++a)
          {
            const int ax = a % 2, ay = a / 2;
            const double lx = ax ? 0.5 * (1.0 + t) : 0.5 * (1.0 - t);
            const double ly = ay ? 0.5 * (1.0 + s) : 0.5 * (1.0 - s);
            const double dlx = ax ? 0.5 : -0.5, dly = ay ? 0.5 : -0.5;
            dt[a] = dlx * ly;
            ds[a] = lx * dly;
          }
          for(int a = 0; a < 4; ++a)
            for(int b = 0; b < 4; ++b) G[q][a][b] = dt[a] * dt[b] + ds[a] * ds[b];
        }
      }
    };

    /// Q1 diffusion matrix on `level` for a coefficient given per cell of `level_c` >= level.
    inline CsrHost q1_diffusion_2d(int level, const std::vector<double>& coef, int level_c)
    {
      static const Q1GaussTable T;
      CsrHost A = q1_laplace(2, level);  // pattern: 9-point element adjacency, ascending columns
      for(auto& v : A.val) v = 0.0;
      const std::uint64_t m = q1_nodes_1d(level), N = m - 1, Nc = std::uint64_t(1) << level_c;
      const double h = 1.0 / double(N);
      auto pos = [&](std::uint64_t r, std::uint64_t c) {
        std::uint64_t k = A.rowptr[r];
        while(A.col[k] != c) ++k;
        return k;
      };
      for(std::uint64_t ey = 0; ey < N; ++ey)
        for(std::uint64_t ex = 0; ex < N; ++ex)
        {
          double cq[4];
          for(int q = 0; q < 4; ++q)
          {
            const double x = (double(ex) + 0.5 + 0.5 * T.pt[q][0]) * h;
            const double y = (double(ey) + 0.5 + 0.5 * T.pt[q][1]) * h;
            std::uint64_t ix = std::uint64_t(x * double(Nc)), iy = std::uint64_t(y * double(Nc));
            if(ix >= Nc) ix = Nc - 1;
            if(iy >= Nc) iy = Nc - 1;
            cq[q] = coef[iy * Nc + ix];
          }
          std::uint64_t node[4];
          for(int a = 0; a < 4; ++a) node[a] = (ey + std::uint64_t(a / 2)) * m + ex + std::uint64_t(a % 2);
          for(int a = 0; a < 4; ++a)
            for(int b = 0; b < 4; ++b)
            {
              double k = 0.0;
              for(int q = 0; q < 4; ++q) k = std::fma(cq[q], T.G[q][a][b], k);
              A.val[pos(node[a], node[b])] += k;
            }
        }
      return A;
    }

    /// 3D Q1 diffusion matrix -div(c K grad u), K = diag(1, 1, kz), on `level` of the unit
    /// cube for a coefficient given per cell of `level_c` >= level (2x2x2 Gauss, every Gauss
    /// point samples the finest cell containing it).  Pattern: the 27-point pattern of
    /// q1_laplace(3, level).
    inline CsrHost q1_diffusion_3d(int level, const std::vector<double>& coef, int level_c, double kz)
    {
      CsrHost A = q1_laplace(3, level);
      for(auto& v : A.val) v = 0.0;
      const std::uint64_t m = q1_nodes_1d(level), N = m - 1, Nc = std::uint64_t(1) << level_c;
      const double h = 1.0 / double(N);
      const double g = 1.0 / std::sqrt(3.0);
      // gradient products of the trilinear basis at the 8 Gauss points (reference cube [-1,1]^3)
      static double G[8][8][8];
      static double lastkz = -1.0;
      if(lastkz != kz)
      {
        for(int q = 0; q < 8; ++q)
        {
          const double p[3] = {(q & 1) ? g : -g, (q & 2) ? g : -g, (q & 4) ? g : -g};
          double d[8][3];
          for(int a = 0; a < 8; ++a)
          {
            double l[3], dl[3];
            for(int k = 0; k < 3; ++k)
            {
              const int bit = (a >> k) & 1;
              l[k] = bit ? 0.5 * (1.0 + p[k]) : 0.5 * (1.0 - p[k]);
              dl[k] = bit ? 0.5 : -0.5;
            }
            d[a][0] = dl[0] * l[1] * l[2];
            d[a][1] = l[0] * dl[1] * l[2];
            d[a][2] = l[0] * l[1] * dl[2];
          }
          for(int a = 0; a < 8; ++a)
            for(int b = 0; b < 8; ++b) G[q][a][b] = d[a][0] * d[b][0] + d[a][1] * d[b][1] + kz * (d[a][2] * d[b][2]);
        }
        lastkz = kz;
      }
      auto pos = [&](std::uint64_t r, std::uint64_t c) {
        std::uint64_t k = A.rowptr[r];
        while(A.col[k] != c) ++k;
        return k;
      };
      // (2/h)^2 from the gradients times (h/2)^3 from the volume
      const double scale = 0.5 * h;
      for(std::uint64_t ez = 0; ez < N; ++ez)
        for(std::uint64_t ey = 0; ey < N; ++ey)
          for(std::uint64_t ex = 0; ex < N; ++ex)
          {
            double cq[8];
            for(int q = 0; q < 8; ++q)
            {
              const double x = (double(ex) + 0.5 + 0.5 * ((q & 1) ? g : -g)) * h;
              const double y = (double(ey) + 0.5 + 0.5 * ((q & 2) ? g : -g)) * h;
              const double z = (double(ez) + 0.5 + 0.5 * ((q & 4) ? g : -g)) * h;
              std::uint64_t ix = std::uint64_t(x * double(Nc)), iy = std::uint64_t(y * double(Nc)),
                            iz = std::uint64_t(z * double(Nc));
              ix = ix < Nc ? ix : Nc - 1;
              iy = iy < Nc ? iy : Nc - 1;
              iz = iz < Nc ? iz : Nc - 1;
              cq[q] = coef[(iz * Nc + iy) * Nc + ix];
            }
            std::uint64_t node[8];
            for(int a = 0; a < 8; ++a)
              node[a] = ((ez + std::uint64_t((a >> 2) & 1)) * m + ey + std::uint64_t((a >> 1) & 1)) * m + ex +
                        std::uint64_t(a & 1);
            for(int a = 0; a < 8; ++a)
              for(int b = 0; b < 8; ++b)
              {
                double k = 0.0;
                for(int q = 0; q < 8; ++q) k = std::fma(cq[q], G[q][a][b], k);
                A.val[pos(node[a], node[b])] += scale * k;
              }
          }
      return A;
    }

    /// node feature of a cell coefficient field: log of the mean coefficient of the
    /// (up to 8) cells around each node of level `level_c`
    inline std::vector<float> node_log_coef_3d(const std::vector<double>& coef, int level_c)
    {
      const std::uint64_t Nc = std::uint64_t(1) << level_c, m = Nc + 1;
      std::vector<float> f(m * m * m);
      for(std::uint64_t z = 0; z < m; ++z)
        for(std::uint64_t y = 0; y < m; ++y)
          for(std::uint64_t x = 0; x < m; ++x)
          {
            double s = 0.0;
            int n = 0;
            for(int dz = -1; dz <= 0; ++dz)
              for(int dy = -1; dy <= 0; ++dy)
                for(int dx = -1; dx <= 0; ++dx)
                {
                  const std::int64_t cx = std::int64_t(x) + dx, cy = std::int64_t(y) + dy, cz = std::int64_t(z) + dz;
                  if(cx < 0 || cy < 0 || cz < 0 || cx >= std::int64_t(Nc) || cy >= std::int64_t(Nc) || cz >= std::int64_t(Nc))
                    continue;
                  s += coef[(std::uint64_t(cz) * Nc + std::uint64_t(cy)) * Nc + std::uint64_t(cx)];
                  ++n;
                }
            f[(z * m + y) * m + x] = float(std::log(s / double(n)));
          }
      return f;
    }

    /// xorshift64* step (kernel/util/random.hpp:84-110)
    inline std::uint64_t rng_next(std::uint64_t& x)
    {
      x ^= x >> 12;
      x ^= x << 25;
      x ^= x >> 27;
      return x * 2685821657736338717ull;
    }

    /// levels of the batch instances, uniform in [lmin, lmax]
    inline std::vector<int> batch_levels(std::uint64_t seed, int ninst, int lmin, int lmax)
    {
      std::uint64_t x = seed ? seed : 28054777172512ull;
      std::vector<int> l(ninst);
      for(int i = 0; i < ninst; ++i) l[i] = lmin + int(rng_next(x) % std::uint64_t(lmax - lmin + 1));
      return l;
    }

    /// cell coefficients exp(sigma * z), z ~ N(0,1) by Box-Muller on uniforms in (0,1]
    inline std::vector<double> lognormal_cells(std::uint64_t seed, std::uint64_t n, double sigma)
    {
      std::vector<double> u(2 * ((n + 1) / 2) + 2);
      random_fill(seed, u.data(), u.size(), 0.0, 1.0);
      std::vector<double> c(n);
      const double two_pi = 6.283185307179586;
      for(std::uint64_t i = 0; i < n; i += 2)
      {
        const double u1 = 1.0 - u[i], u2 = u[i + 1];
        const double r = std::sqrt(-2.0 * std::log(u1 > 0.0 ? u1 : 1e-300));
        c[i] = std::exp(sigma * r * std::cos(two_pi * u2));
        if(i + 1 < n) c[i + 1] = std::exp(sigma * r * std::sin(two_pi * u2));
      }
      return c;
    }

    /// the standard batch: instance i has level levels[i], coefficient seed seed+1000+i,
    /// right-hand side seed seed+2000+i (uniform [-1,1], boundary rows zero)
    struct BatchSpec
    {
      std::uint64_t seed = 28054777172512ull;
      int ninst = 64, lmin = 3, lmax = 8;
      double sigma = 1.0;
      std::uint64_t coef_seed(int i) const { return seed + 1000u + std::uint64_t(i); }
      std::uint64_t rhs_seed(int i) const { return seed + 2000u + std::uint64_t(i); }
    };
  } // namespace Assembly
} // namespace feat3amd
