// feat3amd/batch.hpp -- many independent PCG(MG-V(Richardson(Jacobi))) solves of
// uneven size advanced together.
//
// Every instance runs the reference's method (solver/pcg.hpp:188-288 over
// solver/multigrid.hpp:1270-2000, as the single-instance PCG of solver.hpp):
// its own scalars (gamma, alpha, beta), its own stopping state
// (iterative.hpp:748-885) and its own multigrid hierarchy from its finest
// level down to level 1.  The batch turns this into a fixed sequence of
// launches over all instances at once:
//   * PCG vectors of all instances back to back ("PCG layout"); scalar work per
//     instance through chunked kernels (feat_batch_*); an instance whose status
//     leaves `progress` is frozen (x, r no longer touched);
//   * multigrid level k holds every instance whose finest level is >= k, back to
//     back; its matrix, restriction and prolongation are block-diagonal, so the
//     fused smoother / transfer / defect kernels of the single solve run over all
//     of them in one launch; an instance enters the V-cycle at its own finest
//     level (its residual is copied in there) and its correction is copied out
//     after that level's post-smoothing;
//   * the SELL-64 launch shape of every batch matrix is picked by the autotuner;
//   * one PCG iteration of the whole batch is recorded as a HIP graph and
//     replayed until no instance is active.
// The smoother skip rule (iterative.hpp:761-766: skip when the level's right-hand
// side norm is <= eps^2) is not evaluated per instance inside the batch.
#pragma once

#include <feat3amd/batch_assembly.hpp>
#include <feat3amd/solver.hpp>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

namespace feat3amd
{
  namespace Solver
  {
    struct BatchConfig
    {
      Assembly::BatchSpec spec;
      int nu_pre = 4, nu_post = 4;
      double omega_jac = 0.5, omega_rich = 1.0;
      double tol_rel = 1e-8;
      Index max_iter = 100;
      bool autotune = true;
      bool graphs = true;
      /// levels 1..tail_levels of every instance run in one launch (feat_batch_tail_f64,
      /// one workgroup per instance, bit-identical); lowered to what the workgroup holds; 0: off
      int tail_levels = 5;
      /// the smoother calls of the levels above the tail as tiled launches on the structured
      /// grids (feat_batch_pre_tiled_f64 / feat_batch_post_tiled_f64, bit-identical)
      bool tiled = true;
      /// the tiled post-smoothing reads the prolongation's arrays even where it is bilinear injection
      bool tiled_prol_arrays = false;
      /// the restriction of a tiled level in its pre-smoothing launch (where it is the
      /// transpose of bilinear injection; else a SELL-64 launch of its own)
      bool tiled_restriction = true;
      /// the coarse tail in ELL form even where its levels are structured grids
      bool tail_ell = false;
    };

    class BatchedMgPcg
    {
      typedef LAFEM::SparseMatrixCSR<double, Index> Csr;
      typedef LAFEM::DenseVector<double, Index> Vec;

      // chunked segment structure of a back-to-back layout
      struct Segs
      {
        std::uint32_t ninst = 0, nchunks = 0;
        DeviceArray<std::uint64_t> off, lo, hi;
        DeviceArray<std::uint32_t> cinst, ichunk;
        DeviceArray<double> partials;
        feat_batch_t desc{};
        void build(const std::vector<std::uint64_t>& o)
        {
          ninst = std::uint32_t(o.size() - 1);
          std::vector<std::uint64_t> l, h;
          std::vector<std::uint32_t> ci, ic(ninst + 1, 0);
          for(std::uint32_t i = 0; i < ninst; ++i)
          {
            ic[i] = std::uint32_t(l.size());
            for(std::uint64_t a = o[i]; a < o[i + 1]; a += FEAT_BATCH_CHUNK)
            {
              l.push_back(a);
              h.push_back(std::min<std::uint64_t>(a + FEAT_BATCH_CHUNK, o[i + 1]));
              ci.push_back(i);
            }
          }
          ic[ninst] = std::uint32_t(l.size());
          nchunks = std::uint32_t(l.size());
          off = DeviceArray<std::uint64_t>::from_host(o.data(), o.size());
          lo = DeviceArray<std::uint64_t>::from_host(l.data(), l.size());
          hi = DeviceArray<std::uint64_t>::from_host(h.data(), h.size());
          cinst = DeviceArray<std::uint32_t>::from_host(ci.data(), ci.size());
          ichunk = DeviceArray<std::uint32_t>::from_host(ic.data(), ic.size());
          partials = DeviceArray<double>(nchunks);
          desc = feat_batch_t{ninst, nchunks, off.get(), cinst.get(), lo.get(), hi.get(), ichunk.get(), partials.get()};
        }
      };

      // segment copies between the PCG layout and one level's layout
      struct Copies
      {
        std::uint32_t n = 0;
        DeviceArray<std::uint64_t> lvl_lo, pcg_lo, len;
        void build(const std::vector<std::uint64_t>& a, const std::vector<std::uint64_t>& b,
                   const std::vector<std::uint64_t>& l)
        {
          n = std::uint32_t(l.size());
          if(!n) return;
          lvl_lo = DeviceArray<std::uint64_t>::from_host(a.data(), n);
          pcg_lo = DeviceArray<std::uint64_t>::from_host(b.data(), n);
          len = DeviceArray<std::uint64_t>::from_host(l.data(), n);
        }
      };

      struct Level
      {
        int k = 0;
        std::vector<int> inst;            // instances with finest level >= k
        std::vector<std::uint64_t> loff;  // their offsets in this layout
        Csr A, P, R;                      // P: from level k-1 to k; R = P^T
        Vec inv, rhs, sol, def, cor, tmp;
        DeviceArray<std::uint32_t> mask;
        Copies top;                       // instances whose finest level is k
      };

      BatchConfig _cfg;
      std::vector<int> _lev;
      std::vector<std::uint64_t> _off;
      int _lmax = 0;
      std::vector<std::unique_ptr<Level>> _lv;  // index k (1.._lmax); 0 unused
      Csr _A;                                    // PCG layout: each instance's finest matrix
      DeviceArray<std::uint32_t> _mask;
      Segs _segs;
      Vec _r, _p, _t, _z;
      DeviceArray<double> _gamma, _gold, _tp, _rr, _dinit, _dcur, _hist;
      DeviceArray<double> _ptp, _prr;  // chunk partials of t.p and r.r
      DeviceArray<std::uint64_t> _iters;
      DeviceArray<int> _status, _active, _nactive;
      FusedWorkspace _fw;
      // graph replay state
      feat_stream_t _stream = nullptr;
      feat_handle_t _gh = nullptr;
      feat_graph_t _graph = nullptr;
      int* _host_nactive = nullptr;
      const void* _key_x = nullptr;
      std::vector<double> _tune_us;
      // coarse tail: levels 1.._tail.nlevels in one launch (0 levels: off)
      struct Ell
      {
        std::uint32_t n = 0, w = 0;
        std::vector<double> val;
        std::vector<std::uint16_t> col;
        std::vector<std::uint8_t> len;
        bool ok = true;
        DeviceArray<double> dval;
        DeviceArray<std::uint16_t> dcol;
        DeviceArray<std::uint8_t> dlen;
        void upload()
        {
          dval = DeviceArray<double>::from_host(val.data(), val.size());
          dcol = DeviceArray<std::uint16_t>::from_host(col.data(), col.size());
          dlen = DeviceArray<std::uint8_t>::from_host(len.data(), len.size());
        }
      };
      struct TailLevel
      {
        Ell A, P, R;
        // grid form: nodes per side, 9-point stencil values (empty: not a 9-point grid),
        // transfers bilinear injection and its transpose
        std::uint32_t n1 = 0;
        std::vector<double> sv;
        bool bilinear = false;
        DeviceArray<double> dsv;
      };
      feat_batch_tail_t _tail{};
      std::vector<TailLevel> _tl;  // index k (1..T); 0 unused
      DeviceArray<int> _tfin;
      DeviceArray<std::uint64_t> _tloff, _tpcg;
      // tiled smoother calls of the levels above the tail (index k; empty sval: not tiled)
      struct TiledLevel
      {
        std::uint32_t n1 = 0;
        DeviceArray<double> sval;
        Ell P;
        DeviceArray<std::uint64_t> coff;
        DeviceArray<std::uint64_t> top;  // per position: PCG offset of an instance entering here, else ~0
        bool p_struct = false;           // the prolongation is bilinear injection (no P arrays read)
        bool r_struct = false;           // the restriction is its transpose: fused into the pre-smoothing
      };
      std::vector<TiledLevel> _tiled;

      // true when every row of the one-instance prolongation P1 (level with n1 nodes per side
      // from the level below) is bilinear injection: the coarse nodes (x/2 .. (x+1)/2,
      // y/2 .. (y+1)/2) in lexicographic order with the weights 1, 1/2, 1/4
      static bool bilinear_prolongation(const Assembly::CsrHost& P1, std::uint64_t n1)
      {
        const std::uint64_t nc1 = (n1 + 1) / 2;
        if(P1.rows != n1 * n1 || P1.rowptr.size() != P1.rows + 1) return false;
        for(std::uint64_t r = 0; r < P1.rows; ++r)
        {
          const std::uint64_t x = r % n1, y = r / n1, ox = x & 1, oy = y & 1;
          const std::uint64_t e0 = P1.rowptr[r], len = P1.rowptr[r + 1] - e0;
          if(len != (1 + ox) * (1 + oy)) return false;
          const double w = (ox && oy) ? 0.25 : ((ox || oy) ? 0.5 : 1.0);
          std::uint64_t u = 0;
          for(std::uint64_t iy = 0; iy <= oy; ++iy)
            for(std::uint64_t ix = 0; ix <= ox; ++ix, ++u)
              if(P1.col[e0 + u] != (y / 2 + iy) * nc1 + (x / 2 + ix) || P1.val[e0 + u] != w) return false;
        }
        return true;
      }

      // true when R1 (rows: the coarse grid, nc1 per side; columns: the fine grid, n1 per side)
      // is the transpose of bilinear injection: row (X, Y) holds the fine nodes
      // (2X + dx, 2Y + dy), dx, dy in {-1, 0, 1}, inside the grid, ascending, weights w(dx) w(dy)
      // with w(0) = 1, w(+-1) = 1/2
      static bool bilinear_restriction(const Assembly::CsrHost& R1, std::uint64_t n1)
      {
        const std::int64_t nc1 = std::int64_t(n1 + 1) / 2, w1 = std::int64_t(n1);
        if(R1.rows != std::uint64_t(nc1 * nc1) || R1.rowptr.size() != R1.rows + 1) return false;
        for(std::int64_t r = 0; r < nc1 * nc1; ++r)
        {
          const std::int64_t X = r % nc1, Y = r / nc1;
          std::uint64_t e = R1.rowptr[r];
          for(std::int64_t dy = -1; dy <= 1; ++dy)
            for(std::int64_t dx = -1; dx <= 1; ++dx)
            {
              const std::int64_t fx = 2 * X + dx, fy = 2 * Y + dy;
              if(fx < 0 || fx >= w1 || fy < 0 || fy >= w1) continue;
              if(e >= R1.rowptr[r + 1] || std::int64_t(R1.col[e]) != fy * w1 + fx ||
                 R1.val[e] != (dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0))
                return false;
              ++e;
            }
          if(e != R1.rowptr[r + 1]) return false;
        }
        return true;
      }

      // the nine stencil values of every free row of the diagonal blocks of A (block b: rows
      // b*n1*n1 ..), checked to be the full 9-point neighbourhood in CSR order; empty when not
      static std::vector<double> stencil_values(const Assembly::CsrHost& A, std::size_t nb, std::uint64_t n1,
                                                const std::vector<std::uint64_t>& bnd)
      {
        const std::uint64_t nk = n1 * n1;
        std::vector<char> filt(nk, 0);
        for(auto b : bnd) filt[b] = 1;
        std::vector<double> v(nb * 9 * nk, 0.0);
        const std::int64_t w = std::int64_t(n1);
        const std::int64_t offs[9] = {-w - 1, -w, -w + 1, -1, 0, 1, w - 1, w, w + 1};
        for(std::size_t b = 0; b < nb; ++b)
          for(std::uint64_t r = 0; r < nk; ++r)
          {
            if(filt[r]) continue;
            const std::uint64_t g = b * nk + r, e0 = A.rowptr[g];
            const std::int64_t x = std::int64_t(r % n1), y = std::int64_t(r / n1);
            if(A.rowptr[g + 1] - e0 != 9 || x == 0 || y == 0 || x == w - 1 || y == w - 1) return {};
            for(int u = 0; u < 9; ++u)
            {
              if(std::int64_t(A.col[e0 + u]) != std::int64_t(b * nk + r) + offs[u]) return {};
              v[(b * 9 + u) * nk + r] = A.val[e0 + u];
            }
          }
        return v;
      }

      // ELL form of the diagonal blocks of M: block b has rows roff[b] .. roff[b] + n - 1 and
      // columns from coff[b], the same pattern in every block; values per block (in block
      // order) or, when shared, one set equal (bitwise) in every block.  ok = false when the
      // blocks differ where they must agree or a block does not fit the limits.
      static Ell to_ell(const Assembly::CsrHost& M, const std::vector<std::uint64_t>& roff,
                        const std::vector<std::uint64_t>& coff, std::uint64_t n, std::uint64_t ncols, bool shared)
      {
        Ell E;
        E.n = std::uint32_t(n);
        const std::size_t nb = roff.size();
        for(std::size_t b = 0; b < nb; ++b)
          for(std::uint64_t j = 0; j < n; ++j)
            E.w = std::max<std::uint32_t>(E.w, std::uint32_t(M.rowptr[roff[b] + j + 1] - M.rowptr[roff[b] + j]));
        if(ncols > 65536 || E.w > 255)  // 16-bit columns, 8-bit row lengths
        {
          E.ok = false;
          return E;
        }
        E.col.assign(std::size_t(E.w) * n, 0);
        E.len.assign(n, 0);
        E.val.assign((shared ? 1 : nb) * std::size_t(E.w) * n, 0.0);
        for(std::size_t b = 0; b < nb && E.ok; ++b)
        {
          for(std::uint64_t j = 0; j < n; ++j)
          {
            const std::uint64_t g = roff[b] + j, e0 = M.rowptr[g];
            const std::uint64_t l = M.rowptr[g + 1] - e0;
            if(b == 0)
              E.len[j] = std::uint8_t(l);
            else if(E.len[j] != l)
              E.ok = false;
            for(std::uint64_t u = 0; u < l && E.ok; ++u)
            {
              const std::uint64_t c = M.col[e0 + u] - coff[b];
              if(M.col[e0 + u] < coff[b] || c >= ncols) E.ok = false;
              const std::size_t at = std::size_t(u) * n + j;
              if(b == 0)
                E.col[at] = std::uint16_t(c);
              else if(E.col[at] != c)
                E.ok = false;
              const double v = M.val[e0 + u];
              if(!shared)
                E.val[(b * E.w) * n + at] = v;
              else if(b == 0)
                E.val[at] = v;
              else if(std::memcmp(&E.val[at], &v, sizeof v) != 0)
                E.ok = false;
            }
          }
        }
        return E;
      }

      // the deepest T <= tail_levels whose levels fit one workgroup (rows, LDS vectors, ELL widths)
      void plan_tail()
      {
        _tail = feat_batch_tail_t{};
        int T = std::min<int>(int(_tl.size()) - 1, std::min(_lmax, FEAT_BATCH_TAIL_MAX_LEVELS));
        for(; T >= 1; --T)
        {
          std::uint64_t total = 0;
          bool ok = true;
          for(int k = 1; k <= T; ++k)
          {
            const TailLevel& t = _tl[k];
            total += t.A.n;
            ok = ok && t.A.ok && t.A.n <= (k < T ? 1024u : FEAT_BATCH_TAIL_MAX_ROWS) &&
                 t.A.w <= FEAT_BATCH_TAIL_MAX_ROW_LEN;
            if(k > 1)
              ok = ok && t.P.ok && t.R.ok && t.P.w <= FEAT_BATCH_TAIL_MAX_PROL_LEN &&
                   t.R.w <= FEAT_BATCH_TAIL_MAX_ROW_LEN;
          }
          if(ok && total <= FEAT_BATCH_TAIL_MAX_TOTAL) break;
        }
        if(T < 1)
        {
          _tl.clear();
          return;
        }
        _tl.resize(T + 1);
        const int ni = int(_lev.size());
        std::vector<std::uint64_t> loff(std::size_t(T) * ni, 0), pcg(ni);
        for(int k = 1; k <= T; ++k)
        {
          const Level& L = *_lv[k];
          for(std::size_t j = 0; j < L.inst.size(); ++j) loff[std::size_t(k - 1) * ni + L.inst[j]] = L.loff[j];
        }
        for(int i = 0; i < ni; ++i) pcg[i] = _off[i];
        _tfin = DeviceArray<int>::from_host(_lev.data(), _lev.size());
        _tloff = DeviceArray<std::uint64_t>::from_host(loff.data(), loff.size());
        _tpcg = DeviceArray<std::uint64_t>::from_host(pcg.data(), pcg.size());
        _tail.nlevels = T;
        _tail.ninst = std::uint32_t(ni);
        _tail.nu_pre = _cfg.nu_pre;
        _tail.nu_post = _cfg.nu_post;
        _tail.omega = _cfg.omega_rich;
        _tail.finest = _tfin.get();
        _tail.loff = _tloff.get();
        _tail.pcg_off = _tpcg.get();
        _tail.rhs_deep = _lv[T]->rhs.elements();
        _tail.sol_deep = _lv[T]->sol.elements();
        {
          bool grid = !_cfg.tail_ell;
          std::uint64_t below = 0;
          for(int k = 1; k <= T; ++k)
          {
            grid = grid && !_tl[k].sv.empty() && _tl[k].bilinear;
            if(k < T) below += _tl[k].A.n;
          }
          _tail.grid = (grid && below <= FEAT_BATCH_TAIL_MAX_BELOW) ? 1 : 0;
        }
        for(int k = 1; k <= T; ++k)
        {
          TailLevel& t = _tl[k];
          feat_batch_tail_level_t& d = _tail.lv[k - 1];
          if(_tail.grid)
          {
            t.dsv = DeviceArray<double>::from_host(t.sv.data(), t.sv.size());
            d.n1 = t.n1;
            d.sval = t.dsv.get();
          }
          t.sv.clear();
          t.A.upload();
          d.n = t.A.n;
          d.wa = t.A.w;
          d.a_val = t.A.dval.get();
          d.a_col = t.A.dcol.get();
          d.a_len = t.A.dlen.get();
          if(k > 1)
          {
            t.P.upload();
            t.R.upload();
            d.wp = t.P.w;
            d.p_val = t.P.dval.get();
            d.p_col = t.P.dcol.get();
            d.p_len = t.P.dlen.get();
            d.wr = t.R.w;
            d.r_val = t.R.dval.get();
            d.r_col = t.R.dcol.get();
            d.r_len = t.R.dlen.get();
          }
          d.inv = _lv[k]->inv.elements();
          d.mask = _lv[k]->mask.get();
          // the host copies are no longer needed
          t.A.val.clear();
          t.P.val.clear();
          t.R.val.clear();
        }
      }

      static std::vector<std::uint32_t> unit_mask(Index n, const std::vector<std::uint64_t>& idx)
      {
        std::vector<std::uint32_t> m((n + 31) / 32, 0u);
        for(auto i : idx) m[i >> 5] |= 1u << (i & 31);
        return m;
      }

      static void append(Assembly::CsrHost& B, const Assembly::CsrHost& A, std::uint64_t roff, std::uint64_t coff)
      {
        if(B.rowptr.empty()) B.rowptr.push_back(0);
        for(std::uint64_t r = 0; r < A.rows; ++r)
        {
          for(auto i = A.rowptr[r]; i < A.rowptr[r + 1]; ++i)
          {
            B.col.push_back(A.col[i] + coff);
            B.val.push_back(A.val[i]);
          }
          B.rowptr.push_back(B.col.size());
        }
        (void)roff;
      }

    public:
      explicit BatchedMgPcg(const BatchConfig& cfg) : _cfg(cfg) {}
      ~BatchedMgPcg()
      {
        feat_graph_destroy(_graph);
        if(_gh) feat_handle_destroy(_gh);
        if(_stream) feat_stream_destroy(_stream);
        if(_host_nactive) feat_host_free(_host_nactive);
      }
      BatchedMgPcg(const BatchedMgPcg&) = delete;
      BatchedMgPcg& operator=(const BatchedMgPcg&) = delete;

      const std::vector<int>& levels() const { return _lev; }
      const std::vector<std::uint64_t>& offsets() const { return _off; }
      Index size() const { return _off.back(); }
      const std::vector<double>& tuned_us() const { return _tune_us; }
      /// levels 1..tail_depth() run in one launch per V-cycle (0: none)
      int tail_depth() const { return _tail.nlevels; }
      /// the coarse tail runs in its grid form (structured levels, matrices prefetched)
      bool tail_grid() const { return _tail.grid != 0; }
      std::vector<int> tunes() const
      {
        std::vector<int> t;
        for(int k = 1; k <= _lmax; ++k)
        {
          t.push_back(_lv[k]->A.tune());
          if(k > 1)
          {
            t.push_back(_lv[k]->P.tune());
            t.push_back(_lv[k]->R.tune());
          }
        }
        t.push_back(_A.tune());
        return t;
      }

      /// host assembly of every instance and level, device upload, Jacobi diagonals, autotuning
      void init()
      {
        const auto& S = _cfg.spec;
        _lev = Assembly::batch_levels(S.seed, S.ninst, S.lmin, S.lmax);
        _lmax = *std::max_element(_lev.begin(), _lev.end());
        const int ni = S.ninst;
        std::vector<std::vector<double>> coef(ni);
        for(int i = 0; i < ni; ++i)
        {
          const std::uint64_t nc = std::uint64_t(1) << (2 * _lev[i]);
          coef[i] = Assembly::lognormal_cells(S.coef_seed(i), nc, S.sigma);
        }
        // PCG layout
        _off.assign(1, 0);
        Assembly::CsrHost top;
        std::vector<std::uint64_t> topfilt;
        for(int i = 0; i < ni; ++i)
        {
          const auto Ai = Assembly::q1_diffusion_2d(_lev[i], coef[i], _lev[i]);
          append(top, Ai, _off.back(), _off.back());
          for(auto b : Assembly::q1_boundary(2, _lev[i])) topfilt.push_back(b + _off.back());
          _off.push_back(_off.back() + Ai.rows);
        }
        top.rows = top.cols = _off.back();
        _A = Csr::from_host(top.rows, top.cols, top.val.data(), top.col.data(), top.rowptr.data());
        _mask = DeviceArray<std::uint32_t>::from_host(unit_mask(top.rows, topfilt).data(), (top.rows + 31) / 32);
        _segs.build(_off);
        _ptp = DeviceArray<double>(_segs.nchunks);
        _prr = DeviceArray<double>(_segs.nchunks);
        // multigrid levels 1.._lmax
        _lv.clear();
        _lv.resize(_lmax + 1);
        const int tcap = std::min(std::max(_cfg.tail_levels, 0), std::min(_lmax, FEAT_BATCH_TAIL_MAX_LEVELS));
        _tl.clear();
        _tl.resize(tcap + 1);
        _tiled.clear();
        _tiled.resize(_lmax + 1);
        for(int k = 1; k <= _lmax; ++k)
        {
          auto L = std::make_unique<Level>();
          L->k = k;
          Assembly::CsrHost Ak, Pk;
          std::vector<std::uint64_t> filt, clo, plo, len;
          const std::uint64_t nk = Assembly::ipow(Assembly::q1_nodes_1d(k), 2);
          const auto bnd = Assembly::q1_boundary(2, k);
          std::uint64_t o = 0;
          for(int i = 0; i < ni; ++i)
          {
            if(_lev[i] < k) continue;
            L->inst.push_back(i);
            L->loff.push_back(o);
            append(Ak, Assembly::q1_diffusion_2d(k, coef[i], _lev[i]), o, o);
            for(auto b : bnd) filt.push_back(b + o);
            if(_lev[i] == k)
            {
              clo.push_back(o);
              plo.push_back(_off[i]);
              len.push_back(nk);
            }
            o += nk;
          }
          Ak.rows = Ak.cols = o;
          if(k <= tcap)
          {
            _tl[k].A = to_ell(Ak, L->loff, L->loff, nk, nk, false);
            _tl[k].n1 = std::uint32_t(Assembly::q1_nodes_1d(k));
            _tl[k].sv = stencil_values(Ak, L->inst.size(), _tl[k].n1, bnd);
            _tl[k].bilinear = (k == 1);
          }
          const bool tiled_k = _cfg.tiled && k > tcap && k > 1 && _cfg.nu_pre <= FEAT_BATCH_TILED_MAX_STEPS &&
                               _cfg.nu_post <= FEAT_BATCH_TILED_MAX_STEPS;
          if(tiled_k)
          {
            const auto sv = stencil_values(Ak, L->inst.size(), Assembly::q1_nodes_1d(k), bnd);
            if(!sv.empty())
            {
              _tiled[k].n1 = std::uint32_t(Assembly::q1_nodes_1d(k));
              _tiled[k].sval = DeviceArray<double>::from_host(sv.data(), sv.size());
            }
          }
          L->A = Csr::from_host(o, o, Ak.val.data(), Ak.col.data(), Ak.rowptr.data());
          L->mask = DeviceArray<std::uint32_t>::from_host(unit_mask(o, filt).data(), (o + 31) / 32);
          L->top.build(clo, plo, len);
          for(Vec* v : {&L->inv, &L->rhs, &L->sol, &L->def, &L->cor, &L->tmp}) *v = Vec(o);
          // Jacobi: inv = omega / diag (coarse level: one Jacobi step with omega 1)
          L->A.extract_diag(L->inv);
          L->inv.component_invert(L->inv, k == 1 ? 1.0 : _cfg.omega_jac);
          if(k > 1)
          {
            // prolongation from level k-1: rows = this layout, columns = level k-1 layout
            const Level& C = *_lv[k - 1];
            const auto P1 = Assembly::q1_prolongation(2, k);
            for(std::size_t j = 0; j < L->inst.size(); ++j)
            {
              const int i = L->inst[j];
              const auto it = std::find(C.inst.begin(), C.inst.end(), i);
              append(Pk, P1, L->loff[j], C.loff[std::size_t(it - C.inst.begin())]);
            }
            Pk.rows = o;
            Pk.cols = C.A.rows();
            L->P = Csr::from_host(Pk.rows, Pk.cols, Pk.val.data(), Pk.col.data(), Pk.rowptr.data());
            const auto Rk = Assembly::transpose(Pk);
            if(k <= tcap)
            {
              std::vector<std::uint64_t> cl;  // each instance's offset on level k-1
              for(int i : L->inst) cl.push_back(C.loff[std::size_t(std::find(C.inst.begin(), C.inst.end(), i) - C.inst.begin())]);
              const std::uint64_t nc = C.A.rows() / C.inst.size();
              _tl[k].P = to_ell(Pk, L->loff, cl, nk, nc, true);
              _tl[k].R = to_ell(Rk, cl, L->loff, nc, nk, true);
              _tl[k].bilinear = bilinear_prolongation(P1, _tl[k].n1) &&
                                bilinear_restriction(Assembly::transpose(P1), _tl[k].n1);
            }
            if(_tiled[k].n1)
            {
              std::vector<std::uint64_t> cl;
              for(int i : L->inst) cl.push_back(C.loff[std::size_t(std::find(C.inst.begin(), C.inst.end(), i) - C.inst.begin())]);
              TiledLevel& tk = _tiled[k];
              tk.p_struct = !_cfg.tiled_prol_arrays && bilinear_prolongation(P1, tk.n1);
              tk.r_struct = _cfg.tiled_restriction && bilinear_prolongation(P1, tk.n1) &&
                            bilinear_restriction(Assembly::transpose(P1), tk.n1);
              tk.P = to_ell(Pk, L->loff, cl, nk, C.A.rows() / C.inst.size(), true);
              if(tk.P.ok && tk.P.w <= FEAT_BATCH_TAIL_MAX_PROL_LEN)
              {
                tk.P.upload();
                tk.P.val.clear();
                tk.coff = DeviceArray<std::uint64_t>::from_host(cl.data(), cl.size());
                std::vector<std::uint64_t> tp;
                for(int i : L->inst) tp.push_back(_lev[i] == k ? _off[i] : ~std::uint64_t(0));
                tk.top = DeviceArray<std::uint64_t>::from_host(tp.data(), tp.size());
              }
              else
                tk = TiledLevel{};
            }
            L->R = Csr::from_host(Rk.rows, Rk.cols, Rk.val.data(), Rk.col.data(), Rk.rowptr.data());
          }
          _lv[k] = std::move(L);
        }
        plan_tail();
        for(Vec* v : {&_r, &_p, &_t, &_z}) *v = Vec(size());
        for(DeviceArray<double>* a : {&_gamma, &_gold, &_tp, &_rr, &_dinit, &_dcur}) *a = DeviceArray<double>(ni);
        _hist = DeviceArray<double>(Index(ni) * (_cfg.max_iter + 1));
        _iters = DeviceArray<std::uint64_t>(ni);
        _status = DeviceArray<int>(ni);
        _active = DeviceArray<int>(ni);
        _nactive = DeviceArray<int>(1);
        _fw.create(4);
        // execution plans (and their launch shapes) before any recording
        _tune_us.clear();
        auto plan = [&](const Csr& m) {
          if(_cfg.autotune)
            _tune_us.push_back(m.autotune());
          else
            (void)m.exec();
        };
        for(int k = 1; k <= _lmax; ++k)
        {
          plan(_lv[k]->A);
          if(k > 1)
          {
            plan(_lv[k]->P);
            plan(_lv[k]->R);
          }
        }
        plan(_A);
        for(Vec* v : {&_r, &_p, &_t, &_z}) v->format();
        {
          void* hp = nullptr;
          check(feat_host_malloc(&hp, 64), "batch pinned");
          _host_nactive = static_cast<int*>(hp);
        }
        if(_cfg.graphs)
        {
          check(feat_stream_create(&_stream), "batch stream");
          check(feat_handle_create(&_gh, 0, _stream), "batch handle");
        }
        Runtime::synchronize();
      }

      feat_batch_tiled_t tiled_desc(int k, int nu, const double* r, double* z)
      {
        Level& L = *_lv[k];
        const TiledLevel& t = _tiled[k];
        feat_batch_tiled_t d{};
        d.n1 = t.n1;
        d.ninst = std::uint32_t(L.inst.size());
        d.nu = nu;
        d.omega = _cfg.omega_rich;
        d.sval = t.sval.get();
        d.inv = L.inv.elements();
        d.mask = L.mask.get();
        d.rhs = L.rhs.elements();
        d.sol = L.sol.elements();
        d.def = L.def.elements();
        d.xc = _lv[k - 1]->sol.elements();
        d.coff = t.coff.get();
        d.wp = t.P.w;
        d.p_val = t.P.dval.get();
        d.p_col = t.P.dcol.get();
        d.p_len = t.P.dlen.get();
        d.top = t.top.get();
        d.p_struct = t.p_struct ? 1 : 0;
        d.rest = t.r_struct ? 1 : 0;
        d.crhs = _lv[k - 1]->rhs.elements();
        d.cmask = _lv[k - 1]->mask.get();
        d.r = r;
        d.z = z;
        return d;
      }

    public:
      /// levels whose smoother calls run tiled
      std::vector<int> tiled_levels() const
      {
        std::vector<int> v;
        for(int k = 1; k < int(_tiled.size()); ++k)
          if(_tiled[k].n1) v.push_back(k);
        return v;
      }

      /// z = MG(r) for every instance (fused kernels of the single solve over block-diagonal levels)
      void mg_apply(double* z, const double* r)
      {
        feat_handle_t h = Runtime::handle();
        const feat_reduce_ws_t* ws = &_fw.ws;
        const int T = _tail.nlevels;
        for(int k = _lmax; k >= std::max(2, T + 1); --k)
        {
          Level& L = *_lv[k];
          // a tiled level reads the right-hand sides of its entering instances from r itself
          if(L.top.n && !_tiled[k].n1)
            check(feat_batch_gather_f64(h, L.top.n, L.top.lvl_lo.get(), L.top.pcg_lo.get(), L.top.len.get(),
                                        L.rhs.elements(), r),
                  "batch inject");
          if(_tiled[k].n1)
          {
            const feat_batch_tiled_t d = tiled_desc(k, _cfg.nu_pre, r, z);
            check(feat_batch_pre_tiled_f64(h, &d), "batch pre-smoothing (tiled)");
          }
          else
          {
            const feat_matrix_f64_t& A = L.A.exec();
            check(feat_rj_smooth_f64(h, &A, L.inv.elements(), L.mask.get(), _cfg.nu_pre, _cfg.omega_rich,
                                     L.rhs.elements(), nullptr, L.sol.elements(), L.tmp.elements(), L.def.elements(),
                                     nullptr, nullptr, nullptr, ws),
                  "batch pre-smoothing");
          }
          Level& C = *_lv[k - 1];
          if(!(_tiled[k].n1 && _tiled[k].r_struct))   // else done by the tiled pre-smoothing
          {
            const feat_matrix_f64_t& R = L.R.exec();
            check(feat_apply_masked_f64(h, &R, C.rhs.elements(), L.def.elements(), C.mask.get(), nullptr, nullptr,
                                        nullptr, ws),
                  "batch restriction");
          }
        }
        if(T > 0)
          check(feat_batch_tail_f64(h, &_tail, r, z), "batch coarse tail");
        else
        {
          Level& L = *_lv[1];
          if(L.top.n)
            check(feat_batch_gather_f64(h, L.top.n, L.top.lvl_lo.get(), L.top.pcg_lo.get(), L.top.len.get(),
                                        L.rhs.elements(), r),
                  "batch inject");
          const feat_matrix_f64_t& A = L.A.exec();
          check(feat_rj_smooth_f64(h, &A, L.inv.elements(), L.mask.get(), 1, 1.0, L.rhs.elements(), nullptr,
                                   L.sol.elements(), L.tmp.elements(), nullptr, nullptr, nullptr, nullptr, ws),
                "batch coarse Jacobi");
          if(L.top.n)
            check(feat_batch_gather_f64(h, L.top.n, L.top.pcg_lo.get(), L.top.lvl_lo.get(), L.top.len.get(), z,
                                        L.sol.elements()),
                  "batch extract");
        }
        for(int k = std::max(2, T + 1); k <= _lmax; ++k)
        {
          Level& L = *_lv[k];
          if(_tiled[k].n1)
          {
            const feat_batch_tiled_t d = tiled_desc(k, _cfg.nu_post, r, z);
            check(feat_batch_post_tiled_f64(h, &d), "batch prolongation + defect + post-smoothing (tiled)");
          }
          else
          {
            const feat_matrix_f64_t& P = L.P.exec();
            check(feat_prol_add_f64(h, &P, L.sol.elements(), _lv[k - 1]->sol.elements(), L.mask.get()),
                  "batch prolongation");
            const feat_matrix_f64_t& A = L.A.exec();
            check(feat_defect_f64(h, &A, L.def.elements(), L.rhs.elements(), L.sol.elements(), L.mask.get(), nullptr, ws),
                  "batch defect");
            check(feat_rj_smooth_f64(h, &A, L.inv.elements(), L.mask.get(), _cfg.nu_post, _cfg.omega_rich,
                                     L.def.elements(), nullptr, L.cor.elements(), L.tmp.elements(), nullptr,
                                     L.sol.elements(), nullptr, nullptr, ws),
                  "batch post-smoothing");
          }
          if(L.top.n && !_tiled[k].n1)
            check(feat_batch_gather_f64(h, L.top.n, L.top.pcg_lo.get(), L.top.lvl_lo.get(), L.top.len.get(), z,
                                        L.sol.elements()),
                  "batch extract");
        }
      }

      struct Result
      {
        std::vector<int> status;
        std::vector<std::uint64_t> iters;
        std::vector<double> def_init, def_final, hist;
      };

    private:
      feat_batch_state_t state() const
      {
        return feat_batch_state_t{_dinit.get(), _dcur.get(), _iters.get(), _status.get(), _active.get(), _hist.get(),
                                  _cfg.max_iter + 1, _nactive.get()};
      }
      feat_batch_stop_t stop() const
      {
        const double e = 2.220446049250313e-16;
        return feat_batch_stop_t{_cfg.tol_rel, 1.0 / (e * e), 1.0 / e, 1.0 / (e * e), 0, _cfg.max_iter};
      }

      // t = A p, t.p partials, update (gamma and t.p summed from their partials), status (and
      // gamma_old = gamma), nactive -> host, z = MG(r), r.z partials, direction
      void body(double* x, feat_handle_t h, int* host_nactive)
      {
        const feat_reduce_ws_t* ws = &_fw.ws;
        const feat_batch_t* b = &_segs.desc;
        const feat_batch_stop_t sp = stop();
        const feat_batch_state_t st = state();
        const feat_matrix_f64_t& A = _A.exec();
        // the r.z partials (gamma of this iteration) are in the batch's own partials
        // buffer: written by the previous iteration's (or the start's) r.z dot
        double* gz = _segs.partials.get();
        check(feat_apply_masked_f64(h, &A, _t.elements(), _p.elements(), _mask.get(), nullptr, nullptr, nullptr, ws), "batch A*p");
        check(feat_batch_dot_partials_f64(h, b, _t.elements(), _p.elements(), _ptp.get()), "batch t.p");
        check(feat_batch_pcg_update_partials_f64(h, b, x, _r.elements(), _p.elements(), _t.elements(), gz, _ptp.get(),
                                                 _active.get(), _prr.get()),
              "batch update");
        check(feat_batch_status_partials_f64(h, b, _prr.get(), &sp, &st, _gamma.get(), gz), "batch status");
        check(feat_copy_int_to_host(h, _nactive.get(), host_nactive), "batch nactive");
        mg_apply(_z.elements(), _r.elements());
        check(feat_batch_dot_partials_f64(h, b, _r.elements(), _z.elements(), gz), "batch gamma");
        check(feat_batch_pcg_direction_partials_f64(h, b, _p.elements(), _z.elements(), gz, _gamma.get(), _active.get()),
              "batch direction");
      }

    public:
      /// x (initial guesses, overwritten) and b in the PCG layout (device)
      Result solve(double* x, const double* b)
      {
        feat_handle_t caller = Runtime::handle();
        const bool use_graph = _cfg.graphs && _gh;
        feat_handle_t h = use_graph ? _gh : caller;
        Runtime::Use scope(h);
        if(use_graph) check(feat_handle_wait(h, caller), "batch: order after caller");
        const feat_reduce_ws_t* ws = &_fw.ws;
        const feat_batch_t* bd = &_segs.desc;
        const feat_batch_stop_t sp = stop();
        const feat_batch_state_t st = state();
        int* hn = _host_nactive;
        // r = b - A x, initial defects (pcg.hpp:159-173, iterative.hpp:748-770)
        const feat_matrix_f64_t& A = _A.exec();
        check(feat_defect_f64(h, &A, _r.elements(), b, x, _mask.get(), nullptr, ws), "batch defect");
        check(feat_batch_dot_f64(h, bd, _r.elements(), _r.elements(), _rr.get()), "batch r.r");
        check(feat_batch_status_f64(h, _segs.ninst, _rr.get(), &sp, &st, 1), "batch status");
        // z = MG(r), gamma = r.z, p = z
        mg_apply(_z.elements(), _r.elements());
        check(feat_batch_dot_f64(h, bd, _r.elements(), _z.elements(), _gamma.get()), "batch gamma");
        check(feat_batch_pcg_direction_f64(h, bd, _p.elements(), _z.elements(), nullptr, nullptr, _active.get(), 1),
              "batch direction");
        if(use_graph)
        {
          check(feat_copy_int_to_host(h, _nactive.get(), hn), "batch nactive");
          check(feat_sync(h), "batch sync");
          if(_key_x != x)
          {
            feat_graph_destroy(_graph);
            _graph = nullptr;
            _key_x = x;
          }
          if(!_graph)
          {
            check(feat_graph_begin(h), "batch graph begin");
            try
            {
              body(x, h, hn);
            }
            catch(...)
            {
              feat_graph_t g = nullptr;
              feat_graph_end(h, &g);
              feat_graph_destroy(g);
              throw;
            }
            check(feat_graph_end(h, &_graph), "batch graph end");
          }
          while(*hn > 0)
          {
            check(feat_graph_launch(h, _graph), "batch graph launch");
            check(feat_sync(h), "batch sync");
          }
          check(feat_handle_wait(caller, h), "batch: order before caller");
        }
        else
        {
          check(feat_copy_int_to_host(h, _nactive.get(), hn), "batch nactive");
          check(feat_sync(h), "batch sync");
          while(*hn > 0)
          {
            body(x, h, hn);
            check(feat_sync(h), "batch sync");
          }
        }
        Result R;
        const Index ni = _segs.ninst;
        R.status.resize(ni);
        R.iters.resize(ni);
        R.def_init.resize(ni);
        R.def_final.resize(ni);
        R.hist.resize(ni * (_cfg.max_iter + 1));
        _status.to_host(R.status.data());
        _iters.to_host(R.iters.data());
        _dinit.to_host(R.def_init.data());
        _dcur.to_host(R.def_final.data());
        _hist.to_host(R.hist.data());
        return R;
      }
    };
  } // namespace Solver
} // namespace feat3amd
