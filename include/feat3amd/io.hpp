// feat3amd/io.hpp -- LAFEM file formats for CSR matrices and dense vectors (host).
//
// The files FEAT3's containers read and write (SparseMatrixCSR::read_from /
// write_out, lafem/sparse_matrix_csr.hpp:877-1087; DenseVector::read_from /
// write_out, lafem/dense_vector.hpp:549-723), byte for byte where a file is
// written:
//
//   fm_csr / fm_binary  the container image of Container::_serialize
//                       (lafem/container.hpp:356-586), uncompressed:
//       u64 file size, u64 FileMode, u64 value-type hash, u64 index-type hash,
//       u64 #element arrays, #index arrays, #element sizes, #index sizes,
//       #scalar indices, #scalar values, u64 compression (0x11 = off),
//       element sizes, element byte counts, index sizes, index byte counts,
//       scalar indices, scalar values, element arrays (f64), index arrays (u64),
//       16 zero bytes.
//     A CSR matrix holds one element array (values), two index arrays (columns,
//     row pointers) and the scalar indices {rows*cols, rows, cols, nnz}.
//   fm_dv / fm_binary   the same image of a vector: one element array, scalar
//                       index {size}.
//   fm_mtx              MatrixMarket text: "coordinate real general" (every entry)
//                       or "coordinate real symmetric" (entries with row >= col)
//                       for matrices, "array real general" (n x 1) for vectors;
//                       values as printf("%.6e"), the reference's std::scientific.
//   fm_exp              one value per line, "%.6e" (vectors only).
//
// Reading accepts what the reference reads and rejects, with an Error, what it
// would abort on (wrong header, compressed images) or silently mis-store: a row or
// column index outside the declared size, a truncated image, a vector file with a
// number of values other than declared.  Matrix rows without entries get correct row
// pointers (the reference's reader numbers only the rows that have entries); a
// repeated (row, column) keeps its first value, as the reference's map insert does,
// and is counted once.  fm_mtx and fm_exp store 7 significant digits, as the
// reference's writers do; fm_csr / fm_dv are exact.
#pragma once

#include <feat3amd/assembly.hpp>
#include <feat3amd/lafem.hpp>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace feat3amd
{
  namespace LAFEM
  {
    /// FileMode (lafem/base.hpp:31-47): the numeric values are part of the binary format
    enum class FileMode : int
    {
      fm_exp = 0,
      fm_dv,
      fm_mtx,
      fm_ell,
      fm_csr,
      fm_coo,
      fm_bm,
      fm_dm,
      fm_sv,
      fm_svb,
      fm_dvb,
      fm_bcsr,
      fm_cscr,
      fm_binary
    };

    namespace IO
    {
      // type hashes: byte size | integral << 32 | floating << 33 | signed << 34 (util/type_traits.hpp:35-56)
      constexpr std::uint64_t kHashF64 = 8ull | (1ull << 33) | (1ull << 34);
      constexpr std::uint64_t kHashU64 = 8ull | (1ull << 32);
      constexpr std::uint64_t kUncompressed = 0x11;   // elements_off | indices_off

      /// arrays of one container image
      struct Image
      {
        std::vector<std::vector<double>> elements;
        std::vector<std::vector<std::uint64_t>> indices;
        std::vector<std::uint64_t> scalar_index;
      };

      inline std::vector<char> encode(FileMode mode, const Image& im)
      {
        std::vector<std::uint64_t> head;
        const std::uint64_t ne = im.elements.size(), ni = im.indices.size();
        head.insert(head.end(), {0, std::uint64_t(mode), kHashF64, kHashU64, ne, ni, ne, ni,
                                 std::uint64_t(im.scalar_index.size()), 0, kUncompressed});
        for(const auto& e : im.elements) head.push_back(e.size());
        for(const auto& e : im.elements) head.push_back(8 * e.size());
        for(const auto& i : im.indices) head.push_back(i.size());
        for(const auto& i : im.indices) head.push_back(8 * i.size());
        head.insert(head.end(), im.scalar_index.begin(), im.scalar_index.end());
        std::uint64_t bytes = 8 * head.size();
        for(const auto& e : im.elements) bytes += 8 * e.size();
        for(const auto& i : im.indices) bytes += 8 * i.size();
        bytes += 16;
        head[0] = bytes;
        std::vector<char> out(bytes, 0);
        char* p = out.data();
        auto put = [&p](const void* src, std::size_t n) {
          if(n) std::memcpy(p, src, n);
          p += n;
        };
        put(head.data(), 8 * head.size());
        for(const auto& e : im.elements) put(e.data(), 8 * e.size());
        for(const auto& i : im.indices) put(i.data(), 8 * i.size());
        return out;
      }

      inline Image decode(FileMode mode, const std::vector<char>& buf)
      {
        const std::uint64_t nbytes = buf.size();
        std::uint64_t pos = 0;
        auto word = [&](const char* what) {
          if(pos + 8 > nbytes) throw Error(std::string("LAFEM binary file truncated (") + what + ")");
          std::uint64_t w;
          std::memcpy(&w, buf.data() + pos, 8);
          pos += 8;
          return w;
        };
        const std::uint64_t size = word("size"), magic = word("mode");
        if(size > nbytes) throw Error("LAFEM binary file shorter than its declared size");
        if(magic != std::uint64_t(mode))
          throw Error("LAFEM binary file holds FileMode " + std::to_string(magic) + ", expected " +
                      std::to_string(int(mode)));
        const std::uint64_t hdt = word("value type"), hit = word("index type");
        if((hdt & 0xFFFFFFFFull) != 8 || ((hdt >> 32) & 7) != 6)
          throw Error("LAFEM binary file: values are not 64-bit floating point");
        if((hit & 0xFFFFFFFFull) != 8 || ((hit >> 32) & 7) != 1)
          throw Error("LAFEM binary file: indices are not 64-bit unsigned integers");
        const std::uint64_t ne = word("counts"), ni = word("counts"), nes = word("counts"), nis = word("counts");
        const std::uint64_t nsi = word("counts"), nsd = word("counts"), comp = word("compression");
        if(comp != kUncompressed)
          throw Error("LAFEM binary file is compressed (zlib/zfp); only uncompressed images are supported");
        if(ne != nes || ni != nis || ne > 64 || ni > 64 || nsi > 64 || nsd > 64)
          throw Error("LAFEM binary file: inconsistent array counts");
        Image im;
        std::vector<std::uint64_t> esz(ne), isz(ni);
        for(auto& v : esz) v = word("element sizes");
        for(std::uint64_t i = 0; i < ne; ++i)
          if(word("element bytes") != 8 * esz[i]) throw Error("LAFEM binary file: element byte count mismatch");
        for(auto& v : isz) v = word("index sizes");
        for(std::uint64_t i = 0; i < ni; ++i)
          if(word("index bytes") != 8 * isz[i]) throw Error("LAFEM binary file: index byte count mismatch");
        for(std::uint64_t i = 0; i < nsi; ++i) im.scalar_index.push_back(word("scalar indices"));
        pos += 8 * nsd;   // scalar values: none of the containers here stores one
        auto take = [&](auto& dst, std::uint64_t n) {
          if(n > (size - std::min(size, pos)) / 8) throw Error("LAFEM binary file truncated (arrays)");
          dst.resize(n);
          if(n) std::memcpy(dst.data(), buf.data() + pos, 8 * n);
          pos += 8 * n;
        };
        im.elements.resize(ne);
        for(std::uint64_t i = 0; i < ne; ++i) take(im.elements[i], esz[i]);
        im.indices.resize(ni);
        for(std::uint64_t i = 0; i < ni; ++i) take(im.indices[i], isz[i]);
        return im;
      }

      inline std::vector<char> read_all(std::istream& in)
      {
        return std::vector<char>(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
      }

      inline std::string fmt_value(double v)
      {
        char b[64];
        std::snprintf(b, sizeof(b), "%.6e", v);
        return b;
      }

      // space/tab separated tokens of one line
      inline std::vector<std::string> tokens(const std::string& line)
      {
        std::vector<std::string> t;
        std::size_t i = 0;
        while(i < line.size())
        {
          while(i < line.size() && (line[i] == ' ' || line[i] == '\t' || line[i] == '\r')) ++i;
          std::size_t j = i;
          while(j < line.size() && line[j] != ' ' && line[j] != '\t' && line[j] != '\r') ++j;
          if(j > i) t.push_back(line.substr(i, j - i));
          i = j;
        }
        return t;
      }

      inline std::uint64_t to_index(const std::string& s, const char* what)
      {
        char* end = nullptr;
        const unsigned long long v = std::strtoull(s.c_str(), &end, 10);
        if(end == s.c_str() || s[0] == '-') throw Error(std::string("MatrixMarket file: bad ") + what + " '" + s + "'");
        return std::uint64_t(v);
      }

      // header line, then the first line that is not a comment (the size line)
      inline std::vector<std::string> mtx_sizes(std::istream& in, const std::string& header, bool* symmetric)
      {
        std::string line;
        if(!std::getline(in, line)) throw Error("MatrixMarket file is empty");
        if(symmetric)
        {
          const bool gen = line.find(header + " general") != std::string::npos;
          const bool sym = line.find(header + " symmetric") != std::string::npos;
          if(!gen && !sym) throw Error("not a '" + header + " general/symmetric' MatrixMarket file");
          *symmetric = sym;
        }
        else if(line.find(header + " general") == std::string::npos)
          throw Error("not a '" + header + " general' MatrixMarket file");
        while(std::getline(in, line))
        {
          auto t = tokens(line);
          if(t.empty() || t[0][0] == '%') continue;
          return t;
        }
        throw Error("MatrixMarket file has no size line");
      }
    } // namespace IO

    /// SparseMatrixCSR::write_out (sparse_matrix_csr.hpp:1038-1087)
    inline void write_csr(FileMode mode, std::ostream& out, const Assembly::CsrHost& A, bool symmetric = false)
    {
      if(A.rowptr.size() != A.rows + 1 || A.col.size() != A.rowptr[A.rows] || A.val.size() != A.col.size())
        throw Error("write_csr: inconsistent CSR arrays");
      if(mode == FileMode::fm_csr || mode == FileMode::fm_binary)
      {
        IO::Image im;
        if(A.rows * A.cols != 0 || !A.val.empty())
        {
          im.elements.push_back(A.val);
          im.indices.push_back(A.col);
          im.indices.push_back(A.rowptr);
        }
        im.scalar_index = {A.rows * A.cols, A.rows, A.cols, std::uint64_t(A.val.size())};
        const auto bytes = IO::encode(FileMode::fm_csr, im);
        out.write(bytes.data(), std::streamsize(bytes.size()));
      }
      else if(mode == FileMode::fm_mtx)
      {
        std::uint64_t cnt = 0;
        for(std::uint64_t r = 0; r < A.rows; ++r)
          for(std::uint64_t k = A.rowptr[r]; k < A.rowptr[r + 1]; ++k)
            cnt += (!symmetric || r >= A.col[k]);
        out << "%%MatrixMarket matrix coordinate real " << (symmetric ? "symmetric" : "general") << "\n";
        out << A.rows << " " << A.cols << " " << cnt << "\n";
        for(std::uint64_t r = 0; r < A.rows; ++r)
          for(std::uint64_t k = A.rowptr[r]; k < A.rowptr[r + 1]; ++k)
            if(!symmetric || r >= A.col[k]) out << r + 1 << " " << A.col[k] + 1 << " " << IO::fmt_value(A.val[k]) << "\n";
      }
      else
        throw Error("write_csr: file mode " + std::to_string(int(mode)) + " not supported for CSR matrices");
      if(!out.good()) throw Error("write_csr: write failed");
    }

    /// SparseMatrixCSR::read_from (sparse_matrix_csr.hpp:877-1026)
    inline Assembly::CsrHost read_csr(FileMode mode, std::istream& in)
    {
      Assembly::CsrHost A;
      if(mode == FileMode::fm_csr || mode == FileMode::fm_binary)
      {
        const IO::Image im = IO::decode(FileMode::fm_csr, IO::read_all(in));
        if(im.scalar_index.size() != 4) throw Error("read_csr: not a CSR image");
        A.rows = im.scalar_index[1];
        A.cols = im.scalar_index[2];
        const std::uint64_t nnz = im.scalar_index[3];
        if(im.elements.empty() && nnz == 0)
        {
          A.rowptr.assign(A.rows + 1, 0);
          return A;
        }
        if(im.elements.size() != 1 || im.indices.size() != 2 || im.elements[0].size() != nnz ||
           im.indices[0].size() != nnz || im.indices[1].size() != A.rows + 1)
          throw Error("read_csr: CSR image array sizes do not match its dimensions");
        A.val = im.elements[0];
        A.col = im.indices[0];
        A.rowptr = im.indices[1];
        if(A.rowptr[0] != 0 || A.rowptr[A.rows] != nnz) throw Error("read_csr: bad row pointers");
        for(std::uint64_t r = 0; r < A.rows; ++r)
          if(A.rowptr[r] > A.rowptr[r + 1]) throw Error("read_csr: bad row pointers");
        for(auto c : A.col)
          if(c >= A.cols) throw Error("read_csr: column index out of range");
        return A;
      }
      if(mode != FileMode::fm_mtx)
        throw Error("read_csr: file mode " + std::to_string(int(mode)) + " not supported for CSR matrices");
      bool symmetric = false;
      const auto sz = IO::mtx_sizes(in, "%%MatrixMarket matrix coordinate real", &symmetric);
      if(sz.size() < 2) throw Error("MatrixMarket file: bad size line");
      A.rows = IO::to_index(sz[0], "row count");
      A.cols = IO::to_index(sz[1], "column count");
      struct Ent
      {
        std::uint64_t r, c, seq;
        double v;
      };
      std::vector<Ent> e;
      std::string line;
      while(std::getline(in, line))
      {
        const auto t = IO::tokens(line);
        if(t.empty() || t[0][0] == '%') continue;
        if(t.size() < 3) throw Error("MatrixMarket file: bad entry line '" + line + "'");
        const std::uint64_t r = IO::to_index(t[0], "row"), c = IO::to_index(t[1], "column");
        if(r < 1 || r > A.rows || c < 1 || c > A.cols) throw Error("MatrixMarket file: entry outside the matrix");
        const double v = std::strtod(t[2].c_str(), nullptr);
        e.push_back({r - 1, c - 1, e.size(), v});
        if(symmetric && r != c) e.push_back({c - 1, r - 1, e.size(), v});
      }
      // ordered by (row, column); the first of a repeated position wins
      std::sort(e.begin(), e.end(), [](const Ent& a, const Ent& b) {
        return a.r != b.r ? a.r < b.r : (a.c != b.c ? a.c < b.c : a.seq < b.seq);
      });
      A.rowptr.assign(A.rows + 1, 0);
      for(std::size_t i = 0; i < e.size(); ++i)
      {
        if(i > 0 && e[i].r == e[i - 1].r && e[i].c == e[i - 1].c) continue;
        A.col.push_back(e[i].c);
        A.val.push_back(e[i].v);
        ++A.rowptr[e[i].r + 1];
      }
      for(std::uint64_t r = 0; r < A.rows; ++r) A.rowptr[r + 1] += A.rowptr[r];
      return A;
    }

    /// DenseVector::write_out (dense_vector.hpp:684-723)
    inline void write_dv(FileMode mode, std::ostream& out, const double* v, std::uint64_t n)
    {
      if(mode == FileMode::fm_dv || mode == FileMode::fm_binary)
      {
        IO::Image im;
        if(n) im.elements.push_back(std::vector<double>(v, v + n));
        im.scalar_index = {n};
        const auto bytes = IO::encode(FileMode::fm_dv, im);
        out.write(bytes.data(), std::streamsize(bytes.size()));
      }
      else if(mode == FileMode::fm_mtx)
      {
        out << "%%MatrixMarket matrix array real general\n" << n << " 1\n";
        for(std::uint64_t i = 0; i < n; ++i) out << IO::fmt_value(v[i]) << "\n";
      }
      else if(mode == FileMode::fm_exp)
      {
        for(std::uint64_t i = 0; i < n; ++i) out << IO::fmt_value(v[i]) << "\n";
      }
      else
        throw Error("write_dv: file mode " + std::to_string(int(mode)) + " not supported for vectors");
      if(!out.good()) throw Error("write_dv: write failed");
    }

    /// DenseVector::read_from (dense_vector.hpp:549-659)
    inline std::vector<double> read_dv(FileMode mode, std::istream& in)
    {
      std::vector<double> v;
      if(mode == FileMode::fm_dv || mode == FileMode::fm_binary)
      {
        const IO::Image im = IO::decode(FileMode::fm_dv, IO::read_all(in));
        if(im.scalar_index.size() != 1 || !im.indices.empty() || im.elements.size() > 1)
          throw Error("read_dv: not a vector image");
        const std::uint64_t n = im.scalar_index[0];
        if(n == 0 && im.elements.empty()) return v;
        if(im.elements.size() != 1 || im.elements[0].size() != n) throw Error("read_dv: image size mismatch");
        return im.elements[0];
      }
      std::string line;
      if(mode == FileMode::fm_mtx)
      {
        const auto sz = IO::mtx_sizes(in, "%%MatrixMarket matrix array real", nullptr);
        if(sz.size() < 2 || IO::to_index(sz[1], "column count") != 1)
          throw Error("MatrixMarket file is not an n x 1 array (dense vector)");
        const std::uint64_t n = IO::to_index(sz[0], "row count");
        while(std::getline(in, line))
        {
          const auto t = IO::tokens(line);
          if(t.empty()) continue;
          v.push_back(std::strtod(t[0].c_str(), nullptr));
        }
        if(v.size() != n)
          throw Error("MatrixMarket vector file holds " + std::to_string(v.size()) + " values, declares " + std::to_string(n));
        return v;
      }
      if(mode == FileMode::fm_exp)
      {
        // one value per line; lines holding '#' are comments
        while(std::getline(in, line))
        {
          if(line.find('#') != std::string::npos) continue;
          if(in.eof() && IO::tokens(line).empty()) break;
          const auto t = IO::tokens(line);
          v.push_back(t.empty() ? 0.0 : std::strtod(t[0].c_str(), nullptr));
        }
        return v;
      }
      throw Error("read_dv: file mode " + std::to_string(int(mode)) + " not supported for vectors");
    }

    inline std::ios_base::openmode io_openmode(FileMode mode, bool write)
    {
      const bool text = mode == FileMode::fm_mtx || mode == FileMode::fm_exp;
      return (write ? std::ios_base::out : std::ios_base::in) | (text ? std::ios_base::openmode() : std::ios_base::binary);
    }

    inline void write_csr(FileMode mode, const std::string& filename, const Assembly::CsrHost& A, bool symmetric = false)
    {
      std::ofstream f(filename, io_openmode(mode, true));
      if(!f.is_open()) throw Error("Unable to open matrix file " + filename);
      write_csr(mode, f, A, symmetric);
    }
    inline Assembly::CsrHost read_csr(FileMode mode, const std::string& filename)
    {
      std::ifstream f(filename, io_openmode(mode, false));
      if(!f.is_open()) throw Error("Unable to open matrix file " + filename);
      return read_csr(mode, f);
    }
    inline void write_dv(FileMode mode, const std::string& filename, const double* v, std::uint64_t n)
    {
      std::ofstream f(filename, io_openmode(mode, true));
      if(!f.is_open()) throw Error("Unable to open vector file " + filename);
      write_dv(mode, f, v, n);
    }
    inline std::vector<double> read_dv(FileMode mode, const std::string& filename)
    {
      std::ifstream f(filename, io_openmode(mode, false));
      if(!f.is_open()) throw Error("Unable to open vector file " + filename);
      return read_dv(mode, f);
    }

    /// write_out / read_from of the device containers: the arrays pass through host memory
    inline void write_out(const SparseMatrixCSR<double, std::uint64_t>& A, FileMode mode, const std::string& filename,
                          bool symmetric = false)
    {
      Assembly::CsrHost h;
      h.rows = A.rows();
      h.cols = A.columns();
      h.val.resize(A.used_elements());
      h.col.resize(A.used_elements());
      h.rowptr.resize(A.rows() + 1);
      DeviceArray<double>::wrap(const_cast<double*>(A.val()), A.used_elements()).to_host(h.val.data());
      DeviceArray<std::uint64_t>::wrap(const_cast<std::uint64_t*>(A.col_ind()), A.used_elements()).to_host(h.col.data());
      DeviceArray<std::uint64_t>::wrap(const_cast<std::uint64_t*>(A.row_ptr()), A.rows() + 1).to_host(h.rowptr.data());
      write_csr(mode, filename, h, symmetric);
    }
    inline SparseMatrixCSR<double, std::uint64_t> read_matrix(FileMode mode, const std::string& filename)
    {
      const Assembly::CsrHost h = read_csr(mode, filename);
      return SparseMatrixCSR<double, std::uint64_t>::from_host(h.rows, h.cols, h.val.data(), h.col.data(), h.rowptr.data());
    }
    inline void write_out(const DenseVector<double, std::uint64_t>& x, FileMode mode, const std::string& filename)
    {
      const std::vector<double> h = x.to_host();
      write_dv(mode, filename, h.data(), h.size());
    }
    inline DenseVector<double, std::uint64_t> read_vector(FileMode mode, const std::string& filename)
    {
      return DenseVector<double, std::uint64_t>::from_host(read_dv(mode, filename));
    }
  } // namespace LAFEM
} // namespace feat3amd
