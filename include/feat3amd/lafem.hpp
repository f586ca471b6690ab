// feat3amd/lafem.hpp -- device-resident LAFEM containers on top of feat3amd.h.
//
// Mirrors the member-function API of the reference's containers
// (kernel/lafem/dense_vector.hpp, sparse_matrix_csr.hpp, unit_filter.hpp,
// transfer.hpp).  Every arithmetic member calls one C-ABI entry point; no
// arithmetic happens on the host.
#pragma once

#include <feat3amd.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace feat3amd
{
  typedef std::uint64_t Index;

  class Error : public std::runtime_error
  {
  public:
    explicit Error(const std::string& s) : std::runtime_error(s) {}
  };

  inline void check(int rc, const char* what)
  {
    if(rc != FEAT_OK)
      throw Error(std::string(what) + ": " + feat_last_error());
  }

  /// Process-wide device handle (kernel/runtime.cpp:37-80 analogue): one HIP
  /// stream and the reduction workspace of the C-ABI.
  class Runtime
  {
  public:
    static feat_handle_t& slot()
    {
      static feat_handle_t h = nullptr;
      return h;
    }
    static void initialize(int device = 0, feat_stream_t stream = nullptr)
    {
      if(!slot())
        check(feat_handle_create(&slot(), device, stream), "Runtime::initialize");
    }
    static void finalize()
    {
      if(slot())
      {
        feat_handle_destroy(slot());
        slot() = nullptr;
      }
    }
    /// handle every container/solver call of this thread currently launches on
    static feat_handle_t handle()
    {
      if(override_slot()) return override_slot();
      if(!slot()) initialize(0);
      return slot();
    }
    static feat_handle_t& override_slot()
    {
      static thread_local feat_handle_t h = nullptr;
      return h;
    }
    /// flags of the SELL-64 plans built from now on (FEAT_SELL_NO_DICT: no pair dictionary)
    static int& sell_flags()
    {
      static int f = 0;
      return f;
    }
    /// multigrid grid forms where the levels qualify: d-linear transfers as FEAT_FMT_DLINEAR
    /// descriptors, the fused coarse levels' stencils and transfers formed in the kernel
    /// (off: the plans and CSR transfers are read)
    static bool& grid_forms()
    {
      static bool on = true;
      return on;
    }
    /// Scope in which this thread's calls launch on `h` (graph recording on a solver stream).
    class Use
    {
      feat_handle_t _prev;

    public:
      explicit Use(feat_handle_t h) : _prev(override_slot()) { override_slot() = h; }
      ~Use() { override_slot() = _prev; }
      Use(const Use&) = delete;
      Use& operator=(const Use&) = delete;
    };
    static void synchronize() { check(feat_sync(handle()), "Runtime::synchronize"); }
  };

  /// Owning or non-owning device array.
  template<typename T_>
  class DeviceArray
  {
    std::shared_ptr<T_> _p;
    Index _n = 0;

  public:
    DeviceArray() = default;
    explicit DeviceArray(Index n) : _n(n)
    {
      void* p = nullptr;
      check(feat_malloc(&p, sizeof(T_) * (n ? n : 1)), "DeviceArray");
      _p = std::shared_ptr<T_>(static_cast<T_*>(p), [](T_* q) { feat_free(q); });
    }
    static DeviceArray wrap(T_* p, Index n)
    {
      DeviceArray a;
      a._p = std::shared_ptr<T_>(p, [](T_*) {});
      a._n = n;
      return a;
    }
    static DeviceArray from_host(const T_* h, Index n)
    {
      DeviceArray a(n);
      if(n) check(feat_copy(Runtime::handle(), a.get(), h, sizeof(T_) * n), "DeviceArray::from_host");
      return a;
    }
    void to_host(T_* h) const
    {
      if(_n)
      {
        check(feat_copy(Runtime::handle(), h, get(), sizeof(T_) * _n), "DeviceArray::to_host");
        Runtime::synchronize();
      }
    }
    T_* get() const { return _p.get(); }
    Index size() const { return _n; }
    bool empty() const { return !_p; }
  };

  namespace LAFEM
  {
    enum class CloneMode { Shallow, Layout, Deep };

    /// DenseVector<DT_, IT_> (lafem/dense_vector.hpp:58) on the device.
    template<typename DT_, typename IT_ = Index>
    class DenseVector
    {
      DeviceArray<DT_> _a;

    public:
      typedef DT_ DataType;
      typedef IT_ IndexType;

      DenseVector() = default;
      explicit DenseVector(Index n) : _a(n) {}
      DenseVector(Index n, DT_ value) : _a(n) { format(value); }
      static DenseVector from_host(const DT_* p, Index n)
      {
        DenseVector v;
        v._a = DeviceArray<DT_>::from_host(p, n);
        return v;
      }
      static DenseVector from_host(const std::vector<DT_>& v) { return from_host(v.data(), v.size()); }
      static DenseVector wrap(DT_* dev, Index n)
      {
        DenseVector v;
        v._a = DeviceArray<DT_>::wrap(dev, n);
        return v;
      }
      std::vector<DT_> to_host() const
      {
        std::vector<DT_> h(size());
        _a.to_host(h.data());
        return h;
      }
      Index size() const { return _a.size(); }
      DT_* elements() { return _a.get(); }
      const DT_* elements() const { return _a.get(); }
      bool empty() const { return _a.empty(); }

      DenseVector clone(CloneMode mode = CloneMode::Deep) const
      {
        if(mode == CloneMode::Shallow) return *this;
        DenseVector v(size());
        if(mode == CloneMode::Deep) v.copy(*this);
        return v;
      }

      /// Container::format (lafem/container.hpp:847)
      void format(DT_ value = DT_(0))
      {
        if(std::is_same<DT_, double>::value)
          check(feat_set_memory_f64(Runtime::handle(), (double*)elements(), double(value), size()), "format");
        else
          check(feat_set_memory_f32(Runtime::handle(), (float*)elements(), float(value), size()), "format");
      }

      /// DenseVector::copy (dense_vector.hpp:793)
      void copy(const DenseVector& x)
      {
        if(x.size() != size()) throw Error("DenseVector::copy: size mismatch");
        check(feat_copy_device(Runtime::handle(), elements(), x.elements(), sizeof(DT_) * size()), "copy");
      }

      /// this = alpha*x + y (dense_vector.hpp:807-830)
      void axpy(const DenseVector& x, const DenseVector& y, DT_ alpha = DT_(1))
      {
        if(x.size() != size() || y.size() != size()) throw Error("DenseVector::axpy: size mismatch");
        if(std::is_same<DT_, double>::value)
          check(feat_axpy_f64(Runtime::handle(), (double*)elements(), double(alpha), (const double*)x.elements(),
                              (const double*)y.elements(), size()), "axpy");
        else
          check(feat_axpy_f32(Runtime::handle(), (float*)elements(), float(alpha), (const float*)x.elements(),
                              (const float*)y.elements(), size()), "axpy");
      }

      /// dense_vector.hpp:924-938 (host scalar, synchronises)
      DT_ dot(const DenseVector& x) const
      {
        static_assert(std::is_same<DT_, double>::value, "dot: FP64 only");
        double r = 0.0;
        check(feat_dot_f64_host(Runtime::handle(), &r, (const double*)elements(), (const double*)x.elements(), size()), "dot");
        return DT_(r);
      }

      /// dense_vector.hpp:943-956
      DT_ norm2() const
      {
        static_assert(std::is_same<DT_, double>::value, "norm2: FP64 only");
        double r = 0.0;
        check(feat_norm2_f64_host(Runtime::handle(), &r, (const double*)elements(), size()), "norm2");
        return DT_(r);
      }

      DT_ norm2sqr() const
      {
        DT_ n = norm2();
        return n * n;
      }

      /// this = alpha*x (dense_vector.hpp:879-891)
      void scale(const DenseVector& x, DT_ alpha)
      {
        check(feat_scale_f64(Runtime::handle(), (double*)elements(), (const double*)x.elements(), double(alpha), size()), "scale");
      }

      /// this = x∘y (dense_vector.hpp:837-850)
      void component_product(const DenseVector& x, const DenseVector& y)
      {
        check(feat_component_product_f64(Runtime::handle(), (double*)elements(), (const double*)x.elements(),
                                         (const double*)y.elements(), size()), "component_product");
      }

      /// this = alpha / x (dense_vector.hpp:860-871)
      void component_invert(const DenseVector& x, DT_ alpha = DT_(1))
      {
        check(feat_component_invert_f64(Runtime::handle(), (double*)elements(), (const double*)x.elements(),
                                        double(alpha), size()), "component_invert");
      }
    };

    /// SparseMatrixCSR<DT_, IT_> (lafem/sparse_matrix_csr.hpp:61) on the device.
    template<typename DT_, typename IT_ = Index>
    class SparseMatrixCSR
    {
      Index _rows = 0, _cols = 0, _nnz = 0;
      DeviceArray<DT_> _val;
      DeviceArray<IT_> _col, _rowptr;
      mutable std::shared_ptr<feat_matrix_f64_t> _plan;
      mutable std::shared_ptr<feat_matrix_f32_t> _plan32, _plan16;
      mutable DeviceArray<float> _csr32_val;
      mutable DeviceArray<std::uint32_t> _csr32_col, _csr32_rp;
      bool _use_sell = true;

    public:
      typedef DT_ DataType;
      typedef IT_ IndexType;
      typedef DenseVector<DT_, IT_> VectorTypeL;
      typedef DenseVector<DT_, IT_> VectorTypeR;

      SparseMatrixCSR() = default;
      static SparseMatrixCSR from_host(Index rows, Index cols, const DT_* val, const IT_* col, const IT_* rowptr)
      {
        SparseMatrixCSR m;
        m._rows = rows;
        m._cols = cols;
        m._nnz = Index(rowptr[rows]);
        m._val = DeviceArray<DT_>::from_host(val, m._nnz);
        m._col = DeviceArray<IT_>::from_host(col, m._nnz);
        m._rowptr = DeviceArray<IT_>::from_host(rowptr, rows + 1);
        return m;
      }
      /// uninitialised device arrays of the given shape, to be filled by a device assembly
      static SparseMatrixCSR allocate(Index rows, Index cols, Index nnz)
      {
        SparseMatrixCSR m;
        m._rows = rows;
        m._cols = cols;
        m._nnz = nnz;
        m._val = DeviceArray<DT_>(nnz);
        m._col = DeviceArray<IT_>(nnz);
        m._rowptr = DeviceArray<IT_>(rows + 1);
        return m;
      }
      DT_* val_mut() { return _val.get(); }
      IT_* col_mut() { return _col.get(); }
      IT_* row_ptr_mut() { return _rowptr.get(); }
      static SparseMatrixCSR wrap(Index rows, Index cols, Index nnz, DT_* val, IT_* col, IT_* rowptr)
      {
        SparseMatrixCSR m;
        m._rows = rows;
        m._cols = cols;
        m._nnz = nnz;
        m._val = DeviceArray<DT_>::wrap(val, nnz);
        m._col = DeviceArray<IT_>::wrap(col, nnz);
        m._rowptr = DeviceArray<IT_>::wrap(rowptr, rows + 1);
        return m;
      }
      Index rows() const { return _rows; }
      Index columns() const { return _cols; }
      Index used_elements() const { return _nnz; }
      const DT_* val() const { return _val.get(); }
      const IT_* col_ind() const { return _col.get(); }
      const IT_* row_ptr() const { return _rowptr.get(); }
      bool empty() const { return _rowptr.empty(); }
      VectorTypeL create_vector_l() const { return VectorTypeL(_rows); }
      VectorTypeR create_vector_r() const { return VectorTypeR(_cols); }

      feat_csr_f64_t desc() const
      {
        static_assert(std::is_same<DT_, double>::value && std::is_same<IT_, std::uint64_t>::value, "desc: f64/u64");
        return feat_csr_f64_t{_rows, _cols, _nnz, (const double*)val(), (const std::uint64_t*)col_ind(),
                              (const std::uint64_t*)row_ptr()};
      }

      /// Execution descriptor for the fused kernels: the SELL-64 plan of this
      /// matrix (built on first use, kept until the matrix is destroyed), or the
      /// CSR arrays themselves when no plan can be built.
      const feat_matrix_f64_t& exec() const
      {
        if(!_plan)
        {
          auto p = std::shared_ptr<feat_matrix_f64_t>(new feat_matrix_f64_t(), [](feat_matrix_f64_t* m) {
            if(m->format == FEAT_FMT_SELL || m->format == FEAT_FMT_BANDED) feat_sell_destroy_f64(m);
            delete m;
          });
          const feat_csr_f64_t c = desc();
          if(!_use_sell || feat_sell_create_ex_f64(Runtime::handle(), &c, Runtime::sell_flags(), p.get()) != FEAT_OK)
            check(feat_matrix_from_csr_f64(&c, p.get()), "exec");
          _plan = p;
        }
        return *_plan;
      }

      bool has_plan() const { return bool(_plan); }

      /// Picks the fastest launch shape of the execution plan's row kernels
      /// (feat_autotune_matrix_f64; bit-identical results for every shape).
      /// Returns the time per SpMV of the chosen shape in microseconds (-1: nothing to tune).
      double autotune(int reps = 20) const
      {
        const feat_matrix_f64_t& e = exec();
        DenseVector<DT_, IT_> x(_cols, DT_(1)), r(_rows);
        double us = -1.0;
        check(feat_autotune_matrix_f64(Runtime::handle(), const_cast<feat_matrix_f64_t*>(&e), (const double*)x.elements(),
                                       (double*)r.elements(), reps, &us),
              "autotune");
        return us;
      }
      int tune() const { return exec().tune; }

      /// FP32 SELL-64 plan (values rounded once) for mixed-precision inner solvers.
      const feat_matrix_f32_t& exec32() const
      {
        if(!_plan32)
        {
          auto p = std::shared_ptr<feat_matrix_f32_t>(new feat_matrix_f32_t(), [](feat_matrix_f32_t* m) {
            if(m->format == FEAT_FMT_SELL || m->format == FEAT_FMT_BANDED) feat_sell_destroy_f32(m);
            delete m;
          });
          const feat_csr_f64_t c = desc();
          if(feat_sell_create_ex_f32(Runtime::handle(), &c, Runtime::sell_flags(), p.get()) != FEAT_OK)
          {
            // FP32 CSR copy with 32-bit indices (rows too long for a SELL-64 plan)
            if(_nnz >= (Index(1) << 32) || _cols >= (Index(1) << 32)) throw Error("exec32: matrix too large for u32 indices");
            _csr32_val = DeviceArray<float>(_nnz);
            _csr32_col = DeviceArray<std::uint32_t>(_nnz);
            _csr32_rp = DeviceArray<std::uint32_t>(_rows + 1);
            feat_handle_t h = Runtime::handle();
            check(feat_convert_f64_f32(h, _csr32_val.get(), (const double*)val(), _nnz), "exec32 val");
            check(feat_convert_u64_u32(h, _csr32_col.get(), (const std::uint64_t*)col_ind(), _nnz), "exec32 col");
            check(feat_convert_u64_u32(h, _csr32_rp.get(), (const std::uint64_t*)row_ptr(), _rows + 1), "exec32 rowptr");
            *p = feat_matrix_f32_t{FEAT_FMT_CSR, FEAT_IDX_U32, _rows, _cols, _nnz, _csr32_val.get(), _csr32_col.get(),
                                   _csr32_rp.get(), nullptr, FEAT_VAL_NATIVE, 1.0f};
          }
          _plan32 = p;
        }
        return *_plan32;
      }

      /// FP16-value SELL-64 plan (FP32 arithmetic, exact power-of-two scale); falls
      /// back to exec32() when no SELL-64 plan can be built, and uses it when it is a
      /// pair-dictionary plan (one byte per entry, smaller than FP16 values).
      const feat_matrix_f32_t& exec16() const
      {
        if(!_plan16 && (exec32().index == FEAT_IDX_DICT || exec32().index == FEAT_IDX_DIAG))
          _plan16 = std::shared_ptr<feat_matrix_f32_t>(std::shared_ptr<feat_matrix_f32_t>(), const_cast<feat_matrix_f32_t*>(&exec32()));
        if(!_plan16)
        {
          auto p = std::shared_ptr<feat_matrix_f32_t>(new feat_matrix_f32_t(), [](feat_matrix_f32_t* m) {
            if(m->format == FEAT_FMT_SELL || m->format == FEAT_FMT_BANDED) feat_sell_destroy_f32(m);
            delete m;
          });
          const feat_csr_f64_t c = desc();
          if(feat_sell_create_ex_f16(Runtime::handle(), &c, Runtime::sell_flags(), p.get()) != FEAT_OK)
          {
            _plan16 = std::shared_ptr<feat_matrix_f32_t>(std::shared_ptr<feat_matrix_f32_t>(), const_cast<feat_matrix_f32_t*>(&exec32()));
            return *_plan16;
          }
          _plan16 = p;
        }
        return *_plan16;
      }

      /// The values (not the pattern) were overwritten: the execution plans, which
      /// hold value codes or converted values, are built again on next use.
      void values_changed()
      {
        _plan.reset();
        _plan32.reset();
        _plan16.reset();
      }

      /// Selects CSR (false) or the SELL-64 plan (true, default) for exec().
      void set_sell(bool on)
      {
        _use_sell = on;
        _plan.reset();
      }

      /// r = A*x  (sparse_matrix_csr.hpp:1728-1756)
      void apply(VectorTypeL& r, const VectorTypeR& x) const { _apply(r, x, nullptr, DT_(1), false); }
      /// r = A^T*x
      void apply_transposed(VectorTypeR& r, const VectorTypeL& x) const { _apply(r, x, nullptr, DT_(1), true); }
      /// r = y + alpha*A*x  (sparse_matrix_csr.hpp:1807-1843)
      void apply(VectorTypeL& r, const VectorTypeR& x, const VectorTypeL& y, DT_ alpha) const
      {
        _apply(r, x, &y, alpha, false);
      }

      /// extract_diag (sparse_matrix_csr.hpp:2277-2303)
      void extract_diag(VectorTypeL& diag) const
      {
        check(feat_csr_extract_diag_f64_u64(Runtime::handle(), (double*)diag.elements(), (const double*)val(),
                                            (const std::uint64_t*)col_ind(), (const std::uint64_t*)row_ptr(), _rows),
              "extract_diag");
      }

    private:
      void _apply(DenseVector<DT_, IT_>& r, const DenseVector<DT_, IT_>& x, const DenseVector<DT_, IT_>* y, DT_ alpha,
                  bool tr) const
      {
        const Index nr = tr ? _cols : _rows, nx = tr ? _rows : _cols;
        if(r.size() != nr || x.size() != nx || (y && y->size() != nr)) throw Error("SparseMatrixCSR::apply: size mismatch");
        const DT_ b = y ? DT_(1) : DT_(0);
        const DT_* yp = y ? y->elements() : r.elements();
        int rc;
        if(std::is_same<DT_, double>::value && std::is_same<IT_, std::uint64_t>::value)
          rc = feat_csr_apply_f64_u64(Runtime::handle(), (double*)r.elements(), double(alpha), (const double*)x.elements(),
                                      double(b), (const double*)yp, (const double*)val(), (const std::uint64_t*)col_ind(),
                                      (const std::uint64_t*)row_ptr(), _rows, _cols, _nnz, tr);
        else if(std::is_same<DT_, float>::value && std::is_same<IT_, std::uint32_t>::value)
          rc = feat_csr_apply_f32_u32(Runtime::handle(), (float*)r.elements(), float(alpha), (const float*)x.elements(),
                                      float(b), (const float*)yp, (const float*)val(), (const std::uint32_t*)col_ind(),
                                      (const std::uint32_t*)row_ptr(), _rows, _cols, _nnz, tr);
        else
          throw Error("SparseMatrixCSR::apply: type combination not provided");
        check(rc, "SparseMatrixCSR::apply");
      }
    };

    /// UnitFilter<DT_, IT_> (lafem/unit_filter.hpp) plus the bit mask the
    /// fused kernels use for the same filter.
    template<typename DT_, typename IT_ = Index>
    class UnitFilter
    {
      Index _size = 0, _ue = 0;
      DeviceArray<std::uint64_t> _idx;
      DeviceArray<double> _val;
      DeviceArray<std::uint32_t> _mask;

    public:
      UnitFilter() = default;
      UnitFilter(Index size, const std::uint64_t* idx, const double* values, Index ue) : _size(size), _ue(ue)
      {
        std::vector<double> zeros;
        if(!values)
        {
          zeros.assign(ue, 0.0);
          values = zeros.data();
        }
        _idx = DeviceArray<std::uint64_t>::from_host(idx, ue);
        _val = DeviceArray<double>::from_host(values, ue);
        _mask = DeviceArray<std::uint32_t>((size + 31) / 32);
        check(feat_unit_filter_mask_u64(Runtime::handle(), _mask.get(), size, _idx.get(), ue), "UnitFilter mask");
      }
      /// homogeneous filter on device-resident indices
      UnitFilter(Index size, const DeviceArray<std::uint64_t>& idx) : _size(size), _ue(idx.size()), _idx(idx)
      {
        _val = DeviceArray<double>(_ue);
        if(_ue) check(feat_set_memory_f64(Runtime::handle(), _val.get(), 0.0, _ue), "UnitFilter values");
        _mask = DeviceArray<std::uint32_t>((size + 31) / 32);
        check(feat_unit_filter_mask_u64(Runtime::handle(), _mask.get(), size, _idx.get(), _ue), "UnitFilter mask");
      }
      Index size() const { return _size; }
      Index used_elements() const { return _ue; }
      const std::uint32_t* mask() const { return _mask.get(); }
      const std::uint64_t* indices() const { return _idx.get(); }

      void filter_rhs(DenseVector<DT_, IT_>& v) const
      {
        if(_ue) check(feat_unit_filter_rhs_f64_u64(Runtime::handle(), (double*)v.elements(), _val.get(), _idx.get(), _ue), "filter_rhs");
      }
      void filter_sol(DenseVector<DT_, IT_>& v) const { filter_rhs(v); }
      void filter_def(DenseVector<DT_, IT_>& v) const
      {
        if(_ue) check(feat_unit_filter_def_f64_u64(Runtime::handle(), (double*)v.elements(), _idx.get(), _ue), "filter_def");
      }
      void filter_cor(DenseVector<DT_, IT_>& v) const { filter_def(v); }
    };

    /// Transfer<Matrix_> (lafem/transfer.hpp:28, rest/prol :247-283)
    template<typename Matrix_>
    class Transfer
    {
      Matrix_ _prol, _rest;

    public:
      typedef typename Matrix_::VectorTypeL VectorType;
      Transfer() = default;
      Transfer(Matrix_&& prol, Matrix_&& rest) : _prol(std::move(prol)), _rest(std::move(rest)) {}
      const Matrix_& get_mat_prol() const { return _prol; }
      const Matrix_& get_mat_rest() const { return _rest; }
      Matrix_& get_mat_prol() { return _prol; }
      Matrix_& get_mat_rest() { return _rest; }
      bool rest(const VectorType& fine, VectorType& coarse) const
      {
        _rest.apply(coarse, fine);
        return true;
      }
      bool prol(VectorType& fine, const VectorType& coarse) const
      {
        _prol.apply(fine, coarse);
        return true;
      }
    };
  } // namespace LAFEM
} // namespace feat3amd
