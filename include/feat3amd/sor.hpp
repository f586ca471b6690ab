// feat3amd/sor.hpp -- SORPrecond / SSORPrecond / ILUPrecond on the device.
//
// SolverBase preconditioners over the SOR / SSOR sweeps and the level-scheduled ILU(0)
// of libfeat3amd.so (include/feat3amd.h): the interface and lifecycle of the
// reference's SORPrecond (solver/sor_precond.hpp:741-925), SSORPrecond
// (solver/ssor_precond.hpp:720-899) and ILUPrecond (solver/ilu_precond.hpp:1670-1856):
// init_symbolic builds the level sets (or, with FEAT_SOR_MULTICOLOR, the colouring of
// the CUDA backend), init_numeric takes the values, apply(cor, def) computes
// cor = M^{-1} def and filters it.  In level-set order the results are those of the
// reference's generic backend bit for bit.
#pragma once

#include <feat3amd/lafem.hpp>
#include <feat3amd/solver.hpp>

namespace feat3amd
{
  namespace Solver
  {
    /// SOR (sor_precond.hpp:159-222): cor_i = omega (def_i - sum_{j<i} a_ij cor_j) / a_ii
    template<typename Matrix_, typename Filter_>
    class SORPrecond : public SolverBase<typename Matrix_::VectorTypeL>
    {
    public:
      typedef typename Matrix_::VectorTypeL VectorType;
      typedef typename Matrix_::DataType DataType;

    protected:
      const Matrix_& _matrix;
      const Filter_& _filter;
      DataType _omega;
      feat_sor_t _plan = nullptr;
      int _flags = 0;

    public:
      SORPrecond(const Matrix_& m, const Filter_& f, DataType omega = DataType(1), int flags = 0)
        : _matrix(m), _filter(f), _omega(omega), _flags(flags)
      {
        if(m.rows() != m.columns()) throw Error("SORPrecond: matrix is not square");
      }
      ~SORPrecond() override { done_symbolic(); }
      std::string name() const override { return "SOR"; }
      void set_omega(DataType omega)
      {
        if(!(omega > DataType(0))) throw Error("SORPrecond: omega must be positive");
        _omega = omega;
      }
      void init_symbolic() override
      {
        done_symbolic();
        const feat_csr_f64_t d = _matrix.desc();
        check(feat_sor_init_symbolic(Runtime::handle(), &d, _flags, &_plan), "SORPrecond::init_symbolic");
      }
      void init_numeric() override
      {
        const feat_csr_f64_t d = _matrix.desc();
        check(feat_sor_init_numeric(Runtime::handle(), _plan, &d), "SORPrecond::init_numeric");
      }
      void done_symbolic() override
      {
        if(_plan) feat_sor_done_symbolic(_plan);
        _plan = nullptr;
      }
      feat_sor_t plan() const { return _plan; }
      Status apply(VectorType& cor, const VectorType& def) override
      {
        check(feat_sor_apply(Runtime::handle(), _plan, cor.elements(), def.elements(), _omega), "SORPrecond::apply");
        _filter.filter_cor(cor);
        return Status::success;
      }
    };

    /// SSOR (ssor_precond.hpp:166-231): forward sweep, backward sweep, scaling by omega (2 - omega)
    template<typename Matrix_, typename Filter_>
    class SSORPrecond : public SORPrecond<Matrix_, Filter_>
    {
      typedef SORPrecond<Matrix_, Filter_> BaseClass;

    public:
      typedef typename BaseClass::VectorType VectorType;
      typedef typename BaseClass::DataType DataType;
      using BaseClass::BaseClass;
      std::string name() const override { return "SSOR"; }
      Status apply(VectorType& cor, const VectorType& def) override
      {
        const feat_handle_t h = Runtime::handle();
        check(feat_ssor_forward_apply(h, this->_plan, cor.elements(), def.elements(), this->_omega), "SSORPrecond::apply");
        check(feat_ssor_backward_apply(h, this->_plan, cor.elements(), this->_omega), "SSORPrecond::apply");
        cor.scale(cor, this->_omega * (DataType(2.0) - this->_omega));
        this->_filter.filter_cor(cor);
        return Status::success;
      }
    };

    /// ILU(p) (ilu_precond.hpp:1040-1104, generic backend; the level-p pattern of
    /// ILUCoreSymbolic::factorize_symbolic): cor = (D+U)^{-1}(I+L)^{-1} def
    template<typename Matrix_, typename Filter_>
    class ILUPrecond : public SolverBase<typename Matrix_::VectorTypeL>
    {
    public:
      typedef typename Matrix_::VectorTypeL VectorType;
      typedef typename Matrix_::DataType DataType;

    private:
      const Matrix_& _matrix;
      const Filter_& _filter;
      feat_ilu_t _plan = nullptr;
      int _p = 0;

    public:
      ILUPrecond(const Matrix_& m, const Filter_& f, int p = 0) : _matrix(m), _filter(f), _p(p)
      {
        if(m.rows() != m.columns()) throw Error("ILUPrecond: matrix is not square");
        if(p < 0) throw Error("ILUPrecond: negative fill-in level");
      }
      ~ILUPrecond() override { done_symbolic(); }
      std::string name() const override { return "ILU"; }
      void init_symbolic() override
      {
        done_symbolic();
        const feat_csr_f64_t d = _matrix.desc();
        check(feat_ilu_init_symbolic_p(Runtime::handle(), &d, _p, &_plan), "ILUPrecond::init_symbolic");
      }
      void init_numeric() override
      {
        const feat_csr_f64_t d = _matrix.desc();
        check(feat_ilu_init_numeric(Runtime::handle(), _plan, &d), "ILUPrecond::init_numeric");
      }
      void done_symbolic() override
      {
        if(_plan) feat_ilu_done_symbolic(_plan);
        _plan = nullptr;
      }
      Status apply(VectorType& cor, const VectorType& def) override
      {
        check(feat_ilu_apply(Runtime::handle(), _plan, cor.elements(), def.elements()), "ILUPrecond::apply");
        _filter.filter_cor(cor);
        return Status::success;
      }
    };

    template<typename Matrix_, typename Filter_>
    std::shared_ptr<ILUPrecond<Matrix_, Filter_>> new_ilu_precond(const Matrix_& m, const Filter_& f, int p = 0)
    {
      return std::make_shared<ILUPrecond<Matrix_, Filter_>>(m, f, p);
    }

    template<typename Matrix_, typename Filter_>
    std::shared_ptr<SORPrecond<Matrix_, Filter_>> new_sor_precond(const Matrix_& m, const Filter_& f,
                                                                  typename Matrix_::DataType omega = 1)
    {
      return std::make_shared<SORPrecond<Matrix_, Filter_>>(m, f, omega);
    }

    template<typename Matrix_, typename Filter_>
    std::shared_ptr<SSORPrecond<Matrix_, Filter_>> new_ssor_precond(const Matrix_& m, const Filter_& f,
                                                                    typename Matrix_::DataType omega = 1)
    {
      return std::make_shared<SSORPrecond<Matrix_, Filter_>>(m, f, omega);
    }
  } // namespace Solver
} // namespace feat3amd
