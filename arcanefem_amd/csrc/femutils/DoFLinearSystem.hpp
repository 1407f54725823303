// DoFLinearSystem.hpp -- the linear-system side of ArcaneFEM's drop-in
// boundary, re-expressed over the afem C ABI.
//
//   IDoFLinearSystemImpl     femutils/internal/IDoFLinearSystemImpl.h:37-70
//   IDoFLinearSystemFactory  femutils/IDoFLinearSystemFactory.h:34-44
//   DoFLinearSystem          femutils/DoFLinearSystem.h:154-409 (pimpl facade)
//   CsrFormatMatrixView      femutils/CsrFormatMatrixView.h:135-210
//
// Arcane's types are replaced by small stand-ins (DoFLocalId, a device-array
// "variable", a Runner holding the device index); the method names, argument
// meaning and error behaviour follow the reference.  Mi355xDoFLinearSystemFactory
// creates the implementation that solves with the HIP PCG of libafem; it plays
// the role of HypreDoFLinearSystemFactory / PetscDoFLinearSystemFactory
// (CSR view hand-off, GPU-side Dirichlet transformations, then solve()).
#pragma once

#include "Mesh.hpp"

#include <memory>
#include <sstream>

namespace afem::femutils {

struct DoFLocalId
{
  Int32 v = -1;
  DoFLocalId() = default;
  explicit DoFLocalId(Int32 x)
  : v(x)
  {}
  operator Int32() const { return v; }
};

// Values of a DoF variable held in device memory (VariableDoFReal & co).
template <class T>
class DoFVariable
{
 public:
  DoFVariable() = default;
  DoFVariable(T* device, Int32 n)
  : m_ptr(device)
  , m_n(n)
  {}
  T* devicePointer() const { return m_ptr; }
  Int32 size() const { return m_n; }
  std::vector<T> toHost() const
  {
    std::vector<T> h((size_t)m_n);
    check(afem_memcpy_dtoh(h.data(), m_ptr, sizeof(T) * (size_t)m_n), "afem_memcpy_dtoh");
    return h;
  }
  void fromHost(const std::vector<T>& h)
  {
    if ((Int32)h.size() != m_n)
      throw FatalError(AFEM_ERR_ARGUMENT, "DoF variable size mismatch");
    check(afem_memcpy_htod(m_ptr, h.data(), sizeof(T) * (size_t)m_n), "afem_memcpy_htod");
  }
  void fill(T value) { fromHost(std::vector<T>((size_t)m_n, value)); }

 private:
  T* m_ptr = nullptr;
  Int32 m_n = 0;
};
using VariableDoFReal = DoFVariable<Real>;
using VariableDoFByte = DoFVariable<Byte>;
using VariableDoFBool = DoFVariable<Byte>;

// A device array the caller owns, seen as a DoF variable: the node and cell
// fields of a module (the heat module's old temperature, per-cell lambda).
template <typename T>
class OwnedDoFVariable
{
 public:
  explicit OwnedDoFVariable(Int32 n)
  {
    void* p = nullptr;
    check(afem_malloc(&p, sizeof(T) * (size_t)(n > 0 ? n : 1)), "afem_malloc");
    m_var = DoFVariable<T>(static_cast<T*>(p), n);
  }
  OwnedDoFVariable(const OwnedDoFVariable&) = delete;
  OwnedDoFVariable& operator=(const OwnedDoFVariable&) = delete;
  ~OwnedDoFVariable() { afem_free(m_var.devicePointer()); }
  DoFVariable<T>& variable() { return m_var; }
  const DoFVariable<T>& variable() const { return m_var; }
  T* devicePointer() const { return m_var.devicePointer(); }
  void fill(T value) { m_var.fill(value); }
  void copyFrom(const DoFVariable<T>& src)
  {
    if (src.size() != m_var.size())
      throw FatalError(AFEM_ERR_ARGUMENT, "DoF variable size mismatch");
    check(afem_memcpy_dtod(m_var.devicePointer(), src.devicePointer(), sizeof(T) * (size_t)src.size()),
          "afem_memcpy_dtod");
  }

 private:
  DoFVariable<T> m_var;
};

// Non-owning CSR view of device arrays with the spans of
// femutils/CsrFormatMatrixView.h:145-153: rows holds nb_row row starts (the
// last row ends at nb_value), rows_nb_column the per-row counts (optional).
struct CsrFormatView
{
  const Int32* rows = nullptr;
  const Int32* rows_nb_column = nullptr;
  const Int32* columns = nullptr;
  Real* values = nullptr;
  Int32 nb_row = 0;
  Int64 nb_value = 0;
  afem_bsr_t bsr = nullptr; // set when the view comes from a BSRFormat
  Int32 nbRow() const { return nb_row; }
  Int64 nbValue() const { return nb_value; }
};
using CSRFormatView = CsrFormatView;

struct Runner
{
  int device = 0;
};

struct CommandLineArguments
{
  std::vector<std::string> args;
};

class IDoFLinearSystemImpl
{
 public:
  virtual ~IDoFLinearSystemImpl() = default;
  virtual void matrixAddValue(DoFLocalId row, DoFLocalId column, Real value) = 0;
  virtual void matrixSetValue(DoFLocalId row, DoFLocalId column, Real value) = 0;
  virtual void eliminateRow(DoFLocalId row, Real value) = 0;
  virtual void eliminateRowColumn(DoFLocalId row, Real value) = 0;
  virtual void applyMatrixTransformation() = 0;
  virtual void applyRHSTransformation() = 0;
  virtual void solve() = 0;
  virtual VariableDoFReal& solutionVariable() = 0;
  virtual VariableDoFReal& rhsVariable() = 0;
  virtual void setSolverCommandLineArguments(const CommandLineArguments& args) = 0;
  virtual void clearValues() = 0;
  virtual void setCSRValues(const CSRFormatView& csr_view) = 0;
  virtual CSRFormatView& getCSRValues() = 0;
  virtual VariableDoFReal& getForcedValue() = 0;
  virtual VariableDoFBool& getForcedInfo() = 0;
  virtual VariableDoFByte& getEliminationInfo() = 0;
  virtual VariableDoFReal& getEliminationValue() = 0;
  [[nodiscard]] virtual bool hasSetCSRValues() const = 0;
  virtual void setConstantMatrixSparsity(bool val) = 0;
  virtual void setConstantMatrixValues(bool val) = 0;
  [[nodiscard]] virtual bool isMatrixSparsityConstant() const = 0;
  [[nodiscard]] virtual bool isMatrixValuesConstant() const = 0;
  virtual void setRunner(const Runner& r) = 0;
  [[nodiscard]] virtual Runner runner() const = 0;
};

// The DoF family is represented by its size.
class IDoFLinearSystemFactory
{
 public:
  virtual ~IDoFLinearSystemFactory() = default;
  virtual IDoFLinearSystemImpl* createInstance(Int32 nb_dof, const std::string& solver_name) = 0;
};

// Solver settings of the MI355X implementation.  setSolverCommandLineArguments
// accepts the PETSc-style flags of the reference's PETSc back-end
// (-ksp_rtol, -ksp_atol, -ksp_max_it, -pc_type jacobi|none,
// -ksp_norm_type unpreconditioned|natural|preconditioned) plus
// -afem_precision fp64|mixed.
struct SolverSettings
{
  afem_solver_options options{};
  SolverSettings() { afem_solver_default_options(&options); }
};

class Mi355xDoFLinearSystemImpl : public IDoFLinearSystemImpl
{
 public:
  Mi355xDoFLinearSystemImpl(Int32 nb_dof, const std::string& solver_name)
  : m_name(solver_name)
  {
    check(afem_system_create(nb_dof, &m_sys), "afem_system_create");
    double* rhs = nullptr;
    double* x = nullptr;
    uint8_t* fi = nullptr;
    double* fv = nullptr;
    uint8_t* ei = nullptr;
    double* ev = nullptr;
    check(afem_system_rhs(m_sys, &rhs), "afem_system_rhs");
    check(afem_system_solution(m_sys, &x), "afem_system_solution");
    check(afem_system_forced(m_sys, &fi, &fv), "afem_system_forced");
    check(afem_system_elimination(m_sys, &ei, &ev), "afem_system_elimination");
    m_rhs = VariableDoFReal(rhs, nb_dof);
    m_x = VariableDoFReal(x, nb_dof);
    m_forced_info = VariableDoFBool(fi, nb_dof);
    m_forced_value = VariableDoFReal(fv, nb_dof);
    m_elim_info = VariableDoFByte(ei, nb_dof);
    m_elim_value = VariableDoFReal(ev, nb_dof);
  }
  ~Mi355xDoFLinearSystemImpl() override { afem_system_destroy(m_sys); }

  void matrixAddValue(DoFLocalId row, DoFLocalId column, Real value) override
  {
    check(afem_system_matrix_add_value(m_sys, row, column, value, 0), "matrixAddValue");
  }
  void matrixSetValue(DoFLocalId row, DoFLocalId column, Real value) override
  {
    check(afem_system_matrix_add_value(m_sys, row, column, value, 1), "matrixSetValue");
  }
  // femutils/internal/CsrDoFLinearSystemImpl.h:64-71: not implemented for CSR
  // implementations; the elimination helpers write the DoF arrays instead.
  void eliminateRow(DoFLocalId, Real) override { throw FatalError(AFEM_ERR_NOT_IMPLEMENTED, "eliminateRow"); }
  void eliminateRowColumn(DoFLocalId, Real) override
  {
    throw FatalError(AFEM_ERR_NOT_IMPLEMENTED, "eliminateRowColumn");
  }
  void applyMatrixTransformation() override
  {
    check(afem_system_apply_matrix_transformation(m_sys), "applyMatrixTransformation");
  }
  void applyRHSTransformation() override
  {
    check(afem_system_apply_rhs_transformation(m_sys), "applyRHSTransformation");
  }
  void solve() override
  {
    afem_solve_result r{};
    int rc = afem_system_solve(m_sys, &m_settings.options, &r);
    m_last = r;
    check(rc, ("solve of '" + m_name + "'").c_str());
  }
  VariableDoFReal& solutionVariable() override { return m_x; }
  VariableDoFReal& rhsVariable() override { return m_rhs; }
  void setSolverCommandLineArguments(const CommandLineArguments& args) override
  {
    auto& o = m_settings.options;
    for (size_t k = 0; k + 1 < args.args.size(); k += 2) {
      const std::string& key = args.args[k];
      const std::string& val = args.args[k + 1];
      if (key == "-ksp_rtol")
        o.rtol = std::stod(val);
      else if (key == "-ksp_atol")
        o.atol = std::stod(val);
      else if (key == "-ksp_max_it")
        o.max_iter = std::stoi(val);
      else if (key == "-pc_type")
        o.preconditioner = (val == "none") ? AFEM_PRECOND_NONE : AFEM_PRECOND_JACOBI;
      else if (key == "-ksp_norm_type") { // PETSc KSPNormType names
        if (val != "unpreconditioned" && val != "natural" && val != "preconditioned")
          throw FatalError(AFEM_ERR_ARGUMENT, "unknown -ksp_norm_type '" + val + "'");
        o.norm = (val == "natural") ? 1 : (val == "preconditioned") ? 2 : 0;
      }
      else if (key == "-afem_precision")
        o.method = (val == "mixed") ? AFEM_SOLVER_PCG_MIXED : AFEM_SOLVER_PCG_FP64;
      else
        throw FatalError(AFEM_ERR_ARGUMENT, "unknown solver argument '" + key + "'");
    }
  }
  void clearValues() override
  {
    check(afem_system_clear_values(m_sys), "clearValues");
    m_csr = {};
  }
  void setCSRValues(const CSRFormatView& v) override
  {
    m_csr = v;
    if (v.bsr)
      check(afem_system_set_bsr(m_sys, v.bsr), "setCSRValues");
    else
      check(afem_system_set_csr_values(m_sys, v.nb_row, v.nb_value, v.rows, v.rows_nb_column, v.columns, v.values),
            "setCSRValues");
  }
  CSRFormatView& getCSRValues() override { return m_csr; }
  VariableDoFReal& getForcedValue() override { return m_forced_value; }
  VariableDoFBool& getForcedInfo() override { return m_forced_info; }
  VariableDoFByte& getEliminationInfo() override { return m_elim_info; }
  VariableDoFReal& getEliminationValue() override { return m_elim_value; }
  bool hasSetCSRValues() const override { return true; }
  void setConstantMatrixSparsity(bool v) override { m_constant_sparsity = v; }
  void setConstantMatrixValues(bool v) override { m_constant_values = v; }
  bool isMatrixSparsityConstant() const override { return m_constant_sparsity; }
  bool isMatrixValuesConstant() const override { return m_constant_values; }
  void setRunner(const Runner& r) override
  {
    m_runner = r;
    check(afem_init(r.device), "setRunner");
  }
  Runner runner() const override { return m_runner; }

  afem_system_t handle() const { return m_sys; }
  SolverSettings& settings() { return m_settings; }
  const afem_solve_result& lastSolve() const { return m_last; }

 private:
  std::string m_name;
  afem_system_t m_sys = nullptr;
  VariableDoFReal m_rhs, m_x, m_forced_value, m_elim_value;
  VariableDoFBool m_forced_info;
  VariableDoFByte m_elim_info;
  CSRFormatView m_csr;
  SolverSettings m_settings;
  afem_solve_result m_last{};
  bool m_constant_sparsity = false;
  bool m_constant_values = false;
  Runner m_runner;
};

class Mi355xDoFLinearSystemFactory : public IDoFLinearSystemFactory
{
 public:
  IDoFLinearSystemImpl* createInstance(Int32 nb_dof, const std::string& solver_name) override
  {
    return new Mi355xDoFLinearSystemImpl(nb_dof, solver_name);
  }
};

// Facade of femutils/DoFLinearSystem.h: owns the implementation created by
// the factory (deleted in reset() and in the destructor,
// femutils/DoFLinearSystem.cc:347-352, 595-601).
class DoFLinearSystem
{
 public:
  DoFLinearSystem() = default;
  DoFLinearSystem(const DoFLinearSystem&) = delete;
  DoFLinearSystem& operator=(const DoFLinearSystem&) = delete;
  ~DoFLinearSystem() { reset(); }

  void setLinearSystemFactory(IDoFLinearSystemFactory* factory) { m_factory = factory; }
  IDoFLinearSystemFactory* linearSystemFactory() const { return m_factory; }

  // femutils/DoFLinearSystem.cc:367-385: createInstance then setRunner.
  void initialize(Runner* runner, Int32 nb_dof, const std::string& solver_name)
  {
    if (m_p)
      throw FatalError(AFEM_ERR_STATE, "DoFLinearSystem already initialized: call reset() first");
    IDoFLinearSystemFactory* f = m_factory ? m_factory : &m_default_factory;
    m_p = f->createInstance(nb_dof, solver_name);
    m_p->setRunner(runner ? *runner : Runner{});
  }
  bool isInitialized() const { return m_p != nullptr; }
  void reset()
  {
    delete m_p;
    m_p = nullptr;
  }

  void matrixAddValue(DoFLocalId row, DoFLocalId column, Real value) { impl()->matrixAddValue(row, column, value); }
  void matrixSetValue(DoFLocalId row, DoFLocalId column, Real value) { impl()->matrixSetValue(row, column, value); }
  void applyLinearSystemTransformationAndSolve()
  {
    applyMatrixTransformation();
    applyRHSTransformation();
    solve();
  }
  void applyMatrixTransformation() { impl()->applyMatrixTransformation(); }
  void applyRHSTransformation() { impl()->applyRHSTransformation(); }
  void solve() { impl()->solve(); }
  void clearValues() { impl()->clearValues(); }
  VariableDoFReal& solutionVariable() { return impl()->solutionVariable(); }
  VariableDoFReal& rhsVariable() { return impl()->rhsVariable(); }
  void setSolverCommandLineArguments(const CommandLineArguments& a) { impl()->setSolverCommandLineArguments(a); }
  void setCSRValues(const CSRFormatView& v) { impl()->setCSRValues(v); }
  bool hasSetCSRValues() const { return impl()->hasSetCSRValues(); }
  CSRFormatView& getCSRValues() { return impl()->getCSRValues(); }
  VariableDoFReal& getForcedValue() { return impl()->getForcedValue(); }
  VariableDoFBool& getForcedInfo() { return impl()->getForcedInfo(); }
  VariableDoFByte& getEliminationInfo() { return impl()->getEliminationInfo(); }
  VariableDoFReal& getEliminationValue() { return impl()->getEliminationValue(); }
  void setConstantMatrixSparsity(bool v) { impl()->setConstantMatrixSparsity(v); }
  void setConstantMatrixValues(bool v) { impl()->setConstantMatrixValues(v); }
  bool isMatrixSparsityConstant() { return impl()->isMatrixSparsityConstant(); }
  bool isMatrixValuesConstant() { return impl()->isMatrixValuesConstant(); }
  IDoFLinearSystemImpl* implementation() const { return m_p; }

 private:
  IDoFLinearSystemImpl* impl() const
  {
    // femutils/DoFLinearSystem.cc:_checkInit
    if (!m_p)
      throw FatalError(AFEM_ERR_STATE, "The instance is not initialized. You need to call initialize() before using this class");
    return m_p;
  }
  IDoFLinearSystemImpl* m_p = nullptr;
  IDoFLinearSystemFactory* m_factory = nullptr;
  Mi355xDoFLinearSystemFactory m_default_factory;
};

} // namespace afem::femutils
