// BSRFormat.hpp -- the module-side assembly entry points of ArcaneFEM
// (femutils/BSRFormat.h:143-236) and the GPU boundary-condition helpers
// (femutils/ArcaneFemFunctionsGpu.h:729-816), over the afem C ABI.
//
// The reference's assembleBilinearAtomic / assembleBilinearAtomicFree take a
// device lambda computing the element matrix; a C ABI cannot carry one, so
// the element is named instead (P1 Laplace with an optional per-cell
// conductivity -- the Poisson and Fourier modules -- or P1 linear elasticity).
// Both entry points run the same atomic-free node-wise gather kernels.
#pragma once

#include "DoFLinearSystem.hpp"

#include <cmath>
#include <iomanip>
#include <iostream>
#include <sstream>

namespace afem::femutils {

enum class Element
{
  Laplace,    // modules/poisson/ElementMatrix.h, modules/fourier/ElementMatrix.h
  Elasticity, // modules/elasticity/ElementMatrix.h:151-301
  Heat        // lambda K + M / dt, modules/heat/ElementMatrix.h:28-39, :80-91
};

struct ElementParameters
{
  const Real* cell_lambda = nullptr; // device array, nb_cell values, or null
  Real source_f = 0.0;               // fused constant-source RHS (Laplace)
  Real* rhs = nullptr;               // device array written when source_f is used
  Real lame_lambda = 0.0;
  Real lame_mu = 0.0;
  Real dt = 0.0; // time step of the heat element
};

class BSRFormat
{
 public:
  BSRFormat() = default;
  BSRFormat(const BSRFormat&) = delete;
  BSRFormat& operator=(const BSRFormat&) = delete;
  ~BSRFormat()
  {
    if (m_bsr)
      afem_bsr_destroy(m_bsr);
  }

  // femutils/BSRFormat.cc:350-372
  void initialize(const Mesh& mesh, int nb_dof, bool does_linear_system_use_csr, bool use_atomic_free = false)
  {
    if (m_bsr)
      afem_bsr_destroy(m_bsr);
    m_bsr = nullptr;
    check(afem_bsr_create(mesh.handle(), nb_dof, does_linear_system_use_csr ? 1 : 0, use_atomic_free ? 1 : 0, &m_bsr),
          "BSRFormat(initialize)");
    m_use_csr = does_linear_system_use_csr;
  }
  void computeSparsity() { check(afem_bsr_compute_sparsity(handle()), "BSRFormat(computeSparsity)"); }
  void resetMatrixValues() { check(afem_bsr_reset_values(handle()), "BSRFormat(resetMatrixValues)"); }

  void assembleBilinearAtomic(Element e, const ElementParameters& p) { assemble(e, p); }
  void assembleBilinearAtomicFree(Element e, const ElementParameters& p) { assemble(e, p); }
  void assembleBilinear(Element e, const ElementParameters& p) { assemble(e, p); }

  // femutils/BSRFormat.cc:377-395: hands the (scalar-CSR-ordered) values to a
  // CSR-capable linear system as a non-owning view.
  void toLinearSystem(DoFLinearSystem& ls)
  {
    if (!ls.hasSetCSRValues())
      throw FatalError(AFEM_ERR_ARGUMENT,
                       "BSRFormat(toLinearSystem): Linear system was set to use CSR but is incompatible");
    afem_bsr_info i = info();
    CSRFormatView v;
    v.bsr = handle();
    v.nb_row = i.nb_row * i.nb_block;
    v.nb_value = i.nb_non_zero_value;
    const int32_t* r = nullptr;
    const int32_t* c = nullptr;
    check(afem_bsr_device_arrays(handle(), &r, &c, &v.values, nullptr), "toLinearSystem");
    v.rows = r;
    v.columns = c;
    ls.setCSRValues(v);
  }

  afem_bsr_info info() const
  {
    afem_bsr_info i{};
    check(afem_bsr_get_info(handle(), &i), "BSRFormat(info)");
    return i;
  }
  afem_bsr_t handle() const
  {
    if (!m_bsr)
      throw FatalError(AFEM_ERR_STATE, "BSRFormat: initialize() not called");
    return m_bsr;
  }

 private:
  void assemble(Element e, const ElementParameters& p)
  {
    if (e == Element::Laplace)
      check(afem_bsr_assemble_laplace(handle(), p.cell_lambda, p.source_f, p.rhs), "BSRFormat(assembleBilinear)");
    else if (e == Element::Heat)
      check(afem_bsr_assemble_heat(handle(), p.cell_lambda, p.dt), "BSRFormat(assembleBilinear)");
    else
      check(afem_bsr_assemble_elasticity(handle(), p.lame_lambda, p.lame_mu), "BSRFormat(assembleBilinear)");
  }
  afem_bsr_t m_bsr = nullptr;
  bool m_use_csr = false;
};

namespace BoundaryConditions {

  // femutils/ArcaneFemFunctionsGpu.h:675-708 (node-wise gather form)
  inline void applyConstantSourceToRhs(Real qdot, const Mesh& mesh, int nb_dof, int dof_index, VariableDoFReal& rhs)
  {
    check(afem_rhs_constant_source(mesh.handle(), qdot, nullptr, nb_dof, dof_index, rhs.devicePointer()),
          "applyConstantSourceToRhs");
  }

  // femutils/ArcaneFemFunctions.h:2558-2590 (tri3), :1628-1660 (tet4):
  // rhs += M (scale * field), M the consistent P1 mass matrix; the heat
  // module passes u_old with scale = 1 / dt.
  inline void integrateNodalFieldToRhs(const VariableDoFReal& field, Real scale, const Mesh& mesh,
                                       VariableDoFReal& rhs)
  {
    check(afem_rhs_mass_field(mesh.handle(), field.devicePointer(), scale, 1, 0, rhs.devicePointer()),
          "integrateNodalFieldToRhs");
  }

  // Traction component dof_index of an elasticity system: t |F| / k at each
  // of the k nodes of each face of the group (femutils/ArcaneFemFunctions.h:
  // 2854-2885 edges, :2188-2220 triangles), t int_F N with Gauss points on
  // Quad4 / Hexa8 meshes (:2222-2290 and the Quad4 edge form): the scalar
  // Neumann integral of each component.
  inline void applyTractionToRhs(const Mesh& mesh, const std::string& group, int nb_dof, int dof_index, Real t,
                                 VariableDoFReal& rhs)
  {
    const std::vector<Int32>& faces = mesh.faceGroup(group);
    check(afem_rhs_neumann(mesh.handle(), (int)(faces.size() / mesh.nodesPerFace()), faces.data(), 1, &t, nb_dof,
                           dof_index, rhs.devicePointer()),
          "applyTractionToRhs");
  }

  // modules/elasticity/BodyForce.h:149-161
  inline void applyBodyForceToRhs(const Real f[3], const Mesh& mesh, VariableDoFReal& rhs)
  {
    check(afem_rhs_body_force(mesh.handle(), f, rhs.devicePointer()), "applyBodyForceToRhs");
  }

  // femutils/ArcaneFemFunctionsGpu.cc:269-310 (edges / triangles,
  // femutils/ArcaneFemFunctions.h:2715-2759, :1793-1840; the Gauss-point
  // forms of :741-850 and :1146-1320 on Quad4 / Hexa8 meshes): one
  // value is a scalar flux g, dim values ("NULL" = 0) the flux vector whose
  // component along the outward normal enters the RHS.
  inline void applyNeumannToRhs(const Mesh& mesh, const std::string& group, const std::vector<std::string>& values,
                                VariableDoFReal& rhs)
  {
    std::vector<Real> v;
    if (values.size() == 1) {
      if (values[0] == "NULL")
        return;
      v.push_back(std::stod(values[0]));
    }
    else {
      for (int d = 0; d < mesh.dimension(); ++d)
        v.push_back((size_t)d < values.size() && values[d] != "NULL" ? std::stod(values[d]) : 0.0);
    }
    const std::vector<Int32>& faces = mesh.faceGroup(group);
    check(afem_rhs_neumann(mesh.handle(), (int)(faces.size() / mesh.nodesPerFace()), faces.data(), (int)v.size(),
                           v.data(), 1, 0, rhs.devicePointer()),
          "applyNeumannToRhs");
  }

  // modules/heat/FemModule.cc:302-331, matrix part (first time step only):
  // h times the consistent mass matrix of each face of the group.
  inline void applyConvectionToLhs(const Mesh& mesh, const std::string& group, Real h, BSRFormat& bsr)
  {
    const std::vector<Int32>& faces = mesh.faceGroup(group);
    check(afem_bsr_add_convection(bsr.handle(), (int)(faces.size() / mesh.nodesPerFace()), faces.data(), h),
          "applyConvectionToLhs");
  }

  // modules/heat/FemModule.cc:333-337, RHS part: (h Text) |F| / k at each
  // of the k nodes of each face.
  inline void applyConvectionToRhs(const Mesh& mesh, const std::string& group, Real h, Real t_ext,
                                   VariableDoFReal& rhs)
  {
    const std::vector<Int32>& faces = mesh.faceGroup(group);
    check(afem_rhs_face_uniform(mesh.handle(), (int)(faces.size() / mesh.nodesPerFace()), faces.data(), h * t_ext, 1,
                                0, rhs.devicePointer()),
          "applyConvectionToRhs");
  }

  // femutils/ArcaneFemFunctionsGpu.cc:115-147 for one component: method is
  // "Penalty" (default penalty 1e30, femutils/FemBoundaryConditions.axl:104-109),
  // "RowElimination" or "RowColumnElimination".
  inline void applyDirichletToLhsAndRhs(const std::string& method, const NodeGroup& nodes, int nb_dof_per_node,
                                        int dof_index, Real value, DoFLinearSystem& ls, Real penalty = 1.0e30)
  {
    int m;
    if (method == "Penalty")
      m = AFEM_DIRICHLET_PENALTY;
    else if (method == "RowElimination")
      m = AFEM_DIRICHLET_ROW_ELIMINATION;
    else if (method == "RowColumnElimination")
      m = AFEM_DIRICHLET_ROW_COLUMN_ELIMINATION;
    else
      throw FatalError(AFEM_ERR_ARGUMENT, "Unknown method to enforce Dirichlet BC: '" + method + "'");
    auto* impl = dynamic_cast<Mi355xDoFLinearSystemImpl*>(ls.implementation());
    if (!impl)
      throw FatalError(AFEM_ERR_ARGUMENT, "applyDirichletToLhsAndRhs needs the MI355X linear system");
    check(afem_system_apply_dirichlet(impl->handle(), m, (int32_t)nodes.size(), nodes.data(), nb_dof_per_node,
                                      dof_index, value, penalty),
          "applyDirichletToLhsAndRhs");
  }

} // namespace BoundaryConditions

// femutils/FemUtils.cc:85-171 (checkNodeResultFile): reads "uid v..." lines
// (the first ncomp values of each line: vector results are written as Real3,
// so 2D files carry a zero third component) and counts relative differences
// (Arcane's isNearlyEqualWithEpsilon: equal, or |a-b| / (|a|+|b|) < epsilon),
// skipping pairs below min_value.  Returns the number of differing values and
// throws if any.
inline Int64 checkNodeResultFile(const std::string& filename, const std::vector<Int64>& uids,
                                 const std::vector<Real>& values, int ncomp, double epsilon, double min_value = 0.0)
{
  std::ifstream f(filename);
  if (!f)
    throw FatalError(AFEM_ERR_ARGUMENT, "Invalid result file '" + filename + "'");
  std::map<Int64, std::vector<Real>> ref;
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream in(line);
    Int64 uid;
    if (!(in >> uid))
      continue; // blank line
    std::vector<Real> v((size_t)ncomp);
    for (auto& x : v)
      if (!(in >> x))
        throw FatalError(AFEM_ERR_ARGUMENT, "Error during parsing of file '" + filename + "'");
    ref[uid] = v;
  }
  Int64 nb_error = 0;
  for (size_t n = 0; n < uids.size(); ++n) {
    auto it = ref.find(uids[n]);
    if (it == ref.end())
      continue;
    for (int c = 0; c < ncomp; ++c) {
      const double a = it->second[c], b = values[n * ncomp + c];
      if (std::fabs(a) < min_value && std::fabs(b) < min_value)
        continue;
      const double d = a - b;
      if (d == 0.0)
        continue;
      if (!(std::fabs(d) / (std::fabs(a) + std::fabs(b)) < epsilon)) {
        if (++nb_error < 50)
          std::cout << "ERROR: uid=" << uids[n] << " ref=" << a << " v=" << b << " diff=" << d << "\n";
      }
    }
  }
  if (nb_error > 0)
    throw FatalError(AFEM_ERR_NOT_FOUND, "Error checking values nb_error=" + std::to_string(nb_error));
  return nb_error;
}

} // namespace afem::femutils
