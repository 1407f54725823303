// afem_module -- the Poisson, Fourier, elasticity and heat module flows of
// ArcaneFEM (modules/poisson/FemModule.cc:75-146, modules/fourier/FemModule.cc:
// 160-356, modules/elasticity/FemModule.cc:57-154, modules/heat/FemModule.cc:
// 74-420) driven through the C++ DoFLinearSystem /
// BSRFormat layer and the MI355X linear-system factory, on a reference Gmsh
// mesh, a structured box or an Arcane Cartesian mesh.
//
//   afem_module poisson --mesh circle_cut.msh --f 5.5 --dirichlet horizontal=0.5 \
//       [--neumann curved=-0.35,1.65] \
//       --check poisson_test_ref_circle_2D.txt --epsilon 1e-4
//   afem_module elasticity --mesh bar_dynamic_3D.msh --E 21e5 --nu 0.28 --f -1.0 \
//       --dirichlet surfaceleft=0.0,0.0,0.0 --dirichlet surfaceright=NULL,1.0,NULL \
//       --check bar.3D.Dirichlet.bodyForce.txt --epsilon 1e-3 --min-value 1e-10
//   afem_module fourier --mesh multi-material.msh --lambda 0 --material Mat1=100 \
//       --material Mat2=1 --qdot 15 --dirichlet Left=50 --dirichlet Right=5 \
//       --check test2_results.txt
//   afem_module heat --mesh plate.msh --lambda 1.75 --dt 0.4 --tmax 20 --Tinit 30 \
//       --penalty 1e31 --dirichlet left=10.0 [--convection right=1.0,20.0] \
//       --check 2d_conduction.txt
//   afem_module poisson --mesh mms.quad4_n4.msh --manufactured sine --mms-tolerance 8e-2
//   afem_module elasticity --cartesian 10,2 1.0,0.1 --f -1.0,NULL --dirichlet XMIN=0.0,0.0 \
//       --dirichlet XMAX=NULL,1.0 --check bar.2D.cartesian.Dirichlet.bodyForce.txt
//
// Exit status 0 when the solve converged and the check (if any) passed.
#include "femutils/BSRFormat.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <map>
#include <cstdio>
#include <cstdlib>

using namespace afem::femutils;

namespace {

struct Args
{
  std::string physics, mesh, check, method = "Penalty", format = "BSR", precision = "fp64";
  int box_dim = 0, box_n = 0;
  std::vector<int> cart_n;      // --cartesian nx,ny[,nz] lx,ly[,lz]
  std::vector<double> cart_len;
  double box_jitter = 0.0;
  unsigned long long box_seed = 0;
  std::vector<double> f;
  std::vector<std::pair<std::string, std::vector<std::string>>> dirichlet, neumann, convection, traction;
  std::vector<std::pair<std::string, double>> materials; // fourier: per-volume-group lambda
  bool has_lambda = false;
  // modules/poisson/Fem.axl:47-53: manufactured-solution name and tolerance
  std::string manufactured;
  double mms_tolerance = 1.0e-1;
  double E = 21e5, nu = 0.28, rtol = 1e-12, epsilon = 1e-4, min_value = 0.0;
  // heat (modules/heat/Fem.axl defaults) and the Dirichlet penalty
  // (femutils/FemBoundaryConditions.axl:104-109)
  double lambda = 1.75, dt = 0.1, tmax = 1.0, t_init = 0.0, penalty = 1.0e30;
  bool print = false;
  bool parse_only = false; // read the mesh on the host, print a summary, exit
};

[[noreturn]] void usage(const std::string& msg)
{
  std::fprintf(stderr, "afem_module: %s\n", msg.c_str());
  std::exit(2);
}

std::vector<std::string> split(const std::string& s, char sep)
{
  std::vector<std::string> out;
  std::string cur;
  for (char c : s) {
    if (c == sep) {
      out.push_back(cur);
      cur.clear();
    }
    else
      cur += c;
  }
  out.push_back(cur);
  return out;
}

Args parse(int argc, char** argv)
{
  Args a;
  if (argc < 2)
    usage("expected 'poisson', 'elasticity' or 'heat'");
  a.physics = argv[1];
  for (int k = 2; k < argc; ++k) {
    std::string key = argv[k];
    auto next = [&]() -> std::string {
      if (k + 1 >= argc)
        usage("missing value after " + key);
      return argv[++k];
    };
    if (key == "--mesh")
      a.mesh = next();
    else if (key == "--box") {
      a.box_dim = std::stoi(next());
      a.box_n = std::stoi(next());
      a.box_jitter = std::stod(next());
      a.box_seed = std::stoull(next());
    }
    else if (key == "--cartesian") {
      // Arcane's Cartesian2D / Cartesian3D generator: cells and lengths per axis
      for (auto& t : split(next(), ','))
        a.cart_n.push_back(std::stoi(t));
      for (auto& t : split(next(), ','))
        a.cart_len.push_back(std::stod(t));
      if ((a.cart_n.size() != 2 && a.cart_n.size() != 3) || a.cart_len.size() != a.cart_n.size())
        usage("--cartesian expects nx,ny[,nz] lx,ly[,lz]");
      for (size_t d = 0; d < a.cart_n.size(); ++d)
        if (a.cart_n[d] < 1 || !(a.cart_len[d] > 0.0))
          usage("--cartesian expects positive cell counts and lengths");
    }
    else if (key == "--f") {
      for (auto& s : split(next(), ','))
        a.f.push_back(s == "NULL" ? 0.0 : std::stod(s));
    }
    else if (key == "--neumann") {
      // group=value (scalar flux) or group=gx,gy[,gz] (flux vector)
      const std::string s = next();
      const size_t eq = s.rfind('=');
      if (eq == std::string::npos || eq == 0)
        usage("--neumann expects group=value[,value,value]");
      a.neumann.push_back({ s.substr(0, eq), split(s.substr(eq + 1), ',') });
    }
    else if (key == "--dirichlet" || key == "--dirichlet-point") {
      // point conditions (applyPointDirichletToLhsAndRhs,
      // femutils/ArcaneFemFunctionsGpu.cc:157-184) are node-group conditions
      // on a named Gmsh point
      // group=value[,value,value]; the group may itself be a plane "x=0"
      const std::string s = next();
      const size_t eq = s.rfind('=');
      if (eq == std::string::npos || eq == 0 || eq + 1 == s.size())
        usage("--dirichlet expects group=value[,value,value]");
      a.dirichlet.push_back({ s.substr(0, eq), split(s.substr(eq + 1), ',') });
    }
    else if (key == "--traction") {
      // group=tx,ty[,tz] with NULL for a free component (elasticity)
      const std::string s = next();
      const size_t eq = s.rfind('=');
      if (eq == std::string::npos || eq == 0 || eq + 1 == s.size())
        usage("--traction expects group=tx,ty[,tz]");
      a.traction.push_back({ s.substr(0, eq), split(s.substr(eq + 1), ',') });
    }
    else if (key == "--convection") {
      // group=h,Text (modules/heat/Fem.axl convection-boundary-condition)
      const std::string s = next();
      const size_t eq = s.rfind('=');
      const std::vector<std::string> v = eq == std::string::npos ? std::vector<std::string>{} : split(s.substr(eq + 1), ',');
      if (eq == std::string::npos || eq == 0 || v.size() != 2)
        usage("--convection expects group=h,Text");
      a.convection.push_back({ s.substr(0, eq), v });
    }
    else if (key == "--lambda") {
      a.lambda = std::stod(next());
      a.has_lambda = true;
    }
    else if (key == "--qdot") // the Fourier module's name for the source term
      a.f = { std::stod(next()) };
    else if (key == "--material") {
      // group=lambda (modules/fourier/Fem.axl material-property)
      const std::string s = next();
      const size_t eq = s.rfind('=');
      if (eq == std::string::npos || eq == 0 || eq + 1 == s.size())
        usage("--material expects group=lambda");
      a.materials.push_back({ s.substr(0, eq), std::stod(s.substr(eq + 1)) });
    }
    else if (key == "--manufactured")
      a.manufactured = next();
    else if (key == "--mms-tolerance")
      a.mms_tolerance = std::stod(next());
    else if (key == "--dt")
      a.dt = std::stod(next());
    else if (key == "--tmax")
      a.tmax = std::stod(next());
    else if (key == "--Tinit")
      a.t_init = std::stod(next());
    else if (key == "--penalty")
      a.penalty = std::stod(next());
    else if (key == "--method")
      a.method = next();
    else if (key == "--format")
      a.format = next();
    else if (key == "--precision")
      a.precision = next();
    else if (key == "--E")
      a.E = std::stod(next());
    else if (key == "--nu")
      a.nu = std::stod(next());
    else if (key == "--rtol")
      a.rtol = std::stod(next());
    else if (key == "--check")
      a.check = next();
    else if (key == "--epsilon")
      a.epsilon = std::stod(next());
    else if (key == "--min-value")
      a.min_value = std::stod(next());
    else if (key == "--print")
      a.print = true;
    else if (key == "--parse-only")
      a.parse_only = true;
    else
      usage("unknown option " + key);
  }
  if (a.mesh.empty() && a.box_dim == 0 && a.cart_n.empty())
    usage("give --mesh, --box or --cartesian");
  if (a.physics != "fourier" && !a.materials.empty())
    usage("--material is available for fourier");
  if (a.physics != "elasticity" && !a.traction.empty())
    usage("--traction is available for elasticity");
  if (!a.manufactured.empty() && (!a.dirichlet.empty() || !a.neumann.empty() || !a.f.empty()))
    usage("--manufactured sets its own source and boundary values");
  if (!a.manufactured.empty() && a.physics != "poisson")
    usage("--manufactured is available for poisson");
  if (!a.manufactured.empty() && a.manufactured != "sine" && a.manufactured != "poly" &&
      a.manufactured != "quadratic")
    usage("--manufactured expects sine, poly or quadratic");
  return a;
}

// group name, or "x=0" style plane selectors for boxes
NodeGroup group_of(const Mesh& m, const std::string& name)
{
  if (name.size() > 2 && name[1] == '=' && (name[0] == 'x' || name[0] == 'y' || name[0] == 'z'))
    return m.nodesOnPlane(name[0] - 'x', std::stod(name.substr(2)));
  return m.group(name);
}

// Nodes of the outer faces (faces of exactly one cell): Arcane's
// mesh()->outerFaces().nodeGroup(), used by the manufactured solutions
// (modules/poisson/ManufacturedSolutions.cc:248-262).
NodeGroup outer_face_nodes(const Mesh& mesh)
{
  static const int tri[3][2] = { { 0, 1 }, { 1, 2 }, { 2, 0 } };
  static const int quad[4][2] = { { 0, 1 }, { 1, 2 }, { 2, 3 }, { 3, 0 } };
  static const int tet[4][3] = { { 0, 1, 2 }, { 0, 1, 3 }, { 0, 2, 3 }, { 1, 2, 3 } };
  static const int hexa[6][4] = { { 0, 1, 2, 3 }, { 4, 5, 6, 7 }, { 0, 1, 5, 4 },
                                  { 1, 2, 6, 5 }, { 2, 3, 7, 6 }, { 3, 0, 4, 7 } };
  const int npc = mesh.nodesPerCell(), stride = npc > 4 ? npc : 4;
  int nf, k;
  const int* table;
  if (mesh.dimension() == 2 && npc == 3) { nf = 3; k = 2; table = &tri[0][0]; }
  else if (mesh.dimension() == 2 && npc == 4) { nf = 4; k = 2; table = &quad[0][0]; }
  else if (npc == 4) { nf = 4; k = 3; table = &tet[0][0]; }
  else { nf = 6; k = 4; table = &hexa[0][0]; }
  std::map<std::array<Int32, 4>, int> count;
  const std::vector<Int32>& cells = mesh.cells();
  for (Int32 c = 0; c < mesh.nbCell(); ++c)
    for (int f = 0; f < nf; ++f) {
      std::array<Int32, 4> key = { -1, -1, -1, -1 };
      for (int j = 0; j < k; ++j)
        key[j] = cells[(size_t)c * stride + table[f * k + j]];
      std::sort(key.begin(), key.begin() + k);
      ++count[key];
    }
  std::vector<char> on((size_t)mesh.nbNode(), 0);
  for (auto& kv : count)
    if (kv.second == 1)
      for (int j = 0; j < k; ++j)
        on[(size_t)kv.first[j]] = 1;
  NodeGroup g;
  for (Int32 i = 0; i < mesh.nbNode(); ++i)
    if (on[(size_t)i])
      g.push_back(i);
  return g;
}

// modules/poisson/ManufacturedSolutions.h:91-142 (the exact solutions)
double manufactured_exact(const std::string& name, double x, double y)
{
  const double pi = 3.141592653589793238462643383279502884;
  if (name == "sine")
    return std::sin(pi * x) * std::sin(pi * y);
  if (name == "poly")
    return x * (1.0 - x) * y * (1.0 - y);
  return x * x + y * y;
}

int manufactured_id(const std::string& name)
{
  return name == "sine" ? 0 : name == "poly" ? 1 : 2;
}

void apply_dirichlet(const Args& a, const Mesh& mesh, int nb_dof, DoFLinearSystem& ls)
{
  for (auto& d : a.dirichlet) {
    const NodeGroup& g = group_of(mesh, d.first);
    for (size_t c = 0; c < d.second.size() && (int)c < nb_dof; ++c) {
      if (d.second[c] == "NULL")
        continue;
      BoundaryConditions::applyDirichletToLhsAndRhs(a.method, g, nb_dof, (int)c, std::stod(d.second[c]), ls,
                                                    a.penalty);
    }
  }
}

// modules/heat/FemModule.cc: the matrix lambda K + M/dt (and the convection
// matrix) is assembled at the first step only (:455-488, :302-331 under
// t <= dt - 1e-8); every step rebuilds the RHS -- convection h Text, the old
// temperature M u_old/dt (:350-362), Neumann, Dirichlet -- and solves.  The
// time loop stops after the solve made at the first t >= tmax (:79-82 with
// :128-135).  The BSR values stay in place between steps, so the matrix
// transformation (penalty diagonal, eliminated rows) runs once.
int run_heat(const Args& a, Mesh& mesh, Runner& runner)
{
  if (a.dt <= 0.0)
    usage("--dt must be positive");
  const Int32 n = mesh.nbNode();
  auto t0 = std::chrono::steady_clock::now();
  Mi355xDoFLinearSystemFactory factory;
  DoFLinearSystem ls;
  ls.setLinearSystemFactory(&factory);
  ls.initialize(&runner, n, "Solver");
  char rtol_text[32];
  std::snprintf(rtol_text, sizeof rtol_text, "%.17g", a.rtol);
  CommandLineArguments solver_args;
  solver_args.args = { "-ksp_rtol", rtol_text, "-afem_precision", a.precision };
  if (a.method == "Penalty") {
    solver_args.args.push_back("-ksp_norm_type");
    solver_args.args.push_back("preconditioned");
  }
  ls.setSolverCommandLineArguments(solver_args);
  BSRFormat bsr;
  bsr.initialize(mesh, 1, true, a.format == "AF-BSR");
  bsr.computeSparsity();

  OwnedDoFVariable<Real> cell_lambda(mesh.nbCell()), t_old(n); // m_cell_lambda, m_node_temperature_old
  cell_lambda.fill(a.lambda);
  t_old.fill(a.t_init);
  ElementParameters ep;
  ep.cell_lambda = cell_lambda.devicePointer();
  ep.dt = a.dt;
  bsr.assembleBilinear(Element::Heat, ep);
  bsr.toLinearSystem(ls);
  std::vector<std::pair<double, double>> conv;
  for (auto& c : a.convection) {
    conv.push_back({ std::stod(c.second[0]), std::stod(c.second[1]) });
    BoundaryConditions::applyConvectionToLhs(mesh, c.first, conv.back().first, bsr);
  }

  auto* impl = dynamic_cast<Mi355xDoFLinearSystemImpl*>(ls.implementation());
  VariableDoFReal& rhs = ls.rhsVariable();
  double t = 0.0;
  int nb_solve = 0, iterations = 0;
  for (;;) {
    const bool last = t >= a.tmax;
    rhs.fill(0.0);
    for (size_t k = 0; k < a.convection.size(); ++k)
      BoundaryConditions::applyConvectionToRhs(mesh, a.convection[k].first, conv[k].first, conv[k].second, rhs);
    BoundaryConditions::integrateNodalFieldToRhs(t_old.variable(), 1.0 / a.dt, mesh, rhs);
    for (auto& nm : a.neumann)
      BoundaryConditions::applyNeumannToRhs(mesh, nm.first, nm.second, rhs);
    apply_dirichlet(a, mesh, 1, ls);
    if (nb_solve == 0)
      ls.applyMatrixTransformation();
    ls.applyRHSTransformation();
    ls.solve();
    const afem_solve_result& r = impl->lastSolve();
    iterations += r.iterations;
    if (!r.converged) {
      std::fprintf(stderr, "afem_module heat: no convergence at t=%g (relres %.3e)\n", t, r.relres);
      return 1;
    }
    t_old.copyFrom(ls.solutionVariable());
    ++nb_solve;
    if (last)
      break;
    t += a.dt;
  }
  auto t1 = std::chrono::steady_clock::now();
  std::printf("afem_module heat: nodes=%d dofs=%d steps=%d iterations=%d t=%g time_ms=%.1f\n", n, n, nb_solve,
              iterations, t, std::chrono::duration<double, std::milli>(t1 - t0).count());
  std::vector<Real> u = ls.solutionVariable().toHost();
  if (a.print)
    for (Int32 i = 0; i < n; ++i)
      std::printf("u[%lld] = %.15g\n", (long long)mesh.uniqueIds()[i], u[(size_t)i]);
  if (!a.check.empty()) {
    checkNodeResultFile(a.check, mesh.uniqueIds(), u, 1, a.epsilon, a.min_value);
    std::printf("CHECK OK %s\n", a.check.c_str());
  }
  return 0;
}

} // namespace

int main(int argc, char** argv)
{
  Args a;
  try {
    a = parse(argc, argv);
  }
  catch (const std::exception& e) { // std::stod / std::stoi on a malformed number
    usage(std::string("malformed number in the arguments (") + e.what() + ")");
  }
  try {
    if (a.parse_only) { // modules/poisson/CMakeLists.txt "2D_parse_and_exit"
      Mesh host_mesh;
      host_mesh.readGmsh41(a.mesh, false);
      long long sum_uid = 0;
      for (Int64 u : host_mesh.uniqueIds())
        sum_uid += u;
      double sx = 0.0;
      for (Real x : host_mesh.coords())
        sx += x;
      std::printf("dim=%d nodes=%d cells=%d uid_sum=%lld coord_sum=%.17g\n", host_mesh.dimension(),
                  host_mesh.nbNode(), host_mesh.nbCell(), sum_uid, sx);
      for (auto& g : host_mesh.groups()) {
        long long s = 0;
        for (Int32 v : g.second)
          s += v;
        std::printf("group %s %zu %lld\n", g.first.c_str(), g.second.size(), s);
      }
      return 0;
    }
    Runner runner{ 0 };
    check(afem_init(runner.device), "afem_init");
    Mesh mesh;
    if (!a.mesh.empty())
      mesh.readGmsh41(a.mesh);
    else if (!a.cart_n.empty())
      mesh.buildCartesian((int)a.cart_n.size(), a.cart_n.data(), a.cart_len.data());
    else
      mesh.buildBox(a.box_dim, a.box_n, a.box_jitter, a.box_seed);
    if (a.physics == "heat")
      return run_heat(a, mesh, runner);
    const bool elasticity = a.physics == "elasticity";
    const bool fourier = a.physics == "fourier";
    if (!elasticity && !fourier && a.physics != "poisson")
      usage("physics must be poisson, fourier, elasticity or heat");
    if (!a.convection.empty())
      usage("--convection is available for heat");
    const int nb_dof = elasticity ? mesh.dimension() : 1; // modules/elasticity/FemModule.cc:174
    const Int32 n_dofs = mesh.nbNode() * nb_dof;
    auto t0 = std::chrono::steady_clock::now();

    // compute(): linear system from the factory, BSR format with CSR values
    Mi355xDoFLinearSystemFactory factory;
    DoFLinearSystem ls;
    ls.setLinearSystemFactory(&factory);
    ls.initialize(&runner, n_dofs, "Solver");
    CommandLineArguments solver_args;
    char rtol_text[32]; // std::to_string would print 1e-12 as "0.000000"
    std::snprintf(rtol_text, sizeof rtol_text, "%.17g", a.rtol);
    solver_args.args = { "-ksp_rtol", rtol_text, "-afem_precision", a.precision };
    ls.setSolverCommandLineArguments(solver_args);
    BSRFormat bsr;
    bsr.initialize(mesh, nb_dof, true, a.format == "AF-BSR");
    bsr.computeSparsity();

    // _assembleBilinearOperator
    ElementParameters ep;
    std::unique_ptr<OwnedDoFVariable<Real>> cell_lambda;
    if (elasticity) {
      ep.lame_mu = a.E / (2 * (1 + a.nu)); // modules/elasticity/FemModule.cc:171-172
      ep.lame_lambda = a.E * a.nu / ((1 + a.nu) * (1 - 2 * a.nu));
      bsr.assembleBilinear(Element::Elasticity, ep);
    }
    else if (fourier) {
      // modules/fourier/FemModule.cc:172-194: lambda everywhere, then per material group
      std::vector<Real> lam((size_t)mesh.nbCell(), a.has_lambda ? a.lambda : 1.0);
      for (auto& mt : a.materials)
        for (Int32 c : mesh.cellGroup(mt.first))
          lam[(size_t)c] = mt.second;
      cell_lambda = std::make_unique<OwnedDoFVariable<Real>>(mesh.nbCell());
      cell_lambda->variable().fromHost(lam);
      ep.cell_lambda = cell_lambda->devicePointer();
      bsr.assembleBilinear(Element::Laplace, ep);
    }
    else {
      bsr.assembleBilinear(Element::Laplace, ep);
    }
    bsr.toLinearSystem(ls);

    // _assembleLinearOperator: source / body force, then Dirichlet
    VariableDoFReal& rhs = ls.rhsVariable();
    rhs.fill(0.0);
    const bool simplex_mesh = mesh.nodesPerCell() == mesh.dimension() + 1;
    if (elasticity && (mesh.dimension() == 2 || !simplex_mesh)) {
      // modules/elasticity/BodyForce.h:52-161: f_c area / 3 per node and
      // component on triangles, f_c int N (Gauss points) on Quad4 / Hexa8
      for (int c = 0; c < nb_dof && (size_t)c < a.f.size(); ++c)
        if (a.f[c] != 0.0)
          BoundaryConditions::applyConstantSourceToRhs(a.f[c], mesh, nb_dof, c, rhs);
    }
    else if (elasticity) {
      double f3[3] = { 0.0, 0.0, 0.0 };
      for (size_t k = 0; k < a.f.size() && k < 3; ++k)
        f3[k] = a.f[k];
      BoundaryConditions::applyBodyForceToRhs(f3, mesh, rhs);
    }
    else if (!a.manufactured.empty()) {
      // _applyManufacturedSource then _applyManufacturedDirichlet
      // (modules/poisson/FemModule.cc:181-185): the source replaces the
      // module's RHS; every outer-face node gets the penalty diagonal and
      // penalty * exact(x) in the RHS.
      check(afem_rhs_manufactured(mesh.handle(), manufactured_id(a.manufactured), 1, 0, rhs.devicePointer()),
            "applyManufacturedSource");
      const NodeGroup outer = outer_face_nodes(mesh);
      BoundaryConditions::applyDirichletToLhsAndRhs("Penalty", outer, 1, 0, 0.0, ls, a.penalty);
      std::vector<Real> b = rhs.toHost();
      for (Int32 i : outer)
        b[(size_t)i] = a.penalty * manufactured_exact(a.manufactured, mesh.coords()[3 * (size_t)i],
                                                      mesh.coords()[3 * (size_t)i + 1]);
      rhs.fromHost(b);
    }
    else if (!a.f.empty()) {
      BoundaryConditions::applyConstantSourceToRhs(a.f[0], mesh, 1, 0, rhs);
    }
    // femutils/ArcaneFemFunctions.h:2854-2885 (edges), :2188-2220 (triangles):
    // t_c |F| / k per node and component
    for (auto& tr : a.traction) {
      for (int c = 0; c < nb_dof && (size_t)c < tr.second.size(); ++c) {
        if (tr.second[c] != "NULL")
          BoundaryConditions::applyTractionToRhs(mesh, tr.first, nb_dof, c, std::stod(tr.second[c]), rhs);
      }
    }
    if (!a.neumann.empty() && elasticity)
      usage("--neumann is available for poisson, fourier and heat");
    for (auto& nm : a.neumann)
      BoundaryConditions::applyNeumannToRhs(mesh, nm.first, nm.second, rhs);
    apply_dirichlet(a, mesh, nb_dof, ls);

    // _solve.  Penalty rows (1e30 on the diagonal, 1e30*g in the RHS) make
    // ||b|| meaningless, so those systems stop on the preconditioned norm
    // ||M^-1 r|| / ||M^-1 b||, in which a penalty row weighs like any other.
    auto* impl = dynamic_cast<Mi355xDoFLinearSystemImpl*>(ls.implementation());
    if (a.method == "Penalty" || !a.manufactured.empty())
      ls.setSolverCommandLineArguments(CommandLineArguments{ { "-ksp_norm_type", "preconditioned" } });
    ls.applyLinearSystemTransformationAndSolve();
    auto t1 = std::chrono::steady_clock::now();
    const afem_solve_result& r = impl->lastSolve();
    std::printf("afem_module %s: nodes=%d dofs=%d iterations=%d relres=%.3e converged=%d time_ms=%.1f\n",
                a.physics.c_str(), mesh.nbNode(), n_dofs, r.iterations, r.relres, r.converged,
                std::chrono::duration<double, std::milli>(t1 - t0).count());

    // _updateVariables / _validateResults
    std::vector<Real> u = ls.solutionVariable().toHost();
    if (a.print)
      for (Int32 i = 0; i < mesh.nbNode(); ++i) {
        std::printf("u[%lld] =", (long long)mesh.uniqueIds()[i]);
        for (int c = 0; c < nb_dof; ++c)
          std::printf(" %.15g", u[(size_t)i * nb_dof + c]);
        std::printf("\n");
      }
    if (!a.check.empty()) {
      checkNodeResultFile(a.check, mesh.uniqueIds(), u, nb_dof, a.epsilon, a.min_value);
      std::printf("CHECK OK %s\n", a.check.c_str());
    }
    if (!a.manufactured.empty()) { // modules/poisson/FemModule.cc:469-476
      double l2 = 0.0;
      check(afem_manufactured_l2_error(mesh.handle(), manufactured_id(a.manufactured),
                                       ls.solutionVariable().devicePointer(), &l2),
            "computeManufacturedL2Error");
      std::printf("Manufactured solution L2 error = %.6e\n", l2);
      if (l2 > a.mms_tolerance)
        throw FatalError(AFEM_ERR_ARGUMENT, "Manufactured solution L2 error is greater than tolerance");
    }
  }
  catch (const FatalError& e) {
    std::fprintf(stderr, "afem_module: fatal (status %d): %s\n", e.status, e.what());
    return 1;
  }
  catch (const std::exception& e) { // e.g. a condition value that is not a number
    std::fprintf(stderr, "afem_module: invalid input: %s\n", e.what());
    return 2;
  }
  return 0;
}
