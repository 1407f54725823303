/*
 * afem.h -- C ABI of the MI355X-native P1 assembly + PCG path.
 *
 * This is the drop-in boundary for ArcaneFEM's hot path: the BSR/CSR
 * assembly entry points a physics module calls (femutils/BSRFormat.h:170-236)
 * and the linear-system implementation an IDoFLinearSystemFactory hands back
 * (femutils/IDoFLinearSystemFactory.h:42-43,
 * femutils/internal/IDoFLinearSystemImpl.h:45-69).  Each entry point below
 * cites the reference interface it replaces.
 *
 * Conventions
 *  - Every function returns an afem_status (0 = AFEM_OK).  Errors never cross
 *    the ABI as exceptions; afem_last_error() gives the message of the last
 *    failing call on the calling thread.  The C++ adapter in
 *    arcanefem_amd/csrc/femutils/ turns them back into exceptions, the
 *    reference's convention (ARCANE_FATAL / ARCANE_THROW).
 *  - Pointers documented [host] are host memory; [device] are device memory of
 *    the current device (afem_malloc or any hipMalloc'ed buffer).
 *  - All work is ordered on one library stream; calls return once the work
 *    is enqueued unless documented as synchronous.  afem_synchronize() waits.
 *  - Handles are not thread-safe (as DoFLinearSystem::matrixAddValue is not,
 *    femutils/AlephDoFLinearSystem.cc:125-135).
 *  - Local ids: node lid i has DoFs i*nb_dof + k (femutils/FemDoFsOnNodes.cc:101-112);
 *    the library's own BSR row arrays carry nb_row+1 offsets; a CSR view
 *    handed to afem_system_set_csr_values carries the reference's nb_row.
 */
#ifndef AFEM_H
#define AFEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AFEM_API __attribute__((visibility("default")))

typedef enum afem_status {
  AFEM_OK = 0,
  AFEM_ERR_ARGUMENT = 1,        /* ARCANE_THROW(ArgumentException, ...) */
  AFEM_ERR_NOT_IMPLEMENTED = 2, /* ARCANE_THROW(NotImplementedException, ...) */
  AFEM_ERR_STATE = 3,           /* call made out of order */
  AFEM_ERR_NOT_FOUND = 4,       /* BSRMatrix::findValueIndex "Value not found" */
  AFEM_ERR_MESH = 5,            /* invalid connectivity or a degenerate cell */
  AFEM_ERR_CAPACITY = 6,        /* a row exceeds every kernel variant's capacity */
  AFEM_ERR_NOT_CONVERGED = 7,   /* solver hit max_iter before the tolerance */
  AFEM_ERR_DEVICE = 8,          /* no usable device */
  AFEM_ERR_HIP = 9,             /* a HIP runtime call failed */
  AFEM_ERR_OUT_OF_MEMORY = 10
} afem_status;

/* ---- runtime -------------------------------------------------------------- */

AFEM_API const char* afem_last_error(void);
AFEM_API int afem_version(void);
/* Selects the device and creates the library stream (default priority). */
AFEM_API int afem_init(int device);
AFEM_API int afem_finalize(void);
AFEM_API int afem_synchronize(void);

/* Plain device-memory plumbing for callers without an allocator of their own. */
AFEM_API int afem_malloc(void** dptr, size_t bytes);
AFEM_API int afem_free(void* dptr);
AFEM_API int afem_memcpy_htod(void* dst_device, const void* src_host, size_t bytes);
AFEM_API int afem_memcpy_dtoh(void* dst_host, const void* src_device, size_t bytes);
AFEM_API int afem_memcpy_dtod(void* dst_device, const void* src_device, size_t bytes);
AFEM_API int afem_memset(void* dptr, int value, size_t bytes);
/* Device-to-device copy by the library's own streaming kernel (16-byte
 * non-temporal loads and stores, grid-stride) on the library stream, timed
 * under the profile name "stream_copy": the measured HBM bandwidth the
 * roofline figures are set beside.  dst and src 16-byte aligned, bytes a
 * multiple of 16.  No counterpart in the reference. */
AFEM_API int afem_stream_copy(void* dst, const void* src, size_t bytes);

/* Per-kernel timing with HIP events recorded on the library stream around each
 * launch.  query returns the number of recorded launches of the named kernel
 * and their summed duration in milliseconds (synchronous). */
AFEM_API int afem_profile_enable(int on);
AFEM_API int afem_profile_reset(void);
AFEM_API int afem_profile_query(const char* kernel, int64_t* launches, double* total_ms);

/* ---- synthetic meshes [host] ------------------------------------------------
 * Structured unit square (dim 2: each quad split into 2 triangles along the
 * (0,0)-(1,1) diagonal) or unit cube (dim 3: each hex split into 6 Kuhn tets
 * sharing the (0,0,0)-(1,1,1) diagonal) with n cells per side.  Nodes are
 * numbered lexicographically (x fastest); interior nodes are moved by
 * jitter*h*U(-1,1) per coordinate from a counter-based generator keyed by seed.
 * Cells are written with stride 4 (triangles: slot 3 = -1) and positive
 * orientation.  Stand-in for the reference's mesh files (meshes/msh). */
AFEM_API int afem_box_mesh_size(int dim, int n, int32_t* nb_node, int32_t* nb_cell);
AFEM_API int afem_box_mesh(int dim, int n, double jitter, uint64_t seed,
                           double* coords /* nb_node*3 */, int32_t* cells /* nb_cell*4 */);

/* ---- mesh ---------------------------------------------------------------------
 * Device copy of the node coordinates and the cell->node connectivity, plus
 * the node->cell connectivity built on the device (sorted by cell id).  It
 * stands in for Arcane's IMesh and UnstructuredMeshConnectivityView
 * (cellNode / nodeCell) that BSRFormat::initialize receives
 * (femutils/BSRFormat.cc:350-372). */
typedef struct afem_mesh* afem_mesh_t;

/* coords [host] nb_node*3, cells [host] nb_cell*stride.  nodes_per_cell is 3
 * (tri3) or 4 (quad4) in dim 2, 4 (tet4) or 8 (hexa8) in dim 3; the stride is
 * 4, or 8 for hexa8 (triangles pad slot 3 with -1).  Quad4 nodes go
 * counter-clockwise, Hexa8 the bottom face counter-clockwise then the top
 * face (the Gmsh / Arcane local order, femutils/ShapeFunctions.h:104-129,
 * :291-350).  On Quad4 / Hexa8 meshes the sparsity, the Laplace assembly
 * (element matrices once per cell, then a row gather; variant 0, and rows of
 * more than 128 entries, rebuild each element row per (node, cell) pair; no
 * autotuning) and the constant-source RHS run; the
 * other element kernels return AFEM_ERR_NOT_IMPLEMENTED or AFEM_ERR_ARGUMENT.
 * (Heat element, mass-weighted field, convection, Neumann and the
 * face-uniform RHS also accept Q1 meshes; elasticity and body force do not.)
 * Fails with AFEM_ERR_MESH on out-of-range node ids. */
AFEM_API int afem_mesh_create(int dim, int32_t nb_node, const double* coords,
                              int32_t nb_cell, int32_t nodes_per_cell, const int32_t* cells,
                              afem_mesh_t* mesh);
AFEM_API int afem_mesh_destroy(afem_mesh_t mesh);
/* node->cell lists as (cell*4 + local slot) [host]: offsets nb_node+1, items 4*nb_cell max. */
AFEM_API int afem_mesh_node_cells(afem_mesh_t mesh, int32_t* offsets, int32_t* items);

/* ---- BSR matrix (femutils/BSRFormat.h:143-236, BSRMatrix :77-137) ----------- */
typedef struct afem_bsr* afem_bsr_t;

typedef struct afem_bsr_info {
  int32_t nb_row;            /* block rows = nodes (BSRMatrix::nbRow) */
  int32_t nb_block;          /* block size = DoFs per node (BSRMatrix::nbBlock) */
  int64_t nb_col;            /* stored blocks = N + 2E for P1 (BSRMatrix::nbColumn) */
  int64_t nb_non_zero_value; /* nb_col * nb_block^2 (BSRMatrix::nbNonZero) */
  int32_t order_values_per_block; /* BSRMatrix::orderValuePerBlock */
  int32_t max_row_length;    /* longest block row */
} afem_bsr_info;

/* BSRFormat::initialize(mesh, nb_dof, does_linear_system_use_csr, use_atomic_free)
 * (femutils/BSRFormat.cc:350-372).  Values are ordered per row (scalar CSR
 * order) when does_linear_system_use_csr, per block otherwise.  The
 * atomic-free flag is accepted for API parity; the MI355X path always builds
 * the pattern and assembles without floating-point atomics.  nb_dof is 1
 * (scalar), 2 (2D elasticity) or 3 (3D elasticity). */
AFEM_API int afem_bsr_create(afem_mesh_t mesh, int nb_dof, int does_linear_system_use_csr,
                             int use_atomic_free, afem_bsr_t* bsr);
AFEM_API int afem_bsr_destroy(afem_bsr_t bsr);
/* BSRFormat::computeSparsity (femutils/BSRFormat.cc:1011-1018): rows sorted
 * ascending (the canonical form of the reference's column set). */
AFEM_API int afem_bsr_compute_sparsity(afem_bsr_t bsr);
AFEM_API int afem_bsr_get_info(afem_bsr_t bsr, afem_bsr_info* info);
/* BSRFormat::resetMatrixValues (femutils/BSRFormat.cc:1032-1035). */
AFEM_API int afem_bsr_reset_values(afem_bsr_t bsr);

/* BSRFormat::assembleBilinearAtomic / assembleBilinearAtomicFree with the P1
 * Laplace element: modules/poisson/ElementMatrix.h:41-63, 93-118, with a
 * per-cell conductivity modules/fourier/ElementMatrix.h:38-95.
 *   cell_lambda [device] nb_cell values or NULL (lambda = 1);
 *   rhs [device] nb_row values or NULL: when given, the same kernel also
 *   writes the constant-source right-hand side b_i = sum_T f*|T|/npc
 *   (femutils/ArcaneFemFunctionsGpu.h:675-708), overwriting rhs.
 * Requires nb_dof = 1.  Values are overwritten (no prior reset needed). */
AFEM_API int afem_bsr_assemble_laplace(afem_bsr_t bsr, const double* cell_lambda,
                                       double source_f, double* rhs);

/* Assembly kernel variants of the Laplace element (no counterpart in the
 * reference, which has one lambda per format): 0 searches each column in the
 * LDS-staged row; 1 reads the per-(node, cell) pair plan built by
 * afem_bsr_compute_sparsity; 2 / 3 the sliced block-local plan and 4 / 5 the
 * brick (Morton-ordered) plan, with 8 / 16 plan words in flight per thread;
 * 6 builds each cell's element matrix once per brick and adds it into the
 * rows with LDS atomics (values reproducible to rounding, not bit for bit).
 * Plans 2-6 are built by the first assembly that needs them; a variant whose
 * plan does not fit falls back (cell -> brick -> sliced -> pair plan -> search).
 * autotune times each variant on this matrix, keeps the fastest and reports
 * it; set_assembly_variant pins one (-1 = default, variant 5).
 * The autotuner assembles with lambda = 1 and no RHS: it overwrites the
 * matrix values, so assemble the system afterwards. */
AFEM_API int afem_bsr_autotune_assembly(afem_bsr_t bsr, double* best_ms, int* variant);
AFEM_API int afem_bsr_set_assembly_variant(afem_bsr_t bsr, int variant);

/* Heat element lambda K + M / dt (modules/heat/ElementMatrix.h:28-39 tri,
 * :80-91 tet; M the consistent mass matrix, femutils/FemUtils.h:341-355):
 * the Laplace assembly (cell_lambda [device] or NULL = 1, current variant)
 * followed by a pass adding M / dt.  On Quad4 / Hexa8 meshes the Gauss-point
 * element of modules/heat/ElementMatrixHexQuad.h:33-130 (M_ab = sum_g N_a N_b
 * det J), built once per cell.  Values only; nb_dof = 1, dt > 0. */
AFEM_API int afem_bsr_assemble_heat(afem_bsr_t bsr, const double* cell_lambda, double dt);

/* Same with the 12x12 P1 linear-elasticity element
 * (modules/elasticity/ElementMatrix.h:151-301) into 3x3 blocks; nb_dof = 3.
 * One thread per nonzero block sums the block over the cells of its edge,
 * from cell geometry staged once per workgroup in LDS; assembly variant 0
 * selects the row-per-thread kernel with a column search instead, variant 1
 * the block kernel that rebuilds the geometry for every (block, cell) use.  On a triangle mesh with nb_dof = 2: the 6x6 2D element
 * (modules/elasticity/ElementMatrix.h:41-58), one thread per node's two rows;
 * a linear system takes such a matrix in the per-row (CSR) value order. */
AFEM_API int afem_bsr_assemble_elasticity(afem_bsr_t bsr, double lame_lambda, double lame_mu);

/* Device arrays of the matrix: row_index (nb_row+1), columns (nb_col),
 * values (nb_non_zero_value), nnz_per_row (nb_row).  Owned by the handle. */
AFEM_API int afem_bsr_device_arrays(afem_bsr_t bsr, const int32_t** row_index,
                                    const int32_t** columns, double** values,
                                    const int32_t** nnz_per_row);
/* Copy to host (synchronous); any pointer may be NULL. */
AFEM_API int afem_bsr_download(afem_bsr_t bsr, int32_t* row_index, int32_t* columns,
                               double* values, int32_t* nnz_per_row);
/* BSRMatrix::toCsr (femutils/BSRFormat.cc:110-172) done on the device: scalar
 * CSR rows (nb_row*nb_block+1) and columns (nb_non_zero_value) owned by the
 * handle; the values array is shared unchanged (per-row order required for
 * nb_block > 1). */
AFEM_API int afem_bsr_to_csr(afem_bsr_t bsr, const int32_t** csr_row, const int32_t** csr_col,
                             int32_t* nb_csr_row);

/* ---- right-hand side ---------------------------------------------------------- */

/* FemModulePoisson::_applyManufacturedSource
 * (modules/poisson/ManufacturedSolutions.cc:66-236): rhs[n*nb_dof+dof_index]
 * = the manufactured source integrated against N_n, written (not added).
 * solution: 0 sine (2 pi^2 sin(pi x) sin(pi y)), 1 poly
 * (2 (x(1-x) + y(1-y))), 2 quadratic (-4), ManufacturedSolutions.h:91-142.
 * Triangles: source(barycentre) |T| / 3; Quad4: 2x2 Gauss; Hexa8: 3x3x3 Gauss
 * (ManufacturedSolutions.cc:22-40).
 * Tetrahedra: AFEM_ERR_NOT_IMPLEMENTED (the reference has no branch). */
AFEM_API int afem_rhs_manufactured(afem_mesh_t mesh, int solution, int nb_dof, int dof_index, double* rhs);

/* FemModulePoisson::_computeManufacturedL2Error
 * (modules/poisson/ManufacturedSolutions.cc:270-420): *l2_error [host] =
 * sqrt(sum over cells of the squared error of the nodal solution u [device,
 * one value per node]): triangles at the barycentre (mean of u) times |T|,
 * Quad4 at the 2x2, Hexa8 at the 3x3x3 Gauss points times det J w_g.  The per-cell terms
 * are summed on the host in cell order. */
AFEM_API int afem_manufactured_l2_error(afem_mesh_t mesh, int solution, const double* u, double* l2_error);

/* BoundaryConditions::applyConstantSourceToRhs, node-wise gather
 * (femutils/ArcaneFemFunctionsGpu.h:675-708): rhs[n*nb_dof+dof_index] =
 * sum over cells T of n of f*scale_T*|T|/npc.  cell_scale [device] or NULL. */
AFEM_API int afem_rhs_constant_source(afem_mesh_t mesh, double f, const double* cell_scale,
                                      int nb_dof, int dof_index, double* rhs);
/* Convection term of the heat module on boundary faces (modules/heat/
 * FemModule.cc:302-331, the matrix part, added once before the time loop):
 * for each face F (face_nodes [host]: nb_face * k node ids, k = dim, or 4 on
 * a Hexa8 mesh) and each pair (a, b) of its nodes, A[a][b] += h c m_ab |F|
 * with m_ab = 2 (a = b) or 1 and c = 1/6 (edges), 1/12 (triangles) or 1/20
 * (quadrangles).  Summed per entry in face order (no atomics).  The RHS
 * part, (h Text) |F| / k at each face node, is afem_rhs_face_uniform with
 * g = h * Text.  nb_dof = 1; AFEM_ERR_NOT_FOUND when a pair is outside the
 * pattern (a face that is not a face of the mesh), values untouched. */
AFEM_API int afem_bsr_add_convection(afem_bsr_t bsr, int nb_face, const int32_t* face_nodes, double h);
/* rhs[i*nb_dof+dof_index] += g |F| / k for each node i of each face F
 * (face_nodes [host]: nb_face * k node ids, k = dim, or 4 on a Hexa8 mesh;
 * |F| by computeLengthEdge2 / computeAreaTria3 / computeAreaQuad4): the
 * convection RHS (h Text) measure / num_nodes of modules/heat/FemModule.cc:
 * 333-337 with g = h * Text.  Summed per node in face order. */
AFEM_API int afem_rhs_face_uniform(afem_mesh_t mesh, int nb_face, const int32_t* face_nodes, double g, int nb_dof,
                                   int dof_index, double* rhs);
/* BoundaryConditions::applyNeumannToRhs (femutils/ArcaneFemFunctions.h:
 * 2715-2759 edges, :1793-1840 triangles; GPU femutils/ArcaneFemFunctionsGpu.cc:
 * 679-738, and :741-850 / :1146-1320 for Quad4 / Hexa8 meshes): for each
 * boundary face F (face_nodes [host]: nb_face * k node ids, k = dim, or 4 on
 * a Hexa8 mesh) and each node i of F, rhs[i*nb_dof+dof_index] += g |F| / dim
 * (nb_value = 1) or (g . n_F) |F| / dim (nb_value = dim, values [host] the
 * vector g, n_F the outward unit normal of F) on simplex meshes; on Quad4 /
 * Hexa8 meshes the same integrals of g N_i over F with 2 (edge) or 2x2
 * (quadrangle) Gauss points.  A quadrangle's normal is t1 x t2 of its node
 * order, as the reference takes it from Arcane's face order (first the node
 * of smallest unique id, then towards its neighbour of smaller id): pass the
 * faces in that order.  The faces' values are summed per node in face order
 * (no atomics).  AFEM_ERR_MESH when a face is not a face of any cell. */
AFEM_API int afem_rhs_neumann(afem_mesh_t mesh, int nb_face, const int32_t* face_nodes, int nb_value,
                              const double* values, int nb_dof, int dof_index, double* rhs);
/* integrateNodalFieldToRhs (femutils/ArcaneFemFunctions.h:2558-2590 tri,
 * :1628-1660 tet, :2609-2660 quad, :1680-1740 hexa) on scale * field, gather
 * form: rhs[i*nb_dof+dof_index] += (M (scale * field))_i with M the
 * consistent mass matrix (Gauss points on Quad4 / Hexa8).  field [device],
 * one value per node (the heat module's M u_old / dt with scale = 1 / dt). */
AFEM_API int afem_rhs_mass_field(afem_mesh_t mesh, const double* field, double scale, int nb_dof,
                                 int dof_index, double* rhs);
/* Elasticity body force, modules/elasticity/BodyForce.h:149-161, gather form:
 * rhs[3n+c] += sum over tets T of n of f[c]*|T|/4.  f [host] 3 values. */
AFEM_API int afem_rhs_body_force(afem_mesh_t mesh, const double* f, double* rhs);

/* ---- linear system (IDoFLinearSystemImpl, femutils/internal/IDoFLinearSystemImpl.h:37-70) */
typedef struct afem_system* afem_system_t;

enum afem_dirichlet_method {
  AFEM_DIRICHLET_PENALTY = 0,               /* femutils/ArcaneFemFunctionsGpu.cc:54-78 */
  AFEM_DIRICHLET_ROW_ELIMINATION = 1,       /* :83-105 with ELIMINATE_ROW */
  AFEM_DIRICHLET_ROW_COLUMN_ELIMINATION = 2 /* :83-105 with ELIMINATE_ROW_COLUMN */
};

/* IDoFLinearSystemFactory::createInstance(sd, dof_family, solver_name):
 * nb_dof is the size of the DoF family.  The instance owns the RHS, solution,
 * forced and elimination DoF arrays (femutils/DoFLinearSystemImplBase.cc:26-36),
 * zero-initialised. */
AFEM_API int afem_system_create(int32_t nb_dof, afem_system_t* sys);
AFEM_API int afem_system_destroy(afem_system_t sys);
/* IDoFLinearSystemImpl::setCSRValues(CSRFormatView) with the four spans of
 * the reference's view (femutils/CsrFormatMatrixView.h:145-187):
 *   rows [device] nb_row row starts, the view's rows(); the last row ends at
 *     nnz (= nbColumn()), as CsrFormatMatrixView::rowRange takes it.  nb_row
 *     is nb_dof, or nb_dof + 1 when the array carries a closing entry equal
 *     to nnz (the scalar expansion of BSRMatrix::toCsr for blocks > 1,
 *     femutils/BSRFormat.cc:117-139);
 *   rows_nb_column [device] nb_dof per-row counts (nbColumnForRow) or NULL;
 *     when given they are checked against the row extents;
 *   columns [device] nnz column ids in [0, nb_dof); values [device] nnz.
 * columns and values are kept as a non-owning view that must stay valid until
 * solve (femutils/DoFLinearSystem.h:318-325); the row starts are copied (and
 * closed with nnz) into an array the system owns.  Synchronous: a view whose
 * starts are not non-decreasing from 0, whose counts or closing entry
 * disagree, or whose columns leave the DoF family fails with
 * AFEM_ERR_ARGUMENT. */
AFEM_API int afem_system_set_csr_values(afem_system_t sys, int32_t nb_row, int64_t nnz,
                                        const int32_t* rows, const int32_t* rows_nb_column,
                                        const int32_t* columns, double* values);
/* setCSRValues from a BSR handle (BSRFormat::toLinearSystem, femutils/BSRFormat.cc:377-395).
 * With nb_block > 1 the solver keeps the block structure for its SpMV and
 * block-Jacobi preconditioner. */
AFEM_API int afem_system_set_bsr(afem_system_t sys, afem_bsr_t bsr);
/* rhsVariable(), solutionVariable(), getForcedInfo/Value(), getEliminationInfo/Value()
 * as device arrays of nb_dof entries. */
AFEM_API int afem_system_rhs(afem_system_t sys, double** rhs);
AFEM_API int afem_system_solution(afem_system_t sys, double** x);
AFEM_API int afem_system_forced(afem_system_t sys, uint8_t** info, double** value);
AFEM_API int afem_system_elimination(afem_system_t sys, uint8_t** info, double** value);
/* BoundaryConditions::applyDirichletToLhsAndRhs for one node group and one DoF
 * component (femutils/ArcaneFemFunctionsGpu.cc:115-147):
 *   penalty: forced_info = 1, forced_value = penalty, rhs = penalty*value;
 *   row / row-column elimination: elimination info and value.
 * nodes [host] node local ids; DoF = node*nb_dof_per_node + dof_index. */
AFEM_API int afem_system_apply_dirichlet(afem_system_t sys, int method, int32_t nb_node,
                                         const int32_t* nodes, int nb_dof_per_node,
                                         int dof_index, double value, double penalty);
/* CsrDoFLinearSystemImpl::applyMatrixTransformation / applyRHSTransformation
 * (femutils/CsrDoFLinearSystemImpl.cc:235-253), including the reference's
 * row-column elimination rule that leaves column index 0 untouched
 * (femutils/CsrDoFLinearSystemImpl.cc:111). */
AFEM_API int afem_system_apply_matrix_transformation(afem_system_t sys);
AFEM_API int afem_system_apply_rhs_transformation(afem_system_t sys);
/* CsrDoFLinearSystemImpl::matrixAddValue / matrixSetValue
 * (femutils/internal/CsrDoFLinearSystemImpl.h:54-63): finds (row, column) in
 * the CSR view and adds to (set = 0) or overwrites (set = 1) its value.
 * AFEM_ERR_NOT_FOUND when the entry is not in the pattern. */
AFEM_API int afem_system_matrix_add_value(afem_system_t sys, int32_t row, int32_t column, double value, int set);

/* clearValues (femutils/DoFLinearSystemImplBase.cc:42-48): forced and
 * elimination arrays reset and the CSR view dropped. */
AFEM_API int afem_system_clear_values(afem_system_t sys);

enum afem_solver_method {
  AFEM_SOLVER_PCG_FP64 = 0,      /* FP64 PCG (the Jacobi-PCG of PETSc cg+jacobi) */
  AFEM_SOLVER_PCG_MIXED = 1,     /* FP32 SpMV/preconditioner/vectors, FP64 dots and
                                    FP64 reliable-update residual r = b - A x */
  AFEM_SOLVER_CG_FP64 = 2,       /* unpreconditioned FP64 CG */
  AFEM_SOLVER_FPCG_SPAI = 3      /* flexible PCG (Polak-Ribiere beta) with the sparse
                                    preconditioner matrix given by afem_system_set_spai /
                                    afem_spai_apply, FP32 SpMVs, FP64 reliable updates;
                                    falls back to Jacobi mixed PCG from the current iterate
                                    if the preconditioner is not positive on a residual */
};
enum afem_precond {
  AFEM_PRECOND_JACOBI = 0,       /* diagonal (scalar) or 3x3 block-diagonal inverse */
  AFEM_PRECOND_NONE = 1
};

typedef struct afem_solver_options {
  int method;            /* afem_solver_method */
  int preconditioner;    /* afem_precond */
  double rtol;           /* stop when ||b - A x||_2 <= max(rtol*||b||_2, atol) */
  double atol;
  int max_iter;
  int use_initial_guess; /* 1: start from the solution array; 0: x0 = 0 except
                            eliminated rows, which start at their value */
  int check_every;       /* iterations between host convergence polls (0 = 16) */
  double reliable_delta; /* mixed: residual reduction factor between FP64
                            reliable updates (0 = 0.1) */
  int spmv_lanes;        /* SpMV format: 0 = autotuned, > 0 = CSR with that many
                            lanes per row (4, 8, 16, 32), -1 = SELL-64,
                            -2 = SELL-64 with 16-bit column offsets (scalar
                            matrices whose |col - row| fits; else SELL-64),
                            -3 = the same with 8 slots per step, -4 / -5 =
                            -2 / -3 with 512-thread workgroups */
  int norm;              /* stopping norm: 0 = ||b - A x||_2 <= max(rtol ||b||_2, atol);
                            1 = natural norm sqrt(r.M^-1 r) <= rtol * its initial value;
                            2 = preconditioned norm ||M^-1 r||_2 <= max(rtol ||M^-1 b||_2,
                            atol) (PETSc's default test for CG, KSP_NORM_PRECONDITIONED;
                            it scales penalty rows by 1/1e30, so it is the one to use
                            when penalty rows make ||b|| meaningless).  1 and 2 have no
                            true-residual restart and are FP64 PCG only. */
  int precond_bf16;      /* AFEM_SOLVER_FPCG_SPAI: store the preconditioner values as
                            BF16 (with the 16-bit-offset SELL format; FP32 otherwise) */
} afem_solver_options;

typedef struct afem_solve_result {
  int iterations;
  int converged;
  int reliable_updates;
  double relres;         /* FP64 true relative residual ||b - A x||_2 / ||b||_2
                            (norm 2: ||M^-1 r||_2 / ||M^-1 b||_2 of the last iteration) */
  double solve_ms;       /* wall time of the solve on the library stream */
  int fallback;          /* AFEM_SOLVER_FPCG_SPAI finished with Jacobi PCG */
} afem_solve_result;

AFEM_API void afem_solver_default_options(afem_solver_options* opt);
/* IDoFLinearSystemImpl::solve.  Synchronous.  Returns AFEM_ERR_NOT_CONVERGED
 * (with the result filled) when max_iter is reached. */
AFEM_API int afem_system_solve(afem_system_t sys, const afem_solver_options* opt,
                               afem_solve_result* result);
/* No counterpart in the reference.  For callers that cannot allocate device or
 * pinned memory (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): afem_system_solve for the FP64 methods (AFEM_SOLVER_PCG_FP64,
 * AFEM_SOLVER_CG_FP64; Jacobi, block-Jacobi or no preconditioner; stopping norms
 * 0-2), with the SpMV on the system's own CSR / 3x3 BSR view (16 lanes per row)
 * and every work vector, reduction partial and the solver state in a fixed
 * device array of the library (56 MiB per device, held until the process ends).
 * Allocates nothing.  At most 1048576 DoFs (AFEM_ERR_ARGUMENT beyond).  Every
 * call on one device shares that array: calls that overlap on a device race.
 * Synchronous, like afem_system_solve; the same result and error conventions. */
AFEM_API int afem_system_solve_noalloc(afem_system_t sys, const afem_solver_options* opt,
                                       afem_solve_result* result);

/* ---- sparse kernels exposed for testing and benchmarking ---------------------- */

/* y = A x with the system's matrix (FP64). x, y [device] nb_dof values. */
AFEM_API int afem_system_spmv(afem_system_t sys, const double* x, double* y);
/* y = A32 x32 with the FP32 copy of the system's matrix (made on first use).
 * The copy is in difference form when every non-empty row (block row) has
 * its diagonal entry: off-diagonal entries rounded to FP32, the diagonal
 * entry (block) replaced by the row sum taken in FP64, and the product
 * formed as s_i x_i + sum_{j != i} a_ij (x_j - x_i) -- the same value, without
 * the cancellation a rounded diagonal leaves on smooth vectors. */
AFEM_API int afem_system_spmv_fp32(afem_system_t sys, const float* x, float* y);
/* Installs a sparse preconditioner M with the pattern of the system's matrix:
 * values [device] nnz FP32 in the order of the matrix values; z = M r.  Used
 * by AFEM_SOLVER_FPCG_SPAI until the matrix pattern changes.  No counterpart
 * in the reference (its preconditioners live inside Hypre / PETSc). */
AFEM_API int afem_system_set_spai(afem_system_t sys, const float* values);

/* SpMV format the autotuner chose for the system's matrix: *format = 0 (CSR,
 * *lanes lanes per row, or per 3x3 block row when the matrix has blocks),
 * 1 (SELL-64), 2 (SELL-64 with 16-bit column offsets) or 3 (the same, 8
 * slots per step); -1 while no SpMV or solve has run on the current matrix. */
AFEM_API int afem_system_spmv_format(afem_system_t sys, int* format, int* lanes);

/* Kernel shape of the SELL-64 16-bit-offset SpMV (formats 2 and 3): slots per
 * step and workgroup size (256 or 512 threads) of the FP64 SpMV, and of the
 * FP32 SpMV and preconditioner product, which the mixed solvers and
 * afem_system_spmv_fp32 tune separately.  All zero for the other formats.
 * No counterpart in the reference. */
AFEM_API int afem_system_spmv_tuning(afem_system_t sys, int* fp64_slots, int* fp64_block, int* fp32_slots,
                                     int* fp32_block);

/* ---- heterogeneous coefficients (no counterpart in the reference) ------------------- */

/* Random-Fourier-feature parameters of a Gaussian random field with
 * squared-exponential covariance sigma^2 exp(-|r|^2 / (2 l^2)):
 * omega [host] nfeatures x 3 (w_k ~ N(0, I / l^2); zero z-component in 2D),
 * phase [host] nfeatures (phi_k ~ U[0, 2 pi)).  Host only. */
AFEM_API int afem_rff_parameters(int dim, int nfeatures, double corr_length, uint64_t seed,
                                 double* omega, double* phase);
/* lambda_c = exp(g(x_c)) at the cell centroids x_c [device, nb_cell values],
 * g(x) = sigma sqrt(2/K) sum_k cos(w_k . x + phi_k) with the parameters above
 * (phase evaluated in FP32).  The per-cell conductivity of
 * afem_bsr_assemble_laplace (modules/fourier/ElementMatrix.h:38-95). */
AFEM_API int afem_lognormal_field(afem_mesh_t mesh, double sigma, double corr_length,
                                  int nfeatures, uint64_t seed, double* cell_lambda);

/* ---- learned sparse approximate inverse (no counterpart in the reference) ------------
 * A 3-layer message-passing network (width 32, BF16 weights and activations,
 * FP32 accumulation on MFMA) on the graph of the Jacobi-scaled matrix
 * S = D^-1/2 A D^-1/2 predicts Mh with the pattern of A; the preconditioner
 * is M = D^-1/2 Mh D^-1/2.  Training minimises (1/8N) sum_p ||z_p - Mh S z_p||^2
 * over 8 Rademacher probes per system.  Systems must be scalar, with the
 * Dirichlet transformation applied (those rows are taken as identity). */
typedef struct afem_spai* afem_spai_t;
/* Number of FP32 parameters of the model. */
AFEM_API int afem_spai_num_params(void);
/* New model; the untrained model is the Jacobi preconditioner. */
AFEM_API int afem_spai_create(uint64_t seed, afem_spai_t* spai);
AFEM_API int afem_spai_destroy(afem_spai_t spai);
AFEM_API int afem_spai_get_params(afem_spai_t spai, float* params, int n);
/* Replaces the parameters and resets the optimiser state. */
AFEM_API int afem_spai_set_params(afem_spai_t spai, const float* params, int n);
/* Probe loss of the system (forward only). */
AFEM_API int afem_spai_loss(afem_spai_t spai, afem_system_t sys, uint64_t probe_seed, double* loss);
/* Loss and its gradient (kept for afem_spai_adam_step; copied to grad [host,
 * afem_spai_num_params() floats] when grad is not NULL). */
AFEM_API int afem_spai_loss_grad(afem_spai_t spai, afem_system_t sys, uint64_t probe_seed,
                                 double* loss, float* grad);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the last gradient. */
AFEM_API int afem_spai_adam_step(afem_spai_t spai, double lr);
/* afem_spai_loss_grad followed by afem_spai_adam_step. */
AFEM_API int afem_spai_train_step(afem_spai_t spai, afem_system_t sys, uint64_t probe_seed,
                                  double lr, double* loss);
/* Forward pass on the system and afem_system_set_spai with the resulting M. */
AFEM_API int afem_spai_apply(afem_spai_t spai, afem_system_t sys);
/* Arrays of the last forward pass [host floats], for testing: 0 = S values
 * (nnz), 1 = node features (N x 16), 2 = h3 (N x 32), 3 = u (N x 32),
 * 4 = Mh (nnz), 5 = M (nnz), 6 = D^-1/2 (N), 7 = h0 (N x 32). */
AFEM_API int afem_spai_inspect(afem_spai_t spai, int what, float* out, int64_t count);

/* ---- batched systems (no counterpart in the reference) ------------------------------
 * Independent scalar systems that are the diagonal blocks of one system's
 * matrix (e.g. the Laplace matrix of a disjoint union of meshes, assembled in
 * one launch by afem_bsr_assemble_laplace).  They are solved together by
 * mixed-precision Jacobi PCG with per-system scalars: FP32 matrix and vectors,
 * FP64 x, FP64 reliable-update residual deciding convergence.  Either chunks
 * of iterations of ordinary kernels captured in a hipGraph and replayed (the
 * default), or one persistent kernel launch (captured in a hipGraph and
 * replayed) in which every system gets a team of resident workgroups that
 * synchronize through bounded barrier waits in device memory. */
typedef struct afem_batch* afem_batch_t;
typedef struct afem_batch_options {
  double rtol;            /* per system: ||b_s - A_s x_s|| <= max(rtol ||b_s||, atol) */
  double atol;
  int max_iter;
  double reliable_delta;  /* reliable-update reduction factor (0 = 0.1) */
  int chunk;              /* chunked solve: iterations per captured graph launch (even; 0 = 32) */
  int use_graph;          /* 1: replay a hipGraph (per chunk, or of the persistent launch); 0: plain launches */
  int use_initial_guess;  /* 1: start from the system's solution array */
  int persistent;         /* 0: chunks of iterations (default); 1: one persistent launch, each
                             system's team on one XCD when the device dispatches workgroups to XCDs
                             round-robin (probed once per batch) and the teams fit, exchanging
                             vectors through that XCD's L2, else as 2; 2: one persistent launch,
                             vectors exchanged through the device's coherence point */
  int workgroups;         /* persistent: workgroups of the launch (0 = as many as the device holds
                             at once; capped there; at least one per system) */
  long long spin_limit;   /* persistent: polls before a barrier wait gives up (0 = 2^21); the solve
                             then fails with AFEM_ERR_DEVICE instead of waiting on */
  double timeout_ms;      /* persistent: wall time after which the host raises the cancel flag that
                             every barrier wait polls (0 = 60000) */
} afem_batch_options;
typedef struct afem_batch_result {
  int iterations;
  int converged;
  int reliable_updates;
  double relres;          /* FP64 true relative residual of the system */
} afem_batch_result;
AFEM_API void afem_batch_default_options(afem_batch_options* opt);
/* segment_rows [host] nb_segment + 1 row offsets: 0 = r_0 < r_1 < ... = nb_dof. */
AFEM_API int afem_batch_create(afem_system_t sys, int nb_segment, const int32_t* segment_rows,
                               afem_batch_t* batch);
AFEM_API int afem_batch_destroy(afem_batch_t batch);
/* Lanes per matrix row (4, 8, 16 over the CSR arrays; 1 = one thread per row
 * over a SELL-64 copy whose slices start at every segment's first row, with
 * 16-bit column offsets -- CSR with 8 lanes when an offset does not fit) and
 * 256-row strips per workgroup (1..16). */
AFEM_API int afem_batch_set_tiles(afem_batch_t batch, int lanes, int tiles);
/* Times the (lanes, tiles) candidates on one chunk and keeps the fastest. */
AFEM_API int afem_batch_autotune(afem_batch_t batch, int* lanes, int* tiles, double* ms_per_iteration);
/* Solves every system; results [host] nb_segment entries (may be NULL).
 * AFEM_ERR_NOT_CONVERGED when a system misses the tolerance;
 * AFEM_ERR_ARGUMENT when a non-empty matrix row has no diagonal entry (the
 * FP32 product keeps each row sum in the diagonal slot). */
AFEM_API int afem_batch_solve(afem_batch_t batch, const afem_batch_options* opt,
                              afem_batch_result* results, double* total_ms);
/* Exchange used by the last persistent solve of the batch: 1 XCD-local teams,
 * 2 coherence point, 0 no persistent solve yet. */
AFEM_API int afem_batch_persistent_mode(afem_batch_t batch, int* mode);

#ifdef __cplusplus
}
#endif

#endif /* AFEM_H */
