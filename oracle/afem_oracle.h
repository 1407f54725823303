/*
 * afem_oracle.h -- CPU FP64 oracle for the P1 assembly + PCG path.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  This is a plain-C restatement of
 * ArcaneFEM's CPU algorithms for the P1 bilinear-form assembly into CSR/BSR
 * and the Jacobi-PCG solve behind it.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker or as
 * the timed CPU baseline; the product library never calls it.
 *
 * Every function names the reference file:line it restates (paths relative to
 * the arcanefem repository root).  Arithmetic follows the reference's
 * operation order where that is cheap, so that the oracle and the reference
 * round the same way.
 *
 * Pinning (what checks each function independently of its own formula) is
 * listed in tests/test_oracle.py; functions without such a check are marked
 * "parity unpinned" below.
 */
#ifndef AFEM_ORACLE_H
#define AFEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- geometry and P1 gradients ------------------------------------------ */

/* |T| of a triangle: |(n1-n0) x (n2-n0)| / 2.
 * femutils/ArcaneFemFunctionsGpu.h:76-86, femutils/ArcaneFemFunctions.h:180-189 */
double afo_area_tria3(const double* n0, const double* n1, const double* n2);

/* |T| of a tetrahedron: |(n1-n0) . ((n2-n0) x (n3-n0))| / 6.
 * femutils/ArcaneFemFunctionsGpu.h:113-128, femutils/ArcaneFemFunctions.h:156-168 */
double afo_volume_tetra4(const double* n0, const double* n1, const double* n2, const double* n3);

/* dN/dx, dN/dy of the three P1 shape functions (signed 2A denominator).
 * femutils/ArcaneFemFunctionsGpu.h:241-281 */
void afo_grad_tria3(const double* X /* 3x3 row-major */, double dx[3], double dy[3]);

/* dN/dx, dN/dy, dN/dz of the four P1 shape functions (|6V| denominator).
 * femutils/ArcaneFemFunctionsGpu.h:414-535 */
void afo_grad_tetra4(const double* X /* 4x3 row-major */, double dx[4], double dy[4], double dz[4]);

/* ---- element matrices ----------------------------------------------------- */

/* K = |T| lambda (dx dx^T + dy dy^T), 3x3 row-major.
 * modules/poisson/ElementMatrix.h:41-49 (lambda = 1), modules/fourier/ElementMatrix.h:38-47 */
void afo_elem_laplace_tria3(const double* X, double lambda, double K[9]);

/* K = |T| lambda (dx dx^T + dy dy^T + dz dz^T), 4x4 row-major.
 * modules/poisson/ElementMatrix.h:93-102, modules/fourier/ElementMatrix.h:85-95 */
void afo_elem_laplace_tetra4(const double* X, double lambda, double K[16]);

/* 12x12 linear-elasticity tet matrix, DoF order (node, component) interleaved.
 * modules/elasticity/ElementMatrix.h:151-211 (computeElementMatrixTetra4Base) */
void afo_elem_elasticity_tetra4(const double* X, double lame_lambda, double lame_mu, double K[144]);
/* 2D (tri3, 2 DoFs per node): modules/elasticity/ElementMatrix.h:41-58 */
void afo_elem_elasticity_tria3(const double* X, double lame_lambda, double lame_mu, double K[36]);
/* Quad4 / Hexa8 (2 or 3 DoFs per node, Gauss points): modules/elasticity/
 * ElementMatrixHexQuad.h:50-70, :137-187 */
void afo_elem_elasticity_q1(int dim, const double* X, double lame_lambda, double lame_mu, double* K);

/* Lame parameters from (E, nu).  modules/elasticity/FemModule.cc:171-172 */
void afo_lame(double E, double nu, double* lame_lambda, double* lame_mu);

/* ---- mesh connectivity ---------------------------------------------------- */

/* node -> (cell*4 + local slot) lists sorted by cell id, CSR offsets of size
 * nb_node+1.  Cells are stored with stride 4 (triangles leave slot 3 unused).
 * Returns 0, or -1 if a node id is out of range. */
int afo_node_cells(int32_t nb_node, int32_t nb_cell, int32_t npc, const int32_t* cells,
                   int32_t* offsets, int32_t* items);

/* ---- sparsity --------------------------------------------------------------- */

/* Canonical node-graph pattern: row i holds every node sharing a cell with i
 * (i included), sorted ascending.  row_index has nb_node+1 entries.
 * Column *set* of femutils/BSRFormat.cc:799-1006 (atomic path) and :445-790
 * (atomic-free path); both of those emit the same set in a path-dependent
 * order, the canonical form sorts it.  Returns nnz, or -1 if capacity is too
 * small (row_index is still complete so the caller can size columns). */
int64_t afo_pattern(int32_t nb_node, int32_t nb_cell, int32_t npc, const int32_t* cells,
                    int32_t* row_index, int32_t* columns, int64_t capacity);

/* ---- assembly --------------------------------------------------------------- */

enum afo_element { AFO_LAPLACE = 0, AFO_ELASTICITY = 1, AFO_HEAT = 2 };

/* Cells are stored with stride 4 (tri3 padded with -1, tet4, quad4) or 8
 * (hexa8); Quad4 is (dim 2, 4 nodes), Hexa8 (dim 3, 8 nodes). */
#define AFO_CELL_STRIDE(npc) ((npc) > 4 ? (npc) : 4)
#define AFO_IS_Q1(dim, npc) (((dim) == 2 && (npc) == 4) || ((dim) == 3 && (npc) == 8))

/* Q1 elements: physical shape-function gradients and signed det J at
 * (xi, eta[, zeta]) (femutils/ShapeFunctions.h:104-129, :291-350;
 * femutils/ArcaneFemFunctionsGpu.h:296-350, :553-587); the Laplace element
 * matrix with 2x2(x2) Gauss points times lambda
 * (modules/poisson/ElementMatrixHexQuad.h:33-64, :159-195); the constant
 * source vector (femutils/ArcaneFemFunctionsGpu.cc:393-470, :875-960). */
double afo_q1_gradients(int dim, const double* X, double xi, double eta, double zeta, double N[8], double gx[8],
                        double gy[8], double gz[8]);
void afo_elem_laplace_q1(int dim, const double* X, double lambda, double* K);
void afo_source_q1(int dim, const double* X, double qdot, double* b);

/* Cell-wise assembly with a linear column search in each owned row, values
 * accumulated with '+=' (the reference uses atomics on GPU and map inserts on
 * CPU; the sum is the same up to summation order).
 *   femutils/BSRFormat.h:257-303  (ordered per block: value = block*nb^2 + i*nb + j)
 *   femutils/BSRFormat.h:309-370  (ordered per row = scalar CSR order:
 *        value = row_index[r]*nb^2 + nb*(x + i*nnz_row[r]) + j)
 *   modules/poisson/FemModule.cc:398-421 (DOK/cell loop)
 * nb_dof = 1 for AFO_LAPLACE and AFO_HEAT (2D tri or 3D tet; for AFO_HEAT
 * lame_lambda carries the time step dt), 3 for AFO_ELASTICITY (tet).
 * cell_lambda may be NULL (lambda = 1 for every cell).
 * values (nnz*nb_dof^2) are zeroed first.  Returns 0, or -1 when a column is
 * missing from its row (BSRFormat.cc:105 "Value not found"). */
int afo_assemble(int dim, int32_t nb_node, const double* coords /* nb_node x 3 */,
                 int32_t nb_cell, int32_t npc, const int32_t* cells,
                 const int32_t* row_index, const int32_t* columns,
                 int element, int nb_dof, int order_per_block,
                 const double* cell_lambda, double lame_lambda, double lame_mu,
                 double* values);

/* Scalar CSR of a BSR matrix, femutils/BSRFormat.cc:110-172 (BSRMatrix::toCsr):
 * row_index has nb_node*nb+1 entries, each block row expands to nb scalar rows
 * listing (block col)*nb + k; the values array is reused unchanged, so it must
 * already be in per-row order. */
void afo_bsr_to_csr(int32_t nb_node, int nb, const int32_t* row_index, const int32_t* columns,
                    int32_t* csr_row, int32_t* csr_col);

/* BSRMatrix::findValueIndex / CsrDoFLinearSystemImpl::indexValue over the
 * reference's nb_row-entry row array (last row ends at nb_col): the value
 * index of (row, col) in per-block or per-row order, -1 when absent. */
int64_t afo_find_value_index(int32_t nb_row, int nb_block, int order_per_block, const int32_t* rows,
                             int64_t nb_col, const int32_t* columns, const int32_t* nnz_per_row, int32_t row,
                             int32_t col);

/* Laplace values (lambda = 1) and the constant-source RHS f*|T|/npc on `nthreads`
 * threads with atomic updates (timing baseline; see afem_oracle.c). */
int afo_assemble_laplace_threads(int32_t nb_node, const double* coords, int32_t nb_cell, int32_t npc,
                                 const int32_t* cells, const int32_t* row_index, const int32_t* columns,
                                 double f, double* values, double* rhs, int nthreads);
int afo_max_threads(void);

/* ---- right-hand side ---------------------------------------------------------- */

/* b[dof(n)] += f*|T|/npc for every cell T and node n of T.
 * femutils/ArcaneFemFunctions.h:1601-1611 (tet), :2531-2541 (tri);
 * the GPU gather form is femutils/ArcaneFemFunctionsGpu.h:675-708. */
void afo_rhs_constant_source(int dim, const double* coords, int32_t nb_cell, int32_t npc,
                             const int32_t* cells, const double* cell_scale /* NULL = 1 */,
                             double f, int nb_dof, int dof_index, double* rhs);

/* Mass-matrix product of a nodal field (times scale) into the RHS:
 * femutils/ArcaneFemFunctions.h:2558-2590 (tri), :1628-1660 (tet). */
void afo_rhs_mass_field(int dim, const double* coords, int32_t nb_cell, int32_t npc, const int32_t* cells,
                        const double* field, double scale, int nb_dof, int dof_index, double* rhs);
void afo_elem_heat_tria3(const double* X, double lambda, double dt, double K[9]);
void afo_elem_heat_tetra4(const double* X, double lambda, double dt, double K[16]);

/* b[3n+c] += f[c]*V/4 for every tet and node.  modules/elasticity/BodyForce.h:149-161 */
void afo_rhs_body_force(const double* coords, int32_t nb_cell, const int32_t* cells,
                        const double f[3], double* rhs);

/* Convection term of the heat module on boundary faces of k nodes (2 edges,
 * 3 triangles, 4 quadrangles): A[a][b] += h * c * m_ab * |F| for every pair
 * of nodes of F, with m the matrix of FemUtils.h:584-596 (1 off the diagonal,
 * 2 on it) and c = 1/6, 1/12 or 1/20.  modules/heat/FemModule.cc:302-331.
 * The RHS part of the same condition, (h * Text) |F| / k per face node, is
 * afo_rhs_face_uniform with g = h * Text.  |F|: afo_face_measure
 * (computeLengthEdge2, computeAreaTria3, computeAreaQuad4).  Returns -1 when a
 * pair is not in the pattern.  afo_elem_heat_q1: the Quad4 / Hexa8 heat
 * element lambda K + M / dt with the Gauss-point mass matrix
 * (modules/heat/ElementMatrixHexQuad.h:33-130). */
double afo_face_measure(int k, const double* coords, const int32_t* face_nodes);
void afo_rhs_face_uniform(int k, const double* coords, int32_t nb_face, const int32_t* faces, double g,
                          int nb_dof, int dof_index, double* rhs);
void afo_elem_heat_q1(int dim, const double* X, double lambda, double dt, double* K);
int afo_convection_lhs(int k, const double* coords, int32_t nb_face, const int32_t* faces, double h,
                       const int32_t* row_index, const int32_t* columns, double* values);

/* Manufactured solutions (modules/poisson/ManufacturedSolutions.h:91-142,
 * ManufacturedSolutions.cc:66-420): which = 0 sine, 1 poly, 2 quadratic; the
 * exact solution, the source, the RHS on tri3 / Quad4 / Hexa8 meshes (-1 on
 * other cells) and the L2 error of a nodal solution. */
double afo_manufactured_exact(int which, const double* x);
double afo_manufactured_source(int which, const double* x);
int afo_rhs_manufactured(int dim, const double* coords, int32_t nb_cell, int32_t npc, const int32_t* cells,
                         int which, double* rhs);
double afo_manufactured_l2_error(int dim, const double* coords, int32_t nb_cell, int32_t npc, const int32_t* cells,
                                 int which, const double* u);

/* Neumann condition on the faces of a Quad4 mesh (edges) or a Hexa8 mesh
 * (quadrangles, 4 nodes per face) with Gauss points on the face:
 * femutils/ArcaneFemFunctionsGpu.cc:741-850 (edges, outward normal) and
 * :1146-1320 (quadrangles: normal t1 x t2 of the face node order, or turned
 * away from the cell when outward = 1).  -1 when a face has no cell. */
int afo_rhs_neumann_q1(int dim, const double* coords, int32_t nb_cell, int32_t npc, const int32_t* cells,
                       int32_t nb_face, const int32_t* faces, int nb_value, const double* values, int nb_dof,
                       int dof_index, int outward, double* rhs);

/* Neumann condition on boundary faces (dim nodes each): rhs[dof(n)] += g*|F|/dim
 * (nb_value 1) or (g.n_F)*|F|/dim (nb_value dim, n_F the outward unit normal).
 * femutils/ArcaneFemFunctions.h:2715-2759 (edges), :1793-1840 (triangles).
 * Returns -1 when a face is not a face of any cell. */
int afo_rhs_neumann(int dim, const double* coords, int32_t nb_cell, const int32_t* cells, int32_t nb_face,
                    const int32_t* faces, int nb_value, const double* values, int nb_dof, int dof_index,
                    double* rhs);

/* ---- Dirichlet and linear-system transformations --------------------------- */

enum afo_elim { AFO_ELIMINATE_NONE = 0, AFO_ELIMINATE_ROW = 1, AFO_ELIMINATE_ROW_COLUMN = 2 };

/* Applies, in the reference order, what CsrDoFLinearSystemImpl does in
 * applyMatrixTransformation() and applyRHSTransformation():
 *   femutils/CsrDoFLinearSystemImpl.cc:187-230 (row-column map, filled first)
 *   :126-152 (row elimination), :88-121 (row-column elimination, which skips
 *   column index 0), :50-72 (forced / penalty values on the diagonal),
 *   femutils/DoFLinearSystemImplBase.cc:55-88 (rhs lift), :157-182 (rhs := g).
 * forced_info/forced_value implement the penalty method
 * (femutils/ArcaneFemFunctionsGpu.cc:54-78: the caller sets rhs = P*g).
 * The CSR uses nb_row+1 offsets.  Returns 0, or -1 if a forced diagonal is
 * missing from the pattern. */
int afo_apply_transformations(int32_t nb_row, const int32_t* row_index, const int32_t* columns,
                              double* values, double* rhs,
                              const uint8_t* forced_info, const double* forced_value,
                              const uint8_t* elim_info, const double* elim_value);

/* ---- sparse kernels and the Jacobi-PCG solve ------------------------------ */

/* y = A x (CSR, nb_row+1 offsets). */
void afo_spmv(int32_t nb_row, const int32_t* row_index, const int32_t* columns,
              const double* values, const double* x, double* y);

/* Jacobi-preconditioned CG in FP64 with the relative 2-norm stopping test
 * ||r_k||_2 <= rtol ||b||_2 on the recursively updated residual.  x holds the
 * initial guess on entry (row-eliminated systems need x0 = g on the
 * eliminated rows, which keeps the iteration on the symmetric reduced system)
 * and the solution on exit.
 * Semantics of PETSc cg+jacobi (femutils/PetscDoFLinearSystemFactory.axl:5-10)
 * and of Hypre ParCSR PCG with TwoNorm=1 (femutils/HypreDoFLinearSystem.cc:596-604).
 * Iteration counts and residual histories of those third-party solvers are
 * not reproduced (parity unpinned); the solution is pinned by direct solves.
 * Returns the iteration count (>= 0), or -1 when max_iter was reached;
 * *true_relres receives ||b - A x||_2 / ||b||_2 recomputed at the end. */
int afo_pcg_jacobi(int32_t nb_row, const int32_t* row_index, const int32_t* columns,
                   const double* values, const double* b, double* x,
                   double rtol, int max_iter, double* true_relres);

/* afo_pcg_jacobi on nthreads OpenMP threads (reduction order may differ):
 * the all-cores CPU timing baseline of the solve. */
int afo_pcg_jacobi_threads(int32_t n, const int32_t* row_index, const int32_t* columns, const double* values,
                           const double* b, double* x, double rtol, int max_iter, int nthreads,
                           double* true_relres);

/* ---- result comparison ------------------------------------------------------- */

/* Relative comparison of femutils/FemUtils.cc:85-171 (checkNodeResultFile):
 * a value differs when not (|ref|<min_value and |v|<min_value) and not
 * nearly-equal, nearly-equal(a,b) = (a == b) or |a-b|/(|a|+|b|) < epsilon
 * (Arcane's TypeEqualT<Real>::isNearlyEqualWithEpsilon).  Returns the number
 * of differing entries. */
int64_t afo_count_differences(int64_t n, const double* ref, const double* v,
                              double epsilon, double min_value);

/* Structured-mesh helper used by the CPU baseline: wall-clock seconds. */
double afo_wtime(void);

#ifdef __cplusplus
}
#endif

#endif /* AFEM_ORACLE_H */
