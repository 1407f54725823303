// linsys.hip -- the linear-system instance behind IDoFLinearSystemImpl: DoF
// arrays, the CSR/BSR matrix view, Dirichlet boundary conditions and the
// matrix / RHS transformations of CsrDoFLinearSystemImpl, on the device.
//
// The transformation kernels walk scalar rows.  A scalar row r of a matrix
// viewed with block size NB (1: scalar CSR, 3: BSR with values per block) is
// the node row a = r / NB, component i = r % NB; its entries are
// (column, value index) pairs produced by RowEntries below.
#include "common.h"
#include "device_utils.h"
#include "linsys.h"

namespace afem {

template <int NB>
struct RowEntries
{
  const int32_t* rp;
  const int32_t* ci;
  // entry range of scalar row r, as block range [b, e) and component i
  __device__ __forceinline__ void range(int32_t r, int32_t& b, int32_t& e, int& i) const
  {
    const int32_t a = r / NB;
    i = r - a * NB;
    b = rp[a];
    e = rp[a + 1];
  }
  __device__ __forceinline__ int32_t col(int32_t blk, int j) const { return ci[blk] * NB + j; }
  __device__ __forceinline__ int64_t vidx(int32_t blk, int i, int j) const
  {
    return (int64_t)blk * NB * NB + i * NB + j;
  }
};

// femutils/CsrDoFLinearSystemImpl.cc:126-152
template <int NB>
__global__ void k_row_elimination(RowEntries<NB> m, int32_t n, double* __restrict__ vals,
                                  const uint8_t* __restrict__ elim)
{
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n || elim[r] != kEliminateRow)
    return;
  int32_t b, e;
  int i;
  m.range(r, b, e, i);
  for (int32_t blk = b; blk < e; ++blk)
    for (int j = 0; j < NB; ++j)
      vals[m.vidx(blk, i, j)] = (m.col(blk, j) == r) ? 1.0 : 0.0;
}

// femutils/CsrDoFLinearSystemImpl.cc:88-121, including its 'column_index > 0'
// test (column 0 is never cleared by this step).
template <int NB>
__global__ void k_row_column_elimination(RowEntries<NB> m, int32_t n, double* __restrict__ vals,
                                         const uint8_t* __restrict__ elim)
{
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n)
    return;
  const bool row_elim = elim[r] == kEliminateRow || elim[r] == kEliminateRowColumn;
  int32_t b, e;
  int i;
  m.range(r, b, e, i);
  for (int32_t blk = b; blk < e; ++blk)
    for (int j = 0; j < NB; ++j) {
      const int32_t c = m.col(blk, j);
      if (c > 0) {
        const bool col_elim = elim[c] == kEliminateRow || elim[c] == kEliminateRowColumn;
        if (row_elim || col_elim)
          vals[m.vidx(blk, i, j)] = (c == r) ? 1.0 : 0.0;
      }
    }
}

// femutils/CsrDoFLinearSystemImpl.cc:50-72
template <int NB>
__global__ void k_forced_values(RowEntries<NB> m, int32_t n, double* __restrict__ vals,
                                const uint8_t* __restrict__ forced, const double* __restrict__ forced_value,
                                int32_t* __restrict__ missing)
{
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n || !forced[r])
    return;
  int32_t b, e;
  int i;
  m.range(r, b, e, i);
  for (int32_t blk = b; blk < e; ++blk)
    for (int j = 0; j < NB; ++j)
      if (m.col(blk, j) == r) {
        vals[m.vidx(blk, i, j)] = forced_value[r];
        return;
      }
  atomicOr(missing, 1);
}

// femutils/DoFLinearSystemImplBase.cc:55-88 in gather form: row j collects
// A(i, j) * g_i from every row-column eliminated i != j of its pattern, where
// A(i, j) is the value saved before the matrix transformation.
template <int NB>
__global__ void k_row_column_rhs_lift(RowEntries<NB> m, int32_t n, const double* __restrict__ saved,
                                      const uint8_t* __restrict__ elim, const double* __restrict__ elim_value,
                                      double* __restrict__ rhs)
{
  const int32_t jrow = blockIdx.x * blockDim.x + threadIdx.x;
  if (jrow >= n)
    return;
  int32_t b, e;
  int ij;
  m.range(jrow, b, e, ij);
  double acc = 0.0;
  for (int32_t blk = b; blk < e; ++blk)
    for (int jj = 0; jj < NB; ++jj) {
      const int32_t irow = m.col(blk, jj);
      if (irow == jrow || elim[irow] != kEliminateRowColumn)
        continue;
      // find A(irow, jrow) in row irow
      int32_t b2, e2;
      int ii;
      m.range(irow, b2, e2, ii);
      for (int32_t blk2 = b2; blk2 < e2; ++blk2)
        for (int j2 = 0; j2 < NB; ++j2)
          if (m.col(blk2, j2) == jrow)
            acc += saved[m.vidx(blk2, ii, j2)] * elim_value[irow];
    }
  rhs[jrow] -= acc;
}

// femutils/CsrDoFLinearSystemImpl.cc:157-182
__global__ void k_rhs_elimination(int32_t n, const uint8_t* __restrict__ elim, const double* __restrict__ elim_value,
                                  double* __restrict__ rhs)
{
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n && (elim[r] == kEliminateRow || elim[r] == kEliminateRowColumn))
    rhs[r] = elim_value[r];
}

__global__ void k_any_equal(const uint8_t* __restrict__ a, int32_t n, uint8_t v, int32_t* __restrict__ flag)
{
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n && a[r] == v)
    *flag = 1;
}

// femutils/ArcaneFemFunctionsGpu.cc:54-105 for one node group and component.
__global__ void k_dirichlet(const int32_t* __restrict__ nodes, int32_t count, int nb_dof_per_node, int dof_index,
                            int method, double value, double penalty, uint8_t* __restrict__ forced_info,
                            double* __restrict__ forced_value, uint8_t* __restrict__ elim_info,
                            double* __restrict__ elim_value, double* __restrict__ rhs)
{
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count)
    return;
  const int64_t dof = (int64_t)nodes[k] * nb_dof_per_node + dof_index;
  if (method == AFEM_DIRICHLET_PENALTY) {
    forced_info[dof] = 1;
    forced_value[dof] = penalty;
    rhs[dof] = penalty * value;
  }
  else {
    elim_info[dof] = (method == AFEM_DIRICHLET_ROW_ELIMINATION) ? kEliminateRow : kEliminateRowColumn;
    elim_value[dof] = value;
  }
}

// Closes the row starts of a CSR view (femutils/CsrFormatMatrixView.h:145-187:
// `rows` has nbRow() entries and the last row ends at nbColumn()) into
// nb_dof + 1 offsets, and checks the view on the way: starts in [0, nnz] and
// non-decreasing, a closing entry (nb_given = nb_dof + 1, the scalar expansion
// of femutils/BSRFormat.cc:117-139) equal to nnz, per-row counts (when given)
// equal to the row extents.  bad[0] collects the violations.
__global__ void k_close_rows(const int32_t* __restrict__ rows, int32_t nb_given, const int32_t* __restrict__ counts,
                             int32_t n, int64_t nnz, int32_t* __restrict__ out, int32_t* __restrict__ bad)
{
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r > n)
    return;
  const int64_t b = (r < n) ? rows[r] : nnz;
  if (r == n) {
    out[n] = (int32_t)nnz;
    if (nb_given > n && rows[n] != nnz)
      atomicOr(bad, 1);
    return;
  }
  const int64_t e = (r + 1 < n) ? rows[r + 1] : (int64_t)nnz;
  out[r] = (int32_t)b;
  if (b < 0 || b > e || e > nnz || (r == 0 && b != 0))
    atomicOr(bad, 2);
  if (counts && counts[r] != e - b)
    atomicOr(bad, 4);
}

// Column ids of a CSR view must name DoFs of the family (the solver gathers
// x at them).
__global__ void k_check_columns(const int32_t* __restrict__ cols, int64_t nnz, int32_t n, int32_t* __restrict__ bad)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz && (cols[k] < 0 || cols[k] >= n))
    atomicOr(bad, 8);
}

// One (row, column) entry of the matrix view: add or set; found = 0 when the
// entry is not in the pattern.
template <int NB>
__global__ void k_matrix_entry(RowEntries<NB> m, double* __restrict__ vals, int32_t row, int32_t col, double v,
                               int set, int32_t* __restrict__ found)
{
  if (threadIdx.x != 0 || blockIdx.x != 0)
    return;
  int32_t b, e;
  int i;
  m.range(row, b, e, i);
  for (int32_t blk = b; blk < e; ++blk)
    for (int j = 0; j < NB; ++j)
      if (m.col(blk, j) == col) {
        double& a = vals[m.vidx(blk, i, j)];
        a = set ? v : a + v;
        *found = 1;
        return;
      }
  *found = 0;
}

static bool any_equal(const DBuf<uint8_t>& a, int32_t n, uint8_t v)
{
  DBuf<int32_t> flag(1);
  flag.zero();
  hipLaunchKernelGGL(k_any_equal, dim3(grid_for(n, 256)), dim3(256), 0, stream(), a.p, n, v, flag.p);
  check_launch("any_equal");
  int32_t h = 0;
  flag.download(&h, 1);
  return h != 0;
}

void system_require_matrix(const afem_system* s)
{
  AFEM_REQUIRE(s->has_matrix, AFEM_ERR_STATE, "no matrix: call setCSRValues (afem_system_set_csr_values / set_bsr) first");
}

template <int NB>
static void matrix_transformation(afem_system* s)
{
  RowEntries<NB> m{ s->row_index, s->columns };
  const int32_t n = s->nb_dof;
  const unsigned grid = grid_for(n, 256);
  s->has_row_column = any_equal(s->elim_info, n, kEliminateRowColumn);
  s->dof0_row_column = false;
  if (s->has_row_column) {
    uint8_t info0 = 0;
    AFEM_HIP(hipMemcpyAsync(&info0, s->elim_info.p, 1, hipMemcpyDeviceToHost, stream()));
    AFEM_HIP(hipStreamSynchronize(stream()));
    s->dof0_row_column = n > 0 && info0 == kEliminateRowColumn;
    s->saved_values.reserve((size_t)s->nnz);
    AFEM_HIP(hipMemcpyAsync(s->saved_values.p, s->values, sizeof(double) * s->nnz, hipMemcpyDeviceToDevice, stream()));
  }
  hipLaunchKernelGGL(k_row_elimination<NB>, dim3(grid), dim3(256), 0, stream(), m, n, s->values, s->elim_info.p);
  check_launch("row_elimination");
  if (s->has_row_column) {
    hipLaunchKernelGGL(k_row_column_elimination<NB>, dim3(grid), dim3(256), 0, stream(), m, n, s->values,
                       s->elim_info.p);
    check_launch("row_column_elimination");
  }
  DBuf<int32_t> missing(1);
  missing.zero();
  hipLaunchKernelGGL(k_forced_values<NB>, dim3(grid), dim3(256), 0, stream(), m, n, s->values, s->forced_info.p,
                     s->forced_value.p, missing.p);
  check_launch("forced_values");
  int32_t h = 0;
  missing.download(&h, 1);
  AFEM_REQUIRE(h == 0, AFEM_ERR_NOT_FOUND, "a forced DoF has no diagonal entry in the matrix pattern");
  s->matrix_changed = true;
}

template <int NB>
static void rhs_transformation(afem_system* s)
{
  RowEntries<NB> m{ s->row_index, s->columns };
  const int32_t n = s->nb_dof;
  const unsigned grid = grid_for(n, 256);
  if (s->has_row_column) {
    hipLaunchKernelGGL(k_row_column_rhs_lift<NB>, dim3(grid), dim3(256), 0, stream(), m, n, s->saved_values.p,
                       s->elim_info.p, s->elim_value.p, s->rhs.p);
    check_launch("row_column_rhs_lift");
  }
  hipLaunchKernelGGL(k_rhs_elimination, dim3(grid), dim3(256), 0, stream(), n, s->elim_info.p, s->elim_value.p,
                     s->rhs.p);
  check_launch("rhs_elimination");
}

void system_matrix_transformation(afem_system* s)
{
  system_require_matrix(s);
  if (s->block == 3)
    matrix_transformation<3>(s);
  else
    matrix_transformation<1>(s);
}

void system_rhs_transformation(afem_system* s)
{
  system_require_matrix(s);
  if (s->block == 3)
    rhs_transformation<3>(s);
  else
    rhs_transformation<1>(s);
}

} // namespace afem

using namespace afem;

extern "C" {

AFEM_API int afem_system_create(int32_t nb_dof, afem_system_t* out)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(out, AFEM_ERR_ARGUMENT, "null output handle");
  *out = nullptr;
  AFEM_REQUIRE(nb_dof > 0, AFEM_ERR_ARGUMENT, "nb_dof must be positive");
  auto* s = new afem_system();
  try {
    s->nb_dof = nb_dof;
    s->rhs.alloc(nb_dof);
    s->x.alloc(nb_dof);
    s->forced_value.alloc(nb_dof);
    s->elim_value.alloc(nb_dof);
    s->forced_info.alloc(nb_dof);
    s->elim_info.alloc(nb_dof);
    s->rhs.zero();
    s->x.zero();
    s->forced_value.zero();
    s->elim_value.zero();
    s->forced_info.zero();
    s->elim_info.zero();
  }
  catch (...) {
    delete s;
    throw;
  }
  *out = s;
  AFEM_GUARD_END
}

AFEM_API int afem_system_destroy(afem_system_t sys)
{
  AFEM_GUARD_BEGIN
  if (sys) {
    AFEM_HIP(hipStreamSynchronize(stream()));
    system_release_workspace(sys);
    delete sys;
  }
  AFEM_GUARD_END
}

AFEM_API int afem_system_set_csr_values(afem_system_t sys, int32_t nb_row, int64_t nnz, const int32_t* rows,
                                        const int32_t* rows_nb_column, const int32_t* columns, double* values)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys, AFEM_ERR_ARGUMENT, "null handle");
  const int32_t n = sys->nb_dof;
  AFEM_REQUIRE(nb_row == n || nb_row == n + 1, AFEM_ERR_ARGUMENT,
               "CSR row count differs from the DoF count (nb_dof, or nb_dof + 1 with a closing entry)");
  AFEM_REQUIRE(nnz > 0 && nnz < (1LL << 31), AFEM_ERR_ARGUMENT, "nnz out of range");
  AFEM_REQUIRE(rows && columns && values, AFEM_ERR_ARGUMENT, "null CSR array");
  // closed into a new array first: a rejected view leaves the current one in place
  DBuf<int32_t> closed((size_t)n + 1);
  DBuf<int32_t> bad(1);
  bad.zero();
  hipLaunchKernelGGL(k_close_rows, dim3(grid_for((int64_t)n + 1, 256)), dim3(256), 0, stream(), rows, nb_row,
                     rows_nb_column, n, nnz, closed.p, bad.p);
  check_launch("close_rows");
  hipLaunchKernelGGL(k_check_columns, dim3(grid_for(nnz, 256)), dim3(256), 0, stream(), columns, nnz, n, bad.p);
  check_launch("check_columns");
  int32_t h = 0;
  bad.download(&h, 1);
  AFEM_REQUIRE(!(h & 1), AFEM_ERR_ARGUMENT, "CSR view: the closing row entry differs from the value count");
  AFEM_REQUIRE(!(h & 2), AFEM_ERR_ARGUMENT, "CSR view: row starts are not non-decreasing in [0, nnz] from 0");
  AFEM_REQUIRE(!(h & 4), AFEM_ERR_ARGUMENT, "CSR view: a row's column count differs from its extent");
  AFEM_REQUIRE(!(h & 8), AFEM_ERR_ARGUMENT, "CSR view: a column id is outside the DoF family");
  sys->csr_rows = std::move(closed);
  sys->has_matrix = true;
  sys->nb_row = n;
  sys->nnz = nnz;
  sys->row_index = sys->csr_rows.p;
  sys->columns = columns;
  sys->values = values;
  sys->block = 1;
  sys->bsr = nullptr;
  sys->matrix_changed = true;
  sys->pattern_changed = true; // a caller's arrays may hold a new pattern at the same addresses
  AFEM_GUARD_END
}

AFEM_API int afem_system_set_bsr(afem_system_t sys, afem_bsr_t bsr)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys && bsr, AFEM_ERR_ARGUMENT, "null handle");
  AFEM_REQUIRE(bsr->has_sparsity, AFEM_ERR_STATE, "computeSparsity not called");
  AFEM_REQUIRE((int64_t)bsr->nb_row * bsr->nb_dof == sys->nb_dof, AFEM_ERR_ARGUMENT,
               "matrix size differs from the DoF count");
  AFEM_REQUIRE(bsr->nb_dof != 2 || !bsr->per_block, AFEM_ERR_NOT_IMPLEMENTED,
               "2x2 blocks: use the per-row (CSR) value order (does_linear_system_use_csr)");
  if (sys->bsr != bsr || sys->bsr_gen != bsr->pattern_gen)
    sys->pattern_changed = true;
  sys->has_matrix = true;
  sys->nb_row = bsr->nb_row * bsr->nb_dof;
  sys->nnz = bsr->nb_col * bsr->nb_dof * bsr->nb_dof;
  sys->values = bsr->values.p;
  sys->bsr = bsr;
  sys->bsr_gen = bsr->pattern_gen;
  sys->matrix_changed = true;
  if (bsr->nb_dof == 1) {
    sys->block = 1;
    sys->row_index = bsr->row_index.p;
    sys->columns = bsr->columns.p;
  }
  else if (bsr->per_block) {
    sys->block = bsr->nb_dof;
    sys->row_index = bsr->row_index.p;
    sys->columns = bsr->columns.p;
  }
  else {
    // per-row values are scalar-CSR ordered: hand over the expanded pattern
    const int32_t* r;
    const int32_t* c;
    int32_t nr;
    int rc = afem_bsr_to_csr(bsr, &r, &c, &nr);
    if (rc != AFEM_OK)
      fail(rc, afem_last_error());
    sys->block = 1;
    sys->row_index = r;
    sys->columns = c;
  }
  AFEM_GUARD_END
}

AFEM_API int afem_system_rhs(afem_system_t sys, double** rhs)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys && rhs, AFEM_ERR_ARGUMENT, "null argument");
  *rhs = sys->rhs.p;
  AFEM_GUARD_END
}

AFEM_API int afem_system_solution(afem_system_t sys, double** x)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys && x, AFEM_ERR_ARGUMENT, "null argument");
  *x = sys->x.p;
  AFEM_GUARD_END
}

AFEM_API int afem_system_forced(afem_system_t sys, uint8_t** info, double** value)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys, AFEM_ERR_ARGUMENT, "null handle");
  if (info)
    *info = sys->forced_info.p;
  if (value)
    *value = sys->forced_value.p;
  AFEM_GUARD_END
}

AFEM_API int afem_system_elimination(afem_system_t sys, uint8_t** info, double** value)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys, AFEM_ERR_ARGUMENT, "null handle");
  if (info)
    *info = sys->elim_info.p;
  if (value)
    *value = sys->elim_value.p;
  AFEM_GUARD_END
}

AFEM_API int afem_system_apply_dirichlet(afem_system_t sys, int method, int32_t nb_node, const int32_t* nodes,
                                         int nb_dof_per_node, int dof_index, double value, double penalty)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys, AFEM_ERR_ARGUMENT, "null handle");
  // femutils/ArcaneFemFunctionsGpu.cc:143 "Unknown method to enforce Dirichlet BC"
  AFEM_REQUIRE(method == AFEM_DIRICHLET_PENALTY || method == AFEM_DIRICHLET_ROW_ELIMINATION ||
                 method == AFEM_DIRICHLET_ROW_COLUMN_ELIMINATION,
               AFEM_ERR_ARGUMENT, "Unknown method to enforce Dirichlet BC");
  AFEM_REQUIRE(nb_dof_per_node >= 1 && dof_index >= 0 && dof_index < nb_dof_per_node, AFEM_ERR_ARGUMENT,
               "bad DoF component");
  AFEM_REQUIRE(nb_node >= 0 && (nb_node == 0 || nodes), AFEM_ERR_ARGUMENT, "null node list");
  if (nb_node == 0)
    return AFEM_OK;
  for (int32_t k = 0; k < nb_node; ++k) {
    const int64_t dof = (int64_t)nodes[k] * nb_dof_per_node + dof_index;
    AFEM_REQUIRE(nodes[k] >= 0 && dof < sys->nb_dof, AFEM_ERR_ARGUMENT, "Dirichlet node outside the DoF family");
  }
  DBuf<int32_t> dn((size_t)nb_node);
  dn.upload(nodes, (size_t)nb_node);
  hipLaunchKernelGGL(k_dirichlet, dim3(grid_for(nb_node, 256)), dim3(256), 0, stream(), dn.p, nb_node,
                     nb_dof_per_node, dof_index, method, value, penalty, sys->forced_info.p, sys->forced_value.p,
                     sys->elim_info.p, sys->elim_value.p, sys->rhs.p);
  check_launch("dirichlet");
  AFEM_HIP(hipStreamSynchronize(stream()));
  AFEM_GUARD_END
}

AFEM_API int afem_system_apply_matrix_transformation(afem_system_t sys)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys, AFEM_ERR_ARGUMENT, "null handle");
  system_matrix_transformation(sys);
  AFEM_GUARD_END
}

AFEM_API int afem_system_apply_rhs_transformation(afem_system_t sys)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys, AFEM_ERR_ARGUMENT, "null handle");
  system_rhs_transformation(sys);
  AFEM_GUARD_END
}

AFEM_API int afem_system_matrix_add_value(afem_system_t sys, int32_t row, int32_t column, double value, int set)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys, AFEM_ERR_ARGUMENT, "null handle");
  system_require_matrix(sys);
  AFEM_REQUIRE(row >= 0 && row < sys->nb_dof && column >= 0 && column < sys->nb_dof, AFEM_ERR_ARGUMENT,
               "matrix entry outside the DoF family");
  DBuf<int32_t> found(1);
  if (sys->block == 3)
    hipLaunchKernelGGL(k_matrix_entry<3>, dim3(1), dim3(64), 0, stream(), RowEntries<3>{ sys->row_index, sys->columns },
                       sys->values, row, column, value, set, found.p);
  else
    hipLaunchKernelGGL(k_matrix_entry<1>, dim3(1), dim3(64), 0, stream(), RowEntries<1>{ sys->row_index, sys->columns },
                       sys->values, row, column, value, set, found.p);
  check_launch("matrix_entry");
  int32_t h = 0;
  found.download(&h, 1);
  AFEM_REQUIRE(h == 1, AFEM_ERR_NOT_FOUND, "BSRMatrix(findValueIndex): Value not found");
  sys->matrix_changed = true;
  AFEM_GUARD_END
}

AFEM_API int afem_system_clear_values(afem_system_t sys)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys, AFEM_ERR_ARGUMENT, "null handle");
  sys->forced_info.zero();
  sys->elim_info.zero();
  sys->elim_value.zero();
  sys->has_matrix = false;
  sys->row_index = nullptr;
  sys->columns = nullptr;
  sys->values = nullptr;
  sys->bsr = nullptr;
  sys->matrix_changed = true;
  sys->pattern_changed = true;
  AFEM_GUARD_END
}

} // extern "C"
