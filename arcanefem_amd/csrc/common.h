// common.h -- shared internals of libafem: error handling, the library
// stream, device buffers, per-kernel event timing and the handle structs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "afem.h"

namespace afem {

// Error carried from internal code to the C ABI, where it becomes a status
// code plus the thread-local message of afem_last_error().
struct Error : std::runtime_error
{
  int status;
  Error(int s, const std::string& msg)
  : std::runtime_error(msg)
  , status(s)
  {}
};

[[noreturn]] void fail(int status, const std::string& msg);
void set_last_error(const std::string& msg);

#define AFEM_HIP(call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      ::afem::fail(e_ == hipErrorOutOfMemory ? AFEM_ERR_OUT_OF_MEMORY : AFEM_ERR_HIP,   \
                   std::string(#call) + ": " + hipGetErrorString(e_));                  \
  } while (0)

#define AFEM_REQUIRE(cond, status, msg)                                                 \
  do {                                                                                  \
    if (!(cond))                                                                        \
      ::afem::fail((status), (msg));                                                    \
  } while (0)

// Library runtime: one device, one stream of default priority.
struct Runtime
{
  int device = -1;
  hipStream_t stream = nullptr;
  int num_cu = 0;
  bool profiling = false;
};
Runtime& rt();
void ensure_init();
inline hipStream_t stream()
{
  ensure_init();
  return rt().stream;
}

// Owning device buffer.
template <class T>
struct DBuf
{
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  explicit DBuf(size_t count) { alloc(count); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept
  : p(o.p)
  , n(o.n)
  {
    o.p = nullptr;
    o.n = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept
  {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DBuf() { release(); }
  void alloc(size_t count)
  {
    release();
    if (count) {
      ensure_init();
      AFEM_HIP(hipMalloc(&p, count * sizeof(T)));
    }
    n = count;
  }
  // grow-only reallocation (contents not preserved)
  void reserve(size_t count)
  {
    if (count > n)
      alloc(count);
  }
  void release()
  {
    if (p)
      (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void zero() { zero_n(n); }
  void zero_n(size_t count)
  {
    if (count)
      AFEM_HIP(hipMemsetAsync(p, 0, count * sizeof(T), stream()));
  }
  // synchronous: the host array may be a temporary
  void upload(const T* host, size_t count)
  {
    if (count) {
      AFEM_HIP(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, stream()));
      AFEM_HIP(hipStreamSynchronize(stream()));
    }
  }
  void download(T* host, size_t count) const
  {
    if (count) {
      AFEM_HIP(hipMemcpyAsync(host, p, count * sizeof(T), hipMemcpyDeviceToHost, stream()));
      AFEM_HIP(hipStreamSynchronize(stream()));
    }
  }
};

// Event pair recorded around one launch when profiling is on.
struct ProfScope
{
  const char* name;
  hipEvent_t start = nullptr;
  explicit ProfScope(const char* kernel_name);
  ~ProfScope();
};

// Checked grid size for 1-D launches.
inline unsigned grid_for(int64_t work, int block)
{
  int64_t g = (work + block - 1) / block;
  if (g < 1)
    g = 1;
  if (g > 0x7fffffffLL)
    fail(AFEM_ERR_ARGUMENT, "launch too large");
  return (unsigned)g;
}

inline void check_launch(const char* what)
{
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    fail(AFEM_ERR_HIP, std::string("launch of ") + what + ": " + hipGetErrorString(e));
}

// Exclusive scan of n int32 counts into out[0..n] (out[n] = total).  Returns
// the total (synchronous read-back).  The total must fit in int32.
int64_t exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n);

} // namespace afem

// C-ABI bodies: exceptions become status codes + afem_last_error().
#define AFEM_GUARD_BEGIN try {
#define AFEM_GUARD_END                                                                    \
  }                                                                                       \
  catch (const afem::Error& e)                                                            \
  {                                                                                       \
    afem::set_last_error(e.what());                                                       \
    return e.status;                                                                      \
  }                                                                                       \
  catch (const std::bad_alloc&)                                                           \
  {                                                                                       \
    afem::set_last_error("host allocation failed");                                      \
    return AFEM_ERR_OUT_OF_MEMORY;                                                        \
  }                                                                                       \
  catch (const std::exception& e)                                                         \
  {                                                                                       \
    afem::set_last_error(e.what());                                                       \
    return AFEM_ERR_HIP;                                                                  \
  }                                                                                       \
  return AFEM_OK;

// ---- handle structs --------------------------------------------------------

struct afem_mesh
{
  int dim = 0;
  int npc = 0;
  int stride = 4;                    // cell stride: 4 (tri3, tet4, quad4) or 8 (hexa8)
  int slot_bits = 2;                 // log2(stride)
  bool simplex = true;               // tri3 / tet4; false for Quad4 / Hexa8
  int32_t nb_node = 0;
  int32_t nb_cell = 0;
  int32_t max_degree = 0;            // largest node->cell list
  afem::DBuf<double> coords;         // nb_node * 4 (x, y, z, 0)
  afem::DBuf<int32_t> cells;         // nb_cell * stride
  afem::DBuf<int32_t> nc_offsets;    // nb_node + 1
  afem::DBuf<int32_t> nc_items;      // cell*stride + slot, sorted by cell per node
};

struct afem_bsr
{
  afem_mesh* mesh = nullptr;
  int nb_dof = 1;
  bool per_block = true;             // value order
  bool use_csr = false;
  bool atomic_free = false;
  bool has_sparsity = false;
  uint64_t pattern_gen = 0;          // counts computeSparsity calls (linear systems key derived data on it)
  int32_t nb_row = 0;
  int64_t nb_col = 0;                // stored blocks
  int32_t max_row = 0;
  afem::DBuf<int32_t> row_index;     // nb_row + 1
  afem::DBuf<int32_t> columns;       // nb_col
  afem::DBuf<int32_t> nnz_per_row;   // nb_row
  afem::DBuf<double> values;         // nb_col * nb_dof^2
  afem::DBuf<int32_t> csr_row;       // scalar CSR (toCsr), built on demand
  afem::DBuf<int32_t> csr_col;
  // assembly plan: per (node, cell) pair the cell, the slot and the row
  // positions of the other nodes; per row the diagonal position
  afem::DBuf<uint64_t> plan;
  unsigned plans_tried = 0;          // derived plans built (or found unbuildable) for this pattern
  afem::DBuf<uint64_t> lplan;        // block-local plan (node-table indices + positions): built, sliced, released
  afem::DBuf<uint64_t> lplan_sl;     // the same words sliced per block of rows (coalesced reads)
  afem::DBuf<int32_t> lcell_sl;      // cell id of each sliced word
  afem::DBuf<int64_t> slice;         // per block of rows: first sliced word
  // brick order (Morton order of the nodes, blocks of 256 rows)
  afem::DBuf<int32_t> bperm;         // rows in brick order
  afem::DBuf<int32_t> bval_off;      // LDS offset of each row within its brick
  afem::DBuf<int32_t> btotal;        // values per brick
  afem::DBuf<int32_t> btable, bcount; // per brick: sorted referenced nodes (fixed stride), count
  afem::DBuf<int64_t> bslice;        // per brick: first sliced plan word
  afem::DBuf<uint64_t> bplan;        // sliced local-plan words of the bricks
  afem::DBuf<int32_t> bcell;         // cell id of each brick plan word
  afem::DBuf<int4> brec;             // per brick slot: row, value offset, len | diag << 8 | cells << 16, table index
  afem::DBuf<double> bxyz;           // per brick: its node table's coordinates, xy pairs then z
  afem::DBuf<int64_t> bxyz_off;      // per brick: offset of its coordinates in bxyz (doubles, even)
  int32_t brick_val_max = 0, brick_nodes_max = 0;
  // cell-once brick plan: per brick its cell instances (two words each), their
  // cell ids, the table index -> brick slot map
  afem::DBuf<int64_t> cinst_off;     // nb_brick + 1
  afem::DBuf<uint64_t> cwords;
  afem::DBuf<int32_t> ccell;
  afem::DBuf<uint8_t> cmap;          // nb_brick * 2048
  int32_t cell_inst_max = 0;
  // per nonzero block plan (elasticity): workgroups of whole rows with at most
  // 256 blocks, per block the cells sharing its edge (sliced per workgroup)
  afem::DBuf<int32_t> ewg_row;       // first row of each workgroup (nb_wg + 1)
  afem::DBuf<int64_t> eslice;        // first cell word of each workgroup
  afem::DBuf<int32_t> ecell;         // cell j of block t of workgroup w at eslice[w] + 256 j + t
  afem::DBuf<uint8_t> ecnt;          // cells per block
  afem::DBuf<int32_t> einst_off;     // first instance of each workgroup in einst (nb_wg + 1)
  afem::DBuf<int32_t> einst;         // the distinct cells of each workgroup's rows, sorted
  afem::DBuf<uint16_t> egeo;         // instance | slot of p << 10 | slot of q << 12, eight per block together
  afem::DBuf<int64_t> eslice8;       // per workgroup: first egeo word (widths rounded up to multiples of 8)
  afem::DBuf<int4> einst4;           // the nodes of each instance
  int32_t e_inst_max = 0;            // largest instance count of a workgroup (0: no geometry plan)
  // Quad4 / Hexa8: per cell the upper triangle of the element matrix and the
  // source vector, cell-minor (entry k of cell c at k * nb_cell + c)
  afem::DBuf<double> q1_elem;
  afem::DBuf<uint64_t> q1_plan;      // per (node, cell) pair: the row positions of the cell's nodes, 8 bits each
  afem::DBuf<uint64_t> q1_plan_sl;   // the same words sliced per gather workgroup: pair j of row t at (w*width + j)*BS + t
  afem::DBuf<int32_t> q1_item_sl;    // the pairs' node->cell items in the same layout
  int32_t q1_sl_width = 0, q1_sl_rows = 0; // slice width (largest cell count of a node) and rows per workgroup
  int32_t e_nwg = 0;
  afem::DBuf<int32_t> block_nodes;   // per block of rows: sorted referenced nodes (fixed stride)
  afem::DBuf<int32_t> block_count;   // per block: number of referenced nodes
  int32_t block_nodes_max = 0;       // largest node table
  afem::DBuf<int32_t> diag_pos;
  int variant = -1;                  // assembly kernel variant (-1: default)
};

struct afem_solver_ws;

struct afem_system
{
  int32_t nb_dof = 0;
  afem::DBuf<double> rhs, x, forced_value, elim_value;
  afem::DBuf<uint8_t> forced_info, elim_info;
  // matrix view (non-owning, except the closed copy of a CSR view's row starts)
  bool has_matrix = false;
  int32_t nb_row = 0;
  int64_t nnz = 0;
  const int32_t* row_index = nullptr; // nb_dof + 1 entries, last = nnz
  afem::DBuf<int32_t> csr_rows;       // row_index of a view given by afem_system_set_csr_values
  const int32_t* columns = nullptr;
  double* values = nullptr;
  int block = 1;                     // 3 when the view is a per-block BSR
  afem_bsr* bsr = nullptr;
  uint64_t bsr_gen = 0;              // pattern_gen of bsr when it was set
  bool matrix_changed = true;        // invalidates derived data (FP32 copy, diagonal)
  bool pattern_changed = true;       // invalidates pattern-derived data (SELL layouts)
  bool has_row_column = false;       // row-column eliminations present
  bool dof0_row_column = false;      // DoF 0 is row-column eliminated (column 0 kept: non-symmetric)
  afem::DBuf<double> saved_values;   // values before the row-column elimination
  afem_solver_ws* ws = nullptr;
};

namespace afem {

// Quad4 / Hexa8 meshes (q1.hip): Laplace values (and the fused constant-source
// RHS when rhs is given) and the constant-source RHS alone.
void q1_assemble_laplace(afem_bsr* b, const double* cell_lambda, double f, double* rhs);
void q1_assemble_heat(afem_bsr* b, const double* cell_lambda, double dt);
void q1_assemble_elasticity(afem_bsr* b, double lam, double mu);
void q1_rhs_mass_field(const afem_mesh& m, const double* field, double scale, int nb_dof, int dof_index, double* rhs);
// manufactured-solution source on tri3 / Quad4 / Hexa8 meshes (q1.hip)
void rhs_manufactured(const afem_mesh& m, int which, int nb_dof, int dof_index, double* rhs);
// per-cell squared L2 error of a nodal solution against a manufactured solution
void manufactured_cell_error(const afem_mesh& m, int which, const double* u, double* err2);
void q1_rhs_constant_source(const afem_mesh& m, double f, const double* cell_scale, int nb_dof, int dof_index,
                            double* rhs);

} // namespace afem
