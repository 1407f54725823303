// q1.hip -- Quad4 / Hexa8 (Q1) meshes: node-wise gather assembly of the
// Laplace bilinear form (2x2 / 2x2x2 Gauss points, an isoparametric Jacobian
// per point) and of the constant-source right-hand side.
//
// Reference behaviour: modules/poisson/ElementMatrixHexQuad.h:33-64 (Quad4),
// :159-195 (Hexa8); femutils/ShapeFunctions.h:104-129, :291-350;
// femutils/ArcaneFemFunctionsGpu.h:296-350, :553-587 (gradients and det J);
// femutils/ArcaneFemFunctionsGpu.cc:393-470, :875-960 (source term).  The
// reference scatters each element matrix with atomics (assembleBilinearAtomic)
// or recomputes it per node with a column search (assembleBilinearAtomicFree).
// Here one thread owns one matrix row, so no two threads ever write the same
// value: by default the element matrices are built once per cell into a
// workspace and each row thread gathers the rows of its node's cells; the
// variant 0 form rebuilds the element row of the node's slot per (node, cell)
// pair in registers and needs no workspace.
#include "common.h"
#include "device_utils.h"
#include "q1_point.h"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace afem {

namespace {

__device__ __forceinline__ int q1_find_col(const int32_t* __restrict__ cols, int len, int32_t v)
{
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cols[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// One thread per node i.  MAT: zero row i, then for each cell of i add
// lambda_c * sum_g (grad N_s . grad N_b) det J_g into the entries (i, node b);
// *err = 1 when a column is missing.  RHS: rhs[i * nb_dof + dof_index] =
// sum over the cells of f * scale_c * sum_g N_s det J_g.
template <int DIM, bool MAT>
__global__ void __launch_bounds__(256)
k_assemble_q1(const double* __restrict__ coords, const int32_t* __restrict__ cells, int slot_bits,
              const int32_t* __restrict__ nc_off, const int32_t* __restrict__ nc_items, int32_t nb_node,
              const int32_t* __restrict__ row_index, const int32_t* __restrict__ columns,
              const double* __restrict__ cell_lambda, double* __restrict__ values, double f,
              const double* __restrict__ cell_scale, int nb_dof, int dof_index, double* __restrict__ rhs,
              int32_t* __restrict__ err)
{
  constexpr int N = Q1Point<DIM>::N;
  constexpr int NG = DIM == 2 ? 4 : 8;
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb_node)
    return;
  int32_t a0 = 0, len = 0;
  const int32_t* rc = nullptr;
  if (MAT) {
    a0 = row_index[i];
    len = row_index[i + 1] - a0;
    rc = columns + a0;
    for (int k = 0; k < len; ++k)
      values[a0 + k] = 0.0;
  }
  double src = 0.0;
  for (int32_t e = nc_off[i]; e < nc_off[i + 1]; ++e) {
    const int32_t it = nc_items[e];
    const int s = it & ((1 << slot_bits) - 1);
    const int32_t c = it >> slot_bits;
    const int32_t* cn = cells + (it - s);
    int32_t node[N];
    double x[N][3];
#pragma unroll
    for (int a = 0; a < N; ++a) {
      node[a] = cn[a];
      const double2 xy = *reinterpret_cast<const double2*>(coords + 4 * (int64_t)node[a]);
      x[a][0] = xy.x;
      x[a][1] = xy.y;
      x[a][2] = coords[4 * (int64_t)node[a] + 2];
    }
    double row[N];
#pragma unroll
    for (int b = 0; b < N; ++b)
      row[b] = 0.0;
    double ns = 0.0;
    for (int g = 0; g < NG; ++g) {
      Q1Point<DIM> q;
      q.eval(x, g);
      double sx = 0.0, sy = 0.0, sz = 0.0, sn = 0.0; // the values of slot s, selected without indexing
#pragma unroll
      for (int a = 0; a < N; ++a)
        if (a == s) {
          sx = q.gx[a];
          sy = q.gy[a];
          sz = q.gz[a];
          sn = q.Nv[a];
        }
      if (MAT) {
#pragma unroll
        for (int b = 0; b < N; ++b)
          row[b] += (sx * q.gx[b]) * q.det + (sy * q.gy[b]) * q.det + (sz * q.gz[b]) * q.det;
      }
      ns += sn * q.det;
    }
    if (rhs)
      src += f * (cell_scale ? cell_scale[c] : 1.0) * ns;
    if (MAT) {
      const double lam = cell_lambda ? cell_lambda[c] : 1.0;
#pragma unroll
      for (int b = 0; b < N; ++b) {
        const int pos = q1_find_col(rc, len, node[b]);
        if (pos < len && rc[pos] == node[b])
          values[a0 + pos] += row[b] * lam;
        else
          atomicOr(err, 1);
      }
    }
  }
  if (rhs)
    rhs[(int64_t)i * nb_dof + dof_index] = src;
}

// Element-once form.  Plan (once per pattern): per (node, cell) pair the
// row positions of the cell's nodes.  Pass 1, one thread per cell: the element matrix
// sum_g (grad N_a . grad N_b) det J_g and the source vector sum_g N_a det J_g,
// stored per cell as N rows of N values then the N source values (row s of
// cell c is one contiguous N * 8-byte read).  Pass 2, one thread per row,
// 256 / 128 / 64 rows per workgroup: the workgroup's rows are accumulated in
// LDS in CSR order (each thread owns its row's slots, so no atomics), for each
// cell of the node (slot s) lambda_c times the cell's row s, then written to
// HBM once, coalesced; the RHS receives f times entry s of the source vector.
// Each element matrix is built once instead of once per node (4x / 8x fewer
// Gauss-point evaluations).
// HEAT: lambda_c K + M / dt with M_ab = sum_g N_a N_b det J_g
// (modules/heat/ElementMatrixHexQuad.h:33-130); the gather then applies no lambda.
// The records (N * N + N doubles per cell) leave through LDS in chunks of CH
// values per cell, so that a wave's stores cover contiguous runs instead of
// one double every record length.
constexpr int kQ1ElemBlock = 128;

template <int DIM>
struct Q1Chunk
{
  static constexpr int N = DIM == 2 ? 4 : 8;
  static constexpr int NE = N * N + N;
  static constexpr int CH = DIM == 2 ? NE : (NE + 1) / 2; // values per cell per chunk
  static constexpr int PAD = CH + 1;                      // LDS stride per cell (bank spread)
};

template <int DIM, bool HEAT>
__global__ void __launch_bounds__(kQ1ElemBlock)
k_q1_elements(const double* __restrict__ coords, const int32_t* __restrict__ cells, int32_t nb_cell,
              const double* __restrict__ cell_lambda, double inv_dt, double* __restrict__ elem)
{
  constexpr int N = Q1Point<DIM>::N;
  constexpr int NG = DIM == 2 ? 4 : 8;
  constexpr int NT = N * (N + 1) / 2;
  constexpr int NE = Q1Chunk<DIM>::NE, CH = Q1Chunk<DIM>::CH, PAD = Q1Chunk<DIM>::PAD;
  __shared__ double s_rec[kQ1ElemBlock * PAD];
  const int t = threadIdx.x;
  const int32_t c0 = blockIdx.x * kQ1ElemBlock, c = c0 + t;
  double K[NT], src[N]; // upper triangle, row-major
#pragma unroll
  for (int k = 0; k < NT; ++k)
    K[k] = 0.0;
#pragma unroll
  for (int a = 0; a < N; ++a)
    src[a] = 0.0;
  if (c < nb_cell) {
    constexpr int stride = N > 4 ? N : 4;
    double x[N][3];
#pragma unroll
    for (int a = 0; a < N; ++a) {
      const int32_t n = cells[(int64_t)stride * c + a];
      const double2 xy = *reinterpret_cast<const double2*>(coords + 4 * (int64_t)n);
      x[a][0] = xy.x;
      x[a][1] = xy.y;
      x[a][2] = coords[4 * (int64_t)n + 2];
    }
    const double lam = HEAT && cell_lambda ? cell_lambda[c] : 1.0;
    for (int g = 0; g < NG; ++g) {
      Q1Point<DIM> q;
      q.eval(x, g);
      int k = 0;
#pragma unroll
      for (int a = 0; a < N; ++a) {
#pragma unroll
        for (int b = a; b < N; ++b, ++k) {
          if (HEAT)
            K[k] += (q.gx[a] * q.gx[b]) * q.det * lam + (q.gy[a] * q.gy[b]) * q.det * lam +
                    (q.gz[a] * q.gz[b]) * q.det * lam + (q.Nv[a] * q.Nv[b]) * q.det * inv_dt;
          else
            K[k] += (q.gx[a] * q.gx[b]) * q.det + (q.gy[a] * q.gy[b]) * q.det + (q.gz[a] * q.gz[b]) * q.det;
        }
        src[a] += q.Nv[a] * q.det;
      }
    }
  }
  const int32_t ncell = min(kQ1ElemBlock, nb_cell - c0);
  double* out = elem + (int64_t)NE * c0;
  double* rec = s_rec + t * PAD;
#pragma unroll
  for (int base = 0; base < NE; base += CH) {
    // record entry v: v < N * N is K(a, b) with a = v / N, b = v % N; then src
#pragma unroll
    for (int v = base; v < base + CH && v < NE; ++v) {
      double val;
      if (v < N * N) {
        const int a = v / N, b = v % N, lo = b < a ? b : a, hi = b < a ? a : b;
        val = K[lo * N - lo * (lo - 1) / 2 + (hi - lo)];
      }
      else {
        val = src[v - N * N];
      }
      rec[v - base] = val;
    }
    __syncthreads();
    const int chn = (base + CH <= NE ? CH : NE - base);
    if ((chn & 1) == 0 && (NE & 1) == 0 && (base & 1) == 0) {
      // pairs of values as 16-byte stores (records and chunks start even)
      const int half = chn / 2;
      for (int j = t; j < ncell * half; j += kQ1ElemBlock) {
        const int cell = j / half, k = 2 * (j - cell * half);
        *reinterpret_cast<double2*>(out + (int64_t)cell * NE + base + k) =
            make_double2(s_rec[cell * PAD + k], s_rec[cell * PAD + k + 1]);
      }
    }
    else {
      for (int j = t; j < ncell * chn; j += kQ1ElemBlock) {
        const int cell = j / chn, k = j - cell * chn;
        out[(int64_t)cell * NE + base + k] = s_rec[cell * PAD + k];
      }
    }
    __syncthreads();
  }
}

// Per (node, cell) pair e (node->cell order): the positions in the node's
// row of the cell's N nodes, 8 bits each (rows of at most 255 entries).
// *err = 1 when a column is missing from the pattern.
template <int DIM>
__global__ void __launch_bounds__(256)
k_q1_plan(const int32_t* __restrict__ cells, int slot_bits, const int32_t* __restrict__ nc_off,
          const int32_t* __restrict__ nc_items, int32_t nb_node, const int32_t* __restrict__ row_index,
          const int32_t* __restrict__ columns, uint64_t* __restrict__ plan, int32_t* __restrict__ err)
{
  constexpr int N = Q1Point<DIM>::N;
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb_node)
    return;
  const int32_t a0 = row_index[i], len = row_index[i + 1] - a0;
  const int32_t* rc = columns + a0;
  for (int32_t e = nc_off[i]; e < nc_off[i + 1]; ++e) {
    const int32_t it = nc_items[e];
    const int32_t* cn = cells + (it - (it & ((1 << slot_bits) - 1)));
    uint64_t w = 0;
#pragma unroll
    for (int b = 0; b < N; ++b) {
      const int pos = q1_find_col(rc, len, cn[b]);
      if (!(pos < len && rc[pos] == cn[b]))
        atomicOr(err, 1);
      w |= (uint64_t)(pos & 0xff) << (8 * b);
    }
    plan[e] = w;
  }
}

// Sliced copy of the plan words and node->cell items for the gather: pair j
// of row t of workgroup w at (w * width + j) * BS + t (rows with fewer cells
// leave their tail unused).
__global__ void __launch_bounds__(256)
k_q1_slice(const int32_t* __restrict__ nc_off, const int32_t* __restrict__ nc_items,
           const uint64_t* __restrict__ plan, int32_t nb_node, int width, int rows,
           uint64_t* __restrict__ plan_sl, int32_t* __restrict__ item_sl)
{
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb_node)
    return;
  const int32_t w = i / rows, t = i - w * rows;
  const int64_t base = (int64_t)w * width * rows + t;
  for (int32_t e = nc_off[i]; e < nc_off[i + 1]; ++e) {
    const int64_t at = base + (int64_t)(e - nc_off[i]) * rows;
    plan_sl[at] = plan[e];
    item_sl[at] = nc_items[e];
  }
}

// a += v on an LDS double owned by the calling thread alone (ds_add_f64: the
// LDS unit does the read-modify-write; one writer keeps program order)
__device__ __forceinline__ void q1_lds_add(double* a, double v)
{
  __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Gather from the sliced plan: the (item, word) pairs of four cells are
// loaded together, the cells' element rows (four 16-byte loads each) added
// into the row's LDS slots with ds_add_f64.
template <int DIM, int BS>
__global__ void __launch_bounds__(BS)
k_q1_gather_sl(int slot_bits, const int32_t* __restrict__ nc_off, const int32_t* __restrict__ item_sl,
               const uint64_t* __restrict__ plan_sl, int width, int32_t nb_node, const int32_t* __restrict__ row_index,
               const double* __restrict__ elem, const double* __restrict__ cell_lambda, double* __restrict__ values,
               double f, double* __restrict__ rhs)
{
  constexpr int N = Q1Point<DIM>::N;
  constexpr int NE = N * N + N;
  constexpr int U = 4;
  extern __shared__ double s_val[];
  const int t = threadIdx.x;
  const int32_t r0 = blockIdx.x * BS, i = r0 + t;
  const int32_t rend = min(r0 + BS, nb_node);
  const int32_t base = row_index[r0], total = row_index[rend] - base;
  for (int j = t; j < total; j += BS)
    s_val[j] = 0.0;
  __syncthreads();
  if (i < nb_node) {
    double* row = s_val + (row_index[i] - base);
    const int32_t ncell = nc_off[i + 1] - nc_off[i];
    const int64_t sb = (int64_t)blockIdx.x * width * BS + t;
    double src = 0.0;
    for (int32_t e = 0; e < ncell; e += U) {
      int32_t it[U];
      uint64_t w[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t at = sb + (int64_t)min(e + u, ncell - 1) * BS;
        it[u] = item_sl[at];
        w[u] = plan_sl[at];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (e + u >= ncell)
          break;
        const int s = it[u] & ((1 << slot_bits) - 1);
        const int32_t c = it[u] >> slot_bits;
        const double* ke = elem + (int64_t)NE * c;
        const double lam = cell_lambda ? cell_lambda[c] : 1.0;
#pragma unroll
        for (int b = 0; b < N; b += 2) {
          const double2 p = *reinterpret_cast<const double2*>(ke + s * N + b);
          q1_lds_add(row + ((w[u] >> (8 * b)) & 0xff), p.x * lam);
          q1_lds_add(row + ((w[u] >> (8 * b + 8)) & 0xff), p.y * lam);
        }
        if (rhs)
          src += f * ke[N * N + s];
      }
    }
    if (rhs)
      rhs[i] = src;
  }
  __syncthreads();
  for (int j = t; j < total; j += BS)
    values[base + j] = s_val[j];
}

// rhs[i*nb_dof+dof_index] += (M (scale * field))_i, gathered per node over
// its cells: sum_g N_s (sum_a N_a scale f_a) det J_g
// (integrateNodalFieldToRhsQuad4 / Hexa8, femutils/ArcaneFemFunctions.h:
// 2609-2660, :1680-1740).
template <int DIM>
__global__ void __launch_bounds__(256)
k_q1_mass_field(const double* __restrict__ coords, const int32_t* __restrict__ cells, int slot_bits,
                const int32_t* __restrict__ nc_off, const int32_t* __restrict__ nc_items, int32_t nb_node,
                const double* __restrict__ field, double scale, int nb_dof, int dof_index, double* __restrict__ rhs)
{
  constexpr int N = Q1Point<DIM>::N;
  constexpr int NG = DIM == 2 ? 4 : 8;
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb_node)
    return;
  double acc = 0.0;
  for (int32_t e = nc_off[i]; e < nc_off[i + 1]; ++e) {
    const int32_t it = nc_items[e];
    const int s = it & ((1 << slot_bits) - 1);
    const int32_t* cn = cells + (it - s);
    double x[N][3], fv[N];
#pragma unroll
    for (int a = 0; a < N; ++a) {
      const int32_t n = cn[a];
      const double2 xy = *reinterpret_cast<const double2*>(coords + 4 * (int64_t)n);
      x[a][0] = xy.x;
      x[a][1] = xy.y;
      x[a][2] = coords[4 * (int64_t)n + 2];
      fv[a] = field[n] * scale;
    }
    double part = 0.0;
    for (int g = 0; g < NG; ++g) {
      Q1Point<DIM> q;
      q.eval(x, g);
      double fg = 0.0, ns = 0.0;
#pragma unroll
      for (int a = 0; a < N; ++a) {
        fg += q.Nv[a] * fv[a];
        if (a == s)
          ns = q.Nv[a];
      }
      part += fg * ns * q.det;
    }
    acc += part;
  }
  rhs[(int64_t)i * nb_dof + dof_index] += acc;
}

// Q1 linear elasticity (modules/elasticity/ElementMatrixHexQuad.h:50-70,
// :137-187), one thread per node i and its DIM rows.  For each cell of the
// node (slot s), the Gauss-point sum of the DIM x DIM blocks
//   B_ab[c][e] = lambda g_a[c] g_b[e] + mu g_a[e] g_b[c] + mu (g_a . g_b) delta_ce
// (a = s, g the physical gradients) times det J, added into the node's rows
// (per-block or per-row value order).  *err = 1 when a column is missing.
template <int DIM>
__global__ void __launch_bounds__(256)
k_q1_elasticity(const double* __restrict__ coords, const int32_t* __restrict__ cells, int slot_bits,
                const int32_t* __restrict__ nc_off, const int32_t* __restrict__ nc_items, int32_t nb_node,
                const int32_t* __restrict__ row_index, const int32_t* __restrict__ columns, int per_block, double lam,
                double mu, double* __restrict__ values, int32_t* __restrict__ err)
{
  constexpr int N = Q1Point<DIM>::N;
  constexpr int NG = DIM == 2 ? 4 : 8;
  constexpr int B = DIM;
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb_node)
    return;
  const int32_t a0 = row_index[i], len = row_index[i + 1] - a0;
  const int32_t* rc = columns + a0;
  double* rv = values + (int64_t)a0 * B * B;
  for (int k = 0; k < len * B * B; ++k)
    rv[k] = 0.0;
  for (int32_t e = nc_off[i]; e < nc_off[i + 1]; ++e) {
    const int32_t it = nc_items[e];
    const int s = it & ((1 << slot_bits) - 1);
    const int32_t* cn = cells + (it - s);
    int32_t node[N];
    double x[N][3];
#pragma unroll
    for (int a = 0; a < N; ++a) {
      node[a] = cn[a];
      const double2 xy = *reinterpret_cast<const double2*>(coords + 4 * (int64_t)node[a]);
      x[a][0] = xy.x;
      x[a][1] = xy.y;
      x[a][2] = coords[4 * (int64_t)node[a] + 2];
    }
    double blk[N][B][B];
#pragma unroll
    for (int b = 0; b < N; ++b)
#pragma unroll
      for (int c = 0; c < B; ++c)
#pragma unroll
        for (int d = 0; d < B; ++d)
          blk[b][c][d] = 0.0;
    for (int g = 0; g < NG; ++g) {
      Q1Point<DIM> q;
      q.eval(x, g);
      double gs[3] = { 0.0, 0.0, 0.0 };
#pragma unroll
      for (int a = 0; a < N; ++a)
        if (a == s) {
          gs[0] = q.gx[a];
          gs[1] = q.gy[a];
          gs[2] = q.gz[a];
        }
#pragma unroll
      for (int b = 0; b < N; ++b) {
        const double gb[3] = { q.gx[b], q.gy[b], q.gz[b] };
        double dotv = 0.0;
#pragma unroll
        for (int c = 0; c < B; ++c)
          dotv += gs[c] * gb[c];
#pragma unroll
        for (int c = 0; c < B; ++c)
#pragma unroll
          for (int d = 0; d < B; ++d)
            blk[b][c][d] += (lam * gs[c] * gb[d] + mu * gs[d] * gb[c] + (c == d ? mu * dotv : 0.0)) * q.det;
      }
    }
#pragma unroll
    for (int b = 0; b < N; ++b) {
      const int pos = q1_find_col(rc, len, node[b]);
      if (!(pos < len && rc[pos] == node[b])) {
        atomicOr(err, 1);
        continue;
      }
#pragma unroll
      for (int c = 0; c < B; ++c)
#pragma unroll
        for (int d = 0; d < B; ++d) {
          if (per_block)
            rv[(int64_t)pos * B * B + c * B + d] += blk[b][c][d];
          else // scalar row B i + c holds len * B entries
            rv[(int64_t)c * len * B + pos * B + d] += blk[b][c][d];
        }
    }
  }
}

template <bool MAT>
void launch_q1(const afem_mesh& m, const int32_t* row_index, const int32_t* columns, const double* cell_lambda,
               double* values, double f, const double* cell_scale, int nb_dof, int dof_index, double* rhs,
               int32_t* err)
{
  const dim3 grid(grid_for(m.nb_node, 256)), block(256);
  if (m.dim == 2)
    hipLaunchKernelGGL((k_assemble_q1<2, MAT>), grid, block, 0, stream(), m.coords.p, m.cells.p, m.slot_bits,
                       m.nc_offsets.p, m.nc_items.p, m.nb_node, row_index, columns, cell_lambda, values, f,
                       cell_scale, nb_dof, dof_index, rhs, err);
  else
    hipLaunchKernelGGL((k_assemble_q1<3, MAT>), grid, block, 0, stream(), m.coords.p, m.cells.p, m.slot_bits,
                       m.nc_offsets.p, m.nc_items.p, m.nb_node, row_index, columns, cell_lambda, values, f,
                       cell_scale, nb_dof, dof_index, rhs, err);
  check_launch("assemble_q1");
}

} // namespace

void q1_assemble(afem_bsr* b, const double* cell_lambda, double f, double* rhs, double inv_dt)
{
  const afem_mesh& m = *b->mesh;
  const bool heat = inv_dt > 0.0;
  DBuf<int32_t> err(1);
  err.zero();
  ProfScope prof(heat ? (m.dim == 2 ? "assemble_heat_quad4" : "assemble_heat_hexa8")
                      : (m.dim == 2 ? "assemble_laplace_quad4" : "assemble_laplace_hexa8"));
  // rows longer than 128 entries (a workgroup's values beyond 64 KB of LDS) or
  // 255 (8-bit plan positions) take the per-pair form
  const bool gather_fits = b->max_row <= 128;
  AFEM_REQUIRE(gather_fits || !heat, AFEM_ERR_CAPACITY, "Q1 heat rows longer than 128 entries");
  if ((b->variant == 0 || !gather_fits) && !heat) { // element rows rebuilt per (node, cell) pair, no workspace
    launch_q1<true>(m, b->row_index.p, b->columns.p, cell_lambda, b->values.p, f, nullptr, 1, 0, rhs, err.p);
  }
  else {
    const int n = m.npc;
    const size_t need = (size_t)(n * n + n) * m.nb_cell;
    if (b->q1_elem.n != need)
      b->q1_elem.alloc(need);
    const dim3 cgrid(grid_for(m.nb_cell, kQ1ElemBlock)), block(256), eblock(kQ1ElemBlock);
    if (m.dim == 2 && heat)
      hipLaunchKernelGGL((k_q1_elements<2, true>), cgrid, eblock, 0, stream(), m.coords.p, m.cells.p, m.nb_cell,
                         cell_lambda, inv_dt, b->q1_elem.p);
    else if (m.dim == 2)
      hipLaunchKernelGGL((k_q1_elements<2, false>), cgrid, eblock, 0, stream(), m.coords.p, m.cells.p, m.nb_cell,
                         nullptr, 0.0, b->q1_elem.p);
    else if (heat)
      hipLaunchKernelGGL((k_q1_elements<3, true>), cgrid, eblock, 0, stream(), m.coords.p, m.cells.p, m.nb_cell,
                         cell_lambda, inv_dt, b->q1_elem.p);
    else
      hipLaunchKernelGGL((k_q1_elements<3, false>), cgrid, eblock, 0, stream(), m.coords.p, m.cells.p, m.nb_cell,
                         nullptr, 0.0, b->q1_elem.p);
    const double* gather_lambda = heat ? nullptr : cell_lambda; // the heat element carries lambda already
    check_launch("q1_elements");
    if (b->q1_plan.n != m.nc_items.n) { // once per pattern
      AFEM_REQUIRE(b->max_row <= 255, AFEM_ERR_CAPACITY, "Q1 matrix rows longer than 255 entries");
      b->q1_plan.alloc(m.nc_items.n);
      b->q1_plan_sl.release(); // the sliced copy follows the plan
      b->q1_item_sl.release();
      const dim3 pgrid(grid_for(m.nb_node, 256));
      if (m.dim == 2)
        hipLaunchKernelGGL(k_q1_plan<2>, pgrid, block, 0, stream(), m.cells.p, m.slot_bits, m.nc_offsets.p,
                           m.nc_items.p, m.nb_node, b->row_index.p, b->columns.p, b->q1_plan.p, err.p);
      else
        hipLaunchKernelGGL(k_q1_plan<3>, pgrid, block, 0, stream(), m.cells.p, m.slot_bits, m.nc_offsets.p,
                           m.nc_items.p, m.nb_node, b->row_index.p, b->columns.p, b->q1_plan.p, err.p);
      check_launch("q1_plan");
      int32_t h = 0;
      err.download(&h, 1);
      if (h != 0)
        b->q1_plan.release();
      AFEM_REQUIRE(h == 0, AFEM_ERR_NOT_FOUND, "BSRMatrix(findValueIndex): Value not found");
    }
    // rows per workgroup: the workgroup's values (at most rows * max_row) fit in 64 KB of LDS
    const int rows = b->max_row <= 32 ? 256 : b->max_row <= 64 ? 128 : 64;
    const size_t lds = (size_t)rows * b->max_row * sizeof(double);
    AFEM_REQUIRE(lds <= 64 * 1024, AFEM_ERR_CAPACITY, "Q1 matrix rows too long for the LDS-staged gather");
    const dim3 ngrid(grid_for(m.nb_node, rows));
    if (b->q1_sl_rows != rows || b->q1_plan_sl.n == 0) { // sliced copy, once per pattern
      const int width = std::max(1, m.max_degree);
      const size_t words = (size_t)ngrid.x * rows * width;
      b->q1_plan_sl.alloc(words);
      b->q1_item_sl.alloc(words);
      hipLaunchKernelGGL(k_q1_slice, dim3(grid_for(m.nb_node, 256)), dim3(256), 0, stream(), m.nc_offsets.p,
                         m.nc_items.p, b->q1_plan.p, m.nb_node, width, rows, b->q1_plan_sl.p, b->q1_item_sl.p);
      check_launch("q1_slice");
      b->q1_sl_width = width;
      b->q1_sl_rows = rows;
    }
    auto gather = [&](auto dim_tag, auto bs_tag) {
      constexpr int D = decltype(dim_tag)::value;
      constexpr int BS = decltype(bs_tag)::value;
      hipLaunchKernelGGL((k_q1_gather_sl<D, BS>), ngrid, dim3(BS), lds, stream(), m.slot_bits, m.nc_offsets.p,
                         b->q1_item_sl.p, b->q1_plan_sl.p, b->q1_sl_width, m.nb_node, b->row_index.p, b->q1_elem.p,
                         gather_lambda, b->values.p, f, rhs);
    };
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    if (rows == 256)
      m.dim == 2 ? gather(I2{}, std::integral_constant<int, 256>{}) : gather(I3{}, std::integral_constant<int, 256>{});
    else if (rows == 128)
      m.dim == 2 ? gather(I2{}, std::integral_constant<int, 128>{}) : gather(I3{}, std::integral_constant<int, 128>{});
    else
      m.dim == 2 ? gather(I2{}, std::integral_constant<int, 64>{}) : gather(I3{}, std::integral_constant<int, 64>{});
    check_launch("q1_gather");
  }
  int32_t h = 0;
  err.download(&h, 1);
  AFEM_REQUIRE(h == 0, AFEM_ERR_NOT_FOUND, "BSRMatrix(findValueIndex): Value not found");
}

void q1_assemble_laplace(afem_bsr* b, const double* cell_lambda, double f, double* rhs)
{
  q1_assemble(b, cell_lambda, f, rhs, 0.0);
}

void q1_assemble_heat(afem_bsr* b, const double* cell_lambda, double dt)
{
  q1_assemble(b, cell_lambda, 0.0, nullptr, 1.0 / dt);
}

void q1_assemble_elasticity(afem_bsr* b, double lam, double mu)
{
  const afem_mesh& m = *b->mesh;
  AFEM_REQUIRE(b->nb_dof == m.dim, AFEM_ERR_ARGUMENT, "Q1 elasticity needs nb_dof = dim");
  DBuf<int32_t> err(1);
  err.zero();
  ProfScope prof(m.dim == 2 ? "assemble_elasticity_quad4" : "assemble_elasticity_hexa8");
  const dim3 grid(grid_for(m.nb_node, 256)), block(256);
  if (m.dim == 2)
    hipLaunchKernelGGL(k_q1_elasticity<2>, grid, block, 0, stream(), m.coords.p, m.cells.p, m.slot_bits,
                       m.nc_offsets.p, m.nc_items.p, m.nb_node, b->row_index.p, b->columns.p, b->per_block ? 1 : 0,
                       lam, mu, b->values.p, err.p);
  else
    hipLaunchKernelGGL(k_q1_elasticity<3>, grid, block, 0, stream(), m.coords.p, m.cells.p, m.slot_bits,
                       m.nc_offsets.p, m.nc_items.p, m.nb_node, b->row_index.p, b->columns.p, b->per_block ? 1 : 0,
                       lam, mu, b->values.p, err.p);
  check_launch("q1_elasticity");
  int32_t h = 0;
  err.download(&h, 1);
  AFEM_REQUIRE(h == 0, AFEM_ERR_NOT_FOUND, "BSRMatrix(findValueIndex): Value not found");
}

// rhs[i*nb_dof+dof_index] = sum over the cells of node i of the manufactured
// source against N_i: tri3 source(barycentre) |T| / 3, Quad4 / Hexa8
// sum_g N_i(x_g) source(x_g) det J_g with x_g = sum_a N_a x_a
// (_applyManufacturedSource, modules/poisson/ManufacturedSolutions.cc:66-236).
// DIM = 0 selects tri3.
template <int DIM>
__global__ void __launch_bounds__(256)
k_rhs_manufactured(const double* __restrict__ coords, const int32_t* __restrict__ cells, int slot_bits,
                   const int32_t* __restrict__ nc_off, const int32_t* __restrict__ nc_items, int32_t nb_node,
                   int which, int nb_dof, int dof_index, double* __restrict__ rhs)
{
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb_node)
    return;
  double acc = 0.0;
  for (int32_t e = nc_off[i]; e < nc_off[i + 1]; ++e) {
    const int32_t it = nc_items[e];
    const int s = it & ((1 << slot_bits) - 1);
    const int32_t* cn = cells + (it - s);
    if constexpr (DIM == 0) {
      double x[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d)
          x[a][d] = coords[4 * (int64_t)cn[a] + d];
      const double e1[3] = { x[1][0] - x[0][0], x[1][1] - x[0][1], x[1][2] - x[0][2] };
      const double e2[3] = { x[2][0] - x[0][0], x[2][1] - x[0][1], x[2][2] - x[0][2] };
      const double v0 = e1[1] * e2[2] - e1[2] * e2[1], v1 = e1[2] * e2[0] - e1[0] * e2[2],
                   v2 = e1[0] * e2[1] - e1[1] * e2[0];
      const double area = sqrt(v0 * v0 + v1 * v1 + v2 * v2) / 2.0;
      const double bx = (x[0][0] + x[1][0] + x[2][0]) / 3.0, by = (x[0][1] + x[1][1] + x[2][1]) / 3.0;
      acc += manufactured_source(which, bx, by) * area / 3;
    }
    else {
      constexpr int N = Q1Point<DIM>::N;
      double x[N][3];
#pragma unroll
      for (int a = 0; a < N; ++a) {
        const int32_t n = cn[a];
        x[a][0] = coords[4 * (int64_t)n];
        x[a][1] = coords[4 * (int64_t)n + 1];
        x[a][2] = coords[4 * (int64_t)n + 2];
      }
      for (int g = 0; g < ManufacturedRule<DIM>::NG; ++g) {
        Q1Point<DIM> q;
        const double w = ManufacturedRule<DIM>::point(g, q, x);
        double xg = 0.0, yg = 0.0, ns = 0.0;
#pragma unroll
        for (int a = 0; a < N; ++a) {
          xg += q.Nv[a] * x[a][0];
          yg += q.Nv[a] * x[a][1];
          if (a == s)
            ns = q.Nv[a];
        }
        acc += ns * manufactured_source(which, xg, yg) * (q.det * w);
      }
    }
  }
  rhs[(int64_t)i * nb_dof + dof_index] = acc;
}

// err2[c] = the squared L2 error of cell c, one thread per cell.
template <int DIM>
__global__ void __launch_bounds__(256)
k_manufactured_cell_error(const double* __restrict__ coords, const int32_t* __restrict__ cells, int stride,
                          int32_t nb_cell, int which, const double* __restrict__ u, double* __restrict__ err2)
{
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nb_cell)
    return;
  err2[c] = manufactured_cell_error_sq<DIM>(coords, cells + (int64_t)stride * c, which, u);
}

void manufactured_cell_error(const afem_mesh& m, int which, const double* u, double* err2)
{
  ProfScope prof("manufactured_l2_error");
  const dim3 grid(grid_for(m.nb_cell, 256)), block(256);
  const int stride = 1 << m.slot_bits;
  if (m.simplex)
    hipLaunchKernelGGL(k_manufactured_cell_error<0>, grid, block, 0, stream(), m.coords.p, m.cells.p, stride,
                       m.nb_cell, which, u, err2);
  else if (m.dim == 2)
    hipLaunchKernelGGL(k_manufactured_cell_error<2>, grid, block, 0, stream(), m.coords.p, m.cells.p, stride,
                       m.nb_cell, which, u, err2);
  else
    hipLaunchKernelGGL(k_manufactured_cell_error<3>, grid, block, 0, stream(), m.coords.p, m.cells.p, stride,
                       m.nb_cell, which, u, err2);
  check_launch("manufactured_cell_error");
}

void rhs_manufactured(const afem_mesh& m, int which, int nb_dof, int dof_index, double* rhs)
{
  ProfScope prof("rhs_manufactured");
  const dim3 grid(grid_for(m.nb_node, 256)), block(256);
  if (m.simplex)
    hipLaunchKernelGGL(k_rhs_manufactured<0>, grid, block, 0, stream(), m.coords.p, m.cells.p, m.slot_bits,
                       m.nc_offsets.p, m.nc_items.p, m.nb_node, which, nb_dof, dof_index, rhs);
  else if (m.dim == 2)
    hipLaunchKernelGGL(k_rhs_manufactured<2>, grid, block, 0, stream(), m.coords.p, m.cells.p, m.slot_bits,
                       m.nc_offsets.p, m.nc_items.p, m.nb_node, which, nb_dof, dof_index, rhs);
  else
    hipLaunchKernelGGL(k_rhs_manufactured<3>, grid, block, 0, stream(), m.coords.p, m.cells.p, m.slot_bits,
                       m.nc_offsets.p, m.nc_items.p, m.nb_node, which, nb_dof, dof_index, rhs);
  check_launch("rhs_manufactured");
}

void q1_rhs_mass_field(const afem_mesh& m, const double* field, double scale, int nb_dof, int dof_index, double* rhs)
{
  ProfScope prof("rhs_mass_field_q1");
  const dim3 grid(grid_for(m.nb_node, 256)), block(256);
  if (m.dim == 2)
    hipLaunchKernelGGL(k_q1_mass_field<2>, grid, block, 0, stream(), m.coords.p, m.cells.p, m.slot_bits,
                       m.nc_offsets.p, m.nc_items.p, m.nb_node, field, scale, nb_dof, dof_index, rhs);
  else
    hipLaunchKernelGGL(k_q1_mass_field<3>, grid, block, 0, stream(), m.coords.p, m.cells.p, m.slot_bits,
                       m.nc_offsets.p, m.nc_items.p, m.nb_node, field, scale, nb_dof, dof_index, rhs);
  check_launch("q1_mass_field");
}

void q1_rhs_constant_source(const afem_mesh& m, double f, const double* cell_scale, int nb_dof, int dof_index,
                            double* rhs)
{
  ProfScope prof("rhs_constant_source_q1");
  launch_q1<false>(m, nullptr, nullptr, nullptr, nullptr, f, cell_scale, nb_dof, dof_index, rhs, nullptr);
}

} // namespace afem
