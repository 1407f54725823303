// bsr.hip -- BSRFormat on the MI355X: atomic-free sparsity pattern built from
// the node->cell connectivity, and node-wise gather assembly of the P1
// Laplace / conductivity / elasticity bilinear forms with the row staged in
// LDS (no floating-point atomics anywhere), plus the scalar-CSR expansion.
//
// Reference behaviour: femutils/BSRFormat.h:143-565, femutils/BSRFormat.cc:48-1053.
#include "common.h"
#include "device_utils.h"

#include <algorithm>

#include <hipcub/hipcub.hpp>

namespace afem {

// ---- sparsity -------------------------------------------------------------------
// One thread per node.  The thread walks its node->cell list and inserts the
// other nodes of each cell into a sorted, duplicate-free list held in LDS
// (column-major per block so that lanes hit different banks).  Count pass:
// writes the row length.  Fill pass: writes the sorted row.  A list that
// would exceed CAP raises the overflow flag; the host then retries with a
// larger CAP.

template <int CAP, int BS>
__global__ void __launch_bounds__(BS)
k_pattern_rows(const int32_t* __restrict__ nc_off, const int32_t* __restrict__ nc_items,
               const int32_t* __restrict__ cells, int npc, int slot_bits, int32_t nb_node,
               const int32_t* __restrict__ row_index, int32_t* __restrict__ out,
               int32_t* __restrict__ overflow)
{
  __shared__ int32_t list[CAP * BS];
  const int tid = threadIdx.x;
  const int32_t i = blockIdx.x * BS + tid;
  if (i >= nb_node)
    return;
  int m = 1;
  list[tid] = i; // the diagonal
  bool ovf = false;
  for (int32_t e = nc_off[i]; e < nc_off[i + 1]; ++e) {
    const int32_t it = nc_items[e];
    const int s = it & ((1 << slot_bits) - 1);
    const int32_t* cn = cells + (it - s); // the cell's first node: items are stride * cell + slot
    for (int k = 0; k < npc; ++k) {
      if (k == s)
        continue;
      const int32_t v = cn[k];
      int pos = 0;
      while (pos < m && list[pos * BS + tid] < v)
        ++pos;
      if (pos < m && list[pos * BS + tid] == v)
        continue;
      if (m == CAP) {
        ovf = true;
        continue;
      }
      for (int q = m; q > pos; --q)
        list[q * BS + tid] = list[(q - 1) * BS + tid];
      list[pos * BS + tid] = v;
      ++m;
    }
  }
  if (ovf) {
    atomicOr(overflow, 1);
    return;
  }
  if (row_index == nullptr) {
    out[i] = m;
  }
  else {
    const int32_t base = row_index[i];
    for (int q = 0; q < m; ++q)
      out[base + q] = list[q * BS + tid];
  }
}

// Sixteen lanes per node: lane l takes the node's cells l, l + 16, ... and
// puts the other nodes of each (and, lane 0, the node itself) into its 8
// register slots; the group's 128 keys (padding 0x7fffffff) are sorted by a
// bitonic network -- key g = 8 lane + slot, exchanges within a lane between
// registers and across lanes by shuffles -- and every key that differs from
// its predecessor is a column.  Count pass: the row length; fill pass: the
// sorted row at row_index[i].  *overflow = 1 when a node's candidates do not
// fit its 128 slots (the caller then runs the per-thread kernel).
constexpr int kPatternLanes = 16, kPatternSlots = 8;

__global__ void __launch_bounds__(256)
k_pattern_rows16(const int32_t* __restrict__ nc_off, const int32_t* __restrict__ nc_items,
                 const int32_t* __restrict__ cells, int npc, int slot_bits, int32_t nb_node,
                 const int32_t* __restrict__ row_index, int32_t* __restrict__ out, int32_t* __restrict__ overflow)
{
  constexpr int L = kPatternLanes, S = kPatternSlots;
  const int lane = threadIdx.x & (L - 1);
  const int32_t i = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) / L);
  const bool live = i < nb_node;
  int32_t key[S];
#pragma unroll
  for (int q = 0; q < S; ++q)
    key[q] = 0x7fffffff;
  bool ovf = false;
  if (live) {
    const int32_t e0 = nc_off[i], ncell = nc_off[i + 1] - e0;
    const int rounds = (ncell + L - 1) / L;
    ovf = rounds * (npc - 1) + 1 > S;
    if (!ovf) {
      int q = 0;
      if (lane == 0)
        key[q++] = i; // the diagonal
      for (int r = 0; r < rounds; ++r) {
        const int32_t c = lane + L * r;
        if (c < ncell) {
          const int32_t it = nc_items[e0 + c];
          const int sl = it & ((1 << slot_bits) - 1);
          const int32_t* cn = cells + (it - sl); // items are stride * cell + slot
          for (int k = 0; k < npc; ++k)
            if (k != sl) {
              const int32_t v = cn[k];
#pragma unroll
              for (int z = 0; z < S; ++z)
                if (z == q)
                  key[z] = v;
              ++q;
            }
        }
      }
    }
  }
  // every lane of a group takes the same decision: a group overflows as a whole
  if (ovf)
    atomicOr(overflow, 1);
  // bitonic sort of the group's 128 keys, ascending in g = S * lane + slot
#pragma unroll
  for (int k = 2; k <= L * S; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= S) {
        const int m = j / S; // partner lane = lane ^ m
        const bool lower = (lane & m) == 0;
#pragma unroll
        for (int q = 0; q < S; ++q) {
          const int g = S * lane + q;
          const bool asc = (g & k) == 0;
          const int32_t w = __shfl_xor(key[q], m, 64);
          key[q] = (lower == asc) ? min(key[q], w) : max(key[q], w);
        }
      }
      else {
#pragma unroll
        for (int q = 0; q < S; ++q) {
          if (q & j)
            continue;
          const int g = S * lane + q;
          const bool asc = (g & k) == 0;
          const int32_t a = key[q], b = key[q | j];
          key[q] = asc ? min(a, b) : max(a, b);
          key[q | j] = asc ? max(a, b) : min(a, b);
        }
      }
    }
  }
  // columns: keys that differ from their predecessor (and are not padding)
  int32_t prev = __shfl_up(key[S - 1], 1, L);
  if (lane == 0)
    prev = -1;
  int mine = 0;
#pragma unroll
  for (int q = 0; q < S; ++q) {
    const int32_t before = q == 0 ? prev : key[q - 1];
    mine += (key[q] != before && key[q] != 0x7fffffff) ? 1 : 0;
  }
  int incl = mine;
#pragma unroll
  for (int d = 1; d < L; d <<= 1) {
    const int u = __shfl_up(incl, d, L);
    if (lane >= d)
      incl += u;
  }
  if (!live || ovf)
    return;
  if (row_index == nullptr) {
    if (lane == L - 1)
      out[i] = incl;
  }
  else {
    int32_t at = row_index[i] + incl - mine;
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const int32_t before = q == 0 ? prev : key[q - 1];
      if (key[q] != before && key[q] != 0x7fffffff)
        out[at++] = key[q];
    }
  }
}

// Runs one pass (count if row_index == nullptr, fill otherwise) with the
// smallest LDS capacity that holds every row.  Returns false on overflow at
// the largest capacity.
static bool pattern_pass(const afem_mesh& m, const int32_t* row_index, int32_t* out, int cap_hint)
{
  DBuf<int32_t> ovf(1);
  {
    // sixteen lanes per node when every node's candidates fit 128 slots
    ovf.zero();
    {
      ProfScope prof("pattern_rows");
      const int64_t threads = (int64_t)m.nb_node * kPatternLanes;
      hipLaunchKernelGGL(k_pattern_rows16, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, stream(),
                         m.nc_offsets.p, m.nc_items.p, m.cells.p, m.npc, m.slot_bits, m.nb_node, row_index, out,
                         ovf.p);
    }
    check_launch("pattern_rows16");
    int32_t flag = 0;
    ovf.download(&flag, 1);
    if (flag == 0)
      return true;
  }
  auto run = [&](auto cap_tag, auto bs_tag) {
    constexpr int CAP = decltype(cap_tag)::value;
    constexpr int BS = decltype(bs_tag)::value;
    ovf.zero();
    {
      ProfScope prof("pattern_rows");
      hipLaunchKernelGGL((k_pattern_rows<CAP, BS>), dim3(grid_for(m.nb_node, BS)), dim3(BS), 0, stream(),
                         m.nc_offsets.p, m.nc_items.p, m.cells.p, m.npc, m.slot_bits, m.nb_node, row_index, out,
                         ovf.p);
    }
    check_launch("pattern_rows");
    int32_t flag = 0;
    ovf.download(&flag, 1);
    return flag == 0;
  };
  if (cap_hint <= 32 && run(std::integral_constant<int, 32>{}, std::integral_constant<int, 256>{}))
    return true;
  if (cap_hint <= 128 && run(std::integral_constant<int, 128>{}, std::integral_constant<int, 64>{}))
    return true;
  return run(std::integral_constant<int, 512>{}, std::integral_constant<int, 16>{});
}

static int row_cap_hint(const afem_mesh& m)
{
  int64_t bound = 1 + (int64_t)(m.npc - 1) * m.max_degree;
  return bound > 512 ? 512 : (int)bound;
}

// nb_col of BSRFormat::computeNbColumns (femutils/BSRFormat.cc:256-345): for P1
// simplices the number of node pairs sharing a cell (N + 2E).
static void bsr_count_columns(afem_bsr* b)
{
  const afem_mesh& m = *b->mesh;
  b->nnz_per_row.alloc((size_t)m.nb_node);
  if (!pattern_pass(m, nullptr, b->nnz_per_row.p, row_cap_hint(m)))
    fail(AFEM_ERR_CAPACITY, "a node has more than 511 neighbours");
  b->row_index.alloc((size_t)m.nb_node + 1);
  int64_t total = exclusive_scan_i32(b->nnz_per_row.p, b->row_index.p, m.nb_node);
  b->nb_col = total;
  AFEM_REQUIRE(total * b->nb_dof * b->nb_dof < (1LL << 31), AFEM_ERR_ARGUMENT,
               "matrix too large for 32-bit value offsets");
}

__global__ void k_max_row(const int32_t* __restrict__ len, int32_t n, int32_t* __restrict__ out)
{
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  int32_t v = (i < n) ? len[i] : 0;
  v = max(v, __shfl_xor(v, 32, 64));
  v = max(v, __shfl_xor(v, 16, 64));
  v = max(v, __shfl_xor(v, 8, 64));
  v = max(v, __shfl_xor(v, 4, 64));
  v = max(v, __shfl_xor(v, 2, 64));
  v = max(v, __shfl_xor(v, 1, 64));
  if ((threadIdx.x & 63) == 0)
    atomicMax(out, v);
}

static void bsr_compute_sparsity(afem_bsr* b)
{
  const afem_mesh& m = *b->mesh;
  b->columns.alloc((size_t)b->nb_col);
  if (!pattern_pass(m, b->row_index.p, b->columns.p, row_cap_hint(m)))
    fail(AFEM_ERR_CAPACITY, "a node has more than 511 neighbours");
  DBuf<int32_t> mx(1);
  mx.zero();
  hipLaunchKernelGGL(k_max_row, dim3(grid_for(m.nb_node, 256)), dim3(256), 0, stream(), b->nnz_per_row.p,
                     m.nb_node, mx.p);
  check_launch("max_row");
  mx.download(&b->max_row, 1);
  b->values.alloc((size_t)b->nb_col * b->nb_dof * b->nb_dof);
  b->values.zero();
  b->has_sparsity = true;
  ++b->pattern_gen;
}

static void build_plan(afem_bsr* b);
// derived plans, built on first use (afem_bsr::plans_tried)
constexpr unsigned kPlansSliced = 1, kPlansBrick = 2, kPlansEdge = 4, kPlansCell = 8;
static void ensure_plans(afem_bsr* b, unsigned what);

// ---- element kernels ---------------------------------------------------------------

struct d3
{
  double x, y, z;
};

__device__ __forceinline__ d3 load_node(const double* __restrict__ coords, int32_t n)
{
  const double2 xy = *reinterpret_cast<const double2*>(coords + 4 * (int64_t)n);
  return { xy.x, xy.y, coords[4 * (int64_t)n + 2] };
}

__device__ __forceinline__ d3 sub(d3 a, d3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
__device__ __forceinline__ d3 cross(d3 a, d3 b)
{
  return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
}
__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// 1/x from the hardware FP64 reciprocal estimate refined by two
// Newton-Raphson steps, y <- y + y (1 - x y); within an ulp or two of the
// correctly rounded quotient (x > 0 finite).
__device__ __forceinline__ double fast_rcp(double x)
{
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

// Position of column v inside a sorted row held in LDS.
__device__ __forceinline__ int find_col(const int32_t* __restrict__ cols, int len, int32_t v)
{
  int lo = 0, hi = len;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (cols[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// The other nodes of a cell seen from its node at slot s, in cell order.
__device__ __forceinline__ int3 other_nodes(int4 cq, int s)
{
  return { (s == 0) ? cq.y : cq.x, (s <= 1) ? cq.z : cq.y, (s <= 2) ? cq.w : cq.z };
}

__device__ __forceinline__ void lds_add(const int32_t* __restrict__ rc, double* rv, int len, int32_t col, double v)
{
  const int pos = find_col(rc, len, col);
  if (pos < len)
    rv[pos] += v;
}

// Node-wise gather assembly of the P1 Laplace form, one thread per row.
// The block's rows are consecutive, so their values form one contiguous
// range [row_index[r0], row_index[r1]) that is staged in LDS (columns +
// accumulators), zero-filled, accumulated by each row's owner thread without
// atomics, and written back with coalesced stores.
//
// Each (row, cell) pair recomputes the element row: the element kernel runs
// npc times per cell, which costs FP64 ALU instead of HBM bytes.  The row
// node is taken as the origin: with e_k = p_k - p (the other nodes in cell
// order), C_k = e_{k+1} x e_{k+2} are det * grad N_k, det = e_a . C_a and
// C_p = -(C_a + C_b + C_c); then K_pq = lambda (C_p . C_q) / (6 |det|).  The
// cells of a row are processed U at a time with all their index and
// coordinate loads issued before any arithmetic, so a thread keeps U
// dependent load chains in flight.  The same pass writes the constant-source
// RHS b_p = sum f |T| / npc when rhs is given.
template <int NPC, int U>
__global__ void __launch_bounds__(256)
k_assemble_laplace(const double* __restrict__ coords, const int32_t* __restrict__ cells,
                   const int32_t* __restrict__ nc_off, const int32_t* __restrict__ nc_items,
                   const int32_t* __restrict__ row_index, const int32_t* __restrict__ columns,
                   int32_t nb_node, int rows_per_block,
                   const double* __restrict__ cell_lambda, double source_f,
                   double* __restrict__ rhs, double* __restrict__ values)
{
  extern __shared__ unsigned char smem[];
  const int32_t tile = (int32_t)xcd_tile(blockIdx.x, gridDim.x);
  const int32_t r0 = tile * rows_per_block;
  const int32_t r1 = min(r0 + rows_per_block, nb_node);
  const int32_t a = row_index[r0];
  const int32_t cnt = row_index[r1] - a;
  double* s_val = reinterpret_cast<double*>(smem);
  int32_t* s_col = reinterpret_cast<int32_t*>(s_val + cnt);
  for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
    s_col[k] = columns[a + k];
    s_val[k] = 0.0;
  }
  __syncthreads();
  const int32_t i = r0 + threadIdx.x;
  if (threadIdx.x < rows_per_block && i < r1) {
    const int32_t rb = row_index[i] - a;
    const int len = row_index[i + 1] - row_index[i];
    const int32_t* rc = s_col + rb;
    double* rv = s_val + rb;
    const int dpos = find_col(rc, len, i);
    const d3 p = load_node(coords, i);
    const int32_t e0 = nc_off[i], e1 = nc_off[i + 1];
    double diag = 0.0, b_acc = 0.0;
    for (int32_t e = e0; e < e1; e += U) {
      int32_t it[U];
      int4 cq[U];
      double lam[U];
      d3 pa[U], pb[U], pc[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        it[u] = nc_items[min(e + u, e1 - 1)];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t c = it[u] >> 2;
        cq[u] = *reinterpret_cast<const int4*>(cells + 4 * (int64_t)c);
        lam[u] = cell_lambda ? cell_lambda[c] : 1.0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int3 o = other_nodes(cq[u], it[u] & 3);
        pa[u] = load_node(coords, o.x);
        pb[u] = load_node(coords, o.y);
        if constexpr (NPC == 4)
          pc[u] = load_node(coords, o.z);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (e + u >= e1)
          break;
        const int3 o = other_nodes(cq[u], it[u] & 3);
        if constexpr (NPC == 4) {
          const d3 ea = sub(pa[u], p), eb = sub(pb[u], p), ec = sub(pc[u], p);
          const d3 Ca = cross(eb, ec), Cb = cross(ec, ea), Cc = cross(ea, eb);
          const double adet = fabs(dot(ea, Ca));
          const d3 Cp = { -(Ca.x + Cb.x + Cc.x), -(Ca.y + Cb.y + Cc.y), -(Ca.z + Cb.z + Cc.z) };
          const double scale = lam[u] * fast_rcp(6.0 * adet);
          diag += scale * dot(Cp, Cp);
          lds_add(rc, rv, len, o.x, scale * dot(Cp, Ca));
          lds_add(rc, rv, len, o.y, scale * dot(Cp, Cb));
          lds_add(rc, rv, len, o.z, scale * dot(Cp, Cc));
          b_acc += adet; // b_p = f * sum |det| / 24
        }
        else {
          const d3 ea = sub(pa[u], p), eb = sub(pb[u], p);
          const double det = ea.x * eb.y - eb.x * ea.y;
          const double adet = fabs(det);
          const double gax = eb.y, gay = -eb.x, gbx = -ea.y, gby = ea.x;
          const double gpx = -(gax + gbx), gpy = -(gay + gby);
          const double scale = lam[u] * fast_rcp(2.0 * adet);
          diag += scale * (gpx * gpx + gpy * gpy);
          lds_add(rc, rv, len, o.x, scale * (gpx * gax + gpy * gay));
          lds_add(rc, rv, len, o.y, scale * (gpx * gbx + gpy * gby));
          b_acc += adet; // b_p = f * sum |det| / 6
        }
      }
    }
    if (dpos < len)
      rv[dpos] += diag;
    if (rhs)
      rhs[i] = source_f * b_acc / (NPC == 4 ? 24.0 : 6.0);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < cnt; k += blockDim.x)
    values[a + k] = s_val[k];
}

// Node-wise gather assembly of the 12x12 P1 elasticity element into 3x3
// blocks: K_ab = (1/(6|det|)) [lambda C_a C_b^T + mu C_b C_a^T + mu (C_a.C_b) I],
// the closed form of modules/elasticity/ElementMatrix.h:151-211.  Blocks are
// staged in LDS (9 doubles each, row-major) and written in the requested
// order: per block (block*9 + i*3 + j) or per row (scalar CSR order,
// row_index[a]*9 + 3*(x + i*nnz_row) + j).  Same origin-at-the-row-node
// cofactors and U-cell batching as the Laplace kernel.
__device__ __forceinline__ void lds_add_block(const int32_t* __restrict__ rc, double* rv, int len, int32_t col,
                                              const d3& cp, const d3& cq, double la, double mu)
{
  const int pos = find_col(rc, len, col);
  if (pos >= len)
    return;
  const double a[3] = { cp.x, cp.y, cp.z };
  const double b[3] = { cq.x, cq.y, cq.z };
  const double dd = mu * (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]);
  double* blk = rv + 9 * pos;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int t = 0; t < 3; ++t)
      blk[r * 3 + t] += la * a[r] * b[t] + mu * b[r] * a[t] + (r == t ? dd : 0.0);
}

template <int U>
__global__ void __launch_bounds__(256)
k_assemble_elasticity(const double* __restrict__ coords, const int32_t* __restrict__ cells,
                      const int32_t* __restrict__ nc_off, const int32_t* __restrict__ nc_items,
                      const int32_t* __restrict__ row_index, const int32_t* __restrict__ columns,
                      int32_t nb_node, int rows_per_block, double lame_lambda, double lame_mu,
                      int per_block, double* __restrict__ values)
{
  extern __shared__ unsigned char smem[];
  const int32_t tile = (int32_t)xcd_tile(blockIdx.x, gridDim.x);
  const int32_t r0 = tile * rows_per_block;
  const int32_t r1 = min(r0 + rows_per_block, nb_node);
  const int32_t a = row_index[r0];
  const int32_t cnt = row_index[r1] - a;
  double* s_val = reinterpret_cast<double*>(smem);
  int32_t* s_col = reinterpret_cast<int32_t*>(s_val + 9 * (int64_t)cnt);
  for (int k = threadIdx.x; k < cnt; k += blockDim.x)
    s_col[k] = columns[a + k];
  for (int k = threadIdx.x; k < 9 * cnt; k += blockDim.x)
    s_val[k] = 0.0;
  __syncthreads();
  const int32_t i = r0 + threadIdx.x;
  if (threadIdx.x < rows_per_block && i < r1) {
    const int32_t rb = row_index[i] - a;
    const int len = row_index[i + 1] - row_index[i];
    const int32_t* rc = s_col + rb;
    double* rv = s_val + 9 * (int64_t)rb;
    const d3 p = load_node(coords, i);
    const int32_t e0 = nc_off[i], e1 = nc_off[i + 1];
    for (int32_t e = e0; e < e1; e += U) {
      int32_t it[U];
      int4 cq[U];
      d3 pa[U], pb[U], pc[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        it[u] = nc_items[min(e + u, e1 - 1)];
#pragma unroll
      for (int u = 0; u < U; ++u)
        cq[u] = *reinterpret_cast<const int4*>(cells + 4 * (int64_t)(it[u] >> 2));
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int3 o = other_nodes(cq[u], it[u] & 3);
        pa[u] = load_node(coords, o.x);
        pb[u] = load_node(coords, o.y);
        pc[u] = load_node(coords, o.z);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (e + u >= e1)
          break;
        const int3 o = other_nodes(cq[u], it[u] & 3);
        const d3 ea = sub(pa[u], p), eb = sub(pb[u], p), ec = sub(pc[u], p);
        const d3 Ca = cross(eb, ec), Cb = cross(ec, ea), Cc = cross(ea, eb);
        const double adet = fabs(dot(ea, Ca));
        const d3 Cp = { -(Ca.x + Cb.x + Cc.x), -(Ca.y + Cb.y + Cc.y), -(Ca.z + Cb.z + Cc.z) };
        const double inv6 = fast_rcp(6.0 * adet);
        const double la = lame_lambda * inv6, mu = lame_mu * inv6;
        lds_add_block(rc, rv, len, i, Cp, Cp, la, mu);
        lds_add_block(rc, rv, len, o.x, Cp, Ca, la, mu);
        lds_add_block(rc, rv, len, o.y, Cp, Cb, la, mu);
        lds_add_block(rc, rv, len, o.z, Cp, Cc, la, mu);
      }
    }
  }
  __syncthreads();
  if (per_block) {
    for (int k = threadIdx.x; k < 9 * cnt; k += blockDim.x)
      values[9 * (int64_t)a + k] = s_val[k];
  }
  else {
    // one thread per row: scatter the row's blocks into scalar-CSR order
    if (threadIdx.x < rows_per_block && i < r1) {
      const int32_t rb = row_index[i] - a;
      const int len = row_index[i + 1] - row_index[i];
      const int64_t base = 9 * (int64_t)row_index[i];
      for (int r = 0; r < 3; ++r)
        for (int x = 0; x < len; ++x)
          for (int t = 0; t < 3; ++t)
            values[base + 3 * (x + (int64_t)r * len) + t] = s_val[9 * (int64_t)(rb + x) + r * 3 + t];
    }
  }
}

// ---- assembly plan --------------------------------------------------------------------
// Built once with the pattern: for every (row node i, cell c) pair, in
// node->cell order, one 64-bit word
//   bits 0-31 cell id, 32-33 slot of i in c,
//   34-41 / 42-49 / 50-57 positions inside row i of the other nodes of c
//   (in cell order, as returned by other_nodes),
// and per row the position of the diagonal.  With it the assembly needs no
// column search and no column staging.  Rows longer than 255 entries do
// not get a plan (the searching kernels handle them).

constexpr int kPlanMaxRow = 255;

__device__ __forceinline__ int32_t plan_cell(uint64_t w) { return (int32_t)(w & 0xffffffffULL); }
__device__ __forceinline__ int plan_slot(uint64_t w) { return (int)((w >> 32) & 3); }
__device__ __forceinline__ int plan_pos(uint64_t w, int k) { return (int)((w >> (34 + 8 * k)) & 0xff); }

// One thread per row; the workgroup's 256 rows of columns are staged in LDS
// when they fit (8192 entries), so the binary searches of the row positions
// read LDS instead of re-reading the row from memory for every search.
constexpr int kPairStage = 8192;

template <class RowPtr>
__device__ __forceinline__ bool pair_plan_row(RowPtr rc, int len, int32_t i, const int32_t* __restrict__ nc_off,
                                              const int32_t* __restrict__ nc_items, const int32_t* __restrict__ cells,
                                              int npc, uint64_t* __restrict__ plan, int32_t* __restrict__ diag_pos)
{
  const int d = find_col(rc, len, i);
  diag_pos[i] = d;
  bool bad = (d >= len || rc[d] != i);
  for (int32_t e = nc_off[i]; e < nc_off[i + 1]; ++e) {
    const int32_t it = nc_items[e];
    const int32_t c = it >> 2;
    const int s = it & 3;
    const int4 cq = *reinterpret_cast<const int4*>(cells + 4 * (int64_t)c);
    const int3 o = other_nodes(cq, s);
    const int32_t nodes[3] = { o.x, o.y, o.z };
    uint64_t w = (uint64_t)(uint32_t)c | ((uint64_t)s << 32);
    for (int k = 0; k < npc - 1; ++k) {
      const int pos = find_col(rc, len, nodes[k]);
      bad = bad || pos >= len || rc[pos] != nodes[k] || pos > kPlanMaxRow;
      w |= (uint64_t)(pos & 0xff) << (34 + 8 * k);
    }
    plan[e] = w;
  }
  return bad;
}

__global__ void __launch_bounds__(256)
k_pair_plan(const int32_t* __restrict__ nc_off, const int32_t* __restrict__ nc_items,
            const int32_t* __restrict__ cells, int npc, int32_t nb_node, const int32_t* __restrict__ row_index,
            const int32_t* __restrict__ columns, uint64_t* __restrict__ plan, int32_t* __restrict__ diag_pos,
            int32_t* __restrict__ err)
{
  __shared__ int32_t s_cols[kPairStage];
  const int32_t r0 = blockIdx.x * 256;
  const int32_t r1 = min(r0 + 256, nb_node);
  const int32_t a = row_index[r0];
  const int32_t cnt = row_index[r1] - a;
  const bool staged = cnt <= kPairStage;
  if (staged)
    for (int32_t k = threadIdx.x; k < cnt; k += 256)
      s_cols[k] = columns[a + k];
  __syncthreads();
  const int32_t i = r0 + threadIdx.x;
  if (i >= r1)
    return;
  const int len = row_index[i + 1] - row_index[i];
  const bool bad = staged ? pair_plan_row(s_cols + (row_index[i] - a), len, i, nc_off, nc_items, cells, npc, plan,
                                          diag_pos)
                          : pair_plan_row(columns + row_index[i], len, i, nc_off, nc_items, cells, npc, plan, diag_pos);
  if (bad)
    atomicOr(err, 1);
}

// Row positions stored in 5 bits (the local and brick plans): rows of at most
// 32 entries.
constexpr int kCompactMaxRow = 32;

// ---- block-local plan ---------------------------------------------------------------------
// Rows are split into blocks of kLocalRows consecutive rows.  The nodes a
// block's assembly touches are exactly the union of its rows' columns; that
// sorted set is the block's node table (at most kLocalMaxNodes entries).  The
// local plan word of a (row, cell) pair holds the table indices of the other
// nodes (11 bits each) and their row positions (5 bits each):
//   bits 0-10 / 11-21 / 22-32 table indices, 33-37 / 38-42 / 43-47 positions.
// The assembly stages the block's coordinates in LDS once and then reads one
// word per pair from HBM.
constexpr int kLocalRows = 256;
constexpr int kLocalMaxNodes = 2048;
constexpr int kLocalSort = 8192;

__device__ __forceinline__ int lplan_idx(uint64_t w, int k) { return (int)((w >> (11 * k)) & 2047); }

// a += v on an LDS double owned by the calling thread alone: the LDS unit
// does the read-modify-write (ds_add_f64), so the thread does not wait for
// the old value; with a single writer the summation order is program order.
__device__ __forceinline__ void lds_accumulate(double* a, double v)
{
  __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ int lplan_pos(uint64_t w, int k) { return (int)((w >> (33 + 5 * k)) & 31); }

// LDS coordinates of the other nodes of one local-plan word.
template <int NPC>
struct PairNodes
{
  double2 xy[NPC - 1];
  double z[NPC - 1];
};

template <int NPC>
__device__ __forceinline__ PairNodes<NPC> pair_nodes(const double2* __restrict__ s_xy, const double* __restrict__ s_z,
                                                     uint64_t w)
{
  PairNodes<NPC> n;
#pragma unroll
  for (int k = 0; k < NPC - 1; ++k) {
    const int l = lplan_idx(w, k);
    n.xy[k] = s_xy[l];
    n.z[k] = NPC == 4 ? s_z[l] : 0.0;
  }
  return n;
}

// Row p of one element from a local-plan word: the diagonal entry and |det|
// are summed in registers, the other entries added into the row's LDS slots
// (slot k at rv[stride * k]).  Edges relative to the row node p (as in
// k_assemble_laplace).
template <int NPC>
__device__ __forceinline__ void pair_element(const PairNodes<NPC>& n, d3 p, double lam, uint64_t w, double* rv,
                                             int stride, double& diag, double& det_sum)
{
  const d3 ea = { n.xy[0].x - p.x, n.xy[0].y - p.y, n.z[0] - p.z };
  const d3 eb = { n.xy[1].x - p.x, n.xy[1].y - p.y, n.z[1] - p.z };
  if constexpr (NPC == 4) {
    const d3 ec = { n.xy[2].x - p.x, n.xy[2].y - p.y, n.z[2] - p.z };
    const d3 Ca = cross(eb, ec), Cb = cross(ec, ea), Cc = cross(ea, eb);
    const double adet = fabs(dot(ea, Ca));
    const d3 Cp = { -(Ca.x + Cb.x + Cc.x), -(Ca.y + Cb.y + Cc.y), -(Ca.z + Cb.z + Cc.z) };
    const double scale = lam * fast_rcp(6.0 * adet);
    diag += scale * dot(Cp, Cp);
    lds_accumulate(rv + stride * lplan_pos(w, 0), scale * dot(Cp, Ca));
    lds_accumulate(rv + stride * lplan_pos(w, 1), scale * dot(Cp, Cb));
    lds_accumulate(rv + stride * lplan_pos(w, 2), scale * dot(Cp, Cc));
    det_sum += adet;
  }
  else {
    const double adet = fabs(ea.x * eb.y - eb.x * ea.y);
    const double gax = eb.y, gay = -eb.x, gbx = -ea.y, gby = ea.x;
    const double gpx = -(gax + gbx), gpy = -(gay + gby);
    const double scale = lam * fast_rcp(2.0 * adet);
    diag += scale * (gpx * gpx + gpy * gpy);
    lds_accumulate(rv + stride * lplan_pos(w, 0), scale * (gpx * gax + gpy * gay));
    lds_accumulate(rv + stride * lplan_pos(w, 1), scale * (gpx * gbx + gpy * gby));
    det_sum += adet;
  }
}

// All the (row, cell) pairs of one row from its local-plan words, word j at
// plan[wbase + j * wstride] (cell id at cell_of[...]).  Full batches of U
// words run without branches, the coordinates of pair u + 1 read from LDS
// before pair u is computed; the last partial batch stops at ncell.
template <int NPC, int U, bool CELL_SHIFT>
__device__ __forceinline__ void row_pairs(const uint64_t* __restrict__ plan, const int32_t* __restrict__ cell_of,
                                          int64_t wbase, int wstride, int32_t ncell,
                                          const double* __restrict__ cell_lambda, const double2* __restrict__ s_xy,
                                          const double* __restrict__ s_z, d3 p, double* rv, int stride, double& diag,
                                          double& det_sum)
{
  int32_t e = 0;
  for (; e + U <= ncell; e += U) {
    uint64_t w[U];
    double lam[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t at = wbase + (int64_t)(e + u) * wstride;
      w[u] = plan[at];
      lam[u] = cell_lambda ? cell_lambda[CELL_SHIFT ? (cell_of[at] >> 2) : cell_of[at]] : 1.0;
    }
    PairNodes<NPC> next = pair_nodes<NPC>(s_xy, s_z, w[0]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const PairNodes<NPC> cur = next;
      if (u + 1 < U)
        next = pair_nodes<NPC>(s_xy, s_z, w[u + 1]);
      pair_element<NPC>(cur, p, lam[u], w[u], rv, stride, diag, det_sum);
    }
  }
  if (e < ncell) {
    uint64_t w[U];
    double lam[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t at = wbase + (int64_t)min(e + u, ncell - 1) * wstride;
      w[u] = plan[at];
      lam[u] = cell_lambda ? cell_lambda[CELL_SHIFT ? (cell_of[at] >> 2) : cell_of[at]] : 1.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (e + u >= ncell)
        break;
      pair_element<NPC>(pair_nodes<NPC>(s_xy, s_z, w[u]), p, lam[u], w[u], rv, stride, diag, det_sum);
    }
  }
}

// Sorted unique columns of each block of rows (bitonic sort in LDS, then a
// compaction with a block-wide scan).  count < 0 marks a block whose rows
// hold more than kLocalSort entries.
__global__ void __launch_bounds__(256)
k_block_nodes(const int32_t* __restrict__ row_index, const int32_t* __restrict__ columns, int32_t nb_node,
              int32_t* __restrict__ table, int32_t* __restrict__ count)
{
  __shared__ int32_t buf[kLocalSort];
  __shared__ int32_t scan[256];
  const int tid = threadIdx.x;
  const int32_t blk = blockIdx.x;
  const int32_t r0 = blk * kLocalRows;
  const int32_t r1 = min(r0 + kLocalRows, nb_node);
  const int32_t a = row_index[r0];
  const int32_t cnt = row_index[r1] - a;
  if (cnt > kLocalSort) {
    if (tid == 0)
      count[blk] = -1;
    return;
  }
  int n2 = 256;
  while (n2 < cnt)
    n2 <<= 1;
  for (int k = tid; k < n2; k += 256)
    buf[k] = (k < cnt) ? columns[a + k] : 0x7fffffff;
  __syncthreads();
  for (int size = 2; size <= n2; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int k = tid; k < n2 / 2; k += 256) {
        const int i = (k / stride) * 2 * stride + (k % stride);
        const int j = i + stride;
        const bool asc = (i & size) == 0;
        const int32_t x = buf[i], y = buf[j];
        if ((x > y) == asc) {
          buf[i] = y;
          buf[j] = x;
        }
      }
      __syncthreads();
    }
  // each thread owns a contiguous chunk; count first occurrences, scan, write
  const int chunk = n2 / 256;
  const int k0 = tid * chunk;
  int mine = 0;
  for (int k = k0; k < k0 + chunk; ++k)
    if (k < cnt && (k == 0 || buf[k] != buf[k - 1]))
      ++mine;
  scan[tid] = mine;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int v = (tid >= off) ? scan[tid - off] : 0;
    __syncthreads();
    scan[tid] += v;
    __syncthreads();
  }
  int pos = scan[tid] - mine;
  for (int k = k0; k < k0 + chunk; ++k)
    if (k < cnt && (k == 0 || buf[k] != buf[k - 1])) {
      if (pos < kLocalMaxNodes)
        table[(int64_t)blk * kLocalMaxNodes + pos] = buf[k];
      ++pos;
    }
  if (tid == 255)
    count[blk] = scan[255];
}

__device__ __forceinline__ int find_node(const int32_t* __restrict__ tab, int n, int32_t v)
{
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (tab[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(256)
k_local_plan(const int32_t* __restrict__ nc_off, const uint64_t* __restrict__ plan,
             const int32_t* __restrict__ cells, int npc, int32_t nb_node, const int32_t* __restrict__ table,
             const int32_t* __restrict__ count, uint64_t* __restrict__ lplan, int32_t* __restrict__ err)
{
  __shared__ int32_t tab[kLocalMaxNodes];
  const int32_t blk = blockIdx.x;
  const int tn = count[blk];
  if (tn < 0 || tn > kLocalMaxNodes) {
    if (threadIdx.x == 0)
      atomicOr(err, 1);
    return;
  }
  for (int k = threadIdx.x; k < tn; k += 256)
    tab[k] = table[(int64_t)blk * kLocalMaxNodes + k];
  __syncthreads();
  const int32_t i = blk * kLocalRows + threadIdx.x;
  if (i >= nb_node)
    return;
  bool bad = false;
  for (int32_t e = nc_off[i]; e < nc_off[i + 1]; ++e) {
    const uint64_t w = plan[e];
    const int4 cq = *reinterpret_cast<const int4*>(cells + 4 * (int64_t)plan_cell(w));
    const int3 o = other_nodes(cq, plan_slot(w));
    const int32_t nodes[3] = { o.x, o.y, o.z };
    uint64_t lw = 0;
    for (int k = 0; k < npc - 1; ++k) {
      const int l = find_node(tab, tn, nodes[k]);
      const int pos = plan_pos(w, k);
      bad = bad || l >= tn || tab[l] != nodes[k] || pos >= kCompactMaxRow;
      lw |= (uint64_t)(l & 2047) << (11 * k);
      lw |= (uint64_t)(pos & 31) << (33 + 5 * k);
    }
    lplan[e] = lw;
  }
  if (bad)
    atomicOr(err, 1);
}

static void build_local_plan(afem_bsr* b)
{
  const afem_mesh& m = *b->mesh;
  b->lplan.release();
  b->block_nodes.release();
  b->block_count.release();
  b->block_nodes_max = 0;
  if (!b->plan.p || b->max_row > kCompactMaxRow)
    return;
  const int32_t nblk = (m.nb_node + kLocalRows - 1) / kLocalRows;
  b->block_nodes.alloc((size_t)nblk * kLocalMaxNodes);
  b->block_count.alloc((size_t)nblk);
  hipLaunchKernelGGL(k_block_nodes, dim3(nblk), dim3(256), 0, stream(), b->row_index.p, b->columns.p, m.nb_node,
                     b->block_nodes.p, b->block_count.p);
  check_launch("block_nodes");
  std::vector<int32_t> counts((size_t)nblk);
  b->block_count.download(counts.data(), counts.size());
  int32_t mx = 0;
  for (int32_t c : counts) {
    if (c < 0 || c > kLocalMaxNodes) {
      b->block_nodes.release();
      b->block_count.release();
      return;
    }
    mx = std::max(mx, c);
  }
  b->lplan.alloc(b->plan.n);
  DBuf<int32_t> err(1);
  err.zero();
  hipLaunchKernelGGL(k_local_plan, dim3(nblk), dim3(256), 0, stream(), m.nc_offsets.p, b->plan.p, m.cells.p, m.npc,
                     m.nb_node, b->block_nodes.p, b->block_count.p, b->lplan.p, err.p);
  check_launch("local_plan");
  int32_t h = 0;
  err.download(&h, 1);
  if (h) {
    b->lplan.release();
    b->block_nodes.release();
    b->block_count.release();
    return;
  }
  b->block_nodes_max = mx;
}

// ---- brick-ordered local plan ---------------------------------------------------------------
// Rows are processed in bricks of kLocalRows nodes that are consecutive in
// the Morton order of the node coordinates (8 x 8 x 4 nodes on a structured
// grid), so a brick's node table is its nodes plus a one-layer halo (about
// 600 nodes instead of about 1800 for 256 lexicographically consecutive
// rows).  Row numbering is unchanged: each thread owns one row of its brick,
// accumulates it in LDS at the row's offset in the brick, and writes the
// row's segment of the value array at the end.

// Row offsets of each brick slot in LDS (exclusive scan of the row lengths)
// and the brick's total; total < 0 when it exceeds kLocalSort.
__global__ void __launch_bounds__(256)
k_brick_offsets(const int32_t* __restrict__ perm, const int32_t* __restrict__ row_index, int32_t nb_node,
                int32_t* __restrict__ val_off, int32_t* __restrict__ total)
{
  __shared__ int32_t scan[256];
  const int tid = threadIdx.x;
  const int64_t slot = (int64_t)blockIdx.x * kLocalRows + tid;
  int32_t len = 0;
  if (slot < nb_node) {
    const int32_t r = perm[slot];
    len = row_index[r + 1] - row_index[r];
  }
  scan[tid] = len;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int v = (tid >= off) ? scan[tid - off] : 0;
    __syncthreads();
    scan[tid] += v;
    __syncthreads();
  }
  if (slot < nb_node)
    val_off[slot] = scan[tid] - len;
  if (tid == 255)
    total[blockIdx.x] = scan[255] <= kLocalSort ? scan[255] : -1;
}

// Sorted unique columns of each brick's rows (as k_block_nodes).
__global__ void __launch_bounds__(256)
k_brick_nodes(const int32_t* __restrict__ perm, const int32_t* __restrict__ val_off,
              const int32_t* __restrict__ total, const int32_t* __restrict__ row_index,
              const int32_t* __restrict__ columns, int32_t nb_node, int32_t* __restrict__ table,
              int32_t* __restrict__ count)
{
  __shared__ int32_t buf[kLocalSort];
  __shared__ int32_t scan[256];
  const int tid = threadIdx.x;
  const int32_t blk = blockIdx.x;
  const int32_t cnt = total[blk];
  if (cnt < 0) {
    if (tid == 0)
      count[blk] = -1;
    return;
  }
  int n2 = 256;
  while (n2 < cnt)
    n2 <<= 1;
  for (int k = tid; k < n2; k += 256)
    buf[k] = 0x7fffffff;
  __syncthreads();
  const int64_t slot = (int64_t)blk * kLocalRows + tid;
  if (slot < nb_node) {
    const int32_t r = perm[slot];
    const int32_t o = val_off[slot];
    for (int32_t k = row_index[r]; k < row_index[r + 1]; ++k)
      buf[o + (k - row_index[r])] = columns[k];
  }
  __syncthreads();
  for (int size = 2; size <= n2; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int k = tid; k < n2 / 2; k += 256) {
        const int i = (k / stride) * 2 * stride + (k % stride);
        const int j = i + stride;
        const bool asc = (i & size) == 0;
        const int32_t x = buf[i], y = buf[j];
        if ((x > y) == asc) {
          buf[i] = y;
          buf[j] = x;
        }
      }
      __syncthreads();
    }
  const int chunk = n2 / 256;
  const int k0 = tid * chunk;
  int mine = 0;
  for (int k = k0; k < k0 + chunk; ++k)
    if (k < cnt && (k == 0 || buf[k] != buf[k - 1]))
      ++mine;
  scan[tid] = mine;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int v = (tid >= off) ? scan[tid - off] : 0;
    __syncthreads();
    scan[tid] += v;
    __syncthreads();
  }
  int pos = scan[tid] - mine;
  for (int k = k0; k < k0 + chunk; ++k)
    if (k < cnt && (k == 0 || buf[k] != buf[k - 1])) {
      if (pos < kLocalMaxNodes)
        table[(int64_t)blk * kLocalMaxNodes + pos] = buf[k];
      ++pos;
    }
  if (tid == 255)
    count[blk] = scan[255];
}

// Sliced local-plan words of each brick (table indices + row positions, as
// k_local_plan) and the cell ids, in pairs: words 2q and 2q + 1 of slot t at
// slice[blk] + q * 512 + 2 t (+ 1), so one 16-byte load per lane reads two.
__global__ void __launch_bounds__(256)
k_brick_plan(const int32_t* __restrict__ perm, const int32_t* __restrict__ nc_off, const uint64_t* __restrict__ plan,
             const int32_t* __restrict__ cells, int npc, int32_t nb_node, const int32_t* __restrict__ table,
             const int32_t* __restrict__ count, const int64_t* __restrict__ slice, uint64_t* __restrict__ out_plan,
             int32_t* __restrict__ out_cell, int32_t* __restrict__ err)
{
  __shared__ int32_t tab[kLocalMaxNodes];
  const int32_t blk = blockIdx.x;
  const int tn = count[blk];
  if (tn < 0 || tn > kLocalMaxNodes) {
    if (threadIdx.x == 0)
      atomicOr(err, 1);
    return;
  }
  for (int k = threadIdx.x; k < tn; k += 256)
    tab[k] = table[(int64_t)blk * kLocalMaxNodes + k];
  __syncthreads();
  const int64_t slot = (int64_t)blk * kLocalRows + threadIdx.x;
  if (slot >= nb_node)
    return;
  const int32_t i = perm[slot];
  const int64_t base = slice[blk] + 2 * threadIdx.x;
  bool bad = false;
  for (int32_t e = nc_off[i]; e < nc_off[i + 1]; ++e) {
    const uint64_t w = plan[e];
    const int32_t cell = plan_cell(w);
    const int4 cq = *reinterpret_cast<const int4*>(cells + 4 * (int64_t)cell);
    const int3 o = other_nodes(cq, plan_slot(w));
    const int32_t nodes[3] = { o.x, o.y, o.z };
    uint64_t lw = 0;
    for (int k = 0; k < npc - 1; ++k) {
      const int l = find_node(tab, tn, nodes[k]);
      const int pos = plan_pos(w, k);
      bad = bad || l >= tn || tab[l] != nodes[k] || pos >= kCompactMaxRow;
      lw |= (uint64_t)(l & 2047) << (11 * k);
      lw |= (uint64_t)(pos & 31) << (33 + 5 * k);
    }
    const int32_t j = e - nc_off[i];
    const int64_t at = base + (int64_t)(j >> 1) * (2 * kLocalRows) + (j & 1);
    out_plan[at] = lw;
    out_cell[at] = cell;
  }
  if (bad)
    atomicOr(err, 1);
}

// Per brick slot, what the brick assembly needs of its row, gathered once
// per pattern: the row id, its value offset, its length, diagonal position
// and cell count (packed), and the row node's index in the brick's node
// table.  err is set when a field does not fit its bits or the row node is
// missing from the table.
__global__ void __launch_bounds__(256)
k_brick_records(const int32_t* __restrict__ perm, const int32_t* __restrict__ row_index,
                const int32_t* __restrict__ diag_pos, const int32_t* __restrict__ nc_off,
                const int32_t* __restrict__ table, const int32_t* __restrict__ count, int32_t nb_node,
                int4* __restrict__ rec, int32_t* __restrict__ err)
{
  __shared__ int32_t tab[kLocalMaxNodes];
  const int32_t blk = blockIdx.x;
  const int tn = count[blk];
  if (tn < 0 || tn > kLocalMaxNodes) {
    if (threadIdx.x == 0)
      atomicOr(err, 1);
    return;
  }
  for (int k = threadIdx.x; k < tn; k += 256)
    tab[k] = table[(int64_t)blk * kLocalMaxNodes + k];
  __syncthreads();
  const int64_t slot = (int64_t)blk * kLocalRows + threadIdx.x;
  int4 r = make_int4(0, 0, 0, 0);
  if (slot < nb_node) {
    const int32_t i = perm[slot];
    const int32_t a = row_index[i], len = row_index[i + 1] - a;
    const int32_t d = diag_pos[i];
    const int32_t nc = nc_off[i + 1] - nc_off[i];
    const int li = find_node(tab, tn, i);
    if (li >= tn || tab[li] != i || len > 255 || d < 0 || d > 255 || nc < 0 || nc > 65535)
      atomicOr(err, 1);
    r = make_int4(i, a, len | (d << 8) | (nc << 16), li);
  }
  rec[slot] = r;
}

// Per brick, the coordinates of its node table (xy pairs, then z) at
// off[blk], so the assembly stages them with contiguous loads.
__global__ void __launch_bounds__(256)
k_brick_coords(const int32_t* __restrict__ table, const int32_t* __restrict__ count,
               const double* __restrict__ coords, const int64_t* __restrict__ off, double* __restrict__ out)
{
  const int32_t blk = blockIdx.x;
  const int tn = count[blk];
  double* base = out + off[blk];
  for (int k = threadIdx.x; k < tn; k += 256) {
    const int64_t g = table[(int64_t)blk * kLocalMaxNodes + k];
    reinterpret_cast<double2*>(base)[k] = *reinterpret_cast<const double2*>(coords + 4 * g);
    base[2 * tn + k] = coords[4 * g + 2];
  }
}

__host__ __device__ inline uint64_t morton_spread(uint64_t v) // 21 bits -> every third bit
{
  v &= 0x1fffffULL;
  v = (v | (v << 32)) & 0x1f00000000ffffULL;
  v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
  v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
  v = (v | (v << 2)) & 0x1249249249249249ULL;
  return v;
}

// Per-workgroup min / max of the coordinates (6 doubles per workgroup).
__global__ void __launch_bounds__(256)
k_bbox(const double* __restrict__ coords, int32_t n, double* __restrict__ out)
{
  __shared__ double red[6][4];
  double lo[3] = { 1e300, 1e300, 1e300 }, hi[3] = { -1e300, -1e300, -1e300 };
  for (int32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    for (int d = 0; d < 3; ++d) {
      const double x = coords[4 * (int64_t)i + d];
      lo[d] = fmin(lo[d], x);
      hi[d] = fmax(hi[d], x);
    }
  for (int d = 0; d < 3; ++d)
    for (int off = 32; off > 0; off >>= 1) {
      lo[d] = fmin(lo[d], __shfl_xor(lo[d], off, 64));
      hi[d] = fmax(hi[d], __shfl_xor(hi[d], off, 64));
    }
  if ((threadIdx.x & 63) == 0)
    for (int d = 0; d < 3; ++d) {
      red[d][threadIdx.x >> 6] = lo[d];
      red[3 + d][threadIdx.x >> 6] = hi[d];
    }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int q = threadIdx.x;
    double v = red[q][0];
    for (int w = 1; w < 4; ++w)
      v = q < 3 ? fmin(v, red[q][w]) : fmax(v, red[q][w]);
    out[6 * blockIdx.x + q] = v;
  }
}

// Morton code of each node (21 bits per axis over the bounding box).
__global__ void k_morton(const double* __restrict__ coords, int32_t n, double3 lo, double3 ext,
                         uint64_t* __restrict__ code, int32_t* __restrict__ node)
{
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const double x[3] = { coords[4 * (int64_t)i], coords[4 * (int64_t)i + 1], coords[4 * (int64_t)i + 2] };
  const double l[3] = { lo.x, lo.y, lo.z }, e[3] = { ext.x, ext.y, ext.z };
  uint64_t c = 0;
  for (int d = 0; d < 3; ++d) {
    const double t = e[d] > 0.0 ? (x[d] - l[d]) / e[d] : 0.0;
    const uint64_t q = (uint64_t)fmin(2097151.0, fmax(0.0, t * 2097151.0));
    c |= morton_spread(q) << d;
  }
  code[i] = c;
  node[i] = i;
}

// Sorts each brick's 256 rows into increasing order (bitonic, in LDS), so the
// row segments a wave writes at the end lie in runs of consecutive rows.
__global__ void __launch_bounds__(256)
k_sort_brick_rows(int32_t* __restrict__ perm, int32_t n)
{
  __shared__ int32_t buf[kLocalRows];
  const int64_t slot = (int64_t)blockIdx.x * kLocalRows + threadIdx.x;
  buf[threadIdx.x] = slot < n ? perm[slot] : 0x7fffffff;
  __syncthreads();
  for (int size = 2; size <= kLocalRows; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int t = threadIdx.x;
      const int partner = t ^ stride;
      const int32_t mine = buf[t], other = buf[partner];
      __syncthreads();
      const bool asc = (t & size) == 0;
      const bool lower = t < partner;
      buf[t] = (lower == asc) ? min(mine, other) : max(mine, other);
      __syncthreads();
    }
  if (slot < n)
    perm[slot] = buf[threadIdx.x];
}

// Slice width of each block of 256 rows (rows perm[blk * 256 + t], or
// blk * 256 + t without a permutation): its longest node->cell list, times 256.
__global__ void __launch_bounds__(256)
k_slice_width(const int32_t* __restrict__ perm, const int32_t* __restrict__ nc_off, int32_t n, int even,
              int64_t* __restrict__ width)
{
  __shared__ int32_t red[4];
  const int64_t slot = (int64_t)blockIdx.x * kLocalRows + threadIdx.x;
  int32_t w = 0;
  if (slot < n) {
    const int32_t r = perm ? perm[slot] : (int32_t)slot;
    w = nc_off[r + 1] - nc_off[r];
  }
  for (int off = 32; off > 0; off >>= 1)
    w = max(w, __shfl_xor(w, off, 64));
  if ((threadIdx.x & 63) == 0)
    red[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t wmax = max(max(red[0], red[1]), max(red[2], red[3]));
    if (even)
      wmax = (wmax + 1) & ~1;
    width[blockIdx.x] = (int64_t)wmax * kLocalRows;
  }
}

// slice[0] = 0, slice[k + 1] = slice[k] + width[k] (one workgroup; nblk is
// the number of 256-row blocks, a few thousand to tens of thousands)
__global__ void __launch_bounds__(1024)
k_slice_scan(const int64_t* __restrict__ width, int32_t nblk, int64_t* __restrict__ slice)
{
  __shared__ int64_t carry;
  __shared__ int64_t wsum[16];
  if (threadIdx.x == 0) {
    carry = 0;
    slice[0] = 0;
  }
  __syncthreads();
  for (int32_t base = 0; base < nblk; base += 1024) {
    const int32_t k = base + threadIdx.x;
    int64_t v = k < nblk ? width[k] : 0;
    // inclusive scan within the wave
    for (int off = 1; off < 64; off <<= 1) {
      const int64_t u = __shfl_up(v, off, 64);
      if ((threadIdx.x & 63) >= off)
        v += u;
    }
    if ((threadIdx.x & 63) == 63)
      wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    int64_t before = carry;
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w)
      before += wsum[w];
    if (k < nblk)
      slice[k + 1] = before + v;
    __syncthreads();
    if (threadIdx.x == 1023)
      carry = before + v;
    __syncthreads();
  }
}

// Slices of the blocks of 256 rows on the device (widths rounded up to an
// even count when `even`); returns the total words.
static int64_t device_slices(const int32_t* perm, const afem_mesh& m, int32_t nblk, DBuf<int64_t>& slice,
                             bool even = false)
{
  DBuf<int64_t> width((size_t)nblk);
  slice.alloc((size_t)nblk + 1);
  hipLaunchKernelGGL(k_slice_width, dim3(nblk), dim3(kLocalRows), 0, stream(), perm, m.nc_offsets.p, m.nb_node,
                     even ? 1 : 0, width.p);
  check_launch("slice_width");
  hipLaunchKernelGGL(k_slice_scan, dim3(1), dim3(1024), 0, stream(), width.p, nblk, slice.p);
  check_launch("slice_scan");
  int64_t total = 0;
  AFEM_HIP(hipMemcpyAsync(&total, slice.p + nblk, sizeof(int64_t), hipMemcpyDeviceToHost, stream()));
  AFEM_HIP(hipStreamSynchronize(stream()));
  return total;
}

static void build_brick_plan(afem_bsr* b)
{
  const afem_mesh& m = *b->mesh;
  auto release_all = [b]() {
    b->bperm.release();
    b->bval_off.release();
    b->btotal.release();
    b->btable.release();
    b->bcount.release();
    b->bslice.release();
    b->bplan.release();
    b->bcell.release();
    b->brec.release();
    b->bxyz.release();
    b->bxyz_off.release();
    b->brick_val_max = 0;
    b->brick_nodes_max = 0;
  };
  release_all();
  if (!b->plan.p || b->max_row > kCompactMaxRow)
    return;
  const int32_t n = m.nb_node;
  // Morton order of the node coordinates (21 bits per axis in the bounding
  // box), sorted on the device
  const int nbb = std::min<int>(1024, (int)grid_for(n, 256));
  DBuf<double> bb((size_t)6 * nbb);
  hipLaunchKernelGGL(k_bbox, dim3(nbb), dim3(256), 0, stream(), m.coords.p, n, bb.p);
  check_launch("bbox");
  std::vector<double> hb((size_t)6 * nbb);
  bb.download(hb.data(), hb.size());
  double lo[3] = { 1e300, 1e300, 1e300 }, hi[3] = { -1e300, -1e300, -1e300 };
  for (int k = 0; k < nbb; ++k)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], hb[6 * k + d]);
      hi[d] = std::max(hi[d], hb[6 * k + 3 + d]);
    }
  DBuf<uint64_t> code((size_t)n), code2((size_t)n);
  DBuf<int32_t> node((size_t)n);
  b->bperm.alloc((size_t)n);
  hipLaunchKernelGGL(k_morton, dim3(grid_for(n, 256)), dim3(256), 0, stream(), m.coords.p, n,
                     make_double3(lo[0], lo[1], lo[2]), make_double3(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]),
                     code.p, node.p);
  check_launch("morton");
  size_t tmp_bytes = 0;
  AFEM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, code.p, code2.p, node.p, b->bperm.p, n, 0, 63,
                                              stream()));
  DBuf<unsigned char> tmp(std::max<size_t>(tmp_bytes, 1));
  AFEM_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, code.p, code2.p, node.p, b->bperm.p, n, 0, 63,
                                              stream()));
  const int32_t nblk = (n + kLocalRows - 1) / kLocalRows;
  hipLaunchKernelGGL(k_sort_brick_rows, dim3(nblk), dim3(kLocalRows), 0, stream(), b->bperm.p, n);
  check_launch("sort_brick_rows");
  b->bval_off.alloc(n);
  b->btotal.alloc(nblk);
  hipLaunchKernelGGL(k_brick_offsets, dim3(nblk), dim3(256), 0, stream(), b->bperm.p, b->row_index.p, n,
                     b->bval_off.p, b->btotal.p);
  check_launch("brick_offsets");
  b->btable.alloc((size_t)nblk * kLocalMaxNodes);
  b->bcount.alloc(nblk);
  hipLaunchKernelGGL(k_brick_nodes, dim3(nblk), dim3(256), 0, stream(), b->bperm.p, b->bval_off.p, b->btotal.p,
                     b->row_index.p, b->columns.p, n, b->btable.p, b->bcount.p);
  check_launch("brick_nodes");
  std::vector<int32_t> totals((size_t)nblk), counts((size_t)nblk);
  b->btotal.download(totals.data(), totals.size());
  b->bcount.download(counts.data(), counts.size());
  int32_t vmax = 0, cmax = 0;
  for (int32_t k = 0; k < nblk; ++k) {
    if (totals[k] < 0 || counts[k] < 0 || counts[k] > kLocalMaxNodes) {
      release_all();
      return;
    }
    vmax = std::max(vmax, totals[k]);
    cmax = std::max(cmax, counts[k]);
  }
  // slices: width = the brick's longest node->cell list
  const int64_t words = device_slices(b->bperm.p, m, nblk, b->bslice, true);
  b->bplan.alloc((size_t)words);
  b->bcell.alloc((size_t)words);
  DBuf<int32_t> err(1);
  err.zero();
  hipLaunchKernelGGL(k_brick_plan, dim3(nblk), dim3(256), 0, stream(), b->bperm.p, m.nc_offsets.p, b->plan.p,
                     m.cells.p, m.npc, n, b->btable.p, b->bcount.p, b->bslice.p, b->bplan.p, b->bcell.p, err.p);
  check_launch("brick_plan");
  // per-slot row records and per-brick coordinate blocks (even offsets, so
  // every block starts 16-byte aligned)
  std::vector<int64_t> xoff((size_t)nblk + 1, 0);
  for (int32_t k = 0; k < nblk; ++k)
    xoff[k + 1] = xoff[k] + ((3 * (int64_t)counts[k] + 1) & ~(int64_t)1);
  b->bxyz_off.alloc(xoff.size());
  b->bxyz_off.upload(xoff.data(), xoff.size());
  b->bxyz.alloc((size_t)std::max<int64_t>(xoff[nblk], 2));
  b->brec.alloc((size_t)nblk * kLocalRows);
  hipLaunchKernelGGL(k_brick_records, dim3(nblk), dim3(256), 0, stream(), b->bperm.p, b->row_index.p, b->diag_pos.p,
                     m.nc_offsets.p, b->btable.p, b->bcount.p, n, b->brec.p, err.p);
  check_launch("brick_records");
  hipLaunchKernelGGL(k_brick_coords, dim3(nblk), dim3(256), 0, stream(), b->btable.p, b->bcount.p, m.coords.p,
                     b->bxyz_off.p, b->bxyz.p);
  check_launch("brick_coords");
  int32_t h = 0;
  err.download(&h, 1);
  if (h) {
    release_all();
    return;
  }
  b->brick_val_max = vmax;
  b->brick_nodes_max = cmax;
}

// Sliced copy of the local plan and of the cell ids: word j of row r of
// block blk at slice[blk] + j * 256 + (r - r0).
__global__ void __launch_bounds__(256)
k_slice_plan(const int32_t* __restrict__ nc_off, const int32_t* __restrict__ nc_items,
             const uint64_t* __restrict__ lplan, const int64_t* __restrict__ slice, int32_t nb_node,
             uint64_t* __restrict__ out_plan, int32_t* __restrict__ out_cell)
{
  const int32_t blk = blockIdx.x;
  const int32_t i = blk * kLocalRows + threadIdx.x;
  if (i >= nb_node)
    return;
  const int32_t e0 = nc_off[i], n = nc_off[i + 1] - e0;
  const int64_t base = slice[blk] + threadIdx.x;
  for (int32_t j = 0; j < n; ++j) {
    out_plan[base + (int64_t)j * kLocalRows] = lplan[e0 + j];
    out_cell[base + (int64_t)j * kLocalRows] = nc_items[e0 + j] >> 2;
  }
}

static void build_sliced_plan(afem_bsr* b)
{
  const afem_mesh& m = *b->mesh;
  b->lplan_sl.release();
  b->lcell_sl.release();
  b->slice.release();
  if (!b->lplan.p)
    return;
  const int32_t nblk = (m.nb_node + kLocalRows - 1) / kLocalRows;
  const int64_t words = device_slices(nullptr, m, nblk, b->slice);
  b->lplan_sl.alloc((size_t)words);
  b->lcell_sl.alloc((size_t)words);
  hipLaunchKernelGGL(k_slice_plan, dim3(nblk), dim3(kLocalRows), 0, stream(), m.nc_offsets.p, m.nc_items.p,
                     b->lplan.p, b->slice.p, m.nb_node, b->lplan_sl.p, b->lcell_sl.p);
  check_launch("slice_plan");
}

// Laplace assembly with the block's node coordinates staged in LDS (SoA):
// per (row, cell) pair one word from HBM, then only LDS reads.  The words
// (and cell ids) are stored per block of rows with word j of row r at
// slice[blk] + j * 256 + (r - r0), so one wave load is one contiguous
// 512-byte read.
template <int NPC, int U>
__global__ void __launch_bounds__(256)
k_assemble_laplace_lplan(const double* __restrict__ coords, const int32_t* __restrict__ nc_off,
                         const int32_t* __restrict__ lcell, const uint64_t* __restrict__ lplan,
                         const int64_t* __restrict__ slice, const int32_t* __restrict__ table,
                         const int32_t* __restrict__ count, const int32_t* __restrict__ diag_pos,
                         const int32_t* __restrict__ row_index, int32_t nb_node, int val_cap, int node_cap,
                         const double* __restrict__ cell_lambda, double source_f, double* __restrict__ rhs,
                         double* __restrict__ values)
{
  extern __shared__ unsigned char smem[];
  const int32_t blk = (int32_t)xcd_tile(blockIdx.x, gridDim.x);
  const int32_t r0 = blk * kLocalRows;
  const int32_t r1 = min(r0 + kLocalRows, nb_node);
  const int32_t a = row_index[r0];
  const int32_t cnt = row_index[r1] - a;
  const int tn = count[blk];
  double* s_val = reinterpret_cast<double*>(smem);
  double2* s_xy = reinterpret_cast<double2*>(s_val + val_cap); // val_cap is even
  double* s_z = reinterpret_cast<double*>(s_xy + node_cap);
  const int32_t* tab = table + (int64_t)blk * kLocalMaxNodes;
  for (int k = threadIdx.x; k < tn; k += blockDim.x) {
    const int64_t g = tab[k];
    s_xy[k] = *reinterpret_cast<const double2*>(coords + 4 * g);
    s_z[k] = coords[4 * g + 2];
  }
  for (int k = threadIdx.x; k < cnt; k += blockDim.x)
    s_val[k] = 0.0;
  __syncthreads();
  const int32_t i = r0 + threadIdx.x;
  if (i < r1) {
    double* rv = s_val + (row_index[i] - a);
    const d3 p = load_node(coords, i);
    const int32_t e0 = nc_off[i];
    const int32_t ncell = nc_off[i + 1] - e0;
    // word j of this row (and its cell id for the conductivity)
    const int64_t wbase = slice[blk] + threadIdx.x;
    double diag = 0.0, det_sum = 0.0;
    row_pairs<NPC, U, false>(lplan, lcell, wbase, kLocalRows, ncell, cell_lambda, s_xy, s_z, p, rv, 1, diag,
                             det_sum);
    rv[diag_pos[i]] += diag;
    if (rhs)
      rhs[i] = source_f * det_sum / (NPC == 4 ? 24.0 : 6.0);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < cnt; k += blockDim.x)
    values[a + k] = s_val[k];
}

// Words 2q, 2q + 1 of a paired slice (and their conductivities; the second
// word of a row's last pair may be padding: its cell is not read).
__device__ __forceinline__ void load_word_pair(const uint64_t* __restrict__ plan, const int32_t* __restrict__ cell_of,
                                               const double* __restrict__ cell_lambda, int64_t wbase, int32_t j,
                                               int32_t ncell, uint64_t& w0, uint64_t& w1, double& l0, double& l1)
{
  const int64_t at = wbase + (int64_t)(j >> 1) * (2 * kLocalRows);
  const ulonglong2 v = *reinterpret_cast<const ulonglong2*>(plan + at);
  w0 = v.x;
  w1 = v.y;
  l0 = 1.0;
  l1 = 1.0;
  if (cell_lambda) {
    const int2 c = *reinterpret_cast<const int2*>(cell_of + at);
    l0 = cell_lambda[c.x];
    if (j + 1 < ncell)
      l1 = cell_lambda[c.y];
  }
}

// All the (row, cell) pairs of one row from a paired brick slice (as
// row_pairs: U words in flight, the coordinates of pair u + 1 read from LDS
// before pair u is computed).
template <int NPC, int U>
__device__ __forceinline__ void brick_pairs(const uint64_t* __restrict__ plan, const int32_t* __restrict__ cell_of,
                                            int64_t wbase, int32_t ncell, const double* __restrict__ cell_lambda,
                                            const double2* __restrict__ s_xy, const double* __restrict__ s_z, d3 p,
                                            double* rv, double& diag, double& det_sum)
{
  static_assert(U % 2 == 0, "words are read in pairs");
  int32_t e = 0;
  for (; e + U <= ncell; e += U) {
    uint64_t w[U];
    double lam[U];
#pragma unroll
    for (int u = 0; u < U; u += 2)
      load_word_pair(plan, cell_of, cell_lambda, wbase, e + u, ncell, w[u], w[u + 1], lam[u], lam[u + 1]);
    PairNodes<NPC> next = pair_nodes<NPC>(s_xy, s_z, w[0]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const PairNodes<NPC> cur = next;
      if (u + 1 < U)
        next = pair_nodes<NPC>(s_xy, s_z, w[u + 1]);
      pair_element<NPC>(cur, p, lam[u], w[u], rv, 64, diag, det_sum);
    }
  }
  if (e < ncell) {
    uint64_t w[U];
    double lam[U];
    const int32_t last = (ncell - 1) & ~1; // the row's last pair
#pragma unroll
    for (int u = 0; u < U; u += 2)
      load_word_pair(plan, cell_of, cell_lambda, wbase, min(e + u, last), ncell, w[u], w[u + 1], lam[u], lam[u + 1]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (e + u >= ncell)
        break;
      pair_element<NPC>(pair_nodes<NPC>(s_xy, s_z, w[u]), p, lam[u], w[u], rv, 64, diag, det_sum);
    }
  }
}

// Brick-ordered Laplace assembly: thread t of brick blk owns the row of
// slot blk * 256 + t and writes its segment of the value array at the end;
// otherwise as k_assemble_laplace_lplan (SLICED).  The row accumulators are
// interleaved in LDS -- slot k of row t at (t / 64) * 64 * max_row + k * 64 +
// t % 64 -- so the 64 lanes of a wave always hit 64 consecutive doubles
// whatever their column positions (no bank conflicts).  Everything a row
// needs besides its plan words comes from one 16-byte slot record, and the
// brick's node coordinates from one contiguous block (both gathered once per
// pattern; the mesh's coordinates do not change), so no load depends on
// another before the pair loop.
template <int NPC, int U>
__global__ void __launch_bounds__(256)
k_assemble_laplace_brick(const int4* __restrict__ brec, const double* __restrict__ bxyz,
                         const int64_t* __restrict__ bxyz_off, const int32_t* __restrict__ count,
                         const uint64_t* __restrict__ bplan, const int32_t* __restrict__ bcell,
                         const int64_t* __restrict__ slice, int32_t nb_node, int val_cap,
                         const double* __restrict__ cell_lambda, double source_f, double* __restrict__ rhs,
                         double* __restrict__ values)
{
  extern __shared__ unsigned char smem[];
  const int32_t blk = (int32_t)xcd_tile(blockIdx.x, gridDim.x);
  const int max_row = val_cap / kLocalRows;
  const int tn = count[blk];
  double* s_val = reinterpret_cast<double*>(smem);
  double* s_coord = s_val + val_cap; // val_cap is even: 16-byte aligned
  const double2* s_xy = reinterpret_cast<const double2*>(s_coord);
  const double* s_z = s_coord + 2 * tn;
  const int64_t slot = (int64_t)blk * kLocalRows + threadIdx.x;
  const int4 rec = brec[slot];
  {
    const double2* src = reinterpret_cast<const double2*>(bxyz + bxyz_off[blk]);
    double2* dst = reinterpret_cast<double2*>(s_coord);
    const int n2 = (3 * tn + 1) >> 1; // the block is padded to an even count
    for (int k = threadIdx.x; k < n2; k += blockDim.x)
      dst[k] = src[k];
  }
  for (int k = threadIdx.x; k < val_cap; k += blockDim.x)
    s_val[k] = 0.0;
  __syncthreads();
  if (slot < nb_node) {
    const int32_t i = rec.x, a = rec.y;
    const int len = rec.z & 255, dpos = (rec.z >> 8) & 255;
    const int32_t ncell = (int32_t)((uint32_t)rec.z >> 16);
    double* rv = s_val + (threadIdx.x >> 6) * (64 * max_row) + (threadIdx.x & 63); // slot k at rv[64 k]
    const d3 p = { s_xy[rec.w].x, s_xy[rec.w].y, s_z[rec.w] };
    const int64_t wbase = slice[blk] + 2 * threadIdx.x;
    double diag = 0.0, det_sum = 0.0;
    brick_pairs<NPC, U>(bplan, bcell, wbase, ncell, cell_lambda, s_xy, s_z, p, rv, diag, det_sum);
    rv[64 * dpos] += diag;
    if (rhs)
      rhs[i] = source_f * det_sum / (NPC == 4 ? 24.0 : 6.0);
    // only this thread writes into rv, so its row is complete here; pairs
    // of values go out as 16-byte stores (half the store instructions)
    {
      int k = 0;
      if (a & 1) {
        values[a] = rv[0];
        k = 1;
      }
      for (; k + 1 < len; k += 2)
        *reinterpret_cast<double2*>(values + a + k) = make_double2(rv[64 * k], rv[64 * (k + 1)]);
      if (k < len)
        values[a + k] = rv[64 * k];
    }
  }
}

// ---- cell-once brick plan ---------------------------------------------------------------
// Every cell with a vertex among a brick's rows is listed once for that brick
// (an instance), so the assembly builds each cell's element matrix once per
// brick instead of once per (row, cell) pair.  Instance word 0: the cell's
// nodes as indices into the brick's node table (11 bits each) and, at bit 44,
// the mask of its vertices that are rows of the brick; word 1: for each owned
// vertex v, the row positions of the cell's other nodes, in cell order
// (5 bits each, at bit 5 * ((npc - 1) v + k)).  A brick's instances are ordered
// by mask, so a wave's lanes mostly share one mask and take the same
// branches, and within a mask by a bijective hash of the cell id, so that a
// wave's lanes are scattered cells (neighbouring cells share vertices, and
// their LDS additions into the same rows would collide).  Per brick also the
// table index -> brick slot map (255: not a row of the brick).
constexpr int kCellMaxInst = kLocalSort;

// Bijective scrambling of 32-bit cell ids (odd multiplier, then an xorshift
// by half the width, which is its own inverse) and its inverse.
__device__ __forceinline__ uint32_t cell_hash(uint32_t c)
{
  uint32_t x = c * 0x9E3779B1u;
  return x ^ (x >> 16);
}
__device__ __forceinline__ uint32_t cell_unhash(uint32_t x)
{
  x ^= x >> 16;
  return x * 0x0E8B2F51u; // the inverse of 0x9E3779B1 modulo 2^32
}

__global__ void k_invert_perm(const int32_t* __restrict__ perm, int32_t n, int32_t* __restrict__ inv)
{
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n)
    inv[perm[k]] = k;
}

// Block-wide exclusive scan of one value per thread (256 threads); returns
// the thread's offset, *total the sum.
__device__ __forceinline__ int32_t block_scan256(int32_t v, int32_t* scan, int32_t* total)
{
  const int tid = threadIdx.x;
  scan[tid] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int32_t u = tid >= off ? scan[tid - off] : 0;
    __syncthreads();
    scan[tid] += u;
    __syncthreads();
  }
  const int32_t incl = scan[tid];
  *total = scan[255];
  __syncthreads();
  return incl - v;
}

// WRITE = false: the number of distinct cells of each brick (ninst[blk], -1
// when the brick's (row, cell) pairs exceed kCellMaxInst); WRITE = true: the
// instance words at off[blk], the cell ids, and the slot map.
template <bool WRITE>
__global__ void __launch_bounds__(256)
k_cell_instances(const int32_t* __restrict__ perm, const int32_t* __restrict__ inv,
                 const int32_t* __restrict__ nc_off, const int32_t* __restrict__ nc_items,
                 const int32_t* __restrict__ cells, int npc, int32_t nb_node,
                 const int32_t* __restrict__ row_index, const int32_t* __restrict__ columns,
                 const int32_t* __restrict__ table, const int32_t* __restrict__ count,
                 const int64_t* __restrict__ off, int32_t* __restrict__ ninst, uint64_t* __restrict__ words,
                 int32_t* __restrict__ inst_cell, uint8_t* __restrict__ slot_map, int32_t* __restrict__ err)
{
  extern __shared__ uint32_t s_cells[]; // 2 * kCellMaxInst: the sorted (row, cell) pairs, the distinct cells
  uint32_t* buf = s_cells;               // hashed cell ids, then the owned masks
  uint32_t* uniq = s_cells + kCellMaxInst; // hashed cell ids
  __shared__ int32_t scan[256];
  const int tid = threadIdx.x;
  const int32_t blk = blockIdx.x;
  const int64_t slot = (int64_t)blk * kLocalRows + tid;
  int32_t r = -1, e0 = 0, nc = 0;
  if (slot < nb_node) {
    r = perm[slot];
    e0 = nc_off[r];
    nc = nc_off[r + 1] - e0;
  }
  int32_t total = 0;
  const int32_t start = block_scan256(nc, scan, &total);
  if (total > kCellMaxInst) {
    if (tid == 0) {
      if (WRITE)
        atomicOr(err, 1);
      else
        ninst[blk] = -1;
    }
    return;
  }
  int n2 = 256;
  while (n2 < total)
    n2 <<= 1;
  for (int k = tid; k < n2; k += 256)
    buf[k] = 0xffffffffu;
  __syncthreads();
  for (int k = 0; k < nc; ++k)
    buf[start + k] = cell_hash((uint32_t)(nc_items[e0 + k] >> 2));
  __syncthreads();
  for (int size = 2; size <= n2; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int k = tid; k < n2 / 2; k += 256) {
        const int i = (k / stride) * 2 * stride + (k % stride);
        const int j = i + stride;
        const bool asc = (i & size) == 0;
        const uint32_t x = buf[i], y = buf[j];
        if ((x > y) == asc) {
          buf[i] = y;
          buf[j] = x;
        }
      }
      __syncthreads();
    }
  // distinct cells, in hash order (the padding sorts last: only the first
  // `total` entries are read)
  const int chunk = n2 / 256;
  const int k0 = tid * chunk;
  int mine = 0;
  for (int k = k0; k < k0 + chunk; ++k)
    if (k < total && (k == 0 || buf[k] != buf[k - 1]))
      ++mine;
  int32_t nu = 0;
  int32_t pos = block_scan256(mine, scan, &nu);
  for (int k = k0; k < k0 + chunk; ++k)
    if (k < total && (k == 0 || buf[k] != buf[k - 1]))
      uniq[pos++] = buf[k];
  __syncthreads();
  if (!WRITE) {
    if (tid == 0)
      ninst[blk] = nu;
    return;
  }
  const int tn = count[blk];
  const int32_t* tab = table + (int64_t)blk * kLocalMaxNodes;
  for (int k = tid; k < tn; k += 256)
    slot_map[(int64_t)blk * kLocalMaxNodes + k] = 255;
  __syncthreads();
  if (slot < nb_node) {
    const int l = find_node(tab, tn, r);
    if (l < tn && tab[l] == r)
      slot_map[(int64_t)blk * kLocalMaxNodes + l] = (uint8_t)tid;
    else
      atomicOr(err, 1); // the row's own node is always among its columns
  }
  // owned-vertex mask of each distinct cell (buf reused)
  for (int k = tid; k < nu; k += 256) {
    const int4 cq = *reinterpret_cast<const int4*>(cells + 4 * (int64_t)cell_unhash(uniq[k]));
    const int32_t nodes[4] = { cq.x, cq.y, cq.z, cq.w };
    int m = 0;
    for (int v = 0; v < npc; ++v)
      if (inv[nodes[v]] / kLocalRows == blk)
        m |= 1 << v;
    buf[k] = (uint32_t)m;
  }
  __syncthreads();
  // order: masks with more owned vertices first, then by mask, then by hash
  const int uchunk = (nu + 255) / 256;
  const int u0 = tid * uchunk, u1 = min(nu, u0 + uchunk);
  int32_t base = 0;
  bool bad = false;
  for (int pc = npc; pc >= 1; --pc)
    for (int m = (1 << npc) - 1; m >= 1; --m) {
      if (__popc(m) != pc)
        continue;
      int cnt = 0;
      for (int k = u0; k < u1; ++k)
        cnt += buf[k] == (uint32_t)m;
      int32_t mtotal = 0;
      int32_t at = base + block_scan256(cnt, scan, &mtotal);
      for (int k = u0; k < u1; ++k) {
        if (buf[k] != (uint32_t)m)
          continue;
        const int32_t c = (int32_t)cell_unhash(uniq[k]);
        const int4 cq = *reinterpret_cast<const int4*>(cells + 4 * (int64_t)c);
        const int32_t nodes[4] = { cq.x, cq.y, cq.z, cq.w };
        uint64_t w0 = (uint64_t)m << 44, w1 = 0;
        for (int v = 0; v < npc; ++v) {
          const int l = find_node(tab, tn, nodes[v]);
          bad = bad || l >= tn || tab[l] != nodes[v];
          w0 |= (uint64_t)(l & 2047) << (11 * v);
          if (!((m >> v) & 1))
            continue;
          const int32_t* rc = columns + row_index[nodes[v]];
          const int len = row_index[nodes[v] + 1] - row_index[nodes[v]];
          int kk = 0;
          for (int wv = 0; wv < npc; ++wv) {
            if (wv == v)
              continue;
            const int p = find_col(rc, len, nodes[wv]);
            bad = bad || p >= len || rc[p] != nodes[wv] || p >= kCompactMaxRow;
            w1 |= (uint64_t)(p & 31) << (5 * ((npc - 1) * v + kk));
            ++kk;
          }
        }
        const int64_t g = off[blk] + at;
        words[2 * g] = w0;
        words[2 * g + 1] = w1;
        inst_cell[g] = c;
        ++at;
      }
      base += mtotal;
    }
  if (bad)
    atomicOr(err, 1);
}

static void build_cell_plan(afem_bsr* b)
{
  const afem_mesh& m = *b->mesh;
  auto release_all = [b]() {
    b->cinst_off.release();
    b->cwords.release();
    b->ccell.release();
    b->cmap.release();
    b->cell_inst_max = 0;
  };
  release_all();
  if (!b->bplan.p || !m.simplex || m.npc > 4 || b->max_row > kCompactMaxRow)
    return;
  const int32_t n = m.nb_node;
  const int32_t nblk = (n + kLocalRows - 1) / kLocalRows;
  DBuf<int32_t> inv((size_t)n), cnt((size_t)nblk), err(1);
  hipLaunchKernelGGL(k_invert_perm, dim3(grid_for(n, 256)), dim3(256), 0, stream(), b->bperm.p, n, inv.p);
  check_launch("invert_perm");
  const size_t lds = 2 * kCellMaxInst * sizeof(int32_t);
  hipLaunchKernelGGL(k_cell_instances<false>, dim3(nblk), dim3(256), lds, stream(), b->bperm.p, inv.p, m.nc_offsets.p,
                     m.nc_items.p, m.cells.p, m.npc, n, b->row_index.p, b->columns.p, b->btable.p, b->bcount.p,
                     (const int64_t*)nullptr, cnt.p, (uint64_t*)nullptr, (int32_t*)nullptr, (uint8_t*)nullptr, err.p);
  check_launch("cell_instances_count");
  std::vector<int32_t> c((size_t)nblk);
  cnt.download(c.data(), c.size());
  std::vector<int64_t> off((size_t)nblk + 1, 0);
  int32_t mx = 0;
  for (int32_t k = 0; k < nblk; ++k) {
    if (c[k] < 0)
      return;
    off[k + 1] = off[k] + c[k];
    mx = std::max(mx, c[k]);
  }
  b->cinst_off.alloc(off.size());
  b->cinst_off.upload(off.data(), off.size());
  b->cwords.alloc((size_t)std::max<int64_t>(2 * off[nblk], 2));
  b->ccell.alloc((size_t)std::max<int64_t>(off[nblk], 1));
  b->cmap.alloc((size_t)nblk * kLocalMaxNodes);
  err.zero();
  hipLaunchKernelGGL(k_cell_instances<true>, dim3(nblk), dim3(256), lds, stream(), b->bperm.p, inv.p, m.nc_offsets.p,
                     m.nc_items.p, m.cells.p, m.npc, n, b->row_index.p, b->columns.p, b->btable.p, b->bcount.p,
                     b->cinst_off.p, cnt.p, b->cwords.p, b->ccell.p, b->cmap.p, err.p);
  check_launch("cell_instances");
  int32_t h = 0;
  err.download(&h, 1);
  if (h) {
    release_all();
    return;
  }
  b->cell_inst_max = mx;
}

// a += v on an LDS double that other lanes may update too (ds_add_f64)
__device__ __forceinline__ void lds_atomic_add(double* a, double v)
{
  __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Cell-once Laplace assembly over the bricks: each thread takes instances of
// the brick's cells, builds the cell's gradients once (cofactors of the
// edges from its first vertex) and adds the off-diagonal entries of every
// owned vertex's element row into that row's LDS slots (interleaved as in
// k_assemble_laplace_brick) and |det| into the row's source accumulator.
// After the workgroup barrier each row thread sets its diagonal to minus the
// sum of its off-diagonal entries (every element row sums to zero) and writes
// its row.  The LDS additions of different waves land in any order, so the
// values are reproducible to rounding, not bit for bit.
template <int NPC>
__global__ void __launch_bounds__(256)
k_assemble_laplace_cell(const double* __restrict__ coords, const int32_t* __restrict__ perm,
                        const int64_t* __restrict__ inst_off, const uint64_t* __restrict__ words,
                        const int32_t* __restrict__ inst_cell, const uint8_t* __restrict__ slot_map,
                        const int32_t* __restrict__ table, const int32_t* __restrict__ count,
                        const int32_t* __restrict__ diag_pos, const int32_t* __restrict__ row_index,
                        int32_t nb_node, int max_row, int node_cap, const double* __restrict__ cell_lambda,
                        double source_f, double* __restrict__ rhs, double* __restrict__ values)
{
  extern __shared__ unsigned char smem[];
  const int32_t blk = (int32_t)xcd_tile(blockIdx.x, gridDim.x);
  const int val_cap = kLocalRows * max_row;
  const int tn = count[blk];
  double* s_val = reinterpret_cast<double*>(smem);
  double* s_src = s_val + val_cap;                              // kLocalRows
  double2* s_xy = reinterpret_cast<double2*>(s_src + kLocalRows); // node_cap
  double* s_z = reinterpret_cast<double*>(s_xy + node_cap);
  uint8_t* s_map = reinterpret_cast<uint8_t*>(s_z + node_cap);
  const int32_t* tab = table + (int64_t)blk * kLocalMaxNodes;
  for (int k = threadIdx.x; k < tn; k += blockDim.x) {
    const int64_t g = tab[k];
    s_xy[k] = *reinterpret_cast<const double2*>(coords + 4 * g);
    s_z[k] = coords[4 * g + 2];
    s_map[k] = slot_map[(int64_t)blk * kLocalMaxNodes + k];
  }
  for (int k = threadIdx.x; k < val_cap; k += blockDim.x)
    s_val[k] = 0.0;
  s_src[threadIdx.x] = 0.0;
  __syncthreads();
  const int64_t i1 = inst_off[blk + 1];
  for (int64_t i = inst_off[blk] + threadIdx.x; i < i1; i += kLocalRows) {
    const uint64_t w0 = words[2 * i], w1 = words[2 * i + 1];
    const double lam = cell_lambda ? cell_lambda[inst_cell[i]] : 1.0;
    const int mask = (int)((w0 >> 44) & 15);
    d3 x[NPC];
#pragma unroll
    for (int v = 0; v < NPC; ++v) {
      const int l = (int)((w0 >> (11 * v)) & 2047);
      const double2 xy = s_xy[l];
      x[v] = d3{ xy.x, xy.y, NPC == 4 ? s_z[l] : 0.0 };
    }
    double K[NPC][NPC];
    double adet;
    if constexpr (NPC == 4) {
      const d3 e1 = sub(x[1], x[0]), e2 = sub(x[2], x[0]), e3 = sub(x[3], x[0]);
      d3 C[4];
      C[1] = cross(e2, e3);
      C[2] = cross(e3, e1);
      C[3] = cross(e1, e2);
      adet = fabs(dot(e1, C[1]));
      C[0] = d3{ -(C[1].x + C[2].x + C[3].x), -(C[1].y + C[2].y + C[3].y), -(C[1].z + C[2].z + C[3].z) };
      const double scale = lam * fast_rcp(6.0 * adet);
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = a + 1; c < 4; ++c) {
          K[a][c] = scale * dot(C[a], C[c]);
          K[c][a] = K[a][c];
        }
    }
    else {
      const d3 e1 = sub(x[1], x[0]), e2 = sub(x[2], x[0]);
      const double det = e1.x * e2.y - e2.x * e1.y;
      adet = fabs(det);
      // rotated edges: the gradients times det
      const double g1x = e2.y, g1y = -e2.x, g2x = -e1.y, g2y = e1.x;
      const double g0x = -(g1x + g2x), g0y = -(g1y + g2y);
      const double gx[3] = { g0x, g1x, g2x }, gy[3] = { g0y, g1y, g2y };
      const double scale = lam * fast_rcp(2.0 * adet);
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = a + 1; c < 3; ++c) {
          K[a][c] = scale * (gx[a] * gx[c] + gy[a] * gy[c]);
          K[c][a] = K[a][c];
        }
    }
#pragma unroll
    for (int v = 0; v < NPC; ++v) {
      if (!((mask >> v) & 1))
        continue;
      const int l = (int)((w0 >> (11 * v)) & 2047);
      const int t = s_map[l];
      double* rv = s_val + (t >> 6) * (64 * max_row) + (t & 63);
      int kk = 0;
#pragma unroll
      for (int c = 0; c < NPC; ++c) {
        if (c == v)
          continue;
        const int pos = (int)((w1 >> (5 * ((NPC - 1) * v + kk))) & 31);
        lds_atomic_add(rv + 64 * pos, K[v][c]);
        ++kk;
      }
      lds_atomic_add(s_src + t, adet);
    }
  }
  __syncthreads();
  const int64_t slot = (int64_t)blk * kLocalRows + threadIdx.x;
  if (slot < nb_node) {
    const int32_t r = perm[slot];
    double* rv = s_val + (threadIdx.x >> 6) * (64 * max_row) + (threadIdx.x & 63);
    const int32_t a = row_index[r], len = row_index[r + 1] - a;
    const int d = diag_pos[r];
    double diag = 0.0;
    for (int k = 0; k < len; ++k)
      if (k != d)
        diag -= rv[64 * k];
    rv[64 * d] = diag;
    if (rhs)
      rhs[r] = source_f * s_src[threadIdx.x] / (NPC == 4 ? 24.0 : 6.0);
    for (int k = 0; k < len; ++k)
      values[a + k] = rv[64 * k];
  }
}

// ---------------------------------------------------------------------------
// Edge plan: for every off-diagonal block (p, q) of the pattern, the cells
// that contain both p and q, taken from the pair plan (the row positions of
// the other nodes of each (row, cell) pair).  The rows are packed into
// workgroups of at most 256 blocks; cell j of block t of workgroup w is at
// eslice[w] + 256 j + t.  Each row thread writes only its own row's blocks,
// so counting and filling need no atomics and keep the node->cell order.

__global__ void __launch_bounds__(256)
k_edge_count(const int32_t* __restrict__ nc_off, const uint64_t* __restrict__ plan,
             const int32_t* __restrict__ row_index, int npc, int32_t nb_node, int32_t* __restrict__ cnt)
{
  const int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nb_node)
    return;
  const int32_t r0 = row_index[p];
  for (int32_t e = nc_off[p]; e < nc_off[p + 1]; ++e) {
    const uint64_t w = plan[e];
    for (int k = 0; k < npc - 1; ++k)
      cnt[r0 + plan_pos(w, k)] += 1;
  }
}

// widest block of each workgroup; *bad = 1 when a block has more than 255 cells
__global__ void __launch_bounds__(256)
k_edge_width(const int32_t* __restrict__ wg_row, const int32_t* __restrict__ row_index,
             const int32_t* __restrict__ cnt, int32_t nwg, int32_t* __restrict__ width, int32_t* __restrict__ bad)
{
  const int32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwg)
    return;
  int32_t m = 0;
  for (int32_t k = row_index[wg_row[w]]; k < row_index[wg_row[w + 1]]; ++k)
    m = max(m, cnt[k]);
  width[w] = m;
  if (m > 255)
    atomicOr(bad, 1);
}

__global__ void __launch_bounds__(256)
k_edge_fill(const int32_t* __restrict__ nc_off, const uint64_t* __restrict__ plan,
            const int32_t* __restrict__ row_index, const int32_t* __restrict__ row_wg,
            const int32_t* __restrict__ wg_row, const int64_t* __restrict__ eslice, int npc, int32_t nb_node,
            int32_t* __restrict__ run, int32_t* __restrict__ ecell, uint8_t* __restrict__ ecnt)
{
  const int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nb_node)
    return;
  const int32_t w = row_wg[p];
  const int32_t b0 = row_index[wg_row[w]];
  const int32_t r0 = row_index[p];
  for (int32_t e = nc_off[p]; e < nc_off[p + 1]; ++e) {
    const uint64_t wd = plan[e];
    for (int k = 0; k < npc - 1; ++k) {
      const int32_t blk = r0 + plan_pos(wd, k);
      const int32_t j = run[blk]++;
      ecell[eslice[w] + 256 * (int64_t)j + (blk - b0)] = plan_cell(wd);
    }
  }
  for (int32_t k = r0; k < row_index[p + 1]; ++k)
    ecnt[k] = (uint8_t)run[k];
}

// Geometry plan of the elasticity kernel: for each edge workgroup the
// distinct cells of its rows (its "instances", sorted), and for each (block,
// cell) entry of the edge slices the instance index and the slots of the
// block's two nodes in that cell.
constexpr int kEdgeSortCap = 2048; // (row, cell) pairs of one workgroup the instance builder sorts
constexpr int kEdgeInstCap = 1023; // instances of one workgroup (10-bit index)

// Distinct cells of the rows of edge workgroup w, sorted in LDS (bitonic).
// FILL = false: their number into count[w]; FILL = true: the cells at
// inst + inst_off[w].  *bad = 1 when a workgroup has more pairs than the sort
// holds.
template <bool FILL>
__global__ void __launch_bounds__(256)
k_edge_instances(const int32_t* __restrict__ wg_row, const int32_t* __restrict__ nc_off,
                 const int32_t* __restrict__ nc_items, int slot_bits, const int32_t* __restrict__ inst_off,
                 int32_t* __restrict__ count, int32_t* __restrict__ inst, int32_t* __restrict__ bad)
{
  __shared__ int32_t s_key[kEdgeSortCap];
  __shared__ int32_t s_sum[256];
  const int32_t w = blockIdx.x;
  const int32_t e0 = nc_off[wg_row[w]];
  const int n = nc_off[wg_row[w + 1]] - e0;
  if (n > kEdgeSortCap) {
    if (threadIdx.x == 0)
      atomicOr(bad, 1);
    return;
  }
  int P = 1;
  while (P < n)
    P <<= 1;
  for (int k = threadIdx.x; k < P; k += 256)
    s_key[k] = k < n ? (nc_items[e0 + k] >> slot_bits) : 0x7fffffff;
  __syncthreads();
  for (int size = 2; size <= P; size <<= 1)
    for (int half = size >> 1; half > 0; half >>= 1) {
      for (int k = threadIdx.x; k < P / 2; k += 256) {
        const int i = 2 * half * (k / half) + (k % half), j = i + half;
        const int32_t a = s_key[i], c = s_key[j];
        if ((a > c) == ((i & size) == 0)) {
          s_key[i] = c;
          s_key[j] = a;
        }
      }
      __syncthreads();
    }
  // thread t owns the 8 sorted entries [8t, 8t + 8): count its first occurrences
  const int t = threadIdx.x;
  int own = 0;
  for (int k = 8 * t; k < 8 * t + 8 && k < n; ++k)
    own += (k == 0 || s_key[k] != s_key[k - 1]) ? 1 : 0;
  s_sum[t] = own;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int v = t >= off ? s_sum[t - off] : 0;
    __syncthreads();
    s_sum[t] += v;
    __syncthreads();
  }
  if (!FILL) {
    if (t == 255)
      count[w] = s_sum[255];
    return;
  }
  int at = inst_off[w] + s_sum[t] - own;
  for (int k = 8 * t; k < 8 * t + 8 && k < n; ++k)
    if (k == 0 || s_key[k] != s_key[k - 1])
      inst[at++] = s_key[k];
}

// Instance index and node slots of every (block, cell) entry of workgroup w,
// entry j of block t at eslice8[w] + (j / 8) * 2048 + 8 t + j % 8 (eight
// entries per 16-byte load).
__global__ void __launch_bounds__(256)
k_edge_encode(const int32_t* __restrict__ wg_row, const int32_t* __restrict__ row_index,
              const int32_t* __restrict__ columns, const int64_t* __restrict__ eslice,
              const int32_t* __restrict__ ecell, const uint8_t* __restrict__ ecnt, const int32_t* __restrict__ cells,
              const int32_t* __restrict__ inst_off, const int32_t* __restrict__ inst,
              const int64_t* __restrict__ eslice8, uint16_t* __restrict__ egeo, int32_t* __restrict__ bad)
{
  __shared__ int32_t s_inst[kEdgeInstCap];
  __shared__ int32_t s_rows[257];
  const int32_t w = blockIdx.x;
  const int32_t p0 = wg_row[w], nr = wg_row[w + 1] - p0;
  const int32_t b0 = row_index[p0];
  const int32_t i0 = inst_off[w], ni = inst_off[w + 1] - i0;
  if (ni > kEdgeInstCap) {
    if (threadIdx.x == 0)
      atomicOr(bad, 1);
    return;
  }
  for (int k = threadIdx.x; k < ni; k += 256)
    s_inst[k] = inst[i0 + k];
  for (int k = threadIdx.x; k <= nr; k += 256)
    s_rows[k] = row_index[p0 + k] - b0;
  __syncthreads();
  const int t = threadIdx.x;
  if (t >= s_rows[nr])
    return;
  int lo = 0, hi = nr - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (s_rows[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int32_t p = p0 + lo, q = columns[b0 + t];
  const int64_t base = eslice[w] + t;
  for (int j = 0; j < ecnt[b0 + t]; ++j) {
    const int32_t c = ecell[base + 256 * (int64_t)j];
    const int li = find_col(s_inst, ni, c);
    const int4 cq = *reinterpret_cast<const int4*>(cells + 4 * (int64_t)c);
    const int32_t v[4] = { cq.x, cq.y, cq.z, cq.w };
    int sp = -1, sq = -1;
    for (int k = 0; k < 4; ++k) {
      if (v[k] == p)
        sp = k;
      if (v[k] == q)
        sq = k;
    }
    if (li >= ni || s_inst[li] != c || sp < 0 || sq < 0) {
      atomicOr(bad, 2);
      continue;
    }
    egeo[eslice8[w] + (int64_t)(j >> 3) * (8 * 256) + 8 * t + (j & 7)] = (uint16_t)(li | (sp << 10) | (sq << 12));
  }
}

// The nodes of every instance (its cell's connectivity).
__global__ void k_edge_inst_nodes(const int32_t* __restrict__ inst, int64_t n, const int32_t* __restrict__ cells,
                                  int4* __restrict__ out)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n)
    out[k] = *reinterpret_cast<const int4*>(cells + 4 * (int64_t)inst[k]);
}

static void build_edge_geometry(afem_bsr* b)
{
  const afem_mesh& m = *b->mesh;
  b->einst_off.release();
  b->einst.release();
  b->einst4.release();
  b->egeo.release();
  b->eslice8.release();
  b->e_inst_max = 0;
  const int32_t nwg = b->e_nwg;
  if (nwg <= 0)
    return;
  DBuf<int32_t> cnt((size_t)nwg), bad(1);
  bad.zero();
  hipLaunchKernelGGL(k_edge_instances<false>, dim3(nwg), dim3(256), 0, stream(), b->ewg_row.p, m.nc_offsets.p,
                     m.nc_items.p, m.slot_bits, nullptr, cnt.p, nullptr, bad.p);
  check_launch("edge_instances");
  int32_t h = 0;
  bad.download(&h, 1);
  if (h)
    return;
  std::vector<int32_t> c((size_t)nwg);
  cnt.download(c.data(), c.size());
  std::vector<int32_t> off((size_t)nwg + 1, 0);
  int32_t mx = 0;
  for (int32_t w = 0; w < nwg; ++w) {
    off[w + 1] = off[w] + c[w];
    mx = std::max(mx, c[w]);
  }
  if (mx > kEdgeInstCap)
    return;
  b->einst_off.alloc(off.size());
  b->einst_off.upload(off.data(), off.size());
  b->einst.alloc((size_t)std::max(off[nwg], 1));
  hipLaunchKernelGGL(k_edge_instances<true>, dim3(nwg), dim3(256), 0, stream(), b->ewg_row.p, m.nc_offsets.p,
                     m.nc_items.p, m.slot_bits, b->einst_off.p, nullptr, b->einst.p, bad.p);
  check_launch("edge_instances_fill");
  b->einst4.alloc((size_t)std::max(off[nwg], 1));
  hipLaunchKernelGGL(k_edge_inst_nodes, dim3(grid_for(std::max(off[nwg], 1), 256)), dim3(256), 0, stream(),
                     b->einst.p, (int64_t)off[nwg], m.cells.p, b->einst4.p);
  check_launch("edge_inst_nodes");
  // packed entry slices: each workgroup's width (its widest block's cell
  // count) rounded up to a multiple of 8
  std::vector<int64_t> sl((size_t)nwg + 1), sl8((size_t)nwg + 1, 0);
  b->eslice.download(sl.data(), sl.size());
  for (int32_t w = 0; w < nwg; ++w) {
    const int64_t width = (sl[w + 1] - sl[w]) / 256;
    sl8[w + 1] = sl8[w] + 256 * ((width + 7) / 8 * 8);
  }
  b->eslice8.alloc(sl8.size());
  b->eslice8.upload(sl8.data(), sl8.size());
  b->egeo.alloc((size_t)std::max<int64_t>(sl8[nwg], 8));
  hipLaunchKernelGGL(k_edge_encode, dim3(nwg), dim3(256), 0, stream(), b->ewg_row.p, b->row_index.p, b->columns.p,
                     b->eslice.p, b->ecell.p, b->ecnt.p, m.cells.p, b->einst_off.p, b->einst.p, b->eslice8.p,
                     b->egeo.p, bad.p);
  check_launch("edge_encode");
  bad.download(&h, 1);
  if (h) {
    b->einst_off.release();
    b->einst.release();
    b->einst4.release();
    b->egeo.release();
    b->eslice8.release();
    return;
  }
  b->e_inst_max = mx;
}

static void build_edge_plan(afem_bsr* b)
{
  const afem_mesh& m = *b->mesh;
  auto release_all = [b]() {
    b->ewg_row.release();
    b->eslice.release();
    b->ecell.release();
    b->ecnt.release();
    b->e_nwg = 0;
  };
  release_all();
  if (!b->plan.p || b->nb_dof != 3 || m.npc != 4 || b->max_row > 256)
    return;
  const int32_t n = m.nb_node;
  std::vector<int32_t> rows((size_t)n + 1);
  b->row_index.download(rows.data(), rows.size());
  // whole rows, at most 256 blocks per workgroup
  std::vector<int32_t> wg_row(1, 0), row_wg((size_t)n);
  for (int32_t p = 0; p < n; ++p) {
    if (rows[p + 1] - rows[wg_row.back()] > 256)
      wg_row.push_back(p);
    row_wg[p] = (int32_t)wg_row.size() - 1;
  }
  wg_row.push_back(n);
  const int32_t nwg = (int32_t)wg_row.size() - 1;
  const int64_t nnz = rows[n];
  DBuf<int32_t> cnt((size_t)nnz), width((size_t)nwg), bad(1), d_row_wg((size_t)n);
  cnt.zero();
  bad.zero();
  b->ewg_row.alloc(wg_row.size());
  b->ewg_row.upload(wg_row.data(), wg_row.size());
  d_row_wg.upload(row_wg.data(), row_wg.size());
  hipLaunchKernelGGL(k_edge_count, dim3(grid_for(n, 256)), dim3(256), 0, stream(), m.nc_offsets.p, b->plan.p,
                     b->row_index.p, m.npc, n, cnt.p);
  check_launch("edge_count");
  hipLaunchKernelGGL(k_edge_width, dim3(grid_for(nwg, 256)), dim3(256), 0, stream(), b->ewg_row.p, b->row_index.p,
                     cnt.p, nwg, width.p, bad.p);
  check_launch("edge_width");
  int32_t h = 0;
  bad.download(&h, 1);
  if (h) {
    release_all();
    return;
  }
  std::vector<int32_t> wv((size_t)nwg);
  width.download(wv.data(), wv.size());
  std::vector<int64_t> slice((size_t)nwg + 1, 0);
  for (int32_t w = 0; w < nwg; ++w)
    slice[w + 1] = slice[w] + 256 * (int64_t)wv[w];
  b->eslice.alloc(slice.size());
  b->eslice.upload(slice.data(), slice.size());
  b->ecell.alloc((size_t)std::max<int64_t>(slice[nwg], 1));
  b->ecell.zero();
  b->ecnt.alloc((size_t)nnz);
  cnt.zero(); // running counters of the fill
  hipLaunchKernelGGL(k_edge_fill, dim3(grid_for(n, 256)), dim3(256), 0, stream(), m.nc_offsets.p, b->plan.p,
                     b->row_index.p, d_row_wg.p, b->ewg_row.p, b->eslice.p, m.npc, n, cnt.p, b->ecell.p, b->ecnt.p);
  check_launch("edge_fill");
  AFEM_HIP(hipStreamSynchronize(stream()));
  b->e_nwg = nwg;
  build_edge_geometry(b);
}

// Elasticity assembly, one thread per block (p, q) of a workgroup of whole
// rows: the 3x3 block lambda grad_p (x) grad_q + mu grad_q (x) grad_p +
// mu (grad_p . grad_q) I times the volume, summed over the cells of the edge
// (cofactors with the origin at p, as k_assemble_elasticity) in registers,
// and written once (72 contiguous bytes in per-block order).  The diagonal
// block is minus the sum of the row's other blocks: every row of the element
// matrix annihilates rigid translations (the shape-function gradients of a
// cell sum to zero), K_pp = -sum_{q != p} K_pq, summed in column order.
__global__ void __launch_bounds__(256)
k_assemble_elasticity_edge(const double* __restrict__ coords, const int32_t* __restrict__ cells,
                           const int32_t* __restrict__ row_index, const int32_t* __restrict__ columns,
                           const int32_t* __restrict__ wg_row, const int64_t* __restrict__ eslice,
                           const int32_t* __restrict__ ecell, const uint8_t* __restrict__ ecnt, double lame_lambda,
                           double lame_mu, int per_block, double* __restrict__ values)
{
  __shared__ int32_t s_rows[257];
  __shared__ double s_blk[9 * 256];
  const int32_t w = (int32_t)xcd_tile(blockIdx.x, gridDim.x);
  const int32_t p0 = wg_row[w], nr = wg_row[w + 1] - p0;
  const int32_t b0 = row_index[p0];
  for (int k = threadIdx.x; k <= nr; k += 256)
    s_rows[k] = row_index[p0 + k] - b0;
  __syncthreads();
  const int nb = s_rows[nr];
  const int t = threadIdx.x;
  double K[9];
#pragma unroll
  for (int i = 0; i < 9; ++i)
    K[i] = 0.0;
  int pr = 0;
  bool diag = false;
  if (t < nb) {
    int lo = 0, hi = nr - 1; // the row of block t: the last row starting at or before t
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_rows[mid] <= t)
        lo = mid;
      else
        hi = mid - 1;
    }
    pr = lo;
    const int32_t p = p0 + pr, q = columns[b0 + t];
    diag = p == q;
    const int cnum = ecnt[b0 + t];
    const int64_t base = eslice[w] + t;
    const d3 P = load_node(coords, p);
    const d3 eq = sub(load_node(coords, q), P);
    for (int j = 0; j < cnum; ++j) {
      const int32_t c = ecell[base + 256 * (int64_t)j];
      const int4 cq = *reinterpret_cast<const int4*>(cells + 4 * (int64_t)c);
      // the two vertices other than p and q, in cell order
      const int32_t v[4] = { cq.x, cq.y, cq.z, cq.w };
      int32_t o0 = -1, o1 = -1;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (v[k] != p && v[k] != q) {
          if (o0 < 0)
            o0 = v[k];
          else
            o1 = v[k];
        }
      if (o1 < 0)
        continue; // not a cell of this edge (the plan never lists one)
      const d3 ea = sub(load_node(coords, o0), P), eb = sub(load_node(coords, o1), P);
      const d3 Cq = cross(ea, eb), Ca = cross(eb, eq), Cb = cross(eq, ea);
      const double adet = fabs(dot(eq, Cq));
      const d3 Cp = { -(Cq.x + Ca.x + Cb.x), -(Cq.y + Ca.y + Cb.y), -(Cq.z + Ca.z + Cb.z) };
      const double inv6 = fast_rcp(6.0 * adet);
      const double la = lame_lambda * inv6, mu = lame_mu * inv6;
      const double av[3] = { Cp.x, Cp.y, Cp.z };
      const double bq[3] = { Cq.x, Cq.y, Cq.z };
      const double dd = mu * (av[0] * bq[0] + av[1] * bq[1] + av[2] * bq[2]);
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c2 = 0; c2 < 3; ++c2)
          K[r * 3 + c2] += la * av[r] * bq[c2] + mu * bq[r] * av[c2] + (r == c2 ? dd : 0.0);
    }
  }
#pragma unroll
  for (int i = 0; i < 9; ++i)
    s_blk[i * 256 + t] = K[i];
  __syncthreads();
  if (t >= nb)
    return;
  const int x0 = s_rows[pr], len = s_rows[pr + 1] - x0;
  if (diag) {
    for (int u = x0; u < x0 + len; ++u)
      if (u != t) {
#pragma unroll
        for (int i = 0; i < 9; ++i)
          K[i] -= s_blk[i * 256 + u];
      }
  }
  const int64_t blk = (int64_t)b0 + t;
  if (per_block) {
#pragma unroll
    for (int i = 0; i < 9; ++i)
      values[9 * blk + i] = K[i];
  }
  else {
    const int64_t base = 9 * ((int64_t)b0 + x0);
    const int x = t - x0;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int s = 0; s < 3; ++s)
        values[base + 3 * (x + (int64_t)r * len) + s] = K[r * 3 + s];
  }
}

// Elasticity assembly with the cell geometry computed once per workgroup:
// the workgroup's distinct cells (its instances, given by their four nodes)
// first get their cofactor vectors C_k = det grad N_k (origin at the cell's
// first node: C_1 = e2 x e3, C_2 = e3 x e1, C_3 = e1 x e2, C_0 = -(C_1 + C_2 +
// C_3)) and 1 / (6 |det|) staged in LDS, one thread per instance; then, as
// k_assemble_elasticity_edge, one thread per nonzero block (p, q) sums
// lambda C_p (x) C_q + mu C_q (x) C_p + mu (C_p . C_q) I, times 1 / (6 |det|),
// over the cells of its edge -- 7 LDS reads per (block, cell), the entries
// read eight at a time.  The diagonal block is minus the sum of the row's
// other blocks.  The LDS of the instances is reused for that sum and, in
// per-block order, for a block-major copy that the workgroup writes with
// 16-byte stores (its blocks are one contiguous range of the value array).
__global__ void __launch_bounds__(256)
k_assemble_elasticity_geo(const double* __restrict__ coords, const int4* __restrict__ inst_nodes,
                          const int32_t* __restrict__ row_index, const int32_t* __restrict__ columns,
                          const int32_t* __restrict__ wg_row, const int64_t* __restrict__ eslice8,
                          const uint16_t* __restrict__ egeo, const uint8_t* __restrict__ ecnt,
                          const int32_t* __restrict__ inst_off, int cap, double lame_lambda, double lame_mu,
                          int per_block, double* __restrict__ values)
{
  extern __shared__ double s_geo[]; // 13 * cap doubles (component-major), later 9 * 256
  __shared__ int32_t s_rows[257];
  const int32_t w = (int32_t)xcd_tile(blockIdx.x, gridDim.x);
  const int32_t p0 = wg_row[w], nr = wg_row[w + 1] - p0;
  const int32_t b0 = row_index[p0];
  const int32_t i0 = inst_off[w], ni = inst_off[w + 1] - i0;
  for (int k = threadIdx.x; k <= nr; k += 256)
    s_rows[k] = row_index[p0 + k] - b0;
  for (int k = threadIdx.x; k < ni; k += 256) {
    const int4 cq = inst_nodes[i0 + k];
    const d3 X0 = load_node(coords, cq.x);
    const d3 e1 = sub(load_node(coords, cq.y), X0), e2 = sub(load_node(coords, cq.z), X0),
             e3 = sub(load_node(coords, cq.w), X0);
    const d3 C1 = cross(e2, e3), C2 = cross(e3, e1), C3 = cross(e1, e2);
    const double inv6 = fast_rcp(6.0 * fabs(dot(e1, C1)));
    s_geo[0 * cap + k] = -(C1.x + C2.x + C3.x);
    s_geo[1 * cap + k] = -(C1.y + C2.y + C3.y);
    s_geo[2 * cap + k] = -(C1.z + C2.z + C3.z);
    s_geo[3 * cap + k] = C1.x;
    s_geo[4 * cap + k] = C1.y;
    s_geo[5 * cap + k] = C1.z;
    s_geo[6 * cap + k] = C2.x;
    s_geo[7 * cap + k] = C2.y;
    s_geo[8 * cap + k] = C2.z;
    s_geo[9 * cap + k] = C3.x;
    s_geo[10 * cap + k] = C3.y;
    s_geo[11 * cap + k] = C3.z;
    s_geo[12 * cap + k] = inv6;
  }
  __syncthreads();
  const int nb = s_rows[nr];
  const int t = threadIdx.x;
  double K[9];
#pragma unroll
  for (int i = 0; i < 9; ++i)
    K[i] = 0.0;
  int pr = 0;
  bool diag = false;
  if (t < nb) {
    int lo = 0, hi = nr - 1; // the row of block t: the last row starting at or before t
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_rows[mid] <= t)
        lo = mid;
      else
        hi = mid - 1;
    }
    pr = lo;
    diag = (p0 + pr) == columns[b0 + t];
    const int cnum = ecnt[b0 + t];
    const int64_t base = eslice8[w] + 8 * t;
    for (int j0 = 0; j0 < cnum; j0 += 8) {
      const uint4 g = *reinterpret_cast<const uint4*>(egeo + base + (int64_t)(j0 >> 3) * (8 * 256));
      const uint32_t codes[8] = { g.x & 0xffffu, g.x >> 16, g.y & 0xffffu, g.y >> 16,
                                  g.z & 0xffffu, g.z >> 16, g.w & 0xffffu, g.w >> 16 };
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (j0 + u >= cnum)
          break;
        const uint32_t code = codes[u];
        const int li = code & 1023, sp = (code >> 10) & 3, sq = (code >> 12) & 3;
        const double av[3] = { s_geo[(3 * sp) * cap + li], s_geo[(3 * sp + 1) * cap + li],
                               s_geo[(3 * sp + 2) * cap + li] };
        const double bq[3] = { s_geo[(3 * sq) * cap + li], s_geo[(3 * sq + 1) * cap + li],
                               s_geo[(3 * sq + 2) * cap + li] };
        const double inv6 = s_geo[12 * cap + li];
        const double la = lame_lambda * inv6, mu = lame_mu * inv6;
        const double dd = mu * (av[0] * bq[0] + av[1] * bq[1] + av[2] * bq[2]);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c2 = 0; c2 < 3; ++c2)
            K[r * 3 + c2] += la * av[r] * bq[c2] + mu * bq[r] * av[c2] + (r == c2 ? dd : 0.0);
      }
    }
  }
  __syncthreads(); // the instances are no longer read: the LDS holds the blocks now
  double* s_blk = s_geo;
#pragma unroll
  for (int i = 0; i < 9; ++i)
    s_blk[i * 256 + t] = K[i];
  __syncthreads();
  int x0 = 0, len = 0;
  if (t < nb) {
    x0 = s_rows[pr];
    len = s_rows[pr + 1] - x0;
    if (diag) {
      for (int u = x0; u < x0 + len; ++u)
        if (u != t) {
#pragma unroll
          for (int i = 0; i < 9; ++i)
            K[i] -= s_blk[i * 256 + u];
        }
    }
  }
  if (per_block) {
    // block-major copy in LDS, then the workgroup's 9 nb values leave as
    // 16-byte stores (one 8-byte store first when the range starts odd)
    __syncthreads();
    if (t < nb) {
#pragma unroll
      for (int i = 0; i < 9; ++i)
        s_blk[9 * t + i] = K[i];
    }
    __syncthreads();
    const int64_t start = 9 * (int64_t)b0;
    const int n = 9 * nb;
    const int head = (int)(start & 1);
    if (head && t == 0)
      values[start] = s_blk[0];
    const int pairs = (n - head) >> 1;
    for (int k = t; k < pairs; k += 256)
      *reinterpret_cast<double2*>(values + start + head + 2 * k) =
          make_double2(s_blk[head + 2 * k], s_blk[head + 2 * k + 1]);
    if (((n - head) & 1) && t == 0)
      values[start + n - 1] = s_blk[n - 1];
  }
  else if (t < nb) {
    const int64_t base = 9 * ((int64_t)b0 + x0);
    const int x = t - x0;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int s2 = 0; s2 < 3; ++s2)
        values[base + 3 * (x + (int64_t)r * len) + s2] = K[r * 3 + s2];
  }
}

// ---- Neumann condition on boundary faces ---------------------------------------------
//
// Per face (2 nodes in 2D, 3 in 3D): its measure, its outward unit normal
// (the face normal from its node order, turned away from the vertex of its
// cell that is not on the face) and the value each of its nodes receives,
// g |F| / k or (g . n) |F| / k, with the operation order of
// femutils/ArcaneFemFunctions.h:2715-2759 (edges) and :1793-1840 (triangles).
// *err = 1 when a face is not a face of any cell.
__global__ void __launch_bounds__(256)
k_neumann_faces(int dim, int32_t nface, const int32_t* __restrict__ faces, const double* __restrict__ coords,
                const int32_t* __restrict__ cells, const int32_t* __restrict__ nc_off,
                const int32_t* __restrict__ nc_items, int scalar, double g, double gx, double gy, double gz,
                double* __restrict__ contrib, int32_t* __restrict__ err)
{
  const int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nface)
    return;
  const int k = dim; // nodes per face
  const int32_t* fn = faces + (int64_t)k * f;
  int32_t opp = -1;
  for (int32_t e = nc_off[fn[0]]; e < nc_off[fn[0] + 1] && opp < 0; ++e) {
    const int4 cq = *reinterpret_cast<const int4*>(cells + 4 * (int64_t)(nc_items[e] >> 2));
    const int32_t v[4] = { cq.x, cq.y, cq.z, cq.w };
    int hits = 0;
    for (int a = 1; a < k; ++a)
      for (int s = 0; s < k + 1; ++s)
        hits += v[s] == fn[a];
    if (hits != k - 1)
      continue;
    for (int s = 0; s < k + 1; ++s) {
      bool on_face = false;
      for (int a = 0; a < k; ++a)
        on_face = on_face || v[s] == fn[a];
      if (!on_face)
        opp = v[s];
    }
  }
  if (opp < 0) {
    for (int a = 0; a < k; ++a)
      contrib[(int64_t)k * f + a] = 0.0;
    atomicOr(err, 1);
    return;
  }
  const d3 x0 = load_node(coords, fn[0]), x1 = load_node(coords, fn[1]), xo = load_node(coords, opp);
  double val;
  if (dim == 2) {
    const double dx = x1.x - x0.x, dy = x1.y - x0.y;
    const double length = sqrt(dx * dx + dy * dy);
    double nx = dy / length, ny = -dx / length;
    if (nx * (xo.x - x0.x) + ny * (xo.y - x0.y) > 0.0) { // towards the cell: turn outward
      nx = -nx;
      ny = -ny;
    }
    val = scalar ? g * length / 2.0 : (nx * gx + ny * gy) * length / 2.0;
  }
  else {
    const d3 x2 = load_node(coords, fn[2]);
    d3 c = cross(sub(x1, x0), sub(x2, x0));
    double norm = sqrt(c.x * c.x + c.y * c.y + c.z * c.z);
    const double area = norm / 2.0;
    if (c.x * (xo.x - x0.x) + c.y * (xo.y - x0.y) + c.z * (xo.z - x0.z) > 0.0) {
      // towards the cell: the normal of the face with its first two nodes swapped
      c = cross(sub(x0, x1), sub(x2, x1));
      norm = sqrt(c.x * c.x + c.y * c.y + c.z * c.z);
    }
    const double nx = c.x / norm, ny = c.y / norm, nz = c.z / norm;
    val = scalar ? g * area / 3.0 : (nx * gx + ny * gy + nz * gz) * area / 3.0;
  }
  for (int a = 0; a < k; ++a) // every node of the face receives the same value
    contrib[(int64_t)k * f + a] = val;
}

// rhs[node * nb_dof + dof_index] += the values of the node's faces, in face order
__global__ void __launch_bounds__(256)
k_neumann_gather(int32_t nnode, const int32_t* __restrict__ node_ids, const int32_t* __restrict__ off,
                 const int32_t* __restrict__ pos_of, const double* __restrict__ contrib, int nb_dof, int dof_index,
                 double* __restrict__ rhs)
{
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnode)
    return;
  double* r = rhs + (int64_t)node_ids[t] * nb_dof + dof_index;
  double acc = *r;
  for (int32_t j = off[t]; j < off[t + 1]; ++j)
    acc += contrib[pos_of[j]];
  *r = acc;
}

// Per (face, slot) of a Quad4 mesh (edges) or a Hexa8 mesh (quadrangles, 4
// nodes): the Gauss-point sum of femutils/ArcaneFemFunctionsGpu.cc:741-850
// (2 points on the edge, outward normal as computeNormalFace,
// ArcaneFemFunctionsGpu.h:158-178) and :1146-1320 (2x2 points; the normal is
// t1 x t2 of the face's node order, not turned outward, as the reference
// does).  *err = 1 when a face is not a face of any cell.
__global__ void __launch_bounds__(256)
k_neumann_faces_q1(int dim, int32_t nface, const int32_t* __restrict__ faces, const double* __restrict__ coords,
                   const int32_t* __restrict__ cells, int npc, int slot_bits, const int32_t* __restrict__ nc_off,
                   const int32_t* __restrict__ nc_items, int scalar, double g, double gx, double gy, double gz,
                   double* __restrict__ contrib, int32_t* __restrict__ err)
{
  const int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nface)
    return;
  const int k = dim == 2 ? 2 : 4;
  const int32_t* fn = faces + (int64_t)k * f;
  int32_t inner = -1; // a node of the face's cell that is not on the face
  for (int32_t e = nc_off[fn[0]]; e < nc_off[fn[0] + 1] && inner < 0; ++e) {
    const int32_t it = nc_items[e];
    const int32_t* cn = cells + (it - (it & ((1 << slot_bits) - 1)));
    int hits = 0;
    for (int a = 1; a < k; ++a)
      for (int s = 0; s < npc; ++s)
        hits += cn[s] == fn[a];
    if (hits != k - 1)
      continue;
    for (int s = 0; s < npc && inner < 0; ++s) {
      bool on_face = false;
      for (int a = 0; a < k; ++a)
        on_face = on_face || cn[s] == fn[a];
      if (!on_face)
        inner = cn[s];
    }
  }
  double* out = contrib + (int64_t)k * f;
  for (int a = 0; a < k; ++a)
    out[a] = 0.0;
  if (inner < 0) {
    atomicOr(err, 1);
    return;
  }
  constexpr double gp[2] = { -0.57735026918962576451, 0.57735026918962576451 };
  if (dim == 2) {
    const d3 a = load_node(coords, fn[0]), b = load_node(coords, fn[1]), xi = load_node(coords, inner);
    const double length = sqrt((b.x - a.x) * (b.x - a.x) + (b.y - a.y) * (b.y - a.y));
    double nx = (b.y - a.y) / length, ny = (a.x - b.x) / length;
    if (nx * (xi.x - a.x) + ny * (xi.y - a.y) > 0.0) { // towards the cell: turn outward
      nx = -nx;
      ny = -ny;
    }
    const double v = scalar ? g : nx * gx + ny * gy;
    for (int i = 0; i < 2; ++i) {
      const double iw = 1.0 * length * 0.5;
      out[0] += v * (0.5 * (1.0 - gp[i])) * iw;
      out[1] += v * (0.5 * (1.0 + gp[i])) * iw;
    }
    return;
  }
  d3 x[4];
  for (int a = 0; a < 4; ++a)
    x[a] = load_node(coords, fn[a]);
  for (int ix = 0; ix < 2; ++ix)
    for (int iy = 0; iy < 2; ++iy) {
      const double xi = gp[ix], eta = gp[iy];
      const double N[4] = { 0.25 * (1.0 - xi) * (1.0 - eta), 0.25 * (1.0 + xi) * (1.0 - eta),
                            0.25 * (1.0 + xi) * (1.0 + eta), 0.25 * (1.0 - xi) * (1.0 + eta) };
      const double dxi[4] = { -0.25 * (1.0 - eta), 0.25 * (1.0 - eta), 0.25 * (1.0 + eta), -0.25 * (1.0 + eta) };
      const double deta[4] = { -0.25 * (1.0 - xi), -0.25 * (1.0 + xi), 0.25 * (1.0 + xi), 0.25 * (1.0 - xi) };
      d3 t1 = { 0.0, 0.0, 0.0 }, t2 = { 0.0, 0.0, 0.0 };
      for (int a = 0; a < 4; ++a) {
        t1.x += dxi[a] * x[a].x;
        t1.y += dxi[a] * x[a].y;
        t1.z += dxi[a] * x[a].z;
        t2.x += deta[a] * x[a].x;
        t2.y += deta[a] * x[a].y;
        t2.z += deta[a] * x[a].z;
      }
      const d3 n = cross(t1, t2);
      const double det = sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
      const double v = scalar ? g : (n.x / det) * gx + (n.y / det) * gy + (n.z / det) * gz;
      for (int a = 0; a < 4; ++a)
        out[a] += v * N[a] * (1.0 * 1.0 * det);
    }
}

// |F| of each face of k nodes: computeLengthEdge2 (femutils/ArcaneFemFunctions.h:
// 375-381), computeAreaTria3 (:180-188) or computeAreaQuad4 (:202-212: the
// triangles n2 n1 n0 and n0 n3 n2).
__global__ void __launch_bounds__(256)
k_face_measure(int k, int32_t nface, const int32_t* __restrict__ faces, const double* __restrict__ coords,
               double* __restrict__ measure)
{
  const int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nface)
    return;
  const int32_t* fn = faces + (int64_t)k * f;
  const d3 x0 = load_node(coords, fn[0]), x1 = load_node(coords, fn[1]);
  if (k == 2) {
    const d3 e1 = sub(x1, x0);
    measure[f] = sqrt(e1.x * e1.x + e1.y * e1.y + e1.z * e1.z);
  }
  else if (k == 3) {
    const d3 c = cross(sub(x1, x0), sub(load_node(coords, fn[2]), x0));
    measure[f] = sqrt(c.x * c.x + c.y * c.y + c.z * c.z) / 2.0;
  }
  else {
    const d3 x2 = load_node(coords, fn[2]), x3 = load_node(coords, fn[3]);
    const d3 c = cross(sub(x2, x1), sub(x0, x1)), d = cross(sub(x0, x3), sub(x2, x3));
    measure[f] = 0.5 * (sqrt(c.x * c.x + c.y * c.y + c.z * c.z) + sqrt(d.x * d.x + d.y * d.y + d.z * d.z));
  }
}

// rhs[node * nb_dof + dof_index] += sum over the node's faces (in face order)
// of g |F| / k.
__global__ void __launch_bounds__(256)
k_face_uniform_gather(int32_t nnode, const int32_t* __restrict__ node_ids, const int32_t* __restrict__ off,
                      const int32_t* __restrict__ pos_of, int k, const double* __restrict__ measure, double g,
                      int nb_dof, int dof_index, double* __restrict__ rhs)
{
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnode)
    return;
  double* r = rhs + (int64_t)node_ids[t] * nb_dof + dof_index;
  double acc = *r;
  for (int32_t j = off[t]; j < off[t + 1]; ++j)
    acc += g * measure[pos_of[j] / k] / k;
  *r = acc;
}

// values += hf * m_ab * |F| for the pairs (node, b) of every face of the node,
// in face order (m = 2 on the diagonal, 1 off it; hf = h / 6 or h / 12): the
// convection matrix of modules/heat/FemModule.cc:302-331.  With check_only
// the values are not touched and *err = 1 when a pair is missing from the
// pattern.
__global__ void __launch_bounds__(256)
k_convection_gather(int32_t nnode, const int32_t* __restrict__ node_ids, const int32_t* __restrict__ off,
                    const int32_t* __restrict__ pos_of, const int32_t* __restrict__ faces, int k,
                    const double* __restrict__ measure, double hf, const int32_t* __restrict__ row_index,
                    const int32_t* __restrict__ columns, double* __restrict__ values, int check_only,
                    int32_t* __restrict__ err)
{
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnode)
    return;
  const int32_t i = node_ids[t];
  const int32_t a = row_index[i], len = row_index[i + 1] - a;
  const int32_t* rc = columns + a;
  for (int32_t j = off[t]; j < off[t + 1]; ++j) {
    const int32_t f = pos_of[j] / k;
    const int32_t* fn = faces + (int64_t)k * f;
    for (int b = 0; b < k; ++b) {
      const int pos = find_col(rc, len, fn[b]);
      const bool found = pos < len && rc[pos] == fn[b];
      if (check_only) {
        if (!found)
          atomicOr(err, 1);
      }
      else if (found)
        values[a + pos] += hf * (fn[b] == i ? 2.0 : 1.0) * measure[f];
    }
  }
}

// ---- 2D elasticity (tri3, 2 DoFs per node) ----------------------------------------------

// One thread per node i: rows 2i and 2i+1.  For each cell of the node (slot
// s) and each of its nodes b, the 2x2 block of modules/elasticity/
// ElementMatrix.h:41-58 written out per entry:
//   K00 = (l + 2m) dx_s dx_b A + m dy_s dy_b A,  K01 = l dx_s dy_b A + m dy_s dx_b A,
//   K10 = l dy_s dx_b A + m dx_s dy_b A,         K11 = (l + 2m) dy_s dy_b A + m dx_s dx_b A,
// with the gradients of femutils/ArcaneFemFunctionsGpu.h:241-281 and the area
// of :76-86.  Values in per-block or per-row (scalar CSR) order; the thread
// owns its rows, so there are no atomics.
__global__ void __launch_bounds__(256)
k_assemble_elasticity_tri(const double* __restrict__ coords, const int32_t* __restrict__ cells,
                          const int32_t* __restrict__ nc_off, const int32_t* __restrict__ nc_items, int32_t nb_node,
                          const int32_t* __restrict__ row_index, const int32_t* __restrict__ columns, int per_block,
                          double lam, double mu, double* __restrict__ values, int32_t* __restrict__ err)
{
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb_node)
    return;
  const int32_t a0 = row_index[i], len = row_index[i + 1] - a0;
  const int32_t* rc = columns + a0;
  double* rv = values + (int64_t)a0 * 4;
  for (int k = 0; k < 4 * len; ++k)
    rv[k] = 0.0;
  for (int32_t e = nc_off[i]; e < nc_off[i + 1]; ++e) {
    const int32_t it = nc_items[e];
    const int s = it & 3;
    const int4 cq = *reinterpret_cast<const int4*>(cells + (it - s));
    const int32_t v[3] = { cq.x, cq.y, cq.z };
    const d3 n0 = load_node(coords, v[0]), n1 = load_node(coords, v[1]), n2 = load_node(coords, v[2]);
    const double A2 = (n1.x - n0.x) * (n2.y - n0.y) - (n2.x - n0.x) * (n1.y - n0.y);
    const double dx[3] = { (n1.y - n2.y) / A2, (n2.y - n0.y) / A2, (n0.y - n1.y) / A2 };
    const double dy[3] = { (n2.x - n1.x) / A2, (n0.x - n2.x) / A2, (n1.x - n0.x) / A2 };
    const d3 c = cross(sub(n1, n0), sub(n2, n0));
    const double area = sqrt(c.x * c.x + c.y * c.y + c.z * c.z) / 2.0;
    const double dxs = s == 0 ? dx[0] : s == 1 ? dx[1] : dx[2];
    const double dys = s == 0 ? dy[0] : s == 1 ? dy[1] : dy[2];
    for (int b = 0; b < 3; ++b) {
      const int pos = find_col(rc, len, v[b]);
      if (!(pos < len && rc[pos] == v[b])) {
        atomicOr(err, 1);
        continue;
      }
      const double k00 = (lam + 2 * mu) * dxs * dx[b] * area + mu * dys * dy[b] * area;
      const double k01 = lam * dxs * dy[b] * area + mu * dys * dx[b] * area;
      const double k10 = lam * dys * dx[b] * area + mu * dxs * dy[b] * area;
      const double k11 = (lam + 2 * mu) * dys * dy[b] * area + mu * dxs * dx[b] * area;
      if (per_block) {
        double* blk = rv + 4 * pos;
        blk[0] += k00;
        blk[1] += k01;
        blk[2] += k10;
        blk[3] += k11;
      }
      else { // scalar row 2i holds len * 2 entries, then row 2i + 1
        rv[2 * pos] += k00;
        rv[2 * pos + 1] += k01;
        rv[2 * len + 2 * pos] += k10;
        rv[2 * len + 2 * pos + 1] += k11;
      }
    }
  }
}

// ---- heat: mass terms -----------------------------------------------------------------

// |T| of cell c: |cross(n1 - n0, n2 - n0)| / 2 (tri) or |det| / 6 (tet), as
// MeshOperation::computeAreaTria3 / computeVolumeTetra4.
__device__ __forceinline__ double cell_measure(const double* __restrict__ coords, int4 cq, int npc)
{
  const d3 x0 = load_node(coords, cq.x);
  const d3 e1 = sub(load_node(coords, cq.y), x0), e2 = sub(load_node(coords, cq.z), x0);
  const d3 c = cross(e1, e2);
  if (npc == 3)
    return sqrt(c.x * c.x + c.y * c.y + c.z * c.z) / 2.0;
  return fabs(dot(sub(load_node(coords, cq.w), x0), c)) / 6.0;
}

// values += M / dt, the consistent mass matrix of the heat element
// (modules/heat/ElementMatrix.h:28-39, 80-91: (1/12 or 1/20) * massMatrix * |T| / dt,
// massMatrix 2 on the diagonal and 1 elsewhere), one thread per row.
__global__ void __launch_bounds__(256)
k_add_mass(const double* __restrict__ coords, const int32_t* __restrict__ cells, int npc,
           const int32_t* __restrict__ nc_off, const int32_t* __restrict__ nc_items,
           const int32_t* __restrict__ row_index, const int32_t* __restrict__ columns, int32_t nb_node, double dt,
           double* __restrict__ values)
{
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb_node)
    return;
  const int32_t a = row_index[i], len = row_index[i + 1] - a;
  const int32_t* rc = columns + a;
  const double w = npc == 3 ? 1.0 / 12.0 : 1.0 / 20.0;
  for (int32_t e = nc_off[i]; e < nc_off[i + 1]; ++e) {
    const int4 cq = *reinterpret_cast<const int4*>(cells + 4 * (int64_t)(nc_items[e] >> 2));
    const double m = cell_measure(coords, cq, npc);
    const int32_t v[4] = { cq.x, cq.y, cq.z, cq.w };
    for (int k = 0; k < npc; ++k) {
      const int pos = find_col(rc, len, v[k]);
      if (pos < len && rc[pos] == v[k])
        values[a + pos] += w * (v[k] == i ? 2.0 : 1.0) * m / dt;
    }
  }
}

// rhs[i*nb_dof+dof_index] += (M (scale * field))_i, gathered per node over its
// cells in node->cell order (femutils/ArcaneFemFunctions.h:2558-2590, 1628-1660:
// per cell sum_j c_ij f_j, c = |T|/6 | |T|/12 (tri), |T|/10 | |T|/20 (tet)).
__global__ void __launch_bounds__(256)
k_rhs_mass_field(const double* __restrict__ coords, const int32_t* __restrict__ cells, int npc,
                 const int32_t* __restrict__ nc_off, const int32_t* __restrict__ nc_items, int32_t nb_node,
                 const double* __restrict__ field, double scale, int nb_dof, int dof_index, double* __restrict__ rhs)
{
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb_node)
    return;
  double sum = 0.0;
  for (int32_t e = nc_off[i]; e < nc_off[i + 1]; ++e) {
    const int4 cq = *reinterpret_cast<const int4*>(cells + 4 * (int64_t)(nc_items[e] >> 2));
    const int s = nc_items[e] & 3;
    const double m = cell_measure(coords, cq, npc);
    const double cd = npc == 3 ? m / 6.0 : m / 10.0, co = npc == 3 ? m / 12.0 : m / 20.0;
    const int32_t v[4] = { cq.x, cq.y, cq.z, cq.w };
    double acc = 0.0;
    for (int j = 0; j < npc; ++j)
      acc += (j == s ? cd : co) * (field[v[j]] * scale);
    sum += acc;
  }
  rhs[(int64_t)i * nb_dof + dof_index] += sum;
}

static void build_plan(afem_bsr* b)
{
  const afem_mesh& m = *b->mesh;
  b->plan.release();
  b->diag_pos.release();
  b->q1_plan.release(); // rebuilt by the first Q1 assembly
  b->q1_plan_sl.release();
  b->q1_item_sl.release();
  if (b->max_row > kPlanMaxRow || !m.simplex) // the plans serve the simplex kernels
    return;
  b->plan.alloc(m.nc_items.n);
  b->diag_pos.alloc((size_t)m.nb_node);
  DBuf<int32_t> err(1);
  err.zero();
  hipLaunchKernelGGL(k_pair_plan, dim3(grid_for(m.nb_node, 256)), dim3(256), 0, stream(), m.nc_offsets.p,
                     m.nc_items.p, m.cells.p, m.npc, m.nb_node, b->row_index.p, b->columns.p, b->plan.p,
                     b->diag_pos.p, err.p);
  check_launch("pair_plan");
  int32_t h = 0;
  err.download(&h, 1);
  AFEM_REQUIRE(h == 0, AFEM_ERR_NOT_FOUND, "BSRMatrix: a cell coupling is missing from the pattern");
  b->plans_tried = 0; // the derived plans are built by the first assembly that uses them
}

// Builds, once per pattern, the derived plan a kernel kind needs (kinds that
// cannot be built leave their arrays empty, and the variant falls back).
static void ensure_plans(afem_bsr* b, unsigned what)
{
  const unsigned todo = what & ~b->plans_tried;
  if (!todo || !b->plan.p)
    return;
  b->plans_tried |= todo;
  if (todo & kPlansSliced) {
    ProfScope prof("plan_sliced");
    build_local_plan(b);
    build_sliced_plan(b);
    b->lplan.release(); // only the sliced copy is read by the kernels
  }
  if (todo & kPlansBrick) {
    ProfScope prof("plan_brick");
    build_brick_plan(b);
  }
  if (todo & kPlansEdge) {
    ProfScope prof("plan_edge");
    build_edge_plan(b);
  }
  if (todo & kPlansCell) {
    ProfScope prof("plan_cell");
    build_cell_plan(b); // on the brick plan (ensured first by the caller)
  }
}

// Laplace assembly driven by the plan: only the values are staged in LDS.
template <int NPC, int U>
__global__ void __launch_bounds__(256)
k_assemble_laplace_plan(const double* __restrict__ coords, const int32_t* __restrict__ cells,
                        const int32_t* __restrict__ nc_off, const uint64_t* __restrict__ plan,
                        const int32_t* __restrict__ diag_pos, const int32_t* __restrict__ row_index,
                        int32_t nb_node, int rows_per_block, const double* __restrict__ cell_lambda,
                        double source_f, double* __restrict__ rhs, double* __restrict__ values)
{
  extern __shared__ unsigned char smem[];
  const int32_t tile = (int32_t)xcd_tile(blockIdx.x, gridDim.x);
  const int32_t r0 = tile * rows_per_block;
  const int32_t r1 = min(r0 + rows_per_block, nb_node);
  const int32_t a = row_index[r0];
  const int32_t cnt = row_index[r1] - a;
  double* s_val = reinterpret_cast<double*>(smem);
  for (int k = threadIdx.x; k < cnt; k += blockDim.x)
    s_val[k] = 0.0;
  __syncthreads();
  const int32_t i = r0 + threadIdx.x;
  if (threadIdx.x < rows_per_block && i < r1) {
    double* rv = s_val + (row_index[i] - a);
    const d3 p = load_node(coords, i);
    const int32_t e0 = nc_off[i], e1 = nc_off[i + 1];
    double diag = 0.0, b_acc = 0.0;
    for (int32_t e = e0; e < e1; e += U) {
      uint64_t w[U];
      int4 cq[U];
      double lam[U];
      d3 pa[U], pb[U], pc[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        w[u] = plan[min(e + u, e1 - 1)];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t c = plan_cell(w[u]);
        cq[u] = *reinterpret_cast<const int4*>(cells + 4 * (int64_t)c);
        lam[u] = cell_lambda ? cell_lambda[c] : 1.0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int3 o = other_nodes(cq[u], plan_slot(w[u]));
        pa[u] = load_node(coords, o.x);
        pb[u] = load_node(coords, o.y);
        if constexpr (NPC == 4)
          pc[u] = load_node(coords, o.z);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (e + u >= e1)
          break;
        if constexpr (NPC == 4) {
          const d3 ea = sub(pa[u], p), eb = sub(pb[u], p), ec = sub(pc[u], p);
          const d3 Ca = cross(eb, ec), Cb = cross(ec, ea), Cc = cross(ea, eb);
          const double adet = fabs(dot(ea, Ca));
          const d3 Cp = { -(Ca.x + Cb.x + Cc.x), -(Ca.y + Cb.y + Cc.y), -(Ca.z + Cb.z + Cc.z) };
          const double scale = lam[u] * fast_rcp(6.0 * adet);
          diag += scale * dot(Cp, Cp);
          rv[plan_pos(w[u], 0)] += scale * dot(Cp, Ca);
          rv[plan_pos(w[u], 1)] += scale * dot(Cp, Cb);
          rv[plan_pos(w[u], 2)] += scale * dot(Cp, Cc);
          b_acc += adet; // b_p = f * sum |det| / 24
        }
        else {
          const d3 ea = sub(pa[u], p), eb = sub(pb[u], p);
          const double adet = fabs(ea.x * eb.y - eb.x * ea.y);
          const double gax = eb.y, gay = -eb.x, gbx = -ea.y, gby = ea.x;
          const double gpx = -(gax + gbx), gpy = -(gay + gby);
          const double scale = lam[u] * fast_rcp(2.0 * adet);
          diag += scale * (gpx * gpx + gpy * gpy);
          rv[plan_pos(w[u], 0)] += scale * (gpx * gax + gpy * gay);
          rv[plan_pos(w[u], 1)] += scale * (gpx * gbx + gpy * gby);
          b_acc += adet; // b_p = f * sum |det| / 6
        }
      }
    }
    rv[diag_pos[i]] += diag;
    if (rhs)
      rhs[i] = source_f * b_acc / (NPC == 4 ? 24.0 : 6.0);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < cnt; k += blockDim.x)
    values[a + k] = s_val[k];
}

// Rows per block and LDS bytes for an assembly launch; the block's values
// range is bounded by rows_per_block * max_row.
static bool assembly_shape(const afem_bsr& b, int bytes_per_entry, int max_rows, int* rows_per_block, size_t* lds)
{
  int rows = max_rows;
  const size_t lds_budget = 64 * 1024;
  while (rows > 16 && (size_t)rows * b.max_row * bytes_per_entry > lds_budget)
    rows >>= 1;
  size_t bytes = (size_t)rows * b.max_row * bytes_per_entry;
  *rows_per_block = rows;
  *lds = bytes;
  return bytes <= 160 * 1024;
}

// Laplace assembly variants: the searching kernel (no plan), the pair-plan
// kernel (rows of up to 255 entries), the sliced block-local plan and the
// brick plan (rows of up to 32 entries, node tables of up to 2048 nodes),
// each with U plan words in flight per thread.
enum PlanKind
{
  kSearch = 0,
  kPlan = 1,
  kLocalSliced = 2,
  kBrick = 3,
  kCell = 4
};
struct LaplaceVariant
{
  int kind;
  int u;
  int rows;
};
static const LaplaceVariant kLaplaceVariants[] = {
  { kSearch, 4, 256 },       { kPlan, 8, 256 },   { kLocalSliced, 8, 256 },
  { kLocalSliced, 16, 256 }, { kBrick, 8, 256 },  { kBrick, 16, 256 },
  { kCell, 1, 256 },
};
constexpr int kNumLaplaceVariants = sizeof(kLaplaceVariants) / sizeof(kLaplaceVariants[0]);
constexpr int kDefaultLaplaceVariant = 5;

template <int NPC, int U>
static void launch_laplace_lplan(const afem_bsr* b, const double* lam, double f, double* rhs)
{
  const afem_mesh& m = *b->mesh;
  const int val_cap = (kLocalRows * b->max_row + 1) & ~1; // keeps the coordinate table 16-byte aligned
  const int node_cap = b->block_nodes_max;
  const size_t lds = ((size_t)val_cap + 3 * (size_t)node_cap) * sizeof(double);
  hipLaunchKernelGGL((k_assemble_laplace_lplan<NPC, U>), dim3(grid_for(m.nb_node, kLocalRows)), dim3(kLocalRows),
                     lds, stream(), m.coords.p, m.nc_offsets.p, b->lcell_sl.p, b->lplan_sl.p, b->slice.p,
                     b->block_nodes.p, b->block_count.p, b->diag_pos.p, b->row_index.p, m.nb_node, val_cap, node_cap,
                     lam, f, rhs, b->values.p);
}

template <int NPC, int U>
static void launch_laplace_brick(const afem_bsr* b, const double* lam, double f, double* rhs)
{
  const afem_mesh& m = *b->mesh;
  const int val_cap = kLocalRows * b->max_row; // interleaved row slots (even)
  const int node_cap = b->brick_nodes_max;
  // + 1: the coordinate block is copied in 16-byte pieces, padded to an even count
  const size_t lds = ((size_t)val_cap + 3 * (size_t)node_cap + 1) * sizeof(double);
  hipLaunchKernelGGL((k_assemble_laplace_brick<NPC, U>), dim3(grid_for(m.nb_node, kLocalRows)), dim3(kLocalRows),
                     lds, stream(), b->brec.p, b->bxyz.p, b->bxyz_off.p, b->bcount.p, b->bplan.p, b->bcell.p,
                     b->bslice.p, m.nb_node, val_cap, lam, f, rhs, b->values.p);
}

template <int NPC, int U>
static void launch_laplace_plan(const afem_bsr* b, int rows, size_t lds, const double* lam, double f, double* rhs)
{
  const afem_mesh& m = *b->mesh;
  hipLaunchKernelGGL((k_assemble_laplace_plan<NPC, U>), dim3(grid_for(m.nb_node, rows)), dim3(rows), lds, stream(),
                     m.coords.p, m.cells.p, m.nc_offsets.p, b->plan.p, b->diag_pos.p, b->row_index.p, m.nb_node,
                     rows, lam, f, rhs, b->values.p);
}

static size_t cell_lds(const afem_bsr* b)
{
  return ((size_t)kLocalRows * b->max_row + kLocalRows + 3 * (size_t)b->brick_nodes_max) * sizeof(double) +
         (size_t)b->brick_nodes_max;
}

template <int NPC>
static void launch_laplace_cell(const afem_bsr* b, const double* lam, double f, double* rhs)
{
  const afem_mesh& m = *b->mesh;
  hipLaunchKernelGGL((k_assemble_laplace_cell<NPC>), dim3(grid_for(m.nb_node, kLocalRows)), dim3(kLocalRows),
                     cell_lds(b), stream(), m.coords.p, b->bperm.p, b->cinst_off.p, b->cwords.p, b->ccell.p,
                     b->cmap.p, b->btable.p, b->bcount.p, b->diag_pos.p, b->row_index.p, m.nb_node, b->max_row,
                     b->brick_nodes_max, lam, f, rhs, b->values.p);
}

// The variant actually run: builds the plan the variant needs on first use,
// and falls back to a kind the handle has (brick -> sliced -> pair plan ->
// search).
static LaplaceVariant effective_variant(afem_bsr* b, int v)
{
  LaplaceVariant var = kLaplaceVariants[v];
  if (var.kind == kCell) {
    ensure_plans(b, kPlansBrick);
    ensure_plans(b, kPlansCell);
    if (!b->cwords.p || cell_lds(b) > 160 * 1024) {
      var.kind = kBrick;
      var.u = 16;
    }
  }
  if (var.kind == kBrick) {
    ensure_plans(b, kPlansBrick);
    const size_t brick_lds = ((size_t)kLocalRows * b->max_row + 3 * (size_t)b->brick_nodes_max + 1) * sizeof(double);
    if (!b->bplan.p || !b->brec.p || brick_lds > 160 * 1024)
      var.kind = kLocalSliced;
  }
  if (var.kind == kLocalSliced) {
    ensure_plans(b, kPlansSliced);
    const size_t local_lds = ((size_t)kLocalRows * b->max_row + 3 * (size_t)b->block_nodes_max) * sizeof(double);
    if (!b->lplan_sl.p || local_lds > 160 * 1024)
      var.kind = kPlan;
  }
  if (var.kind == kPlan && !b->plan.p)
    var.kind = kSearch;
  return var;
}

static void assemble_laplace_variant(afem_bsr* b, int v, const double* cell_lambda, double f, double* rhs)
{
  const afem_mesh& m = *b->mesh;
  const LaplaceVariant var = effective_variant(b, v);
  int rows;
  size_t lds;
  const bool fits = assembly_shape(*b, var.kind == kSearch ? 12 : 8, var.rows, &rows, &lds);
  AFEM_REQUIRE(fits, AFEM_ERR_CAPACITY, "matrix rows too long for the LDS-staged assembly");
  ProfScope prof(m.npc == 4 ? "assemble_laplace_tet" : "assemble_laplace_tri");
  if (var.kind == kCell) {
    if (m.npc == 4)
      launch_laplace_cell<4>(b, cell_lambda, f, rhs);
    else
      launch_laplace_cell<3>(b, cell_lambda, f, rhs);
  }
  else if (var.kind == kBrick) {
    if (m.npc == 4) {
      switch (var.u) {
      case 4: launch_laplace_brick<4, 4>(b, cell_lambda, f, rhs); break;
      case 16: launch_laplace_brick<4, 16>(b, cell_lambda, f, rhs); break;
      default: launch_laplace_brick<4, 8>(b, cell_lambda, f, rhs); break;
      }
    }
    else {
      switch (var.u) {
      case 4: launch_laplace_brick<3, 4>(b, cell_lambda, f, rhs); break;
      case 16: launch_laplace_brick<3, 16>(b, cell_lambda, f, rhs); break;
      default: launch_laplace_brick<3, 8>(b, cell_lambda, f, rhs); break;
      }
    }
  }
  else if (var.kind == kLocalSliced) {
    if (m.npc == 4) {
      switch (var.u) {
      case 4: launch_laplace_lplan<4, 4>(b, cell_lambda, f, rhs); break;
      case 16: launch_laplace_lplan<4, 16>(b, cell_lambda, f, rhs); break;
      default: launch_laplace_lplan<4, 8>(b, cell_lambda, f, rhs); break;
      }
    }
    else {
      switch (var.u) {
      case 4: launch_laplace_lplan<3, 4>(b, cell_lambda, f, rhs); break;
      case 16: launch_laplace_lplan<3, 16>(b, cell_lambda, f, rhs); break;
      default: launch_laplace_lplan<3, 8>(b, cell_lambda, f, rhs); break;
      }
    }
  }
  else if (var.kind == kSearch) {
    const unsigned grid = grid_for(m.nb_node, rows);
    if (m.npc == 4)
      hipLaunchKernelGGL((k_assemble_laplace<4, 4>), dim3(grid), dim3(rows), lds, stream(), m.coords.p, m.cells.p,
                         m.nc_offsets.p, m.nc_items.p, b->row_index.p, b->columns.p, m.nb_node, rows, cell_lambda, f,
                         rhs, b->values.p);
    else
      hipLaunchKernelGGL((k_assemble_laplace<3, 4>), dim3(grid), dim3(rows), lds, stream(), m.coords.p, m.cells.p,
                         m.nc_offsets.p, m.nc_items.p, b->row_index.p, b->columns.p, m.nb_node, rows, cell_lambda, f,
                         rhs, b->values.p);
  }
  else if (m.npc == 4) {
    launch_laplace_plan<4, 8>(b, rows, lds, cell_lambda, f, rhs);
  }
  else {
    launch_laplace_plan<3, 8>(b, rows, lds, cell_lambda, f, rhs);
  }
  check_launch("assemble_laplace");
}

static void assemble_laplace(afem_bsr* b, const double* cell_lambda, double f, double* rhs)
{
  AFEM_REQUIRE(b->has_sparsity, AFEM_ERR_STATE, "computeSparsity must be called before assembly");
  AFEM_REQUIRE(b->nb_dof == 1, AFEM_ERR_ARGUMENT, "the Laplace element needs nb_dof = 1");
  if (!b->mesh->simplex) {
    q1_assemble_laplace(b, cell_lambda, f, rhs);
    return;
  }
  assemble_laplace_variant(b, b->variant >= 0 ? b->variant : kDefaultLaplaceVariant, cell_lambda, f, rhs);
}

// Times every Laplace variant (best of 3 after one warm-up launch) on the
// handle's own matrix and keeps the fastest.  Values are overwritten.
static void autotune_laplace(afem_bsr* b, double* best_ms, int* best_variant)
{
  hipEvent_t e0, e1;
  AFEM_HIP(hipEventCreate(&e0));
  AFEM_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int pick = kDefaultLaplaceVariant;
  for (int v = 0; v < kNumLaplaceVariants; ++v) {
    if (effective_variant(b, v).kind != kLaplaceVariants[v].kind)
      continue; // this handle lacks the plan the variant needs
    float t_best = 1e30f;
    for (int rep = 0; rep < 4; ++rep) {
      AFEM_HIP(hipEventRecord(e0, stream()));
      assemble_laplace_variant(b, v, nullptr, 1.0, nullptr);
      AFEM_HIP(hipEventRecord(e1, stream()));
      AFEM_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      AFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (rep > 0 && ms < t_best)
        t_best = ms;
    }
    if (t_best < best) {
      best = t_best;
      pick = v;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->variant = pick;
  if (best_ms)
    *best_ms = best;
  if (best_variant)
    *best_variant = pick;
}

static void assemble_elasticity(afem_bsr* b, double lam, double mu)
{
  AFEM_REQUIRE(b->has_sparsity, AFEM_ERR_STATE, "computeSparsity must be called before assembly");
  if (!b->mesh->simplex) {
    q1_assemble_elasticity(b, lam, mu);
    return;
  }
  if (b->nb_dof == 2 && b->mesh->dim == 2 && b->mesh->npc == 3) {
    const afem_mesh& m = *b->mesh;
    DBuf<int32_t> err(1);
    err.zero();
    ProfScope prof("assemble_elasticity_tri");
    hipLaunchKernelGGL(k_assemble_elasticity_tri, dim3(grid_for(m.nb_node, 256)), dim3(256), 0, stream(), m.coords.p,
                       m.cells.p, m.nc_offsets.p, m.nc_items.p, m.nb_node, b->row_index.p, b->columns.p,
                       b->per_block ? 1 : 0, lam, mu, b->values.p, err.p);
    check_launch("assemble_elasticity_tri");
    int32_t h = 0;
    err.download(&h, 1);
    AFEM_REQUIRE(h == 0, AFEM_ERR_NOT_FOUND, "BSRMatrix(findValueIndex): Value not found");
    return;
  }
  AFEM_REQUIRE(b->nb_dof == 3 && b->mesh->npc == 4 && b->mesh->dim == 3, AFEM_ERR_ARGUMENT,
               "the elasticity element needs a triangle mesh and nb_dof = 2, or a tetrahedral mesh and nb_dof = 3");
  const afem_mesh& m = *b->mesh;
  if (b->variant != 0)
    ensure_plans(b, kPlansEdge);
  if (b->egeo.p && b->variant != 0 && b->variant != 1) {
    // cap: component stride of the staged geometry (odd, to spread the LDS banks)
    const int cap = b->e_inst_max | 1;
    const size_t lds = sizeof(double) * std::max<size_t>((size_t)13 * cap, (size_t)9 * 256);
    ProfScope prof("assemble_elasticity_tet");
    hipLaunchKernelGGL(k_assemble_elasticity_geo, dim3(b->e_nwg), dim3(256), lds, stream(), m.coords.p, b->einst4.p,
                       b->row_index.p, b->columns.p, b->ewg_row.p, b->eslice8.p, b->egeo.p, b->ecnt.p,
                       b->einst_off.p, cap, lam, mu, b->per_block ? 1 : 0, b->values.p);
    check_launch("assemble_elasticity_geo");
    return;
  }
  if (b->ecell.p && b->variant != 0) {
    ProfScope prof("assemble_elasticity_tet");
    hipLaunchKernelGGL(k_assemble_elasticity_edge, dim3(b->e_nwg), dim3(256), 0, stream(), m.coords.p, m.cells.p,
                       b->row_index.p, b->columns.p, b->ewg_row.p, b->eslice.p, b->ecell.p, b->ecnt.p, lam, mu,
                       b->per_block ? 1 : 0, b->values.p);
    check_launch("assemble_elasticity_edge");
    return;
  }
  int rows;
  size_t lds;
  AFEM_REQUIRE(assembly_shape(*b, 9 * 8 + 4, 256, &rows, &lds), AFEM_ERR_CAPACITY,
               "matrix rows too long for the LDS-staged assembly");
  const unsigned grid = grid_for(m.nb_node, rows);
  ProfScope prof("assemble_elasticity_tet");
  hipLaunchKernelGGL(k_assemble_elasticity<2>, dim3(grid), dim3(rows), lds, stream(), m.coords.p, m.cells.p,
                     m.nc_offsets.p, m.nc_items.p, b->row_index.p, b->columns.p, m.nb_node, rows, lam, mu,
                     b->per_block ? 1 : 0, b->values.p);
  check_launch("assemble_elasticity");
}

// ---- right-hand side (node-wise gather) -----------------------------------------------

__global__ void k_rhs_source(const double* __restrict__ coords, const int32_t* __restrict__ cells, int npc,
                             const int32_t* __restrict__ nc_off, const int32_t* __restrict__ nc_items,
                             int32_t nb_node, double f, const double* __restrict__ cell_scale, int nb_dof,
                             int dof_index, double* __restrict__ rhs)
{
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb_node)
    return;
  double sum = 0.0;
  for (int32_t e = nc_off[i]; e < nc_off[i + 1]; ++e) {
    const int32_t c = nc_items[e] >> 2;
    const int32_t* cn = cells + 4 * (int64_t)c;
    double measure;
    if (npc == 4) {
      const d3 p0 = load_node(coords, cn[0]);
      const d3 e1 = sub(load_node(coords, cn[1]), p0);
      const d3 e2 = sub(load_node(coords, cn[2]), p0);
      const d3 e3 = sub(load_node(coords, cn[3]), p0);
      measure = fabs(dot(e1, cross(e2, e3))) / 6.0;
    }
    else {
      const d3 p0 = load_node(coords, cn[0]);
      const d3 e1 = sub(load_node(coords, cn[1]), p0);
      const d3 e2 = sub(load_node(coords, cn[2]), p0);
      const d3 v = cross(e1, e2);
      measure = sqrt(dot(v, v)) / 2.0;
    }
    const double q = cell_scale ? f * cell_scale[c] : f;
    sum += q * measure / npc;
  }
  rhs[(int64_t)i * nb_dof + dof_index] = sum;
}

__global__ void k_rhs_body_force(const double* __restrict__ coords, const int32_t* __restrict__ cells,
                                 const int32_t* __restrict__ nc_off, const int32_t* __restrict__ nc_items,
                                 int32_t nb_node, double fx, double fy, double fz, double* __restrict__ rhs)
{
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb_node)
    return;
  double vsum = 0.0;
  for (int32_t e = nc_off[i]; e < nc_off[i + 1]; ++e) {
    const int32_t* cn = cells + 4 * (int64_t)(nc_items[e] >> 2);
    const d3 p0 = load_node(coords, cn[0]);
    const d3 e1 = sub(load_node(coords, cn[1]), p0);
    const d3 e2 = sub(load_node(coords, cn[2]), p0);
    const d3 e3 = sub(load_node(coords, cn[3]), p0);
    vsum += fabs(dot(e1, cross(e2, e3))) / 6.0 / 4;
  }
  rhs[3 * (int64_t)i + 0] += fx * vsum;
  rhs[3 * (int64_t)i + 1] += fy * vsum;
  rhs[3 * (int64_t)i + 2] += fz * vsum;
}

// ---- BSRMatrix::toCsr on the device ----------------------------------------------------

__global__ void k_csr_rows(const int32_t* __restrict__ row_index, int32_t nb_node, int nb, int32_t* __restrict__ csr_row)
{
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nb_node)
    return;
  if (i == nb_node) {
    csr_row[(int64_t)nb * nb_node] = row_index[nb_node] * nb * nb;
    return;
  }
  const int32_t len = row_index[i + 1] - row_index[i];
  for (int r = 0; r < nb; ++r)
    csr_row[(int64_t)nb * i + r] = row_index[i] * nb * nb + r * len * nb;
}

__global__ void k_csr_cols(const int32_t* __restrict__ row_index, const int32_t* __restrict__ columns, int32_t nb_node,
                           int nb, int32_t* __restrict__ csr_col)
{
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb_node)
    return;
  const int32_t b0 = row_index[i], len = row_index[i + 1] - b0;
  for (int r = 0; r < nb; ++r)
    for (int x = 0; x < len; ++x)
      for (int k = 0; k < nb; ++k)
        csr_col[(int64_t)b0 * nb * nb + (int64_t)r * len * nb + x * nb + k] = columns[b0 + x] * nb + k;
}

static void bsr_to_csr(afem_bsr* b)
{
  AFEM_REQUIRE(b->has_sparsity, AFEM_ERR_STATE, "computeSparsity must be called first");
  const int nb = b->nb_dof;
  const int32_t n = b->nb_row;
  if (b->csr_row.p)
    return;
  b->csr_row.alloc((size_t)n * nb + 1);
  b->csr_col.alloc((size_t)b->nb_col * nb * nb);
  hipLaunchKernelGGL(k_csr_rows, dim3(grid_for((int64_t)n + 1, 256)), dim3(256), 0, stream(), b->row_index.p, n, nb,
                     b->csr_row.p);
  check_launch("csr_rows");
  hipLaunchKernelGGL(k_csr_cols, dim3(grid_for(n, 256)), dim3(256), 0, stream(), b->row_index.p, b->columns.p, n, nb,
                     b->csr_col.p);
  check_launch("csr_cols");
}

// Boundary faces of k nodes on the device with, for each node they touch (ids,
// sorted), its (face, slot) positions f * k + slot in increasing order:
// pos_of[off[t]..off[t+1]).  Gather kernels walk it so that every sum runs in
// face order, with no atomics.
struct FaceIndex
{
  int32_t nb_node = 0;
  DBuf<int32_t> faces, ids, off, pos_of;
  FaceIndex(const int32_t* face_nodes, int nb_face, int k)
  {
    const size_t nf = (size_t)nb_face * k;
    std::vector<int32_t> h_ids(face_nodes, face_nodes + nf);
    std::sort(h_ids.begin(), h_ids.end());
    h_ids.erase(std::unique(h_ids.begin(), h_ids.end()), h_ids.end());
    std::vector<int32_t> h_off(h_ids.size() + 1, 0), h_pos_of(nf);
    auto slot = [&](int32_t v) { return std::lower_bound(h_ids.begin(), h_ids.end(), v) - h_ids.begin(); };
    for (size_t j = 0; j < nf; ++j)
      ++h_off[slot(face_nodes[j]) + 1];
    for (size_t t = 0; t < h_ids.size(); ++t)
      h_off[t + 1] += h_off[t];
    std::vector<int32_t> fill(h_off.begin(), h_off.end() - 1);
    for (size_t j = 0; j < nf; ++j)
      h_pos_of[fill[slot(face_nodes[j])]++] = (int32_t)j;
    nb_node = (int32_t)h_ids.size();
    faces.alloc(nf);
    ids.alloc(h_ids.size());
    off.alloc(h_off.size());
    pos_of.alloc(nf);
    faces.upload(face_nodes, nf);
    ids.upload(h_ids.data(), h_ids.size());
    off.upload(h_off.data(), h_off.size());
    pos_of.upload(h_pos_of.data(), nf);
  }
};

} // namespace afem

using namespace afem;

extern "C" {

AFEM_API int afem_bsr_create(afem_mesh_t mesh, int nb_dof, int does_linear_system_use_csr, int use_atomic_free,
                             afem_bsr_t* out)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(mesh && out, AFEM_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  // femutils/BSRFormat.cc:51-55
  AFEM_REQUIRE(nb_dof > 0, AFEM_ERR_ARGUMENT, "BSRMatrix(initialize): block_size should be positive");
  AFEM_REQUIRE(nb_dof <= mesh->nb_node, AFEM_ERR_ARGUMENT, "BSRMatrix(initialize): block_size should be less than nb_row");
  AFEM_REQUIRE(nb_dof >= 1 && nb_dof <= 3, AFEM_ERR_NOT_IMPLEMENTED, "block sizes 1, 2 and 3 are supported");
  auto* b = new afem_bsr();
  try {
    b->mesh = mesh;
    b->nb_dof = nb_dof;
    b->use_csr = does_linear_system_use_csr != 0;
    b->per_block = !b->use_csr;
    b->atomic_free = use_atomic_free != 0;
    b->nb_row = mesh->nb_node;
    bsr_count_columns(b);
  }
  catch (...) {
    delete b;
    throw;
  }
  *out = b;
  AFEM_GUARD_END
}

AFEM_API int afem_bsr_destroy(afem_bsr_t bsr)
{
  AFEM_GUARD_BEGIN
  if (bsr) {
    AFEM_HIP(hipStreamSynchronize(stream()));
    delete bsr;
  }
  AFEM_GUARD_END
}

AFEM_API int afem_bsr_compute_sparsity(afem_bsr_t bsr)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(bsr, AFEM_ERR_ARGUMENT, "null handle");
  bsr_compute_sparsity(bsr);
  build_plan(bsr);
  AFEM_GUARD_END
}

AFEM_API int afem_bsr_autotune_assembly(afem_bsr_t bsr, double* best_ms, int* variant)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(bsr && bsr->has_sparsity, AFEM_ERR_STATE, "computeSparsity must be called first");
  AFEM_REQUIRE(bsr->nb_dof == 1, AFEM_ERR_NOT_IMPLEMENTED, "assembly autotuning covers the Laplace element");
  AFEM_REQUIRE(bsr->mesh->simplex, AFEM_ERR_NOT_IMPLEMENTED, "Quad4 / Hexa8 meshes have one assembly kernel");
  autotune_laplace(bsr, best_ms, variant);
  AFEM_GUARD_END
}

AFEM_API int afem_bsr_set_assembly_variant(afem_bsr_t bsr, int variant)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(bsr, AFEM_ERR_ARGUMENT, "null handle");
  AFEM_REQUIRE(variant >= -1 && variant < kNumLaplaceVariants, AFEM_ERR_ARGUMENT, "unknown assembly variant");
  bsr->variant = variant;
  AFEM_GUARD_END
}

AFEM_API int afem_bsr_get_info(afem_bsr_t bsr, afem_bsr_info* info)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(bsr && info, AFEM_ERR_ARGUMENT, "null argument");
  info->nb_row = bsr->nb_row;
  info->nb_block = bsr->nb_dof;
  info->nb_col = bsr->nb_col;
  info->nb_non_zero_value = bsr->nb_col * bsr->nb_dof * bsr->nb_dof;
  info->order_values_per_block = bsr->per_block ? 1 : 0;
  info->max_row_length = bsr->max_row;
  AFEM_GUARD_END
}

AFEM_API int afem_bsr_reset_values(afem_bsr_t bsr)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(bsr && bsr->has_sparsity, AFEM_ERR_STATE, "no sparsity yet");
  bsr->values.zero();
  AFEM_GUARD_END
}

AFEM_API int afem_bsr_assemble_laplace(afem_bsr_t bsr, const double* cell_lambda, double source_f, double* rhs)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(bsr, AFEM_ERR_ARGUMENT, "null handle");
  assemble_laplace(bsr, cell_lambda, source_f, rhs);
  AFEM_GUARD_END
}

AFEM_API int afem_bsr_assemble_elasticity(afem_bsr_t bsr, double lame_lambda, double lame_mu)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(bsr, AFEM_ERR_ARGUMENT, "null handle");
  assemble_elasticity(bsr, lame_lambda, lame_mu);
  AFEM_GUARD_END
}

AFEM_API int afem_bsr_device_arrays(afem_bsr_t bsr, const int32_t** row_index, const int32_t** columns, double** values,
                                    const int32_t** nnz_per_row)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(bsr, AFEM_ERR_ARGUMENT, "null handle");
  if (row_index)
    *row_index = bsr->row_index.p;
  if (columns)
    *columns = bsr->columns.p;
  if (values)
    *values = bsr->values.p;
  if (nnz_per_row)
    *nnz_per_row = bsr->nnz_per_row.p;
  AFEM_GUARD_END
}

AFEM_API int afem_bsr_download(afem_bsr_t bsr, int32_t* row_index, int32_t* columns, double* values,
                               int32_t* nnz_per_row)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(bsr, AFEM_ERR_ARGUMENT, "null handle");
  if (row_index)
    bsr->row_index.download(row_index, (size_t)bsr->nb_row + 1);
  if (nnz_per_row)
    bsr->nnz_per_row.download(nnz_per_row, (size_t)bsr->nb_row);
  if (columns || values)
    AFEM_REQUIRE(bsr->has_sparsity, AFEM_ERR_STATE, "computeSparsity not called");
  if (columns)
    bsr->columns.download(columns, (size_t)bsr->nb_col);
  if (values)
    bsr->values.download(values, bsr->values.n);
  AFEM_GUARD_END
}

AFEM_API int afem_bsr_to_csr(afem_bsr_t bsr, const int32_t** csr_row, const int32_t** csr_col, int32_t* nb_csr_row)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(bsr, AFEM_ERR_ARGUMENT, "null handle");
  bsr_to_csr(bsr);
  if (csr_row)
    *csr_row = bsr->csr_row.p;
  if (csr_col)
    *csr_col = bsr->csr_col.p;
  if (nb_csr_row)
    *nb_csr_row = bsr->nb_row * bsr->nb_dof;
  AFEM_GUARD_END
}

AFEM_API int afem_rhs_constant_source(afem_mesh_t mesh, double f, const double* cell_scale, int nb_dof, int dof_index,
                                      double* rhs)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(mesh && rhs, AFEM_ERR_ARGUMENT, "null argument");
  AFEM_REQUIRE(nb_dof >= 1 && dof_index >= 0 && dof_index < nb_dof, AFEM_ERR_ARGUMENT, "bad dof index");
  if (!mesh->simplex) {
    q1_rhs_constant_source(*mesh, f, cell_scale, nb_dof, dof_index, rhs);
    return AFEM_OK;
  }
  ProfScope prof("rhs_constant_source");
  hipLaunchKernelGGL(k_rhs_source, dim3(grid_for(mesh->nb_node, 256)), dim3(256), 0, stream(), mesh->coords.p,
                     mesh->cells.p, mesh->npc, mesh->nc_offsets.p, mesh->nc_items.p, mesh->nb_node, f, cell_scale,
                     nb_dof, dof_index, rhs);
  check_launch("rhs_source");
  AFEM_GUARD_END
}

AFEM_API int afem_rhs_manufactured(afem_mesh_t mesh, int solution, int nb_dof, int dof_index, double* rhs)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(mesh && rhs, AFEM_ERR_ARGUMENT, "null argument");
  AFEM_REQUIRE(nb_dof >= 1 && dof_index >= 0 && dof_index < nb_dof, AFEM_ERR_ARGUMENT, "bad dof index");
  AFEM_REQUIRE(solution >= 0 && solution <= 2, AFEM_ERR_ARGUMENT, "unknown manufactured solution");
  AFEM_REQUIRE(!mesh->simplex || (mesh->dim == 2 && mesh->npc == 3), AFEM_ERR_NOT_IMPLEMENTED,
               "Manufactured source is not implemented for tetrahedra");
  rhs_manufactured(*mesh, solution, nb_dof, dof_index, rhs);
  AFEM_GUARD_END
}

AFEM_API int afem_manufactured_l2_error(afem_mesh_t mesh, int solution, const double* u, double* l2_error)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(mesh && u && l2_error, AFEM_ERR_ARGUMENT, "null argument");
  AFEM_REQUIRE(solution >= 0 && solution <= 2, AFEM_ERR_ARGUMENT, "unknown manufactured solution");
  AFEM_REQUIRE(!mesh->simplex || (mesh->dim == 2 && mesh->npc == 3), AFEM_ERR_NOT_IMPLEMENTED,
               "Manufactured L2 error is not implemented for tetrahedra");
  DBuf<double> err2((size_t)mesh->nb_cell);
  manufactured_cell_error(*mesh, solution, u, err2.p);
  std::vector<double> h((size_t)mesh->nb_cell);
  err2.download(h.data(), h.size());
  double sum = 0.0; // in cell order, as the reference's ENUMERATE_ (Cell, ...) loop
  for (double v : h)
    sum += v;
  *l2_error = std::sqrt(sum);
  AFEM_GUARD_END
}

AFEM_API int afem_bsr_add_convection(afem_bsr_t bsr, int nb_face, const int32_t* face_nodes, double h)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(bsr && (nb_face == 0 || face_nodes), AFEM_ERR_ARGUMENT, "null argument");
  AFEM_REQUIRE(nb_face >= 0, AFEM_ERR_ARGUMENT, "negative face count");
  AFEM_REQUIRE(bsr->nb_dof == 1, AFEM_ERR_ARGUMENT, "the convection term needs nb_dof = 1");
  AFEM_REQUIRE(bsr->has_sparsity, AFEM_ERR_STATE, "computeSparsity must be called first");
  const afem_mesh& m = *bsr->mesh;
  if (nb_face == 0)
    return AFEM_OK;
  const int k = m.npc == 8 ? 4 : m.dim; // nodes per face
  for (size_t j = 0; j < (size_t)nb_face * k; ++j)
    AFEM_REQUIRE(face_nodes[j] >= 0 && face_nodes[j] < m.nb_node, AFEM_ERR_ARGUMENT, "convection face node out of range");
  FaceIndex fx(face_nodes, nb_face, k);
  DBuf<double> measure((size_t)nb_face);
  DBuf<int32_t> err(1);
  err.zero();
  ProfScope prof("assemble_convection");
  hipLaunchKernelGGL(k_face_measure, dim3(grid_for(nb_face, 256)), dim3(256), 0, stream(), k, nb_face, fx.faces.p,
                     m.coords.p, measure.p);
  check_launch("face_measure");
  const double hf = h * (k == 2 ? 1 / 6. : k == 3 ? 1 / 12. : 1 / 20.);
  for (int check_only = 1; check_only >= 0; --check_only) { // a failed call leaves the values as they were
    hipLaunchKernelGGL(k_convection_gather, dim3(grid_for(fx.nb_node, 256)), dim3(256), 0, stream(), fx.nb_node,
                       fx.ids.p, fx.off.p, fx.pos_of.p, fx.faces.p, k, measure.p, hf, bsr->row_index.p,
                       bsr->columns.p, bsr->values.p, check_only, err.p);
    check_launch("convection_gather");
    if (check_only) {
      int32_t e = 0;
      err.download(&e, 1);
      AFEM_REQUIRE(e == 0, AFEM_ERR_NOT_FOUND, "a convection face pair is not in the matrix pattern");
    }
  }
  AFEM_GUARD_END
}

AFEM_API int afem_rhs_neumann(afem_mesh_t mesh, int nb_face, const int32_t* face_nodes, int nb_value,
                              const double* values, int nb_dof, int dof_index, double* rhs)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(mesh && rhs && (nb_face == 0 || face_nodes) && values, AFEM_ERR_ARGUMENT, "null argument");
  AFEM_REQUIRE(nb_face >= 0, AFEM_ERR_ARGUMENT, "negative face count");
  AFEM_REQUIRE(nb_value == 1 || nb_value == mesh->dim, AFEM_ERR_ARGUMENT,
               "a Neumann value is a scalar or a vector of the mesh dimension");
  AFEM_REQUIRE(nb_dof >= 1 && dof_index >= 0 && dof_index < nb_dof, AFEM_ERR_ARGUMENT, "bad dof index");
  if (nb_face == 0)
    return AFEM_OK;
  const int k = mesh->npc == 8 ? 4 : mesh->dim; // nodes per face
  const size_t nf = (size_t)nb_face * k;
  for (size_t j = 0; j < nf; ++j)
    AFEM_REQUIRE(face_nodes[j] >= 0 && face_nodes[j] < mesh->nb_node, AFEM_ERR_ARGUMENT,
                 "Neumann face node out of range");
  FaceIndex fx(face_nodes, nb_face, k);
  DBuf<int32_t> err(1);
  DBuf<double> contrib((size_t)nb_face * k);
  err.zero();
  const bool scalar = nb_value == 1;
  ProfScope prof("rhs_neumann");
  const double gz = (!scalar && mesh->dim == 3) ? values[2] : 0.0;
  if (!mesh->simplex) {
    hipLaunchKernelGGL(k_neumann_faces_q1, dim3(grid_for(nb_face, 256)), dim3(256), 0, stream(), mesh->dim, nb_face,
                       fx.faces.p, mesh->coords.p, mesh->cells.p, mesh->npc, mesh->slot_bits, mesh->nc_offsets.p,
                       mesh->nc_items.p, scalar ? 1 : 0, values[0], scalar ? 0.0 : values[0],
                       scalar ? 0.0 : values[1], gz, contrib.p, err.p);
  }
  else {
    hipLaunchKernelGGL(k_neumann_faces, dim3(grid_for(nb_face, 256)), dim3(256), 0, stream(), mesh->dim, nb_face,
                       fx.faces.p, mesh->coords.p, mesh->cells.p, mesh->nc_offsets.p, mesh->nc_items.p,
                       scalar ? 1 : 0, values[0], scalar ? 0.0 : values[0], scalar ? 0.0 : values[1], gz, contrib.p,
                       err.p);
  }
  check_launch("neumann_faces");
  int32_t h = 0;
  err.download(&h, 1); // before rhs is touched: a failed call leaves it as it was
  AFEM_REQUIRE(h == 0, AFEM_ERR_MESH, "a Neumann face is not a face of any cell");
  hipLaunchKernelGGL(k_neumann_gather, dim3(grid_for(fx.nb_node, 256)), dim3(256), 0, stream(), fx.nb_node,
                     fx.ids.p, fx.off.p, fx.pos_of.p, contrib.p, nb_dof, dof_index, rhs);
  check_launch("neumann_gather");
  AFEM_GUARD_END
}

AFEM_API int afem_rhs_face_uniform(afem_mesh_t mesh, int nb_face, const int32_t* face_nodes, double g, int nb_dof,
                                   int dof_index, double* rhs)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(mesh && rhs && (nb_face == 0 || face_nodes), AFEM_ERR_ARGUMENT, "null argument");
  AFEM_REQUIRE(nb_face >= 0, AFEM_ERR_ARGUMENT, "negative face count");
  AFEM_REQUIRE(nb_dof >= 1 && dof_index >= 0 && dof_index < nb_dof, AFEM_ERR_ARGUMENT, "bad dof index");
  if (nb_face == 0)
    return AFEM_OK;
  const int k = mesh->npc == 8 ? 4 : mesh->dim; // nodes per face
  for (size_t j = 0; j < (size_t)nb_face * k; ++j)
    AFEM_REQUIRE(face_nodes[j] >= 0 && face_nodes[j] < mesh->nb_node, AFEM_ERR_ARGUMENT, "face node out of range");
  FaceIndex fx(face_nodes, nb_face, k);
  DBuf<double> measure((size_t)nb_face);
  ProfScope prof("rhs_face_uniform");
  hipLaunchKernelGGL(k_face_measure, dim3(grid_for(nb_face, 256)), dim3(256), 0, stream(), k, nb_face, fx.faces.p,
                     mesh->coords.p, measure.p);
  check_launch("face_measure");
  hipLaunchKernelGGL(k_face_uniform_gather, dim3(grid_for(fx.nb_node, 256)), dim3(256), 0, stream(), fx.nb_node,
                     fx.ids.p, fx.off.p, fx.pos_of.p, k, measure.p, g, nb_dof, dof_index, rhs);
  check_launch("face_uniform_gather");
  AFEM_GUARD_END
}

AFEM_API int afem_bsr_assemble_heat(afem_bsr_t bsr, const double* cell_lambda, double dt)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(bsr, AFEM_ERR_ARGUMENT, "null handle");
  AFEM_REQUIRE(bsr->nb_dof == 1, AFEM_ERR_ARGUMENT, "the heat element needs nb_dof = 1");
  AFEM_REQUIRE(dt > 0.0, AFEM_ERR_ARGUMENT, "the time step must be positive");
  if (!bsr->mesh->simplex) {
    AFEM_REQUIRE(bsr->has_sparsity, AFEM_ERR_STATE, "computeSparsity must be called before assembly");
    q1_assemble_heat(bsr, cell_lambda, dt);
    return AFEM_OK;
  }
  assemble_laplace(bsr, cell_lambda, 0.0, nullptr);
  const afem_mesh& m = *bsr->mesh;
  ProfScope prof("assemble_heat_mass");
  hipLaunchKernelGGL(k_add_mass, dim3(grid_for(m.nb_node, 256)), dim3(256), 0, stream(), m.coords.p, m.cells.p,
                     m.npc, m.nc_offsets.p, m.nc_items.p, bsr->row_index.p, bsr->columns.p, m.nb_node, dt,
                     bsr->values.p);
  check_launch("add_mass");
  AFEM_GUARD_END
}

AFEM_API int afem_rhs_mass_field(afem_mesh_t mesh, const double* field, double scale, int nb_dof, int dof_index,
                                 double* rhs)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(mesh && field && rhs, AFEM_ERR_ARGUMENT, "null argument");
  if (!mesh->simplex) {
    AFEM_REQUIRE(nb_dof >= 1 && dof_index >= 0 && dof_index < nb_dof, AFEM_ERR_ARGUMENT, "bad dof index");
    q1_rhs_mass_field(*mesh, field, scale, nb_dof, dof_index, rhs);
    return AFEM_OK;
  }
  AFEM_REQUIRE(nb_dof >= 1 && dof_index >= 0 && dof_index < nb_dof, AFEM_ERR_ARGUMENT, "bad dof index");
  hipLaunchKernelGGL(k_rhs_mass_field, dim3(grid_for(mesh->nb_node, 256)), dim3(256), 0, stream(), mesh->coords.p,
                     mesh->cells.p, mesh->npc, mesh->nc_offsets.p, mesh->nc_items.p, mesh->nb_node, field, scale,
                     nb_dof, dof_index, rhs);
  check_launch("rhs_mass_field");
  AFEM_GUARD_END
}

AFEM_API int afem_rhs_body_force(afem_mesh_t mesh, const double* f, double* rhs)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(mesh && f && rhs, AFEM_ERR_ARGUMENT, "null argument");
  AFEM_REQUIRE(mesh->npc == 4 && mesh->dim == 3, AFEM_ERR_NOT_IMPLEMENTED, "body force needs a tetrahedral mesh");
  hipLaunchKernelGGL(k_rhs_body_force, dim3(grid_for(mesh->nb_node, 256)), dim3(256), 0, stream(), mesh->coords.p,
                     mesh->cells.p, mesh->nc_offsets.p, mesh->nc_items.p, mesh->nb_node, f[0], f[1], f[2], rhs);
  check_launch("rhs_body_force");
  AFEM_GUARD_END
}

} // extern "C"
