// mesh.hip -- synthetic structured meshes (host) and the device mesh:
// coordinates, cell->node connectivity and the node->cell connectivity built
// on the GPU (count, scan, fill, per-node sort).
#include "common.h"
#include "device_utils.h"

#include <cmath>
#include <limits>

namespace afem {

// ---- synthetic meshes ----------------------------------------------------------

static inline uint64_t splitmix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Uniform in [-1, 1) from (seed, key), counter-based so any node can be
// generated independently.
static inline double jitter_unit(uint64_t seed, uint64_t key)
{
  uint64_t r = splitmix64(splitmix64(seed) + key);
  return (double)(r >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

static void box_sizes(int dim, int n, int32_t* nb_node, int32_t* nb_cell)
{
  AFEM_REQUIRE(dim == 2 || dim == 3, AFEM_ERR_ARGUMENT, "dim must be 2 or 3");
  AFEM_REQUIRE(n >= 1, AFEM_ERR_ARGUMENT, "n must be >= 1");
  int64_t np = (int64_t)n + 1;
  int64_t nn = (dim == 2) ? np * np : np * np * np;
  int64_t nc = (dim == 2) ? 2LL * n * n : 6LL * n * n * n;
  AFEM_REQUIRE(nn < (1LL << 31) && nc < (1LL << 29), AFEM_ERR_ARGUMENT, "mesh too large for 32-bit ids");
  *nb_node = (int32_t)nn;
  *nb_cell = (int32_t)nc;
}

static void box_mesh(int dim, int n, double jitter, uint64_t seed, double* coords, int32_t* cells)
{
  int32_t nb_node, nb_cell;
  box_sizes(dim, n, &nb_node, &nb_cell);
  AFEM_REQUIRE(jitter >= 0.0 && jitter < 0.5, AFEM_ERR_ARGUMENT, "jitter must be in [0, 0.5)");
  const int64_t np = (int64_t)n + 1;
  const double h = 1.0 / n;
  if (dim == 2) {
    for (int64_t j = 0; j <= n; ++j)
      for (int64_t i = 0; i <= n; ++i) {
        int64_t id = i + np * j;
        double x = i * h, y = j * h;
        if (i > 0 && i < n && j > 0 && j < n) {
          x += jitter * h * jitter_unit(seed, (uint64_t)id * 3 + 0);
          y += jitter * h * jitter_unit(seed, (uint64_t)id * 3 + 1);
        }
        coords[3 * id + 0] = x;
        coords[3 * id + 1] = y;
        coords[3 * id + 2] = 0.0;
      }
    for (int64_t j = 0; j < n; ++j)
      for (int64_t i = 0; i < n; ++i) {
        int64_t q = i + (int64_t)n * j;
        int32_t v00 = (int32_t)(i + np * j), v10 = v00 + 1;
        int32_t v01 = (int32_t)(v00 + np), v11 = v01 + 1;
        int32_t* t0 = cells + 8 * q;
        t0[0] = v00, t0[1] = v10, t0[2] = v11, t0[3] = -1;
        int32_t* t1 = t0 + 4;
        t1[0] = v00, t1[1] = v11, t1[2] = v01, t1[3] = -1;
      }
  }
  else {
    for (int64_t k = 0; k <= n; ++k)
      for (int64_t j = 0; j <= n; ++j)
        for (int64_t i = 0; i <= n; ++i) {
          int64_t id = i + np * (j + np * k);
          double x = i * h, y = j * h, z = k * h;
          if (i > 0 && i < n && j > 0 && j < n && k > 0 && k < n) {
            x += jitter * h * jitter_unit(seed, (uint64_t)id * 3 + 0);
            y += jitter * h * jitter_unit(seed, (uint64_t)id * 3 + 1);
            z += jitter * h * jitter_unit(seed, (uint64_t)id * 3 + 2);
          }
          coords[3 * id + 0] = x;
          coords[3 * id + 1] = y;
          coords[3 * id + 2] = z;
        }
    // Kuhn split: one tet per axis permutation (a,b,c), path
    // v000 -> +e_a -> +e_a+e_b -> v111; odd permutations swap slots 1 and 2
    // to keep a positive orientation.
    static const int perm[6][3] = { { 0, 1, 2 }, { 0, 2, 1 }, { 1, 0, 2 }, { 1, 2, 0 }, { 2, 0, 1 }, { 2, 1, 0 } };
    static const bool odd[6] = { false, true, true, false, false, true };
    const int64_t stride[3] = { 1, np, np * np };
    for (int64_t k = 0; k < n; ++k)
      for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) {
          int64_t hx = i + (int64_t)n * (j + (int64_t)n * k);
          int64_t v000 = i + np * (j + np * k);
          int64_t v111 = v000 + stride[0] + stride[1] + stride[2];
          for (int t = 0; t < 6; ++t) {
            int64_t a = v000 + stride[perm[t][0]];
            int64_t b = a + stride[perm[t][1]];
            int32_t* c = cells + 4 * (6 * hx + t);
            c[0] = (int32_t)v000;
            c[1] = (int32_t)(odd[t] ? b : a);
            c[2] = (int32_t)(odd[t] ? a : b);
            c[3] = (int32_t)v111;
          }
        }
  }
}

// ---- node -> cell connectivity on the device ----------------------------------------

__global__ void k_count_node_cells(const int32_t* __restrict__ cells, int32_t nb_cell, int npc, int stride,
                                   int32_t* __restrict__ counts)
{
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nb_cell * npc)
    return;
  int64_t c = t / npc;
  int s = (int)(t - c * npc);
  atomicAdd(&counts[cells[stride * c + s]], 1);
}

__global__ void k_fill_node_cells(const int32_t* __restrict__ cells, int32_t nb_cell, int npc, int stride,
                                  const int32_t* __restrict__ offsets, int32_t* __restrict__ cursor,
                                  int32_t* __restrict__ items)
{
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nb_cell * npc)
    return;
  int64_t c = t / npc;
  int s = (int)(t - c * npc);
  int32_t n = cells[stride * c + s];
  int32_t pos = offsets[n] + atomicAdd(&cursor[n], 1);
  items[pos] = (int32_t)(stride * c + s);
}

// Insertion sort of each node's list (deterministic cell order) and the
// largest list length.
__global__ void k_sort_node_cells(const int32_t* __restrict__ offsets, int32_t nb_node,
                                  int32_t* __restrict__ items, int32_t* __restrict__ max_degree)
{
  int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nb_node)
    return;
  int32_t b = offsets[n], e = offsets[n + 1];
  for (int32_t i = b + 1; i < e; ++i) {
    int32_t v = items[i];
    int32_t j = i - 1;
    while (j >= b && items[j] > v) {
      items[j + 1] = items[j];
      --j;
    }
    items[j + 1] = v;
  }
  atomicMax(max_degree, e - b);
}

static afem_mesh* mesh_create(int dim, int32_t nb_node, const double* coords, int32_t nb_cell, int32_t npc,
                              const int32_t* cells)
{
  AFEM_REQUIRE(dim == 2 || dim == 3, AFEM_ERR_ARGUMENT, "dim must be 2 or 3");
  AFEM_REQUIRE(npc == dim + 1 || (dim == 2 && npc == 4) || (dim == 3 && npc == 8), AFEM_ERR_NOT_IMPLEMENTED,
               "supported cells: tri3 and quad4 in 2D, tet4 and hexa8 in 3D");
  const int stride = npc > 4 ? 8 : 4;
  AFEM_REQUIRE(nb_node > 0 && nb_cell > 0, AFEM_ERR_ARGUMENT, "empty mesh");
  AFEM_REQUIRE(nb_cell < (1 << 28), AFEM_ERR_ARGUMENT, "too many cells for 32-bit (cell, slot) ids");
  AFEM_REQUIRE(coords && cells, AFEM_ERR_ARGUMENT, "null mesh arrays");
  // validate connectivity on the host: every kernel indexes with these ids
  for (int64_t c = 0; c < nb_cell; ++c)
    for (int s = 0; s < npc; ++s) {
      int32_t v = cells[stride * c + s];
      if (v < 0 || v >= nb_node)
        fail(AFEM_ERR_MESH, "cell " + std::to_string(c) + " references node " + std::to_string(v) +
                              " outside [0, " + std::to_string(nb_node) + ")");
    }
  auto* m = new afem_mesh();
  try {
    m->dim = dim;
    m->npc = npc;
    m->stride = stride;
    m->slot_bits = stride == 8 ? 3 : 2;
    m->simplex = npc == dim + 1;
    m->nb_node = nb_node;
    m->nb_cell = nb_cell;
    std::vector<double> packed((size_t)nb_node * 4);
    for (int64_t i = 0; i < nb_node; ++i) {
      packed[4 * i + 0] = coords[3 * i + 0];
      packed[4 * i + 1] = coords[3 * i + 1];
      packed[4 * i + 2] = coords[3 * i + 2];
      packed[4 * i + 3] = 0.0;
    }
    m->coords.alloc(packed.size());
    m->coords.upload(packed.data(), packed.size());
    std::vector<int32_t> cl((size_t)nb_cell * stride);
    for (int64_t c = 0; c < nb_cell; ++c)
      for (int s = 0; s < stride; ++s)
        cl[stride * c + s] = (s < npc) ? cells[stride * c + s] : -1;
    m->cells.alloc(cl.size());
    m->cells.upload(cl.data(), cl.size());

    DBuf<int32_t> counts((size_t)nb_node);
    counts.zero();
    int64_t work = (int64_t)nb_cell * npc;
    hipLaunchKernelGGL(k_count_node_cells, dim3(grid_for(work, 256)), dim3(256), 0, stream(), m->cells.p, nb_cell,
                       npc, stride, counts.p);
    check_launch("count_node_cells");
    m->nc_offsets.alloc((size_t)nb_node + 1);
    int64_t total = exclusive_scan_i32(counts.p, m->nc_offsets.p, nb_node);
    AFEM_REQUIRE(total == work, AFEM_ERR_MESH, "node->cell count mismatch");
    m->nc_items.alloc((size_t)total);
    counts.zero();
    hipLaunchKernelGGL(k_fill_node_cells, dim3(grid_for(work, 256)), dim3(256), 0, stream(), m->cells.p, nb_cell,
                       npc, stride, m->nc_offsets.p, counts.p, m->nc_items.p);
    check_launch("fill_node_cells");
    DBuf<int32_t> maxdeg(1);
    maxdeg.zero();
    hipLaunchKernelGGL(k_sort_node_cells, dim3(grid_for(nb_node, 256)), dim3(256), 0, stream(), m->nc_offsets.p,
                       nb_node, m->nc_items.p, maxdeg.p);
    check_launch("sort_node_cells");
    maxdeg.download(&m->max_degree, 1);
  }
  catch (...) {
    delete m;
    throw;
  }
  return m;
}

} // namespace afem

using namespace afem;

extern "C" {

AFEM_API int afem_box_mesh_size(int dim, int n, int32_t* nb_node, int32_t* nb_cell)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(nb_node && nb_cell, AFEM_ERR_ARGUMENT, "null output");
  box_sizes(dim, n, nb_node, nb_cell);
  AFEM_GUARD_END
}

AFEM_API int afem_box_mesh(int dim, int n, double jitter, uint64_t seed, double* coords, int32_t* cells)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(coords && cells, AFEM_ERR_ARGUMENT, "null output");
  box_mesh(dim, n, jitter, seed, coords, cells);
  AFEM_GUARD_END
}

AFEM_API int afem_mesh_create(int dim, int32_t nb_node, const double* coords, int32_t nb_cell, int32_t nodes_per_cell,
                              const int32_t* cells, afem_mesh_t* mesh)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(mesh, AFEM_ERR_ARGUMENT, "null output handle");
  *mesh = nullptr;
  *mesh = mesh_create(dim, nb_node, coords, nb_cell, nodes_per_cell, cells);
  AFEM_GUARD_END
}

AFEM_API int afem_mesh_destroy(afem_mesh_t mesh)
{
  AFEM_GUARD_BEGIN
  if (mesh) {
    AFEM_HIP(hipStreamSynchronize(stream()));
    delete mesh;
  }
  AFEM_GUARD_END
}

AFEM_API int afem_mesh_node_cells(afem_mesh_t mesh, int32_t* offsets, int32_t* items)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(mesh, AFEM_ERR_ARGUMENT, "null mesh");
  if (offsets)
    mesh->nc_offsets.download(offsets, (size_t)mesh->nb_node + 1);
  if (items)
    mesh->nc_items.download(items, mesh->nc_items.n);
  AFEM_GUARD_END
}

} // extern "C"
