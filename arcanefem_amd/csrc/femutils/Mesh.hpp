// Mesh.hpp -- C++ stand-in for the Arcane mesh that ArcaneFEM's modules hand to
// BSRFormat (IMesh + UnstructuredMeshConnectivityView + node groups), over the
// afem C ABI, plus a reader for the binary Gmsh 4.1 files the reference's test
// cases use (meshes/msh/*.msh, "$MeshFormat 4.1 1 8").
//
// Header-only; errors are thrown as afem::femutils::FatalError, the
// counterpart of ARCANE_FATAL.
#pragma once

#include "afem.h"

#include <algorithm>
#include <cstdint>
#include <iterator>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace afem::femutils {

using Real = double;
using Int32 = std::int32_t;
using Int64 = std::int64_t;
using Byte = std::uint8_t;

struct FatalError : std::runtime_error
{
  int status;
  FatalError(int s, const std::string& m)
  : std::runtime_error(m)
  , status(s)
  {}
};

inline void check(int rc, const char* where)
{
  if (rc != AFEM_OK)
    throw FatalError(rc, std::string(where) + ": " + afem_last_error());
}

// Node ids of one named boundary group (the nodes of the faces carrying a
// physical name), local ids.
using NodeGroup = std::vector<Int32>;

class Mesh
{
 public:
  Mesh() = default;
  Mesh(const Mesh&) = delete;
  Mesh& operator=(const Mesh&) = delete;
  ~Mesh()
  {
    if (m_handle)
      afem_mesh_destroy(m_handle);
  }

  // Host arrays: coordinates (N x 3), cells (C x stride: 4, or 8 for Hexa8),
  // node unique ids; npc = nodes per cell (0: dim + 1, a simplex mesh);
  // upload = false keeps the mesh on the host (parse-only use).
  void build(int dim, std::vector<Real> coords, std::vector<Int32> cells, std::vector<Int64> uids,
             bool upload = true, int npc = 0)
  {
    m_dim = dim;
    m_npc = npc > 0 ? npc : dim + 1;
    m_coords = std::move(coords);
    m_cells = std::move(cells);
    m_uids = std::move(uids);
    if (m_handle)
      afem_mesh_destroy(m_handle);
    m_handle = nullptr;
    if (upload)
      check(afem_mesh_create(dim, nbNode(), m_coords.data(), nbCell(), m_npc, m_cells.data(), &m_handle),
            "afem_mesh_create");
  }
  const std::map<std::string, NodeGroup>& groups() const { return m_groups; }

  // Structured unit box (see afem_box_mesh); uids = lid + 1 like Gmsh tags.
  void buildBox(int dim, int n, double jitter, std::uint64_t seed)
  {
    Int32 nn = 0, nc = 0;
    check(afem_box_mesh_size(dim, n, &nn, &nc), "afem_box_mesh_size");
    std::vector<Real> x((size_t)nn * 3);
    std::vector<Int32> c((size_t)nc * 4);
    check(afem_box_mesh(dim, n, jitter, seed, x.data(), c.data()), "afem_box_mesh");
    std::vector<Int64> uid((size_t)nn);
    for (Int32 i = 0; i < nn; ++i)
      uid[i] = i + 1;
    build(dim, std::move(x), std::move(c), std::move(uid));
  }

  // Arcane's Cartesian2D / Cartesian3D generator (the meshes of
  // modules/elasticity/inputs/bar.{2D,3D}.cartesian.Dirichlet.bodyForce.arc):
  // n[d] cells of length len[d] / n[d] per axis from the origin, node unique
  // ids from 0 with x fastest, then y, then z; Quad4 cells counter-clockwise
  // from the lower-left corner, Hexa8 cells the bottom face then the top face;
  // node groups XMIN, XMAX, YMIN, YMAX (, ZMIN, ZMAX) of the boundary planes
  // (generate-sod-groups).
  void buildCartesian(int dim, const int* n, const double* len)
  {
    const int nz = dim == 3 ? n[2] : 0;
    const Int64 sx = n[0] + 1, sy = n[1] + 1, sz = nz + 1;
    const Int64 nn = sx * sy * sz;
    std::vector<Real> x((size_t)nn * 3);
    std::vector<Int64> uid((size_t)nn);
    for (Int64 k = 0; k < sz; ++k)
      for (Int64 j = 0; j < sy; ++j)
        for (Int64 i = 0; i < sx; ++i) {
          const Int64 v = i + sx * (j + sy * k);
          x[3 * v] = len[0] * (double)i / n[0];
          x[3 * v + 1] = len[1] * (double)j / n[1];
          x[3 * v + 2] = dim == 3 ? len[2] * (double)k / nz : 0.0;
          uid[(size_t)v] = v;
        }
    const int npc = dim == 3 ? 8 : 4;
    std::vector<Int32> c;
    for (Int64 k = 0; k < (dim == 3 ? nz : 1); ++k)
      for (Int64 j = 0; j < n[1]; ++j)
        for (Int64 i = 0; i < n[0]; ++i) {
          const Int32 v = (Int32)(i + sx * (j + sy * k));
          const Int32 quad[4] = { v, v + 1, (Int32)(v + sx + 1), (Int32)(v + sx) };
          for (int a = 0; a < 4; ++a)
            c.push_back(quad[a]);
          if (dim == 3)
            for (int a = 0; a < 4; ++a)
              c.push_back((Int32)(quad[a] + sx * sy));
        }
    m_groups.clear();
    const char* names[3][2] = { { "XMIN", "XMAX" }, { "YMIN", "YMAX" }, { "ZMIN", "ZMAX" } };
    for (Int64 v = 0; v < nn; ++v) {
      const Int64 idx[3] = { v % sx, (v / sx) % sy, v / (sx * sy) };
      const Int64 hi[3] = { n[0], n[1], nz };
      for (int d = 0; d < dim; ++d) {
        if (idx[d] == 0)
          m_groups[names[d][0]].push_back((Int32)v);
        if (idx[d] == hi[d])
          m_groups[names[d][1]].push_back((Int32)v);
      }
    }
    build(dim, std::move(x), std::move(c), std::move(uid), true, npc);
  }

  // Binary Gmsh 4.1 (size_t = 8): nodes, the cells of the highest dimension
  // (tri3 or quad4 in 2D, tet4 or hexa8 in 3D, one kind per mesh), the node
  // and face groups of the (dim-1) physical names and the node groups of the
  // named physical points.
  void readGmsh41(const std::string& path, bool upload = true)
  {
    std::ifstream f(path, std::ios::binary);
    if (!f)
      throw FatalError(AFEM_ERR_ARGUMENT, "cannot open mesh file '" + path + "'");
    std::string data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    size_t p = 0;
    auto line = [&]() {
      size_t e = data.find('\n', p);
      if (e == std::string::npos)
        throw FatalError(AFEM_ERR_ARGUMENT, "truncated mesh file");
      std::string s = data.substr(p, e - p);
      p = e + 1;
      while (!s.empty() && (s.back() == '\r' || s.back() == ' '))
        s.pop_back();
      return s;
    };
    auto take = [&](void* dst, size_t n) {
      if (n > data.size() - p)
        throw FatalError(AFEM_ERR_ARGUMENT, "truncated mesh file");
      if (n == 0)
        return; // dst may be the null data() of an empty vector
      std::memcpy(dst, data.data() + p, n);
      p += n;
    };
    auto i32 = [&]() {
      int32_t v;
      take(&v, 4);
      return v;
    };
    auto u64 = [&]() {
      uint64_t v;
      take(&v, 8);
      return v;
    };
    // a count of `each`-byte records read next must fit in what is left
    auto fits = [&](uint64_t count, uint64_t each) {
      if (count > (data.size() - p) / each)
        throw FatalError(AFEM_ERR_ARGUMENT, "corrupt mesh file: a record count exceeds the file");
      return (size_t)count;
    };
    auto to_int = [&](const std::string& t) {
      try {
        size_t used = 0;
        const int v = std::stoi(t, &used);
        if (used == 0)
          throw std::invalid_argument(t);
        return v;
      }
      catch (const std::exception&) {
        throw FatalError(AFEM_ERR_ARGUMENT, "corrupt mesh file: '" + t + "' is not a number");
      }
    };
    std::map<std::pair<int, int>, std::string> names;
    std::map<std::pair<int, int>, std::vector<int>> entity_phys;
    std::vector<Int64> tags;
    std::vector<Real> xyz;
    struct Block
    {
      int dim, tag, k;
      std::vector<uint64_t> conn;
    };
    std::vector<Block> blocks;
    // Gmsh element types: point, line, triangle, quadrangle, tetrahedron, hexahedron
    static const std::map<int, int> nodes_of = { { 15, 1 }, { 1, 2 }, { 2, 3 }, { 3, 4 }, { 4, 4 }, { 5, 8 } };
    while (p < data.size()) {
      std::string head = line();
      if (head.empty())
        continue;
      if (head == "$MeshFormat") {
        std::string fmt = line();
        if (fmt.rfind("4.1 1 8", 0) != 0)
          throw FatalError(AFEM_ERR_NOT_IMPLEMENTED, "only binary Gmsh 4.1 with 8-byte size_t is supported");
        if (i32() != 1)
          throw FatalError(AFEM_ERR_NOT_IMPLEMENTED, "big-endian Gmsh file");
        line();
        line(); // $EndMeshFormat
      }
      else if (head == "$PhysicalNames") {
        int n = to_int(line());
        for (int k = 0; k < n; ++k) {
          std::string s = line();
          size_t a = s.find(' '), b = a == std::string::npos ? a : s.find(' ', a + 1);
          if (b == std::string::npos)
            throw FatalError(AFEM_ERR_ARGUMENT, "corrupt mesh file: physical name line '" + s + "'");
          int d = to_int(s.substr(0, a)), t = to_int(s.substr(a + 1, b - a - 1));
          std::string nm = s.substr(b + 1);
          if (nm.size() >= 2 && nm.front() == '"')
            nm = nm.substr(1, nm.size() - 2);
          names[{ d, t }] = nm;
        }
        line();
      }
      else if (head == "$Entities") {
        uint64_t counts[4];
        for (auto& c : counts)
          c = u64();
        for (int d = 0; d < 4; ++d)
          for (uint64_t k = 0; k < counts[d]; ++k) {
            int tag = i32();
            double box[6];
            take(box, d == 0 ? 24 : 48);
            const size_t np = fits(u64(), 4);
            std::vector<int> phys(np);
            for (auto& v : phys)
              v = i32();
            entity_phys[{ d, tag }] = phys;
            if (d > 0) {
              const size_t nb = fits(u64(), 4);
              p += 4 * nb;
            }
          }
        line();
        line();
      }
      else if (head == "$Nodes") {
        uint64_t nblocks = u64(), nnodes = u64();
        u64();
        u64();
        for (uint64_t b = 0; b < nblocks; ++b) {
          i32();
          i32();
          if (i32() != 0)
            throw FatalError(AFEM_ERR_NOT_IMPLEMENTED, "parametric Gmsh nodes");
          const size_t n = fits(u64(), 32); // a tag and three coordinates per node
          size_t t0 = tags.size();
          for (size_t k = 0; k < n; ++k)
            tags.push_back((Int64)u64());
          xyz.resize((t0 + n) * 3);
          take(xyz.data() + t0 * 3, n * 24);
        }
        if (tags.size() != nnodes)
          throw FatalError(AFEM_ERR_MESH, "node count mismatch");
        line();
        line();
      }
      else if (head == "$Elements") {
        uint64_t nblocks = u64();
        u64();
        u64();
        u64();
        for (uint64_t b = 0; b < nblocks; ++b) {
          Block blk;
          blk.dim = i32();
          blk.tag = i32();
          int type = i32();
          auto it = nodes_of.find(type);
          if (it == nodes_of.end())
            throw FatalError(AFEM_ERR_NOT_IMPLEMENTED, "Gmsh element type " + std::to_string(type));
          blk.k = it->second;
          const size_t n = fits(u64(), 8 * (uint64_t)(blk.k + 1));
          blk.conn.resize(n * (blk.k + 1));
          take(blk.conn.data(), blk.conn.size() * 8);
          blocks.push_back(std::move(blk));
        }
        line();
        line();
      }
      else {
        std::string end = "$End" + head.substr(1);
        while (line() != end) {
        }
      }
    }
    // nodes ordered by tag; local id = rank of the tag
    std::vector<size_t> order(tags.size());
    for (size_t k = 0; k < order.size(); ++k)
      order[k] = k;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return tags[a] < tags[b]; });
    std::map<Int64, Int32> lid;
    auto node_of = [&lid](uint64_t tag) {
      auto it = lid.find((Int64)tag);
      if (it == lid.end())
        throw FatalError(AFEM_ERR_MESH, "corrupt mesh file: an element names node " + std::to_string(tag) +
                                          ", which the file does not define");
      return it->second;
    };
    std::vector<Real> coords(tags.size() * 3);
    std::vector<Int64> uids(tags.size());
    for (size_t k = 0; k < order.size(); ++k) {
      lid[tags[order[k]]] = (Int32)k;
      uids[k] = tags[order[k]];
      for (int d = 0; d < 3; ++d)
        coords[3 * k + d] = xyz[3 * order[k] + d];
    }
    int dim = 0;
    for (auto& b : blocks)
      dim = std::max(dim, b.dim);
    std::vector<Int32> cells;
    m_groups.clear();
    m_face_groups.clear();
    m_face_nodes = 0;
    int npc = 0;
    for (auto& b : blocks)
      if (b.dim == dim) {
        if (npc != 0 && npc != b.k)
          throw FatalError(AFEM_ERR_NOT_IMPLEMENTED, "meshes mixing cell types are not supported");
        npc = b.k;
      }
    const int stride = npc > 4 ? npc : 4;
    m_cell_groups.clear();
    for (auto& b : blocks) {
      const size_t n = b.conn.size() / (b.k + 1);
      if (b.dim == dim) {
        const Int32 first = (Int32)(cells.size() / stride);
        for (int ph : entity_phys[{ b.dim, b.tag }]) { // volume physical names (materials)
          auto nm = names.find({ b.dim, ph });
          if (nm == names.end())
            continue;
          auto& cg = m_cell_groups[nm->second];
          for (size_t e = 0; e < n; ++e)
            cg.push_back(first + (Int32)e);
        }
        for (size_t e = 0; e < n; ++e)
          for (int s = 0; s < stride; ++s)
            cells.push_back(s < b.k ? node_of(b.conn[e * (b.k + 1) + 1 + s]) : -1);
      }
      else if (b.dim == dim - 1 || (b.dim == 0 && dim > 1)) { // faces, or named points
        for (int ph : entity_phys[{ b.dim, b.tag }]) {
          auto nm = names.find({ b.dim, ph });
          if (nm == names.end())
            continue;
          auto& g = m_groups[nm->second];
          for (size_t e = 0; e < n; ++e)
            for (int s = 0; s < b.k; ++s)
              g.push_back(node_of(b.conn[e * (b.k + 1) + 1 + s]));
          if (b.dim == dim - 1 && (b.k == dim || (dim == 3 && b.k == 4))) {
            // faces in file order: edges (2D), triangles or quadrangles (3D)
            if (m_face_nodes != 0 && m_face_nodes != b.k)
              throw FatalError(AFEM_ERR_NOT_IMPLEMENTED, "meshes mixing face types are not supported");
            m_face_nodes = b.k;
            auto& fg = m_face_groups[nm->second];
            for (size_t e = 0; e < n; ++e) {
              Int32 f[4];
              for (int s = 0; s < b.k; ++s)
                f[s] = node_of(b.conn[e * (b.k + 1) + 1 + s]);
              if (b.k == 4) { // quadrangles in Arcane's face order (smallest uid first, towards its smaller neighbour)
                const Int64* u = uids.data();
                int i0 = 0;
                for (int s = 1; s < 4; ++s)
                  if (u[f[s]] < u[f[i0]])
                    i0 = s;
                const int step = u[f[(i0 + 1) % 4]] < u[f[(i0 + 3) % 4]] ? 1 : 3;
                for (int s = 0; s < 4; ++s)
                  fg.push_back(f[(i0 + step * s) % 4]);
              }
              else {
                for (int s = 0; s < b.k; ++s)
                  fg.push_back(f[s]);
              }
            }
          }
        }
      }
    }
    for (auto& kv : m_groups) {
      std::sort(kv.second.begin(), kv.second.end());
      kv.second.erase(std::unique(kv.second.begin(), kv.second.end()), kv.second.end());
    }
    if (cells.empty() || dim < 2)
      throw FatalError(AFEM_ERR_MESH, "mesh file without 2D or 3D cells");
    build(dim, std::move(coords), std::move(cells), std::move(uids), upload, npc);
  }

  afem_mesh_t handle() const { return m_handle; }
  int dimension() const { return m_dim; }
  Int32 nbNode() const { return (Int32)(m_coords.size() / 3); }
  Int32 nbCell() const { return (Int32)(m_cells.size() / (m_npc > 4 ? m_npc : 4)); }
  int nodesPerCell() const { return m_npc; }
  // nodes per boundary face of the face groups: dimension() for simplices and
  // Quad4, 4 for Hexa8
  int nodesPerFace() const { return m_face_nodes != 0 ? m_face_nodes : (m_npc == 8 ? 4 : m_dim); }
  const std::vector<Real>& coords() const { return m_coords; }
  const std::vector<Int32>& cells() const { return m_cells; }
  const std::vector<Int64>& uniqueIds() const { return m_uids; }
  const NodeGroup& group(const std::string& name) const
  {
    auto it = m_groups.find(name);
    if (it == m_groups.end())
      throw FatalError(AFEM_ERR_ARGUMENT, "no boundary group '" + name + "'");
    return it->second;
  }
  // cells of a named volume group (a material), local cell ids
  const std::vector<Int32>& cellGroup(const std::string& name) const
  {
    auto it = m_cell_groups.find(name);
    if (it == m_cell_groups.end())
      throw FatalError(AFEM_ERR_ARGUMENT, "no cell group '" + name + "'");
    return it->second;
  }
  // faces of a named boundary group: nodesPerFace() node ids per face
  const std::vector<Int32>& faceGroup(const std::string& name) const
  {
    auto it = m_face_groups.find(name);
    if (it == m_face_groups.end())
      throw FatalError(AFEM_ERR_ARGUMENT, "no boundary face group '" + name + "'");
    return it->second;
  }
  // nodes whose coordinate `axis` equals `value` (structured boxes)
  NodeGroup nodesOnPlane(int axis, double value) const
  {
    NodeGroup g;
    for (Int32 i = 0; i < nbNode(); ++i)
      if (m_coords[3 * (size_t)i + axis] == value)
        g.push_back(i);
    return g;
  }

 private:
  int m_dim = 0;
  int m_npc = 0;
  int m_face_nodes = 0;
  std::vector<Real> m_coords;
  std::vector<Int32> m_cells;
  std::vector<Int64> m_uids;
  std::map<std::string, NodeGroup> m_groups;
  std::map<std::string, std::vector<Int32>> m_face_groups;
  std::map<std::string, std::vector<Int32>> m_cell_groups;
  afem_mesh_t m_handle = nullptr;
};

} // namespace afem::femutils
