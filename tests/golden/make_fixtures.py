"""Build the golden fixtures under tests/golden/ from the reference's own data.

Run in a checkout that has the arcanefem sources at /root/reference (or pass
--reference).  It reads three binary Gmsh 4.1 meshes and the reference's
golden solution files, and writes one .npz per case with:

  coords      (N, 3) float64   node coordinates, nodes ordered by Gmsh tag
  node_uid    (N,)   int64     Gmsh node tag = Arcane node uniqueId
  cells       (C, s) int32     cell -> local node ids, s = 4 (triangles: slot 3 = -1;
                               tets; quads) or 8 (hexas)
  npc         ()     int32     nodes per cell (3, 4 or 8)
  dim         ()     int32
  group_<name>(k,)   int32     local ids of the nodes of the boundary faces
                               carrying the physical name <name>
  faces_<name>(f, dim) int32   those faces (local node ids, file order)
  cellgroup_<name> (c,) int32  cells carrying the volume physical name <name>
  ref_uid     (M,)   int64     uids listed in the golden file
  ref_value   (M, d) float64   golden values (d = 1 or 3)
  ref_uid_<t>, ref_value_<t>   further golden files on the same mesh (EXTRA)

Golden cases (test names and tolerances from the reference's CMake files):
  circle_2D   modules/poisson/inputs/circle.2D.arc, check/poisson_test_ref_circle_2D.txt
              (f=5.5, Dirichlet 0.5 on 'horizontal', eps 1e-4: FemModule.cc:508)
  sphere_3D   modules/poisson/inputs/sphere.3D.arc, check/poisson_test_ref_sphere_3D.txt
  bar_3D      modules/elasticity/inputs/bar.3D.Dirichlet.bodyForce.arc,
              check/bar.3D.Dirichlet.bodyForce.txt (eps 1e-3, min 1e-10)
  circle_neumann_2D  modules/poisson/inputs/circle.neumann.2D.arc,
              check/poisson_test_ref_circle_neumann_2D.txt (f=5.5, Dirichlet 0.5 on
              'horizontal', Neumann flux (-0.35, 1.65) on 'curved', eps 1e-4)
  sphere_neumann_3D  modules/poisson/inputs/sphere.neumann.3D.arc,
              check/poisson_test_ref_sphere_neumann_3D.txt (flux (0.35, 1.65, 3.75))
  point_dirichlet_2D modules/poisson/inputs/perforatedSquare.pointDirichlet.2D.arc,
              check/poisson_test_point_dirichlet_2D.txt (no source; penalty at the
              points topLeftCorner 50, topRightCorner 20, botLeftCorner 20,
              botRightCorner 50)
  heat_plate_2D  modules/heat/inputs/conduction.arc, check/2d_conduction.txt (lambda 1.75,
              dt 0.4, tmax 20, Tinit 30, penalty 1e31 Dirichlet 10 on 'left', eps 1e-4)
  heat_cube_3D   modules/heat/inputs/3d_conduction.arc, check/3d_conduction.txt (lambda 1.75,
              dt 0.1, tmax 1, Tinit 30, Dirichlet 100.6 on 'top' and 1.6 on 'bottom')
  heat_convection_2D  modules/heat/inputs/conduction.convection.arc,
              check/2d_conduction_convection.txt (as heat_plate_2D plus convection
              h 1, Text 20 on 'right', 'top', 'bottom')
  heat_neumann_point_2D  modules/heat/inputs/conduction.neumann.point-dirichlet.arc,
              check/2d_conduction_neumann_pointBC.txt (penalty 1e31 at the points
              topLeft 1.8 and botRight 31, Neumann 9.6 on 'left')
  heat_convection_point_3D  modules/heat/inputs/3d_conduction.pointBc.convection.arc,
              check/3d_conduction_convection_pointBC.txt (convection h 1.9, Text 30 on
              'top', Dirichlet 2.6 on 'bottom', point 'center' 100.6, penalty 1e30)

  circle_2D_quad  modules/poisson/inputs/circle.2D.quad.arc,
              check/poisson_test_ref_circle_2D_quad.txt (Quad4; f=5.5, Dirichlet 0.5 on
              'horizontal', penalty)
  sphere_3D_hexa  modules/poisson/inputs/sphere.3D.hexa.arc,
              check/poisson_test_ref_sphere_3D_hexa.txt (Hexa8; the same conditions)
  circle_2D_quad extra 'neumann': inputs/circle.neumann.2D.quad.arc,
              check/poisson_test_ref_circle_neumann_2D_quad.txt (flux (-0.35, 1.65) on
              'curved'); 'scalar_neumann': the same input with value 2,
              check/poisson_test_ref_circle_scalar_neumann_2D_quad.txt
  sphere_3D_hexa extra 'neumann': inputs/sphere.neumann.3D.hexa.arc,
              check/poisson_test_ref_sphere_neumann_3D_hexa.txt (flux (0.35, 1.65, 3.75));
              'scalar_neumann': value 3.1, check/poisson_test_ref_sphere_scalar_neumann_3D_hexa.txt
  cube_3D_hexa8  modules/poisson/inputs/cube.3D.hexa.arc on meshes/3x3x3_cube_hexa8.msh,
              check/poisson_test_ref_cube_3D_hexa8.txt (f=9.8, Dirichlet 0.5 on 'left',
              scalar Neumann 13.9 on 'right')

  heat_plate_2D_quad  modules/heat/inputs/conduction.quad.arc (plate.quad.msh),
              check/2d_conduction.quad.txt (lambda 1.75, dt 0.4, tmax 20, Tinit 30,
              penalty 1e31 Dirichlet 10 on 'left'); extra 'convection':
              conduction.convection.quad.arc (plus h 5, Text 15 on 'right'); extra
              'neumann_point': conduction.neumann.point-dirichlet.quad.arc (points
              topLeft 1.8, botRight 31, Neumann 9.6 on 'left')
  heat_cube_3D_hexa  modules/heat/inputs/3d_conduction.convection.hexa.arc
              (truncated_cube.hexa.msh), check/3d_conduction_convection.hexa.txt
              (dt 0.4, tmax 20, Dirichlet 10 on 'left' with penalty 1e30, h 5, Text 15
              on 'right'); extra 'neumann': 3d_conduction.neumann.hexa.arc (Dirichlet
              10.6 on 'top', Neumann 1.2 on 'bottom')

  fourier_*  modules/fourier (-div(lambda grad u) = qdot): inputs/conduction.arc
              on plancher.msh (extra 'fourier' of point_dirichlet_2D, check/test1_results.txt:
              lambda 1.75, qdot 1e5, penalty 1e12 Dirichlet Cercle 50, Bas 5, Haut 21,
              Neumann Droite 15, Gauche 0); conduction.heterogeneous.arc on
              multi-material.msh (fourier_multimat_2D, test2_results.txt: lambda 100 on
              Mat1 and 1 on Mat2, qdot 15, Dirichlet Left 50, Right 5); conduction.3D.arc
              (extra 'fourier' of bar_3D, test_conduction_3D.txt: lambda 0.023, qdot
              1.123e-2, Dirichlet surfaceleft 55, surfaceright 12); conduction.quad.arc on
              plancher.quad4.msh (fourier_plancher_quad, conduction_quad.txt);
              conduction.heterogeneous.quad.arc (fourier_multimat_quad,
              conduction_multi-mat_quad.txt); conduction.sphere-cut.hexa.arc (extra
              'fourier' of sphere_3D_hexa, conduction_hexa.txt: lambda 23.5, qdot 1.123e-2,
              Dirichlet horizontal 55, Neumann curved 1003.67)

  bar_2D  modules/elasticity/inputs/bar.2D.Dirichlet.bodyForce.arc (bar.msh, tri3, 2 DoFs
              per node; E 21e5, nu 0.28, f (NULL, -1), u = 0 on 'left'),
              check/bar.2D.Dirichlet.bodyForce.txt; extras 'traction' (bar.2D.Dirichlet.traction
              .arc: traction (1, NULL) on 'right', no body force), 'traction_bodyforce'
              (bar.2D.traction.bodyforce.arc: f (3.33, -6.66), traction (1.33, 2.13)), 'point'
              (bar.2D.PointDirichlet.Dirichlet.bodyForce.arc: ux = 0 on 'left', ux = 1 on
              'right', u = 0 at botLeft, uy = 0 at botRight)

  elasticity_plate_quad / elasticity_five_quads / elasticity_cube_hexa  modules/elasticity
              Quad4 / Hexa8 inputs: bar.2D.Dirichlet.bodyForce.quad.arc and
              bar.2D.traction.bodyforce.quad.arc (plate.quad.msh), 2D.dirichlet.bodyforce.quad.arc
              and 2D.dirichlet.traction.bodyforce.quad.arc (five_quads.msh),
              3D.dirichlet.bodyforce.hexa.arc and 3D.dirichlet.traction.bodyforce.hexa.arc
              (truncated_cube.hexa.msh); the second of each pair is extra 'traction'

The raw .msh files and the golden solution files are copied beside the .npz
files as data fixtures (the C++ module driver reads them directly).
"""

import argparse
import os
import shutil
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = {
    "circle_2D": ("meshes/msh/circle_cut.msh", "modules/poisson/check/poisson_test_ref_circle_2D.txt", 1),
    "sphere_3D": ("meshes/msh/sphere_cut.msh", "modules/poisson/check/poisson_test_ref_sphere_3D.txt", 1),
    "bar_3D": ("meshes/msh/bar_dynamic_3D.msh", "modules/elasticity/check/bar.3D.Dirichlet.bodyForce.txt", 3),
    "circle_neumann_2D": ("meshes/msh/circle_cut.msh", "modules/poisson/check/poisson_test_ref_circle_neumann_2D.txt",
                          1),
    "sphere_neumann_3D": ("meshes/msh/sphere_cut.msh", "modules/poisson/check/poisson_test_ref_sphere_neumann_3D.txt",
                          1),
    "point_dirichlet_2D": ("meshes/msh/plancher.msh", "modules/poisson/check/poisson_test_point_dirichlet_2D.txt", 1),
    "heat_plate_2D": ("meshes/msh/plate.msh", "modules/heat/check/2d_conduction.txt", 1),
    "heat_cube_3D": ("meshes/msh/truncated_cube.msh", "modules/heat/check/3d_conduction.txt", 1),
    "heat_convection_2D": ("meshes/msh/plate.msh", "modules/heat/check/2d_conduction_convection.txt", 1),
    "heat_neumann_point_2D": ("meshes/msh/plate.msh", "modules/heat/check/2d_conduction_neumann_pointBC.txt", 1),
    "heat_convection_point_3D": ("meshes/msh/truncated_cube.msh",
                                 "modules/heat/check/3d_conduction_convection_pointBC.txt", 1),
    "circle_2D_quad": ("meshes/msh/circle_cut.quad.msh", "modules/poisson/check/poisson_test_ref_circle_2D_quad.txt", 1),
    "sphere_3D_hexa": ("meshes/msh/sphere_cut.hexa.msh", "modules/poisson/check/poisson_test_ref_sphere_3D_hexa.txt", 1),
    "cube_3D_hexa8": ("meshes/msh/3x3x3_cube_hexa8.msh", "modules/poisson/check/poisson_test_ref_cube_3D_hexa8.txt", 1),
    "heat_plate_2D_quad": ("meshes/msh/plate.quad.msh", "modules/heat/check/2d_conduction.quad.txt", 1),
    "fourier_multimat_2D": ("meshes/msh/multi-material.msh", "modules/fourier/check/test2_results.txt", 1),
    "bar_2D": ("meshes/msh/bar.msh", "modules/elasticity/check/bar.2D.Dirichlet.bodyForce.txt", 2),
    "elasticity_plate_quad": ("meshes/msh/plate.quad.msh", "modules/elasticity/check/bar.2D.Dirichlet.bodyForce.quad.txt",
                              2),
    "elasticity_five_quads": ("meshes/msh/five_quads.msh", "modules/elasticity/check/2D.dirichlet.bodyforce.quad.txt",
                              2),
    "elasticity_cube_hexa": ("meshes/msh/truncated_cube.hexa.msh",
                             "modules/elasticity/check/3D.dirichlet.bodyforce.hexa.txt", 3),
    "fourier_multimat_quad": ("meshes/msh/multi-material.quad.msh",
                              "modules/fourier/check/conduction_multi-mat_quad.txt", 1),
    "fourier_plancher_quad": ("meshes/msh/plancher.quad4.msh", "modules/fourier/check/conduction_quad.txt", 1),
    "heat_cube_3D_hexa": ("meshes/msh/truncated_cube.hexa.msh", "modules/heat/check/3d_conduction_convection.hexa.txt",
                          1),
}

EXTRA = {
    "elasticity_plate_quad": {"traction": "modules/elasticity/check/bar.2D.traction.bodyforce.quad.txt"},
    "elasticity_five_quads": {"traction": "modules/elasticity/check/2D.dirichlet.traction.bodyforce.quad.txt"},
    "elasticity_cube_hexa": {"traction": "modules/elasticity/check/3D.dirichlet.traction.bodyforce.hexa.txt"},
    "bar_2D": {
        "traction": "modules/elasticity/check/bar.2D.Dirichlet.traction.txt",
        "traction_bodyforce": "modules/elasticity/check/bar.2D.traction.bodyforce.txt",
        "point": "modules/elasticity/check/bar.2D.PointDirichlet.Dirichlet.bodyForce.txt",
    },
    "point_dirichlet_2D": {"fourier": "modules/fourier/check/test1_results.txt"},
    "bar_3D": {"fourier": "modules/fourier/check/test_conduction_3D.txt"},
    "circle_2D_quad": {
        "neumann": "modules/poisson/check/poisson_test_ref_circle_neumann_2D_quad.txt",
        "scalar_neumann": "modules/poisson/check/poisson_test_ref_circle_scalar_neumann_2D_quad.txt",
    },
    "heat_plate_2D_quad": {
        "convection": "modules/heat/check/2d_conduction_convection.quad.txt",
        "neumann_point": "modules/heat/check/2d_conduction_neumann_pointBC.quad.txt",
    },
    "heat_cube_3D_hexa": {
        "neumann": "modules/heat/check/3d_conduction_neumann.hexa.txt",
    },
    "sphere_3D_hexa": {
        "fourier": "modules/fourier/check/conduction_hexa.txt",
        "neumann": "modules/poisson/check/poisson_test_ref_sphere_neumann_3D_hexa.txt",
        "scalar_neumann": "modules/poisson/check/poisson_test_ref_sphere_scalar_neumann_3D_hexa.txt",
    },
}

# Gmsh element type -> number of nodes (point, line, triangle, quad, tet, hexa)
GMSH_NODES = {15: 1, 1: 2, 2: 3, 3: 4, 4: 4, 5: 8}


class _Reader:
    def __init__(self, data):
        self.d = data
        self.p = 0

    def line(self):
        e = self.d.index(b"\n", self.p)
        s = self.d[self.p:e].decode("ascii").strip()
        self.p = e + 1
        return s

    def take(self, fmt, count=1):
        size = struct.calcsize("<" + fmt) * count
        out = np.frombuffer(self.d, dtype=np.dtype("<" + fmt), count=count, offset=self.p)
        self.p += size
        return out

    def i32(self):
        return int(self.take("i")[0])

    def u64(self):
        return int(self.take("Q")[0])


def read_msh41_binary(path):
    """Binary Gmsh 4.1 reader (size_t = 8) for point/line/triangle/tet elements."""
    with open(path, "rb") as f:
        r = _Reader(f.read())
    names = {}
    entity_phys = {}  # (dim, tag) -> list of physical tags
    nodes_tag = []
    nodes_xyz = []
    blocks = []  # (dim, entity_tag, etype, tags, conn)
    while r.p < len(r.d):
        head = r.line()
        if head == "":
            continue
        if head == "$MeshFormat":
            ver, ftype, dsize = r.line().split()
            assert ver == "4.1" and ftype == "1" and dsize == "8", (ver, ftype, dsize)
            one = r.i32()
            assert one == 1, "endianness"
            r.p += 1  # newline after the int
            assert r.line() == "$EndMeshFormat"
        elif head == "$PhysicalNames":
            n = int(r.line())
            for _ in range(n):
                parts = r.line().split(maxsplit=2)
                names[(int(parts[0]), int(parts[1]))] = parts[2].strip('"')
            assert r.line() == "$EndPhysicalNames"
        elif head == "$Entities":
            counts = [r.u64() for _ in range(4)]
            for dim, cnt in enumerate(counts):
                for _ in range(cnt):
                    tag = r.i32()
                    if dim == 0:
                        r.take("d", 3)
                    else:
                        r.take("d", 6)
                    nphys = r.u64()
                    phys = [int(v) for v in r.take("i", nphys)] if nphys else []
                    entity_phys[(dim, tag)] = phys
                    if dim > 0:
                        nb = r.u64()
                        if nb:
                            r.take("i", nb)
            r.line()
            assert r.line() == "$EndEntities"
        elif head == "$Nodes":
            nblocks, nnodes, _, _ = (r.u64() for _ in range(4))
            for _ in range(nblocks):
                _edim = r.i32()
                _etag = r.i32()
                param = r.i32()
                assert param == 0, "parametric nodes unsupported"
                n = r.u64()
                tags = r.take("Q", n).astype(np.int64)
                xyz = r.take("d", 3 * n).reshape(n, 3)
                nodes_tag.append(tags)
                nodes_xyz.append(xyz)
            r.line()
            assert r.line() == "$EndNodes"
            assert sum(len(t) for t in nodes_tag) == nnodes
        elif head == "$Elements":
            nblocks, _, _, _ = (r.u64() for _ in range(4))
            for _ in range(nblocks):
                edim = r.i32()
                etag = r.i32()
                etype = r.i32()
                n = r.u64()
                k = GMSH_NODES[etype]
                raw = r.take("Q", n * (1 + k)).reshape(n, 1 + k).astype(np.int64)
                blocks.append((edim, etag, etype, raw[:, 0], raw[:, 1:]))
            r.line()
            assert r.line() == "$EndElements"
        else:
            # skip an unknown section
            end = "$End" + head[1:]
            while r.line() != end:
                pass
    tags = np.concatenate(nodes_tag)
    xyz = np.concatenate(nodes_xyz)
    order = np.argsort(tags, kind="stable")
    tags = tags[order]
    xyz = xyz[order]
    return names, entity_phys, tags, xyz, blocks


def build_case(ref_root, name, mesh_rel, check_rel, ncomp, out_dir):
    mesh_path = os.path.join(ref_root, mesh_rel)
    names, entity_phys, tags, xyz, blocks = read_msh41_binary(mesh_path)
    tag_to_lid = {int(t): i for i, t in enumerate(tags)}
    dim = max(b[0] for b in blocks)
    cells, cell_groups, ncell = [], {}, 0
    for edim, etag, _etype, _etags, conn in blocks:
        if edim == dim:
            cells.append(conn)
            for ptag in entity_phys.get((edim, etag), []):
                gname = names.get((edim, ptag))
                if gname is not None:
                    cell_groups.setdefault(gname, []).append(np.arange(ncell, ncell + len(conn)))
            ncell += len(conn)
    cells = np.concatenate(cells)
    npc = cells.shape[1]
    lid = np.vectorize(tag_to_lid.__getitem__, otypes=[np.int64])(cells)
    cells4 = -np.ones((lid.shape[0], max(4, npc)), dtype=np.int32)
    cells4[:, :npc] = lid
    groups, faces = {}, {}
    for edim, etag, _etype, _etags, conn in blocks:
        if edim != dim - 1 and not (edim == 0 and dim > 1):  # faces, or named points
            continue
        for ptag in entity_phys.get((edim, etag), []):
            gname = names.get((edim, ptag))
            if gname is None:
                continue
            ids = np.vectorize(tag_to_lid.__getitem__, otypes=[np.int64])(conn.ravel())
            groups.setdefault(gname, []).append(ids)
            if edim == dim - 1:
                faces.setdefault(gname, []).append(ids.reshape(len(conn), -1))
    payload = {
        "coords": xyz.astype(np.float64),
        "node_uid": tags.astype(np.int64),
        "cells": cells4,
        "npc": np.int32(npc),
        "dim": np.int32(dim),
    }
    for gname, lst in groups.items():
        payload["group_" + gname] = np.unique(np.concatenate(lst)).astype(np.int32)
    for gname, lst in faces.items():
        payload["faces_" + gname] = np.concatenate(lst).astype(np.int32)
    for gname, lst in cell_groups.items():
        payload["cellgroup_" + gname] = np.concatenate(lst).astype(np.int32)
    ref = np.loadtxt(os.path.join(ref_root, check_rel), ndmin=2)
    payload["ref_uid"] = ref[:, 0].astype(np.int64)
    payload["ref_value"] = ref[:, 1:1 + ncomp].astype(np.float64)
    for tag, rel in EXTRA.get(name, {}).items():
        ex = np.loadtxt(os.path.join(ref_root, rel), ndmin=2)
        payload["ref_uid_" + tag] = ex[:, 0].astype(np.int64)
        payload["ref_value_" + tag] = ex[:, 1:1 + ncomp].astype(np.float64)
        shutil.copyfile(os.path.join(ref_root, rel), os.path.join(out_dir, os.path.basename(rel)))
    np.savez(os.path.join(out_dir, name + ".npz"), **payload)
    shutil.copyfile(mesh_path, os.path.join(out_dir, os.path.basename(mesh_rel)))
    shutil.copyfile(os.path.join(ref_root, check_rel), os.path.join(out_dir, os.path.basename(check_rel)))
    print(f"{name}: N={len(tags)} cells={cells4.shape[0]} npc={npc} groups={sorted(groups)}")


# files used as they are: the manufactured-solution mesh (no golden file) and
# the golden files of the cases on Arcane's generated Cartesian meshes
MESH_ONLY = ["meshes/msh/mms.quad4_n4.msh",
             "modules/elasticity/check/bar.2D.cartesian.Dirichlet.bodyForce.txt",
             "modules/elasticity/check/bar.3D.cartesian.Dirichlet.bodyForce.txt"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--out", default=HERE)
    args = ap.parse_args()
    for name, (mesh_rel, check_rel, ncomp) in CASES.items():
        build_case(args.reference, name, mesh_rel, check_rel, ncomp, args.out)
    for rel in MESH_ONLY:
        shutil.copyfile(os.path.join(args.reference, rel), os.path.join(args.out, os.path.basename(rel)))


if __name__ == "__main__":
    main()
