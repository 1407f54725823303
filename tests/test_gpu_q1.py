"""GPU parity of the Quad4 / Hexa8 (Q1) path through the C ABI against the
oracle: node->cell lists and sparsity pattern bit-exact; Laplace values and
constant-source RHS within 1e-12 of the row's / vector's largest entry (both
assembly forms); Neumann on edges and quadrangles; heat and elasticity; the
reference's Q1 golden files (circle_2D_quad, sphere_3D_hexa: f = 5.5,
u = 0.5 on 'horizontal', eps 1e-4) through a GPU solve; the element kernels
that have no Q1 form refuse Q1 meshes."""

import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle as O

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

ROW_TOL = 1e-12


def row_scaled_error(row, ref, got):
    lens = np.diff(row)
    rid = np.repeat(np.arange(len(lens)), lens)
    scale = np.zeros(len(lens))
    np.maximum.at(scale, rid, np.abs(ref))
    s = scale[rid]
    return (np.abs(got - ref) / np.where(s > 0, s, 1.0)).max() if len(ref) else 0.0


def q1_mesh(spec):
    import arcanefem_amd as A
    name, dim, n, jitter = spec
    if dim is None:
        d = np.load(os.path.join(GOLDEN, name + ".npz"))
        return d, int(d["dim"]), d["coords"], d["cells"], int(d["npc"])
    X, C = A.box_mesh_q1(dim, n, jitter, seed=13)
    return None, dim, X, C, C.shape[1]


SPECS = [("quad_box", 2, 19, 0.25), ("hexa_box", 3, 7, 0.2), ("circle_2D_quad", None, None, None),
         ("sphere_3D_hexa", None, None, None)]


@pytest.mark.parametrize("spec", SPECS, ids=[s[0] for s in SPECS])
def test_q1_pattern_values_rhs_match_oracle(afem_gpu, spec):
    A = afem_gpu
    print(f"[case] q1 parity {spec[0]}", flush=True)
    _, dim, X, C, npc = q1_mesh(spec)
    N = X.shape[0]
    m = A.Mesh(dim, X, C, npc=npc)
    off, items = m.node_cells()
    ooff, oitems = O.node_cells(N, C, npc)
    assert np.array_equal(off, ooff) and np.array_equal(items, oitems)
    b = A.BSRFormat(m, 1, use_csr=True)
    b.compute_sparsity()
    orow, ocol = O.pattern(N, C, npc)
    row, col, _, _ = b.download()
    assert np.array_equal(row, orow) and np.array_equal(col, ocol)
    lam = np.random.default_rng(5).uniform(0.2, 5.0, C.shape[0])
    oval = O.assemble(dim, X, C, npc, orow, ocol, cell_lambda=lam)
    orhs = O.rhs_constant_source(dim, X, C, npc, 2.5)
    dlam = A.DeviceArray.from_numpy(lam)
    # default: element matrices built once per cell, then gathered per row;
    # variant 0: element rows rebuilt per (node, cell) pair
    for variant in (-1, 0, -1):
        b.set_assembly_variant(variant)
        rhs = A.DeviceArray(N, np.float64)
        b.assemble_laplace(dlam, 2.5, rhs)
        _, _, val, _ = b.download()
        assert row_scaled_error(orow, oval, val) <= ROW_TOL, variant
        assert np.abs(rhs.numpy() - orhs).max() <= 1e-12 * np.abs(orhs).max(), variant
    # the standalone source kernel, with a per-cell scale, into component 1 of 2
    scale = np.random.default_rng(6).uniform(0.5, 1.5, C.shape[0])
    rhs2 = A.DeviceArray(2 * N, np.float64)
    rhs2.zero()
    A.check(A.lib().afem_rhs_constant_source(m.h, 1.5, A.DeviceArray.from_numpy(scale).c, 2, 1, rhs2.c))
    o2 = O.rhs_constant_source(dim, X, C, npc, 1.5, nb_dof=2, dof_index=1, cell_scale=scale)
    assert np.abs(rhs2.numpy() - o2).max() <= 1e-12 * np.abs(o2).max()


@pytest.mark.parametrize("dim,n", [(2, 24), (3, 8)])
@pytest.mark.parametrize("method", ["fp64", "mixed"])
def test_q1_solve_matches_direct(afem_gpu, dim, n, method):
    """u = 0 on x = 0, f = 1: PCG on the GPU against a direct solve of the
    oracle's system (1e-8 relative)."""
    A = afem_gpu
    print(f"[case] q1 solve dim={dim} {method}", flush=True)
    X, C = A.box_mesh_q1(dim, n, 0.2, seed=3)
    N, npc = X.shape[0], C.shape[1]
    m = A.Mesh(dim, X, C, npc=npc)
    b = A.BSRFormat(m, 1, use_csr=True)
    b.compute_sparsity()
    s = A.DoFLinearSystem(N)
    b.assemble_laplace(None, 1.0, s.rhs())
    s.set_bsr(b)
    left = np.nonzero(X[:, 0] == 0.0)[0].astype(np.int32)
    s.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, left, 1, 0, 0.0, 0.0)
    s.apply_matrix_transformation()
    s.apply_rhs_transformation()
    code = A.SOLVER_PCG_FP64 if method == "fp64" else A.SOLVER_PCG_MIXED
    r = s.solve(method=code, rtol=1e-11, max_iter=20000)
    assert r.converged
    row, col = O.pattern(N, C, npc)
    val = O.assemble(dim, X, C, npc, row, col)
    rhs = O.rhs_constant_source(dim, X, C, npc, 1.0)
    ei, ev = np.zeros(N, np.uint8), np.zeros(N)
    ei[left] = 1
    O.apply_transformations(row, col, val, rhs, np.zeros(N, np.uint8), np.zeros(N), ei, ev)
    ref = spl.spsolve(sp.csr_matrix((val, col, row), shape=(N, N)).tocsc(), rhs)
    x = s.solution().numpy()
    assert np.abs(x - ref).max() <= 1e-8 * np.abs(ref).max()


def by_uid(d, x):
    mp = dict(zip(d["node_uid"].tolist(), x))
    return np.array([mp[u] for u in d["ref_uid"].tolist()])


@pytest.mark.parametrize("case", ["circle_2D_quad", "sphere_3D_hexa"])
@pytest.mark.parametrize("method", ["fp64", "mixed"])
def test_q1_golden_poisson_gpu(afem_gpu, case, method):
    """modules/poisson tests 2D_quad / 3D_hexa: f = 5.5, u = 0.5 on
    'horizontal' (row elimination), eps 1e-4."""
    A = afem_gpu
    print(f"[case] q1 golden {case} {method}", flush=True)
    d, dim, X, C, npc = q1_mesh((case, None, None, None))
    N = X.shape[0]
    m = A.Mesh(dim, X, C, npc=npc)
    b = A.BSRFormat(m, 1, use_csr=True)
    b.compute_sparsity()
    s = A.DoFLinearSystem(N)
    b.assemble_laplace(None, 5.5, s.rhs())
    s.set_bsr(b)
    s.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, d["group_horizontal"], 1, 0, 0.5, 0.0)
    s.apply_matrix_transformation()
    s.apply_rhs_transformation()
    code = A.SOLVER_PCG_FP64 if method == "fp64" else A.SOLVER_PCG_MIXED
    r = s.solve(method=code, rtol=1e-12, max_iter=20000)
    assert r.converged
    assert O.count_differences(d["ref_value"][:, 0], by_uid(d, s.solution().numpy()), 1e-4) == 0


def test_q1_unsupported_kernels_refuse(afem_gpu):
    A = afem_gpu
    print("[case] q1 refusals", flush=True)
    X, C = A.box_mesh_q1(3, 3, 0.1, seed=1)
    N = X.shape[0]
    m = A.Mesh(3, X, C)
    b = A.BSRFormat(m, 1, use_csr=True)
    b.compute_sparsity()
    with pytest.raises(A.AfemError):
        b.autotune_assembly()
    b1 = A.BSRFormat(m, 1, use_csr=True)
    b1.compute_sparsity()
    with pytest.raises(A.AfemError):  # elasticity needs dim DoFs per node
        b1.assemble_elasticity(1.0, 1.0)
    X2, C2 = A.box_mesh_q1(2, 3, 0.1, seed=1)
    m2 = A.Mesh(2, X2, C2, npc=4)
    f = np.array([0.0, 0.0, -1.0])
    rhs3 = A.DeviceArray(3 * X2.shape[0], np.float64)
    with pytest.raises(A.AfemError):
        A.check(A.lib().afem_rhs_body_force(m2.h, f.ctypes.data_as(A.ctypes.c_void_p), rhs3.c))


def _box_faces(C, dim):
    loc = [(0, 1), (1, 2), (2, 3), (3, 0)] if dim == 2 else \
        [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]
    cnt = {}
    for c in C:
        for l in loc:
            f = tuple(int(v) for v in c[list(l)])
            key = tuple(sorted(f))
            n, _ = cnt.get(key, (0, f))
            cnt[key] = (n + 1, f)
    return np.array([f for n, f in cnt.values() if n == 1], np.int32)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("value", ["scalar", "vector"])
def test_q1_neumann_matches_oracle(afem_gpu, dim, value):
    """Gauss-point Neumann on the edges / quadrangles of a jittered Q1 box,
    whole boundary and one side, on top of a source term (1e-13 of the
    largest entry); a face of no cell fails and leaves the RHS untouched."""
    A = afem_gpu
    print(f"[case] q1 neumann dim={dim} {value}", flush=True)
    X, C = A.box_mesh_q1(dim, 5, 0.2, seed=7)
    N, npc = X.shape[0], C.shape[1]
    m = A.Mesh(dim, X, C, npc=npc)
    F = _box_faces(C, dim)
    g = 1.75 if value == "scalar" else [0.35, -1.65, 3.75][:dim]
    for faces in (F, F[np.all(X[F][:, :, 0] == 0.0, axis=1)]):
        rhs = A.DeviceArray(N, np.float64)
        A.check(A.lib().afem_rhs_constant_source(m.h, 5.5, None, 1, 0, rhs.c))
        A.rhs_neumann(m, faces, g, rhs)
        ref = O.rhs_constant_source(dim, X, C, npc, 5.5)
        ref = O.rhs_neumann_q1(dim, X, C, npc, faces, g, rhs=ref)
        assert np.abs(rhs.numpy() - ref).max() <= 1e-13 * np.abs(ref).max()
    before = rhs.numpy().copy()
    bad = np.array([[0, N - 1] + ([1, 2] if dim == 3 else [])], np.int32)
    with pytest.raises(A.AfemError):
        A.rhs_neumann(m, bad, g, rhs)
    assert np.array_equal(rhs.numpy(), before)


Q1_NEUMANN = [
    ("circle_2D_quad", "neumann", 5.5, "curved", [-0.35, 1.65], "horizontal", 0.5),
    ("circle_2D_quad", "scalar_neumann", 5.5, "curved", 2.0, "horizontal", 0.5),
    ("sphere_3D_hexa", "neumann", 5.5, "curved", [0.35, 1.65, 3.75], "horizontal", 0.5),
    ("sphere_3D_hexa", "scalar_neumann", 5.5, "curved", 3.1, "horizontal", 0.5),
    ("cube_3D_hexa8", "", 9.8, "right", 13.9, "left", 0.5),
]


@pytest.mark.parametrize("case,tag,f,ngroup,value,dgroup,dvalue", Q1_NEUMANN,
                         ids=[c[0] + "-" + (c[1] or "golden") for c in Q1_NEUMANN])
def test_q1_golden_neumann_gpu(afem_gpu, case, tag, f, ngroup, value, dgroup, dvalue):
    """The reference's Q1 Neumann golden files through the GPU path (faces in
    Arcane's order, Dirichlet by row elimination, FP64 PCG), eps 1e-4."""
    A = afem_gpu
    print(f"[case] q1 golden neumann {case} {tag}", flush=True)
    d, dim, X, C, npc = q1_mesh((case, None, None, None))
    N = X.shape[0]
    m = A.Mesh(dim, X, C, npc=npc)
    b = A.BSRFormat(m, 1, use_csr=True)
    b.compute_sparsity()
    s = A.DoFLinearSystem(N)
    b.assemble_laplace(None, f, s.rhs())
    A.rhs_neumann(m, O.arcane_face_order(d["faces_" + ngroup], d["node_uid"]), value, s.rhs())
    s.set_bsr(b)
    s.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, d["group_" + dgroup], 1, 0, dvalue, 0.0)
    s.apply_matrix_transformation()
    s.apply_rhs_transformation()
    r = s.solve(method=A.SOLVER_PCG_FP64, rtol=1e-12, max_iter=20000)
    assert r.converged
    sfx = "_" + tag if tag else ""
    mp = dict(zip(d["node_uid"].tolist(), s.solution().numpy()))
    got = np.array([mp[u] for u in d["ref_uid" + sfx].tolist()])
    assert O.count_differences(d["ref_value" + sfx][:, 0], got, 1e-4) == 0


@pytest.mark.parametrize("dim", [2, 3])
def test_q1_heat_pieces_match_oracle(afem_gpu, dim):
    """Q1 heat element lambda K + M/dt (per-cell lambda), the convection matrix
    on one side (edges / quadrangles), the face-uniform convection RHS and the
    Gauss mass-weighted field RHS against the oracle (1e-12 / 1e-13)."""
    A = afem_gpu
    print(f"[case] q1 heat dim={dim}", flush=True)
    X, C = A.box_mesh_q1(dim, 6 if dim == 2 else 4, 0.2, seed=17)
    N, npc, dt = X.shape[0], C.shape[1], 0.3
    m = A.Mesh(dim, X, C, npc=npc)
    b = A.BSRFormat(m, 1, use_csr=True)
    b.compute_sparsity()
    orow, ocol = O.pattern(N, C, npc)
    lam = np.random.default_rng(4).uniform(0.5, 3.0, C.shape[0])
    b.assemble_heat(A.DeviceArray.from_numpy(lam), dt)
    _, _, val, _ = b.download()
    oval = O.assemble(dim, X, C, npc, orow, ocol, element=O.HEAT, cell_lambda=lam, lame_lambda=dt)
    assert row_scaled_error(orow, oval, val) <= ROW_TOL
    side = _box_faces(C, dim)
    side = side[np.all(X[side][:, :, 0] == 1.0, axis=1)]
    b.add_convection(side, 2.5)
    O.convection_lhs(dim, X, side, 2.5, orow, ocol, oval)
    _, _, val, _ = b.download()
    assert row_scaled_error(orow, oval, val) <= ROW_TOL
    u = 20.0 + np.cos(3.0 * X[:, 0]) * X[:, 1]
    rhs = A.DeviceArray(N, np.float64)
    rhs.zero()
    A.rhs_face_uniform(m, side, 2.5 * 15.0, rhs)
    A.rhs_mass_field(m, A.DeviceArray.from_numpy(u), 1.0 / dt, rhs)
    ref = O.rhs_face_uniform(X, side, 2.5 * 15.0)
    ref = O.rhs_mass_field(X, C, npc, u, 1.0 / dt, rhs=ref, dim=dim)
    assert np.abs(rhs.numpy() - ref).max() <= 1e-13 * np.abs(ref).max()


def test_q1_long_rows_fall_back(afem_gpu):
    """A Quad4 fan whose centre row has 141 entries: the Laplace assembly takes
    the per-pair form and still matches the oracle; the heat element, which
    has only the gathered form, refuses it."""
    A = afem_gpu
    print("[case] q1 long rows", flush=True)
    k = 70
    ang = 2 * np.pi * np.arange(2 * k) / (2 * k)
    X = np.zeros((2 * k + 1, 3))
    X[1:, 0], X[1:, 1] = np.cos(ang), np.sin(ang)
    C = np.array([[0, 1 + 2 * i, 1 + (2 * i + 1) % (2 * k), 1 + (2 * i + 2) % (2 * k)] for i in range(k)], np.int32)
    N = X.shape[0]
    m = A.Mesh(2, X, C, npc=4)
    b = A.BSRFormat(m, 1, use_csr=True)
    b.compute_sparsity()
    orow, ocol = O.pattern(N, C, 4)
    rhs = A.DeviceArray(N, np.float64)
    b.assemble_laplace(None, 1.0, rhs)
    row, col, val, _ = b.download()
    assert np.array_equal(row, orow) and np.array_equal(col, ocol) and orow[1] - orow[0] == 2 * k + 1
    oval = O.assemble(2, X, C, 4, orow, ocol)
    assert row_scaled_error(orow, oval, val) <= ROW_TOL
    orhs = O.rhs_constant_source(2, X, C, 4, 1.0)
    assert np.abs(rhs.numpy() - orhs).max() <= 1e-12 * np.abs(orhs).max()
    with pytest.raises(A.AfemError):
        b.assemble_heat(None, 0.1)


@pytest.mark.parametrize("dim", [2, 3])
def test_q1_elasticity_matches_oracle(afem_gpu, dim):
    """Q1 elasticity blocks against the oracle in both value orders (1e-12 of
    the scalar row's largest entry)."""
    A = afem_gpu
    print(f"[case] q1 elasticity dim={dim}", flush=True)
    X, C = A.box_mesh_q1(dim, 5 if dim == 2 else 3, 0.2, seed=23)
    N, npc = X.shape[0], C.shape[1]
    m = A.Mesh(dim, X, C, npc=npc)
    lam, mu = O.lame(200e9, 0.3)
    orow, ocol = O.pattern(N, C, npc)
    crow, _ = O.bsr_to_csr(N, dim, orow, ocol)
    for use_csr in (True, False):
        b = A.BSRFormat(m, dim, use_csr=use_csr)
        b.compute_sparsity()
        b.assemble_elasticity(lam, mu)
        _, _, val, _ = b.download()
        oval = O.assemble(dim, X, C, npc, orow, ocol, element=O.ELASTICITY, nb_dof=dim, order_per_block=False,
                          lame_lambda=lam, lame_mu=mu)
        if not use_csr:  # per-block order: compare through the oracle's per-block layout
            oval = O.assemble(dim, X, C, npc, orow, ocol, element=O.ELASTICITY, nb_dof=dim, order_per_block=True,
                              lame_lambda=lam, lame_mu=mu)
            assert np.abs(val - oval).max() <= 1e-12 * np.abs(oval).max()
        else:
            assert row_scaled_error(crow, oval, val) <= ROW_TOL
