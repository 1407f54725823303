"""GPU edge cases of the path: the smallest meshes, a node that belongs to no
cell, a node whose row is longer than any assembly plan allows (the column
search kernel runs), a zero right-hand side, and a system whose DoFs are all
eliminated.  Checked against the oracle (pattern bit-exact, values within
1e-12 of the row's largest entry), scipy and closed forms."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O

pytestmark = pytest.mark.gpu

ROW_TOL = 1e-12
N_LAPLACE_VARIANTS = 7


def row_scaled_error(row, ref, got):
    lens = np.diff(row)
    rid = np.repeat(np.arange(len(lens)), lens)
    scale = np.zeros(len(lens))
    np.maximum.at(scale, rid, np.abs(ref))
    s = scale[rid]
    return (np.abs(got - ref) / np.where(s > 0, s, 1.0)).max() if len(ref) else 0.0


def tri_cells(t):
    c = np.full((len(t), 4), -1, np.int32)
    c[:, :3] = t
    return c


def check_laplace_all_variants(A, dim, X, C, lam=None):
    """Pattern bit-exact; every assembly variant (with its fallbacks) and the
    RHS against the oracle."""
    N = X.shape[0]
    m = A.Mesh(dim, X, C)
    b = A.BSRFormat(m, 1, use_csr=True)
    b.compute_sparsity()
    orow, ocol = O.pattern(N, C, dim + 1)
    row, col, _, _ = b.download()
    assert np.array_equal(row, orow) and np.array_equal(col, ocol)
    oval = O.assemble(dim, X, C, dim + 1, orow, ocol, cell_lambda=lam)
    orhs = O.rhs_constant_source(dim, X, C, dim + 1, 1.5)
    dlam = A.DeviceArray.from_numpy(lam) if lam is not None else None
    for v in range(-1, N_LAPLACE_VARIANTS):
        b.set_assembly_variant(v)
        rhs = A.DeviceArray(N, np.float64)
        b.assemble_laplace(dlam, 1.5, rhs)
        _, _, val, _ = b.download()
        assert row_scaled_error(orow, oval, val) <= ROW_TOL, v
        assert np.abs(rhs.numpy() - orhs).max() <= 1e-12 * max(np.abs(orhs).max(), 1e-300), v
    return m, b


def test_single_tetrahedron_and_triangle(afem_gpu):
    A = afem_gpu
    print("[case] single tet / single triangle", flush=True)
    X = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    check_laplace_all_variants(A, 3, X, np.array([[0, 1, 2, 3]], np.int32), lam=np.array([3.0]))
    X2 = np.array([[0.0, 0, 0], [2, 0, 0], [0, 1, 0]])
    check_laplace_all_variants(A, 2, X2, tri_cells([[0, 1, 2]]))


def test_isolated_node(afem_gpu):
    """A node of no cell: an empty row in the reference's pattern, and no
    contribution anywhere."""
    A = afem_gpu
    print("[case] isolated node", flush=True)
    X = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [5, 5, 5]])
    C = np.array([[0, 1, 2, 3], [1, 2, 3, 4]], np.int32)
    check_laplace_all_variants(A, 3, X, C)


def fan_mesh(k):
    """Node 0 at the centre of k triangles around the unit circle: its row
    has k + 1 entries."""
    ang = 2 * np.pi * np.arange(k) / k
    X = np.zeros((k + 1, 3))
    X[1:, 0], X[1:, 1] = np.cos(ang), np.sin(ang)
    t = [[0, 1 + i, 1 + (i + 1) % k] for i in range(k)]
    return X, tri_cells(t)


@pytest.mark.parametrize("k", [40, 60, 300])
def test_high_valence_node(afem_gpu, k):
    """Rows longer than 32 (no local plans) and than 255 (no plan at all:
    the column-search kernel) still assemble and solve.  The pattern of the
    40-cell fan comes from the sixteen-lane sorting kernel (7 candidates per
    lane), those of 60 and 300 cells from the per-thread fallback (128 and
    512 slots)."""
    A = afem_gpu
    print(f"[case] fan k={k}", flush=True)
    X, C = fan_mesh(k)
    m, b = check_laplace_all_variants(A, 2, X, C)
    N = X.shape[0]
    s = A.DoFLinearSystem(N)
    b.set_assembly_variant(-1)
    b.assemble_laplace(None, 1.0, s.rhs())
    s.set_bsr(b)
    ring = np.arange(1, N, dtype=np.int32)
    s.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, ring, 1, 0, 0.0, 0.0)
    s.apply_matrix_transformation()
    s.apply_rhs_transformation()
    row, col, val, _ = b.download()
    S = sp.csr_matrix((val, col, row), shape=(N, N))
    rhs = s.rhs().numpy()
    ref = spla.spsolve(S.tocsc(), rhs)
    for method in (A.SOLVER_PCG_FP64, A.SOLVER_PCG_MIXED):
        r = s.solve(method=method, rtol=1e-10, max_iter=100)
        assert r.converged and r.relres <= 1e-10
        assert np.abs(s.solution().numpy() - ref).max() <= 1e-8 * np.abs(ref).max()


@pytest.mark.parametrize("method", ["fp64", "mixed", "cg"])
def test_zero_rhs_gives_zero_solution(afem_gpu, method):
    A = afem_gpu
    print(f"[case] zero rhs {method}", flush=True)
    X, C = A.box_mesh(3, 4, 0.15, seed=2)
    N = X.shape[0]
    m = A.Mesh(3, X, C)
    b = A.BSRFormat(m, 1, use_csr=True)
    b.compute_sparsity()
    s = A.DoFLinearSystem(N)
    b.assemble_laplace(None, 0.0, s.rhs())
    s.set_bsr(b)
    bottom = np.nonzero(X[:, 2] == 0.0)[0].astype(np.int32)
    s.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, bottom, 1, 0, 0.0, 0.0)
    s.apply_matrix_transformation()
    s.apply_rhs_transformation()
    code = {"fp64": A.SOLVER_PCG_FP64, "mixed": A.SOLVER_PCG_MIXED, "cg": A.SOLVER_CG_FP64}[method]
    r = s.solve(method=code, rtol=1e-8, max_iter=100)
    assert r.converged and r.iterations == 0
    x = s.solution().numpy()
    assert np.all(np.isfinite(x)) and np.abs(x).max() == 0.0


@pytest.mark.parametrize("method", ["fp64", "mixed"])
def test_every_dof_eliminated(afem_gpu, method):
    """Row elimination of every DoF turns the system into x = g."""
    A = afem_gpu
    print(f"[case] all dofs eliminated {method}", flush=True)
    X, C = A.box_mesh(2, 3, 0.2, seed=4)
    N = X.shape[0]
    m = A.Mesh(2, X, C)
    b = A.BSRFormat(m, 1, use_csr=True)
    b.compute_sparsity()
    s = A.DoFLinearSystem(N)
    b.assemble_laplace(None, 1.0, s.rhs())
    s.set_bsr(b)
    s.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, np.arange(N, dtype=np.int32), 1, 0, 2.5, 0.0)
    s.apply_matrix_transformation()
    s.apply_rhs_transformation()
    code = {"fp64": A.SOLVER_PCG_FP64, "mixed": A.SOLVER_PCG_MIXED}[method]
    r = s.solve(method=code, rtol=1e-10, max_iter=10)
    assert r.converged
    assert np.abs(s.solution().numpy() - 2.5).max() <= 1e-12


@pytest.mark.parametrize("nbytes", [16, 16 * 1024, 16 * 1024 + 16, (1 << 20) + 48])
def test_stream_copy_exact(afem_gpu, nbytes):
    """The reference copy the roofline figures are set beside: one 16 KB chunk
    per workgroup, a partial last chunk, every byte copied and none beyond."""
    A = afem_gpu
    print(f"[case] stream copy {nbytes} B", flush=True)
    rng = np.random.default_rng(nbytes)
    src = rng.integers(0, 256, nbytes + 64, dtype=np.uint8)
    d_src = A.DeviceArray.from_numpy(src)
    d_dst = A.DeviceArray.from_numpy(np.full(nbytes + 64, 7, np.uint8))
    A.check(A.lib().afem_stream_copy(d_dst.c, d_src.c, nbytes))
    out = d_dst.numpy()
    assert np.array_equal(out[:nbytes], src[:nbytes])
    assert np.all(out[nbytes:] == 7)


def test_stream_copy_rejects_unaligned(afem_gpu):
    A = afem_gpu
    d = A.DeviceArray(64, np.uint8)
    assert A.lib().afem_stream_copy(d.c, d.c, 24) == A.ERR_ARGUMENT
