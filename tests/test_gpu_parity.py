"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bars (stated per test):
  * node->cell lists, sparsity pattern and DoF numbering: bit-exact;
  * assembled values and RHS: |gpu - oracle| <= 1e-12 * (largest |entry| of
    the row) per entry, FP64;
  * solutions: FP64 true relative residual <= tolerance, and the solution
    within 1e-8 (relative, max norm) of a direct solve of the same system;
  * golden files of the reference: its own tolerances (1e-4 Poisson,
    1e-3 / min 1e-10 elasticity).
"""

import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle as O

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

ROW_TOL = 1e-12


def row_scaled_error(row, ref, got, nb=1):
    """max over entries of |got - ref| / max |ref| in the same (scalar) row."""
    ref = ref.reshape(-1, nb * nb) if nb > 1 else ref
    got = got.reshape(-1, nb * nb) if nb > 1 else got
    worst = 0.0
    lens = np.diff(row)
    rid = np.repeat(np.arange(len(lens)), lens)
    scale = np.zeros(len(lens))
    np.maximum.at(scale, rid, np.abs(ref).reshape(len(rid), -1).max(axis=1))
    err = np.abs(got - ref).reshape(len(rid), -1).max(axis=1)
    s = scale[rid]
    worst = (err / np.where(s > 0, s, 1.0)).max()
    return worst


def box(dim, n, alpha, seed=11):
    import arcanefem_amd as A
    return A.box_mesh(dim, n, alpha, seed)


def load_case(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"))
    return d, d["coords"], d["cells"], int(d["npc"]), int(d["dim"])


MESHES = [("box2", 2, 17, 0.2), ("box3", 3, 7, 0.15), ("circle_2D", None, None, None), ("sphere_3D", None, None, None)]


def get_mesh(spec):
    name, dim, n, alpha = spec
    if dim is None:
        _, X, C, npc, dim = load_case(name)
        return dim, X, C
    X, C = box(dim, n, alpha)
    return dim, X, C


@pytest.mark.parametrize("spec", MESHES, ids=[m[0] for m in MESHES])
def test_node_cells_and_pattern_bit_exact(afem_gpu, spec):
    A = afem_gpu
    dim, X, C = get_mesh(spec)
    print(f"[case] pattern {spec[0]}", flush=True)
    N = X.shape[0]
    m = A.Mesh(dim, X, C)
    off, items = m.node_cells()
    ooff, oitems = O.node_cells(N, C, dim + 1)
    assert np.array_equal(off, ooff) and np.array_equal(items, oitems)
    b = A.BSRFormat(m, 1)
    b.compute_sparsity()
    row, col, _, nzr = b.download()
    orow, ocol = O.pattern(N, C, dim + 1)
    assert np.array_equal(row, orow) and np.array_equal(col, ocol)
    assert np.array_equal(nzr, np.diff(orow))
    info = b.info()
    assert info.nb_col == len(ocol) and info.max_row_length == np.diff(orow).max()


N_LAPLACE_VARIANTS = 7
CELL_ONCE_VARIANT = 6


@pytest.mark.parametrize("spec", MESHES, ids=[m[0] for m in MESHES])
@pytest.mark.parametrize("with_lambda", [False, True])
@pytest.mark.parametrize("variant", list(range(N_LAPLACE_VARIANTS)))
def test_laplace_assembly_and_fused_rhs(afem_gpu, spec, with_lambda, variant):
    A = afem_gpu
    dim, X, C = get_mesh(spec)
    print(f"[case] laplace {spec[0]} lambda={with_lambda} variant={variant}", flush=True)
    N = X.shape[0]
    lam = np.exp(np.random.default_rng(5).normal(size=C.shape[0])) if with_lambda else None
    m = A.Mesh(dim, X, C)
    b = A.BSRFormat(m, 1)
    b.compute_sparsity()
    b.set_assembly_variant(variant)
    dlam = A.DeviceArray.from_numpy(lam) if lam is not None else None
    rhs = A.DeviceArray(N, np.float64)
    b.assemble_laplace(dlam, 2.5, rhs)
    row, col, val, _ = b.download()
    oval = O.assemble(dim, X, C, dim + 1, row, col, cell_lambda=lam)
    assert row_scaled_error(row, oval, val) <= ROW_TOL
    orhs = O.rhs_constant_source(dim, X, C, dim + 1, 2.5)
    assert np.abs(rhs.numpy() - orhs).max() <= 1e-12 * np.abs(orhs).max()
    # the stand-alone source kernel (gather form of the reference GPU path)
    rhs2 = A.DeviceArray(N, np.float64)
    A.check(A.lib().afem_rhs_constant_source(m.h, 2.5, None, 1, 0, rhs2.c))
    assert np.abs(rhs2.numpy() - orhs).max() <= 1e-12 * np.abs(orhs).max()
    # repeated assembly overwrites (no stale accumulation); the cell-once
    # variant adds with LDS atomics in no fixed order: equal to rounding
    b.assemble_laplace(dlam, 0.0, None)
    _, _, val2, _ = b.download()
    if variant == CELL_ONCE_VARIANT:
        assert row_scaled_error(row, val, val2) <= 1e-14
    else:
        assert np.array_equal(val, val2)


def test_assembly_autotune_keeps_parity(afem_gpu):
    A = afem_gpu
    print("[case] assembly autotune", flush=True)
    X, C = box(3, 12, 0.15)
    m = A.Mesh(3, X, C)
    b = A.BSRFormat(m, 1)
    b.compute_sparsity()
    variant, ms = b.autotune_assembly()
    assert 0 <= variant < N_LAPLACE_VARIANTS and ms > 0
    rhs = A.DeviceArray(X.shape[0], np.float64)
    b.assemble_laplace(None, 1.0, rhs)
    row, col, val, _ = b.download()
    assert row_scaled_error(row, O.assemble(3, X, C, 4, row, col), val) <= ROW_TOL


# geometry-once block kernel (default), column-search kernel, per-use-geometry block kernel
@pytest.mark.parametrize("variant", [-1, 0, 1])
@pytest.mark.parametrize("per_block", [True, False])
@pytest.mark.parametrize("case", ["box3", "bar_3D"])
def test_elasticity_assembly(afem_gpu, per_block, case, variant):
    A = afem_gpu
    print(f"[case] elasticity {case} per_block={per_block} variant={variant}", flush=True)
    if case == "bar_3D":
        _, X, C, _, _ = load_case(case)
    else:
        X, C = box(3, 5, 0.15)
    N = X.shape[0]
    lam, mu = O.lame(21e5, 0.28)
    m = A.Mesh(3, X, C)
    b = A.BSRFormat(m, 3, use_csr=not per_block)
    b.compute_sparsity()
    b.set_assembly_variant(variant)
    b.assemble_elasticity(lam, mu)
    row, col, val, _ = b.download()
    oval = O.assemble(3, X, C, 4, row, col, element=O.ELASTICITY, nb_dof=3, order_per_block=per_block,
                      lame_lambda=lam, lame_mu=mu)
    if per_block:
        assert row_scaled_error(row, oval, val, nb=3) <= ROW_TOL
    else:
        crow, ccol = O.bsr_to_csr(N, 3, row, col)
        assert row_scaled_error(crow, oval, val) <= ROW_TOL
        grow, gcol = b.to_csr()
        assert np.array_equal(grow, crow) and np.array_equal(gcol, ccol)
    f = np.array([0.3, -1.0, 2.0])
    rhs = A.DeviceArray(3 * N, np.float64)
    rhs.zero()
    A.check(A.lib().afem_rhs_body_force(m.h, A._ptr(f), rhs.c))
    orhs = O.rhs_body_force(X, C, f)
    assert np.abs(rhs.numpy() - orhs).max() <= 1e-12 * np.abs(orhs).max()


def build_poisson(A, dim, X, C, f=1.0, lam=None):
    N = X.shape[0]
    m = A.Mesh(dim, X, C)
    b = A.BSRFormat(m, 1, use_csr=True)
    b.compute_sparsity()
    s = A.DoFLinearSystem(N)
    dlam = A.DeviceArray.from_numpy(lam) if lam is not None else None
    b.assemble_laplace(dlam, f, s.rhs())
    s.set_bsr(b)
    return m, b, s, dlam


@pytest.mark.parametrize("method", ["penalty", "row", "row_column"])
@pytest.mark.parametrize("spec", [("box3", 3, 5, 0.15), ("circle_2D", None, None, None)], ids=["box3", "circle"])
def test_dirichlet_transformations_match_oracle(afem_gpu, method, spec):
    A = afem_gpu
    dim, X, C = get_mesh(spec)
    print(f"[case] transformations {spec[0]} {method}", flush=True)
    N = X.shape[0]
    m, b, s, _ = build_poisson(A, dim, X, C, f=1.5)
    if spec[0] == "circle_2D":
        d, *_ = load_case("circle_2D")
        nodes = d["group_horizontal"]
    else:
        nodes = np.nonzero(X[:, 2] == 0.0)[0]  # includes node 0: exercises the column-0 rule
    g = 0.5
    code = {"penalty": A.DIRICHLET_PENALTY, "row": A.DIRICHLET_ROW_ELIMINATION,
            "row_column": A.DIRICHLET_ROW_COLUMN_ELIMINATION}[method]
    s.apply_dirichlet(code, nodes, 1, 0, g, 1e30)
    row, col, val0, _ = b.download()
    rhs0 = s.rhs().numpy()
    s.apply_matrix_transformation()
    s.apply_rhs_transformation()
    _, _, val, _ = b.download()
    rhs = s.rhs().numpy()
    # oracle: same starting A, b and the reference transformation rules
    orhs0 = O.rhs_constant_source(dim, X, C, dim + 1, 1.5)
    fi = np.zeros(N, np.uint8)
    fv = np.zeros(N)
    ei = np.zeros(N, np.uint8)
    ev = np.zeros(N)
    if method == "penalty":
        fi[nodes], fv[nodes] = 1, 1e30
        orhs0[nodes] = 1e30 * g
    else:
        ei[nodes] = 1 if method == "row" else 2
        ev[nodes] = g
    assert np.abs(rhs0 - orhs0).max() <= 1e-12 * np.abs(orhs0).max()
    oval = O.assemble(dim, X, C, dim + 1, row, col)
    O.apply_transformations(row, col, oval, orhs0, fi, fv, ei, ev)
    assert row_scaled_error(row, oval, val) <= ROW_TOL
    assert np.allclose(rhs, orhs0, rtol=1e-12, atol=1e-12 * np.abs(orhs0).max())


def test_block_dirichlet_transformations(afem_gpu):
    A = afem_gpu
    print("[case] block transformations", flush=True)
    X, C = box(3, 4, 0.15)
    N = X.shape[0]
    lam, mu = O.lame(21e5, 0.28)
    m = A.Mesh(3, X, C)
    b = A.BSRFormat(m, 3, use_csr=False)
    b.compute_sparsity()
    b.assemble_elasticity(lam, mu)
    s = A.DoFLinearSystem(3 * N)
    s.set_bsr(b)
    left = np.nonzero(X[:, 0] == 0.0)[0]
    for c in range(3):
        s.apply_dirichlet(A.DIRICHLET_ROW_COLUMN_ELIMINATION, left, 3, c, 0.1 * c, 0.0)
    right = np.nonzero(X[:, 0] == 1.0)[0]
    s.apply_dirichlet(A.DIRICHLET_PENALTY, right, 3, 1, 1.0, 1e30)
    row, col, val0, _ = b.download()
    rhs_in = np.random.default_rng(2).normal(size=3 * N)
    s.rhs().upload(rhs_in + s.rhs().numpy())
    rhs0 = s.rhs().numpy()
    s.apply_matrix_transformation()
    s.apply_rhs_transformation()
    _, _, val, _ = b.download()
    # oracle on the equivalent scalar CSR
    crow, ccol = O.bsr_to_csr(N, 3, row, col)
    A0 = sp.bsr_matrix((val0.reshape(-1, 3, 3), col, row), shape=(3 * N, 3 * N)).tocsr()
    cval = np.array(A0[np.repeat(np.arange(3 * N), np.diff(crow)), ccol]).ravel()
    fi = np.zeros(3 * N, np.uint8)
    fv = np.zeros(3 * N)
    ei = np.zeros(3 * N, np.uint8)
    ev = np.zeros(3 * N)
    for c in range(3):
        ei[3 * left + c], ev[3 * left + c] = 2, 0.1 * c
    fi[3 * right + 1], fv[3 * right + 1] = 1, 1e30
    orhs = rhs0.copy()
    O.apply_transformations(crow, ccol, cval, orhs, fi, fv, ei, ev)
    A1 = sp.bsr_matrix((val.reshape(-1, 3, 3), col, row), shape=(3 * N, 3 * N)).tocsr()
    gval = np.array(A1[np.repeat(np.arange(3 * N), np.diff(crow)), ccol]).ravel()
    assert row_scaled_error(crow, cval, gval) <= ROW_TOL
    assert np.allclose(s.rhs().numpy(), orhs, rtol=1e-12, atol=1e-12 * np.abs(orhs).max())


@pytest.mark.parametrize("spec", [("box2", 2, 40, 0.2), ("box3", 3, 9, 0.15)], ids=["box2", "box3"])
def test_spmv_fp64_fp32(afem_gpu, spec):
    A = afem_gpu
    dim, X, C = get_mesh(spec)
    print(f"[case] spmv {spec[0]}", flush=True)
    N = X.shape[0]
    m, b, s, _ = build_poisson(A, dim, X, C)
    row, col, val, _ = b.download()
    x = np.random.default_rng(9).normal(size=N)
    dx = A.DeviceArray.from_numpy(x)
    dy = A.DeviceArray(N, np.float64)
    s.spmv(dx, dy)
    ref = O.spmv(row, col, val, x)
    scale = np.abs(val).max() * np.abs(x).max()
    assert np.abs(dy.numpy() - ref).max() <= 1e-13 * scale
    dx32 = A.DeviceArray.from_numpy(x.astype(np.float32))
    dy32 = A.DeviceArray(N, np.float32)
    s.spmv_fp32(dx32, dy32)
    assert np.abs(dy32.numpy() - ref).max() <= 2e-5 * scale
    check_fp32_difference_form(A, s, row, col, val)


def poisson_bottom_system(A, dim, n, alpha, lam=None):
    X, C = box(dim, n, alpha)
    N = X.shape[0]
    m, b, s, dl = build_poisson(A, dim, X, C, 1.0, lam)
    axis = dim - 1
    nodes = np.nonzero(X[:, axis] == 0.0)[0]
    s.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, nodes, 1, 0, 0.0, 0.0)
    s.apply_matrix_transformation()
    s.apply_rhs_transformation()
    return X, C, m, b, s, dl


# autotuned, SELL-64, SELL-64 with 16-bit offsets (4 / 8 slots per step; 256- or
# 512-thread workgroups), CSR 8
SPMV_FORMATS = [0, -1, -2, -3, -4, -5, 8]


@pytest.mark.parametrize("fmt", [-1, -2, -3, -4, -5, 8])
def test_spmv_forced_formats(afem_gpu, fmt):
    """Each SpMV format against the oracle (the format is pinned by a one-iteration solve)."""
    A = afem_gpu
    print(f"[case] spmv format {fmt}", flush=True)
    X, C = box(3, 12, 0.15)
    N = X.shape[0]
    m, b, s, _ = build_poisson(A, 3, X, C)
    s.solve(raise_on_failure=False, method=A.SOLVER_PCG_MIXED, max_iter=1, spmv_lanes=fmt)
    want = {-1: ("sell", 64), -2: ("sell16", 64), -3: ("sell16x8", 64), -4: ("sell16", 64), -5: ("sell16x8", 64),
            8: ("csr", 8)}[fmt]
    assert s.spmv_format() == want
    if fmt <= -2:
        slots, block = (4 if fmt in (-2, -4) else 8), (512 if fmt <= -4 else 256)
        assert s.spmv_tuning() == ((slots, block), (slots, block))
    row, col, val, _ = b.download()
    x = np.random.default_rng(3).normal(size=N)
    dy = A.DeviceArray(N, np.float64)
    s.spmv(A.DeviceArray.from_numpy(x), dy)
    ref = O.spmv(row, col, val, x)
    scale = np.abs(val).max() * np.abs(x).max()
    assert np.abs(dy.numpy() - ref).max() <= 1e-13 * scale
    dy32 = A.DeviceArray(N, np.float32)
    s.spmv_fp32(A.DeviceArray.from_numpy(x.astype(np.float32)), dy32)
    assert np.abs(dy32.numpy() - ref).max() <= 2e-5 * scale
    check_fp32_difference_form(A, s, row, col, val)


def check_fp32_difference_form(A, s, row, col, val):
    """The FP32 copy is in difference form (diagonal slot = row sum, other
    slots times x_j - x_i): on a constant vector the product is the row sums
    (zero up to FP64 rounding for these Neumann operators), without the
    cancellation error of about 1e-8 of |a_ii| c that a plain FP32 product
    leaves."""
    N = row.size - 1
    c = 1000.0
    dy32 = A.DeviceArray(N, np.float32)
    s.spmv_fp32(A.DeviceArray.from_numpy(np.full(N, c, np.float32)), dy32)
    ref = O.spmv(row, col, val, np.full(N, c))
    assert np.abs(dy32.numpy() - ref).max() <= 1e-12 * np.abs(val).max() * c


@pytest.mark.parametrize("fmt", SPMV_FORMATS)
@pytest.mark.parametrize("method", ["fp64", "mixed", "cg"])
def test_pcg_solves_match_direct(afem_gpu, method, fmt):
    A = afem_gpu
    print(f"[case] pcg {method} spmv_lanes={fmt}", flush=True)
    lam = np.exp(np.random.default_rng(4).normal(size=6 * 10 ** 3))
    X, C, m, b, s, _ = poisson_bottom_system(A, 3, 10, 0.15, lam)
    N = X.shape[0]
    code = {"fp64": A.SOLVER_PCG_FP64, "mixed": A.SOLVER_PCG_MIXED, "cg": A.SOLVER_CG_FP64}[method]
    r = s.solve(method=code, rtol=1e-10, max_iter=5000, spmv_lanes=fmt)
    assert r.converged == 1 and r.relres <= 1e-10
    row, col, val, _ = b.download()
    rhs = s.rhs().numpy()
    x = s.solution().numpy()
    S = sp.csr_matrix((val, col, row), shape=(N, N))
    assert np.linalg.norm(rhs - S @ x) / np.linalg.norm(rhs) <= 1.01e-10  # recomputed on the host
    direct = spl.spsolve(S.tocsc(), rhs)
    assert np.abs(x - direct).max() <= 1e-8 * np.abs(direct).max()
    if method == "fp64":
        _, it, _ = O.pcg_jacobi(row, col, val, rhs, rtol=1e-10)
        assert abs(r.iterations - it) <= max(3, 0.05 * it)  # same algorithm, same iteration count
    if method == "mixed":
        assert r.reliable_updates >= 1


@pytest.mark.parametrize("fmt", [-1, 16])
def test_block_jacobi_pcg_elasticity(afem_gpu, fmt):
    A = afem_gpu
    print(f"[case] elasticity pcg spmv_lanes={fmt}", flush=True)
    X, C = box(3, 6, 0.15)
    N = X.shape[0]
    lam, mu = O.lame(21e5, 0.28)
    m = A.Mesh(3, X, C)
    b = A.BSRFormat(m, 3, use_csr=False)
    b.compute_sparsity()
    b.assemble_elasticity(lam, mu)
    s = A.DoFLinearSystem(3 * N)
    s.set_bsr(b)
    rhs = s.rhs()
    rhs.zero()
    A.check(A.lib().afem_rhs_body_force(m.h, A._ptr(np.array([0.0, 0.0, -1.0])), rhs.c))
    left = np.nonzero(X[:, 0] == 0.0)[0]
    for c in range(3):
        s.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, left, 3, c, 0.0, 0.0)
    s.apply_matrix_transformation()
    s.apply_rhs_transformation()
    for method in (A.SOLVER_PCG_FP64, A.SOLVER_PCG_MIXED):
        r = s.solve(method=method, rtol=1e-8, max_iter=20000, spmv_lanes=fmt)
        assert r.converged == 1 and r.relres <= 1e-8
        row, col, val, _ = b.download()
        S = sp.bsr_matrix((val.reshape(-1, 3, 3), col, row), shape=(3 * N, 3 * N)).tocsr()
        x = s.solution().numpy()
        bb = s.rhs().numpy()
        assert np.linalg.norm(bb - S @ x) / np.linalg.norm(bb) <= 1.01e-8
        direct = spl.spsolve(S.tocsc(), bb)
        assert np.abs(x - direct).max() <= 1e-6 * np.abs(direct).max()


@pytest.mark.parametrize("fmt", [-1, 16])
def test_block_fp32_difference_form(afem_gpu, fmt):
    """The FP32 copy of a 3x3-block matrix is in difference form (diagonal
    block = block row sum, other blocks times x_c - x_a): a rigid translation
    meets the block row sums (zero up to FP64 rounding without Dirichlet
    rows) to FP64 accuracy, where a plain FP32 product leaves about 1e-8 of
    the diagonal blocks' size; a random vector matches the FP64 product to
    FP32 accuracy in SELL-3 and BSR-16 order."""
    A = afem_gpu
    print(f"[case] elasticity fp32 difference form spmv_lanes={fmt}", flush=True)
    X, C = box(3, 6, 0.15)
    N = X.shape[0]
    lam, mu = O.lame(21e5, 0.28)
    m = A.Mesh(3, X, C)
    b = A.BSRFormat(m, 3, use_csr=False)
    b.compute_sparsity()
    b.assemble_elasticity(lam, mu)
    s = A.DoFLinearSystem(3 * N)
    s.set_bsr(b)
    s.solve(raise_on_failure=False, method=A.SOLVER_PCG_MIXED, max_iter=1, spmv_lanes=fmt)
    assert s.spmv_format() == ({-1: "sell", 16: "csr"}[fmt], {-1: 64, 16: 16}[fmt])
    row, col, val, _ = b.download()
    S = sp.bsr_matrix((val.reshape(-1, 3, 3), col, row), shape=(3 * N, 3 * N)).tocsr()
    dy32 = A.DeviceArray(3 * N, np.float32)
    t = np.tile(np.array([1000.0, -500.0, 250.0]), N)
    s.spmv_fp32(A.DeviceArray.from_numpy(t.astype(np.float32)), dy32)
    assert np.abs(dy32.numpy() - S @ t).max() <= 1e-12 * np.abs(val).max() * 1000.0
    x = np.random.default_rng(5).normal(size=3 * N)
    s.spmv_fp32(A.DeviceArray.from_numpy(x.astype(np.float32)), dy32)
    assert np.abs(dy32.numpy() - S @ x).max() <= 2e-5 * np.abs(val).max() * np.abs(x).max()


# ---- the reference's golden solutions, end to end on the GPU ---------------------------------

def by_uid(d, x, ncomp=1):
    mp = dict(zip(d["node_uid"].tolist(), x.reshape(-1, ncomp)))
    return np.array([mp[u] for u in d["ref_uid"].tolist()])


@pytest.mark.parametrize("case", ["circle_2D", "sphere_3D"])
@pytest.mark.parametrize("method", ["fp64", "mixed"])
def test_golden_poisson_gpu(afem_gpu, case, method):
    """2D_bsr / 3D_bsr of modules/poisson: f = 5.5, u = 0.5 on 'horizontal', eps 1e-4.
    Dirichlet by row elimination (the penalty A, b are checked against the
    oracle in test_dirichlet_transformations_match_oracle)."""
    A = afem_gpu
    print(f"[case] golden {case} {method}", flush=True)
    d, X, C, npc, dim = load_case(case)
    m, b, s, _ = build_poisson(A, dim, X, C, 5.5)
    s.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, d["group_horizontal"], 1, 0, 0.5, 0.0)
    s.apply_matrix_transformation()
    s.apply_rhs_transformation()
    code = A.SOLVER_PCG_FP64 if method == "fp64" else A.SOLVER_PCG_MIXED
    r = s.solve(method=code, rtol=1e-12, max_iter=10000)
    assert r.converged == 1
    x = s.solution().numpy()
    assert O.count_differences(d["ref_value"][:, 0], by_uid(d, x)[:, 0], 1e-4) == 0


@pytest.mark.parametrize("case", ["circle_2D", "sphere_3D"])
@pytest.mark.parametrize("norm", [1, 2])
def test_golden_poisson_penalty_gpu(afem_gpu, case, norm):
    """The reference's default Dirichlet method (Penalty, 1e30) on the golden
    cases.  With x0 = 0 the penalty rows dominate both ||b|| and r.z, so the
    2-norm and natural tests stop early; the preconditioned norm
    ||M^-1 r|| / ||M^-1 b|| (PETSc's CG default) weighs a penalty row like
    any other.  The natural test must therefore fail the golden check and
    the preconditioned one pass it."""
    A = afem_gpu
    print(f"[case] golden penalty {case} norm={norm}", flush=True)
    d, X, C, npc, dim = load_case(case)
    m, b, s, _ = build_poisson(A, dim, X, C, 5.5)
    s.apply_dirichlet(A.DIRICHLET_PENALTY, d["group_horizontal"], 1, 0, 0.5, 1e30)
    s.apply_matrix_transformation()
    s.apply_rhs_transformation()
    r = s.solve(method=A.SOLVER_PCG_FP64, rtol=1e-12, max_iter=10000, norm=norm)
    assert r.converged == 1
    nerr = O.count_differences(d["ref_value"][:, 0], by_uid(d, s.solution().numpy())[:, 0], 1e-4)
    if norm == 2:
        assert nerr == 0 and r.iterations > 5 and r.relres <= 1e-12
    else:
        assert nerr > 0


def test_preconditioned_norm_rejected_for_mixed(afem_gpu):
    A = afem_gpu
    d, X, C, npc, dim = load_case("circle_2D")
    m, b, s, _ = build_poisson(A, dim, X, C, 5.5)
    with pytest.raises(A.AfemError):
        s.solve(method=A.SOLVER_PCG_MIXED, norm=2)


@pytest.mark.parametrize("per_block", [True, False])
def test_golden_elasticity_bar_gpu(afem_gpu, per_block):
    """3D_Dirichlet_bodyforce_bsr of modules/elasticity: eps 1e-3, min 1e-10."""
    A = afem_gpu
    print(f"[case] golden bar per_block={per_block}", flush=True)
    d, X, C, _, _ = load_case("bar_3D")
    N = X.shape[0]
    lam, mu = O.lame(21e5, 0.28)
    m = A.Mesh(3, X, C)
    b = A.BSRFormat(m, 3, use_csr=not per_block)
    b.compute_sparsity()
    b.assemble_elasticity(lam, mu)
    s = A.DoFLinearSystem(3 * N)
    s.set_bsr(b)
    rhs = s.rhs()
    rhs.zero()
    A.check(A.lib().afem_rhs_body_force(m.h, A._ptr(np.array([-1.0, 0.0, 0.0])), rhs.c))
    for c in range(3):
        s.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, d["group_surfaceleft"], 3, c, 0.0, 0.0)
    s.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, d["group_surfaceright"], 3, 1, 1.0, 0.0)
    s.apply_matrix_transformation()
    s.apply_rhs_transformation()
    # entries ~1e6 against ||b|| ~ 1: the FP64 true residual floors near 3e-10
    r = s.solve(method=A.SOLVER_PCG_FP64, rtol=1e-9, max_iter=20000)
    assert r.converged == 1
    got = by_uid(d, s.solution().numpy(), 3)
    assert O.count_differences(d["ref_value"], got, 1e-3, 1e-10) == 0


# ---- error paths ------------------------------------------------------------------------------

def test_error_paths(afem_gpu):
    A = afem_gpu
    print("[case] error paths", flush=True)
    X, C = box(2, 3, 0.1)
    m = A.Mesh(2, X, C)
    b = A.BSRFormat(m, 1)
    with pytest.raises(A.AfemError) as e:
        b.assemble_laplace()  # before computeSparsity
    assert e.value.status == A.ERR_STATE
    s = A.DoFLinearSystem(X.shape[0])
    with pytest.raises(A.AfemError) as e:
        s.solve()
    assert e.value.status == A.ERR_STATE
    with pytest.raises(A.AfemError) as e:
        s.apply_dirichlet(7, [0])
    assert e.value.status == A.ERR_ARGUMENT
    with pytest.raises(A.AfemError) as e:
        s.apply_dirichlet(A.DIRICHLET_PENALTY, [X.shape[0]])
    assert e.value.status == A.ERR_ARGUMENT
    bad = C.copy()
    bad[0, 0] = X.shape[0] + 5
    with pytest.raises(A.AfemError) as e:
        A.Mesh(2, X, bad)
    assert e.value.status == A.ERR_MESH
    with pytest.raises(A.AfemError) as e:
        A.BSRFormat(m, 4)
    assert e.value.status == A.ERR_NOT_IMPLEMENTED
    b.compute_sparsity()
    s.set_bsr(b)
    b.assemble_laplace()
    s.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, [0], 1, 0, 0.0, 0.0)
    s.apply_matrix_transformation()
    s.rhs().upload(np.ones(X.shape[0]))
    s.apply_rhs_transformation()
    r = s.solve(raise_on_failure=False, max_iter=2, rtol=1e-14)
    assert r.converged == 0 and r.iterations <= 2


def _boundary_faces(C, dim):
    from collections import Counter
    import itertools
    cnt = Counter()
    for c in C:
        for f in itertools.combinations(sorted(c[:dim + 1]), dim):
            cnt[f] += 1
    return np.array([f for f, k in cnt.items() if k == 1], np.int32)


@pytest.mark.parametrize("dim,n", [(2, 9), (3, 5)])
@pytest.mark.parametrize("value", ["scalar", "vector"])
def test_neumann_matches_oracle(afem_gpu, dim, n, value):
    """afem_rhs_neumann against the oracle on the whole boundary and on one
    side, added on top of a source term."""
    A = afem_gpu
    print(f"[case] neumann dim={dim} {value}", flush=True)
    X, C = box(dim, n, 0.2)
    N = X.shape[0]
    m = A.Mesh(dim, X, C)
    F = _boundary_faces(C, dim)
    g = 1.75 if value == "scalar" else [0.35, -1.65, 3.75][:dim]
    for faces in (F, F[np.all(X[F][:, :, 0] == 0.0, axis=1)]):
        rhs = A.DeviceArray(N, np.float64)
        A.check(A.lib().afem_rhs_constant_source(m.h, 5.5, None, 1, 0, rhs.c))
        A.rhs_neumann(m, faces, g, rhs)
        ref = O.rhs_constant_source(dim, X, C, dim + 1, 5.5)
        ref = O.rhs_neumann(dim, X, C, faces, g, rhs=ref)
        assert np.abs(rhs.numpy() - ref).max() <= 1e-13 * np.abs(ref).max()
    before = rhs.numpy().copy()
    with pytest.raises(A.AfemError):
        A.rhs_neumann(m, np.array([[0, N - 1] + ([1] if dim == 3 else [])], np.int32), g, rhs)
    assert np.array_equal(rhs.numpy(), before)


@pytest.mark.parametrize("dim,n", [(2, 9), (3, 5)])
def test_heat_assembly_matches_oracle(afem_gpu, dim, n):
    """lambda K + M/dt with a per-cell lambda, every assembly variant, then the
    convection matrix on part of the boundary and the mass-weighted old field
    RHS: values and RHS within 1e-12 of the row's / vector's largest entry."""
    A = afem_gpu
    print(f"[case] heat dim={dim}", flush=True)
    X, C = box(dim, n, 0.2)
    N, dt = X.shape[0], 0.4
    m = A.Mesh(dim, X, C)
    b = A.BSRFormat(m, 1, use_csr=True)
    b.compute_sparsity()
    orow, ocol = O.pattern(N, C, dim + 1)
    lam = 1.0 + np.random.default_rng(3).random(C.shape[0])
    dlam = A.DeviceArray.from_numpy(lam)
    oval = O.assemble(dim, X, C, dim + 1, orow, ocol, element=O.HEAT, cell_lambda=lam, lame_lambda=dt)
    for v in range(-1, N_LAPLACE_VARIANTS):
        b.set_assembly_variant(v)
        b.assemble_heat(dlam, dt)
        _, _, val, _ = b.download()
        assert row_scaled_error(orow, oval, val) <= 1e-12, v
    F = _boundary_faces(C, dim)
    side = F[np.all(X[F][:, :, 0] == 0.0, axis=1)]
    b.add_convection(side, 1.9)
    O.convection_lhs(dim, X, side, 1.9, orow, ocol, oval)
    _, _, val, _ = b.download()
    assert row_scaled_error(orow, oval, val) <= 1e-12
    with pytest.raises(A.AfemError):
        b.add_convection(np.array([[0, N - 1] + ([1] if dim == 3 else [])], np.int32), 1.0)
    assert np.array_equal(b.download()[2], val)
    u = 30.0 + np.sin(X[:, 0] * 3.0) * X[:, 1]
    rhs = A.DeviceArray(N, np.float64)
    rhs.zero()  # the Neumann and mass-field terms add into it
    A.rhs_neumann(m, side, 1.9 * 20.0, rhs)
    A.rhs_mass_field(m, A.DeviceArray.from_numpy(u), 1.0 / dt, rhs)
    ref = O.rhs_neumann(dim, X, C, side, 1.9 * 20.0)
    ref = O.rhs_mass_field(X, C, dim + 1, u, 1.0 / dt, rhs=ref)
    assert np.abs(rhs.numpy() - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("use_csr", [True, False])
def test_elasticity_2d_matches_oracle(afem_gpu, use_csr):
    """2D elasticity (tri3, 2x2 blocks) against the oracle in both value orders
    (1e-12 of the row's largest entry); a linear system refuses the per-block
    2x2 order and takes the per-row one."""
    A = afem_gpu
    print(f"[case] elasticity 2D csr={use_csr}", flush=True)
    X, C = box(2, 12, 0.2)
    N = X.shape[0]
    m = A.Mesh(2, X, C)
    b = A.BSRFormat(m, 2, use_csr=use_csr)
    b.compute_sparsity()
    lam, mu = O.lame(21e5, 0.28)
    b.assemble_elasticity(lam, mu)
    row, col, val, _ = b.download()
    orow, ocol = O.pattern(N, C, 3)
    assert np.array_equal(row, orow) and np.array_equal(col, ocol)
    oval = O.assemble(2, X, C, 3, orow, ocol, element=O.ELASTICITY, nb_dof=2, order_per_block=not use_csr,
                      lame_lambda=lam, lame_mu=mu)
    if use_csr:
        crow, _ = O.bsr_to_csr(N, 2, orow, ocol)
        assert row_scaled_error(crow, oval, val) <= ROW_TOL
    else:
        assert row_scaled_error(orow, oval, val, nb=2) <= ROW_TOL
    s = A.DoFLinearSystem(2 * N)
    if use_csr:
        s.set_bsr(b)
    else:
        with pytest.raises(A.AfemError):
            s.set_bsr(b)
