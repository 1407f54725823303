"""GPU: the allocation-free solve (afem_system_solve_noalloc) against
afem_system_solve.

afem_system_solve_noalloc has no counterpart in the reference; it is the FP64
PCG of afem_system_solve with its work vectors, partials and state in a fixed
device array.  With the existing entry point pinned to the same SpMV kernel
(CSR / 3x3 BSR, 16 lanes per row) both run the same kernels in the same
order, so the bar is bit-identical solutions, iteration counts and residuals
-- on two calls in a row, the second on a smaller system (the array keeps the
first call's contents), for scalar and 3x3-block systems and for the three
stopping norms.  Sizes beyond the array are refused before any GPU work.
"""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def poisson_system(A, n, seed, method):
    X, C = A.box_mesh(3, n, 0.15, seed)
    N = X.shape[0]
    m = A.Mesh(3, X, C)
    b = A.BSRFormat(m, 1, use_csr=True)
    b.compute_sparsity()
    s = A.DoFLinearSystem(N)
    b.assemble_laplace(None, 1.0, s.rhs())
    s.set_bsr(b)
    bottom = np.nonzero(X[:, 2] == 0.0)[0]
    s.apply_dirichlet(method, bottom, 1, 0, 0.5, 1e30)
    s.apply_matrix_transformation()
    s.apply_rhs_transformation()
    return s, (m, b)


def elasticity_system(A, n, seed):
    X, C = A.box_mesh(3, n, 0.15, seed)
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
    return s, (m, b)


def compare(A, s, **kw):
    """Runs afem_system_solve (pinned to the 16-lane CSR / BSR SpMV) and the
    allocation-free solve on the same system; returns both solutions."""
    r_ref = s.solve(spmv_lanes=16, **kw)
    x_ref = s.solution().numpy().copy()
    s.solution().zero()
    r_na = s.solve_noalloc(**kw)
    x_na = s.solution().numpy().copy()
    assert r_na.converged == r_ref.converged == 1
    assert r_na.iterations == r_ref.iterations, (r_na.iterations, r_ref.iterations)
    assert r_na.relres == r_ref.relres, (r_na.relres, r_ref.relres)
    assert np.array_equal(x_na, x_ref), np.abs(x_na - x_ref).max()
    return r_na


@pytest.mark.parametrize("norm", [0, 1, 2])
def test_noalloc_poisson_two_calls_bit_identical(afem_gpu, norm):
    A = afem_gpu
    method = A.DIRICHLET_PENALTY if norm == 2 else A.DIRICHLET_ROW_ELIMINATION
    for n, seed in ((12, 3), (7, 4)):  # the second call is smaller
        print(f"[case] noalloc poisson n={n} norm={norm}", flush=True)
        s, keep = poisson_system(A, n, seed, method)
        r = compare(A, s, method=A.SOLVER_PCG_FP64, rtol=1e-10, norm=norm)
        assert r.iterations > 5


def test_noalloc_block3_and_plain_cg(afem_gpu):
    A = afem_gpu
    for n, seed in ((7, 5), (4, 6)):
        print(f"[case] noalloc elasticity n={n}", flush=True)
        s, keep = elasticity_system(A, n, seed)
        compare(A, s, method=A.SOLVER_PCG_FP64, rtol=1e-9)
    print("[case] noalloc plain CG", flush=True)
    s, keep = poisson_system(A, 9, 7, A.DIRICHLET_ROW_ELIMINATION)
    compare(A, s, method=A.SOLVER_CG_FP64, rtol=1e-10)


def test_noalloc_repeated_call_same_result(afem_gpu):
    A = afem_gpu
    print("[case] noalloc repeated call", flush=True)
    s, keep = poisson_system(A, 8, 9, A.DIRICHLET_ROW_ELIMINATION)
    r1 = s.solve_noalloc(method=A.SOLVER_PCG_FP64, rtol=1e-10)
    x1 = s.solution().numpy().copy()
    r2 = s.solve_noalloc(method=A.SOLVER_PCG_FP64, rtol=1e-10)
    x2 = s.solution().numpy().copy()
    assert r1.iterations == r2.iterations and np.array_equal(x1, x2)


def test_noalloc_refuses_what_it_does_not_run(afem_gpu):
    A = afem_gpu
    print("[case] noalloc limits", flush=True)
    big = A.DoFLinearSystem((1 << 20) + 1)
    with pytest.raises(A.AfemError) as e:
        big.solve_noalloc(method=A.SOLVER_PCG_FP64)
    assert e.value.status == A.ERR_ARGUMENT
    s, keep = poisson_system(A, 4, 1, A.DIRICHLET_ROW_ELIMINATION)
    with pytest.raises(A.AfemError) as e:
        s.solve_noalloc(method=A.SOLVER_PCG_MIXED)
    assert e.value.status == A.ERR_NOT_IMPLEMENTED
