"""GPU: the setCSRValues hand-off with the reference's CSR view layout.

The reference's CsrFormatMatrixView (femutils/CsrFormatMatrixView.h:145-187)
carries nbRow() row starts and ends its last row at nbColumn(); BSRMatrix::toCsr
(femutils/BSRFormat.cc:110-172) keeps a closing entry only for blocks > 1.
afem_system_set_csr_values takes that layout.  Bars:
  * values after the Dirichlet transformations, the transformed RHS, SpMV
    results and solutions through a CSR view are bit-identical to the same
    system handed over with afem_system_set_bsr;
  * matrixAddValue / matrixSetValue (femutils/internal/CsrDoFLinearSystemImpl.h:
    56-63) change exactly the value the oracle's findValueIndex names,
    bit-exact, and an entry outside the pattern is AFEM_ERR_NOT_FOUND;
  * clearValues (femutils/DoFLinearSystemImplBase.cc:42-48) resets the forced
    and elimination arrays and drops the view;
  * malformed views are rejected with AFEM_ERR_ARGUMENT and leave the current
    view in place.
"""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def poisson(A, n=6, alpha=0.15, seed=11):
    X, C = A.box_mesh(3, n, alpha, seed)
    m = A.Mesh(3, X, C)
    b = A.BSRFormat(m, 1, use_csr=True)
    b.compute_sparsity()
    return X, C, m, b


def reference_view(A, row, col, val, closing=False, counts=False, poison_tail=None):
    """Device copies of a CSR matrix in the reference's layout: nb_row starts
    (plus a closing entry when asked), optionally the per-row counts.  With
    poison_tail the rows buffer holds one more entry (an in-bounds but wrong
    end for the last row) that a correct reader never looks at."""
    starts = row[:-1] if not closing else row
    if poison_tail is not None:
        buf = A.DeviceArray.from_numpy(np.append(row[:-1], np.int32(poison_tail)).astype(np.int32))
        rows = A.DeviceArray.view(buf.ptr, len(row) - 1, np.int32)
        rows._keep = buf
    else:
        rows = A.DeviceArray.from_numpy(np.ascontiguousarray(starts, np.int32))
    cnt = A.DeviceArray.from_numpy(np.diff(row).astype(np.int32)) if counts else None
    return rows, A.DeviceArray.from_numpy(col.astype(np.int32)), A.DeviceArray.from_numpy(val.astype(np.float64)), cnt


def pin_spmv_format(A, s, lanes=8):
    """Fixes the SpMV format with a one-iteration solve, so that two systems
    run the same kernel (the autotuner picks by timing)."""
    s.solve(raise_on_failure=False, method=A.SOLVER_PCG_FP64, max_iter=1, spmv_lanes=lanes)


def dirichlet(A, s, X, row_column=True):
    bottom = np.nonzero(X[:, 2] == 0.0)[0]  # holds node 0: the column-0 rule of the reference
    s.apply_dirichlet(A.DIRICHLET_ROW_COLUMN_ELIMINATION if row_column else A.DIRICHLET_ROW_ELIMINATION,
                      bottom, 1, 0, 0.25, 0.0)
    top = np.nonzero(X[:, 2] == 1.0)[0]
    if row_column:
        s.apply_dirichlet(A.DIRICHLET_PENALTY, top, 1, 0, 1.0, 1e30)


@pytest.mark.parametrize("layout", ["nb_row", "closing", "counts", "poisoned_tail"])
def test_csr_view_transformations_bit_identical_to_bsr(afem_gpu, layout):
    A = afem_gpu
    print(f"[case] csr view transformations layout={layout}", flush=True)
    X, C, m, b = poisson(A)
    N = X.shape[0]
    s_bsr = A.DoFLinearSystem(N)
    b.assemble_laplace(None, 1.5, s_bsr.rhs())
    row, col, val0, _ = b.download()
    # the same matrix and RHS behind a CSR view in the reference's layout
    rows, dcol, dval, cnt = reference_view(A, row, col, val0, closing=(layout == "closing"),
                                           counts=(layout == "counts"),
                                           poison_tail=(row[-1] - 1 if layout == "poisoned_tail" else None))
    s_csr = A.DoFLinearSystem(N)
    s_csr.rhs().copy_from(s_bsr.rhs())
    s_csr.set_csr_values(rows, dcol, dval, rows_nb_column=cnt)
    s_bsr.set_bsr(b)
    for s in (s_bsr, s_csr):
        pin_spmv_format(A, s)
        dirichlet(A, s, X)
        s.apply_matrix_transformation()
        s.apply_rhs_transformation()
    _, _, val_bsr, _ = b.download()
    assert np.array_equal(dval.numpy(), val_bsr)
    assert np.array_equal(s_csr.rhs().numpy(), s_bsr.rhs().numpy())
    # and both equal the oracle's transformation of the same matrix
    fi, fv = (a.numpy() for a in s_bsr.forced())
    ei, ev = (a.numpy() for a in s_bsr.elimination())
    oval = val0.copy()
    orhs = O.rhs_constant_source(3, X, C, 4, 1.5)
    top = np.nonzero(X[:, 2] == 1.0)[0]
    orhs[top] = 1e30 * 1.0
    O.apply_transformations(row, col, oval, orhs, fi, fv, ei, ev)
    assert np.array_equal(oval, val_bsr)
    assert np.allclose(s_csr.rhs().numpy(), orhs, rtol=1e-12, atol=1e-12 * np.abs(orhs).max())
    x = np.random.default_rng(1).normal(size=N)
    y1, y2 = A.DeviceArray(N, np.float64), A.DeviceArray(N, np.float64)
    s_bsr.spmv(A.DeviceArray.from_numpy(x), y1)
    s_csr.spmv(A.DeviceArray.from_numpy(x), y2)
    assert np.array_equal(y1.numpy(), y2.numpy())


@pytest.mark.parametrize("method", ["fp64", "mixed"])
def test_csr_view_solution_bit_identical_to_bsr(afem_gpu, method):
    A = afem_gpu
    print(f"[case] csr view solve {method}", flush=True)
    X, C, m, b = poisson(A, n=8)
    N = X.shape[0]
    s_bsr = A.DoFLinearSystem(N)
    b.assemble_laplace(None, 1.0, s_bsr.rhs())
    row, col, val0, _ = b.download()
    rows, dcol, dval, _ = reference_view(A, row, col, val0)
    s_csr = A.DoFLinearSystem(N)
    s_csr.rhs().copy_from(s_bsr.rhs())
    s_csr.set_csr_values(rows, dcol, dval)
    s_bsr.set_bsr(b)
    code = A.SOLVER_PCG_FP64 if method == "fp64" else A.SOLVER_PCG_MIXED
    res = []
    for s in (s_bsr, s_csr):
        dirichlet(A, s, X, row_column=False)
        s.apply_matrix_transformation()
        s.apply_rhs_transformation()
        res.append(s.solve(method=code, rtol=1e-10, max_iter=5000, spmv_lanes=-2))
    assert res[0].converged and res[1].converged
    assert res[0].iterations == res[1].iterations
    assert np.array_equal(s_bsr.solution().numpy(), s_csr.solution().numpy())


def test_csr_view_of_expanded_blocks(afem_gpu):
    """An elasticity matrix in per-row order behind its scalar CSR expansion,
    with and without toCsr's closing entry: same SpMV and transformations as
    set_bsr."""
    A = afem_gpu
    print("[case] csr view of 3x3 blocks", flush=True)
    X, C = A.box_mesh(3, 4, 0.15, 11)
    N = X.shape[0]
    lam, mu = O.lame(21e5, 0.28)
    m = A.Mesh(3, X, C)
    b = A.BSRFormat(m, 3, use_csr=True)
    b.compute_sparsity()
    b.assemble_elasticity(lam, mu)
    crow, ccol = b.to_csr()
    _, _, val0, _ = b.download()
    left = np.nonzero(X[:, 0] == 0.0)[0]
    x = np.random.default_rng(2).normal(size=3 * N)
    s_bsr = A.DoFLinearSystem(3 * N)
    s_bsr.set_bsr(b)
    pin_spmv_format(A, s_bsr, 16)
    y_ref = A.DeviceArray(3 * N, np.float64)
    s_bsr.spmv(A.DeviceArray.from_numpy(x), y_ref)
    for closing in (False, True):
        rows, dcol, dval, cnt = reference_view(A, crow, ccol, val0, closing=closing, counts=True)
        s = A.DoFLinearSystem(3 * N)
        s.set_csr_values(rows, dcol, dval, rows_nb_column=cnt)
        pin_spmv_format(A, s, 16)
        y = A.DeviceArray(3 * N, np.float64)
        s.spmv(A.DeviceArray.from_numpy(x), y)
        assert np.array_equal(y.numpy(), y_ref.numpy())
        for c in range(3):
            s.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, left, 3, c, 0.0, 0.0)
        s.apply_matrix_transformation()
        ov = val0.copy()
        ei = np.zeros(3 * N, np.uint8)
        for c in range(3):
            ei[3 * left + c] = O.ELIMINATE_ROW
        O.apply_transformations(crow, ccol, ov, np.zeros(3 * N), elim_info=ei, elim_value=np.zeros(3 * N))
        assert np.array_equal(dval.numpy(), ov)


def entries_to_touch(row, col, N):
    """(row, column) pairs: first and last rows (the last one ends at the value
    count), diagonal and off-diagonal entries, and a few random ones."""
    rng = np.random.default_rng(7)
    out = [(0, 0), (0, int(col[row[1] - 1])), (N - 1, N - 1), (N - 1, int(col[row[N - 1]]))]
    for r in rng.integers(0, N, size=8):
        out.append((int(r), int(col[rng.integers(row[r], row[r + 1])])))
    return out


def test_matrix_add_set_value_against_oracle(afem_gpu):
    A = afem_gpu
    print("[case] matrixAddValue / matrixSetValue scalar", flush=True)
    X, C, m, b = poisson(A)
    N = X.shape[0]
    b.assemble_laplace(None, 0.0, None)
    row, col, val0, _ = b.download()
    rows, dcol, dval, _ = reference_view(A, row, col, val0)
    s = A.DoFLinearSystem(N)
    s.set_csr_values(rows, dcol, dval)
    ref = val0.copy()
    for k, (r, c) in enumerate(entries_to_touch(row, col, N)):
        v = 0.5 + k
        setv = (k % 3 == 2)
        A.check(A.lib().afem_system_matrix_add_value(s.h, r, c, v, 1 if setv else 0))
        assert O.matrix_add_value(row[:-1], col, ref, r, c, v, set_value=setv)
    assert np.array_equal(dval.numpy(), ref)
    # outside the pattern: NOT_FOUND, values untouched
    far = int(np.setdiff1d(np.arange(N), col[row[0]:row[1]])[-1])
    assert O.find_value_index(row[:-1], col, 0, far) == -1
    rc = A.lib().afem_system_matrix_add_value(s.h, 0, far, 1.0, 0)
    assert rc == A.ERR_NOT_FOUND
    assert np.array_equal(dval.numpy(), ref)


@pytest.mark.parametrize("per_block", [True, False])
def test_matrix_add_value_blocks_against_oracle(afem_gpu, per_block):
    """findValueIndex with 3x3 blocks: per-block order through set_bsr, per-row
    order through the expanded CSR view."""
    A = afem_gpu
    print(f"[case] matrixAddValue blocks per_block={per_block}", flush=True)
    X, C = A.box_mesh(3, 3, 0.15, 11)
    N = X.shape[0]
    lam, mu = O.lame(21e5, 0.28)
    m = A.Mesh(3, X, C)
    b = A.BSRFormat(m, 3, use_csr=not per_block)
    b.compute_sparsity()
    b.assemble_elasticity(lam, mu)
    row, col, val0, nzr = b.download()
    s = A.DoFLinearSystem(3 * N)
    if per_block:
        s.set_bsr(b)
        current = lambda: b.download()[2]  # noqa: E731
    else:
        crow, ccol = b.to_csr()
        rows, dcol, dval, _ = reference_view(A, crow, ccol, val0)
        s.set_csr_values(rows, dcol, dval)
        current = dval.numpy
    ref = val0.copy()
    rng = np.random.default_rng(3)
    touched = [(0, 0), (3 * N - 1, 3 * N - 1), (3 * N - 1, 3 * int(col[row[N - 1]]) + 1)]
    for _ in range(10):
        a = int(rng.integers(0, N))
        blk = int(rng.integers(row[a], row[a + 1]))
        touched.append((3 * a + int(rng.integers(0, 3)), 3 * int(col[blk]) + int(rng.integers(0, 3))))
    for k, (r, c) in enumerate(touched):
        setv = (k % 2 == 1)
        A.check(A.lib().afem_system_matrix_add_value(s.h, r, c, 1.25 * k, 1 if setv else 0))
        assert O.matrix_add_value(row[:-1], col, ref, r, c, 1.25 * k, set_value=setv, nb_block=3,
                                  order_per_block=per_block, nnz_per_row=nzr)
    assert np.array_equal(current(), ref)
    assert A.lib().afem_system_matrix_add_value(s.h, 0, 3 * N - 1, 1.0, 0) == A.ERR_NOT_FOUND


def test_clear_values_resets_state(afem_gpu):
    A = afem_gpu
    print("[case] clearValues", flush=True)
    X, C, m, b = poisson(A)
    N = X.shape[0]
    s = A.DoFLinearSystem(N)
    b.assemble_laplace(None, 1.0, s.rhs())
    row, col, val0, _ = b.download()
    rows, dcol, dval, _ = reference_view(A, row, col, val0)
    s.set_csr_values(rows, dcol, dval)
    dirichlet(A, s, X)
    fv_before = s.forced()[1].numpy()
    s.clear_values()
    fi, fv = (a.numpy() for a in s.forced())
    ei, ev = (a.numpy() for a in s.elimination())
    assert not fi.any() and not ei.any() and not ev.any()
    assert np.array_equal(fv, fv_before)  # the reference leaves the forced values as they are
    with pytest.raises(A.AfemError) as e:
        s.solve(rtol=1e-8)
    assert e.value.status == A.ERR_STATE
    assert A.lib().afem_system_matrix_add_value(s.h, 0, 0, 1.0, 0) == A.ERR_STATE
    # a new view after the clear solves
    s.set_csr_values(rows, dcol, dval)
    s.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, np.nonzero(X[:, 2] == 0.0)[0], 1, 0, 0.0, 0.0)
    s.apply_matrix_transformation()
    s.apply_rhs_transformation()
    r = s.solve(method=A.SOLVER_PCG_FP64, rtol=1e-10)
    assert r.converged and r.relres <= 1e-10


def test_malformed_views_rejected(afem_gpu):
    A = afem_gpu
    print("[case] malformed CSR views", flush=True)
    X, C, m, b = poisson(A, n=3)
    N = X.shape[0]
    b.assemble_laplace(None, 1.0, None)
    row, col, val0, _ = b.download()
    nnz = len(col)
    s = A.DoFLinearSystem(N)
    rows, dcol, dval, _ = reference_view(A, row, col, val0)
    s.set_csr_values(rows, dcol, dval)
    pin_spmv_format(A, s)
    good_y = A.DeviceArray(N, np.float64)
    x = A.DeviceArray.from_numpy(np.ones(N))
    s.spmv(x, good_y)

    def rejected(rows_h, cols_h, counts_h=None):
        r = A.DeviceArray.from_numpy(np.asarray(rows_h, np.int32))
        c = A.DeviceArray.from_numpy(np.asarray(cols_h, np.int32))
        k = A.DeviceArray.from_numpy(np.asarray(counts_h, np.int32)) if counts_h is not None else None
        cn = k.c if k is not None else None
        return A.lib().afem_system_set_csr_values(s.h, len(rows_h), nnz, r.c, cn, c.c, dval.c)

    assert rejected(row[:-2], col) == A.ERR_ARGUMENT                       # too few rows
    bad_close = row.copy()
    bad_close[-1] = nnz - 1
    assert rejected(bad_close, col) == A.ERR_ARGUMENT                      # closing entry != nnz
    cnt = np.diff(row).astype(np.int32)
    cnt[-1] += 1
    assert rejected(row[:-1], col, cnt) == A.ERR_ARGUMENT                  # counts disagree (last row)
    desc = row[:-1].copy()
    desc[3], desc[4] = desc[4], desc[3]
    assert rejected(desc, col) == A.ERR_ARGUMENT                           # not non-decreasing
    shifted = row[:-1] + 1
    assert rejected(shifted, col) == A.ERR_ARGUMENT                        # does not start at 0
    badcol = col.copy()
    badcol[-1] = N
    assert rejected(row[:-1], badcol) == A.ERR_ARGUMENT                    # column outside the family
    # the accepted view is still the one in use
    y = A.DeviceArray(N, np.float64)
    s.spmv(x, y)
    assert np.array_equal(y.numpy(), good_y.numpy())


def test_fp32_copy_without_diagonal_entry(afem_gpu):
    """A CSR view with a non-empty row that lacks its diagonal entry cannot
    hold the difference form (the row sum needs the diagonal slot): the
    single-system FP32 product keeps the plain rounded copy and still equals
    the FP64 product to FP32 accuracy, and the batched solver refuses the
    matrix with AFEM_ERR_ARGUMENT."""
    A = afem_gpu
    print("[case] fp32 copy of a view without a diagonal entry", flush=True)
    rng = np.random.default_rng(3)
    N = 40
    dense = np.zeros((N, N))
    for i in range(N):
        for j in (i - 1, i, i + 1):
            if 0 <= j < N and (i < 20) == (j < 20):  # two diagonal blocks, rows 0-19 and 20-39
                dense[i, j] = rng.uniform(1.0, 2.0) * (4.0 if i == j else -1.0)
    dense[7, 7] = 0.0  # row 7 keeps its off-diagonal entries only
    row = np.zeros(N + 1, np.int32)
    cols, vals = [], []
    for i in range(N):
        nz = np.nonzero(dense[i])[0]
        cols += nz.tolist()
        vals += dense[i, nz].tolist()
        row[i + 1] = row[i] + len(nz)
    col, val = np.array(cols, np.int32), np.array(vals)
    rows, dcol, dval, _ = reference_view(A, row, col, val)
    s = A.DoFLinearSystem(N)
    s.set_csr_values(rows, dcol, dval)
    x = rng.normal(size=N)
    y32 = A.DeviceArray(N, np.float32)
    s.spmv_fp32(A.DeviceArray.from_numpy(x.astype(np.float32)), y32)
    ref = dense @ x
    assert np.abs(y32.numpy() - ref).max() <= 2e-6 * np.abs(dense).max() * np.abs(x).max()
    bs = A.BatchSolver(s, np.array([0, 20, N], np.int32))
    with pytest.raises(A.AfemError) as e:
        bs.solve(rtol=1e-8, max_iter=10)
    assert e.value.status == A.ERR_ARGUMENT
