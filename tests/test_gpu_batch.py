"""GPU tests of the batched solver: independent 2D heat/Poisson systems with
8x8 checkerboard conductivities on meshes of different sizes, assembled in
one launch on the disjoint union of the meshes and solved together with
per-system scalars (segmented mixed-precision PCG: hipGraph chunks of
ordinary kernels by default, one persistent launch with persistent=1).

Checked against: a direct solve of each diagonal block (scipy), the
single-system mixed PCG of the same block (iteration counts), the
plain-launch path (bit-identical to the graph replays), and the persistent
and chunked solves against each other.
"""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu


def checkerboard(X, C, values):
    cen = X[C[:, :3]].mean(axis=1)
    ix = np.minimum((cen[:, 0] * 8).astype(int), 7)
    iy = np.minimum((cen[:, 1] * 8).astype(int), 7)
    return values[iy * 8 + ix]


def build_batch(A, ns, seed):
    rng = np.random.default_rng(seed)
    parts, lams, bnd = [], [], []
    for k, n in enumerate(ns):
        X, C = A.box_mesh(2, n, 0.2, seed=seed + k)
        vals = np.exp(rng.uniform(np.log(0.1), np.log(10.0), 64))
        parts.append((X, C))
        lams.append(checkerboard(X, C, vals))
        on = (X[:, 0] == 0.0) | (X[:, 0] == 1.0) | (X[:, 1] == 0.0) | (X[:, 1] == 1.0)
        bnd.append(np.nonzero(on)[0])
    X, C, offs = A.union_mesh(parts)
    nodes = np.concatenate([b + offs[k] for k, b in enumerate(bnd)]).astype(np.int32)
    mesh = A.Mesh(2, X, C)
    bsr = A.BSRFormat(mesh, 1, use_csr=True)
    bsr.compute_sparsity()
    lam = A.DeviceArray.from_numpy(np.concatenate(lams))
    sysm = A.DoFLinearSystem(X.shape[0])
    bsr.assemble_laplace(lam, 1.0, sysm.rhs())
    sysm.set_bsr(bsr)
    sysm.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, nodes, 1, 0, 0.0, 0.0)
    sysm.apply_matrix_transformation()
    sysm.apply_rhs_transformation()
    return dict(parts=parts, lams=lams, bnd=bnd, offs=offs, mesh=mesh, bsr=bsr, lam=lam, sys=sysm)


def block_direct(case, s):
    row, col, val, _ = case["bsr"].download()
    o0, o1 = case["offs"][s], case["offs"][s + 1]
    n = row.size - 1
    M = sp.csr_matrix((val, col, row), shape=(n, n))[o0:o1, o0:o1]
    b = case["sys"].rhs().numpy()[o0:o1]
    return spla.spsolve(M.tocsc(), b)


def test_batch_matches_direct_solves(afem_gpu):
    A = afem_gpu
    print("[case] batch 5 systems vs direct", flush=True)
    case = build_batch(A, [6, 9, 13, 20, 32], seed=1)
    bs = A.BatchSolver(case["sys"], case["offs"])
    res, ms = bs.solve(rtol=1e-10, max_iter=5000)
    x = case["sys"].solution().numpy()
    for s, r in enumerate(res):
        assert r.converged and r.relres <= 1e-10, (s, r.relres)
        ref = block_direct(case, s)
        xs = x[case["offs"][s]:case["offs"][s + 1]]
        assert np.linalg.norm(xs - ref) <= 1e-8 * np.linalg.norm(ref), s
    # bigger systems need more iterations (condition number ~ n^2)
    its = [r.iterations for r in res]
    assert its[-1] > its[0]


def test_batch_graph_equals_plain_launches(afem_gpu):
    A = afem_gpu
    print("[case] batch graph vs plain", flush=True)
    case = build_batch(A, [7, 16, 11], seed=4)
    bs = A.BatchSolver(case["sys"], case["offs"])
    for persistent in (0, 1):
        r1, _ = bs.solve(rtol=1e-8, use_graph=1, chunk=8, persistent=persistent)
        x1 = case["sys"].solution().numpy()
        r2, _ = bs.solve(rtol=1e-8, use_graph=0, chunk=8, persistent=persistent)
        x2 = case["sys"].solution().numpy()
        assert [r.iterations for r in r1] == [r.iterations for r in r2]
        assert np.array_equal(x1, x2)


def test_batch_iterations_match_single_solver(afem_gpu):
    A = afem_gpu
    print("[case] batch vs single-system mixed PCG", flush=True)
    ns = [10, 24]
    case = build_batch(A, ns, seed=9)
    bs = A.BatchSolver(case["sys"], case["offs"])
    bs.set_tiles(8, 1)
    res, _ = bs.solve(rtol=1e-8, persistent=0)  # chunked, CSR with 8 lanes: the single solver's product
    res_p, _ = bs.solve(rtol=1e-8, persistent=1)  # SELL product and per-strip partial sums
    for s, (X, C) in enumerate(case["parts"]):
        m = A.Mesh(2, X, C)
        b = A.BSRFormat(m, 1, use_csr=True)
        b.compute_sparsity()
        single = A.DoFLinearSystem(X.shape[0])
        b.assemble_laplace(A.DeviceArray.from_numpy(case["lams"][s]), 1.0, single.rhs())
        single.set_bsr(b)
        single.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, case["bnd"][s].astype(np.int32), 1, 0, 0.0, 0.0)
        single.apply_matrix_transformation()
        single.apply_rhs_transformation()
        # the same SpMV as the batch kernels (CSR, 8 lanes per row): with FP32
        # products the iteration count moves with the summation order (42-46
        # and 109-118 across the SpMV formats here)
        r = single.solve(method=A.SOLVER_PCG_MIXED, rtol=1e-8, max_iter=5000, spmv_lanes=8)
        assert abs(r.iterations - res[s].iterations) <= max(3, 0.05 * r.iterations), (s, r.iterations,
                                                                                       res[s].iterations)
        assert abs(r.iterations - res_p[s].iterations) <= max(6, 0.12 * r.iterations), (s, r.iterations,
                                                                                         res_p[s].iterations)


def test_batch_autotune_and_tiles(afem_gpu):
    A = afem_gpu
    print("[case] batch autotune", flush=True)
    case = build_batch(A, [12, 40, 25, 8], seed=2)
    bs = A.BatchSolver(case["sys"], case["offs"])
    lanes, tiles, ms = bs.autotune()
    assert lanes in (1, 4, 8, 16) and 1 <= tiles <= 16 and ms > 0
    res, _ = bs.solve(rtol=1e-8)
    assert all(r.converged and r.relres <= 1e-8 for r in res)
    for lanes, tiles in ((4, 1), (16, 2), (1, 1), (1, 4)):
        bs.set_tiles(lanes, tiles)
        res2, _ = bs.solve(rtol=1e-8)
        assert all(r.converged for r in res2)


def test_batch_sell_and_shrinking_grid(afem_gpu):
    """One large system among many small ones: once the small ones converge
    the grid shrinks to the large one's workgroups.  The SELL-64 layout
    (lanes = 1) and CSR (8 lanes) both match the direct solves, and graph
    replays equal plain launches."""
    A = afem_gpu
    print("[case] batch sell + shrinking grid", flush=True)
    case = build_batch(A, [40, 6, 7, 8, 9, 6, 7, 8, 9], seed=5)
    bs = A.BatchSolver(case["sys"], case["offs"])
    sols = {}
    for lanes in (1, 8):
        bs.set_tiles(lanes, 1)
        res, _ = bs.solve(rtol=1e-10, max_iter=5000, persistent=0)
        x = case["sys"].solution().numpy().copy()
        for s_, r in enumerate(res):
            assert r.converged and r.relres <= 1e-10, (lanes, s_, r.relres)
            ref = block_direct(case, s_)
            xs = x[case["offs"][s_]:case["offs"][s_ + 1]]
            assert np.linalg.norm(xs - ref) <= 1e-8 * np.linalg.norm(ref), (lanes, s_)
        assert res[0].iterations > 2 * max(r.iterations for r in res[1:])  # the grid did shrink
        sols[lanes] = (res, x)
    bs.set_tiles(1, 1)
    r2, _ = bs.solve(rtol=1e-10, max_iter=5000, use_graph=0, persistent=0)
    assert [r.iterations for r in r2] == [r.iterations for r in sols[1][0]]
    assert np.array_equal(case["sys"].solution().numpy(), sols[1][1])


def test_batch_argument_errors(afem_gpu):
    A = afem_gpu
    case = build_batch(A, [5, 6], seed=3)
    n = case["offs"][-1]
    for bad in ([0, n - 1], [1, n], [0, 5, 5, n]):
        with pytest.raises(A.AfemError):
            A.BatchSolver(case["sys"], bad)
    bs = A.BatchSolver(case["sys"], case["offs"])
    with pytest.raises(A.AfemError):
        bs.set_tiles(3, 1)
    for bad in (-1, 3):
        with pytest.raises(A.AfemError) as e:
            bs.solve(persistent=bad)
        assert e.value.status == A.ERR_ARGUMENT
    assert bs.persistent_mode() == 0


def test_batch_heat_time_loop_matches_oracle(afem_gpu):
    """The heat module's backward-Euler loop on every system of a batch at
    once (lambda K + M/dt assembled on the union mesh, rhs = f + M u_old/dt,
    u = 0 rows eliminated, warm-started batched solves) against the oracle's
    loop on each mesh alone with direct solves: 1e-8 relative per system."""
    import oracle as O
    A = afem_gpu
    print("[case] batch heat time loop", flush=True)
    ns, dt, steps = [6, 11, 17, 24], 0.02, 6
    rng = np.random.default_rng(8)
    parts, lams, bnd = [], [], []
    for k, n in enumerate(ns):
        X, C = A.box_mesh(2, n, 0.2, seed=40 + k)
        parts.append((X, C))
        lams.append(checkerboard(X, C, np.exp(rng.uniform(np.log(0.1), np.log(10.0), 64))))
        on = (X[:, 0] == 0.0) | (X[:, 0] == 1.0) | (X[:, 1] == 0.0) | (X[:, 1] == 1.0)
        bnd.append(np.nonzero(on)[0])
    X, C, offs = A.union_mesh(parts)
    N = X.shape[0]
    nodes = np.concatenate([b + offs[k] for k, b in enumerate(bnd)]).astype(np.int32)
    mesh = A.Mesh(2, X, C)
    bsr = A.BSRFormat(mesh, 1, use_csr=True)
    bsr.compute_sparsity()
    sysm = A.DoFLinearSystem(N)
    bsr.assemble_heat(A.DeviceArray.from_numpy(np.concatenate(lams)), dt)
    sysm.set_bsr(bsr)
    sysm.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, nodes, 1, 0, 0.0, 0.0)
    sysm.apply_matrix_transformation()
    source = A.DeviceArray(N, np.float64)
    A.check(A.lib().afem_rhs_constant_source(mesh.h, 1.0, None, 1, 0, source.c))
    sysm.solution().zero()  # Tinit = 0
    bs = A.BatchSolver(sysm, offs)
    out = A.heat_batch_steps(mesh, sysm, bs, source, dt, steps, rtol=1e-11, max_iter=4000)
    assert len(out) == steps and all(r.converged for res, _ in out for r in res)
    x = sysm.solution().numpy()
    for s, (Xs, Cs) in enumerate(parts):
        n = Xs.shape[0]
        row, col = O.pattern(n, Cs, 3)
        val = O.assemble(2, Xs, Cs, 3, row, col, element=O.HEAT, cell_lambda=lams[s], lame_lambda=dt)
        ei, ev = np.zeros(n, np.uint8), np.zeros(n)
        ei[bnd[s]] = 1
        O.apply_transformations(row, col, val, np.zeros(n), np.zeros(n, np.uint8), np.zeros(n), ei, ev)
        lu = spla.splu(sp.csr_matrix((val, col, row), shape=(n, n)).tocsc())
        f = O.rhs_constant_source(2, Xs, Cs, 3, 1.0)
        u = np.zeros(n)
        for _ in range(steps):
            b = O.rhs_mass_field(Xs, Cs, 3, u, 1.0 / dt, rhs=f.copy())
            b[bnd[s]] = 0.0
            u = lu.solve(b)
        xs = x[offs[s]:offs[s + 1]]
        assert np.linalg.norm(xs - u) <= 1e-8 * np.linalg.norm(u), s
    # warm starts: the last step needs no more iterations than the first
    its = [max(r.iterations for r in res) for res, _ in out]
    assert its[-1] <= its[0]


def test_persistent_matches_chunked_and_direct(afem_gpu):
    """One persistent launch against the chunked graphs of ordinary kernels on
    the same batch: both meet 1e-10 per system, match the direct solves, and
    need about the same iterations (the partial sums are taken over
    different workgroup strips, so FP32 rounding differs; at 1e-10, below
    FP32's resolution, the reliable updates' timing moves the count by up to
    6 % here, e.g. 136 and 144 for the 25x25 system)."""
    A = afem_gpu
    print("[case] persistent vs chunked", flush=True)
    case = build_batch(A, [40, 6, 7, 8, 9, 6, 7, 8, 9, 25], seed=6)
    bs = A.BatchSolver(case["sys"], case["offs"])
    out = {}
    for persistent in (1, 2, 0):
        res, _ = bs.solve(rtol=1e-10, max_iter=5000, persistent=persistent)
        if persistent:
            mode = bs.persistent_mode()
            print(f"[case] persistent={persistent} exchange mode {mode}", flush=True)
            assert mode == 2 if persistent == 2 else mode in (1, 2)
        x = case["sys"].solution().numpy().copy()
        for s_, r in enumerate(res):
            assert r.converged and r.relres <= 1e-10, (persistent, s_, r.relres)
            ref = block_direct(case, s_)
            xs = x[case["offs"][s_]:case["offs"][s_ + 1]]
            assert np.linalg.norm(xs - ref) <= 1e-8 * np.linalg.norm(ref), (persistent, s_)
        out[persistent] = res
    for k in (1, 2):
        for rp, rc in zip(out[k], out[0]):
            assert abs(rp.iterations - rc.iterations) <= max(3, 0.08 * rc.iterations), (k, rp.iterations, rc.iterations)


@pytest.mark.parametrize("persistent", [1, 2])
@pytest.mark.parametrize("workgroups", [0, 10, 64, 1024])
def test_persistent_team_sizes(afem_gpu, workgroups, persistent):
    """Teams from one workgroup per system (10 systems) to as many as the
    device holds, with XCD-local teams (where the device allows them) and
    with the coherent exchange: every size converges to the same tolerance,
    and a graph replay of the launch equals a plain launch bit for bit."""
    A = afem_gpu
    print(f"[case] persistent={persistent} workgroups={workgroups}", flush=True)
    case = build_batch(A, [30, 6, 7, 8, 9, 6, 7, 8, 9, 64], seed=12)
    bs = A.BatchSolver(case["sys"], case["offs"])
    r1, _ = bs.solve(rtol=1e-9, max_iter=5000, workgroups=workgroups, use_graph=1, persistent=persistent)
    x1 = case["sys"].solution().numpy()
    assert all(r.converged and r.relres <= 1e-9 for r in r1)
    r2, _ = bs.solve(rtol=1e-9, max_iter=5000, workgroups=workgroups, use_graph=0, persistent=persistent)
    x2 = case["sys"].solution().numpy()
    assert [r.iterations for r in r1] == [r.iterations for r in r2]
    assert np.array_equal(x1, x2)


def test_persistent_wait_limit_gives_up_then_recovers(afem_gpu):
    """A spin limit of one poll makes the first barrier wait of a large team
    give up (forced from the host; every workgroup still leaves the kernel
    normally): the solve fails with AFEM_ERR_DEVICE, and the next solve with
    the default limit converges."""
    A = afem_gpu
    print("[case] persistent spin limit", flush=True)
    case = build_batch(A, [200, 150], seed=13)
    bs = A.BatchSolver(case["sys"], case["offs"])
    with pytest.raises(A.AfemError) as e:
        bs.solve(rtol=1e-10, max_iter=5000, spin_limit=1, workgroups=256, use_graph=0, persistent=1)
    assert e.value.status == A.ERR_DEVICE
    res, _ = bs.solve(rtol=1e-10, max_iter=20000, persistent=1)
    assert all(r.converged and r.relres <= 1e-10 for r in res)
