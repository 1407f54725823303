"""GPU tests of the heterogeneous-coefficient workload and the learned
preconditioner, through the C ABI:

* the log-normal field kernel against the CPU oracle;
* the BF16-MFMA network: activations, Mh and the loss gradient of the GPU
  forward/backward kernels against the CPU oracle (oracle/spai.py, BF16
  emulation; the oracle itself is pinned by finite differences in
  test_spai_oracle.py);
* flexible PCG with the learned preconditioner: converges to the FP64 true
  residual, matches a direct solve, equals Jacobi when untrained, falls back
  to Jacobi when the network's M is not positive;
* training on the GPU lowers the probe loss and the iteration count
  (the gain of the network over the method without it).
"""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
from oracle import field as OF
from oracle import spai as OS

pytestmark = pytest.mark.gpu


def build_system(A, dim, n, seed, sigma=1.0, ell=0.1, jitter=0.15, f=1.0):
    """Unit box with a log-normal conductivity, f = 1, u = 0 by row elimination
    on the face x_dim = 0."""
    X, C = A.box_mesh(dim, n, jitter, seed=seed)
    m = A.Mesh(dim, X, C)
    b = A.BSRFormat(m, 1, use_csr=True)
    b.compute_sparsity()
    lam = A.lognormal_field(m, sigma, ell, 512, seed)
    s = A.DoFLinearSystem(X.shape[0])
    b.assemble_laplace(lam, f, s.rhs())
    s.set_bsr(b)
    nodes = np.nonzero(X[:, dim - 1] == 0.0)[0].astype(np.int32)
    s.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, nodes, 1, 0, 0.0, 0.0)
    s.apply_matrix_transformation()
    s.apply_rhs_transformation()
    return dict(X=X, C=C, mesh=m, bsr=b, sys=s, lam=lam, dim=dim)


def reassemble(A, case, seed, sigma=1.0, ell=0.1, f=1.0):
    """New conductivity on the same mesh and pattern (Dirichlet flags kept)."""
    A.lognormal_field(case["mesh"], sigma, ell, 512, seed, out=case["lam"])
    case["bsr"].assemble_laplace(case["lam"], f, case["sys"].rhs())
    case["sys"].set_bsr(case["bsr"])
    case["sys"].apply_matrix_transformation()
    case["sys"].apply_rhs_transformation()


def host_csr(case):
    row, col, val, _ = case["bsr"].download()
    info, _ = case["sys"].elimination()
    return row, col, val, info.numpy()


def test_lognormal_field_matches_oracle(afem_gpu):
    A = afem_gpu
    print("[case] lognormal field 3D n=8", flush=True)
    X, C = A.box_mesh(3, 8, 0.15, seed=2)
    m = A.Mesh(3, X, C)
    lam = A.lognormal_field(m, 1.0, 0.1, 512, 5).numpy()
    ref = OF.lognormal_field(X, C, 4, 1.0, 0.1, 512, 5)
    # FP32 phase and native cosine: |log lam - g| within 1e-3 (sigma = 1)
    assert np.max(np.abs(np.log(lam) - np.log(ref))) < 1e-3
    X2, C2 = A.box_mesh(2, 16, 0.2, seed=1)
    m2 = A.Mesh(2, X2, C2)
    lam2 = A.lognormal_field(m2, 0.5, 0.2, 256, 3).numpy()
    ref2 = OF.lognormal_field(X2, C2, 3, 0.5, 0.2, 256, 3)
    assert np.max(np.abs(np.log(lam2) - np.log(ref2))) < 1e-3


def random_params(seed, scale=0.3):
    p = OS.init_params_jacobi(seed=seed, scale=scale)
    p[OS.oC] = -0.3
    return p.astype(np.float32)


def test_forward_matches_oracle(afem_gpu):
    A = afem_gpu
    print("[case] spai forward 3D n=5", flush=True)
    case = build_system(A, 3, 5, seed=3)
    row, col, val, elim = host_csr(case)
    net = A.LearnedSpai(seed=1)
    p = random_params(2)
    net.set_params(p)
    net.loss(case["sys"], probe_seed=4)
    g = OS.Graph(row, col, val, elim, exact=False)
    act = OS.forward(p.astype(np.float64), g, exact=False)
    n, nnz = g.n, len(col)
    assert np.allclose(net.inspect("scale", n), g.s, rtol=1e-7)
    assert np.allclose(net.inspect("shat", nnz), g.shat, rtol=1e-6, atol=1e-7)
    feats = net.inspect("features", n * 16).reshape(n, 16)
    assert np.mean(feats != g.x) < 0.01 and np.allclose(feats, g.x, rtol=2 ** -7, atol=1e-6)
    # MFMA layer 0: identical BF16 values up to accumulation-order ties
    h0 = net.inspect("h0", n * 32).reshape(n, 32)
    assert np.mean(h0 != act["h"][0]) < 0.02
    assert np.allclose(h0, act["h"][0], rtol=2 ** -7, atol=1e-3)
    h3 = net.inspect("h3", n * 32).reshape(n, 32)
    scale = np.abs(act["h"][-1]).max()
    assert np.max(np.abs(h3 - act["h"][-1])) <= 2e-2 * scale
    mh = net.inspect("mhat", nnz)
    assert np.max(np.abs(mh - act["mhat"])) <= 2e-2 * np.abs(act["mhat"]).max()


def test_gradient_matches_oracle(afem_gpu):
    A = afem_gpu
    print("[case] spai gradient 3D n=5", flush=True)
    case = build_system(A, 3, 5, seed=7)
    row, col, val, elim = host_csr(case)
    net = A.LearnedSpai(seed=1)
    p = random_params(5)
    net.set_params(p)
    loss_gpu, grad_gpu = net.loss_grad(case["sys"], probe_seed=9)
    g = OS.Graph(row, col, val, elim, exact=False)
    loss_ref, grad_ref = OS.loss_grad(p.astype(np.float64), g, 9, exact=False)
    assert loss_gpu == pytest.approx(loss_ref, rel=5e-3)
    groups = {
        "W0": slice(OS.oW0, OS.oB0), "b0": slice(OS.oB0, OS.oLayer),
        "W1": slice(OS.layer_w(1), OS.layer_b(1)), "W3": slice(OS.layer_w(3), OS.layer_b(3)),
        "b3": slice(OS.layer_b(3), OS.layer_b(3) + OS.H), "Wu": slice(OS.oWu, OS.oWd),
        "wd": slice(OS.oWd, OS.oBd), "bd_c": slice(OS.oBd, OS.N_PARAMS),
    }
    for name, sl in groups.items():
        ref = grad_ref[sl]
        err = np.linalg.norm(grad_gpu[sl] - ref) / max(np.linalg.norm(ref), 1e-30)
        assert err < 5e-2, (name, err)
    # the padded input features get no gradient
    assert np.all(grad_gpu[OS.oW0 + OS.USED * OS.H:OS.oB0] == 0.0)


def test_untrained_model_equals_jacobi(afem_gpu):
    A = afem_gpu
    print("[case] fpcg untrained = jacobi 3D n=12", flush=True)
    case = build_system(A, 3, 12, seed=11)
    s = case["sys"]
    net = A.LearnedSpai(seed=0)
    p = net.params()
    p[OS.oWu:OS.oWu + OS.H * OS.H] = 0.0
    net.set_params(p)
    net.apply(s)
    row, col, val, elim = host_csr(case)
    rid = np.repeat(np.arange(len(row) - 1), np.diff(row))
    diag = col == rid
    m = net.inspect("m", len(col))
    free = elim == 0
    assert np.allclose(m[diag][free], 1.0 / val[diag][free], rtol=1e-5)
    assert np.all(m[~diag] == 0.0)
    r_j = s.solve(method=A.SOLVER_PCG_MIXED, rtol=1e-8, max_iter=5000)
    r_l = s.solve(method=A.SOLVER_FPCG_SPAI, rtol=1e-8, max_iter=5000)
    assert r_l.converged and r_l.relres <= 1e-8 and not r_l.fallback
    assert abs(r_l.iterations - r_j.iterations) <= max(3, 0.05 * r_j.iterations)


def test_fpcg_matches_direct_solve(afem_gpu):
    A = afem_gpu
    print("[case] fpcg vs direct 2D n=24", flush=True)
    case = build_system(A, 2, 24, seed=5)
    s = case["sys"]
    net = A.LearnedSpai(seed=3)
    for k in range(20):
        net.train_step(s, probe_seed=k, lr=3e-3)
    net.apply(s)
    r = s.solve(method=A.SOLVER_FPCG_SPAI, rtol=1e-10, max_iter=5000)
    assert r.converged and r.relres <= 1e-10
    row, col, val, _ = case["bsr"].download()
    n = len(row) - 1
    Asp = sp.csr_matrix((val, col, row), shape=(n, n))
    x_ref = spla.spsolve(Asp.tocsc(), s.rhs().numpy())
    x = s.solution().numpy()
    assert np.linalg.norm(x - x_ref) <= 1e-7 * np.linalg.norm(x_ref)


@pytest.mark.parametrize("fmt,bf16", [(-1, 0), (-2, 0), (-2, 1), (-5, 1), (8, 1)])
def test_fpcg_formats_and_bf16(afem_gpu, fmt, bf16):
    """The preconditioner SpMV in every format, with FP32 or BF16 values
    (BF16 is used with the 16-bit-offset SELL format only)."""
    A = afem_gpu
    print(f"[case] fpcg format {fmt} bf16 {bf16}", flush=True)
    case = build_system(A, 3, 12, seed=21)
    s = case["sys"]
    net = A.LearnedSpai(seed=2)
    for k in range(30):
        net.train_step(s, probe_seed=k, lr=3e-3)
    net.apply(s)
    r = s.solve(method=A.SOLVER_FPCG_SPAI, rtol=1e-9, max_iter=5000, spmv_lanes=fmt, precond_bf16=bf16)
    assert r.converged and r.relres <= 1e-9 and not r.fallback
    row, col, val, _ = case["bsr"].download()
    n = len(row) - 1
    x_ref = spla.spsolve(sp.csr_matrix((val, col, row), shape=(n, n)).tocsc(), s.rhs().numpy())
    assert np.linalg.norm(s.solution().numpy() - x_ref) <= 1e-6 * np.linalg.norm(x_ref)


def test_fpcg_formats_at_scale(afem_gpu):
    """Flexible PCG on a grid of 1024 workgroups in SELL-64 and in the 16-bit
    offset layout (256- and 512-thread A products): same convergence, no
    fallback, bit-identical repeats."""
    A = afem_gpu
    print("[case] fpcg formats 3D n=64", flush=True)
    case = build_system(A, 3, 64, seed=23)
    s = case["sys"]
    net = A.LearnedSpai(seed=4)
    for k in range(40):
        net.train_step(s, probe_seed=k, lr=3e-3)
    net.apply(s)
    its = {}
    for fmt in (-1, -2, -5):
        r = s.solve(method=A.SOLVER_FPCG_SPAI, rtol=1e-8, max_iter=5000, spmv_lanes=fmt, precond_bf16=0)
        assert r.converged and r.relres <= 1e-8 and not r.fallback, (fmt, r.iterations, r.fallback)
        its[fmt] = r.iterations
        if fmt != -1:
            # deterministic: a race on the reduction partials would show here
            x1 = s.solution().numpy().copy()
            r2 = s.solve(method=A.SOLVER_FPCG_SPAI, rtol=1e-8, max_iter=5000, spmv_lanes=fmt, precond_bf16=0)
            assert r2.iterations == r.iterations and np.array_equal(s.solution().numpy(), x1), fmt
    # FP32 products in another summation order move the count by a few percent
    for fmt in (-2, -5):
        assert abs(its[fmt] - its[-1]) <= 0.08 * its[-1], its


def test_fallback_to_jacobi(afem_gpu):
    A = afem_gpu
    print("[case] fpcg fallback 2D n=16", flush=True)
    case = build_system(A, 2, 16, seed=2)
    s = case["sys"]
    # an indefinite preconditioner: -D^-1 on every row
    row, col, val, elim = host_csr(case)
    rid = np.repeat(np.arange(len(row) - 1), np.diff(row))
    diag = col == rid
    mv = np.zeros(len(col), np.float32)
    mv[diag] = -1.0 / val[diag]
    s.set_spai(A.DeviceArray.from_numpy(mv))
    r = s.solve(method=A.SOLVER_FPCG_SPAI, rtol=1e-8, max_iter=5000)
    assert r.fallback == 1 and r.converged and r.relres <= 1e-8


def test_fpcg_needs_preconditioner(afem_gpu):
    A = afem_gpu
    case = build_system(A, 2, 8, seed=1)
    with pytest.raises(A.AfemError) as e:
        case["sys"].solve(method=A.SOLVER_FPCG_SPAI)
    assert e.value.status == A.ERR_STATE


def test_training_reduces_loss_and_iterations(afem_gpu):
    """What the trained network gains over the method without it (Jacobi)."""
    A = afem_gpu
    print("[case] spai training 3D n=16", flush=True)
    case = build_system(A, 3, 16, seed=100)
    s = case["sys"]
    net = A.LearnedSpai(seed=0)
    heldout = 12345
    reassemble(A, case, heldout)
    loss0 = net.loss(s, probe_seed=1)
    for step in range(120):
        reassemble(A, case, 1000 + step % 24)
        net.train_step(s, probe_seed=step, lr=3e-3)
    reassemble(A, case, heldout)
    loss1 = net.loss(s, probe_seed=1)
    r_j = s.solve(method=A.SOLVER_PCG_MIXED, rtol=1e-8, max_iter=10000)
    net.apply(s)
    r_l = s.solve(method=A.SOLVER_FPCG_SPAI, rtol=1e-8, max_iter=10000)
    print(f"[spai] loss {loss0:.4f} -> {loss1:.4f}; iterations jacobi {r_j.iterations} learned {r_l.iterations}",
          flush=True)
    assert loss1 < 0.9 * loss0
    assert r_l.converged and r_l.relres <= 1e-8 and not r_l.fallback
    assert r_l.iterations < r_j.iterations
