"""Pins the CPU oracles of the learned preconditioner (oracle/spai.py) and of
the log-normal field generator (oracle/field.py), which have no counterpart
in the reference:

* the oracle's hand-written backward against central finite differences of
  its own FP64 loss (a dropped term or a transposed operand fails it);
* the untrained-model identity (Mh = I, M = D^-1), the symmetry of Mh and the
  loss of Mh = I against dense linear algebra;
* BF16 rounding against known bit patterns;
* the field's random-feature parameters against the library's host code, and
  their statistics against the squared-exponential covariance.
"""

import numpy as np
import pytest

import arcanefem_amd as A
import oracle as O
from oracle import field as OF
from oracle import spai as OS


def small_system(n=4, seed=0, dirichlet=True):
    """Jittered 2D box, per-cell conductivity, row elimination on x = 0."""
    X, C = A.box_mesh(2, n, 0.2, seed=seed)
    npc = 3
    row, col = O.pattern(X.shape[0], C, npc)
    lam = np.exp(np.random.default_rng(seed).normal(size=C.shape[0]))
    val = O.assemble(2, X, C, npc, row, col, cell_lambda=lam)
    elim = np.zeros(X.shape[0], np.uint8)
    if dirichlet:
        elim[X[:, 0] == 0.0] = O.ELIMINATE_ROW
    rhs = np.zeros(X.shape[0])
    O.apply_transformations(row, col, val, rhs, elim_info=elim, elim_value=np.zeros(X.shape[0]))
    return row, col, val, elim


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -7, -2.5, 0.0, 1e-40], np.float32)
    got = OS.bf16(x)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 is a tie -> even (1 + 2^-6)
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7, -2.5, 0.0, 0.0], np.float32)
    assert np.array_equal(got[:6], want[:6])
    assert got[6] == np.float32(np.float32(1e-40).view(np.uint32) & 0xFFFF0000).view(np.float32) or got[6] >= 0


def test_finite_differences_exact_mode():
    row, col, val, elim = small_system(4, seed=1)
    g = OS.Graph(row, col, val, elim, exact=True)
    p = OS.init_params_jacobi(seed=3, scale=0.4)
    p[OS.oC] = -0.3
    loss0, grad = OS.loss_grad(p, g, seed=7, exact=True)
    assert loss0 > 0
    rng = np.random.default_rng(0)
    head = list(range(OS.oWd, OS.N_PARAMS)) + list(range(OS.oWu, OS.oWu + 40))
    body = rng.choice(OS.oWu, 120, replace=False).tolist()
    eps = 1e-6
    bad = []
    for q in head + body:
        if OS.oW0 + OS.USED * OS.H <= q < OS.oW0 + OS.F * OS.H:
            continue  # padded features: zero input
        pp, pm = p.copy(), p.copy()
        pp[q] += eps
        pm[q] -= eps
        fd = (OS.loss(pp, g, 7)[0] - OS.loss(pm, g, 7)[0]) / (2 * eps)
        if abs(fd - grad[q]) > 1e-7 + 1e-4 * abs(grad[q]):
            bad.append((q, fd, grad[q]))
    assert len(bad) <= 2, bad[:5]  # a ReLU kink inside [p - eps, p + eps] can spoil one or two


def test_untrained_model_is_jacobi():
    row, col, val, elim = small_system(5, seed=2)
    g = OS.Graph(row, col, val, elim, exact=True)
    p = OS.init_params_jacobi(seed=1)
    p[OS.oWu:OS.oWu + OS.H * OS.H] = 0.0
    p[OS.oWd:OS.oWd + OS.H] = 0.0
    p[OS.oC] = 0.0
    act = OS.forward(p, g, exact=True)
    rid = np.repeat(np.arange(g.n), np.diff(row))
    diag = col == rid
    assert np.allclose(act["mhat"][diag], 1.0, rtol=1e-14)
    assert np.all(act["mhat"][~diag] == 0.0)
    d = val[diag]
    free = elim == 0
    assert np.allclose(act["m"][diag][free], 1.0 / d[free], rtol=1e-13)


def test_mhat_symmetric_and_loss_dense():
    row, col, val, elim = small_system(4, seed=4)
    g = OS.Graph(row, col, val, elim, exact=True)
    p = OS.init_params_jacobi(seed=5, scale=0.5)
    p[OS.oC] = 0.2
    act = OS.forward(p, g, exact=True)
    n = g.n
    rid = np.repeat(np.arange(n), np.diff(row))
    Md = np.zeros((n, n))
    Md[rid, col] = act["mhat"]
    assert np.allclose(Md, Md.T, rtol=0, atol=1e-12)
    # loss against dense algebra
    S = np.zeros((n, n))
    S[rid, col] = g.shat
    z = OS.probe_signs(11, n)
    res = z - Md @ (S @ z)
    assert OS.loss(p, g, 11)[0] == pytest.approx(np.sum(res ** 2) / (8 * n), rel=1e-12)
    # S is symmetric with unit diagonal on the free rows
    assert np.allclose(S, S.T, atol=1e-14)
    assert np.allclose(np.diag(S), 1.0)


def test_probe_signs_balanced():
    z = OS.probe_signs(3, 20000)
    assert set(np.unique(z)) == {-1.0, 1.0}
    assert abs(z.mean()) < 0.02
    c = np.corrcoef(z.T)
    assert np.max(np.abs(c - np.eye(8))) < 0.05


def test_rff_parameters_match_library():
    om, ph = A.rff_parameters(3, 64, 0.1, 42)
    om2, ph2 = OF.rff_parameters(3, 64, 0.1, 42)
    assert np.allclose(om, om2, rtol=1e-14, atol=0)
    assert np.allclose(ph, ph2, rtol=1e-14, atol=0)
    om, _ = A.rff_parameters(2, 16, 0.2, 1)
    assert np.all(om[:, 2] == 0.0)


def test_rff_statistics():
    ell = 0.1
    om, ph = OF.rff_parameters(3, 4096, ell, 9)
    # mean_k cos(w_k . r) -> exp(-|r|^2 / (2 l^2)), the characteristic function of N(0, I/l^2)
    for r in (0.0, 0.05, 0.1, 0.2):
        v = np.array([r, 0.0, 0.0])
        emp = np.cos(om @ v).mean()
        assert emp == pytest.approx(np.exp(-r * r / (2 * ell * ell)), abs=0.05)
    assert ph.min() >= 0 and ph.max() <= 2 * np.pi and abs(ph.mean() - np.pi) < 0.15
    # marginal of g: zero mean, unit variance.  One realization's spatial mean
    # over the unit cube has a standard deviation of about
    # ((2 pi)^(3/2) l^3)^(1/2) = 0.13, so average over 8 realizations.
    pts = np.random.default_rng(0).uniform(size=(8000, 3))
    means, variances = [], []
    for seed in range(8):
        om_s, ph_s = OF.rff_parameters(3, 1024, ell, 100 + seed)
        g = OF.gaussian_field(pts, 1.0, om_s, ph_s)
        means.append(g.mean())
        variances.append(g.var())
    assert abs(np.mean(means)) < 0.15
    assert 0.75 < np.mean(variances) < 1.25
