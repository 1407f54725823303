"""CPU oracle of the learned sparse approximate inverse (numpy).

TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/ may import this module,
as the checker of the GPU kernels in arcanefem_amd/csrc/spai.hip.  The
learned preconditioner has no counterpart in the reference (its
preconditioners are Hypre's / PETSc's, femutils/HypreDoFLinearSystem.cc:596-769),
so this restates the model that the library defines:

  S = D^-1/2 A D^-1/2 with Dirichlet rows/columns replaced by the identity
  x_i = 8 row features of S (padded to 16)
  h0 = relu(x W0 + b0);  h_l = h_{l-1} + relu([h_{l-1}, |S|_off h_{l-1}] W_l + b_l), l = 1..3
  u = h3 Wu;  Mh_ii = softplus(h3_i . wd + bd) (1 on Dirichlet rows);  Mh_ij = S_ij (c + u_i . u_j)
  L = 1/(8N) sum_p ||z_p - Mh S z_p||^2 over 8 Rademacher probes.

Two arithmetic modes: exact=True computes everything in FP64 (pinned by
finite differences and closed forms in tests/test_spai_oracle.py);
exact=False rounds to BF16 exactly where the GPU kernels store BF16 (weights
used by the dense layers, x, h, a, u, pre-activations, and the gradient
operands of the dense-layer backward), to check the GPU within BF16 error.
"""

import numpy as np

F, H, L, P = 16, 32, 3, 8
USED = 8

oW0 = 0
oB0 = oW0 + F * H
LAYER = 2 * H * H + H
oLayer = oB0 + H
oWu = oLayer + L * LAYER
oWd = oWu + H * H
oBd = oWd + H
oC = oBd + 1
N_PARAMS = oC + 1


def layer_w(l):
    return oLayer + (l - 1) * LAYER


def layer_b(l):
    return layer_w(l) + 2 * H * H


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    out = u.astype(np.uint32).view(np.float32).copy()
    out[nan] = np.nan
    return out


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W0 = p[oW0:oW0 + F * H].reshape(F, H)
    b0 = p[oB0:oB0 + H]
    Ws = [p[layer_w(l):layer_w(l) + 2 * H * H].reshape(2 * H, H) for l in range(1, L + 1)]
    bs = [p[layer_b(l):layer_b(l) + H] for l in range(1, L + 1)]
    Wu = p[oWu:oWu + H * H].reshape(H, H)
    wd = p[oWd:oWd + H]
    return W0, b0, Ws, bs, Wu, wd, p[oBd], p[oC]


def probe_signs(seed, n):
    """Rademacher probes (n, 8): bit p of a splitmix64-style hash of (seed, row)."""
    with np.errstate(over="ignore"):
        j = np.arange(n, dtype=np.uint64)
        z = np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + j + np.uint64(0x632BE59BD9B4E019)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    bits = (z & np.uint64(0xFF)).astype(np.int64)
    return np.where((bits[:, None] >> np.arange(P)[None, :]) & 1, 1.0, -1.0)


class Graph:
    """Scaled matrix S and features from a CSR matrix and its Dirichlet flags."""

    def __init__(self, row, col, val, elim=None, forced=None, exact=True):
        n = len(row) - 1
        self.n, self.row, self.col = n, np.asarray(row), np.asarray(col)
        self.rid = np.repeat(np.arange(n), np.diff(self.row))
        val = np.asarray(val, dtype=np.float64)
        elim = np.zeros(n, np.uint8) if elim is None else np.asarray(elim)
        forced = np.zeros(n, np.uint8) if forced is None else np.asarray(forced)
        diag = np.zeros(n)
        dmask = self.col == self.rid
        diag[self.rid[dmask]] = val[dmask]
        self.ident = (elim != 0) | (forced != 0) | ~(diag > 0)
        s = np.where(self.ident, 1.0, 1.0 / np.sqrt(np.where(diag > 0, diag, 1.0)))
        if not exact:
            s = s.astype(np.float32).astype(np.float64)
        self.s = s
        i, j = self.rid, self.col
        shat = np.where(self.ident[i] | self.ident[j], 0.0, val * s[i] * s[j])
        shat[dmask] = np.where(self.ident[i[dmask]], 1.0, shat[dmask])
        if not exact:
            shat = shat.astype(np.float32).astype(np.float64)
        self.shat = shat
        self.off = (~dmask) & (shat != 0.0)
        self.dmask = dmask
        # features
        a = np.where(self.off, shat, 0.0)
        m = np.abs(a)
        f = np.zeros((n, F))
        f[:, 0] = 1.0
        f[:, 1] = np.bincount(i, a, n)
        f[:, 2] = np.bincount(i, m, n)
        mx = np.zeros(n)
        np.maximum.at(mx, i, m)
        f[:, 3] = mx
        f[:, 4] = np.bincount(i, a * a, n)
        cnt = np.bincount(i, self.off.astype(float), n)
        f[:, 5] = cnt / 16.0
        f[:, 6] = self.ident.astype(float)
        mn = np.full(n, np.inf)
        np.minimum.at(mn, i[self.off], m[self.off])
        f[:, 7] = np.where(cnt > 0, mn, 0.0)
        self.x = f if exact else bf16(f).astype(np.float64)
        # |S| off the diagonal as a dense-free operator
        self.absoff = np.where(self.off, np.abs(shat), 0.0)

    def agg(self, h):
        """sum_{j != i} |S_ij| h_j."""
        out = np.zeros_like(h)
        np.add.at(out, self.rid, self.absoff[:, None] * h[self.col])
        return out

    def matvec(self, vals, v):
        """sum_j vals_ij v_j (vals in CSR order)."""
        out = np.zeros((self.n,) + v.shape[1:])
        np.add.at(out, self.rid, vals.reshape((-1,) + (1,) * (v.ndim - 1)) * v[self.col])
        return out


def _softplus(x):
    return np.where(x > 20, x, np.log1p(np.exp(np.minimum(x, 20))))


def _sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def forward(params, g, exact=True):
    """Activations dict: h (list), a (list), pre (list), u, pre_d, mhat, m."""
    W0, b0, Ws, bs, Wu, wd, bd, c = unpack(params)
    r = (lambda t: t) if exact else (lambda t: bf16(t).astype(np.float64))
    W0r, Wsr, Wur = r(W0), [r(W) for W in Ws], r(Wu)
    act = {"h": [], "a": [None], "pre": []}
    pre = g.x @ W0r + b0
    act["pre"].append(r(pre))
    act["h"].append(r(np.maximum(pre, 0.0)))
    for l in range(1, L + 1):
        hp = act["h"][-1]
        a = r(g.agg(hp))
        pre = np.concatenate([hp, a], axis=1) @ Wsr[l - 1] + bs[l - 1]
        act["a"].append(a)
        act["pre"].append(r(pre))
        act["h"].append(r(np.maximum(pre, 0.0) + hp))
    h3 = act["h"][-1]
    u = r(h3 @ Wur)
    pre_d = h3 @ wd + bd
    i, j = g.rid, g.col
    mh = np.where(g.off, g.shat * (c + np.sum(u[i] * u[j], axis=1)), 0.0)
    mh[g.dmask] = np.where(g.ident[i[g.dmask]], 1.0, _softplus(pre_d[i[g.dmask]]))
    act.update(u=u, pre_d=pre_d, mhat=mh, m=g.s[i] * mh * g.s[j], weights=(W0r, Wsr, Wur))
    return act


def loss(params, g, seed, exact=True):
    act = forward(params, g, exact)
    z = probe_signs(seed, g.n)
    y = g.matvec(g.shat, z)
    res = z - g.matvec(act["mhat"], y)
    return float(np.sum(res * res) / (P * g.n)), act, y, res


def loss_grad(params, g, seed, exact=True):
    """Loss and its gradient by the chain rule the GPU kernels implement."""
    val, act, y, res = loss(params, g, seed, exact)
    W0, b0, Ws, bs, Wu, wd, bd, c = unpack(params)
    W0r, Wsr, Wur = act["weights"]
    r = (lambda t: t) if exact else (lambda t: bf16(t).astype(np.float64))
    n = g.n
    coef = -2.0 / (P * n)
    grad = np.zeros(N_PARAMS)
    i, j = g.rid, g.col
    # G_ij = dL/dMh_ij
    gij = coef * np.sum(res[i] * y[j], axis=1)
    gji = coef * np.sum(res[j] * y[i], axis=1)
    h3, u = act["h"][-1], act["u"]
    # diagonal head
    dsel = g.dmask & ~g.ident[i]
    rows_d = i[dsel]
    dpre_d = np.zeros(n)
    dpre_d[rows_d] = gij[dsel] * _sigmoid(act["pre_d"][rows_d])
    dh = dpre_d[:, None] * wd[None, :]
    grad[oWd:oWd + H] = dpre_d @ h3
    grad[oBd] = dpre_d.sum()
    # off-diagonal head
    osel = g.off & ~g.ident[i]
    grad[oC] = np.sum(gij[osel] * g.shat[osel])
    w = np.where(osel, (gij + gji) * g.shat, 0.0)
    du = np.zeros((n, H))
    np.add.at(du, i, w[:, None] * u[j])
    # u = h3 Wu
    dur = r(du)
    grad[oWu:oWu + H * H] = (h3.T @ dur).ravel()
    dh = dh + dur @ Wur.T
    for l in range(L, 0, -1):
        hp, a, pre = act["h"][l - 1], act["a"][l], act["pre"][l]
        dpre = np.where(pre > 0, dh, 0.0)
        dpr = r(dpre)
        X = np.concatenate([hp, a], axis=1)
        grad[layer_w(l):layer_w(l) + 2 * H * H] = (X.T @ dpr).ravel()
        grad[layer_b(l):layer_b(l) + H] = dpre.sum(axis=0)
        dX = dpr @ Wsr[l - 1].T
        dh = dh + dX[:, :H] + g.agg(dX[:, H:])
    dpre0 = np.where(act["pre"][0] > 0, dh, 0.0)
    grad[oW0:oW0 + F * H] = (g.x.T @ r(dpre0)).ravel()
    grad[oB0:oB0 + H] = dpre0.sum(axis=0)
    return val, grad


def init_params_jacobi(n_params=N_PARAMS, seed=0, scale=0.3):
    """Random parameters with a non-trivial head (for gradient checks)."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(-scale, scale, n_params)
    p[oW0 + USED * H:oW0 + F * H] = 0.0  # padded input features
    p[oBd] = np.log(np.e - 1.0)
    return p
