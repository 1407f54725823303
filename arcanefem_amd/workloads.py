"""Synthetic inputs of the benchmark configurations (host side).

The same builders serve bench.py and the full-size parity tests, so that a
test checks exactly the systems the benchmark times.  Meshes come from
afem_box_mesh (structured unit square / cube, 2 triangles per quad or 6 Kuhn
tets per hex, interior nodes jittered from a counter-based generator).
"""

import numpy as np


def poisson2d(A, n=256, seed=0, jitter=0.2):
    """2D Poisson P1: unit square, n x n quads x 2 triangles, f = 1, u = 0 on
    the whole boundary.  Returns (X, C, boundary node ids)."""
    X, C = A.box_mesh(2, n, jitter, seed=seed)
    on = np.nonzero((X[:, 0] == 0.0) | (X[:, 0] == 1.0) | (X[:, 1] == 0.0) | (X[:, 1] == 1.0))[0]
    return X, C, on.astype(np.int32)


def poisson3d(A, n=128, seed=0, jitter=0.15):
    """3D Poisson P1: unit cube, n^3 hexes x 6 Kuhn tets, f = 1, u = 0 on
    z = 0.  Returns (X, C, bottom node ids)."""
    X, C = A.box_mesh(3, n, jitter, seed=seed)
    return X, C, np.nonzero(X[:, 2] == 0.0)[0].astype(np.int32)


def elasticity3d(A, n=96, seed=0, jitter=0.15):
    """3D linear elasticity P1: n^3 hexes x 6 tets, E = 21e5, nu = 0.28, body
    force (0, 0, -1), x = 0 clamped.  Returns (X, C, clamped node ids,
    lame_lambda, lame_mu)."""
    X, C = A.box_mesh(3, n, jitter, seed=seed)
    E, nu = 21e5, 0.28
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    return X, C, np.nonzero(X[:, 0] == 0.0)[0].astype(np.int32), lam, mu


def batch2d(A, nsys=64, seed=0):
    """nsys independent unit squares with n x n quads x 2 triangles, n
    log-uniform in [64, 512], interior nodes jittered by 0.2h, conductivity
    piecewise constant on an 8 x 8 checkerboard with values log-uniform in
    [0.1, 10], f = 1, u = 0 on the boundary.  Returns (ns, parts [(X, C)],
    per-cell conductivities, boundary node ids per part)."""
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(np.log(64), np.log(512), nsys))).astype(int)
    parts, lams, bnds = [], [], []
    for k, n in enumerate(ns):
        X, C = A.box_mesh(2, int(n), 0.2, seed=seed + 100 + k)
        vals = np.exp(rng.uniform(np.log(0.1), np.log(10.0), 64))
        cen = X[C[:, :3]].mean(axis=1)
        ix = np.minimum((cen[:, 0] * 8).astype(int), 7)
        iy = np.minimum((cen[:, 1] * 8).astype(int), 7)
        parts.append((X, C))
        lams.append(vals[iy * 8 + ix])
        on = (X[:, 0] == 0.0) | (X[:, 0] == 1.0) | (X[:, 1] == 0.0) | (X[:, 1] == 1.0)
        bnds.append(np.nonzero(on)[0])
    return ns, parts, lams, bnds


# Log-normal conductivity of the heterogeneous configuration: g a Gaussian
# random field with squared-exponential covariance, sigma 1, correlation
# length 0.1, from 512 random Fourier features.
LOGNORMAL = dict(sigma=1.0, corr_length=0.1, nfeatures=512)
