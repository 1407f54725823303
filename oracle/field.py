"""CPU oracle of the log-normal coefficient generator (numpy).

TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/ may import this module,
as the checker of arcanefem_amd/csrc/field.hip.  The generator has no
counterpart in the reference (the Fourier module reads one conductivity per
cell group, modules/fourier/FemModule.cc:175-190); it samples a Gaussian
random field with squared-exponential covariance sigma^2 exp(-|r|^2/(2 l^2))
by K random Fourier features (Rahimi & Recht 2007):
  g(x) = sigma sqrt(2/K) sum_k cos(w_k . x + phi_k),  w_k ~ N(0, I/l^2),  phi_k ~ U[0, 2 pi)
with w and phi drawn from a counter-based splitmix64 stream and Box-Muller.
"""

import numpy as np

M64 = (1 << 64) - 1


def _mix64(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _unit_open(seed, key):
    r = _mix64((_mix64(seed ^ 0x5DEECE66D) + key) & M64)
    return ((r >> 11) + 1) * (1.0 / 9007199254740992.0)


def rff_parameters(dim, nfeatures, corr_length, seed):
    omega = np.zeros((nfeatures, 3))
    phase = np.zeros(nfeatures)
    for k in range(nfeatures):
        for d in range(3):
            u1 = _unit_open(seed, k * 8 + 2 * d)
            u2 = _unit_open(seed, k * 8 + 2 * d + 1)
            gauss = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
            omega[k, d] = gauss / corr_length if d < dim else 0.0
        phase[k] = 2.0 * np.pi * _unit_open(seed, k * 8 + 6)
    return omega, phase


def centroids(coords, cells, npc):
    return coords[cells[:, :npc]].mean(axis=1)


def gaussian_field(points, sigma, omega, phase):
    """g at the given points (FP64)."""
    k = omega.shape[0]
    return sigma * np.sqrt(2.0 / k) * np.cos(points @ omega.T + phase[None, :]).sum(axis=1)


def lognormal_field(coords, cells, npc, sigma, corr_length, nfeatures, seed):
    dim = 3 if npc == 4 else 2
    om, ph = rff_parameters(dim, nfeatures, corr_length, seed)
    return np.exp(gaussian_field(centroids(coords, cells, npc), sigma, om, ph))
