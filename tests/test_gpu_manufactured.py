"""Manufactured solutions of the Poisson module
(modules/poisson/ManufacturedSolutions.{h,cc}) on the device, through the
C-ABI: the source RHS (afem_rhs_manufactured) and the L2 error
(afem_manufactured_l2_error) against the oracle on tri3, Quad4 and Hexa8
meshes, relative tolerance 1e-12."""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

NAMES = ["sine", "poly", "quadratic"]


def meshes(A):
    X, C = A.box_mesh(2, 13, 0.2, seed=4)
    yield "tri", 2, X, C, 3
    X, C = A.box_mesh_q1(2, 17, 0.25, seed=5)
    yield "quad", 2, X, C, 4
    X, C = A.box_mesh_q1(3, 6, 0.2, seed=6)
    yield "hexa", 3, X, C, 8


@pytest.mark.parametrize("name", NAMES)
def test_manufactured_rhs_and_error_match_oracle(afem_gpu, name):
    A = afem_gpu
    for tag, dim, X, C, npc in meshes(A):
        print(f"[case] manufactured {name} {tag}", flush=True)
        m = A.Mesh(dim, X, C, npc)
        N = X.shape[0]
        ref = O.rhs_manufactured(dim, X, C, npc, name)
        rhs = A.DeviceArray(N, np.float64)
        rhs.zero()
        A.rhs_manufactured(m, name, rhs)
        got = rhs.numpy()
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), tag
        # component 1 of 2: the other component is left alone
        rhs2 = A.DeviceArray.from_numpy(np.full(2 * N, 7.0))
        A.rhs_manufactured(m, name, rhs2, nb_dof=2, dof_index=1)
        got2 = rhs2.numpy()
        assert np.all(got2[0::2] == 7.0)
        assert np.abs(got2[1::2] - ref).max() <= 1e-12 * np.abs(ref).max(), tag
        u = np.random.default_rng(9).uniform(-1.0, 2.0, N)
        e_ref = O.manufactured_l2_error(dim, X, C, npc, name, u)
        e = A.manufactured_l2_error(m, name, A.DeviceArray.from_numpy(u))
        assert abs(e - e_ref) <= 1e-12 * e_ref, tag


def test_manufactured_refusals(afem_gpu):
    A = afem_gpu
    X, C = A.box_mesh(3, 2)
    m = A.Mesh(3, X, C)
    rhs = A.DeviceArray(X.shape[0], np.float64)
    with pytest.raises(A.AfemError, match="not implemented"):
        A.rhs_manufactured(m, "sine", rhs)
    with pytest.raises(A.AfemError, match="not implemented"):
        A.manufactured_l2_error(m, "sine", rhs)
    with pytest.raises(ValueError):
        A.rhs_manufactured(m, "cosine", rhs)
    X, C = A.box_mesh_q1(2, 3)
    m2 = A.Mesh(2, X, C, 4)
    with pytest.raises(A.AfemError, match="unknown manufactured"):
        A.check(A.lib().afem_rhs_manufactured(m2.h, 3, 1, 0, rhs.c))
