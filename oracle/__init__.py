"""ctypes front-end of the CPU FP64 oracle (oracle/liboracle.so).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this package, and only as the checker or as the
timed CPU baseline.  The product library never calls it.

Each wrapper mirrors one C function of afem_oracle.h, which cites the
reference file:line it restates.
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

ELIMINATE_NONE, ELIMINATE_ROW, ELIMINATE_ROW_COLUMN = 0, 1, 2
LAPLACE, ELASTICITY, HEAT = 0, 1, 2


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _P(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        L.afo_area_tria3.restype = dbl
        L.afo_area_tria3.argtypes = [vp, vp, vp]
        L.afo_volume_tetra4.restype = dbl
        L.afo_volume_tetra4.argtypes = [vp, vp, vp, vp]
        L.afo_grad_tria3.argtypes = [vp, vp, vp]
        L.afo_grad_tetra4.argtypes = [vp, vp, vp, vp]
        L.afo_elem_laplace_tria3.argtypes = [vp, dbl, vp]
        L.afo_elem_laplace_tetra4.argtypes = [vp, dbl, vp]
        L.afo_elem_elasticity_tetra4.argtypes = [vp, dbl, dbl, vp]
        L.afo_elem_elasticity_tria3.argtypes = [vp, dbl, dbl, vp]
        L.afo_elem_elasticity_q1.argtypes = [ctypes.c_int, vp, dbl, dbl, vp]
        L.afo_lame.argtypes = [dbl, dbl, vp, vp]
        L.afo_node_cells.restype = ctypes.c_int
        L.afo_node_cells.argtypes = [i32, i32, i32, vp, vp, vp]
        L.afo_pattern.restype = i64
        L.afo_pattern.argtypes = [i32, i32, i32, vp, vp, vp, i64]
        L.afo_assemble.restype = ctypes.c_int
        L.afo_assemble.argtypes = [ctypes.c_int, i32, vp, i32, i32, vp, vp, vp,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, dbl, dbl, vp]
        L.afo_bsr_to_csr.argtypes = [i32, ctypes.c_int, vp, vp, vp, vp]
        L.afo_find_value_index.restype = i64
        L.afo_pcg_jacobi_threads.restype = ctypes.c_int
        L.afo_pcg_jacobi_threads.argtypes = [i32, vp, vp, vp, vp, vp, dbl, ctypes.c_int, ctypes.c_int, vp]
        L.afo_find_value_index.argtypes = [i32, ctypes.c_int, ctypes.c_int, vp, i64, vp, vp, i32, i32]
        L.afo_rhs_constant_source.argtypes = [ctypes.c_int, vp, i32, i32, vp, vp, dbl,
                                              ctypes.c_int, ctypes.c_int, vp]
        L.afo_rhs_body_force.argtypes = [vp, i32, vp, vp, vp]
        L.afo_rhs_mass_field.argtypes = [ctypes.c_int, vp, i32, i32, vp, vp, dbl, ctypes.c_int, ctypes.c_int, vp]
        L.afo_face_measure.restype = dbl
        L.afo_face_measure.argtypes = [ctypes.c_int, vp, vp]
        L.afo_rhs_face_uniform.argtypes = [ctypes.c_int, vp, i32, vp, dbl, ctypes.c_int, ctypes.c_int, vp]
        L.afo_elem_heat_q1.argtypes = [ctypes.c_int, vp, dbl, dbl, vp]
        L.afo_rhs_neumann.restype = ctypes.c_int
        L.afo_rhs_neumann.argtypes = [ctypes.c_int, vp, i32, vp, i32, vp, ctypes.c_int, vp, ctypes.c_int,
                                      ctypes.c_int, vp]
        L.afo_elem_laplace_q1.argtypes = [ctypes.c_int, vp, dbl, vp]
        L.afo_source_q1.argtypes = [ctypes.c_int, vp, dbl, vp]
        L.afo_q1_gradients.restype = dbl
        L.afo_q1_gradients.argtypes = [ctypes.c_int, vp, dbl, dbl, dbl, vp, vp, vp, vp]
        L.afo_rhs_neumann_q1.restype = ctypes.c_int
        L.afo_rhs_neumann_q1.argtypes = [ctypes.c_int, vp, i32, i32, vp, i32, vp, ctypes.c_int, vp, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, vp]
        L.afo_manufactured_exact.restype = dbl
        L.afo_manufactured_exact.argtypes = [ctypes.c_int, vp]
        L.afo_rhs_manufactured.restype = ctypes.c_int
        L.afo_rhs_manufactured.argtypes = [ctypes.c_int, vp, i32, i32, vp, ctypes.c_int, vp]
        L.afo_manufactured_l2_error.restype = dbl
        L.afo_manufactured_l2_error.argtypes = [ctypes.c_int, vp, i32, i32, vp, ctypes.c_int, vp]
        L.afo_convection_lhs.restype = ctypes.c_int
        L.afo_convection_lhs.argtypes = [ctypes.c_int, vp, i32, vp, dbl, vp, vp, vp]
        L.afo_apply_transformations.restype = ctypes.c_int
        L.afo_apply_transformations.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, vp]
        L.afo_spmv.argtypes = [i32, vp, vp, vp, vp, vp]
        L.afo_pcg_jacobi.restype = ctypes.c_int
        L.afo_pcg_jacobi.argtypes = [i32, vp, vp, vp, vp, vp, dbl, ctypes.c_int, vp]
        L.afo_count_differences.restype = i64
        L.afo_count_differences.argtypes = [i64, vp, vp, dbl, dbl]
        L.afo_wtime.restype = dbl
        L.afo_assemble_laplace_threads.restype = ctypes.c_int
        L.afo_assemble_laplace_threads.argtypes = [i32, vp, i32, i32, vp, vp, vp, dbl, vp, vp, ctypes.c_int]
        L.afo_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def area_tria3(n0, n1, n2):
    return lib().afo_area_tria3(_P(_f64(n0)), _P(_f64(n1)), _P(_f64(n2)))


def volume_tetra4(n0, n1, n2, n3):
    return lib().afo_volume_tetra4(_P(_f64(n0)), _P(_f64(n1)), _P(_f64(n2)), _P(_f64(n3)))


def grad_tria3(X):
    X = _f64(X)
    dx, dy = np.zeros(3), np.zeros(3)
    lib().afo_grad_tria3(_P(X), _P(dx), _P(dy))
    return dx, dy


def grad_tetra4(X):
    X = _f64(X)
    dx, dy, dz = np.zeros(4), np.zeros(4), np.zeros(4)
    lib().afo_grad_tetra4(_P(X), _P(dx), _P(dy), _P(dz))
    return dx, dy, dz


def elem_laplace(X, lam=1.0):
    X = _f64(X)
    n = X.shape[0]
    K = np.zeros((n, n))
    if n == 3:
        lib().afo_elem_laplace_tria3(_P(X), lam, _P(K))
    else:
        lib().afo_elem_laplace_tetra4(_P(X), lam, _P(K))
    return K


def elem_elasticity(X, lame_lambda, lame_mu):
    """tet4 (X (4, 3)): 12 x 12; tri3 (X (3, 3)): 6 x 6, DoFs interleaved per node."""
    X = _f64(X)
    if X.shape[0] == 3:
        K = np.zeros((6, 6))
        lib().afo_elem_elasticity_tria3(_P(X), lame_lambda, lame_mu, _P(K))
        return K
    K = np.zeros((12, 12))
    lib().afo_elem_elasticity_tetra4(_P(X), lame_lambda, lame_mu, _P(K))
    return K


def elem_elasticity_q1(dim, X, lame_lambda, lame_mu):
    X = _f64(X)
    nd = (4 if dim == 2 else 8) * dim
    K = np.zeros((nd, nd))
    lib().afo_elem_elasticity_q1(dim, _P(X), lame_lambda, lame_mu, _P(K))
    return K


def lame(E, nu):
    lam, mu = np.zeros(1), np.zeros(1)
    lib().afo_lame(E, nu, _P(lam), _P(mu))
    return float(lam[0]), float(mu[0])


def node_cells(nb_node, cells, npc):
    cells = _i32(cells)
    off = np.zeros(nb_node + 1, np.int32)
    items = np.zeros(max(1, cells.shape[0] * npc), np.int32)
    rc = lib().afo_node_cells(nb_node, cells.shape[0], npc, _P(cells), _P(off), _P(items))
    if rc != 0:
        raise ValueError("node id out of range")
    return off, items[: off[-1]]


def pattern(nb_node, cells, npc):
    cells = _i32(cells)
    row = np.zeros(nb_node + 1, np.int32)
    cap = int(cells.shape[0]) * npc * npc + nb_node
    cols = np.zeros(max(cap, 1), np.int32)
    nnz = lib().afo_pattern(nb_node, cells.shape[0], npc, _P(cells), _P(row), _P(cols), cap)
    if nnz < 0:
        raise ValueError("pattern build failed")
    return row, cols[:nnz].copy()


def assemble(dim, coords, cells, npc, row, cols, element=LAPLACE, nb_dof=1, order_per_block=True,
             cell_lambda=None, lame_lambda=0.0, lame_mu=0.0):
    coords, cells, row, cols = _f64(coords), _i32(cells), _i32(row), _i32(cols)
    lamv = None if cell_lambda is None else _f64(cell_lambda)
    vals = np.zeros(int(row[-1]) * nb_dof * nb_dof)
    rc = lib().afo_assemble(dim, coords.shape[0], _P(coords), cells.shape[0], npc, _P(cells),
                            _P(row), _P(cols), element, nb_dof, 1 if order_per_block else 0,
                            _P(lamv), lame_lambda, lame_mu, _P(vals))
    if rc == -2:
        raise ValueError("assembly: Q1 cells carry the Laplace element only")
    if rc != 0:
        raise ValueError("assembly: value not found in pattern")
    return vals


def bsr_to_csr(nb_node, nb, row, cols):
    row, cols = _i32(row), _i32(cols)
    csr_row = np.zeros(nb_node * nb + 1, np.int32)
    csr_col = np.zeros(int(row[-1]) * nb * nb, np.int32)
    lib().afo_bsr_to_csr(nb_node, nb, _P(row), _P(cols), _P(csr_row), _P(csr_col))
    return csr_row, csr_col


def rhs_constant_source(dim, coords, cells, npc, f, nb_dof=1, dof_index=0, cell_scale=None, rhs=None):
    coords, cells = _f64(coords), _i32(cells)
    if rhs is None:
        rhs = np.zeros(coords.shape[0] * nb_dof)
    sc = None if cell_scale is None else _f64(cell_scale)
    lib().afo_rhs_constant_source(dim, _P(coords), cells.shape[0], npc, _P(cells), _P(sc), f,
                                  nb_dof, dof_index, _P(rhs))
    return rhs


def rhs_neumann(dim, coords, cells, faces, value, nb_dof=1, dof_index=0, rhs=None):
    """faces: (nf, dim) node ids; value: a scalar g or a dim-vector."""
    coords, cells, faces = _f64(coords), _i32(cells), _i32(faces)
    vals = _f64(np.atleast_1d(np.asarray(value, dtype=np.float64)))
    if rhs is None:
        rhs = np.zeros(coords.shape[0] * nb_dof)
    rc = lib().afo_rhs_neumann(dim, _P(coords), cells.shape[0], _P(cells), faces.shape[0], _P(faces), vals.size,
                               _P(vals), nb_dof, dof_index, _P(rhs))
    if rc != 0:
        raise ValueError("a Neumann face is not a face of any cell")
    return rhs


def elem_laplace_q1(dim, X, lam=1.0):
    """Quad4 (dim 2, X (4, 3)) or Hexa8 (dim 3, X (8, 3)) Laplace element matrix."""
    X = _f64(X)
    n = 4 if dim == 2 else 8
    K = np.zeros(n * n)
    lib().afo_elem_laplace_q1(dim, _P(X), lam, _P(K))
    return K.reshape(n, n)


def source_q1(dim, X, qdot):
    X = _f64(X)
    b = np.zeros(4 if dim == 2 else 8)
    lib().afo_source_q1(dim, _P(X), qdot, _P(b))
    return b


def convection_lhs(dim, coords, faces, h, row, cols, values):
    """Adds the heat module's boundary convection matrix h * M_F to values (in place)."""
    coords, faces, row, cols = _f64(coords), _i32(faces), _i32(row), _i32(cols)
    assert values.dtype == np.float64 and values.flags.c_contiguous
    rc = lib().afo_convection_lhs(faces.shape[1], _P(coords), faces.shape[0], _P(faces), h, _P(row), _P(cols),
                                  _P(values))
    if rc != 0:
        raise ValueError("convection: a face pair is not in the pattern")
    return values


def arcane_face_order(faces, node_uid):
    """Face node lists in Arcane's face order: start at the node of smallest
    unique id, then walk towards its neighbour (in the face's cycle) of
    smaller unique id.  Arcane builds faces this way when it reads a mesh;
    the reference's Hexa8 Neumann normal (t1 x t2 of that order,
    femutils/ArcaneFemFunctionsGpu.cc:1146-1320) depends on it.  Arcane is a
    third-party dependency absent here; the rule is pinned by the reference's
    poisson_test_ref_sphere_neumann_3D_hexa.txt."""
    faces = np.asarray(faces)
    uid = np.asarray(node_uid)
    out = np.empty_like(faces)
    k = faces.shape[1]
    for r, f in enumerate(faces):
        u = uid[f]
        i = int(np.argmin(u))
        step = 1 if uid[f[(i + 1) % k]] < uid[f[(i - 1) % k]] else -1
        out[r] = [f[(i + step * j) % k] for j in range(k)]
    return out


def rhs_neumann_q1(dim, coords, cells, npc, faces, value, nb_dof=1, dof_index=0, outward=False, rhs=None):
    """Quad4 / Hexa8 meshes: faces (nf, 2) edges or (nf, 4) quadrangles."""
    coords, cells, faces = _f64(coords), _i32(cells), _i32(faces)
    vals = _f64(np.atleast_1d(np.asarray(value, dtype=np.float64)))
    if rhs is None:
        rhs = np.zeros(coords.shape[0] * nb_dof)
    rc = lib().afo_rhs_neumann_q1(dim, _P(coords), cells.shape[0], npc, _P(cells), faces.shape[0], _P(faces),
                                  vals.size, _P(vals), nb_dof, dof_index, 1 if outward else 0, _P(rhs))
    if rc != 0:
        raise ValueError("a Neumann face is not a face of any cell")
    return rhs


def rhs_face_uniform(coords, faces, g, nb_dof=1, dof_index=0, rhs=None):
    """rhs += g |F| / k at the k nodes of every face (the convection RHS)."""
    coords, faces = _f64(coords), _i32(faces)
    if rhs is None:
        rhs = np.zeros(coords.shape[0] * nb_dof)
    lib().afo_rhs_face_uniform(faces.shape[1], _P(coords), faces.shape[0], _P(faces), g, nb_dof, dof_index, _P(rhs))
    return rhs


def elem_heat_q1(dim, X, lam, dt):
    X = _f64(X)
    n = 4 if dim == 2 else 8
    K = np.zeros(n * n)
    lib().afo_elem_heat_q1(dim, _P(X), lam, dt, _P(K))
    return K.reshape(n, n)


MANUFACTURED = {"sine": 0, "poly": 1, "quadratic": 2}


def manufactured_exact(name, X):
    X = _f64(X)
    return np.array([lib().afo_manufactured_exact(MANUFACTURED[name], _P(np.ascontiguousarray(x))) for x in X])


def rhs_manufactured(dim, coords, cells, npc, name, rhs=None):
    coords, cells = _f64(coords), _i32(cells)
    if rhs is None:
        rhs = np.zeros(coords.shape[0])
    if lib().afo_rhs_manufactured(dim, _P(coords), cells.shape[0], npc, _P(cells), MANUFACTURED[name], _P(rhs)):
        raise ValueError("manufactured solutions cover tri3, Quad4 and Hexa8 cells")
    return rhs


def manufactured_l2_error(dim, coords, cells, npc, name, u):
    coords, cells, u = _f64(coords), _i32(cells), _f64(u)
    return float(lib().afo_manufactured_l2_error(dim, _P(coords), cells.shape[0], npc, _P(cells), MANUFACTURED[name],
                                                 _P(u)))


def max_threads():
    return int(lib().afo_max_threads())


def assemble_laplace_threads(coords, cells, npc, row, cols, f, nthreads):
    """Laplace values and constant-source RHS on nthreads threads (timing baseline)."""
    coords, cells, row, cols = _f64(coords), _i32(cells), _i32(row), _i32(cols)
    vals = np.empty(int(row[-1]))
    rhs = np.empty(coords.shape[0])
    rc = lib().afo_assemble_laplace_threads(coords.shape[0], _P(coords), cells.shape[0], npc, _P(cells), _P(row),
                                            _P(cols), f, _P(vals), _P(rhs), int(nthreads))
    if rc != 0:
        raise ValueError("assembly: value not found in pattern")
    return vals, rhs


def rhs_mass_field(coords, cells, npc, field, scale, nb_dof=1, dof_index=0, rhs=None, dim=None):
    """rhs += M (scale * field); dim defaults to 2 for 3-node and 3 for 4- and
    8-node cells (pass dim = 2 for Quad4)."""
    if dim is None:
        dim = 2 if npc == 3 else 3
    coords, cells, field = _f64(coords), _i32(cells), _f64(field)
    if rhs is None:
        rhs = np.zeros(coords.shape[0] * nb_dof)
    lib().afo_rhs_mass_field(dim, _P(coords), cells.shape[0], npc, _P(cells), _P(field), scale, nb_dof, dof_index,
                             _P(rhs))
    return rhs


def rhs_body_force(coords, cells, f, rhs=None):
    coords, cells = _f64(coords), _i32(cells)
    if rhs is None:
        rhs = np.zeros(coords.shape[0] * 3)
    fv = _f64(f)
    lib().afo_rhs_body_force(_P(coords), cells.shape[0], _P(cells), _P(fv), _P(rhs))
    return rhs


def apply_transformations(row, cols, vals, rhs, forced_info=None, forced_value=None,
                          elim_info=None, elim_value=None):
    """In place on vals and rhs (copies are made for non-contiguous inputs)."""
    n = len(row) - 1
    row, cols = _i32(row), _i32(cols)
    fi = None if forced_info is None else np.ascontiguousarray(forced_info, np.uint8)
    fv = None if forced_value is None else _f64(forced_value)
    ei = None if elim_info is None else np.ascontiguousarray(elim_info, np.uint8)
    ev = None if elim_value is None else _f64(elim_value)
    rc = lib().afo_apply_transformations(n, _P(row), _P(cols), _P(vals), _P(rhs),
                                         _P(fi), _P(fv), _P(ei), _P(ev))
    if rc != 0:
        raise ValueError("forced diagonal missing from the pattern")
    return vals, rhs


def find_value_index(rows, cols, row, col, nb_block=1, order_per_block=True, nnz_per_row=None):
    """Value index of scalar entry (row, col) as BSRMatrix::findValueIndex
    (femutils/BSRFormat.cc:79-106) finds it: rows holds one start per block
    row (the reference's layout, no closing entry); -1 when absent."""
    rows, cols = _i32(rows), _i32(cols)
    nzr = _i32(nnz_per_row) if nnz_per_row is not None else None
    return int(lib().afo_find_value_index(len(rows), nb_block, 1 if order_per_block else 0, _P(rows), len(cols),
                                          _P(cols), _P(nzr), int(row), int(col)))


def matrix_add_value(rows, cols, vals, row, col, value, set_value=False, **kw):
    """matrixAddValue / matrixSetValue (femutils/internal/CsrDoFLinearSystemImpl.h:56-63)
    in place on vals; returns False when the entry is not in the pattern."""
    k = find_value_index(rows, cols, row, col, **kw)
    if k < 0:
        return False
    vals[k] = value if set_value else vals[k] + value
    return True


def spmv(row, cols, vals, x):
    row, cols, vals, x = _i32(row), _i32(cols), _f64(vals), _f64(x)
    y = np.zeros(len(row) - 1)
    lib().afo_spmv(len(row) - 1, _P(row), _P(cols), _P(vals), _P(x), _P(y))
    return y


def pcg_jacobi(row, cols, vals, b, rtol=1e-10, max_iter=100000, x0=None):
    row, cols, vals, b = _i32(row), _i32(cols), _f64(vals), _f64(b)
    x = np.zeros(len(b)) if x0 is None else np.array(x0, dtype=np.float64)
    rr = np.zeros(1)
    it = lib().afo_pcg_jacobi(len(b), _P(row), _P(cols), _P(vals), _P(b), _P(x), rtol, max_iter, _P(rr))
    return x, it, float(rr[0])


def pcg_jacobi_threads(row, cols, vals, b, nthreads, rtol=1e-10, max_iter=100000, x0=None):
    """Jacobi-PCG on nthreads OpenMP threads (CPU timing baseline); returns
    (x, iterations or -1 when max_iter was reached, true relative residual)."""
    row, cols, vals, b = _i32(row), _i32(cols), _f64(vals), _f64(b)
    x = np.zeros(len(b)) if x0 is None else np.array(x0, dtype=np.float64)
    rr = np.zeros(1)
    it = lib().afo_pcg_jacobi_threads(len(b), _P(row), _P(cols), _P(vals), _P(b), _P(x), rtol, max_iter,
                                      int(nthreads), _P(rr))
    return x, it, float(rr[0])


def count_differences(ref, v, epsilon, min_value=0.0):
    ref, v = _f64(ref).ravel(), _f64(v).ravel()
    return int(lib().afo_count_differences(ref.size, _P(ref), _P(v), epsilon, min_value))


def wtime():
    return lib().afo_wtime()
