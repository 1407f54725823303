"""arcanefem_amd -- MI355X-native P1 assembly + PCG path of ArcaneFEM.

The compute path is libafem.so (hand-written HIP kernels for gfx950 behind the
C ABI of include/afem.h).  This module is a thin ctypes front-end used by the
tests and bench.py; it mirrors the reference's names (BSRFormat,
DoFLinearSystem) and raises AfemError on any non-zero status.  There is no
fallback: if libafem.so is missing or no GPU is present, calls fail loudly.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libafem.so")
_lib = None

# status codes (afem_status)
OK = 0
ERR_ARGUMENT, ERR_NOT_IMPLEMENTED, ERR_STATE, ERR_NOT_FOUND, ERR_MESH = 1, 2, 3, 4, 5
ERR_CAPACITY, ERR_NOT_CONVERGED, ERR_DEVICE, ERR_HIP, ERR_OUT_OF_MEMORY = 6, 7, 8, 9, 10

DIRICHLET_PENALTY, DIRICHLET_ROW_ELIMINATION, DIRICHLET_ROW_COLUMN_ELIMINATION = 0, 1, 2
SOLVER_PCG_FP64, SOLVER_PCG_MIXED, SOLVER_CG_FP64, SOLVER_FPCG_SPAI = 0, 1, 2, 3
PRECOND_JACOBI, PRECOND_NONE = 0, 1


class AfemError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"afem status {status}: {msg}")
        self.status = status


class BsrInfo(ctypes.Structure):
    _fields_ = [("nb_row", ctypes.c_int32), ("nb_block", ctypes.c_int32), ("nb_col", ctypes.c_int64),
                ("nb_non_zero_value", ctypes.c_int64), ("order_values_per_block", ctypes.c_int32),
                ("max_row_length", ctypes.c_int32)]


class SolverOptions(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("preconditioner", ctypes.c_int), ("rtol", ctypes.c_double),
                ("atol", ctypes.c_double), ("max_iter", ctypes.c_int), ("use_initial_guess", ctypes.c_int),
                ("check_every", ctypes.c_int), ("reliable_delta", ctypes.c_double), ("spmv_lanes", ctypes.c_int),
                ("norm", ctypes.c_int), ("precond_bf16", ctypes.c_int)]


class SolveResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int), ("reliable_updates", ctypes.c_int),
                ("relres", ctypes.c_double), ("solve_ms", ctypes.c_double), ("fallback", ctypes.c_int)]


class BatchOptions(ctypes.Structure):
    _fields_ = [("rtol", ctypes.c_double), ("atol", ctypes.c_double), ("max_iter", ctypes.c_int),
                ("reliable_delta", ctypes.c_double), ("chunk", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("use_initial_guess", ctypes.c_int), ("persistent", ctypes.c_int), ("workgroups", ctypes.c_int),
                ("spin_limit", ctypes.c_longlong), ("timeout_ms", ctypes.c_double)]


class BatchResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int), ("reliable_updates", ctypes.c_int),
                ("relres", ctypes.c_double)]


def _declare(L):
    vp, i32, i64, dbl, ci = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_int
    pp = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "afem_last_error": (ctypes.c_char_p, []),
        "afem_version": (ci, []),
        "afem_init": (ci, [ci]),
        "afem_finalize": (ci, []),
        "afem_synchronize": (ci, []),
        "afem_malloc": (ci, [pp, ctypes.c_size_t]),
        "afem_free": (ci, [vp]),
        "afem_memcpy_htod": (ci, [vp, vp, ctypes.c_size_t]),
        "afem_memcpy_dtoh": (ci, [vp, vp, ctypes.c_size_t]),
        "afem_memcpy_dtod": (ci, [vp, vp, ctypes.c_size_t]),
        "afem_memset": (ci, [vp, ci, ctypes.c_size_t]),
        "afem_stream_copy": (ci, [vp, vp, ctypes.c_size_t]),
        "afem_profile_enable": (ci, [ci]),
        "afem_profile_reset": (ci, []),
        "afem_profile_query": (ci, [ctypes.c_char_p, ctypes.POINTER(i64), ctypes.POINTER(dbl)]),
        "afem_box_mesh_size": (ci, [ci, ci, ctypes.POINTER(i32), ctypes.POINTER(i32)]),
        "afem_box_mesh": (ci, [ci, ci, dbl, ctypes.c_uint64, vp, vp]),
        "afem_mesh_create": (ci, [ci, i32, vp, i32, i32, vp, pp]),
        "afem_mesh_destroy": (ci, [vp]),
        "afem_mesh_node_cells": (ci, [vp, vp, vp]),
        "afem_bsr_create": (ci, [vp, ci, ci, ci, pp]),
        "afem_bsr_destroy": (ci, [vp]),
        "afem_bsr_compute_sparsity": (ci, [vp]),
        "afem_bsr_get_info": (ci, [vp, ctypes.POINTER(BsrInfo)]),
        "afem_bsr_reset_values": (ci, [vp]),
        "afem_bsr_assemble_laplace": (ci, [vp, vp, dbl, vp]),
        "afem_bsr_assemble_elasticity": (ci, [vp, dbl, dbl]),
        "afem_bsr_autotune_assembly": (ci, [vp, ctypes.POINTER(dbl), ctypes.POINTER(ci)]),
        "afem_bsr_set_assembly_variant": (ci, [vp, ci]),
        "afem_bsr_device_arrays": (ci, [vp, pp, pp, pp, pp]),
        "afem_bsr_download": (ci, [vp, vp, vp, vp, vp]),
        "afem_bsr_to_csr": (ci, [vp, pp, pp, ctypes.POINTER(i32)]),
        "afem_rhs_constant_source": (ci, [vp, dbl, vp, ci, ci, vp]),
        "afem_rhs_body_force": (ci, [vp, vp, vp]),
        "afem_rhs_neumann": (ci, [vp, ci, vp, ci, vp, ci, ci, vp]),
        "afem_rhs_mass_field": (ci, [vp, vp, ctypes.c_double, ci, ci, vp]),
        "afem_bsr_assemble_heat": (ci, [vp, vp, ctypes.c_double]),
        "afem_bsr_add_convection": (ci, [vp, ci, vp, ctypes.c_double]),
        "afem_rhs_face_uniform": (ci, [vp, ci, vp, ctypes.c_double, ci, ci, vp]),
        "afem_rhs_manufactured": (ci, [vp, ci, ci, ci, vp]),
        "afem_manufactured_l2_error": (ci, [vp, ci, vp, vp]),
        "afem_system_create": (ci, [i32, pp]),
        "afem_system_destroy": (ci, [vp]),
        "afem_system_set_csr_values": (ci, [vp, i32, i64, vp, vp, vp, vp]),
        "afem_system_set_bsr": (ci, [vp, vp]),
        "afem_system_rhs": (ci, [vp, pp]),
        "afem_system_solution": (ci, [vp, pp]),
        "afem_system_forced": (ci, [vp, pp, pp]),
        "afem_system_elimination": (ci, [vp, pp, pp]),
        "afem_system_apply_dirichlet": (ci, [vp, ci, i32, vp, ci, ci, dbl, dbl]),
        "afem_system_apply_matrix_transformation": (ci, [vp]),
        "afem_system_apply_rhs_transformation": (ci, [vp]),
        "afem_system_clear_values": (ci, [vp]),
        "afem_system_matrix_add_value": (ci, [vp, i32, i32, dbl, ci]),
        "afem_solver_default_options": (None, [ctypes.POINTER(SolverOptions)]),
        "afem_system_solve": (ci, [vp, ctypes.POINTER(SolverOptions), ctypes.POINTER(SolveResult)]),
        "afem_system_solve_noalloc": (ci, [vp, ctypes.POINTER(SolverOptions), ctypes.POINTER(SolveResult)]),
        "afem_system_spmv": (ci, [vp, vp, vp]),
        "afem_system_spmv_fp32": (ci, [vp, vp, vp]),
        "afem_system_spmv_format": (ci, [vp, ctypes.POINTER(ci), ctypes.POINTER(ci)]),
        "afem_system_spmv_tuning": (ci, [vp] + [ctypes.POINTER(ci)] * 4),
        "afem_system_set_spai": (ci, [vp, vp]),
        "afem_rff_parameters": (ci, [ci, ci, dbl, ctypes.c_uint64, vp, vp]),
        "afem_lognormal_field": (ci, [vp, dbl, dbl, ci, ctypes.c_uint64, vp]),
        "afem_spai_num_params": (ci, []),
        "afem_spai_create": (ci, [ctypes.c_uint64, pp]),
        "afem_spai_destroy": (ci, [vp]),
        "afem_spai_get_params": (ci, [vp, vp, ci]),
        "afem_spai_set_params": (ci, [vp, vp, ci]),
        "afem_spai_loss": (ci, [vp, vp, ctypes.c_uint64, ctypes.POINTER(dbl)]),
        "afem_spai_loss_grad": (ci, [vp, vp, ctypes.c_uint64, ctypes.POINTER(dbl), vp]),
        "afem_spai_adam_step": (ci, [vp, dbl]),
        "afem_spai_train_step": (ci, [vp, vp, ctypes.c_uint64, dbl, ctypes.POINTER(dbl)]),
        "afem_spai_apply": (ci, [vp, vp]),
        "afem_spai_inspect": (ci, [vp, ci, vp, i64]),
        "afem_batch_default_options": (None, [ctypes.POINTER(BatchOptions)]),
        "afem_batch_create": (ci, [vp, ci, vp, pp]),
        "afem_batch_destroy": (ci, [vp]),
        "afem_batch_persistent_mode": (ci, [vp, ctypes.POINTER(ci)]),
        "afem_batch_set_tiles": (ci, [vp, ci, ci]),
        "afem_batch_autotune": (ci, [vp, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(dbl)]),
        "afem_batch_solve": (ci, [vp, ctypes.POINTER(BatchOptions), vp, ctypes.POINTER(dbl)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return sig


def lib():
    """Loads libafem.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AfemError(ERR_DEVICE, f"{LIB_PATH} is missing: run python -m arcanefem_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        raise AfemError(rc, lib().afem_last_error().decode(errors="replace"))


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


# ---- device memory --------------------------------------------------------------------

class DeviceArray:
    """Device allocation of a fixed dtype/shape with numpy upload/download."""

    def __init__(self, shape, dtype, ptr=None, owner=True):
        self.shape = tuple(np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.owner = owner
        if ptr is None:
            p = ctypes.c_void_p()
            check(lib().afem_malloc(ctypes.byref(p), self.nbytes))
            self.ptr = p.value
        else:
            self.ptr = ptr

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    @classmethod
    def view(cls, ptr, shape, dtype):
        return cls(shape, dtype, ptr=ptr, owner=False)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        check(lib().afem_memcpy_htod(ctypes.c_void_p(self.ptr), _ptr(a), self.nbytes))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().afem_memcpy_dtoh(_ptr(out), ctypes.c_void_p(self.ptr), self.nbytes))
        return out

    def zero(self):
        check(lib().afem_memset(ctypes.c_void_p(self.ptr), 0, self.nbytes))

    def copy_from(self, other):
        assert other.nbytes == self.nbytes, (other.nbytes, self.nbytes)
        check(lib().afem_memcpy_dtod(ctypes.c_void_p(self.ptr), ctypes.c_void_p(other.ptr), self.nbytes))

    @property
    def c(self):
        return ctypes.c_void_p(self.ptr)

    def free(self):
        if self.owner and self.ptr:
            check(lib().afem_free(ctypes.c_void_p(self.ptr)))
        self.ptr = None

    def __del__(self):
        try:
            if self.owner and self.ptr and _lib is not None:
                _lib.afem_free(ctypes.c_void_p(self.ptr))
        except Exception:
            pass


def init(device=0):
    check(lib().afem_init(device))


def synchronize():
    check(lib().afem_synchronize())


# ---- synthetic meshes (host) -------------------------------------------------------------

def box_mesh(dim, n, jitter=0.0, seed=0):
    """Jittered structured tri (dim 2) / Kuhn-tet (dim 3) mesh of the unit box."""
    nn, nc = ctypes.c_int32(), ctypes.c_int32()
    check(lib().afem_box_mesh_size(dim, n, ctypes.byref(nn), ctypes.byref(nc)))
    coords = np.zeros((nn.value, 3), np.float64)
    cells = np.zeros((nc.value, 4), np.int32)
    check(lib().afem_box_mesh(dim, n, float(jitter), ctypes.c_uint64(seed), _ptr(coords), _ptr(cells)))
    return coords, cells


def box_mesh_q1(dim, n, jitter=0.0, seed=0):
    """Jittered structured Quad4 (dim 2) / Hexa8 (dim 3) mesh of the unit box,
    n cells per side: nodes x-fastest, quads counter-clockwise from the
    lower-left corner, hexas the bottom face counter-clockwise then the top
    face (Gmsh / Arcane local order).  Interior nodes move uniformly by up to
    jitter * h per coordinate."""
    rng = np.random.default_rng(seed)
    g = np.linspace(0.0, 1.0, n + 1)
    h = 1.0 / n
    if dim == 2:
        y, x = np.meshgrid(g, g, indexing="ij")
        X = np.stack([x.ravel(), y.ravel(), np.zeros(x.size)], axis=1)
        i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
        v = (j * (n + 1) + i).ravel()
        C = np.stack([v, v + 1, v + n + 2, v + n + 1], axis=1)
    else:
        z, y, x = np.meshgrid(g, g, g, indexing="ij")
        X = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
        k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        v = (k * (n + 1) ** 2 + j * (n + 1) + i).ravel()
        s, p = n + 1, (n + 1) ** 2
        C = np.stack([v, v + 1, v + s + 1, v + s, v + p, v + p + 1, v + p + s + 1, v + p + s], axis=1)
    if jitter > 0.0:
        inner = np.all((X > 0.0) & (X < 1.0), axis=1) if dim == 3 else np.all((X[:, :2] > 0.0) & (X[:, :2] < 1.0),
                                                                                axis=1)
        d = rng.uniform(-jitter * h, jitter * h, size=(X.shape[0], 3))
        if dim == 2:
            d[:, 2] = 0.0
        X[inner] += d[inner]
    return X, C.astype(np.int32)


# ---- mesh / BSRFormat / DoFLinearSystem ------------------------------------------------------

class Mesh:
    """Device mesh: tri3 / tet4 cells (dim + 1 nodes, stride 4), Quad4 (dim 2,
    npc 4) or Hexa8 (dim 3, npc 8, stride 8).  npc defaults to dim + 1, or 8
    for 8-column cells."""

    def __init__(self, dim, coords, cells, npc=None):
        coords = np.ascontiguousarray(coords, np.float64)
        cells = np.ascontiguousarray(cells, np.int32)
        if npc is None:
            npc = 8 if cells.shape[1] == 8 else dim + 1
        assert coords.ndim == 2 and coords.shape[1] == 3 and cells.ndim == 2 and cells.shape[1] == max(4, npc)
        self.dim = dim
        self.nb_node = coords.shape[0]
        self.nb_cell = cells.shape[0]
        self.npc = npc
        h = ctypes.c_void_p()
        check(lib().afem_mesh_create(dim, self.nb_node, _ptr(coords), self.nb_cell, self.npc, _ptr(cells),
                                     ctypes.byref(h)))
        self.h = h

    def node_cells(self):
        off = np.zeros(self.nb_node + 1, np.int32)
        check(lib().afem_mesh_node_cells(self.h, _ptr(off), None))
        items = np.zeros(max(1, int(off[-1])), np.int32)
        check(lib().afem_mesh_node_cells(self.h, None, _ptr(items)))
        return off, items[: off[-1]]

    def __del__(self):
        try:
            if getattr(self, "h", None) and _lib is not None:
                _lib.afem_mesh_destroy(self.h)
        except Exception:
            pass


class BSRFormat:
    """Mirror of femutils/BSRFormat.h: initialize in the constructor."""

    def __init__(self, mesh, nb_dof=1, use_csr=False, use_atomic_free=True):
        self.mesh = mesh
        h = ctypes.c_void_p()
        check(lib().afem_bsr_create(mesh.h, nb_dof, 1 if use_csr else 0, 1 if use_atomic_free else 0,
                                    ctypes.byref(h)))
        self.h = h

    def info(self):
        i = BsrInfo()
        check(lib().afem_bsr_get_info(self.h, ctypes.byref(i)))
        return i

    def compute_sparsity(self):
        check(lib().afem_bsr_compute_sparsity(self.h))

    def reset_values(self):
        check(lib().afem_bsr_reset_values(self.h))

    def assemble_laplace(self, cell_lambda=None, source_f=0.0, rhs=None):
        lam = cell_lambda.c if cell_lambda is not None else None
        check(lib().afem_bsr_assemble_laplace(self.h, lam, float(source_f), rhs.c if rhs is not None else None))

    def assemble_heat(self, cell_lambda, dt):
        """lambda K + M / dt (cell_lambda a DeviceArray or None for 1)."""
        lam = cell_lambda.c if cell_lambda is not None else None
        check(lib().afem_bsr_assemble_heat(self.h, lam, float(dt)))

    def add_convection(self, faces, h):
        """Adds h times the consistent mass matrix of each boundary face
        ((nf, dim) node ids): the heat module's convection matrix."""
        faces = np.ascontiguousarray(faces, np.int32)
        check(lib().afem_bsr_add_convection(self.h, int(faces.shape[0]), _ptr(faces), float(h)))

    def assemble_elasticity(self, lame_lambda, lame_mu):
        check(lib().afem_bsr_assemble_elasticity(self.h, float(lame_lambda), float(lame_mu)))

    def autotune_assembly(self):
        ms, v = ctypes.c_double(), ctypes.c_int()
        check(lib().afem_bsr_autotune_assembly(self.h, ctypes.byref(ms), ctypes.byref(v)))
        return v.value, ms.value

    def set_assembly_variant(self, variant):
        check(lib().afem_bsr_set_assembly_variant(self.h, int(variant)))

    def download(self):
        i = self.info()
        row = np.zeros(i.nb_row + 1, np.int32)
        col = np.zeros(i.nb_col, np.int32)
        val = np.zeros(i.nb_non_zero_value, np.float64)
        nzr = np.zeros(i.nb_row, np.int32)
        check(lib().afem_bsr_download(self.h, _ptr(row), _ptr(col), _ptr(val), _ptr(nzr)))
        return row, col, val, nzr

    def device_arrays(self):
        r, c, v, z = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().afem_bsr_device_arrays(self.h, ctypes.byref(r), ctypes.byref(c), ctypes.byref(v),
                                           ctypes.byref(z)))
        return r.value, c.value, v.value, z.value

    def to_csr(self):
        r, c, n = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_int32()
        check(lib().afem_bsr_to_csr(self.h, ctypes.byref(r), ctypes.byref(c), ctypes.byref(n)))
        i = self.info()
        row = DeviceArray.view(r.value, n.value + 1, np.int32).numpy()
        col = DeviceArray.view(c.value, i.nb_non_zero_value, np.int32).numpy()
        return row, col

    def __del__(self):
        try:
            if getattr(self, "h", None) and _lib is not None:
                _lib.afem_bsr_destroy(self.h)
        except Exception:
            pass


def solver_options(**kw):
    o = SolverOptions()
    lib().afem_solver_default_options(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class DoFLinearSystem:
    """Mirror of the IDoFLinearSystemImpl created by the MI355X factory."""

    def __init__(self, nb_dof):
        h = ctypes.c_void_p()
        check(lib().afem_system_create(nb_dof, ctypes.byref(h)))
        self.h = h
        self.nb_dof = nb_dof

    def _arr(self, fn, dtype):
        p = ctypes.c_void_p()
        check(fn(self.h, ctypes.byref(p)))
        return DeviceArray.view(p.value, self.nb_dof, dtype)

    def rhs(self):
        return self._arr(lib().afem_system_rhs, np.float64)

    def solution(self):
        return self._arr(lib().afem_system_solution, np.float64)

    def forced(self):
        i, v = ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().afem_system_forced(self.h, ctypes.byref(i), ctypes.byref(v)))
        return DeviceArray.view(i.value, self.nb_dof, np.uint8), DeviceArray.view(v.value, self.nb_dof, np.float64)

    def elimination(self):
        i, v = ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().afem_system_elimination(self.h, ctypes.byref(i), ctypes.byref(v)))
        return DeviceArray.view(i.value, self.nb_dof, np.uint8), DeviceArray.view(v.value, self.nb_dof, np.float64)

    def set_csr_values(self, rows, col, val, rows_nb_column=None):
        """setCSRValues with the reference's view layout: rows holds one start
        per row (the last row ends at the value count; a closing entry equal
        to it is accepted), rows_nb_column the per-row counts or None.  All
        DeviceArray; columns and values are kept alive by the caller until
        solve, the row starts are copied."""
        nb_row = int(np.prod(rows.shape))
        cnt = rows_nb_column.c if rows_nb_column is not None else None
        check(lib().afem_system_set_csr_values(self.h, nb_row, int(np.prod(col.shape)), rows.c, cnt, col.c, val.c))

    def set_bsr(self, bsr):
        check(lib().afem_system_set_bsr(self.h, bsr.h))

    def apply_dirichlet(self, method, nodes, nb_dof_per_node=1, dof_index=0, value=0.0, penalty=1e30):
        nodes = np.ascontiguousarray(nodes, np.int32)
        check(lib().afem_system_apply_dirichlet(self.h, method, len(nodes), _ptr(nodes), nb_dof_per_node,
                                                dof_index, float(value), float(penalty)))

    def apply_matrix_transformation(self):
        check(lib().afem_system_apply_matrix_transformation(self.h))

    def apply_rhs_transformation(self):
        check(lib().afem_system_apply_rhs_transformation(self.h))

    def clear_values(self):
        check(lib().afem_system_clear_values(self.h))

    def solve(self, raise_on_failure=True, **kw):
        o = solver_options(**kw)
        r = SolveResult()
        rc = lib().afem_system_solve(self.h, ctypes.byref(o), ctypes.byref(r))
        if rc != OK and (raise_on_failure or rc != ERR_NOT_CONVERGED):
            check(rc)
        return r

    def solve_noalloc(self, raise_on_failure=True, **kw):
        """The allocation-free FP64 solve (afem_system_solve_noalloc): at most
        1048576 DoFs, FP64 methods only, CSR / 3x3 BSR SpMV on the view."""
        o = solver_options(**kw)
        r = SolveResult()
        rc = lib().afem_system_solve_noalloc(self.h, ctypes.byref(o), ctypes.byref(r))
        if rc != OK and (raise_on_failure or rc != ERR_NOT_CONVERGED):
            check(rc)
        return r

    def spmv(self, x, y):
        check(lib().afem_system_spmv(self.h, x.c, y.c))

    def spmv_fp32(self, x, y):
        check(lib().afem_system_spmv_fp32(self.h, x.c, y.c))

    def set_spai(self, values):
        """Installs a sparse preconditioner with the matrix pattern (DeviceArray of nnz float32)."""
        check(lib().afem_system_set_spai(self.h, values.c))

    def spmv_format(self):
        """('csr', lanes), ('sell', 64), ('sell16', 64) or ('sell16x8', 64) as chosen by the SpMV autotuner;
        None before any SpMV."""
        f, lanes = ctypes.c_int(), ctypes.c_int()
        check(lib().afem_system_spmv_format(self.h, ctypes.byref(f), ctypes.byref(lanes)))
        if f.value < 0:
            return None
        return {1: ("sell", 64), 2: ("sell16", 64), 3: ("sell16x8", 64)}.get(f.value, ("csr", lanes.value))

    def spmv_tuning(self):
        """((slots, workgroup) of the FP64 SpMV, (slots, workgroup) of the FP32 SpMV) for the
        16-bit-offset SELL formats; zeros otherwise."""
        v = [ctypes.c_int() for _ in range(4)]
        check(lib().afem_system_spmv_tuning(self.h, *[ctypes.byref(c) for c in v]))
        return (v[0].value, v[1].value), (v[2].value, v[3].value)

    def __del__(self):
        try:
            if getattr(self, "h", None) and _lib is not None:
                _lib.afem_system_destroy(self.h)
        except Exception:
            pass


def rhs_neumann(mesh, faces, value, rhs, nb_dof=1, dof_index=0):
    """Adds the Neumann condition on the boundary faces `faces` ((nf, dim) node
    ids) to the device vector `rhs`: a scalar flux or a dim-vector g."""
    faces = np.ascontiguousarray(faces, np.int32)
    vals = np.ascontiguousarray(np.atleast_1d(np.asarray(value, np.float64)))
    check(lib().afem_rhs_neumann(mesh.h, int(faces.shape[0]), _ptr(faces), int(vals.size), _ptr(vals), nb_dof,
                                 dof_index, rhs.c))


def rhs_face_uniform(mesh, faces, g, rhs, nb_dof=1, dof_index=0):
    """rhs += g |F| / k at the k nodes of every face (the convection RHS)."""
    faces = np.ascontiguousarray(faces, np.int32)
    check(lib().afem_rhs_face_uniform(mesh.h, int(faces.shape[0]), _ptr(faces), float(g), nb_dof, dof_index,
                                      rhs.c))


MANUFACTURED = {"sine": 0, "poly": 1, "quadratic": 2}


def rhs_manufactured(mesh, solution, rhs, nb_dof=1, dof_index=0):
    """rhs = the manufactured source `solution` ("sine", "poly", "quadratic")
    integrated against the shape functions (tri3, Quad4, Hexa8 meshes)."""
    if solution not in MANUFACTURED:
        raise ValueError(f"unknown manufactured solution {solution!r}")
    check(lib().afem_rhs_manufactured(mesh.h, MANUFACTURED[solution], nb_dof, dof_index, rhs.c))


def manufactured_l2_error(mesh, solution, u):
    """L2 error of the nodal solution u (a DeviceArray) against `solution`."""
    if solution not in MANUFACTURED:
        raise ValueError(f"unknown manufactured solution {solution!r}")
    out = ctypes.c_double(0.0)
    check(lib().afem_manufactured_l2_error(mesh.h, MANUFACTURED[solution], u.c, ctypes.byref(out)))
    return out.value


def rhs_mass_field(mesh, field, scale, rhs, nb_dof=1, dof_index=0):
    """rhs += M (scale * field), field a DeviceArray of one value per node."""
    check(lib().afem_rhs_mass_field(mesh.h, field.c, float(scale), nb_dof, dof_index, rhs.c))


def rff_parameters(dim, nfeatures, corr_length, seed):
    """(omega (K, 3), phase (K,)) of the random-Fourier-feature field generator."""
    om = np.zeros((nfeatures, 3))
    ph = np.zeros(nfeatures)
    check(lib().afem_rff_parameters(dim, nfeatures, corr_length, seed, _ptr(om), _ptr(ph)))
    return om, ph


def lognormal_field(mesh, sigma=1.0, corr_length=0.1, nfeatures=512, seed=0, out=None):
    """Per-cell log-normal conductivity exp(g) on the device (DeviceArray of nb_cell doubles)."""
    if out is None:
        out = DeviceArray(mesh.nb_cell, np.float64)
    check(lib().afem_lognormal_field(mesh.h, sigma, corr_length, nfeatures, seed, out.c))
    return out


class LearnedSpai:
    """The learned sparse approximate inverse (message-passing network on the
    Jacobi-scaled matrix graph; BF16 MFMA forward/backward kernels)."""

    INSPECT = {"shat": 0, "features": 1, "h3": 2, "u": 3, "mhat": 4, "m": 5, "scale": 6, "h0": 7}

    def __init__(self, seed=0):
        h = ctypes.c_void_p()
        check(lib().afem_spai_create(seed, ctypes.byref(h)))
        self.h = h
        self.n_params = lib().afem_spai_num_params()

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        check(lib().afem_spai_get_params(self.h, _ptr(p), self.n_params))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().afem_spai_set_params(self.h, _ptr(p), p.size))

    def loss(self, sysm, probe_seed=0):
        out = ctypes.c_double()
        check(lib().afem_spai_loss(self.h, sysm.h, probe_seed, ctypes.byref(out)))
        return out.value

    def loss_grad(self, sysm, probe_seed=0):
        out = ctypes.c_double()
        g = np.zeros(self.n_params, np.float32)
        check(lib().afem_spai_loss_grad(self.h, sysm.h, probe_seed, ctypes.byref(out), _ptr(g)))
        return out.value, g

    def adam_step(self, lr):
        check(lib().afem_spai_adam_step(self.h, lr))

    def train_step(self, sysm, probe_seed, lr):
        out = ctypes.c_double()
        check(lib().afem_spai_train_step(self.h, sysm.h, probe_seed, lr, ctypes.byref(out)))
        return out.value

    def apply(self, sysm):
        check(lib().afem_spai_apply(self.h, sysm.h))

    def inspect(self, what, count, dtype=np.float32):
        out = np.zeros(count, np.float32)
        check(lib().afem_spai_inspect(self.h, self.INSPECT[what], _ptr(out), count))
        return out

    def __del__(self):
        try:
            if getattr(self, "h", None) and _lib is not None:
                _lib.afem_spai_destroy(self.h)
        except Exception:
            pass
        self.h = None


def union_mesh(meshes):
    """Disjoint union of (coords, cells) pairs: concatenated arrays with the
    node ids of each mesh offset, and the node offsets of the parts."""
    coords, cells, offs = [], [], [0]
    for X, C in meshes:
        Cn = C.copy()
        Cn[Cn >= 0] += offs[-1]
        coords.append(X)
        cells.append(Cn)
        offs.append(offs[-1] + X.shape[0])
    return np.concatenate(coords), np.concatenate(cells), np.array(offs, np.int32)


class BatchSolver:
    """Independent systems on the diagonal blocks of one system's matrix,
    solved together (segmented mixed-precision PCG: hipGraph chunks of
    iterations by default; one persistent launch with persistent=1, which
    keeps each system's team on one XCD where the device allows it, or
    persistent=2, which exchanges through the coherence point)."""

    def __init__(self, sysm, segment_rows):
        seg = np.ascontiguousarray(segment_rows, dtype=np.int32)
        h = ctypes.c_void_p()
        check(lib().afem_batch_create(sysm.h, len(seg) - 1, _ptr(seg), ctypes.byref(h)))
        self.h = h
        self.nb_segment = len(seg) - 1
        self._sys = sysm

    def set_tiles(self, lanes, tiles):
        check(lib().afem_batch_set_tiles(self.h, lanes, tiles))

    def autotune(self):
        lanes, tiles, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(lib().afem_batch_autotune(self.h, ctypes.byref(lanes), ctypes.byref(tiles), ctypes.byref(ms)))
        return lanes.value, tiles.value, ms.value

    def solve(self, raise_on_failure=True, **kw):
        o = BatchOptions()
        lib().afem_batch_default_options(ctypes.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        res = (BatchResult * self.nb_segment)()
        ms = ctypes.c_double()
        rc = lib().afem_batch_solve(self.h, ctypes.byref(o), ctypes.cast(res, ctypes.c_void_p), ctypes.byref(ms))
        if rc != OK and (raise_on_failure or rc != ERR_NOT_CONVERGED):
            check(rc)
        return list(res), ms.value

    def persistent_mode(self):
        """Exchange of the last persistent solve: 1 XCD-local teams, 2
        coherence point, 0 none yet."""
        m = ctypes.c_int()
        check(lib().afem_batch_persistent_mode(self.h, ctypes.byref(m)))
        return m.value

    def __del__(self):
        try:
            if getattr(self, "h", None) and _lib is not None:
                _lib.afem_batch_destroy(self.h)
        except Exception:
            pass
        self.h = None


def heat_batch_steps(mesh, sysm, batch, source, dt, nb_step, **solve_kw):
    """Backward-Euler time loop of the heat module (modules/heat/FemModule.cc:
    74-110, 344-362) over every system of a batch at once.  The matrix
    lambda K + M/dt is assembled on the union mesh, set on sysm and
    transformed (Dirichlet) by the caller; the solution array holds the
    initial temperature.  Each step: rhs = source + M u_old / dt, the RHS
    transformation, one batched solve started from u_old.  source: a
    DeviceArray (the constant f integrated once) or None.  Returns the batch
    results and solve time of every step."""
    out = []
    rhs, x = sysm.rhs(), sysm.solution()
    for _ in range(nb_step):
        if source is not None:
            rhs.copy_from(source)
        else:
            rhs.zero()
        rhs_mass_field(mesh, x, 1.0 / dt, rhs)
        sysm.apply_rhs_transformation()
        out.append(batch.solve(use_initial_guess=1, **solve_kw))
    return out


def profile_enable(on=True):
    check(lib().afem_profile_enable(1 if on else 0))


def profile_reset():
    check(lib().afem_profile_reset())


def profile_query(kernel):
    n, t = ctypes.c_int64(), ctypes.c_double()
    check(lib().afem_profile_query(kernel.encode(), ctypes.byref(n), ctypes.byref(t)))
    return n.value, t.value
