"""ctypes binding of ``libdiffphd.so`` (the C-ABI declared in include/diffphd.h).

The library is required: importing a compute entry point without it raises
``NativeLibraryMissing``; there is no fallback path.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIBRARY = Path(__file__).resolve().parent / "libdiffphd.so"

DPHD_OK = 0
ERRORS = {1: ValueError, 2: RuntimeError, 3: RuntimeError, 4: NotImplementedError, 5: MemoryError}


class NativeLibraryMissing(RuntimeError):
    """libdiffphd.so has not been built (run __graft_entry__.build())."""


class Config(ctypes.Structure):
    _fields_ = [
        ("element_kind", ctypes.c_int),
        ("material", ctypes.c_int),
        ("anderson_window", ctypes.c_int),
        ("max_iterations", ctypes.c_int),
        ("abs_tol", ctypes.c_double),
        ("rel_tol", ctypes.c_double),
        ("factor_precision", ctypes.c_int),
        ("integrate_pinned_velocities", ctypes.c_int),
        ("jacobian_mode", ctypes.c_int),
        ("iteration_chunk", ctypes.c_int),
        ("slice_rows", ctypes.c_int),
    ]


class Info(ctypes.Structure):
    _fields_ = [
        ("bodies", ctypes.c_int),
        ("vertices", ctypes.c_int),
        ("elements", ctypes.c_int),
        ("samples_per_element", ctypes.c_int),
        ("corners", ctypes.c_int),
        ("factor_nnz", ctypes.c_int64),
        ("matrix_nnz", ctypes.c_int64),
        ("jacobian_bytes", ctypes.c_int64),
        ("setup_seconds", ctypes.c_double),
        ("factor_values", ctypes.c_int64),
        ("factor_slots", ctypes.c_int64),
        ("slice_rows", ctypes.c_int),
        ("slice_switch", ctypes.c_int),
        ("tune_seconds", ctypes.c_double),
        ("split_quads", ctypes.c_int),
        ("split_switch", ctypes.c_int),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double

SIGNATURES = {
    "dphd_last_error": ([], ctypes.c_char_p),
    "dphd_version": ([], _I),
    "dphd_factor_create": ([_I, _P, _P, _P, ctypes.POINTER(_P)], _I),
    "dphd_factor_sizes": ([_P, _P, _P, _P], _I),
    "dphd_factor_export": ([_P, _P, _P, _P, _P, _P, _P, _P], _I),
    "dphd_factor_destroy": ([_P], None),
    "dphd_build_incidence": ([_I, _I, _I, _P, _P, _P], _I),
    "dphd_tet_geometry": ([_I, _P, _P, _P], _I),
    "dphd_hex_gradients": ([_D, _P], _I),
    "dphd_assemble": ([_I, _I, _I, _I, _P, _P, _I, _P, _P, _P, _P, _D, _I, _P, _P, _P, _P, _P, _P, _I], _I),
    "dphd_create": ([ctypes.POINTER(Config), _I, ctypes.POINTER(_P)], _I),
    "dphd_destroy": ([_P], None),
    "dphd_set_bodies": ([_P, _I, _P, _P], _I),
    "dphd_set_tet4": ([_P, _P, _P], _I),
    "dphd_set_hex8": ([_P, _P, _P], _I),
    "dphd_set_particles": ([_P, _P, _P, _P], _I),
    "dphd_set_materials": ([_P, _P], _I),
    "dphd_set_muscles": ([_P, _P, _P], _I),
    "dphd_set_planar": ([_P, _I, _P, _P, _P], _I),
    "dphd_set_mesh_contact": ([_P, _D, _D, _D, _D, _D, _I, _P, _I, _P], _I),
    "dphd_mesh_contact_force": ([_P, _P, _P, _P, _P], _I),
    "dphd_mesh_contact_pairs": ([_P, ctypes.POINTER(_I)], _I),
    "dphd_mesh_contact_pair_counts": ([_P, _P], _I),
    "dphd_set_frictional_boundary": ([_P, _I, _P, _I, _P, _D, _D, _I, _I, _I, _P], _I),
    "dphd_boundary_parameters": ([_P, _P, _P], _I),
    "dphd_contact_state": ([_P, _P, _P, _P, _P, _P], _I),
    "dphd_contact_record": ([_P, _P, _P, _P], _I),
    "dphd_set_tolerances": ([_P, _I, _D, _D], _I),
    "dphd_build": ([_P, _D], _I),
    "dphd_get_info": ([_P, ctypes.POINTER(Info)], _I),
    "dphd_body_nnz": ([_P, _P], _I),
    "dphd_loop_stats": ([_P, _P], _I),
    "dphd_step": ([_P, _P, _P, _P, _P, _P, _P, _P, _P, _P], _I),
    "dphd_backward": ([_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P], _I),
    "dphd_backward_static": ([_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P], _I),
    "dphd_read_status": ([_P, _I, ctypes.POINTER(_I)], _I),
    "dphd_iterations": ([_P, _I, _P], _I),
    "dphd_synchronize": ([_P], _I),
    "dphd_apply_inverse": ([_P, _P, _P, _P], _I),
    "dphd_deformation_determinants": ([_P, _P, _P, _P], _I),
    "dphd_eval_projections": ([_I, _P, _P, _D, _D, _D, _D, _P, _P, _P, _P, _P, _P, _P], _I),
    "dphd_eval_jvps": ([_I, _P, _P, _P, _D, _D, _D, _P, _P, _P, _P, _P], _I),
}

_lib = None


ABI_VERSION = 6  # DPHD_ABI_VERSION of include/diffphd.h


def lib():
    """Load the native library once; raise loudly when it is absent."""
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            raise NativeLibraryMissing(f"{LIBRARY} is missing; build it with __graft_entry__.build()")
        handle = ctypes.CDLL(str(LIBRARY))
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = res
        if handle.dphd_version() != ABI_VERSION:
            raise NativeLibraryMissing(
                f"{LIBRARY} has ABI {handle.dphd_version()}, this package needs {ABI_VERSION}; rebuild it"
            )
        _lib = handle
    return _lib


def check(code: int, what: str = "") -> None:
    if code != DPHD_OK:
        message = lib().dphd_last_error().decode()
        raise ERRORS.get(code, RuntimeError)(f"{what}: {message}" if what else message)


def ptr(array) -> ctypes.c_void_p | None:
    """Host address of a contiguous NumPy array, or device address of a torch tensor."""
    if array is None:
        return None
    if isinstance(array, np.ndarray):
        if not array.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return ctypes.c_void_p(array.ctypes.data)
    if not array.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(array.data_ptr())


class Factor:
    """Exact inverse factor of one SPD matrix (replaces ``SparseInverse`` setup, DP/sparse.py:291-320)."""

    def __init__(self, offsets, columns, values):
        offsets = np.ascontiguousarray(offsets, dtype=np.int32)
        columns = np.ascontiguousarray(columns, dtype=np.int32)
        values = np.ascontiguousarray(values, dtype=np.float64)
        n = len(offsets) - 1
        handle = ctypes.c_void_p()
        check(lib().dphd_factor_create(n, ptr(offsets), ptr(columns), ptr(values), ctypes.byref(handle)), "factor")
        try:
            z, fz, mz = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
            check(lib().dphd_factor_sizes(handle, ctypes.byref(z), ctypes.byref(fz), ctypes.byref(mz)))
            z = z.value
            self.permutation = np.zeros(n, np.int32)
            self.s = (np.zeros(n + 1, np.int32), np.zeros(z, np.int32), np.zeros(z))
            self.st = (np.zeros(n + 1, np.int32), np.zeros(z, np.int32), np.zeros(z))
            check(lib().dphd_factor_export(handle, ptr(self.permutation), *map(ptr, self.s), *map(ptr, self.st)))
            self.nnz, self.factorization_nnz, self.matrix_nnz = z, fz.value, mz.value
        finally:
            lib().dphd_factor_destroy(handle)

    def as_dict(self):
        return {
            "permutation": self.permutation,
            "s": self.s,
            "st": self.st,
            "factorization_nnz": self.factorization_nnz,
            "matrix_nnz": self.matrix_nnz,
        }


def incidence(vertices, indices):
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    elements, corners = indices.shape
    offsets = np.zeros(vertices + 1, np.int32)
    entries = np.zeros(elements * corners, np.int32)
    check(lib().dphd_build_incidence(vertices, elements, corners, ptr(indices), ptr(offsets), ptr(entries)))
    return offsets, entries


def tet_geometry(poses):
    poses = np.ascontiguousarray(poses, dtype=np.float32)
    e = len(poses)
    grads = np.zeros((e, 4, 3))
    volumes = np.zeros(e)
    check(lib().dphd_tet_geometry(e, ptr(poses), ptr(grads), ptr(volumes)))
    return grads, volumes


def hex_gradients(dx):
    out = np.zeros((8, 8, 3))
    check(lib().dphd_hex_gradients(float(dx), ptr(out)))
    return out


def assemble(vertices, indices, grads, shared, weights, free, inertia, fibre_dir=None, fibre_w=None):
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    grads = np.ascontiguousarray(grads, dtype=np.float64)
    weights = np.ascontiguousarray(weights, dtype=np.float64)
    free = np.ascontiguousarray(free, dtype=np.uint8)
    fibre_dir = None if fibre_dir is None else np.ascontiguousarray(fibre_dir, dtype=np.float64)
    fibre_w = None if fibre_w is None else np.ascontiguousarray(fibre_w, dtype=np.float64)
    elements, corners = indices.shape
    samples = 8 if shared else 1
    nnz = ctypes.c_int()
    args = [vertices, elements, corners, samples, ptr(indices), ptr(grads), int(shared), ptr(weights),
            ptr(fibre_dir), ptr(fibre_w), ptr(free), float(inertia), 0, None, None]
    check(lib().dphd_assemble(*args, ctypes.byref(nnz), None, None, None, 0))
    off = np.zeros(vertices + 1, np.int32)
    col = np.zeros(nnz.value, np.int32)
    val = np.zeros(nnz.value)
    check(lib().dphd_assemble(*args, ctypes.byref(nnz), ptr(off), ptr(col), ptr(val), nnz.value))
    return off, col, val
