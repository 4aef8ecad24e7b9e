"""ctypes binding of libtatva_amd.so (the C ABI declared in include/tatva_amd.h).

The library is required: importing a compute path without it raises at once,
there is no Python or CPU fallback for any kernel.
"""

from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtatva_amd.so")

OK, EINVAL, EHIP, ENOMEM, EBREAKDOWN, ENOTCONV = 0, 1, 2, 3, 4, 5
TRI3, TET4, HEX8 = 0, 1, 2
POISSON, LINEAR_ELASTIC, NEO_HOOKEAN = 1, 2, 3

c_int, c_i32, c_i64, c_dbl, c_void = (ctypes.c_int, ctypes.c_int32, ctypes.c_int64,
                                       ctypes.c_double, ctypes.c_void_p)


class Density(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("points_per_element", ctypes.c_int32),
                ("p0", ctypes.c_double), ("p1", ctypes.c_double), ("p2", ctypes.c_double),
                ("elem_params", ctypes.c_void_p)]


class PcgStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("status", ctypes.c_int32),
                ("rel_residual", ctypes.c_double), ("b_norm", ctypes.c_double),
                ("op_ms", ctypes.c_double), ("op_samples", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("op_main_ms", ctypes.c_double)]


# name -> (restype, argtypes)
_SIGS = {
    "tatva_last_error": (ctypes.c_char_p, []),
    "tatva_version": (c_int, []),
    "tatva_device_count": (c_int, [ctypes.POINTER(c_int)]),
    "tatva_set_profiling": (c_int, [c_int]),
    "tatva_set_option": (c_int, [c_i32, c_i32]),
    "tatva_mesh_info": (c_int, [c_void, ctypes.POINTER(c_i64), c_int]),
    "tatva_mesh_create": (c_int, [c_int, c_int, c_void, c_i64, c_int, c_void, c_i64, c_int, c_int,
                                  c_void, c_void, ctypes.POINTER(c_void)]),
    "tatva_mesh_destroy": (c_int, [c_void]),
    "tatva_weights": (c_int, [c_void, c_void, c_void]),
    "tatva_eval": (c_int, [c_void, c_int, c_void, c_void, c_void]),
    "tatva_grad": (c_int, [c_void, c_int, c_void, c_void, c_void]),
    "tatva_integrate": (c_int, [c_void, c_int, c_int, c_void, c_void, c_void, c_void]),
    "tatva_density_eval": (c_int, [c_int, c_i64, ctypes.POINTER(Density), c_void, c_void, c_void,
                                   c_void]),
    "tatva_energy": (c_int, [c_void, ctypes.POINTER(Density), c_void, c_void, c_void, c_void]),
    "tatva_residual": (c_int, [c_void, ctypes.POINTER(Density), c_void, c_void, c_void, c_void,
                               c_void]),
    "tatva_hvp": (c_int, [c_void, ctypes.POINTER(Density), c_void, c_void, c_void, c_void, c_void]),
    "tatva_hvp_f32": (c_int, [c_void, ctypes.POINTER(Density), c_void, c_void, c_void, c_void,
                              c_void]),
    "tatva_hessian_diag": (c_int, [c_void, ctypes.POINTER(Density), c_void, c_void, c_void, c_void]),
    "tatva_hessian_diag_f32": (c_int, [c_void, ctypes.POINTER(Density), c_void, c_void, c_void,
                                       c_void]),
    "tatva_colored_hessian_workspace": (c_i64, [c_void, c_int]),
    "tatva_colored_hessian": (c_int, [c_void, ctypes.POINTER(Density), c_void, c_void, c_void, c_i64,
                                      c_void, c_i32, c_void, c_void, c_i64, c_void, c_void, c_i64,
                                      c_void]),
    "tatva_colored_hessian_mapped": (c_int, [c_void, ctypes.POINTER(Density), c_void, c_void, c_void,
                                             c_void, c_i64, c_void, c_i32, c_void, c_void, c_i64,
                                             c_i32, c_void, c_void, c_i64, c_void]),
    "tatva_mass_matrix": (c_int, [c_void, c_i32, c_void, c_void, c_i64, c_void, c_i32, c_void, c_void,
                                  c_i64, c_void, c_void, c_i64, c_void]),
    "tatva_project_rhs": (c_int, [c_void, c_i32, c_void, c_void, c_void]),
    "tatva_interpolate": (c_int, [c_void, c_i64, c_void, c_i32, c_void, c_void, ctypes.POINTER(c_i32),
                                  c_void]),
    "tatva_rows_cols": (c_int, [c_i64, c_void, c_void, c_void, c_void, c_void, c_void]),
    "tatva_spmv_csr": (c_int, [c_i64, c_i64, c_void, c_void, c_void, c_void, c_void, c_void]),
    "tatva_pcg_csr_workspace": (c_i64, [c_i64, c_void]),
    "tatva_pcg_csr": (c_int, [c_i64, c_i64, c_void, c_void, c_void, c_void, c_i32, c_void, c_void,
                              c_dbl, c_i32, c_i32, c_void, c_i64, ctypes.POINTER(PcgStats),
                              c_void]),
    "tatva_csr_block_view": (c_int, [c_i64, c_i64, c_void, c_void, c_i32, c_void,
                                     ctypes.POINTER(c_i32), c_void]),
    "tatva_spmv_bcsr": (c_int, [c_i64, c_void, c_void, c_i32, c_void, c_void, c_void, c_void]),
    "tatva_pcg_matfree_workspace": (c_i64, [c_void, ctypes.POINTER(Density), c_i64]),
    "tatva_pcg_matfree_f32": (c_int, [c_void, ctypes.POINTER(Density), c_void, c_void, c_i64, c_void,
                                      c_void, c_void, c_dbl, c_i32, c_i32, c_void, c_i64,
                                      ctypes.POINTER(PcgStats), c_void]),
    "tatva_pcg_matfree_mapped_workspace": (c_i64, [c_void, c_i64]),
    "tatva_pcg_matfree_f32_mapped": (c_int, [c_void, ctypes.POINTER(Density), c_void, c_void,
                                             c_void, c_void, c_i64, c_void, c_void, c_void, c_dbl,
                                             c_i32, c_i32, c_void, c_i64,
                                             ctypes.POINTER(PcgStats), c_void]),
    "tatva_batch_create": (c_int, [c_int, c_i32, ctypes.POINTER(c_void)]),
    "tatva_batch_destroy": (c_int, [c_void]),
    "tatva_batch_slot_workspace": (c_i64, [c_void, c_i32, c_i64]),
    "tatva_batch_start": (c_int, [c_void, c_i32, c_void, ctypes.POINTER(Density), c_void, c_void,
                                  c_i64, c_void, c_void, c_void, c_dbl, c_i32, c_void, c_i64,
                                  c_void]),
    "tatva_batch_release": (c_int, [c_void, c_i32]),
    "tatva_batch_iterate": (c_int, [c_void, c_i32, c_void]),
    "tatva_batch_poll": (c_int, [c_void, c_void, ctypes.POINTER(c_i32), c_void]),
    "tatva_scatter_free": (c_int, [c_i64, c_void, c_void, c_void, c_void]),
    "tatva_set_dofs": (c_int, [c_i64, c_void, c_void, c_dbl, c_void, c_void]),
    "tatva_gather_free": (c_int, [c_i64, c_void, c_void, c_void, c_void]),
    "tatva_lift_map": (c_int, [c_i64, c_void, c_void, c_void, c_void]),
    "tatva_reduce_rows": (c_int, [c_i64, c_void, c_void, c_void, c_void, c_void]),
    "tatva_axpy_free_f32": (c_int, [c_i64, c_void, c_dbl, c_void, c_void, c_void]),
    "tatva_axpy_free": (c_int, [c_i64, c_void, c_dbl, c_void, c_void, c_void]),
    "tatva_cast_f64_f32": (c_int, [c_i64, c_dbl, c_void, c_void, c_void]),
    "tatva_cast_f32_f64": (c_int, [c_i64, c_dbl, c_void, c_void, c_void]),
    "tatva_dot_workspace": (c_i64, [c_i64]),
    "tatva_dot_noalloc": (c_int, [c_i64, c_void, c_void, c_void, c_void]),
    "tatva_valu_calibration": (c_int, [c_i32, c_i32, c_i32, c_void, c_void]),
    "tatva_mass_matrix_mapped": (c_int, [c_void, c_i32, c_void, c_void, c_void, c_i64, c_void,
                                         c_i32, c_void, c_void, c_i64, c_i32, c_void, c_void,
                                         c_i64, c_void]),
    "tatva_bicgstab_csr_workspace": (c_i64, [c_i64]),
    "tatva_bicgstab_csr": (c_int, [c_i64, c_i64, c_void, c_void, c_void, c_void, c_void, c_dbl,
                                   c_i32, c_void, c_i64, ctypes.POINTER(PcgStats), c_void]),
    "tatva_check_distance2_colors": (c_int, [c_i64, c_void, c_void, c_void,
                                             ctypes.POINTER(c_i64)]),
    "tatva_dot": (c_int, [c_i64, c_void, c_void, c_void, c_void, c_i64, c_void]),
    "tatva_pattern_from_mesh": (c_int, [c_void, c_i64, c_int, c_i64, c_int, c_void, c_i64,
                                        ctypes.POINTER(c_void)]),
    "tatva_pattern_from_element_dofs": (c_int, [c_void, c_i64, c_int, c_i64, c_void, c_i64,
                                                ctypes.POINTER(c_void)]),
    "tatva_csr_submatrix": (c_int, [c_i64, c_void, c_void, c_void, c_i64, ctypes.POINTER(c_void)]),
    "tatva_csr_host_rows": (c_i64, [c_void]),
    "tatva_csr_host_nnz": (c_i64, [c_void]),
    "tatva_csr_host_copy": (c_int, [c_void, c_void, c_void]),
    "tatva_csr_host_free": (None, [c_void]),
    "tatva_distance2_colors": (c_int, [c_i64, c_void, c_void, c_void, ctypes.POINTER(c_i32)]),
}

EXPORTED = tuple(_SIGS)

_lib = None
_lock = threading.Lock()


class TatvaError(RuntimeError):
    """Error reported by the native library."""


class SolverBreakdown(TatvaError):
    """CG breakdown: p.Ap <= 0 or a non-finite value."""


class NotConverged(TatvaError):
    """Iteration limit reached before the tolerance."""


def lib():
    """Load the native library (raises if it has not been built).

    torch is imported first: its wheel carries its own HIP and HSA runtimes
    under the same sonames as ROCm's, so the library then binds to the copy
    torch already mapped.  Loaded the other way round, the process maps both
    copies, and the second HSA runtime sees no GPU (hipSetDevice reports no
    ROCm-capable device); device pointers from torch would also belong to a
    runtime the library does not use."""
    global _lib
    if _lib is not None:
        return _lib
    torch()
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"tatva_amd native library not found at {LIB_PATH}; build it with "
                "`python -m tatva_amd._build` (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    """Raise the Python exception matching a native status code."""
    if rc == OK:
        return
    msg = lib().tatva_last_error().decode(errors="replace")
    if rc == EINVAL:
        raise ValueError(msg)
    if rc == ENOMEM:
        raise MemoryError(msg)
    if rc == EBREAKDOWN:
        raise SolverBreakdown(msg)
    if rc == ENOTCONV:
        raise NotConverged(msg)
    raise TatvaError(msg or f"native status {rc}")


def host_ptr(a: np.ndarray) -> int:
    return a.ctypes.data


# ----------------------------------------------------------------- torch ----

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t  # noqa: PLC0415
        _torch = t
    return _torch


def device():
    """The GPU all device arrays live on (raises without a ROCm GPU)."""
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("tatva_amd compute calls need a ROCm GPU (torch.cuda is unavailable)")
    return t.device("cuda", t.cuda.current_device())


def stream() -> int:
    return torch().cuda.current_stream().cuda_stream


def dptr(t) -> int:
    """Device pointer of a contiguous CUDA tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a GPU tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return t.data_ptr()


def to_device(x, dtype=None):
    """numpy / torch / scalar -> contiguous tensor on the GPU."""
    t = torch()
    dt = dtype or t.float64
    if isinstance(x, t.Tensor):
        out = x.to(device=device(), dtype=dt)
    else:
        out = t.as_tensor(np.asarray(x), device=device(), dtype=dt)
    return out.contiguous()
