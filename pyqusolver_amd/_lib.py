"""ctypes binding of libqes_mi.so (C ABI in include/qes_mi.h).

torch is imported first so that the library binds to the HIP runtime torch
already loaded (one runtime per process).  A missing library is an error:
there is no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import torch  # noqa: F401  (loads the HIP runtime before the library)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqes_mi.so")

_lib = None


class QesError(RuntimeError):
    pass


class GroupSpec(C.Structure):
    _fields_ = [
        ("ns", C.c_int32),
        ("n_up", C.c_int32),
        ("z2", C.c_int32),
        ("nbase", C.c_int32),
        ("base_ops", (C.c_uint8 * 4) * 4),
        ("ntx", C.c_int32),
        ("nty", C.c_int32),
        ("row_bits", C.c_int32),
        ("cell_bits", C.c_int32),
        ("chi_re", C.POINTER(C.c_double)),
        ("chi_im", C.POINTER(C.c_double)),
        ("cross_masks", C.POINTER(C.c_uint64)),
        ("bphase", C.POINTER(C.c_double)),
        ("perm", C.POINTER(C.c_int32)),
    ]


_P = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
_PI32 = C.POINTER(C.c_int32)
_PI64 = C.POINTER(C.c_int64)
_PD = C.POINTER(C.c_double)
_PU64 = C.POINTER(C.c_uint64)

SIGNATURES = {
    "qes_basis_create": [C.POINTER(GroupSpec), C.POINTER(_P), _PI64],
    "qes_basis_get": [_P, _PU64, _PD],
    "qes_basis_device_ptrs": [_P, C.POINTER(_P), C.POINTER(_P)],
    "qes_basis_map_states": [_P, _PU64, _I64, _PI64, _PI32],
    "qes_basis_destroy": [_P],
    "qes_basis_expand": [_P, _P, _P, _I32, _P],
    "qes_basis_reduce": [_P, _P, _P, _I32, _P],
    "qes_op_create": [_P, _PI32, _PI32, _PD, _PD, _I32, _I32, _PI32, C.POINTER(_P)],
    "qes_op_destroy": [_P],
    "qes_op_needs_complex": [_P, _PI32],
    "qes_matvec": [_P, _P, _P, _I64, _I32, _I32, _P],
    "qes_matvec_ws": [_P, _P, _P, _I64, _I32, _I32, _P, _I64, _P],
    "qes_matvec_couplings": [_P, _P, _P, _I32, _P],
    "qes_projected_matvec": [_P, _P, _P, _P, _P],
    "qes_op_n_couplings": [_P, _PI32],
    "qes_op_count_offdiag": [_P, _PI64],
    "qes_op_count_outside": [_P, _PI64],
    "qes_op_assemble": [_P, _I32, C.c_double, _P],
    "qes_op_assembly_info": [_P, _PI64, _PI64, _PI32, _PI32, _PI32],
    "qes_op_panel_info": [_P, _PI32, _PI64, _PI64, _PI64],
    "qes_op_table_coo": [_P, _P, _P, _P, _I64, _PI64, _P],
    "qes_tridiag_ritz": [_PD, _PD, _I32, _PD, _PD, _PD],
    "qes_batch_create_scaled": [C.POINTER(_P), _I32, _I32, _PD, _I32, _I32, C.POINTER(_P)],
    "qes_op_matvec_scaled": [_P, _PD, _I32, _P, _P, _I32, _P],
    "qes_op_coupling_dots": [_P, _P, _P, _I32, _PD, _P],
    "qes_op_coupling_dots_noalloc": [_P, _P, _P, _I32, _PD, _P],
    "qes_op_diagonal": [_P, _P, _P, _P],
    "qes_op_set_canonicalizer": [_P, _I32],
    "qes_op_set_option": [_P, C.c_char_p, _I64],
    "qes_op_get_plan": [_P, _PI32, _PI32, _PI32],
    "qes_basis_canon_check": [_P, _I32, _PU64, _I64, _PI64],
    "qes_basis_set_lookup": [_P, _I32],
    "qes_basis_set_bucket": [_P, _I32],
    "qes_lanczos_create": [_P, _I32, _I32, _I32, C.POINTER(_P)],
    "qes_lanczos_start": [_P, _P, _P],
    "qes_lanczos_run": [_P, _I32, _I32, _I32],
    "qes_lanczos_get": [_P, _PD, _PD, _PI32],
    "qes_lanczos_ritz": [_P, _PD, _I32, _P, _P],
    "qes_lanczos_accumulate": [_P, _PD, _I32, _P, _P],
    "qes_lanczos_destroy": [_P],
    "qes_lanczos_stats": [_P, _PI32],
    "qes_lanczos_ground_state": [_P, _P, _I32, _I32, _I32, C.c_double, _I32, _PD, _PD, _PI32, _PD, _P],
    "qes_batch_create": [C.POINTER(_P), _I32, _I32, _I32, C.POINTER(_P)],
    "qes_batch_create_blocks": [C.POINTER(_P), _I32, C.POINTER(_P), _PI32, C.POINTER(_P), _I32, _I32, _I32,
                                C.POINTER(_P)],
    "qes_batch_offsets": [_P, _PI64],
    "qes_batch_start": [_P, _P, _P],
    "qes_batch_run": [_P, _I32, _I32],
    "qes_batch_get": [_P, _PD, _PD, _PI32],
    "qes_batch_set_active": [_P, _PI32],
    "qes_batch_matvec": [_P, _P, _P, _P],
    "qes_batch_set_option": [_P, C.c_char_p, _I64],
    "qes_batch_ground_states": [_P, C.c_double, _I32, _PD, _PI32],
    "qes_batch_destroy": [_P],
    "qes_block_dot_ws_bytes": [_I32, _I64, _PI64],
    "qes_block_dot": [_P, _I64, _I32, _P, _I64, _I32, _I64, _I32, _P, _P, _I64, _P],
    "qes_block_combine": [_P, _I64, _I32, _P, C.c_double, _P, _I64, _I32, _I64, _I32, _P],
    "qes_last_error": [],
    "qes_version": [],
}


def lib():
    """Load the library (once).  Raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m pyqusolver_amd.build_ext` "
            "(hipcc, gfx950). There is no CPU fallback.")
    L = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = C.c_char_p if name in ("qes_last_error", "qes_version") else C.c_int
    _lib = L
    return L


def exported_symbols():
    """Names the library exports among those declared in include/qes_mi.h."""
    L = lib()
    return [n for n in SIGNATURES if getattr(L, n, None) is not None]


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().qes_last_error().decode(errors="replace")
        raise QesError(f"{what}: {msg}" if what else msg)


def require_gpu() -> None:
    if not torch.cuda.is_available():
        raise QesError("pyqusolver_amd needs a ROCm GPU (gfx950); torch.cuda.is_available() is False")


def dptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def stream_ptr(stream=None) -> C.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)
