"""libqes_mi.so loads and exports every entry point declared in
include/qes_mi.h (no compute calls: CPU only)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "qes_mi.h")).read()
    return sorted(set(re.findall(r"\b(qes_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = _declared()
    for n in ("qes_basis_create", "qes_matvec", "qes_lanczos_ground_state", "qes_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from pyqusolver_amd import _lib
    L = _lib.lib()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    assert L.qes_version().decode().startswith("qes_mi")


def test_binding_signatures_cover_header():
    from pyqusolver_amd import _lib
    missing = [n for n in _declared() if n not in _lib.SIGNATURES]
    assert not missing, missing


def test_error_path_without_gpu_work():
    from pyqusolver_amd import _lib
    L = _lib.lib()
    assert L.qes_basis_create(None, None, None) == -1
    assert "null" in L.qes_last_error().decode()


def test_every_entry_point_rejects_null_arguments():
    """Every entry point of the C ABI called with null handles / pointers and
    zero sizes returns an error status (destroy functions accept null) and
    leaves a message, without touching a device."""
    import ctypes as C
    from pyqusolver_amd import _lib
    L = _lib.lib()
    for name, argtypes in _lib.SIGNATURES.items():
        if name in ("qes_version", "qes_last_error"):
            continue
        args = []
        for t in argtypes:
            if t in (C.c_int32, C.c_int64):
                args.append(0)
            elif t is C.c_double:
                args.append(0.0)
            else:
                args.append(None)
        rc = getattr(L, name)(*args)
        if name.endswith("_destroy"):
            assert rc == 0, name
        else:
            assert rc == -1, name
            assert L.qes_last_error().decode(), name
