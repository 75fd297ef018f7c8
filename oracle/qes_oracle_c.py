"""ctypes wrapper of oracle/qes_oracle.c (test/baseline infrastructure only)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from . import qes_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libqes_oracle.so")


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


class _OGroup(C.Structure):
    _fields_ = [("ns", C.c_int), ("n_group", C.c_int), ("max_ops", C.c_int),
                ("n_ops", C.c_void_p), ("op_code", C.c_void_p), ("arg0", C.c_void_p), ("arg1", C.c_void_p),
                ("trans_perm", C.c_void_p), ("refl_perm", C.c_void_p), ("parity_axis", C.c_void_p),
                ("chi_re", C.c_void_p), ("chi_im", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.oracle_push_matvec.restype = C.c_int64
        L.oracle_push_matvec.argtypes = [C.POINTER(_OGroup), C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                         C.c_void_p, C.c_int, C.c_int64, C.c_int64]
        L.oracle_scan_window.restype = C.c_int64
        L.oracle_scan_window.argtypes = [C.POINTER(_OGroup), C.c_int, C.c_uint64, C.c_int64, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


class CGroup:
    """Flattened group tables of an oracle Group for the C restatement."""

    def __init__(self, group: O.Group):
        G = group.n_group
        mo = max(1, max((len(e) for e in group.elements), default=1))
        self.n_ops = np.zeros(G, np.int32)
        self.op_code = np.zeros((G, mo), np.int32)
        self.arg0 = np.zeros((G, mo), np.int64)
        self.arg1 = np.zeros((G, mo), np.int64)
        for g, e in enumerate(group.elements):
            self.n_ops[g] = len(e)
            for j, (code, a0, a1) in enumerate(e):
                self.op_code[g, j], self.arg0[g, j], self.arg1[g, j] = code, a0, a1
        ns = group.lat.ns
        self.tp = np.ascontiguousarray(np.array(group.trans_perm, np.int64).reshape(-1, ns) if group.trans_perm
                                       else np.zeros((1, ns), np.int64))
        self.rp = np.ascontiguousarray(np.array(group.refl_perm, np.int64).reshape(-1, ns) if group.refl_perm
                                       else np.zeros((1, ns), np.int64))
        self.pa = np.ascontiguousarray(np.array(group.parity_axis or [0], np.int32))
        self.cr = np.ascontiguousarray(group.chi.real)
        self.ci = np.ascontiguousarray(group.chi.imag)
        self.s = _OGroup(ns, G, mo, self.n_ops.ctypes.data, self.op_code.ctypes.data, self.arg0.ctypes.data,
                         self.arg1.ctypes.data, self.tp.ctypes.data, self.rp.ctypes.data, self.pa.ctypes.data,
                         self.cr.ctypes.data, self.ci.ctypes.data)


def push_matvec(group: O.Group, reps, norms, codes, sites, coeffs, is_complex, x, k_range=None, cgroup=None,
                y=None):
    """Accumulate the push-form matvec over columns k_range (default all)
    into y (zeros of x's shape and dtype when None).  Returns (y, n_gathers)."""
    cg = cgroup or CGroup(group)
    reps = np.ascontiguousarray(reps, dtype=np.int64)
    norms = np.ascontiguousarray(norms, dtype=np.float64)
    n = len(reps)
    codes = np.ascontiguousarray(codes, np.int32)
    sites = np.ascontiguousarray(np.asarray(sites, np.int32).reshape(-1, 2))
    co = np.asarray(coeffs, dtype=np.complex128)
    cre, cim = np.ascontiguousarray(co.real), np.ascontiguousarray(co.imag)
    xc = np.iscomplexobj(x)
    xa = np.ascontiguousarray(x, dtype=np.complex128 if xc else np.float64)
    if y is None:
        y = np.zeros_like(xa)
    elif y.shape != xa.shape or y.dtype != xa.dtype or not y.flags.c_contiguous:
        raise ValueError("y must be a C-contiguous array of x's shape and dtype")
    k0, k1 = (0, n) if k_range is None else k_range
    g = lib().oracle_push_matvec(C.byref(cg.s), reps.ctypes.data, norms.ctypes.data, n, codes.ctypes.data,
                                 sites.ctypes.data, cre.ctypes.data, cim.ctypes.data, len(codes),
                                 1 if is_complex else 0, xa.ctypes.data, y.ctypes.data, 1 if xc else 0, k0, k1)
    return y, int(g)


def scan_window(group: O.Group, rank0: int, count: int, cgroup=None, threads: int = 1):
    """Representatives and norms among the states of the group's filling
    with colex rank in [rank0, rank0 + count) (oracle_scan_window), split
    over ``threads`` host threads (ctypes releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    cg = cgroup or CGroup(group)
    n_up = int(group.n_up)
    parts = max(1, int(threads))
    bounds = [rank0 + (count * i) // parts for i in range(parts + 1)]

    def one(i):
        r0, c = bounds[i], bounds[i + 1] - bounds[i]
        reps = np.zeros(max(c, 1), np.int64)
        norms = np.zeros(max(c, 1), np.float64)
        m = lib().oracle_scan_window(C.byref(cg.s), n_up, r0, c, reps.ctypes.data, norms.ctypes.data)
        return reps[:m], norms[:m]
    with ThreadPoolExecutor(max_workers=parts) as ex:
        out = list(ex.map(one, range(parts)))
    return np.concatenate([o[0] for o in out]), np.concatenate([o[1] for o in out])
