"""Spin-1/2 Hamiltonians on a symmetry-reduced sector, matvec on the GPU.

Host-side mirror of the reference path
  Hamiltonian.add -> SpecialOperator instruction arrays
      (Algebra/Operator/special_operator.py:1113-1231, :1392-1417)
  SpecialOperator.matvec_fun (special_operator.py:593-709)
  Operator.matvec(vecs, hilbert_in=...) (Algebra/Operator/operator.py:1559-1602)
  GeneralMatrix as scipy LinearOperator (Algebra/Operator/matrix.py:269-408)
  Hamiltonian.diagonalize(method='lanczos') (Algebra/hamil.py:1809-2114)
with every matvec and Lanczos step executed by libqes_mi on the device.

Differences from the reference, on purpose:
  * the LinearOperator has shape (nh, nh) and the Hilbert context of the
    Hamiltonian (the reference leaves (0, 0) and passes hilbert=None,
    matrix.py:334,381 and operator.py:495-503);
  * a float64 vector with complex matrix elements (complex couplings or a
    complex momentum sector) raises instead of silently dropping Im.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib

# SpinLookupCodes (Algebra/Operator/impl/operators_spin.py:1084-1118)
SIG_X_L, SIG_Y_L, SIG_Z_L, SIG_P_L, SIG_M_L = 11, 12, 13, 14, 15
SIG_X_C, SIG_Y_C, SIG_Z_C, SIG_P_C, SIG_M_C = 20, 21, 22, 23, 24
SIG_XY_C, SIG_YX_C, SIG_XZ_C, SIG_ZX_C, SIG_YZ_C, SIG_ZY_C = 25, 26, 27, 28, 29, 30
SIG_PM_C, SIG_MP_C = 31, 32

_LOCAL = {"x": SIG_X_L, "y": SIG_Y_L, "z": SIG_Z_L, "p": SIG_P_L, "m": SIG_M_L}
_CORR = {"x": SIG_X_C, "y": SIG_Y_C, "z": SIG_Z_C, "p": SIG_P_C, "m": SIG_M_C, "xy": SIG_XY_C,
         "yx": SIG_YX_C, "xz": SIG_XZ_C, "zx": SIG_ZX_C, "yz": SIG_YZ_C, "zy": SIG_ZY_C,
         "pm": SIG_PM_C, "mp": SIG_MP_C}


@dataclass(frozen=True)
class SpinOperator:
    """Lightweight stand-in for the reference's spin Operator objects: only
    the instruction code matters on this path."""
    name: str
    code: int
    type_acting: str       # "local" | "correlation"
    modifies: bool


def spin_op(kind: str, type_act: str = "local") -> SpinOperator:
    t = str(getattr(type_act, "name", type_act)).lower()
    if t.startswith("loc"):
        code = _LOCAL[kind]
        t = "local"
    elif t.startswith("cor"):
        code = _CORR[kind]
        t = "correlation"
    else:
        raise ValueError("type_act must be 'local' or 'correlation'")
    return SpinOperator(f"S{kind}/{t[0].upper()}", code, t, kind not in ("z",))


class _SpinOps:
    """ham.operators namespace (sig_x(ns=..., type_act=...), ...)."""

    def __getattr__(self, name):
        if not name.startswith("sig_"):
            raise AttributeError(name)
        kind = name[4:]

        def make(ns=None, type_act="local", sites=None, **kw):
            return spin_op(kind, type_act)
        return make


class Hamiltonian:
    """Many-body spin-1/2 Hamiltonian on a HilbertSpace sector."""

    def __init__(self, hilbert_space, dtype=np.float64, name: str = "Hamiltonian", lattice=None, **kwargs):
        self._hilbert_space = hilbert_space
        self._lattice = lattice if lattice is not None else getattr(hilbert_space, "lattice", None)
        self._dtype = np.dtype(dtype)
        self._name = name
        self.ns = hilbert_space.ns
        self._instr_codes: List[int] = []
        self._instr_coeffs: List[complex] = []
        self._instr_sites: List[List[int]] = []
        self._instr_coupling: List[int] = []
        self._op_handle = None
        self._cached_matvec_fun = None
        self._eig_val = None
        self._eig_vec = None
        self.operators = _SpinOps()
        self.last_lanczos = None

    # ---- term list ----------------------------------------------------
    def _dtype_is_complex(self) -> bool:
        return np.issubdtype(self._dtype, np.complexfloating)

    @property
    def dtype(self):
        return self._dtype

    @property
    def hilbert_space(self):
        return self._hilbert_space

    @property
    def nh(self) -> int:
        return self._hilbert_space.nh

    def add(self, operator: SpinOperator, multiplier, modifies: bool = False, sites=None, coupling: int = 0):
        """special_operator.py:1113-1231: append (code, coeff, sites)."""
        if abs(multiplier) < 1e-15:
            return
        s = list(sites) if sites else [0]
        if len(s) > 2:
            raise ValueError("spin-1/2 instructions act on at most two sites")
        s = s + [-1] * (2 - len(s))
        self._instr_codes.append(int(operator.code))
        self._instr_coeffs.append(multiplier)
        self._instr_sites.append(s)
        self._instr_coupling.append(int(coupling))
        self._invalidate()

    def _invalidate(self):
        if self._op_handle is not None:
            try:
                _lib.lib().qes_op_destroy(self._op_handle)
            except Exception:
                pass
        self._op_handle = None
        self._cached_matvec_fun = None

    def instruction_arrays(self):
        """(sites int32[T,2], coeffs, codes int32[T]) like
        _finalize_instruction_arrays (special_operator.py:1392-1417)."""
        n = len(self._instr_codes)
        sites = np.asarray(self._instr_sites, dtype=np.int32).reshape(n, 2)
        coeffs = np.asarray(self._instr_coeffs, dtype=np.complex128 if self._dtype_is_complex() else np.float64)
        codes = np.asarray(self._instr_codes, dtype=np.int32)
        return sites, coeffs, codes

    @property
    def n_coupling_ids(self) -> int:
        """Max coupling id + 1 over the instructions (= qes_op_n_couplings)."""
        return max(self._instr_coupling, default=0) + 1

    def set_coupling_scale(self, scales: Sequence[float]):
        """Rescale the coefficients of each coupling id (used by autograd)."""
        self._coupling_scale = np.asarray(scales, dtype=np.float64)
        self._invalidate()

    # ---- device operator ----------------------------------------------
    def _device_op(self):
        if self._op_handle is not None:
            return self._op_handle
        err = getattr(self._hilbert_space.group, "matvec_error", None)
        if err:
            raise _lib.QesError(err)
        sites, coeffs, codes = self.instruction_arrays()
        cpl = np.asarray(self._instr_coupling, dtype=np.int32)
        coeffs = coeffs.astype(np.complex128)
        scale = getattr(self, "_coupling_scale", None)
        if scale is not None and len(cpl):
            coeffs = coeffs * scale[cpl]
        cre = np.ascontiguousarray(coeffs.real)
        cim = np.ascontiguousarray(coeffs.imag)
        h = C.c_void_p()
        T = len(codes)
        _lib.check(_lib.lib().qes_op_create(
            self._hilbert_space.handle,
            codes.ctypes.data_as(_lib._PI32) if T else None,
            np.ascontiguousarray(sites).ctypes.data_as(_lib._PI32) if T else None,
            cre.ctypes.data_as(_lib._PD) if T else None,
            cim.ctypes.data_as(_lib._PD) if T else None,
            T, 1 if self._dtype_is_complex() else 0,
            cpl.ctypes.data_as(_lib._PI32) if T else None,
            C.byref(h)), "qes_op_create")
        self._op_handle = h
        return h

    @property
    def needs_complex(self) -> bool:
        v = C.c_int32()
        _lib.check(_lib.lib().qes_op_needs_complex(self._device_op(), C.byref(v)))
        return bool(v.value)

    @property
    def vector_dtype(self):
        return np.complex128 if self.needs_complex else np.float64

    @property
    def n_couplings(self) -> int:
        v = C.c_int32()
        _lib.check(_lib.lib().qes_op_n_couplings(self._device_op(), C.byref(v)))
        return int(v.value)

    CANON_KINDS = {"auto": -1, "brute": 0, "cyclic": 1, "rows": 2}

    def set_canonicalizer(self, kind: str = "auto"):
        """Select the canonicalisation kernel (tuning / cross-checks)."""
        _lib.check(_lib.lib().qes_op_set_canonicalizer(self._device_op(), self.CANON_KINDS[kind]),
                   "qes_op_set_canonicalizer")

    def set_option(self, name: str, value: int):
        """Kernel tuning knob (bond_order_uniform, direct_lookup, xcd_remap, top_row_test, direct_batch,
        asm_panel_cols: columns per panel of the next assembled table, -1 automatic, 0 off; rank_gather:
        1 builds the basis's rank-indexed gather vector for the matrix-free product, 0 frees it; lean:
        1 (default) runs the kernels compiled for groups of translations and one spin flip where they
        apply, 0 the general ones -- the same products either way; sell_grid, sell_lds: launch shape
        of the table product)."""
        _lib.check(_lib.lib().qes_op_set_option(self._device_op(), name.encode(), int(value)), "qes_op_set_option")

    @property
    def plan(self) -> dict:
        c, nc, ng = C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(_lib.lib().qes_op_get_plan(self._device_op(), C.byref(c), C.byref(nc), C.byref(ng)))
        names = {v: k for k, v in self.CANON_KINDS.items()}
        return {"canonicalizer": names[c.value], "exchange_classes": nc.value, "generic_groups": ng.value}

    def assemble(self, enable: bool = True, max_gb: float = 0.0, dictionary: bool = True, couplings: bool = False):
        """Tabulate H on the device (sliced ELLPACK, built once by the same
        row visitor) so that single-vector matvecs stream the table instead
        of canonicalising; ``enable=False`` frees it.  ``dictionary`` lets
        the table store value indices instead of values.  ``couplings``
        keys the dictionary by coupling id as well (one diagonal per
        coupling), which ``matvec_scaled_device`` and
        ``coupling_dots_device`` (the differentiable H(J)) need; it raises
        when the values do not fit a dictionary.  Returns the table size in
        bytes."""
        mode = 0 if not enable else (3 if couplings else (1 if dictionary else 2))
        _lib.check(_lib.lib().qes_op_assemble(self._device_op(), mode, float(max_gb), _lib.stream_ptr()),
                   "qes_op_assemble")
        return self.assembly_bytes

    @property
    def assembly(self) -> dict:
        e, b, f, nv, tg = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(_lib.lib().qes_op_assembly_info(self._device_op(), C.byref(e), C.byref(b), C.byref(f),
                                                   C.byref(nv), C.byref(tg)))
        names = {-1: None, 0: "plain", 1: "packed32", 2: "index16"}
        np_, pc, pe, pb = C.c_int32(), C.c_int64(), C.c_int64(), C.c_int64()
        _lib.check(_lib.lib().qes_op_panel_info(self._device_op(), C.byref(np_), C.byref(pc), C.byref(pe),
                                                C.byref(pb)))
        return {"entries": int(e.value), "bytes": int(b.value), "format": names[int(f.value)],
                "n_values": int(nv.value), "couplings": bool(tg.value),
                "panels": int(np_.value), "panel_cols": int(pc.value), "panel_entries": int(pe.value),
                "panel_bytes": int(pb.value)}

    @property
    def assembly_bytes(self) -> int:
        return self.assembly["bytes"]

    def build(self, use_numpy: bool = False, max_gb: float = 0.0, **kwargs):
        """Explicit sparse matrix of the reduced H (Hamiltonian.build,
        Algebra/hamil.py:1434-1557): the assembled table is built on the
        device and converted to scipy CSR (duplicate entries summed, padding
        dropped).  ``use_numpy`` returns a dense array instead."""
        import scipy.sparse as sp
        built = False
        if self.assembly_bytes == 0:
            self.assemble(True, max_gb=max_gb)
            built = True
        try:
            L = _lib.lib()
            cnt = C.c_int64()
            _lib.check(L.qes_op_table_coo(self._device_op(), None, None, None, 0, C.byref(cnt), None),
                       "qes_op_table_coo")
            m = int(cnt.value)
            cplx = self.needs_complex
            rows = torch.full((m,), -1, dtype=torch.int32, device="cuda")
            cols = torch.zeros((m,), dtype=torch.int32, device="cuda")
            vals = torch.zeros((m, 2) if cplx else (m,), dtype=torch.float64, device="cuda")
            _lib.check(L.qes_op_table_coo(self._device_op(), _lib.dptr(rows), _lib.dptr(cols), _lib.dptr(vals), m,
                                          C.byref(cnt), _lib.stream_ptr()), "qes_op_table_coo")
            v = torch.view_as_complex(vals) if cplx else vals
            keep = (rows >= 0) & (v != 0)
            r, c_, v = rows[keep].cpu().numpy(), cols[keep].cpu().numpy(), v[keep].cpu().numpy()
        finally:
            if built:
                self.assemble(False)
        n = self.nh
        A = sp.csr_matrix((v, (r, c_)), shape=(n, n))
        A.sum_duplicates()
        self._hamil_sparse = A
        return A.toarray() if use_numpy else A

    def count_outside(self) -> int:
        """Nonzero off-diagonal elements whose target lies outside the sector
        (skipped by the product, as the reference's kernel does): 0 for an
        operator that conserves the sector's quantum numbers."""
        v = C.c_int64()
        _lib.check(_lib.lib().qes_op_count_outside(self._device_op(), C.byref(v)), "qes_op_count_outside")
        return int(v.value)

    def count_offdiag(self) -> int:
        v = C.c_int64()
        _lib.check(_lib.lib().qes_op_count_offdiag(self._device_op(), C.byref(v)), "qes_op_count_offdiag")
        return int(v.value)

    # ---- matvec ---------------------------------------------------------
    def matvec_device(self, x: torch.Tensor, out: Optional[torch.Tensor] = None, accumulate: bool = False,
                      stream=None) -> torch.Tensor:
        """y = H x for device tensors of shape (nh,) or (nh, nb)."""
        n = self.nh
        if x.device.type != "cuda":
            raise ValueError("matvec_device expects a device tensor")
        if x.shape[0] != n or x.dim() > 2:
            raise ValueError(f"vector shape {tuple(x.shape)} does not match nh={n}")
        if x.dtype not in (torch.float64, torch.complex128):
            raise TypeError("vectors must be float64 or complex128")
        x = x.contiguous()
        nb = 1 if x.dim() == 1 else x.shape[1]
        y = out if out is not None else torch.empty_like(x)
        if y.shape != x.shape or y.dtype != x.dtype or not y.is_contiguous():
            raise ValueError("out must match x in shape/dtype and be contiguous")
        nbytes = x.numel() * x.element_size()
        if x.numel() and (x.data_ptr() < y.data_ptr() + nbytes and y.data_ptr() < x.data_ptr() + nbytes):
            raise ValueError("out must not overlap x: the product reads x at other rows while it writes out")
        dtype = 1 if x.dtype == torch.complex128 else 0
        if nb == 1 and self.assembly_bytes == 0:
            # matrix-free: gather from N .* x in a workspace (one random line
            # fewer per neighbour that is a representative)
            ws = torch.empty_like(x)
            if stream is not None:
                ws.record_stream(stream)      # freed only after the product on that stream
            _lib.check(_lib.lib().qes_matvec_ws(self._device_op(), _lib.dptr(x), _lib.dptr(y), 1, dtype,
                                                1 if accumulate else 0, _lib.dptr(ws), ws.numel() * ws.element_size(),
                                                _lib.stream_ptr(stream)), "qes_matvec_ws")
            return y
        _lib.check(_lib.lib().qes_matvec(self._device_op(), _lib.dptr(x), _lib.dptr(y), nb, dtype,
                                         1 if accumulate else 0, _lib.stream_ptr(stream)), "qes_matvec")
        return y

    def matvec_couplings_device(self, x: torch.Tensor) -> torch.Tensor:
        """Stack of H_c x for every coupling id c: shape (n_couplings, nh)."""
        ncpl = self.n_couplings
        y = torch.empty((ncpl,) + tuple(x.shape), dtype=x.dtype, device=x.device)
        dtype = 1 if x.dtype == torch.complex128 else 0
        _lib.check(_lib.lib().qes_matvec_couplings(self._device_op(), _lib.dptr(x.contiguous()), _lib.dptr(y), dtype,
                                                   _lib.stream_ptr()), "qes_matvec_couplings")
        return y

    def matvec_scaled_device(self, scale, x: torch.Tensor) -> torch.Tensor:
        """sum_c scale[c] H_c x in one pass over a coupling-tagged table
        (``assemble(couplings=True)``)."""
        sc = np.ascontiguousarray(np.asarray(scale, dtype=np.float64).reshape(-1))
        x = x.contiguous()
        y = torch.empty_like(x)
        dtype = 1 if x.dtype == torch.complex128 else 0
        _lib.check(_lib.lib().qes_op_matvec_scaled(self._device_op(), sc.ctypes.data_as(_lib._PD), len(sc),
                                                   _lib.dptr(x), _lib.dptr(y), dtype, _lib.stream_ptr()),
                   "qes_op_matvec_scaled")
        return y

    def coupling_dots_device(self, v: torch.Tensor, g: torch.Tensor) -> np.ndarray:
        """<g, H_c v> = sum_a conj(g_a) (H_c v)_a for every coupling id c
        (complex numpy array), one pass over a coupling-tagged table."""
        ncpl = self.n_couplings
        out = np.zeros(2 * ncpl, dtype=np.float64)
        v, g = v.contiguous(), g.contiguous()
        dtype = 1 if v.dtype == torch.complex128 else 0
        _lib.check(_lib.lib().qes_op_coupling_dots(self._device_op(), _lib.dptr(v), _lib.dptr(g), dtype,
                                                   out.ctypes.data_as(_lib._PD), _lib.stream_ptr()),
                   "qes_op_coupling_dots")
        return out[0::2] + 1j * out[1::2]

    def _matvec_any(self, x):
        if isinstance(x, torch.Tensor):
            if x.device.type == "cuda":
                return self.matvec_device(x)
            return self.matvec_device(x.to("cuda")).cpu()
        xa = np.asarray(x)
        if xa.dtype not in (np.float64, np.complex128):
            xa = xa.astype(np.complex128 if np.iscomplexobj(xa) else np.float64)
        return self._matvec_host(np.ascontiguousarray(xa))

    def _matvec_host(self, xa: np.ndarray) -> np.ndarray:
        """H x for a host array: x is copied (multi-threaded) into a cached
        page-locked staging buffer and sent to a cached device vector; the
        result comes back over PCIe straight into a freshly allocated
        page-locked array, returned as the caller's numpy array (no second
        host copy).  Pageable numpy memory cannot be registered here, so one
        host copy of x remains."""
        import threading
        if xa.shape[0] != self.nh or xa.ndim > 2:
            raise ValueError(f"vector shape {xa.shape} does not match nh={self.nh}")
        lock = self.__dict__.setdefault("_host_lock", threading.Lock())
        with lock:
            dt = torch.from_numpy(xa[:0]).dtype
            key = (tuple(xa.shape), dt)
            st = self.__dict__.get("_host_stage")
            if st is None or st[0] != key:
                st = (key, torch.empty(xa.shape, dtype=dt, pin_memory=True),
                      torch.empty(xa.shape, dtype=dt, device="cuda"), torch.empty(xa.shape, dtype=dt, device="cuda"))
                self._host_stage = st
            _, pin_in, xd, yd = st
            pin_in.copy_(torch.from_numpy(xa))
            xd.copy_(pin_in, non_blocking=True)
            self.matvec_device(xd, out=yd)
            out = torch.empty(xa.shape, dtype=dt, pin_memory=True)
            out.copy_(yd, non_blocking=True)
            torch.cuda.current_stream().synchronize()
            return out.numpy()

    @property
    def matvec_fun(self):
        """Callable x -> H x (special_operator.py:593-709): output has the
        dtype and shape of x; numpy in -> numpy out, device tensor in ->
        device tensor out."""
        if self._cached_matvec_fun is None:
            self._device_op()
            self._cached_matvec_fun = self._matvec_any
        return self._cached_matvec_fun

    def projected_matvec(self, vecs, hilbert_out):
        """y = U_out^dagger O U_in x into another sector of the same symmetry
        group (adapters.py:193-206, matrix_builder_jit.py:51-108); complex128
        result, numpy in -> numpy out, device tensor in -> device tensor out."""
        as_np = not isinstance(vecs, torch.Tensor)
        x = torch.as_tensor(np.ascontiguousarray(vecs)) if as_np else vecs
        if x.dim() != 1 or x.shape[0] != self.nh:
            raise ValueError(f"vector shape {tuple(x.shape)} does not match nh={self.nh}")
        x = x.to(device="cuda", dtype=torch.complex128).contiguous()
        y = torch.empty(hilbert_out.nh, dtype=torch.complex128, device="cuda")
        if hilbert_out.nh:
            _lib.check(_lib.lib().qes_projected_matvec(self._device_op(), hilbert_out.handle, _lib.dptr(x), _lib.dptr(y),
                                                       _lib.stream_ptr()), "qes_projected_matvec")
        return y.cpu().numpy() if as_np else y

    def matvec(self, vecs, *args, hilbert_in=None, hilbert_out=None, out=None, **kwargs):
        """Operator.matvec (operator.py:1559-1602): same sector, or into
        ``hilbert_out`` (projected product) when it is another sector."""
        if hilbert_in is not None and hilbert_in is not self._hilbert_space:
            raise ValueError("hilbert_in differs from the operator's sector")
        if hilbert_out is not None and hilbert_out is not self._hilbert_space:
            y = self.projected_matvec(vecs, hilbert_out)
        else:
            y = self._matvec_any(vecs)
        if out is not None:
            if isinstance(out, torch.Tensor):
                out.copy_(y if isinstance(y, torch.Tensor) else torch.from_numpy(y))
            else:
                out[...] = y
            return out
        return y

    # ---- overlaps and matrix elements (operator.py:926-1075) --------------
    @staticmethod
    def _as_columns(vec, name: str):
        """(dim, ncol) view of a vector or column block and whether it was a vector."""
        if isinstance(vec, (list, tuple)):
            vec = np.asarray(vec)
        if not hasattr(vec, "ndim"):
            raise TypeError(f"{name} must be array-like, got {type(vec)!r}.")
        if vec.ndim == 1:
            return vec[:, None], True
        if vec.ndim == 2:
            return vec, False
        raise ValueError(f"{name} must be 1D or 2D, got shape {tuple(vec.shape)}.")

    @staticmethod
    def _column_overlap(left, right, left_is_vec: bool, right_is_vec: bool, pairwise: bool):
        """<left|right> with the reference's column conventions: pairwise
        column overlaps when both sides hold several columns (pairwise=True),
        else the full left^dagger right, squeezed where an input was a vector."""
        is_t = isinstance(left, torch.Tensor)
        nl, nr = int(left.shape[1]), int(right.shape[1])
        if pairwise and nl > 1 and nr > 1:
            if nl != nr:
                raise ValueError(f"Pairwise overlap requires equal column counts, got {nl} and {nr}.")
            return (left.conj() * right).sum(dim=0) if is_t else (left.conj() * right).sum(axis=0)
        out = left.conj().transpose(0, 1) @ right if is_t else left.conj().T @ right
        if left_is_vec and right_is_vec:
            return out[0, 0]
        if left_is_vec:
            return out[0, :]
        if right_is_vec:
            return out[:, 0]
        return out

    def overlap(self, vec1, vec2=None, *, pairwise: bool = True):
        """<vec1|vec2> for vectors or column blocks (operator.py:975-1011)."""
        if vec2 is None:
            vec2 = vec1
        l, lv = self._as_columns(vec1, "vec1")
        r, rv = self._as_columns(vec2, "vec2")
        if l.shape[0] != r.shape[0]:
            raise ValueError(f"State size mismatch: {l.shape[0]} (vec1) vs {r.shape[0]} (vec2).")
        return self._overlap_cols(l, lv, r, rv, pairwise)

    def expectation_value(self, state, other_state=None, *, hilbert_in=None, hilbert_out=None, pairwise: bool = True,
                          hilbert=None, **kwargs):
        """<state|O|other_state> (operator.py:1013-1075): O applied to
        other_state by the device matvec (projected into ``hilbert_out`` when
        that is another sector; then ``state`` lives there), overlaps with the
        reference's column conventions.  numpy in -> numpy out, device tensors
        -> device tensors."""
        if other_state is None:
            other_state = state
        left, lv = self._as_columns(state, "state")
        right, rv = self._as_columns(other_state, "other_state")
        hin = hilbert_in if hilbert_in is not None else hilbert
        if hilbert_out is None and left.shape[0] != right.shape[0]:
            raise ValueError(f"State size mismatch: {left.shape[0]} (state) vs {right.shape[0]} (other_state).")
        if hilbert_out is not None and hilbert_out is not self._hilbert_space:
            cols = [self.matvec(right[:, j], hilbert_in=hin, hilbert_out=hilbert_out) for j in range(right.shape[1])]
            operated = (torch.stack(cols, dim=1) if isinstance(cols[0], torch.Tensor) else np.stack(cols, axis=1))
        else:
            src = right[:, 0] if rv else right
            if isinstance(src, torch.Tensor):
                src = src.contiguous()
            elif not src.flags.c_contiguous:
                src = np.ascontiguousarray(src)
            operated = self.matvec(src, hilbert_in=hin)
            if operated.ndim == 1:
                operated = operated[:, None]
        return self._overlap_cols(left, lv, operated, rv, pairwise)

    def _overlap_cols(self, left, lv, right, rv, pairwise):
        if isinstance(left, torch.Tensor) != isinstance(right, torch.Tensor):
            raise TypeError("both sides must be numpy arrays or both torch tensors")
        if not isinstance(left, torch.Tensor):
            dt = np.result_type(left.dtype, right.dtype)
            left, right = left.astype(dt, copy=False), right.astype(dt, copy=False)
        else:
            dt = torch.promote_types(left.dtype, right.dtype)
            left, right = left.to(dt), right.to(dt)
        return self._column_overlap(left, right, lv, rv, pairwise)

    def __matmul__(self, x):
        return self._matvec_any(x)

    def as_linear_operator(self, dtype=None):
        """scipy LinearOperator with the right shape and adjoint = self."""
        from scipy.sparse.linalg import LinearOperator
        n = self.nh
        dt = np.dtype(dtype) if dtype is not None else np.dtype(self.vector_dtype)
        f = self._matvec_any
        return LinearOperator((n, n), matvec=f, rmatvec=f, matmat=f, rmatmat=f, dtype=dt)

    @property
    def hamil(self):
        return self.as_linear_operator()

    # ---- diagonalisation ------------------------------------------------
    EXACT_AUTO_MAX = 4096      # method='auto': dense below, Lanczos above

    def diagonalize(self, method: str = "exact", k: Optional[int] = None, tol: float = 1e-10, maxiter: int = 300,
                    v0=None, seed: int = 0, reorth: bool = True, store_basis: Optional[bool] = None,
                    compute_eigvec: bool = False, check_every: int = 10, assemble: Optional[bool] = None,
                    which: str = "smallest", **kwargs):
        """Eigenvalues of the reduced H (hamil.py:1809-2114).

        method='exact' (the reference's default): the whole spectrum of the
        dense reduced matrix, built on the device from the assembled table
        and diagonalised there (FP64 ``torch.linalg.eigh``); N^2 doubles
        must fit in a quarter of the free device memory.  'auto': exact up
        to N = 4096, Lanczos above.

        method='block_lanczos' (hamil_diag_engine.py:59): the k extreme
        pairs by block Lanczos with ``block_size`` vectors per step
        (default max(4, k)) and full reorthogonalisation on FP64 matrix
        cores; ``maxiter`` counts block steps.

        method='lanczos': extreme eigenvalue(s) by device Lanczos, k
        defaulting to 6 and which to 'smallest' as in hamil_diag_engine.py:
        183,254-256 ('largest' and 'both' as there).  With the basis kept
        (reorth, the default) all k wanted Ritz pairs are converged to the
        residual test; without it only the extreme one(s).  One start vector
        finds one copy of a degenerate level.

        ``compute_eigvec``: eig_vec is (n,) for one Lanczos pair, (n, k)
        otherwise, (n, n) for 'exact'.  Keyword arguments of the reference's
        other solvers are accepted and ignored, except those that would
        change the result: ``sigma`` (shift-invert) and ``hermitian=False``
        raise.

        ``assemble``: tabulate H for the solve (see ``assemble``); None picks
        it when the table (with its column panels, where x exceeds the
        cache) fits in 40 % of the free device memory (at most 128 GB) and
        N >= 65536.  A table built here is freed afterwards."""
        from .lanczos import block_lanczos, lanczos_ground_state, which_code
        if method not in ("lanczos", "exact", "auto", "block_lanczos"):
            raise NotImplementedError(f"method={method!r}: 'exact', 'lanczos', 'block_lanczos' and 'auto' run on "
                                      f"the device")
        if kwargs.get("sigma") is not None:
            raise NotImplementedError("shift-invert (sigma) is not part of the device Lanczos path")
        if not kwargs.get("hermitian", True):
            raise NotImplementedError("the device Lanczos path is for Hermitian operators")
        which_code(which)
        k = 6 if k is None else int(k)
        if self.nh == 0:
            self._eig_val = np.zeros(0)
            return self._eig_val
        if method == "exact" or (method == "auto" and self.nh <= self.EXACT_AUTO_MAX):
            return self._diagonalize_exact(compute_eigvec)
        built = False
        if self.assembly_bytes == 0 and (assemble or (assemble is None and self.nh >= 65536)):
            free, _ = torch.cuda.mem_get_info()
            try:
                self.assemble(True, max_gb=min(128.0, 0.4 * free / 1e9))
                built = True
            except _lib.QesError:
                if assemble:
                    raise
        try:
            if method == "block_lanczos":
                res = block_lanczos(self, k=k, block_size=kwargs.get("block_size"), tol=tol, maxiter=maxiter, v0=v0,
                                    seed=seed, which=which, compute_eigvec=compute_eigvec)
            else:
                res = lanczos_ground_state(self, k=k, tol=tol, maxiter=maxiter, v0=v0, seed=seed, reorth=reorth,
                                           store_basis=store_basis, compute_eigvec=compute_eigvec,
                                           check_every=check_every, which=which)
        finally:
            if built:
                self.assemble(False)
        res.assembled = built or self.assembly_bytes > 0
        self.last_lanczos = res
        self._eig_val = res.eigenvalues
        self._eig_vec = res.eigenvectors
        return self._eig_val

    def dense_device(self) -> torch.Tensor:
        """The reduced H as a dense (N, N) device matrix (float64 or
        complex128), filled from the assembled table on the device."""
        n = self.nh
        cplx = bool(self.needs_complex)
        itemsize = 16 if cplx else 8
        free, _ = torch.cuda.mem_get_info()
        if n * n * itemsize > 0.25 * free:
            raise _lib.QesError(f"dense H of N={n} takes {n * n * itemsize / 1e9:.1f} GB, more than a quarter "
                                f"of the free device memory; use method='lanczos'")
        built = False
        if self.assembly_bytes == 0:
            self.assemble(True)
            built = True
        try:
            L = _lib.lib()
            cnt = C.c_int64()
            _lib.check(L.qes_op_table_coo(self._device_op(), None, None, None, 0, C.byref(cnt), None),
                       "qes_op_table_coo")
            m = int(cnt.value)
            rows = torch.full((m,), -1, dtype=torch.int32, device="cuda")
            cols = torch.zeros((m,), dtype=torch.int32, device="cuda")
            vals = torch.zeros((m, 2) if cplx else (m,), dtype=torch.float64, device="cuda")
            _lib.check(L.qes_op_table_coo(self._device_op(), _lib.dptr(rows), _lib.dptr(cols), _lib.dptr(vals), m,
                                          C.byref(cnt), _lib.stream_ptr()), "qes_op_table_coo")
        finally:
            if built:
                self.assemble(False)
        v = torch.view_as_complex(vals) if cplx else vals
        keep = rows >= 0                       # padding slots carry row -1
        A = torch.zeros((n, n), dtype=v.dtype, device="cuda")
        A.index_put_((rows[keep].long(), cols[keep].long()), v[keep], accumulate=True)
        return A

    def _diagonalize_exact(self, compute_eigvec: bool):
        A = self.dense_device()
        if compute_eigvec:
            w, V = torch.linalg.eigh(A)
            self._eig_vec = V
        else:
            w = torch.linalg.eigvalsh(A)
            self._eig_vec = None
        del A
        self._eig_val = w.cpu().numpy()
        self.last_lanczos = None
        return self._eig_val

    @property
    def eig_val(self):
        return self._eig_val

    @property
    def eigenvalues(self):
        return self._eig_val

    @property
    def eig_vec(self):
        return self._eig_vec

    def __del__(self):
        try:
            if self._op_handle is not None:
                _lib.lib().qes_op_destroy(self._op_handle)
        except Exception:
            pass


class Operator(Hamiltonian):
    """A general spin-1/2 operator (term list on a sector), e.g. a
    momentum-resolved S^z_q for sector-changing products; same API."""

    def __init__(self, hilbert_space, dtype=np.complex128, name: str = "Operator", **kwargs):
        super().__init__(hilbert_space, dtype=dtype, name=name, **kwargs)
