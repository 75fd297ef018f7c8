"""HilbertSpace: a symmetry-reduced spin-1/2 sector whose basis lives on the GPU.

Mirrors the constructor and accessors of the reference HilbertSpace that the
ED path uses (Algebra/hilbert.py:132-386, :1012-1207; accessors in
Algebra/Hilbert/hilbert_base.py:276-376): ``nh``, ``nhfull``, ``ns``,
``representative_list``, ``normalization``, ``sym_group``, plus the
compact-map view (``repr_map``/``phase_idx`` entries for given states).

The basis is built by libqes_mi on the device: no 2^Ns map is ever formed.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from .symmetry import (GlobalSymmetry, SectorGroup, build_sector_group, get_u1_sym,
                       normalize_symmetry_generators)


class HilbertSpace:
    def __init__(self, lattice=None, ns: Optional[int] = None, sym_gen=None, global_syms=None,
                 is_manybody: bool = True, gen_mapping: bool = True, boundary_flux=None, **kwargs):
        if lattice is None and ns is None:
            raise ValueError("HilbertSpace needs a lattice or ns")
        if boundary_flux is not None:           # hilbert_base.py:106-108: set on the lattice
            if lattice is None:
                raise ValueError("boundary_flux needs a lattice")
            lattice.flux = boundary_flux
        self.lattice = lattice
        self._ns = int(ns if ns is not None else lattice.ns)
        if lattice is not None and lattice.ns != self._ns:
            raise ValueError("ns does not match the lattice")
        if isinstance(global_syms, GlobalSymmetry):
            global_syms = [global_syms]
        self.global_syms = list(global_syms or [])
        specs = normalize_symmetry_generators(sym_gen)
        self.group: SectorGroup = build_sector_group(lattice, self._ns, specs, self.global_syms)
        _lib.require_gpu()
        L = _lib.lib()
        h = C.c_void_p()
        n = C.c_int64()
        spec = self.group.spec()
        _lib.check(L.qes_basis_create(C.byref(spec), C.byref(h), C.byref(n)), "qes_basis_create")
        self._h = h
        self._nh = int(n.value)
        self._reps = None
        self._norms = None

    # ---- reference accessors -----------------------------------------
    @property
    def ns(self) -> int:
        return self._ns

    @property
    def Ns(self) -> int:
        return self._ns

    @property
    def nh(self) -> int:
        return self._nh

    @property
    def dim(self) -> int:
        return self._nh

    @property
    def nhfull(self) -> int:
        return 1 << self._ns

    @property
    def handle(self):
        return self._h

    @property
    def sym_group(self):
        return self.group

    @property
    def is_complex(self) -> bool:
        return not self.group.is_real

    def _fetch(self):
        if self._reps is None:
            reps = np.zeros(self._nh, dtype=np.uint64)
            norms = np.zeros(self._nh, dtype=np.float64)
            if self._nh:
                _lib.check(_lib.lib().qes_basis_get(self._h, reps.ctypes.data_as(_lib._PU64),
                                                    norms.ctypes.data_as(_lib._PD)), "qes_basis_get")
            self._reps, self._norms = reps, norms

    @property
    def representative_list(self) -> np.ndarray:
        """Sorted representatives as int64 (symmetry_container.py:1698-1702)."""
        self._fetch()
        return self._reps.astype(np.int64)

    @property
    def normalization(self) -> np.ndarray:
        self._fetch()
        return self._norms

    def map_states(self, states):
        """(repr_map, phase_idx) entries of the reference compact map for
        ``states`` (symmetry_container_jit.py:248-255): rep index (-1 outside
        the sector) and phase-table index (255 outside)."""
        st = np.ascontiguousarray(np.asarray(states, dtype=np.uint64))
        idx = np.zeros(st.shape, dtype=np.int64)
        g = np.zeros(st.shape, dtype=np.int32)
        if st.size:
            _lib.check(_lib.lib().qes_basis_map_states(self._h, st.ctypes.data_as(_lib._PU64), st.size,
                                                       idx.ctypes.data_as(_lib._PI64),
                                                       g.ctypes.data_as(_lib._PI32)), "qes_basis_map_states")
        pidx = np.where(idx >= 0, self.group.g_to_pidx[np.maximum(g, 0)], 255).astype(np.uint8)
        return idx, pidx

    def canon_check(self, states, kind: str) -> int:
        """Number of states where the fast canonicaliser ``kind`` ('cyclic' or
        'rows') disagrees with the brute-force one."""
        st = np.ascontiguousarray(np.asarray(states, dtype=np.uint64))
        bad = C.c_int64()
        k = {"cyclic": 1, "rows": 2}[kind]
        _lib.check(_lib.lib().qes_basis_canon_check(self._h, k, st.ctypes.data_as(_lib._PU64), st.size,
                                                    C.byref(bad)), "qes_basis_canon_check")
        return int(bad.value)

    LOOKUP_MODES = {"sorted": 0, "hash": 1, "packed": 2, "rank": 3}

    def set_lookup(self, mode: str = "packed", bucket_rows_log2: Optional[int] = None):
        """Representative lookup of the matvec: 'sorted' (32/64-bit keys),
        'hash', 'packed' (64-bit key | stabiliser size) or 'rank' (bitmap
        over the colex ranks of a U(1) filling, ns <= 32); optionally rebuild
        the bucket table for about 2^bucket_rows_log2 representatives per bucket."""
        _lib.check(_lib.lib().qes_basis_set_lookup(self._h, self.LOOKUP_MODES[mode]), "qes_basis_set_lookup")
        if bucket_rows_log2 is not None:
            _lib.check(_lib.lib().qes_basis_set_bucket(self._h, int(bucket_rows_log2)), "qes_basis_set_bucket")

    # ---- reduced <-> full -----------------------------------------------
    def _full_op(self, vec, reduce: bool):
        import torch
        as_np = not isinstance(vec, torch.Tensor)
        t = torch.as_tensor(np.ascontiguousarray(vec)) if as_np else vec
        if t.dim() != 1:
            raise ValueError("one vector at a time")
        cplx = t.is_complex() or not self.group.is_real
        dt = torch.complex128 if cplx else torch.float64
        t = t.to(device="cuda", dtype=dt).contiguous()
        n_out = self._nh if reduce else (1 << self._ns)
        n_in = (1 << self._ns) if reduce else self._nh
        if t.shape[0] != n_in:
            raise ValueError(f"vector length {t.shape[0]} does not match {n_in}")
        out = torch.empty(n_out, dtype=dt, device="cuda")
        L = _lib.lib()
        if reduce:
            _lib.check(L.qes_basis_reduce(self._h, _lib.dptr(t), _lib.dptr(out), 1 if cplx else 0, _lib.stream_ptr()),
                       "qes_basis_reduce")
        else:
            _lib.check(L.qes_basis_expand(self._h, _lib.dptr(t), _lib.dptr(out), 1 if cplx else 0, _lib.stream_ptr()),
                       "qes_basis_expand")
        return out.cpu().numpy() if as_np else out

    def expand_state(self, vec_reduced, vec_full=None):
        """Reduced -> full space (hilbert.py:1659-1715): a 2^ns vector,
        numpy in -> numpy out, device tensor in -> device tensor out."""
        out = self._full_op(vec_reduced, reduce=False)
        if vec_full is not None:
            vec_full[...] = out
            return vec_full
        return out

    def transform_to_reduced_space(self, vec_full):
        """Full -> reduced space (hilbert_base.py:485-563)."""
        return self._full_op(vec_full, reduce=True)

    def device_arrays(self):
        """Device pointers (reps uint64, norms float64) owned by this space."""
        r = C.c_void_p()
        nm = C.c_void_p()
        _lib.check(_lib.lib().qes_basis_device_ptrs(self._h, C.byref(r), C.byref(nm)), "qes_basis_device_ptrs")
        return r.value, nm.value

    def __len__(self):
        return self._nh

    def __repr__(self):
        gens = ",".join(f"{g.name}={s}" for g, s in self.group.generators)
        return f"HilbertSpace(ns={self._ns}, nh={self._nh}, |G|={self.group.order}, {gens})"

    def __del__(self):
        try:
            if getattr(self, "_h", None) and self._h.value:
                _lib.lib().qes_basis_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass


__all__ = ["HilbertSpace", "get_u1_sym"]
