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
from itertools import product
from typing import Dict, Iterator, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .symmetry import (GlobalSymmetry, SectorGroup, SymmetryGenerators, build_sector_group, get_u1_sym,
                       normalize_symmetry_generators)

_TRANSLATIONS = (SymmetryGenerators.Translation_x, SymmetryGenerators.Translation_y, SymmetryGenerators.Translation_z)
_SPATIAL_PARITIES = (SymmetryGenerators.Reflection, SymmetryGenerators.Inversion)


def _lattice_extents(lattice) -> Tuple[int, int, int]:
    lx = int(getattr(lattice, "lx", None) or getattr(lattice, "Lx", 0) or 0)
    ly = int(getattr(lattice, "ly", None) or getattr(lattice, "Ly", 1) or 1)
    lz = int(getattr(lattice, "lz", None) or getattr(lattice, "Lz", 1) or 1)
    return lx, ly, lz


def _momentum_is_real(lattice, sector: Dict) -> bool:
    """k = -k (mod L) in every translated direction: the momenta at which a
    spatial parity (P T P^-1 = T^-1) commutes with the translations, i.e. at
    which (k, parity) is a one-dimensional irrep and the sector block is
    Hermitian."""
    ext = _lattice_extents(lattice) if lattice is not None else (0, 0, 0)
    for gen, L in zip(_TRANSLATIONS, ext):
        if gen in sector and L:
            k = int(sector[gen]) % L
            if (2 * k) % L != 0:
                return False
    return True


def sector_specs(symmetry_types: Sequence[Tuple[SymmetryGenerators, Sequence]], lattice=None,
                 physical_only: bool = True) -> List[Tuple[Dict, List[Tuple[SymmetryGenerators, object]]]]:
    """(sector_dict, sym_gen) of every sector that iter_symmetry_sectors
    visits, in the reference's order: itertools.product over the sector lists
    (Algebra/hilbert.py:520-543).  With ``physical_only`` a spatial parity
    (Reflection, Inversion) is kept only at momenta with k = -k; at other
    momenta the sector is listed once, without that generator and with
    parity value 0 in sector_dict (the reference lists both parities there;
    those blocks are not Hermitian and their dimensions double-count the
    momentum sector)."""
    if not symmetry_types:
        raise ValueError("At least one symmetry type must be provided.")
    gens = [g for g, _ in symmetry_types]
    lists = [list(v) for _, v in symmetry_types]
    out, seen = [], set()
    for values in product(*lists):
        sector = dict(zip(gens, values))
        sym_gen = list(zip(gens, values))
        if physical_only and any(g in sector for g in _SPATIAL_PARITIES) and not _momentum_is_real(lattice, sector):
            for g in _SPATIAL_PARITIES:
                if g in sector:
                    sector[g] = 0
            sym_gen = [(g, v) for g, v in sym_gen if g not in _SPATIAL_PARITIES]
            key = tuple(sector.items())
            if key in seen:
                continue
            seen.add(key)
        out.append((sector, sym_gen))
    return out


def momentum_symmetry_types(lattice, include_parity: bool = False) -> List[Tuple[SymmetryGenerators, Sequence]]:
    """Translation sector lists of a lattice (x, then y and z where the
    lattice extends that way) plus reflection parities on request
    (Algebra/hilbert.py:626-642)."""
    lx, ly, lz = _lattice_extents(lattice)
    if lx <= 0:
        raise ValueError("Lattice must have positive lx dimension.")
    types = [(SymmetryGenerators.Translation_x, range(lx))]
    if ly > 1:
        types.append((SymmetryGenerators.Translation_y, range(ly)))
    if lz > 1:
        types.append((SymmetryGenerators.Translation_z, range(lz)))
    if include_parity:
        types.append((SymmetryGenerators.Reflection, [1, -1]))
    return types


def momentum_key(lattice, sector: Dict, include_parity: bool = False) -> Tuple:
    """(kx[, ky[, kz]][, parity]) as iter_momentum_sectors yields it."""
    lx, ly, lz = _lattice_extents(lattice)
    key = [sector[SymmetryGenerators.Translation_x]]
    if ly > 1:
        key.append(sector[SymmetryGenerators.Translation_y])
    if lz > 1:
        key.append(sector[SymmetryGenerators.Translation_z])
    if include_parity:
        key.append(sector[SymmetryGenerators.Reflection])
    return tuple(key)


_ALL_SECTOR_NAMES = {
    "translation": "translation", "momentum": "translation", "trans": "translation",
    "reflection": "reflection", "parityspatial": "reflection", "spatial": "reflection", "mirror": "reflection",
    "inversion": "inversion", "inv": "inversion", "spatialinversion": "inversion",
    "parityz": "parity_z", "spinparity": "parity_z", "pz": "parity_z", "spinflip": "parity_z",
    "parityx": "parity_x", "px": "parity_x",
    "parityy": "parity_y", "py": "parity_y",
}


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

    # ---- sector iteration (Algebra/hilbert.py:412-993) -------------------
    @staticmethod
    def iter_symmetry_sectors(symmetry_types, *, lattice=None, ns: Optional[int] = None, gen_mapping: bool = True,
                              verbose: bool = False, physical_only: bool = True, **hilbert_kwargs) -> Iterator:
        """Yield (sector_dict, HilbertSpace) for the Cartesian product of
        the (generator, sector list) pairs (Algebra/hilbert.py:412-543), one
        space at a time.  ``physical_only`` (default) keeps spatial parities
        only at momenta k = -k, see :func:`sector_specs`; False gives the
        reference's literal product."""
        if lattice is None and ns is None:
            raise ValueError("Either 'lattice' or 'ns' must be provided.")
        specs = sector_specs(symmetry_types, lattice, physical_only)
        for i, (sector, sym_gen) in enumerate(specs):
            if verbose:
                txt = ", ".join(f"{g.name}={v}" for g, v in sector.items())
                print(f"[{i + 1}/{len(specs)}] Building sector: {txt}")
            yield sector, HilbertSpace(lattice=lattice, ns=ns, sym_gen=sym_gen, gen_mapping=gen_mapping,
                                       **hilbert_kwargs)

    @staticmethod
    def iter_momentum_sectors(lattice, *, include_parity: bool = False, gen_mapping: bool = True,
                              verbose: bool = False, physical_only: bool = True, **hilbert_kwargs) -> Iterator:
        """Yield ((kx[, ky[, kz]][, parity]), HilbertSpace) over every
        momentum of the lattice (Algebra/hilbert.py:546-652); with
        ``include_parity`` the reflection parities +1, -1 at momenta k = -k
        and parity 0 (no reflection) elsewhere unless ``physical_only`` is
        False."""
        types = momentum_symmetry_types(lattice, include_parity)
        for sector, hs in HilbertSpace.iter_symmetry_sectors(types, lattice=lattice, gen_mapping=gen_mapping,
                                                             verbose=verbose, physical_only=physical_only,
                                                             **hilbert_kwargs):
            yield momentum_key(lattice, sector, include_parity), hs

    @staticmethod
    def iter_reflection_sectors(lattice, *, momentum_sector: Optional[int] = None, gen_mapping: bool = True,
                                verbose: bool = False, **hilbert_kwargs) -> Iterator:
        """Yield (parity, HilbertSpace) for reflection parity +1, -1
        (Algebra/hilbert.py:655-712)."""
        types = [] if momentum_sector is None else [(SymmetryGenerators.Translation_x, [momentum_sector])]
        types.append((SymmetryGenerators.Reflection, [1, -1]))
        for sector, hs in HilbertSpace.iter_symmetry_sectors(types, lattice=lattice, gen_mapping=gen_mapping,
                                                             verbose=verbose, physical_only=False, **hilbert_kwargs):
            yield sector[SymmetryGenerators.Reflection], hs

    @staticmethod
    def iter_inversion_sectors(lattice, *, momentum_sector: Optional[int] = None, gen_mapping: bool = True,
                               verbose: bool = False, **hilbert_kwargs) -> Iterator:
        """Yield (parity, HilbertSpace) for spatial inversion +1, -1
        (Algebra/hilbert.py:714-783)."""
        types = [] if momentum_sector is None else [(SymmetryGenerators.Translation_x, [momentum_sector])]
        types.append((SymmetryGenerators.Inversion, [1, -1]))
        for sector, hs in HilbertSpace.iter_symmetry_sectors(types, lattice=lattice, gen_mapping=gen_mapping,
                                                             verbose=verbose, physical_only=False, **hilbert_kwargs):
            yield sector[SymmetryGenerators.Inversion], hs

    @staticmethod
    def iter_parity_sectors(lattice, axis: str = "z", *, gen_mapping: bool = True, verbose: bool = False,
                            **hilbert_kwargs) -> Iterator:
        """Yield (parity, HilbertSpace) for the spin-flip parity along
        ``axis`` (Algebra/hilbert.py:785-844)."""
        gens = {"x": SymmetryGenerators.ParityX, "y": SymmetryGenerators.ParityY, "z": SymmetryGenerators.ParityZ}
        if axis.lower() not in gens:
            raise ValueError(f"Unknown parity axis '{axis}'. Use 'x', 'y', or 'z'.")
        g = gens[axis.lower()]
        for sector, hs in HilbertSpace.iter_symmetry_sectors([(g, [1, -1])], lattice=lattice, gen_mapping=gen_mapping,
                                                             verbose=verbose, **hilbert_kwargs):
            yield sector[g], hs

    @staticmethod
    def all_sector_types(lattice, symmetries: Optional[List[str]] = None):
        """(symmetry_types, friendly names) of iter_all_sectors
        (Algebra/hilbert.py:903-978)."""
        symmetries = ["translation"] if symmetries is None else list(symmetries)
        lx, ly, lz = _lattice_extents(lattice)
        types, names = [], {}
        for sym in symmetries:
            kind = _ALL_SECTOR_NAMES.get(sym.lower().replace("_", "").replace("-", ""))
            if kind == "translation":
                for gen, L, nm in zip(_TRANSLATIONS, (lx, ly, lz), ("kx", "ky", "kz")):
                    if gen is SymmetryGenerators.Translation_x or L > 1:
                        types.append((gen, list(range(L))))
                        names[gen] = nm
            elif kind == "reflection":
                types.append((SymmetryGenerators.Reflection, [1, -1]))
                names[SymmetryGenerators.Reflection] = "reflection"
            elif kind == "inversion":
                types.append((SymmetryGenerators.Inversion, [1, -1]))
                names[SymmetryGenerators.Inversion] = "inversion"
            elif kind in ("parity_x", "parity_y", "parity_z"):
                gen = {"parity_x": SymmetryGenerators.ParityX, "parity_y": SymmetryGenerators.ParityY,
                       "parity_z": SymmetryGenerators.ParityZ}[kind]
                types.append((gen, [1, -1]))
                names[gen] = kind
            else:
                raise ValueError(f"Unknown symmetry '{sym}'. Options: translation, reflection, inversion, "
                                 f"parity_z, parity_x, parity_y")
        return types, names

    @staticmethod
    def iter_all_sectors(lattice, symmetries: Optional[List[str]] = None, *, gen_mapping: bool = True,
                         verbose: bool = False, physical_only: bool = True, **hilbert_kwargs) -> Iterator:
        """Yield ({name: value}, HilbertSpace) over every combination of the
        named symmetries' sectors (Algebra/hilbert.py:846-993): names kx,
        ky, kz, reflection, inversion, parity_x, parity_y, parity_z."""
        types, names = HilbertSpace.all_sector_types(lattice, symmetries)
        for sector, hs in HilbertSpace.iter_symmetry_sectors(types, lattice=lattice, gen_mapping=gen_mapping,
                                                             verbose=verbose, physical_only=physical_only,
                                                             **hilbert_kwargs):
            yield {names[g]: v for g, v in sector.items()}, hs

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


__all__ = ["HilbertSpace", "get_u1_sym", "sector_specs", "momentum_symmetry_types", "momentum_key"]
