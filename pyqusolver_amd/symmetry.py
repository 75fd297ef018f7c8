"""Symmetry generators, sector parsing and the device group descriptor.

Mirrors the reference's user-facing symmetry API:
  * ``SymmetryGenerators`` (Algebra/Operator/operator.py:85-152),
  * ``parse_symmetry_spec`` / ``normalize_symmetry_generators``
    (Algebra/Symmetries/symmetry_container.py:2036-2224),
  * the global-symmetry filter (Algebra/Symmetries/parity.py:283-321),
  * ``build_group`` / ``get_character`` / ``build_phases_map``
    (symmetry_container.py:866-1174, :1399-1442, :1767-1804).

Deliberate deviation: two sector kinds keep their (bit-exact) basis but
refuse the matvec, because the reference block is not Hermitian there:
  * reflection together with a translation at a generic momentum (k not 0
    or L/2): the reference accepts it (is_real_sector is always True,
    base.py:581-609) but r*exp(2 pi i k n/L) is not a representation of the
    dihedral group;
  * parity y: its +-i intrinsic phase (parity.py:55-63) enters the
    normalisation (symmetry_container_jit.py:154-174) but not the matrix
    elements (phase_table holds characters only, symmetry_container.py:1783-1803).
"""
from __future__ import annotations

import ctypes as C
import itertools
from dataclasses import dataclass, field
from enum import Enum
from typing import List, Optional, Sequence, Tuple

import numpy as np

SYM_NORM_THRESHOLD = 1e-7

BOP_NONE, BOP_REFLECT, BOP_FLIP, BOP_FLIP_Y, BOP_PERM = 0, 1, 2, 3, 4


class SymmetryGenerators(Enum):
    Translation_x = "Translation_x"
    Translation_y = "Translation_y"
    Translation_z = "Translation_z"
    Reflection = "Reflection"
    ParityX = "ParityX"
    ParityY = "ParityY"
    ParityZ = "ParityZ"
    Inversion = "Inversion"


class GlobalSymmetries(Enum):
    U1 = 1
    Z2_PARITY = 2


@dataclass
class GlobalSymmetry:
    name: GlobalSymmetries
    val: float
    lat: object = None

    def get_val(self):
        return self.val


def get_u1_sym(lat, val) -> GlobalSymmetry:
    """U(1): popcount == val (Algebra/globals.py:164-189)."""
    return GlobalSymmetry(GlobalSymmetries.U1, val, lat)


def get_z2_sym(lat, val) -> GlobalSymmetry:
    """Popcount parity (+1 even, -1 odd) (Algebra/globals.py:336-339)."""
    return GlobalSymmetry(GlobalSymmetries.Z2_PARITY, val, lat)


def parse_symmetry_spec(name: str, value) -> List[Tuple[SymmetryGenerators, object]]:
    """symmetry_container.py:2036-2144 (the spin-1/2 subset)."""
    key = name.lower().replace("_", "").replace("-", "")
    out = []
    if key in {"translation", "translations", "trans", "momentum"}:
        if isinstance(value, dict):
            for kname, gen in (("kx", SymmetryGenerators.Translation_x), ("ky", SymmetryGenerators.Translation_y),
                               ("kz", SymmetryGenerators.Translation_z)):
                alt = kname[0] + "_" + kname[1]
                if kname in value or alt in value:
                    out.append((gen, value.get(kname, value.get(alt))))
        else:
            out.append((SymmetryGenerators.Translation_x, value))
    elif key in {"parity", "parityx", "parityy", "parityz", "spin"}:
        if isinstance(value, dict):
            for ax, gen in (("x", SymmetryGenerators.ParityX), ("y", SymmetryGenerators.ParityY),
                            ("z", SymmetryGenerators.ParityZ)):
                if ax in value or "p" + ax in value:
                    out.append((gen, value.get(ax, value.get("p" + ax))))
        else:
            out.append((SymmetryGenerators.ParityZ, int(value)))
    elif key in {"reflection", "reflect", "mirror"}:
        out.append((SymmetryGenerators.Reflection, value.get("sector", 1) if isinstance(value, dict) else value))
    elif key in {"inversion", "inv", "spatial", "spatialinversion"}:
        out.append((SymmetryGenerators.Inversion, value.get("sector", 1) if isinstance(value, dict) else value))
    else:
        raise ValueError(f"Unknown symmetry name: '{name}'")
    return out


def normalize_symmetry_generators(sym_gen) -> List[Tuple[SymmetryGenerators, object]]:
    """symmetry_container.py:2147-2224."""
    if sym_gen is None:
        return []
    if isinstance(sym_gen, dict):
        out = []
        for k, v in sym_gen.items():
            out.extend(parse_symmetry_spec(k, v))
        return out
    if isinstance(sym_gen, (list, tuple)):
        out = []
        for item in sym_gen:
            g, v = item
            if isinstance(g, SymmetryGenerators):
                out.append((g, v))
            else:
                out.extend(parse_symmetry_spec(str(g), v))
        return out
    raise ValueError(f"sym_gen must be dict, list or None, got {type(sym_gen)}")


@dataclass
class SectorGroup:
    """A compiled symmetry sector: the device group descriptor plus the
    reference-ordered characters and phase table."""
    ns: int
    n_up: int                      # -1 if no U(1)
    z2: int                        # 0 if none
    generators: List[Tuple[SymmetryGenerators, object]]
    nbase: int
    base_ops: List[List[int]]
    ntx: int
    nty: int
    row_bits: int
    cell_bits: int
    chi: np.ndarray                # complex128, reference element order
    g_to_pidx: np.ndarray = field(default=None)
    phase_table: np.ndarray = field(default=None)
    matvec_error: Optional[str] = None    # sector where the reference block is not Hermitian
    # twisted translations: crossing masks (ntx + nty, bit = site) and the
    # boundary phase table (2, ns + 1) of the x and y directions
    cross_masks: Optional[np.ndarray] = None
    bphase: Optional[np.ndarray] = None
    # inversion that is not the full site reversal: its site permutation,
    # applied to bits as the reference does (bit i -> bit perm[i])
    perm: Optional[np.ndarray] = None

    @property
    def order(self) -> int:
        return self.ntx * self.nty * self.nbase

    @property
    def is_real(self) -> bool:
        return bool(np.all(np.abs(self.chi.imag) <= 1e-12))

    def spec(self):
        from ._lib import GroupSpec
        s = GroupSpec()
        s.ns, s.n_up, s.z2, s.nbase = self.ns, self.n_up, self.z2, self.nbase
        for b in range(4):
            for k in range(4):
                s.base_ops[b][k] = self.base_ops[b][k] if b < self.nbase and k < len(self.base_ops[b]) else 0
        s.ntx, s.nty, s.row_bits, s.cell_bits = self.ntx, self.nty, self.row_bits, self.cell_bits
        self._chi_re = np.ascontiguousarray(self.chi.real, dtype=np.float64)
        self._chi_im = np.ascontiguousarray(self.chi.imag, dtype=np.float64)
        s.chi_re = self._chi_re.ctypes.data_as(C.POINTER(C.c_double))
        s.chi_im = self._chi_im.ctypes.data_as(C.POINTER(C.c_double))
        if self.cross_masks is not None:
            self._cm = np.ascontiguousarray(self.cross_masks, dtype=np.uint64)
            bp = np.empty((2, self.ns + 1, 2))
            bp[..., 0], bp[..., 1] = self.bphase.real, self.bphase.imag
            self._bp = np.ascontiguousarray(bp.reshape(-1))
            s.cross_masks = self._cm.ctypes.data_as(C.POINTER(C.c_uint64))
            s.bphase = self._bp.ctypes.data_as(C.POINTER(C.c_double))
        if self.perm is not None:
            self._perm = np.ascontiguousarray(self.perm, dtype=np.int32)
            s.perm = self._perm.ctypes.data_as(C.POINTER(C.c_int32))
        return s


def inversion_permutation(lattice, tol: float = 1e-7) -> np.ndarray:
    """Site map of spatial inversion through the centre of the bounding box
    of the site positions (InversionSymmetry, Algebra/Symmetries/
    inversion.py:135-230): perm[i] = the site at 2 r0 - r_i, periodic images
    included.  The reference applies it to bit positions (bit i -> bit
    perm[i], inversion.py:53-86)."""
    r = np.asarray(lattice.rvectors, dtype=float)
    ns = r.shape[0]
    r0 = 0.5 * (r.min(axis=0) + r.max(axis=0))
    periods = [lattice.lx * lattice.a1, lattice.ly * lattice.a2, lattice.lz * lattice.a3]
    flags = lattice.periodic_flags()
    shifts = [np.zeros(3)]
    for d in range(3):
        if flags[d]:
            shifts = [s + m * periods[d] for s in shifts for m in (-1, 0, 1)]
    perm = np.empty(ns, dtype=np.int64)
    for i in range(ns):
        t = 2.0 * r0 - r[i]
        best_j, best_d = -1, np.inf
        for s in shifts:
            d = np.linalg.norm(r - (t + s), axis=1)
            j = int(np.argmin(d))
            if d[j] < best_d:
                best_j, best_d = j, float(d[j])
        if best_d > tol:
            raise ValueError(f"inversion image of site {i} is not a lattice site (distance {best_d:.3g})")
        perm[i] = best_j
    if len(np.unique(perm)) != ns:
        raise ValueError("inversion is not a bijection of the lattice sites")
    return perm


def _twist_tables(lattice, ns: int, ntx: int, nty: int):
    """Crossing masks and boundary phases of twisted translations.  Bit i is
    site i for the group action (translation.py:50-89); a site crosses the
    x boundary during a x steps when its Bravais x >= lx - a (translation.py:
    316-437 marks x = lx - 1 for one step) and the y boundary during b y
    steps when y >= ly - b.  Phases: Lattice.boundary_phases(), exp(i phi_d n)
    for n crossing up spins (rows x and y)."""
    nb = getattr(lattice, "n_basis", 1)
    masks = np.zeros(ntx + nty, dtype=np.uint64)
    for site in range(ns):
        cell = site // nb
        x, y = cell % lattice.lx, (cell // lattice.lx) % max(lattice.ly, 1)
        for a in range(ntx):
            if ntx > 1 and x >= lattice.lx - a:
                masks[a] |= np.uint64(1) << np.uint64(site)
        for b in range(nty):
            if nty > 1 and y >= lattice.ly - b:
                masks[ntx + b] |= np.uint64(1) << np.uint64(site)
    ph = np.asarray(lattice.boundary_phases(), dtype=np.complex128)[:2]
    return masks, ph


def _is_real_momentum(k, L) -> bool:
    k = int(k) % L
    return k == 0 or (L % 2 == 0 and k == L // 2)


def build_sector_group(lattice, ns: int, specs, global_syms: Sequence[GlobalSymmetry] = ()) -> SectorGroup:
    n_up, z2 = -1, 0
    for g in global_syms or []:
        if g.name == GlobalSymmetries.U1:
            n_up = int(g.val)
        elif g.name == GlobalSymmetries.Z2_PARITY:
            z2 = int(g.val)
    # global-symmetry filter (parity.py:283-321) and pairwise rules
    gens = []
    for gen, sec in specs:
        if gen in (SymmetryGenerators.ParityX, SymmetryGenerators.ParityY, SymmetryGenerators.ParityZ) and n_up >= 0:
            if gen == SymmetryGenerators.ParityZ:
                continue          # acts trivially with U(1): dropped
            if n_up != ns // 2 or ns % 2 != 0:
                continue          # X/Y need half filling
        if gen == SymmetryGenerators.Translation_z:
            raise NotImplementedError(f"{gen.name} is not supported (spin chains and 2D lattices only)")
        if gen == SymmetryGenerators.Inversion:
            if lattice is None:
                raise ValueError("inversion symmetry requires a lattice")
            if int(sec) not in (1, -1):
                raise ValueError(f"Inversion sector must be +1 or -1, got {sec}")
        if any(g == gen for g, _ in gens):
            raise ValueError(f"duplicate generator {gen.name}")
        gens.append((gen, sec))
    if any(g in (SymmetryGenerators.Translation_x, SymmetryGenerators.Translation_y) for g, _ in gens) and lattice is None:
        raise ValueError("translation symmetry requires a lattice")
    lx = getattr(lattice, "lx", ns) if lattice is not None else ns
    ly = getattr(lattice, "ly", 1) if lattice is not None else 1
    nb = getattr(lattice, "n_basis", 1) if lattice is not None else 1
    trans = {g: s for g, s in gens if g in (SymmetryGenerators.Translation_x, SymmetryGenerators.Translation_y)}
    flux = getattr(lattice, "flux", None) or {}
    twisted = []
    for g in trans:
        d = "x" if g == SymmetryGenerators.Translation_x else "y"
        phi = float(flux.get(d, 0.0))
        if phi != 0.0:
            # the twisted translation T' = (boundary phase per crossing up
            # spin) T (symmetry_container_jit.py:91-98): T'^L multiplies by
            # exp(i phi N_up), so the momenta exp(2 pi i k / L) label sectors
            # only when phi N_up is a multiple of 2 pi
            if n_up < 0 or abs(np.exp(1j * phi * n_up) - 1.0) > 1e-9:
                raise NotImplementedError(
                    f"translation along {d} with a boundary flux phi={phi} along {d} needs a U(1) filling N_up with "
                    f"phi * N_up a multiple of 2 pi (got N_up={n_up}): otherwise the twisted translation's L-th "
                    "power is not the identity and exp(2 pi i k / L) does not label its sectors")
            twisted.append(d)
    others = [(g, s) for g, s in gens if g not in trans]
    spatial = [g for g, _ in others if g in (SymmetryGenerators.Reflection, SymmetryGenerators.Inversion)]
    matvec_error = None
    if spatial:
        # P T P^-1 = T^-1: chi(T^n P) = exp(ikn) p is a representation only at k = 0, pi
        for g, k in trans.items():
            L = lx if g == SymmetryGenerators.Translation_x else ly
            if not _is_real_momentum(k, L):
                name = spatial[0].name.lower()
                matvec_error = (f"{name} is only a symmetry of momentum sectors k=0 or k=L/2 (got k={k}, L={L}); "
                                "the reference block is not Hermitian there -- use the translation sector alone")
    # base elements: identity + combinations of the other generators in order (:918-928)
    if any(g == SymmetryGenerators.ParityY for g, _ in others):
        # the reference's parity-y phase (+i for an even number of down spins,
        # parity.py:55-63) does not square to one on every state, so the
        # elements do not form a representation: no consistent block
        matvec_error = ("parity-y sectors: the reference's +-i intrinsic phase (parity.py:55-63) does not make the "
                        "group elements a representation, so the block is not Hermitian -- use parity x")
    base = [()]
    for r in range(1, len(others) + 1):
        for combo in itertools.combinations(range(len(others)), r):
            base.append(tuple(others[i] for i in combo))
    if len(base) > 4:
        raise NotImplementedError("at most two non-translation generators")
    base_ops, base_chi = [], []
    inv_perm = None
    for elem in base:
        ops, chi = [], 1.0
        for g, sec in elem:
            if g == SymmetryGenerators.Reflection:
                ops.append(BOP_REFLECT)
            elif g == SymmetryGenerators.Inversion:
                perm = inversion_permutation(lattice)
                if np.array_equal(perm, np.arange(ns)[::-1]):
                    ops.append(BOP_REFLECT)      # bit i -> bit perm[i] = ns-1-i
                else:
                    # any other inversion: the permutation itself, on bits
                    # (inversion.py:53-86 moves bit i to bit perm[i])
                    if np.any(perm[perm] != np.arange(ns)):
                        raise ValueError("the inversion permutation is not an involution")
                    inv_perm = perm
                    ops.append(BOP_PERM)
            elif g == SymmetryGenerators.ParityY:
                ops.append(BOP_FLIP_Y)
            else:
                ops.append(BOP_FLIP)
            chi = chi * (sec ** 1)                    # base.py:556
        base_ops.append(ops)
        base_chi.append(chi)
    ntx = lx if SymmetryGenerators.Translation_x in trans else 1
    nty = ly if SymmetryGenerators.Translation_y in trans else 1
    kx = trans.get(SymmetryGenerators.Translation_x, 0)
    ky = trans.get(SymmetryGenerators.Translation_y, 0)
    if lattice is None:
        row_bits, cell_bits = ns, 1
    else:
        row_bits, cell_bits = nb * lx, nb
        if ns % row_bits != 0:
            raise ValueError("lattice size does not match ns")
    # characters in reference order g = (a*nty + b)*nbase + B, i.e. element
    # base_elem + (Tx^a, Ty^b) with character built in tuple order (:1420-1442)
    chis = []
    for a in range(ntx):
        for b in range(nty):
            for B in range(len(base)):
                ch = 1.0
                ch = ch * base_chi[B] if base[B] else ch
                if a != 0:
                    ch = ch * np.exp(2j * np.pi * kx * a / lx)   # translation.py:519
                if b != 0:
                    ch = ch * np.exp(2j * np.pi * ky * b / ly)
                chis.append(np.complex128(ch))
    chi = np.asarray(chis, dtype=np.complex128)
    grp = SectorGroup(ns=ns, n_up=n_up, z2=z2, generators=gens, nbase=len(base), base_ops=base_ops, ntx=ntx,
                      nty=nty, row_bits=row_bits, cell_bits=cell_bits, chi=chi)
    grp.matvec_error = matvec_error
    grp.perm = inv_perm
    if twisted:
        grp.cross_masks, grp.bphase = _twist_tables(lattice, ns, ntx, nty)
    grp.g_to_pidx, grp.phase_table = phases_map(chi)
    return grp


def phases_map(chi: np.ndarray):
    """build_phases_map, symmetry_container.py:1767-1804."""
    uniq: List[complex] = []
    g_to_pidx = np.zeros(len(chi), dtype=np.uint8)
    for g, ch in enumerate(chi):
        found = None
        for p, e in enumerate(uniq):
            if abs(ch - e) < SYM_NORM_THRESHOLD:
                found = p
                break
        if found is None:
            if len(uniq) >= 255:
                raise MemoryError("Exceeded maximum number of unique phases (255) for phase indexing.")
            found = len(uniq)
            uniq.append(ch)
        g_to_pidx[g] = found
    return g_to_pidx, np.asarray(uniq, dtype=np.complex128)
