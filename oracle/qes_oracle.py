"""CPU oracle for the symmetric-sector spin-1/2 matvec and Lanczos ground state.

TEST INFRASTRUCTURE ONLY.  This module is a plain NumPy restatement of the
reference algorithm (makskliczkowski/pyqusolver, package ``QES``).  Only the
test-suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker.  The product path in
``pyqusolver_amd`` never calls into it.

Every function cites the reference file:line it restates (paths relative to
the reference root, under ``Python/QES/``).

Parity pins (what checks each function independently of this file):

* ``translation_perm`` / ``reflection_perm`` / ``Group``: the reference's own
  orbit known-answers (tests/algebra/test_momentum_sectors.py:38-80) and the
  closed-form sector dimension (character formula, ``sector_dimension``)
  computed from cycle structures, independent of the orbit scan.
* ``scan_basis``: closed-form sector dimensions; normalisations against the
  orbit-size closed form N_r^2 = |Stab(r)| for real sectors; brute force of
  the projector P = (1/|G|) sum_g conj(chi_g) g on tiny systems.
* ``vm_apply``: the reference kernel known-answers
  (tests/algebra/test_custom_hamiltonian_workflows.py:76-126) and a brute
  force against Kronecker products of 2x2 spin matrices.
* ``push_matvec``: brute force U^dagger H U with the symmetric-basis rotation
  matrix (Algebra/Hilbert/matrix_builder.py:1650-1776), hermiticity, and the
  union-of-sectors spectrum against the full-space spectrum.
* model term builders: exact closed forms (ferromagnet E0 = -n_bonds*J/4, the
  N=4 Heisenberg ring E0 = -2, examples/algebra/example_exact_diagonalization.py:87-97)
  and an independent full-space sparse eigensolve.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

# Algebra/Symmetries/base.py:159-163
OP_IDENTITY, OP_TRANSLATION, OP_REFLECTION, OP_PARITY, OP_INVERSION = 0, 1, 2, 3, 4
# Algebra/Symmetries/jit/symmetry_container_jit.py:53-55
INVALID_REPR_IDX = np.uint32(0xFFFFFFFF)
INVALID_PHASE_IDX = np.uint8(0xFF)
SYM_NORM_THRESHOLD = 1e-7
# Algebra/Operator/impl/operators_spin.py:1084-1118 (lookup codes)
SIG_X_L, SIG_Y_L, SIG_Z_L, SIG_P_L, SIG_M_L = 11, 12, 13, 14, 15
SIG_X_C, SIG_Y_C, SIG_Z_C, SIG_P_C, SIG_M_C = 20, 21, 22, 23, 24
SIG_XY_C, SIG_YX_C, SIG_XZ_C, SIG_ZX_C, SIG_YZ_C, SIG_ZY_C = 25, 26, 27, 28, 29, 30
SIG_PM_C, SIG_MP_C = 31, 32
SPIN = 0.5  # BACKEND_REPR spin value pinned by test_custom_hamiltonian_workflows.py:91-92


# ---------------------------------------------------------------------------
# Lattices (the reference keeps them in the absent general_python submodule;
# these follow the conventions its symmetry code relies on:
# translation.py:369-436 and tests/algebra/test_momentum_sectors.py:5-35)
# ---------------------------------------------------------------------------

@dataclass
class OLattice:
    kind: str          # "chain" | "square" | "honeycomb"
    lx: int
    ly: int = 1
    lz: int = 1

    @property
    def dim(self) -> int:
        return 1 if self.kind == "chain" else 2

    @property
    def n_basis(self) -> int:
        return 2 if self.kind == "honeycomb" else 1

    @property
    def ns(self) -> int:
        return self.lx * self.ly * self.lz * self.n_basis

    # Bravais coordinates, site = x + lx*y (test_momentum_sectors.py:24-31)
    def get_coordinates(self, site: int) -> Tuple[int, int, int]:
        return (site % self.lx, (site // self.lx) % self.ly, site // (self.lx * self.ly))

    def site_index(self, x: int, y: int, z: int) -> int:
        return x + self.lx * y + self.lx * self.ly * z

    def cell_site(self, x: int, y: int, sub: int) -> int:
        return self.n_basis * ((x % self.lx) + self.lx * (y % self.ly)) + sub

    # Geometry used by the inversion permutation (inversion.py:135-230 reads
    # lat.rvectors, a1/a2/a3, basis, fracs, subs).  Site s = n_basis*cell +
    # sub; honeycomb: a1 = (sqrt3, 0), a2 = (sqrt3/2, 3/2), B at (a1+a2)/3,
    # so the three forward bonds of oracle_bonds have unit length.
    def geometry(self):
        if self.kind == "honeycomb":
            a1 = np.array([np.sqrt(3.0), 0.0, 0.0])
            a2 = np.array([np.sqrt(3.0) / 2, 1.5, 0.0])
            basis = np.array([np.zeros(3), (a1 + a2) / 3.0])
        else:
            a1, a2 = np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0])
            basis = np.zeros((1, 3))
        a3 = np.array([0.0, 0.0, 1.0])
        nb = self.n_basis
        fracs, subs, r = [], [], []
        for s in range(self.ns):
            cell, sub = s // nb, s % nb
            x, y = cell % self.lx, (cell // self.lx) % self.ly
            fracs.append((x, y, 0))
            subs.append(sub)
            r.append(x * a1 + y * a2 + basis[sub])
        return a1, a2, a3, basis, np.array(fracs), np.array(subs), np.array(r)


def oracle_bonds(lat: OLattice, order: int = 1) -> List[Tuple[int, int, int]]:
    """Forward bond list ``(i, j, nn_index)`` in site order.

    Stands in for ``get_nn_forward``/``get_nnn_forward`` of the absent lattice
    module (used at Model/Interacting/Spin/j1j2.py:241-290, xxz.py:400-405,
    heisenberg_kitaev.py:533-541).  Chain: i->i+1 (NN), i->i+2 (NNN).
    Square: (x+1,y), (x,y+1) (NN); (x+1,y+1), (x-1,y+1) (NNN).
    Honeycomb (A=sub 0, B=sub 1): A(x,y) -> B(x-1,y) [nn 0, x bond],
    B(x,y-1) [nn 1, y bond], B(x,y) [nn 2, z bond]; B sites have no forward
    bonds.
    """
    out = []
    if lat.kind == "chain":
        L = lat.lx
        for i in range(L):
            j = (i + order) % L
            if j != i:
                out.append((i, j, 0))
    elif lat.kind == "square":
        for i in range(lat.ns):
            x, y, _ = lat.get_coordinates(i)
            nbrs = [(x + 1, y), (x, y + 1)] if order == 1 else [(x + 1, y + 1), (x - 1, y + 1)]
            for n, (xx, yy) in enumerate(nbrs):
                j = lat.site_index(xx % lat.lx, yy % lat.ly, 0)
                if j != i:
                    out.append((i, j, n))
    elif lat.kind == "honeycomb":
        if order != 1:
            raise ValueError("honeycomb: only nearest-neighbour bonds")
        for c in range(lat.lx * lat.ly):
            x, y = c % lat.lx, c // lat.lx
            i = lat.cell_site(x, y, 0)
            for n, (xx, yy) in enumerate([(x - 1, y), (x, y - 1), (x, y)]):
                out.append((i, lat.cell_site(xx, yy, 1), n))
    else:
        raise ValueError(lat.kind)
    return out


# ---------------------------------------------------------------------------
# Symmetry generators and the group
# ---------------------------------------------------------------------------

def translation_perm(lat: OLattice, axis: int, shift: int = 1):
    """Algebra/Symmetries/translation.py:316-437 (perm[src] -> dest, crossing mask)."""
    ns = lat.ns
    perm = np.empty(ns, dtype=np.int64)
    cross = np.zeros(ns, dtype=bool)
    dims = (lat.lx, lat.ly if lat.dim > 1 else 1, lat.lz if lat.dim > 2 else 1)
    size_axis = dims[axis]
    nb = lat.n_basis
    for site in range(ns):
        if nb > 1:  # non-Bravais branch :367-408
            cell, sub = site // nb, site % nb
            c = [cell % lat.lx, (cell // lat.lx) % lat.ly if lat.dim >= 2 else 0, 0]
        else:       # Bravais branch :410-435
            cell, sub = site, 0
            c = list(lat.get_coordinates(site))
        c[axis] += shift
        wrap = False
        if c[axis] >= size_axis:
            c[axis] -= size_axis
            wrap = True
        elif c[axis] < 0:
            c[axis] += size_axis
            wrap = True
        if nb > 1:
            perm[site] = (c[0] + c[1] * lat.lx + c[2] * lat.lx * lat.ly) * nb + sub
        else:
            perm[site] = lat.site_index(c[0], c[1], c[2])
        cross[site] = wrap
    return perm, cross


def reflection_perm(ns: int) -> np.ndarray:
    """Algebra/Symmetries/reflection.py:143 -- full bit reversal."""
    return np.arange(ns)[::-1].copy()


def inversion_perm(lat: OLattice, tol: float = 1e-7) -> np.ndarray:
    """Algebra/Symmetries/inversion.py:135-230 (``_build_inversion_permutation``,
    center "bbox"): perm[i] = site whose position is 2 r0 - r_i, r0 the centre
    of the bounding box of all positions, matched over sublattices by
    rounding the fractional Bravais coordinates and wrapping periodic
    directions."""
    a1, a2, a3, basis, fracs, subs, r = lat.geometry()
    ns = lat.ns
    r0 = 0.5 * (r.min(axis=0) + r.max(axis=0))
    Ainv = np.linalg.inv(np.column_stack([a1, a2, a3]))
    dims = (lat.lx, max(lat.ly, 1), max(lat.lz, 1))
    lookup = {(int(f[0]), int(f[1]), int(f[2]), int(sb)): i for i, (f, sb) in enumerate(zip(fracs, subs))}
    perm = np.full(ns, -1, dtype=np.int64)
    for i in range(ns):
        target = 2.0 * r0 - r[i]
        best = None
        for sb in range(len(basis)):
            fr = Ainv @ (target - basis[sb])
            n = [int(np.rint(fr[0])) % dims[0], int(np.rint(fr[1])) % dims[1], int(np.rint(fr[2])) % dims[2]]
            err = np.linalg.norm(n[0] * a1 + n[1] * a2 + n[2] * a3 + basis[sb] - target)
            if best is None or err < best[0]:
                best = (err, n[0], n[1], n[2], sb)
        err, nx, ny, nz, sb = best
        if err > tol:
            raise ValueError(f"inversion target of site {i} matches no site (err {err})")
        perm[i] = lookup[(nx, ny, nz, sb)]
    if len(np.unique(perm)) != ns:
        raise ValueError("inversion permutation is not a bijection")
    return perm


@dataclass
class Generator:
    kind: str      # "tx" | "ty" | "tz" | "R" | "I" | "Px" | "Py" | "Pz"
    sector: object
    # "I" only: an explicit site permutation (InversionSymmetry's op.perm,
    # inversion.py:135-230) in place of the lattice's bounding-box inversion
    perm: Optional[np.ndarray] = None


@dataclass
class Group:
    """Restatement of SymmetryContainer.build_group / build_compiled_group.

    Algebra/Symmetries/symmetry_container.py:866-1174 and :1399-1442.
    ``elements[g]`` is a list of primitive ops ``(code, table_idx, power)``
    applied in sequence (symmetry_container_jit.py:89-116); ``chi[g]`` is the
    character in the chosen sector.
    """
    lat: OLattice
    generators: List[Generator]
    n_up: Optional[int] = None          # U(1) sector (popcount), None if absent
    z2: Optional[int] = None            # Z2 popcount-parity global, None if absent
    elements: list = field(default_factory=list)
    chi: np.ndarray = None
    trans_perm: list = field(default_factory=list)
    refl_perm: list = field(default_factory=list)
    inv_perm: list = field(default_factory=list)
    parity_axis: list = field(default_factory=list)
    # lattice.boundary_phases(): (3, ns + 1) phase per number of up spins
    # crossing the boundary in one translation step, per direction; None: 1
    bound_phase: Optional[np.ndarray] = None
    trans_cross: list = field(default_factory=list)
    trans_dir: list = field(default_factory=list)

    def __post_init__(self):
        ns = self.lat.ns
        gens = []
        # Global-symmetry filter, parity.py:283-321: Pz is redundant with U(1);
        # Px/Py need half filling (even ns).
        for g in self.generators:
            if g.kind in ("Px", "Py", "Pz") and self.n_up is not None:
                if g.kind == "Pz":
                    continue
                if int(self.n_up) != ns // 2 or ns % 2 != 0:
                    continue
            gens.append(g)
        self.generators = gens
        # build_group :898-959
        trans = {}
        others = []
        for g in gens:
            if g.kind in ("tx", "ty", "tz"):
                trans[g.kind[1]] = g
            else:
                others.append(g)
        base = [()]
        for r in range(1, len(others) + 1):
            for combo in itertools.combinations(range(len(others)), r):
                base.append(tuple(others[i] for i in combo))
        periods = {"x": self.lat.lx, "y": self.lat.ly, "z": self.lat.lz}
        dirs = sorted(trans.keys())
        if dirs:
            telems = []
            for powers in itertools.product(*(range(periods[d]) for d in dirs)):
                telems.append(tuple((trans[d], p) for d, p in zip(dirs, powers) if p != 0))
            full = [b + t for t in telems for b in base]
        else:
            full = base
        # tables, build_compiled_group :981-1045
        tidx, ridx, pidx, iidx = {}, {}, {}, {}
        for g in gens:
            if g.kind in ("tx", "ty", "tz"):
                tidx[id(g)] = len(self.trans_perm)
                perm, cross = translation_perm(self.lat, "xyz".index(g.kind[1]))
                self.trans_perm.append(perm)
                self.trans_cross.append(cross)
                self.trans_dir.append("xyz".index(g.kind[1]))
            elif g.kind == "R":
                ridx[id(g)] = len(self.refl_perm)
                self.refl_perm.append(reflection_perm(ns))
            elif g.kind == "I":
                iidx[id(g)] = len(self.inv_perm)
                self.inv_perm.append(np.asarray(g.perm, dtype=np.int64) if g.perm is not None
                                     else inversion_perm(self.lat))
            else:
                pidx[id(g)] = len(self.parity_axis)
                self.parity_axis.append("xyz".index(g.kind[1]))
        self.elements = []
        chis = []
        for elem in full:
            ops = []
            chi = 1.0
            for entry in elem:
                if isinstance(entry, tuple):
                    g, power = entry
                    count = power
                else:
                    g, power, count = entry, 0, 1
                if g.kind in ("tx", "ty", "tz"):
                    ops.append((OP_TRANSLATION, tidx[id(g)], power))
                    L = periods[g.kind[1]]
                    chi = chi * np.exp(2j * np.pi * g.sector * count / L)  # translation.py:519
                elif g.kind == "R":
                    ops.append((OP_REFLECTION, ridx[id(g)], power))
                    chi = chi * (g.sector ** count)                        # base.py:556
                elif g.kind == "I":
                    ops.append((OP_INVERSION, iidx[id(g)], power))
                    chi = chi * (g.sector ** (count % 2))                  # inversion.py:500-523
                else:
                    ops.append((OP_PARITY, pidx[id(g)], power))
                    chi = chi * (g.sector ** count)
            self.elements.append(ops)
            chis.append(np.complex128(chi))
        self.chi = np.asarray(chis, dtype=np.complex128)

    @property
    def n_group(self) -> int:
        return len(self.elements)

    def apply(self, states: np.ndarray, g: int):
        """symmetry_container_jit.py:63-118 on an int64 array of states.

        Returns (new_states, intrinsic_phase): every translation step
        multiplies bound_phase[direction, occ] with occ the number of up spins
        (set bits, bit = site) on the sites that cross the boundary
        (symmetry_container_jit.py:91-98, translation.py:80-89); 1 without flux.
        """
        ns = self.lat.ns
        cur = states.astype(np.int64, copy=True)
        phase = np.ones(cur.shape, dtype=np.complex128)
        for code, a0, a1 in self.elements[g]:
            if code == OP_TRANSLATION:
                cmask = 0
                for site in np.nonzero(self.trans_cross[a0])[0]:
                    cmask |= 1 << int(site)
                for _ in range(a1):
                    if self.bound_phase is not None:
                        occ = _popcount(cur & np.int64(cmask))
                        phase = phase * self.bound_phase[self.trans_dir[a0]][occ]
                    cur = _apply_perm(cur, ns, self.trans_perm[a0])
            elif code == OP_REFLECTION:
                cur = _apply_perm(cur, ns, self.refl_perm[a0])
            elif code == OP_INVERSION:    # _apply_inversion_permutation, inversion.py:53-86 (base 2)
                cur = _apply_perm(cur, ns, self.inv_perm[a0])
            elif code == OP_PARITY:
                cur, ph = _apply_parity(cur, ns, self.parity_axis[a0])
                phase = phase * ph
        return cur, phase


def _apply_perm(states: np.ndarray, ns: int, perm: np.ndarray) -> np.ndarray:
    """reflection.py:35-44 / translation.py:50-89 (bit src -> bit perm[src])."""
    out = np.zeros_like(states)
    for src in range(ns):
        out |= ((states >> src) & 1) << int(perm[src])
    return out


def _popcount(states: np.ndarray) -> np.ndarray:
    v = states.astype(np.uint64)
    c = np.zeros(v.shape, dtype=np.int64)
    for b in range(64):
        c += ((v >> np.uint64(b)) & np.uint64(1)).astype(np.int64)
    return c


def _apply_parity(states: np.ndarray, ns: int, axis: int):
    """parity.py:42-66 -- every axis complements all ns bits; y adds +-i."""
    mask = (1 << ns) - 1
    new = (~states) & mask
    if axis in (0, 2):
        return new, np.ones(states.shape, dtype=np.complex128)
    ups = ns - _popcount(states)
    return new, np.where((ups & 1) == 0, 1j, -1j).astype(np.complex128)


# ---------------------------------------------------------------------------
# Basis scan (K1-K4)
# ---------------------------------------------------------------------------

def candidate_states(ns: int, n_up: Optional[int], z2: Optional[int] = None) -> np.ndarray:
    """All states passing the global filters, ascending (globals.py:314-343)."""
    if n_up is None:
        s = np.arange(1 << ns, dtype=np.int64)
    else:
        if not 0 <= n_up <= ns:
            return np.zeros(0, dtype=np.int64)
        out = []
        for combo in itertools.combinations(range(ns), n_up):
            v = 0
            for b in combo:
                v |= 1 << b
            out.append(v)
        s = np.sort(np.asarray(out, dtype=np.int64))
    if z2 is not None:
        par = _popcount(s) & 1
        s = s[np.where(par == 0, 1, -1) == int(z2)]
    return s


def phases_map(group: Group):
    """symmetry_container.py:1767-1804 (unique characters within 1e-7)."""
    uniq: List[complex] = []
    g_to_pidx = np.zeros(group.n_group, dtype=np.uint8)
    for g, ch in enumerate(group.chi):
        found = None
        for p, e in enumerate(uniq):
            if abs(ch - e) < SYM_NORM_THRESHOLD:
                found = p
                break
        if found is None:
            if len(uniq) >= 255:
                raise MemoryError("more than 255 distinct phases")
            found = len(uniq)
            uniq.append(ch)
        g_to_pidx[g] = found
    return g_to_pidx, np.asarray(uniq, dtype=np.complex128)


@dataclass
class Basis:
    reps: np.ndarray           # int64, sorted
    norms: np.ndarray          # float64
    states: np.ndarray         # scanned states (ascending)
    state_rep: np.ndarray      # repr_map restricted to ``states``: rep index or -1
    state_pidx: np.ndarray     # phase_idx restricted to ``states``
    state_gmin: np.ndarray     # first minimising group element
    phase_table: np.ndarray
    g_to_pidx: np.ndarray

    def lookup_g(self, s: np.ndarray) -> np.ndarray:
        """First minimising group element of states of the sector."""
        pos = np.searchsorted(self.states, s)
        pos = np.minimum(pos, len(self.states) - 1)
        return self.state_gmin[pos]

    def lookup(self, s: np.ndarray):
        """repr_map/phase_idx lookup for arbitrary states (-1 if not in sector)."""
        pos = np.searchsorted(self.states, s)
        pos = np.minimum(pos, len(self.states) - 1) if len(self.states) else pos
        ok = (len(self.states) > 0) & (self.states[pos] == s) if len(self.states) else np.zeros(s.shape, bool)
        idx = np.where(ok, self.state_rep[pos], -1)
        pidx = np.where(ok, self.state_pidx[pos], INVALID_PHASE_IDX)
        return idx, pidx


def scan_basis(group: Group, chunk: int = 1 << 18) -> Basis:
    """generate_symmetric_basis, symmetry_container.py:1575-1765, with
    scan_chunk_find_representatives (jit :204-255), fill_representatives
    (:193-199), compute_normalization (:154-187) and the state->index
    conversion (symmetry_container.py:184-195)."""
    ns = group.lat.ns
    states = candidate_states(ns, group.n_up, group.z2)
    G = group.n_group
    chi_c = np.conj(group.chi)
    min_state = np.empty_like(states)
    min_g = np.empty(states.shape, dtype=np.int64)
    valid = np.zeros(states.shape, dtype=bool)
    for lo in range(0, len(states), chunk):
        s = states[lo:lo + chunk]
        best = np.full(s.shape, np.iinfo(np.int64).max, dtype=np.int64)
        bg = np.full(s.shape, -1, dtype=np.int64)
        proj = np.zeros(s.shape, dtype=np.complex128)
        for g in range(G):
            t, ph = group.apply(s, g)
            lt = t < best
            best = np.where(lt, t, best)
            bg = np.where(lt, g, bg)
            eq = t == s
            proj = np.where(eq, proj + chi_c[g] * ph, proj)
        min_state[lo:lo + chunk] = best
        min_g[lo:lo + chunk] = bg
        valid[lo:lo + chunk] = np.abs(proj) > SYM_NORM_THRESHOLD
    is_rep = valid & (min_state == states) & (min_g != -1)
    reps = states[is_rep].copy()
    norms = normalization(group, reps)
    g_to_pidx, phase_table = phases_map(group)
    rep_idx = np.searchsorted(reps, min_state)
    rep_idx = np.where(valid, rep_idx, -1)
    if len(reps):
        ok = valid & (reps[np.minimum(rep_idx, len(reps) - 1)] == min_state)
        rep_idx = np.where(ok, rep_idx, -1)
    pidx = np.where(valid, g_to_pidx[np.maximum(min_g, 0)], INVALID_PHASE_IDX).astype(np.uint8)
    return Basis(reps, norms, states, rep_idx.astype(np.int64), pidx, min_g, phase_table, g_to_pidx)


def normalization(group: Group, reps: np.ndarray) -> np.ndarray:
    """jit/symmetry_container_jit.py:154-174: sqrt(|sum_{g r = r} conj chi_g phase_g|)."""
    proj = np.zeros(reps.shape, dtype=np.complex128)
    chi_c = np.conj(group.chi)
    for g in range(group.n_group):
        t, ph = group.apply(reps, g)
        proj = np.where(t == reps, proj + chi_c[g] * ph, proj)
    val = np.abs(proj)
    return np.where(val < SYM_NORM_THRESHOLD, 0.0, np.sqrt(val))


def sector_dimension(group: Group) -> int:
    """Closed form N = (1/|G|) sum_g conj(chi_g) Tr_{U(1)/Z2}(g) from cycle
    structures -- independent of the orbit scan.  Only for groups without
    intrinsic phases (parity x/z)."""
    ns = group.lat.ns
    total = 0.0 + 0.0j
    for g in range(group.n_group):
        # compose the element's site map and flip flag
        perm = np.arange(ns)
        flip = False
        for code, a0, a1 in group.elements[g]:
            if code == OP_TRANSLATION:
                for _ in range(a1):
                    perm = group.trans_perm[a0][perm]
            elif code == OP_REFLECTION:
                perm = group.refl_perm[a0][perm]
            elif code == OP_INVERSION:
                perm = group.inv_perm[a0][perm]
            elif code == OP_PARITY:
                if group.parity_axis[a0] == 1:
                    raise ValueError("closed form excludes parity y")
                flip = not flip
        # cycles of perm; count fixed states via generating polynomial in #ups
        seen = np.zeros(ns, bool)
        poly = np.zeros(ns + 1)
        poly[0] = 1.0
        for i in range(ns):
            if seen[i]:
                continue
            c = 0
            j = i
            while not seen[j]:
                seen[j] = True
                j = perm[j]
                c += 1
            new = np.zeros(ns + 1)
            if not flip:      # cycle all 0 or all 1
                new[:] += poly
                new[c:] += poly[:ns + 1 - c]
            else:             # alternating, needs even length, c/2 ups, 2 ways
                if c % 2 == 1:
                    new[:] = 0
                else:
                    new[c // 2:] += 2 * poly[:ns + 1 - c // 2]
            poly = new
        # fixed states in the U(1) and/or Z2 popcount-parity sector
        # (z2 = +1: even popcount, as candidate_states)
        want = None if group.z2 is None else (0 if int(group.z2) == 1 else 1)
        if group.n_up is not None:
            tr = poly[group.n_up] if (want is None or group.n_up % 2 == want) else 0.0
        elif want is not None:
            tr = poly[want::2].sum()
        else:
            tr = poly.sum()
        total += np.conj(group.chi[g]) * tr
    return int(round((total / group.n_group).real))


# ---------------------------------------------------------------------------
# Spin-1/2 instruction VM (K6)
# ---------------------------------------------------------------------------

def _site_bit(states, ns, site):
    return (states >> (ns - 1 - site)) & 1


def _flip(states, ns, site):
    return states ^ (1 << (ns - 1 - site))


def _single(states, ns, code, s1, s2):
    """operators_spin.py:2165-2313 (one instruction on an array of states).

    Returns (new_states, complex value, is_diagonal)."""
    one = np.ones(states.shape, dtype=np.complex128)

    def op(kind, st, site, spin=SPIN):
        b = _site_bit(st, ns, site)
        if kind == "x":      # _sigma_x_core :147-157
            return _flip(st, ns, site), one * spin
        if kind == "y":      # _sigma_y_core :269-281
            return _flip(st, ns, site), 1j * (2.0 * b - 1.0) * spin
        if kind == "z":      # _sigma_z_core :439-447
            return st, (2.0 * b - 1.0) * spin + 0j
        if kind == "p":      # sigma_plus_int_np :627-645 (annihilates set bit)
            return np.where(b == 1, st, _flip(st, ns, site)), np.where(b == 1, 0.0, spin) + 0j
        if kind == "m":      # sigma_minus_int_np :698-737
            return np.where(b == 0, st, _flip(st, ns, site)), np.where(b == 0, 0.0, spin) + 0j
        raise ValueError(kind)

    def seq(kinds, sites):
        st, c = states, one
        alive = np.ones(states.shape, dtype=bool)
        for k, site in zip(kinds, sites):
            st2, c2 = op(k, st, site)
            if k in "pm":     # ladder ops stop at the first annihilation
                dead = alive & (c2 == 0)
                st = np.where(alive & ~dead, st2, st)
                c = np.where(alive, c * c2, c)
                alive = alive & ~dead
            else:
                st = np.where(alive, st2, st)
                c = np.where(alive, c * c2, c)
        return st, c

    if code == SIG_X_L:
        return (*op("x", states, s1), False)
    if code == SIG_Y_L:
        return (*op("y", states, s1), False)
    if code == SIG_Z_L:
        return (*op("z", states, s1), True)
    if code == SIG_P_L:
        return (*seq("p", [s1]), False)
    if code == SIG_M_L:
        return (*seq("m", [s1]), False)
    if code == SIG_X_C:
        return (*seq("xx", [s1, s2]), False)
    if code == SIG_Y_C:
        return (*seq("yy", [s1, s2]), False)
    if code == SIG_Z_C:
        return (*seq("zz", [s1, s2]), True)
    if code == SIG_P_C:
        return (*seq("pp", [s1, s2]), False)
    if code == SIG_M_C:
        return (*seq("mm", [s1, s2]), False)
    mixed = {SIG_XY_C: "xy", SIG_YX_C: "yx", SIG_XZ_C: "xz", SIG_ZX_C: "zx",
             SIG_YZ_C: "yz", SIG_ZY_C: "zy"}
    if code in mixed:     # create_sigma_mixed_int_core :1722-1759
        k1, k2 = mixed[code]
        st1, c1 = op(k1, states, s1)
        st2, c2 = op(k2, st1, s2)
        return np.where(c1 == 0, st1, st2), np.where(c1 == 0, 0j, c1 * c2), False
    if code == SIG_PM_C:  # sigma_pm_int_np :812-852 (+ on even index)
        return (*seq("pm", [s1, s2]), False)
    if code == SIG_MP_C:  # sigma_mp_int_np :939-962
        return (*seq("mp", [s1, s2]), False)
    raise ValueError(f"unsupported code {code}")


def vm_apply(states: np.ndarray, ns: int, codes, sites, coeffs, is_complex: bool):
    """sigma_operator_composition_int, operators_spin.py:2318-2352.

    Vectorised over ``states``; returns a list of (new_states, values, keep)
    per output slot in the reference's order: off-diagonal outputs in term
    order, then the summed diagonal.  ``keep`` masks the entries the reference
    emits (|value| >= 1e-15)."""
    states = np.asarray(states, dtype=np.int64)
    dt = np.complex128 if is_complex else np.float64
    outs = []
    diag = np.zeros(states.shape, dtype=dt)
    for k in range(len(codes)):
        st, c, is_diag = _single(states, ns, int(codes[k]), int(sites[k][0]),
                                 int(sites[k][1]) if len(sites[k]) > 1 else -1)
        c = c if is_complex else c.real  # dtype(c_s.real) in the real build
        val = (c * coeffs[k]).astype(dt) if is_complex else (c * np.real(coeffs[k])).astype(dt)
        keep = np.abs(val) >= 1e-15
        if is_diag:
            diag = np.where(keep, diag + val, diag)
        else:
            outs.append((st, val, keep))
    outs.append((states.copy(), diag, np.abs(diag) > 1e-15))
    return outs


def push_matvec(basis: Basis, ns: int, codes, sites, coeffs, is_complex: bool, x: np.ndarray,
                group: Optional[Group] = None) -> np.ndarray:
    """_apply_op_batch_compact_seq_jit, Algebra/Hilbert/matrix_builder.py:1347-1377.

    y[idx] += val * conj(phase_table[phase_idx[s']]) * N_idx / N_k * x[k]
    (output overwritten, as SpecialOperator.matvec_fun allocates zeros,
    special_operator.py:656-663).  With ``group`` each element also carries
    the intrinsic phase ph of the first minimising element g* on s' (the
    basis vectors of expand_to_full_state_jit carry it, jit :261-315, so
    U^dagger H U has val * conj(chi(g*)) * ph_{g*}(s') * N_idx / N_k; the
    reference's kernel leaves ph out -- the same where no element has a
    state-dependent phase)."""
    x2 = x[:, None] if x.ndim == 1 else x
    dt = np.result_type(x2.dtype, np.complex128)
    y = np.zeros(x2.shape, dtype=dt)
    reps, norms = basis.reps, basis.norms
    inv_nk = np.where(norms > 0, 1.0 / np.where(norms > 0, norms, 1.0), 0.0)
    for st, val, keep in vm_apply(reps, ns, codes, sites, coeffs, is_complex):
        idx, pidx = basis.lookup(st)
        m = keep & (idx >= 0)
        k = np.nonzero(m)[0]
        i = idx[m]
        phase = basis.phase_table[pidx[m]]
        factor = val[m] * np.conj(phase) * norms[i] * inv_nk[k]
        if group is not None:
            gmin = basis.lookup_g(st[m])
            ph = np.ones(len(k), dtype=np.complex128)
            for g in np.unique(gmin):
                sel = gmin == g
                ph[sel] = group.apply(st[m][sel], int(g))[1]
            factor = factor * ph
        np.add.at(y, i, factor[:, None] * x2[k])
    if not np.iscomplexobj(x):
        if np.abs(y.imag).max(initial=0.0) > 1e-12 * max(1.0, np.abs(y).max(initial=0.0)):
            raise ValueError("complex matrix elements with a real vector")
        y = y.real
    return y[:, 0] if x.ndim == 1 else y


def reduced_dense(basis: Basis, ns, codes, sites, coeffs, is_complex, group: Optional[Group] = None) -> np.ndarray:
    """Dense reduced H from the push-form matvec on unit vectors."""
    n = len(basis.reps)
    return push_matvec(basis, ns, codes, sites, coeffs, is_complex, np.eye(n, dtype=np.complex128), group=group)


# ---------------------------------------------------------------------------
# Brute-force pins (no symmetry code involved)
# ---------------------------------------------------------------------------

_S = {
    "x": np.array([[0, 1], [1, 0]], dtype=np.complex128) * SPIN,
    # <1|Sy|0> = -i/2 and <0|Sy|1> = +i/2 (operators_spin.py:278, bit 1 = up)
    "y": np.array([[0, 1j], [-1j, 0]], dtype=np.complex128) * SPIN,
    "z": np.array([[-1, 0], [0, 1]], dtype=np.complex128) * SPIN,   # basis (bit 0, bit 1)
    "p": np.array([[0, 0], [1, 0]], dtype=np.complex128) * SPIN,    # |0> -> |1> (bit set)
    "m": np.array([[0, 1], [0, 0]], dtype=np.complex128) * SPIN,
}
_CODE_KINDS = {SIG_X_L: "x", SIG_Y_L: "y", SIG_Z_L: "z", SIG_P_L: "p", SIG_M_L: "m",
               SIG_X_C: "xx", SIG_Y_C: "yy", SIG_Z_C: "zz", SIG_P_C: "pp", SIG_M_C: "mm",
               SIG_XY_C: "xy", SIG_YX_C: "yx", SIG_XZ_C: "xz", SIG_ZX_C: "zx",
               SIG_YZ_C: "yz", SIG_ZY_C: "zy", SIG_PM_C: "pm", SIG_MP_C: "mp"}


def full_dense_hamiltonian(ns: int, codes, sites, coeffs, is_complex: bool) -> np.ndarray:
    """Full-space dense H from Kronecker products of 2x2 spin matrices.

    Basis index = integer state; site i is bit ns-1-i (operators_spin.py:153)
    so site 0 is the leftmost Kronecker factor.  For two-site operators the
    first site's matrix acts first (applied right-most in the product), which
    only matters for ladder pairs on equal sites (never used)."""
    dim = 1 << ns
    H = np.zeros((dim, dim), dtype=np.complex128)
    for k in range(len(codes)):
        kinds = _CODE_KINDS[int(codes[k])]
        s = [int(v) for v in sites[k] if int(v) >= 0][:len(kinds)]
        mats = [np.eye(2, dtype=np.complex128)] * ns
        mats = list(mats)
        for kind, site in zip(kinds, s):
            mats[site] = _S[kind] @ mats[site]
        M = np.ones((1, 1), dtype=np.complex128)
        for m in mats:
            M = np.kron(M, m)
        c = coeffs[k]
        if not is_complex:
            # the real build keeps Re of each per-term operator value
            M = M.real.astype(np.complex128)
            c = np.real(c)
        H += c * M
    return H


def rotation_matrix(group: Group, basis: Basis) -> np.ndarray:
    """Symmetric basis vectors as columns, following
    expand_to_full_state_jit (jit/symmetry_container_jit.py:261-315):
    |k> = sum_g conj(chi_g) phase_g g|r_k> / (N_k sqrt|G|)."""
    ns = group.lat.ns
    U = np.zeros((1 << ns, len(basis.reps)), dtype=np.complex128)
    G = group.n_group
    for g in range(G):
        t, ph = group.apply(basis.reps, g)
        w = np.conj(group.chi[g]) * ph / (basis.norms * np.sqrt(G))
        np.add.at(U, (t, np.arange(len(basis.reps))), w)
    return U


def projected_matvec(group_in: Group, basis_in: Basis, group_out: Group, basis_out: Basis, ns: int, codes, sites,
                     coeffs, is_complex: bool, x: np.ndarray) -> np.ndarray:
    """_apply_op_batch_projected_compact_seq_jit
    (Algebra/Symmetries/jit/matrix_builder_jit.py:51-108): for every input
    representative k and group element g, apply the operator to g r_k, then
    for every h with h(s') a representative of the output sector add
        val * conj(chi_in(g)) ph_g / N_k * conj(chi_out(h)) conj(ph_h) / (|G| N_out)
    (both sectors share the group elements; only the characters differ)."""
    G = group_in.n_group
    out = np.zeros(len(basis_out.reps), dtype=np.complex128)
    x = np.asarray(x, dtype=np.complex128)
    for g in range(G):
        s, ph_g = group_in.apply(basis_in.reps, g)
        w_g = np.conj(group_in.chi[g]) * ph_g / basis_in.norms
        for st, val, keep in vm_apply(s, ns, codes, sites, coeffs, is_complex):
            m = keep & (np.abs(val) >= 1e-15)
            k = np.nonzero(m)[0]
            if len(k) == 0:
                continue
            vw = val[m] * w_g[m]
            for h in range(G):
                so, ph_h = group_out.apply(st[m], h)
                pos = np.searchsorted(basis_out.reps, so)
                pos = np.minimum(pos, len(basis_out.reps) - 1)
                hit = (len(basis_out.reps) > 0) & (basis_out.reps[pos] == so)
                if not np.any(hit):
                    continue
                w_h = np.conj(group_out.chi[h]) * np.conj(ph_h[hit]) / (G * basis_out.norms[pos[hit]])
                np.add.at(out, pos[hit], vw[hit] * w_h * x[k[hit]])
    return out


def expand_state(group: Group, basis: Basis, x: np.ndarray) -> np.ndarray:
    """expand_to_full_state_jit (jit/symmetry_container_jit.py:261-315),
    push form over every group element."""
    ns = group.lat.ns
    out = np.zeros(1 << ns, dtype=np.complex128)
    G = group.n_group
    for g in range(G):
        t, ph = group.apply(basis.reps, g)
        w = np.conj(group.chi[g]) * ph / (basis.norms * np.sqrt(G))
        np.add.at(out, t, w * x)
    return out


def transform_to_reduced(basis: Basis, ns: int, psi: np.ndarray) -> np.ndarray:
    """transform_to_reduced_space (Algebra/Hilbert/hilbert_base.py:485-563):
    red[k] = sum over states s of the sector with repr_map[s] = k of
    conj(phase_table[phase_idx[s]]) psi[s], then divided by N_k."""
    states = np.arange(1 << ns, dtype=np.int64)
    idx, pidx = basis.lookup(states)
    ok = idx >= 0
    red = np.zeros(len(basis.reps), dtype=np.complex128)
    np.add.at(red, idx[ok], np.conj(basis.phase_table[pidx[ok]]) * psi[ok])
    return red / basis.norms


# ---------------------------------------------------------------------------
# Model term lists (instruction arrays as SpecialOperator.add builds them,
# special_operator.py:1113-1231 and _finalize_instruction_arrays :1392-1417)
# ---------------------------------------------------------------------------

class Terms:
    def __init__(self, ns, is_complex):
        self.ns = ns
        self.is_complex = is_complex
        self.codes: List[int] = []
        self.sites: List[List[int]] = []
        self.coeffs: List[complex] = []

    def add(self, code, mult, sites):
        if abs(mult) < 1e-15:          # special_operator.py:1156-1162
            return
        s = list(sites) + [-1] * (2 - len(sites))
        self.codes.append(code)
        self.sites.append(s)
        self.coeffs.append(mult)

    def arrays(self):
        dt = np.complex128 if self.is_complex else np.float64
        return (np.asarray(self.codes, dtype=np.int32), np.asarray(self.sites, dtype=np.int32).reshape(-1, 2),
                np.asarray(self.coeffs, dtype=dt))


def _isclose0(v):
    return bool(np.isclose(v, 0.0))


def terms_j1j2(lat: OLattice, J1, J2, hx=None, hy=None, hz=None) -> Terms:
    """Model/Interacting/Spin/j1j2.py:119-292 (spin-1/2 family): fields
    -h*S_i, then -J1 per NN forward bond and -J2 per NNN forward bond, each
    added as Sx/C, Sy/C (hamiltonian_spin.py:188-218) then Sz/C."""
    ns = lat.ns
    t = Terms(ns, is_complex=hy is not None)
    arr = lambda v: None if v is None else (np.full(ns, float(v)) if np.isscalar(v) else np.asarray(v, float))
    j1, j2, hx_, hy_, hz_ = arr(J1), arr(J2), arr(hx), arr(hy), arr(hz)
    ok = lambda v, i: v is not None and not _isclose0(v[i])   # _ADD_CONDITION (hamil.py:123-206)
    for i in range(ns):
        if ok(hx_, i):
            t.add(SIG_X_L, -hx_[i], [i])
        if ok(hy_, i):
            t.add(SIG_Y_L, -hy_[i], [i])
        if ok(hz_, i):
            t.add(SIG_Z_L, -hz_[i], [i])
    for order, J in ((1, j1), (2, j2)):
        if J is None:
            continue
        if order == 2 and lat.kind == "honeycomb":
            continue
        by_site = {}
        for (i, j, n) in oracle_bonds(lat, order):
            by_site.setdefault(i, []).append(j)
        for i in range(ns):
            for j in by_site.get(i, []):
                jj = -1.0 * J[i]
                if _isclose0(jj):
                    continue
                t.add(SIG_X_C, jj, [i, j])
                t.add(SIG_Y_C, jj, [i, j])
                t.add(SIG_Z_C, jj, [i, j])
    return t


def terms_xxz(lat: OLattice, jxy=1.0, jz=1.0, hx=0.0, hz=0.0) -> Terms:
    """Model/Interacting/Spin/xxz.py:278-431 (dtype forced complex, :131):
    -hx Sx_i, -hz Sz_i per site, then per NN forward bond -Jxy (SxSx+SySy)
    and -Jz SzSz."""
    ns = lat.ns
    t = Terms(ns, is_complex=True)
    arr = lambda v: np.full(ns, float(v)) if np.isscalar(v) else np.asarray(v, float)
    jxy_, jz_, hx_, hz_ = arr(jxy), arr(jz), arr(hx), arr(hz)
    for i in range(ns):
        if not _isclose0(hx_[i]):
            t.add(SIG_X_L, -hx_[i], [i])
    for i in range(ns):
        if not _isclose0(hz_[i]):
            t.add(SIG_Z_L, -hz_[i], [i])
    by_site = {}
    for (i, j, n) in oracle_bonds(lat, 1):
        by_site.setdefault(i, []).append(j)
    for i in range(ns):
        for j in by_site.get(i, []):
            if not _isclose0(jxy_[i]):
                t.add(SIG_X_C, -jxy_[i], [i, j])
                t.add(SIG_Y_C, -jxy_[i], [i, j])
            if not _isclose0(jz_[i]):
                t.add(SIG_Z_C, -jz_[i], [i, j])
    return t


def bond_windings(lat: OLattice) -> List[Tuple[int, int, Tuple[int, int]]]:
    """Forward NN bonds (i, j) of oracle_bonds with their winding (wx, wy):
    +1 where the forward step leaves the cell range along that direction
    (chain: i -> i+1 at i = L-1; square: (x+1, y) at x = lx-1, (x, y+1) at
    y = ly-1)."""
    out = []
    for (i, j, n) in oracle_bonds(lat, 1):
        if lat.kind == "chain":
            w = (1 if i == lat.lx - 1 else 0, 0)
        elif lat.kind == "square":
            x, y, _ = lat.get_coordinates(i)
            w = (1 if (n == 0 and x == lat.lx - 1) else 0, 1 if (n == 1 and y == lat.ly - 1) else 0)
        else:
            raise ValueError("twisted bonds for chains and square lattices")
        out.append((i, j, w))
    return out


def terms_xxz_twisted(lat: OLattice, jxy, jz, flux) -> Terms:
    """Twisted-boundary XXZ (no reference model): per forward bond
    -2 jxy (e^{i theta} S+_i S-_j + e^{-i theta} S-_i S+_j) (codes 31, 32;
    XX + YY = 2 (PM + MP) in the VM's units) and -jz SzSz, theta = w . flux."""
    t = Terms(lat.ns, is_complex=True)
    fx, fy = flux
    for (i, j, (wx, wy)) in bond_windings(lat):
        ph = np.exp(1j * (wx * fx + wy * fy))
        t.add(SIG_PM_C, -2.0 * jxy * ph, [i, j])
        t.add(SIG_MP_C, -2.0 * jxy * np.conj(ph), [i, j])
        t.add(SIG_Z_C, -jz, [i, j])
    return t


def boundary_phase_table(ns: int, flux) -> np.ndarray:
    """(3, ns + 1) table exp(i phi_d n) (the lattice's boundary_phases())."""
    n = np.arange(ns + 1)
    fx, fy = flux
    return np.stack([np.exp(1j * fx * n), np.exp(1j * fy * n), np.ones(ns + 1, dtype=np.complex128)])


def honeycomb_bond_flux_phase(lat: OLattice, i: int, n: int, flux) -> float:
    """Hermitian boundary phase Re exp(i w.phi) of honeycomb forward bond n
    from A site i (heisenberg_kitaev.py:560-571 with use_forward=True): the x
    bond (to cell x-1) winds -1 in x when x = 0, the y bond (to cell y-1)
    winds -1 in y when y = 0, the z bond stays in its cell."""
    phx, phy = flux
    x, y = (i // 2) % lat.lx, (i // 2) // lat.lx
    w = (-1 if (n == 0 and x == 0) else 0, -1 if (n == 1 and y == 0) else 0)
    return float(np.real(np.exp(1j * (w[0] * phx + w[1] * phy))))


def terms_heisenberg_kitaev(lat: OLattice, K, J, hx=0.0, hy=0.0, hz=0.0, dlt=1.0, flux=None) -> Terms:
    """Model/Interacting/Spin/heisenberg_kitaev.py:378-639 (spin-1/2):
    fields scaled by 2 (SINGLE_TERM_MULT, :74) in order z, y, x per site;
    per forward bond SzSz, SxSx, SySy scaled by 4 (CORR_TERM_MULT, :76) with
    the Kitaev coupling subtracted on its bond direction (:581-586); with a
    boundary flux (phi_x, phi_y) each bond's couplings carry the Hermitian part
    of its boundary phase (:487-571)."""
    ns = lat.ns
    Kx, Ky, Kz = (K, K, K) if np.isscalar(K) else K
    is_c = not _isclose0(hy)
    t = Terms(ns, is_complex=is_c)
    bonds = {}
    for (i, j, n) in oracle_bonds(lat, 1):
        bonds.setdefault(i, []).append((j, n))
    for i in range(ns):
        if not _isclose0(hz):
            t.add(SIG_Z_L, -2.0 * hz, [i])
        if not _isclose0(hy):
            t.add(SIG_Y_L, -2.0 * hy, [i])
        if not _isclose0(hx):
            t.add(SIG_X_L, -2.0 * hx, [i])
        for (j, n) in bonds.get(i, []):
            ph = 1.0 if flux is None else honeycomb_bond_flux_phase(lat, i, n, flux)
            szsz = ph * 4.0 * J * dlt
            sxsx = ph * 4.0 * J
            sysy = ph * 4.0 * J
            if n == 2:
                szsz -= ph * 4.0 * Kz
            elif n == 1:
                sysy -= ph * 4.0 * Ky
            elif n == 0:
                sxsx -= ph * 4.0 * Kx
            if not _isclose0(szsz):
                t.add(SIG_Z_C, szsz, [i, j])
            if not _isclose0(sxsx):
                t.add(SIG_X_C, sxsx, [i, j])
            if not _isclose0(sysy):
                t.add(SIG_Y_C, sysy, [i, j])
    return t


def full_sparse_hamiltonian(ns: int, codes, sites, coeffs, is_complex: bool):
    """Sparse full-space H from Kronecker products (scipy.sparse.kron), the
    same construction as ``full_dense_hamiltonian`` for larger ns."""
    import scipy.sparse as sp
    H = None
    for k in range(len(codes)):
        kinds = _CODE_KINDS[int(codes[k])]
        s = [int(v) for v in sites[k] if int(v) >= 0][:len(kinds)]
        mats = {}
        for kind, site in zip(kinds, s):
            m = _S[kind]
            mats[site] = m @ mats[site] if site in mats else m
        M = None
        for site in range(ns):
            f = sp.csr_matrix(mats[site]) if site in mats else sp.identity(2, format="csr", dtype=np.complex128)
            M = f if M is None else sp.kron(M, f, format="csr")
        c = coeffs[k]
        if not is_complex:
            M = sp.csr_matrix(M.real.astype(np.complex128))
            c = np.real(c)
        H = c * M if H is None else H + c * M
    return H.tocsr()


def canonicalize(group: Group, states: np.ndarray):
    """min_g g(s), first minimiser, number of minimisers (jit :234-246)."""
    best = np.full(states.shape, np.iinfo(np.int64).max, dtype=np.int64)
    bg = np.full(states.shape, -1, dtype=np.int64)
    cnt = np.zeros(states.shape, dtype=np.int64)
    for g in range(group.n_group):
        t, _ = group.apply(states, g)
        lt = t < best
        eq = t == best
        cnt = np.where(lt, 1, np.where(eq, cnt + 1, cnt))
        best = np.where(lt, t, best)
        bg = np.where(lt, g, bg)
    return best, bg, cnt


def sampled_rows(group: Group, reps: np.ndarray, norms: np.ndarray, sample: np.ndarray, ns: int, codes, sites,
                 coeffs, is_complex: bool, x: np.ndarray) -> np.ndarray:
    """(H x)[a] for the sampled row indices, built from the push-form column
    of each sampled representative (matrix_builder.py:1358-1377):
    H[idx, k] += val * conj(chi(g*)) * N_idx / N_k, and (H x)[a] =
    sum_idx conj(H[idx, a]) x[idx] for Hermitian H.  Lookups use binary search
    in the given representative list (symmetry_container.py:72-102); states
    whose canonical form is not listed are outside the sector."""
    reps = np.asarray(reps, dtype=np.int64)
    r = reps[sample]
    na = norms[sample]
    acc = np.zeros(len(sample), dtype=np.complex128)
    for st, val, keep in vm_apply(r, ns, codes, sites, coeffs, is_complex):
        best, bg, _ = canonicalize(group, st)
        pos = np.searchsorted(reps, best)
        pos_c = np.minimum(pos, len(reps) - 1)
        found = keep & (reps[pos_c] == best)
        # the reference additionally requires |proj(s')| > 1e-7; a state whose
        # canonical form is a listed representative always satisfies it
        chi = group.chi[np.maximum(bg, 0)]
        h = val * np.conj(chi) * norms[pos_c] / na
        acc += np.where(found, np.conj(h) * x[pos_c], 0.0)
    return acc
