"""Lattices used by the symmetric-sector path.

The reference keeps its lattices in the un-vendored ``general_python``
submodule; its symmetry and model code relies on this interface:
``lx/ly/lz``, ``ns``/``Ns``, ``dim``, ``_basis`` (sites per unit cell),
``get_coordinates``, ``site_index`` (Algebra/Symmetries/translation.py:354-436),
``get_nn_forward_num``/``get_nn_forward``, ``get_nnn_forward_num``/
``get_nnn_forward``, ``wrong_nei``, ``bond_winding`` and
``boundary_phase_from_winding`` (Model/Interacting/Spin/j1j2.py:241-290,
xxz.py:400-408, heisenberg_kitaev.py:533-545).

Conventions (fixed here, since the submodule is absent):
  * Bravais lattices: site = x + lx*y (tests/algebra/test_momentum_sectors.py:24-31).
  * Honeycomb: site = 2*(x + lx*y) + sublattice (translation.py:369-404).
  * Forward neighbours -- chain: i+1 (NN), i+2 (NNN); square: (x+1,y),
    (x,y+1) (NN) and (x+1,y+1), (x-1,y+1) (NNN); honeycomb: the A site of cell
    (x,y) bonds to B(x-1,y) [x bond, nn index 0], B(x,y-1) [y bond, 1] and
    B(x,y) [z bond, 2].
  * Geometry (read by the inversion permutation, Algebra/Symmetries/
    inversion.py:135-230): Bravais a1 = x, a2 = y; honeycomb a1 = (sqrt3, 0),
    a2 = (sqrt3/2, 3/2) with B at (a1 + a2)/3 (unit-length bonds).
  * Boundary flux ``flux = {"x": phi_x, "y": phi_y}`` (radians): a bond that
    winds w_d times around periodic direction d carries
    exp(i sum_d w_d phi_d) (``boundary_phase_from_winding``), and
    ``boundary_phases()[d][n] = exp(i phi_d n)`` for n particles crossing the
    boundary (the table the reference's translation reads,
    symmetry_container.py:1047-1051, translation.py:667-668).  Forward bonds
    that leave the cell range in the negative direction wind -1, in the
    positive direction +1.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

X_BOND_NEI, Y_BOND_NEI, Z_BOND_NEI = 0, 1, 2


class Lattice:
    kind = "lattice"
    n_basis = 1

    def __init__(self, lx: int, ly: int = 1, lz: int = 1, bc: str = "pbc", flux=None):
        if bc.lower() != "pbc":
            raise ValueError("only periodic boundary conditions are supported")
        self.lx, self.ly, self.lz = int(lx), int(ly), int(lz)
        self.bc = "pbc"
        self._basis = [np.zeros(3)] * self.n_basis
        self._winding = {}
        self._nn = self._build_forward(1)
        self._nnn = self._build_forward(2)
        self.flux = flux

    # ---- boundary flux ----------------------------------------------------
    @property
    def flux(self) -> dict:
        return dict(self._flux)

    @flux.setter
    def flux(self, value):
        """None, a number (direction x) or {"x"|"y"|"z": phi} in radians."""
        f = {"x": 0.0, "y": 0.0, "z": 0.0}
        if value is None:
            pass
        elif np.isscalar(value):
            f["x"] = float(value)
        else:
            for k, v in dict(value).items():
                key = str(getattr(k, "name", k)).lower()[-1]
                if key not in f:
                    raise ValueError(f"unknown flux direction {k!r}")
                f[key] = float(v)
        self._flux = f

    @property
    def has_flux(self) -> bool:
        return any(abs(v) > 0.0 for v in self._flux.values())

    # ---- geometry -------------------------------------------------------
    @property
    def ns(self) -> int:
        return self.lx * self.ly * self.lz * self.n_basis

    Ns = ns

    @property
    def dim(self) -> int:
        return 1 if self.ly == 1 and self.lz == 1 else (2 if self.lz == 1 else 3)

    # ---- positions ------------------------------------------------------
    @property
    def a1(self) -> np.ndarray:
        return np.array([1.0, 0.0, 0.0])

    @property
    def a2(self) -> np.ndarray:
        return np.array([0.0, 1.0, 0.0])

    @property
    def a3(self) -> np.ndarray:
        return np.array([0.0, 0.0, 1.0])

    @property
    def basis(self) -> np.ndarray:
        return np.zeros((1, 3))

    @property
    def rvectors(self) -> np.ndarray:
        """Position of every site (ns, 3)."""
        out = np.zeros((self.ns, 3))
        for s in range(self.ns):
            x, y, z = self.get_coordinates(s)
            out[s] = x * self.a1 + y * self.a2 + z * self.a3 + self.basis[s % self.n_basis]
        return out

    def periodic_flags(self):
        return (True, self.ly > 1, self.lz > 1)

    def get_coordinates(self, site: int) -> Tuple[int, int, int]:
        cell = site // self.n_basis
        return (cell % self.lx, (cell // self.lx) % self.ly, cell // (self.lx * self.ly))

    def site_index(self, x: int, y: int = 0, z: int = 0, sub: int = 0) -> int:
        x, y, z = x % self.lx, y % self.ly, z % self.lz
        return self.n_basis * (x + self.lx * y + self.lx * self.ly * z) + sub

    # ---- neighbours -----------------------------------------------------
    def _forward(self, site: int, order: int) -> List[int]:
        raise NotImplementedError

    def _forward_cells(self, site: int, order: int):
        """Forward neighbours as unwrapped (x, y, sub) of the target."""
        raise NotImplementedError

    def _forward(self, site: int, order: int) -> List[int]:
        out = []
        for xx, yy, sb in self._forward_cells(site, order):
            j = self.site_index(xx, yy, sub=sb)
            if j != site:
                out.append(j)
        return out

    def _build_forward(self, order: int):
        out = []
        for i in range(self.ns):
            try:
                cells = self._forward_cells(i, order)
            except NotImplementedError:
                cells = []
            js = []
            for xx, yy, sb in cells:
                j = self.site_index(xx, yy, sub=sb)
                if j == i:
                    continue
                js.append(j)
                w = (xx // self.lx, yy // self.ly if self.ly > 1 else 0, 0)
                self._winding.setdefault((i, j), w)
            out.append(js)
        return out

    def get_nn_forward_num(self, i: int) -> int:
        return len(self._nn[i])

    def get_nn_forward(self, i: int, num: int) -> int:
        return self._nn[i][num]

    def get_nnn_forward_num(self, i: int) -> int:
        return len(self._nnn[i])

    def get_nnn_forward(self, i: int, num: int) -> int:
        return self._nnn[i][num]

    def wrong_nei(self, j) -> bool:
        return j is None or int(j) < 0

    def bond_winding(self, i: int, j: int):
        """Times the forward bond i -> j winds around each periodic direction."""
        return self._winding.get((int(i), int(j)), (0, 0, 0))

    def boundary_phase_from_winding(self, wx, wy, wz):
        """exp(i (wx phi_x + wy phi_y + wz phi_z)); a real 1.0 without flux."""
        ang = wx * self._flux["x"] + wy * self._flux["y"] + wz * self._flux["z"]
        if ang == 0.0:
            return 1.0
        return complex(np.exp(1j * ang))

    def boundary_phases(self) -> np.ndarray:
        """(3, ns + 1) table: exp(i phi_d n) for n particles crossing direction d."""
        n = np.arange(self.ns + 1)
        return np.stack([np.exp(1j * self._flux[d] * n) for d in "xyz"])

    def bonds(self, order: int = 1):
        """Forward bonds (i, j, nn_index) in site order."""
        src = self._nn if order == 1 else self._nnn
        return [(i, j, n) for i in range(self.ns) for n, j in enumerate(src[i])]


class ChainLattice(Lattice):
    kind = "chain"

    def __init__(self, L: int, bc: str = "pbc", flux=None):
        super().__init__(L, 1, 1, bc, flux=flux)

    def _forward_cells(self, site, order):
        return [(site + order, 0, 0)]


class SquareLattice(Lattice):
    kind = "square"

    def _forward_cells(self, site, order):
        x, y, _ = self.get_coordinates(site)
        cand = [(x + 1, y), (x, y + 1)] if order == 1 else [(x + 1, y + 1), (x - 1, y + 1)]
        return [(xx, yy, 0) for xx, yy in cand]


class HoneycombLattice(Lattice):
    kind = "honeycomb"
    n_basis = 2

    @property
    def a1(self) -> np.ndarray:
        return np.array([np.sqrt(3.0), 0.0, 0.0])

    @property
    def a2(self) -> np.ndarray:
        return np.array([np.sqrt(3.0) / 2.0, 1.5, 0.0])

    @property
    def basis(self) -> np.ndarray:
        return np.array([np.zeros(3), (self.a1 + self.a2) / 3.0])

    def _forward_cells(self, site, order):
        if order != 1:
            raise NotImplementedError
        if site % 2 == 1:
            return []
        x, y, _ = self.get_coordinates(site)
        return [(x - 1, y, 1), (x, y - 1, 1), (x, y, 1)]
