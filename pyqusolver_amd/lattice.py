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
  * No flux: boundary phases are 1.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

X_BOND_NEI, Y_BOND_NEI, Z_BOND_NEI = 0, 1, 2


class Lattice:
    kind = "lattice"
    n_basis = 1

    def __init__(self, lx: int, ly: int = 1, lz: int = 1, bc: str = "pbc"):
        if bc.lower() != "pbc":
            raise ValueError("only periodic boundary conditions are supported")
        self.lx, self.ly, self.lz = int(lx), int(ly), int(lz)
        self.bc = "pbc"
        self._basis = [np.zeros(3)] * self.n_basis
        self._nn = self._build_forward(1)
        self._nnn = self._build_forward(2)

    # ---- geometry -------------------------------------------------------
    @property
    def ns(self) -> int:
        return self.lx * self.ly * self.lz * self.n_basis

    Ns = ns

    @property
    def dim(self) -> int:
        return 1 if self.ly == 1 and self.lz == 1 else (2 if self.lz == 1 else 3)

    def get_coordinates(self, site: int) -> Tuple[int, int, int]:
        cell = site // self.n_basis
        return (cell % self.lx, (cell // self.lx) % self.ly, cell // (self.lx * self.ly))

    def site_index(self, x: int, y: int = 0, z: int = 0, sub: int = 0) -> int:
        x, y, z = x % self.lx, y % self.ly, z % self.lz
        return self.n_basis * (x + self.lx * y + self.lx * self.ly * z) + sub

    # ---- neighbours -----------------------------------------------------
    def _forward(self, site: int, order: int) -> List[int]:
        raise NotImplementedError

    def _build_forward(self, order: int):
        out = []
        for i in range(self.ns):
            try:
                js = self._forward(i, order)
            except NotImplementedError:
                js = []
            out.append([j for j in js])
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
        return (0, 0, 0)

    def boundary_phase_from_winding(self, wx, wy, wz) -> float:
        return 1.0

    def bonds(self, order: int = 1):
        """Forward bonds (i, j, nn_index) in site order."""
        src = self._nn if order == 1 else self._nnn
        return [(i, j, n) for i in range(self.ns) for n, j in enumerate(src[i])]


class ChainLattice(Lattice):
    kind = "chain"

    def __init__(self, L: int, bc: str = "pbc"):
        super().__init__(L, 1, 1, bc)

    def _forward(self, site, order):
        j = (site + order) % self.lx
        return [j] if j != site else []


class SquareLattice(Lattice):
    kind = "square"

    def _forward(self, site, order):
        x, y, _ = self.get_coordinates(site)
        cand = [(x + 1, y), (x, y + 1)] if order == 1 else [(x + 1, y + 1), (x - 1, y + 1)]
        out = []
        for xx, yy in cand:
            j = self.site_index(xx, yy)
            if j != site:
                out.append(j)
        return out


class HoneycombLattice(Lattice):
    kind = "honeycomb"
    n_basis = 2

    def _forward(self, site, order):
        if order != 1:
            raise NotImplementedError
        if site % 2 == 1:
            return []
        x, y, _ = self.get_coordinates(site)
        return [self.site_index(x - 1, y, sub=1), self.site_index(x, y - 1, sub=1), self.site_index(x, y, sub=1)]
