"""Spin-1/2 models with the reference's term conventions.

Each builder appends instructions in the reference's order and with its
signs/scales (these are not in the docstrings there):
  * J1J2Model adds -J1 / -J2 per forward bond (Model/Interacting/Spin/j1j2.py:263,291),
    fields -h S_i (:217-225); XY exchange = Sx/C + Sy/C (hamiltonian_spin.py:188-218).
  * XXZ has an overall minus: -Jxy (SxSx + SySy) - Jz SzSz - hx Sx - hz Sz,
    dtype forced to complex128 (xxz.py:131, :278-431).
  * HeisenbergKitaev uses Pauli units: fields x2, bonds x4 (heisenberg_kitaev.py:72-76),
    Kitaev subtracted on its bond direction (:575-586).

``coupling`` ids tag every instruction with the parameter it belongs to, so
``Hamiltonian.matvec_couplings_device`` returns dH/dJ_c x for the autograd
path.  With ``split_couplings=True`` terms that the reference merges (e.g.
4J - 4K on one SzSz bond) are added as separate instructions, same matrix.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from .operator import Hamiltonian, spin_op


def _arr(v, ns):
    if v is None:
        return None
    if np.isscalar(v):
        return np.full(ns, float(np.real(v)))
    return np.asarray(v, dtype=float)


def _cv(v) -> float:
    """Reference value of one coupling id for the differentiable path: the
    scalar (or uniform array) value it was compiled at; 1.0 (relative scale)
    for site-dependent arrays and for absent or zero couplings."""
    if v is None:
        return 1.0
    a = np.atleast_1d(np.asarray(v, dtype=float))
    if a.size and np.all(a == a.flat[0]) and a.flat[0] != 0.0:
        return float(a.flat[0])
    return 1.0


def _bond_phase(lattice, i, j, hermitize: bool):
    """Boundary phase of the forward bond i -> j (lattice.boundary_phase_from_winding
    of its winding).  ``hermitize``: its Hermitian part Re exp(i phi), what the
    reference uses for two-site spin bilinears listed once per bond
    (heisenberg_kitaev.py:487-571).  Otherwise a complex phase would make the
    bond's real spin bilinear non-Hermitian (the reference multiplies the
    couplings by it, xxz.py:393-431, j1j2.py:261-291), which is refused."""
    ph = lattice.boundary_phase_from_winding(*lattice.bond_winding(i, j))
    if np.iscomplexobj(ph) or isinstance(ph, complex):
        if hermitize:
            return float(np.real(0.5 * (ph + np.conj(ph))))
        if abs(np.imag(ph)) > 1e-14:
            raise ValueError(f"bond ({i}, {j}) carries the complex boundary phase {ph}: this model's bond would not "
                             "be Hermitian (use a flux of 0 or pi along that direction)")
        return float(np.real(ph))
    return ph


def _nz(v, i=None) -> bool:
    """_ADD_CONDITION (Algebra/hamil.py:123-206) for scalars/arrays."""
    if v is None:
        return False
    x = v[i] if i is not None else v
    return not bool(np.isclose(x, 0.0))


class J1J2Model(Hamiltonian):
    """H = sum_<ij> (-J1) S_i.S_j + sum_<<ij>> (-J2) S_i.S_j - sum_i h.S_i in the
    reference's sign convention: pass J1=-1 for the antiferromagnet."""

    COUPLINGS = ("J1", "J2", "hx", "hy", "hz")

    def __init__(self, lattice, J1=1.0, J2=0.0, *, hilbert_space=None, hx=None, hy=None, hz=None,
                 dtype=np.float64, **kwargs):
        if lattice is None:
            raise ValueError("J1J2 requires a lattice.")
        complex_required = hy is not None
        super().__init__(hilbert_space, dtype=np.complex128 if complex_required else dtype, name="J1-J2",
                         lattice=lattice)
        ns = lattice.ns
        self._j1, self._j2 = _arr(J1, ns), _arr(J2, ns)
        self._hx, self._hy, self._hz = _arr(hx, ns), _arr(hy, ns), _arr(hz, ns)
        sxl, syl, szl = spin_op("x", "local"), spin_op("y", "local"), spin_op("z", "local")
        sxx, syy, szz = spin_op("x", "correlation"), spin_op("y", "correlation"), spin_op("z", "correlation")
        for i in range(ns):
            if _nz(self._hx, i):
                self.add(sxl, -self._hx[i], True, [i], coupling=2)
            if _nz(self._hy, i):
                self.add(syl, -self._hy[i], True, [i], coupling=3)
            if _nz(self._hz, i):
                self.add(szl, -self._hz[i], False, [i], coupling=4)
        for cpl, J, getn, getj in ((0, self._j1, lattice.get_nn_forward_num, lattice.get_nn_forward),
                                  (1, self._j2, lattice.get_nnn_forward_num, lattice.get_nnn_forward)):
            if J is None:
                continue
            for i in range(ns):
                for nidx in range(getn(i)):
                    j = getj(i, num=nidx)
                    if lattice.wrong_nei(j) or int(j) == i:
                        continue
                    jj = -_bond_phase(lattice, i, int(j), False) * J[i]
                    if not _nz(jj):
                        continue
                    self.add(sxx, jj, True, [i, int(j)], coupling=cpl)
                    self.add(syy, jj, True, [i, int(j)], coupling=cpl)
                    self.add(szz, jj, False, [i, int(j)], coupling=cpl)
        self.coupling_values = np.array([_cv(J1), _cv(J2), _cv(hx), _cv(hy), _cv(hz)])[: self.n_coupling_ids]


class XXZ(Hamiltonian):
    """H = -sum_<ij> [Jxy (SxSx+SySy) + Jz SzSz] - sum_i (hx Sx_i + hz Sz_i).

    ``twist=True`` (no reference counterpart; the reference multiplies the
    whole bond by the raw flux phase, xxz.py:393-431, which is Hermitian only
    for phases +-1): a lattice flux enters as a twisted boundary,
    -Jxy (SxSx + SySy) -> -2 Jxy (e^{i theta} S+_i S-_j + e^{-i theta} S-_i S+_j)
    in the instruction VM's units (codes 31, 32; theta = w . phi of the
    forward bond i -> j), which commutes with the twisted translations that
    multiply exp(i phi) per up spin crossing the boundary
    (Lattice.boundary_phases), so momentum sectors stay symmetry sectors."""

    COUPLINGS = ("jxy", "jz", "hx", "hz")

    def __init__(self, lattice, hilbert_space=None, jxy=1.0, jz=1.0, delta=None, hx=0.0, hz=0.0, twist: bool = False,
                 **kwargs):
        if lattice is None:
            raise ValueError("XXZ requires a 'lattice' object to define neighbors.")
        super().__init__(hilbert_space, dtype=np.complex128, name="XXZ Model", lattice=lattice)
        ns = lattice.ns
        self._jxy = _arr(jxy, ns)
        self._jz = self._jxy * _arr(delta, ns) if delta is not None else _arr(jz, ns)
        self._hx, self._hz = _arr(hx, ns), _arr(hz, ns)
        sxl, szl = spin_op("x", "local"), spin_op("z", "local")
        sxx, syy, szz = spin_op("x", "correlation"), spin_op("y", "correlation"), spin_op("z", "correlation")
        for i in range(ns):
            if _nz(self._hx, i):
                self.add(sxl, -self._hx[i], True, [i], coupling=2)
        for i in range(ns):
            if _nz(self._hz, i):
                self.add(szl, -self._hz[i], False, [i], coupling=3)
        for i in range(ns):
            for nidx in range(lattice.get_nn_forward_num(i)):
                j = lattice.get_nn_forward(i, num=nidx)
                if lattice.wrong_nei(j):
                    continue
                if twist:
                    ph = complex(lattice.boundary_phase_from_winding(*lattice.bond_winding(i, int(j))))
                    if _nz(self._jxy[i]):
                        self.add(spin_op("pm", "correlation"), -2.0 * self._jxy[i] * ph, True, [i, int(j)],
                                 coupling=0)
                        self.add(spin_op("mp", "correlation"), -2.0 * self._jxy[i] * np.conj(ph), True,
                                 [i, int(j)], coupling=0)
                    if _nz(self._jz[i]):
                        self.add(szz, -self._jz[i], False, [i, int(j)], coupling=1)
                    continue
                phase = _bond_phase(lattice, i, int(j), False)
                cxy = self._jxy[i] * phase
                if _nz(cxy):
                    self.add(sxx, -cxy, True, [i, int(j)], coupling=0)
                    self.add(syy, -cxy, True, [i, int(j)], coupling=0)
                cz = self._jz[i] * phase
                if _nz(cz):
                    self.add(szz, -cz, False, [i, int(j)], coupling=1)
        self.coupling_values = np.array([_cv(self._jxy), _cv(self._jz), _cv(hx), _cv(hz)])[: self.n_coupling_ids]


class HeisenbergKitaev(Hamiltonian):
    """Kitaev-Heisenberg honeycomb model in the reference's Pauli scaling."""

    SINGLE_TERM_MULT = 2.0
    CORR_TERM_MULT = 4.0
    COUPLINGS = ("K", "J", "h")

    def __init__(self, lattice, K=1.0, J=0.0, *, hilbert_space=None, hx=0.0, hy=0.0, hz=0.0, dlt=1.0,
                 dtype=np.float64, split_couplings: bool = False, **kwargs):
        from .lattice import X_BOND_NEI, Y_BOND_NEI, Z_BOND_NEI
        if lattice is None:
            raise ValueError("HeisenbergKitaev requires a lattice.")
        complex_required = _nz(hy)
        super().__init__(hilbert_space, dtype=np.complex128 if complex_required else dtype,
                         name="Heisenberg-Kitaev", lattice=lattice)
        kx, ky, kz = (K, K, K) if np.isscalar(K) else tuple(K)
        self._k = (kx, ky, kz)
        self._j, self._dlt = J, dlt
        self._h = (hx, hy, hz)
        ls, cs = self.SINGLE_TERM_MULT, self.CORR_TERM_MULT
        sxl, syl, szl = spin_op("x", "local"), spin_op("y", "local"), spin_op("z", "local")
        sxx, syy, szz = spin_op("x", "correlation"), spin_op("y", "correlation"), spin_op("z", "correlation")
        ns = lattice.ns
        for i in range(ns):
            if _nz(hz):
                self.add(szl, -ls * hz, False, [i], coupling=2)
            if _nz(hy):
                self.add(syl, -ls * hy, True, [i], coupling=2)
            if _nz(hx):
                self.add(sxl, -ls * hx, True, [i], coupling=2)
            for nn in range(lattice.get_nn_forward_num(i)):
                nei = lattice.get_nn_forward(i, num=nn)
                if lattice.wrong_nei(nei):
                    continue
                nei = int(nei)
                # forward bonds only: Hermitian part of a flux phase (heisenberg_kitaev.py:487-571)
                phase = _bond_phase(lattice, i, nei, bool(getattr(lattice, "has_flux", False)))
                if not split_couplings:
                    szsz = phase * cs * J * dlt
                    sxsx = phase * cs * J
                    sysy = phase * cs * J
                    if nn == Z_BOND_NEI:
                        szsz -= phase * cs * kz
                    elif nn == Y_BOND_NEI:
                        sysy -= phase * cs * ky
                    elif nn == X_BOND_NEI:
                        sxsx -= phase * cs * kx
                    if _nz(szsz):
                        self.add(szz, szsz, False, [i, nei])
                    if _nz(sxsx):
                        self.add(sxx, sxsx, True, [i, nei])
                    if _nz(sysy):
                        self.add(syy, sysy, True, [i, nei])
                else:
                    kk = {Z_BOND_NEI: (szz, kz), Y_BOND_NEI: (syy, ky), X_BOND_NEI: (sxx, kx)}[nn]
                    if _nz(J):
                        self.add(szz, phase * cs * J * dlt, False, [i, nei], coupling=1)
                        self.add(sxx, phase * cs * J, True, [i, nei], coupling=1)
                        self.add(syy, phase * cs * J, True, [i, nei], coupling=1)
                    if _nz(kk[1]):
                        self.add(kk[0], -phase * cs * kk[1], kk[0] is not szz, [i, nei], coupling=0)
        hmag = float(np.sqrt(sum(float(np.real(c)) ** 2 for c in (hx, hy, hz))))
        kval = _cv(kx) if (kx == ky == kz) else 1.0
        if split_couplings:
            vals = [kval, _cv(J), hmag if hmag > 0 else 1.0]
        else:
            vals = [1.0, 1.0, hmag if hmag > 0 else 1.0]   # K and J share the bond instructions
        self.coupling_values = np.array(vals)[: self.n_coupling_ids]
