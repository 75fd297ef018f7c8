"""Host-side logic of pyqusolver_amd that needs no GPU: sector parsing,
characters in reference order, model instruction lists, refused sectors."""
import numpy as np
import pytest

from oracle import qes_oracle as O
from pyqusolver_amd import ChainLattice, HoneycombLattice, SquareLattice
from pyqusolver_amd.symmetry import (SymmetryGenerators, build_sector_group, get_u1_sym,
                                     normalize_symmetry_generators)


def _grp(lat, sym, n_up=None):
    specs = normalize_symmetry_generators(sym)
    gl = [get_u1_sym(lat, n_up)] if n_up is not None else []
    return build_sector_group(lat, lat.ns, specs, gl)


@pytest.mark.parametrize("lat,olat,sym,ogens,n_up", [
    (ChainLattice(12), O.OLattice("chain", 12), {"translation": 5}, [("tx", 5)], 6),
    (ChainLattice(12), O.OLattice("chain", 12), {"translation": 6, "reflection": -1, "parity": {"x": -1}},
     [("tx", 6), ("R", -1), ("Px", -1)], 6),
    (SquareLattice(4, 3), O.OLattice("square", 4, 3), {"translation": {"kx": 1, "ky": 2}, "parity": {"x": 1}},
     [("tx", 1), ("ty", 2), ("Px", 1)], 6),
    (HoneycombLattice(3, 2), O.OLattice("honeycomb", 3, 2), {"translation": {"kx": 2, "ky": 1}},
     [("tx", 2), ("ty", 1)], None),
])
def test_characters_bit_exact_in_reference_order(lat, olat, sym, ogens, n_up):
    g = _grp(lat, sym, n_up)
    og = O.Group(olat, [O.Generator(k, s) for k, s in ogens], n_up=n_up)
    assert g.order == og.n_group
    np.testing.assert_array_equal(g.chi, og.chi)
    gp, pt = O.phases_map(og)
    np.testing.assert_array_equal(g.g_to_pidx, gp)
    np.testing.assert_array_equal(g.phase_table, pt)


def test_parity_z_dropped_with_u1_and_x_needs_half_filling():
    lat = ChainLattice(10)
    g = _grp(lat, {"translation": 0, "parity": 1}, 5)           # int -> ParityZ, dropped
    assert g.nbase == 1
    g = _grp(lat, {"translation": 0, "parity": {"x": 1}}, 4)    # not half filling
    assert g.nbase == 1
    g = _grp(lat, {"translation": 0, "parity": {"x": 1}}, 5)
    assert g.nbase == 2


def test_non_hermitian_reference_sectors_are_flagged():
    lat = ChainLattice(8)
    assert _grp(lat, {"translation": 2, "reflection": 1}, 4).matvec_error
    assert _grp(lat, {"translation": 4, "reflection": 1}, 4).matvec_error is None
    assert _grp(lat, {"translation": 0, "parity": {"y": 1}}, 4).matvec_error
    # the flagged sectors are indeed non-Hermitian in the reference semantics
    olat = O.OLattice("chain", 8)
    t = O.terms_xxz(olat, 1.0, 0.7)
    c, s, co = t.arrays()
    for gens in ([O.Generator("tx", 2), O.Generator("R", 1)], [O.Generator("tx", 0), O.Generator("Py", 1)]):
        grp = O.Group(olat, gens, n_up=4)
        b = O.scan_basis(grp)
        H = O.reduced_dense(b, 8, c, s, co, True)
        assert np.abs(H - H.conj().T).max() > 1e-3


def _dummy_space(lat):
    class S:
        pass
    s = S()
    s.ns = lat.ns
    s.lattice = lat
    return s


def test_model_instruction_lists_match_oracle():
    from pyqusolver_amd.models import HeisenbergKitaev, J1J2Model, XXZ
    for lat, olat in ((ChainLattice(10), O.OLattice("chain", 10)), (SquareLattice(4, 3), O.OLattice("square", 4, 3))):
        H = J1J2Model(lat, J1=-1.0, J2=-0.5, hilbert_space=_dummy_space(lat), hz=0.3)
        sites, coeffs, codes = H.instruction_arrays()
        c, s, co = O.terms_j1j2(olat, -1.0, -0.5, hz=0.3).arrays()
        np.testing.assert_array_equal(codes, c)
        np.testing.assert_array_equal(sites, s)
        np.testing.assert_array_equal(coeffs, co)
        H = XXZ(lat, hilbert_space=_dummy_space(lat), jxy=1.0, jz=0.7, hz=np.linspace(-1, 1, lat.ns))
        sites, coeffs, codes = H.instruction_arrays()
        c, s, co = O.terms_xxz(olat, 1.0, 0.7, hz=np.linspace(-1, 1, lat.ns)).arrays()
        np.testing.assert_array_equal(codes, c)
        np.testing.assert_array_equal(sites, s)
        np.testing.assert_array_equal(coeffs, co)
    lat, olat = HoneycombLattice(3, 2), O.OLattice("honeycomb", 3, 2)
    h = 0.05 / np.sqrt(3)
    H = HeisenbergKitaev(lat, K=-1.0, J=0.2, hilbert_space=_dummy_space(lat), hx=h, hy=h, hz=h)
    sites, coeffs, codes = H.instruction_arrays()
    c, s, co = O.terms_heisenberg_kitaev(olat, K=-1.0, J=0.2, hx=h, hy=h, hz=h).arrays()
    np.testing.assert_array_equal(codes, c)
    np.testing.assert_array_equal(sites, s)
    np.testing.assert_array_equal(coeffs, co)


def test_split_couplings_same_matrix_bruteforce():
    """HeisenbergKitaev(split_couplings=True) gives the same full-space
    matrix as the reference's merged instruction list."""
    from pyqusolver_amd.models import HeisenbergKitaev
    lat = HoneycombLattice(2, 2)
    kw = dict(K=-1.0, J=0.2, hx=0.02, hy=0.03, hz=0.04)
    A = HeisenbergKitaev(lat, hilbert_space=_dummy_space(lat), **kw)
    B = HeisenbergKitaev(lat, hilbert_space=_dummy_space(lat), split_couplings=True, **kw)
    sa, ca, ka = A.instruction_arrays()
    sb, cb, kb = B.instruction_arrays()
    Ha = O.full_dense_hamiltonian(8, ka, sa, ca, True)
    Hb = O.full_dense_hamiltonian(8, kb, sb, cb, True)
    np.testing.assert_allclose(Ha, Hb, atol=1e-14)


# ------------------------------------------------------------- inversion

@pytest.mark.parametrize("lat,olat", [(ChainLattice(9), O.OLattice("chain", 9)),
                                      (SquareLattice(4, 3), O.OLattice("square", 4, 3)),
                                      (HoneycombLattice(3, 2), O.OLattice("honeycomb", 3, 2))])
def test_inversion_permutation_and_characters_match_oracle(lat, olat):
    from pyqusolver_amd.symmetry import inversion_permutation
    np.testing.assert_array_equal(inversion_permutation(lat), O.inversion_perm(olat))
    tr = {"kx": 0, "ky": 0} if lat.ly > 1 else 0
    otr = [("tx", 0), ("ty", 0)] if lat.ly > 1 else [("tx", 0)]
    for sec in (1, -1):
        g = _grp(lat, {"translation": tr, "inversion": sec})
        og = O.Group(olat, [O.Generator(k, s) for k, s in otr + [("I", sec)]])
        assert g.order == og.n_group and g.matvec_error is None
        np.testing.assert_array_equal(g.chi, og.chi)
    assert _grp(lat, {"translation": {"kx": 1, "ky": 0} if lat.ly > 1 else 1, "inversion": 1}).matvec_error


# ------------------------------------------------------------- boundary flux

def test_bond_windings_and_flux_phases():
    """Forward bonds leaving the cell range wind -1 (negative step) or +1;
    with flux phi along d they carry exp(i w phi); a chain's closing bond is
    the only winding one; every honeycomb row is crossed by lx y bonds."""
    ch = ChainLattice(6, flux=np.pi)
    w = {(i, j): ch.bond_winding(i, j) for i, j, _ in ch.bonds(1)}
    assert [k for k, v in w.items() if v != (0, 0, 0)] == [(5, 0)] and w[(5, 0)] == (1, 0, 0)
    assert ch.has_flux and ch.boundary_phase_from_winding(1, 0, 0) == pytest.approx(-1.0)
    hc = HoneycombLattice(3, 2, flux={"y": 2 * np.pi / 3})
    ws = [hc.bond_winding(i, j) for i, j, _ in hc.bonds(1)]
    assert sum(1 for v in ws if v == (0, -1, 0)) == 3 and sum(1 for v in ws if v == (-1, 0, 0)) == 2
    assert hc.boundary_phase_from_winding(0, -1, 0) == pytest.approx(np.exp(-2j * np.pi / 3))
    assert hc.boundary_phase_from_winding(-1, 0, 0) == 1.0
    tab = hc.boundary_phases()
    assert tab.shape == (3, hc.ns + 1)
    np.testing.assert_allclose(tab[1, 2], np.exp(2j * 2 * np.pi / 3))
    assert not HoneycombLattice(2, 2).has_flux


def test_kitaev_flux_terms_match_oracle_and_are_hermitian():
    """HeisenbergKitaev with a boundary flux: forward bonds carry the
    Hermitian part of their phase (heisenberg_kitaev.py:487-571), so the
    reference's own check holds (tests/models/test_heisenberg_kitaev_flux.py:
    honeycomb 2x2, flux_y = 2 pi/3, J = 1, K = 0: H Hermitian); the
    instruction list equals the oracle's; fluxes phi, -phi and phi + 2 pi
    give the same matrix."""
    from pyqusolver_amd.models import HeisenbergKitaev
    lat, olat = HoneycombLattice(2, 2, flux={"x": 0.0, "y": 2 * np.pi / 3}), O.OLattice("honeycomb", 2, 2)
    H = HeisenbergKitaev(lat, K=(0.0, 0.0, 0.0), J=1.0, hilbert_space=_dummy_space(lat), dtype=np.complex128)
    s, c, k = H.instruction_arrays()
    Hf = O.full_dense_hamiltonian(8, k, s, c, True)
    assert np.linalg.norm(Hf - Hf.conj().T) < 1e-10
    oc, os_, oco = O.terms_heisenberg_kitaev(olat, K=0.0, J=1.0, flux=(0.0, 2 * np.pi / 3)).arrays()
    np.testing.assert_array_equal(k, oc)
    np.testing.assert_array_equal(s, os_)
    np.testing.assert_allclose(c, oco, rtol=1e-15, atol=0)
    ref = O.full_dense_hamiltonian(8, oc, os_, oco, True)
    for phi in (-2 * np.pi / 3, 2 * np.pi / 3 + 2 * np.pi):
        o2 = O.terms_heisenberg_kitaev(olat, K=0.0, J=1.0, flux=(0.0, phi)).arrays()
        np.testing.assert_allclose(O.full_dense_hamiltonian(8, *o2, True), ref, atol=1e-13)
    # the flux changes the spectrum (it is not a no-op)
    o0 = O.terms_heisenberg_kitaev(olat, K=0.0, J=1.0).arrays()
    assert abs(np.linalg.eigvalsh(ref)[0] - np.linalg.eigvalsh(O.full_dense_hamiltonian(8, *o0, True))[0]) > 1e-3


def test_flux_refusals():
    """A complex bond phase would make an XXZ / J1-J2 bond non-Hermitian;
    translation along a fluxed direction is not a symmetry of the couplings."""
    from pyqusolver_amd.models import XXZ, J1J2Model
    lat = ChainLattice(8, flux=0.3)
    with pytest.raises(ValueError):
        XXZ(lat, hilbert_space=_dummy_space(lat), jxy=1.0, jz=0.5)
    with pytest.raises(ValueError):
        J1J2Model(lat, J1=1.0, J2=0.0, hilbert_space=_dummy_space(lat))
    XXZ(ChainLattice(8, flux=np.pi), hilbert_space=_dummy_space(lat), jxy=1.0, jz=0.5)   # real phase -1
    with pytest.raises(NotImplementedError):
        _grp(ChainLattice(8, flux=np.pi), {"translation": 0})
    hc = HoneycombLattice(3, 2, flux={"y": 1.0})
    with pytest.raises(NotImplementedError):
        _grp(hc, {"translation": {"kx": 0, "ky": 0}})
    assert _grp(hc, {"translation": {"kx": 1}}).order == 3


def test_overlap_column_conventions():
    """Operator.overlap follows the reference's column rules
    (operator.py:926-973): vectors give a scalar, two blocks of equal width
    give pairwise overlaps (or the full product with pairwise=False), a
    vector against a block gives a row; numpy and torch alike."""
    import torch
    from pyqusolver_amd.operator import Hamiltonian
    H = Hamiltonian.__new__(Hamiltonian)
    rng = np.random.default_rng(0)
    a = rng.standard_normal((6, 3)) + 1j * rng.standard_normal((6, 3))
    b = rng.standard_normal((6, 3))
    np.testing.assert_allclose(H.overlap(a, b), np.einsum("ij,ij->j", a.conj(), b))
    np.testing.assert_allclose(H.overlap(a, b, pairwise=False), a.conj().T @ b)
    assert np.isclose(H.overlap(a[:, 0], b[:, 1]), np.vdot(a[:, 0], b[:, 1]))
    np.testing.assert_allclose(H.overlap(a[:, 0], b), a[:, 0].conj() @ b)
    np.testing.assert_allclose(H.overlap(torch.from_numpy(a), torch.from_numpy(b)).numpy(),
                               np.einsum("ij,ij->j", a.conj(), b))
    with pytest.raises(ValueError):
        H.overlap(a, b[:4])
    with pytest.raises(ValueError):
        H.overlap(a, rng.standard_normal((6, 2)))


class _ShuffledRing(ChainLattice):
    """A ring of L positions whose sites are numbered in a shuffled order
    (site s at position order[s]): its inversion through the ring's centre
    is an involution other than the full site reversal."""

    def __init__(self, L, order):
        super().__init__(L)
        self._order = np.asarray(order)

    @property
    def rvectors(self):
        out = np.zeros((self.ns, 3))
        out[:, 0] = self._order
        return out


def test_general_inversion_compiles_to_a_permutation_base_op():
    """An inversion that is not the full reversal becomes a permutation base
    element carrying its site permutation (applied to bits, as the reference
    does); the full reversal stays the reflection op."""
    from pyqusolver_amd.symmetry import BOP_PERM, BOP_REFLECT, inversion_permutation
    L = 10
    order = np.random.default_rng(13).permutation(L)
    lat = _ShuffledRing(L, order)
    perm = inversion_permutation(lat)
    at = np.argsort(order)
    np.testing.assert_array_equal(perm, at[L - 1 - order])
    grp = build_sector_group(lat, L, [(SymmetryGenerators.Inversion, -1)], [get_u1_sym(lat, 5)])
    assert grp.base_ops[1] == [BOP_PERM]
    np.testing.assert_array_equal(grp.perm, perm)
    assert np.allclose(grp.chi, [1, -1])
    g2 = build_sector_group(ChainLattice(L), L, [(SymmetryGenerators.Inversion, 1)], [])
    assert g2.base_ops[1] == [BOP_REFLECT] and g2.perm is None



class _Terms:
    """Instruction arrays of an oracle term list, in the Hamiltonian's form."""

    def __init__(self, t):
        self._codes, self._sites, self._coeffs = t.arrays()

    def instruction_arrays(self):
        return self._sites, self._coeffs, self._codes


def test_momentum_classes_follow_the_model_symmetry():
    """Point-group classes of momentum sectors: the 6x6 J1-J2 model (D4)
    has 10 classes of the 36 momenta; a site-dependent field leaves every
    momentum its own class; a 4x6 rectangle keeps the 4 reflections only; a
    chain pairs k with -k."""
    from pyqusolver_amd.batched import momentum_classes
    keys = [(kx, ky) for kx in range(6) for ky in range(6)]
    cl = momentum_classes(SquareLattice(6, 6), _Terms(O.terms_j1j2(O.OLattice("square", 6, 6), -1.0, -0.5)), keys)
    assert [len(m) for _, m in cl] == [1, 4, 4, 2, 4, 8, 4, 4, 4, 1]
    assert sorted(k for _, m in cl for k in m) == sorted(keys)
    assert dict(cl)[(1, 2)] == [(1, 2), (1, 4), (2, 1), (2, 5), (4, 1), (4, 5), (5, 2), (5, 4)]
    hz = np.random.default_rng(3).uniform(-1, 1, 36)
    cl = momentum_classes(SquareLattice(6, 6), _Terms(O.terms_j1j2(O.OLattice("square", 6, 6), -1.0, -0.5,
                                                                       hz=hz)), keys)
    assert len(cl) == 36
    keys46 = [(kx, ky) for kx in range(4) for ky in range(6)]
    cl = momentum_classes(SquareLattice(4, 6), _Terms(O.terms_j1j2(O.OLattice("square", 4, 6), -1.0, -0.5)), keys46)
    assert max(len(m) for _, m in cl) == 4 and len(cl) == 3 * 4
    cl = momentum_classes(ChainLattice(8), _Terms(O.terms_xxz(O.OLattice("chain", 8), 1.0, 0.7)), [(k,) for k in range(8)])
    assert [m for _, m in cl] == [[(0,)], [(1,), (7,)], [(2,), (6,)], [(3,), (5,)], [(4,)]]


def test_momentum_classes_of_the_model_classes():
    """The same classes from the J1-J2 model class's own instruction list
    (the sweep's caller), with and without a uniform field."""
    from pyqusolver_amd.batched import momentum_classes
    from pyqusolver_amd.models import J1J2Model
    lat = SquareLattice(6, 6)
    keys = [(kx, ky) for kx in range(6) for ky in range(6)]
    for hz in (None, 0.3):
        H = J1J2Model(lat, J1=-1.0, J2=-0.5, hilbert_space=_dummy_space(lat), hz=hz)
        cl = momentum_classes(lat, H, keys)
        assert [len(m) for _, m in cl] == [1, 4, 4, 2, 4, 8, 4, 4, 4, 1]


def test_momentum_class_members_have_equal_spectra():
    """The equivalence the classes rest on: on the 4x4 J1-J2 model at
    half filling, every momentum sector of a class has the same spectrum
    (dense blocks of the oracle), and sectors of different classes differ."""
    from pyqusolver_amd.batched import momentum_classes
    olat = O.OLattice("square", 4, 4)
    t = O.terms_j1j2(olat, -1.0, -0.45)
    c, s, co = t.arrays()
    keys = [(kx, ky) for kx in range(4) for ky in range(4)]
    spec = {}
    for k in keys:
        grp = O.Group(olat, [O.Generator("tx", k[0]), O.Generator("ty", k[1])], n_up=8)
        b = O.scan_basis(grp)
        spec[k] = np.linalg.eigvalsh(O.reduced_dense(b, 16, c, s, co, True))
    cl = momentum_classes(SquareLattice(4, 4), _Terms(t), keys)
    reps = []
    for rep, members in cl:
        for m in members:
            np.testing.assert_allclose(spec[m], spec[rep], atol=1e-11)
        reps.append(rep)
    for i, a in enumerate(reps):
        for b_ in reps[i + 1:]:
            assert len(spec[a]) != len(spec[b_]) or not np.allclose(spec[a], spec[b_], atol=1e-8)
