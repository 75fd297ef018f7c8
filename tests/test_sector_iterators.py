"""Sector enumeration of HilbertSpace.iter_*_sectors (Algebra/hilbert.py:412-993):
ordering, the parity handling at generic momenta, and completeness -- the
sector dimensions (character formula of the oracle) add up to the dimension
of the U(1) filling.  CPU only: the specs are built without a device."""
from math import comb

import pytest

from oracle import qes_oracle as O
from pyqusolver_amd import ChainLattice, HoneycombLattice, SquareLattice
from pyqusolver_amd.hilbert import HilbertSpace, momentum_key, momentum_symmetry_types, sector_specs
from pyqusolver_amd.symmetry import SymmetryGenerators as SG

_KIND = {SG.Translation_x: "tx", SG.Translation_y: "ty", SG.Translation_z: "tz", SG.Reflection: "R",
         SG.Inversion: "I", SG.ParityX: "Px", SG.ParityY: "Py", SG.ParityZ: "Pz"}


def _olat(lat):
    return O.OLattice(lat.kind, lat.lx, lat.ly)


def _dim(lat, sym_gen, n_up):
    return O.sector_dimension(O.Group(_olat(lat), [O.Generator(_KIND[g], v) for g, v in sym_gen], n_up=n_up))


@pytest.mark.parametrize("lat,n_up", [(ChainLattice(10), 5), (ChainLattice(9), 4), (SquareLattice(4, 4), 8),
                                      (SquareLattice(3, 4), 6), (HoneycombLattice(2, 3), 6)])
@pytest.mark.parametrize("parity", [False, True])
def test_momentum_sectors_partition_the_filling(lat, n_up, parity):
    specs = sector_specs(momentum_symmetry_types(lat, parity), lat)
    assert sum(_dim(lat, sg, n_up) for _, sg in specs) == comb(lat.ns, n_up)


@pytest.mark.parametrize("lat,n_up", [(ChainLattice(8), 4), (SquareLattice(4, 2), 4)])
def test_all_sectors_with_spin_flip_partition_the_filling(lat, n_up):
    types, names = HilbertSpace.all_sector_types(lat, ["translation", "reflection", "parity_x"])
    specs = sector_specs(types, lat)
    assert set(names.values()) >= {"kx", "reflection", "parity_x"}
    assert sum(_dim(lat, sg, n_up) for _, sg in specs) == comb(lat.ns, n_up)


def test_reference_order_and_parity_handling():
    lat = ChainLattice(24)
    specs = sector_specs(momentum_symmetry_types(lat, True), lat)
    keys = [momentum_key(lat, s, True) for s, _ in specs]
    # itertools.product order: k outer, parity inner; parities only at k in {0, 12}
    assert keys[:3] == [(0, 1), (0, -1), (1, 0)]
    assert (12, 1) in keys and (12, -1) in keys and len(keys) == 26
    assert all(p == 0 for k, p in keys if k not in (0, 12))
    # the reference's literal product keeps both parities everywhere
    lit = sector_specs(momentum_symmetry_types(lat, True), lat, physical_only=False)
    assert len(lit) == 48
    # 6x6: 36 momenta; with reflection (= inversion on the torus) four real momenta doubled
    sq = SquareLattice(6, 6)
    assert len(sector_specs(momentum_symmetry_types(sq, False), sq)) == 36
    assert len(sector_specs(momentum_symmetry_types(sq, True), sq)) == 40


def test_sector_specs_feed_the_batched_helper():
    from pyqusolver_amd.batched import momentum_sectors
    specs = momentum_sectors(24, reflection=True)
    assert len(specs) == 26 and specs[0] == {"translation": 0, "reflection": 1}
    assert specs[2] == {"translation": 1}


@pytest.mark.gpu
@pytest.mark.parametrize("lat,n_up,parity", [(ChainLattice(12), 6, True), (SquareLattice(4, 4), 8, True),
                                             (HoneycombLattice(2, 3), 6, False)])
def test_iter_momentum_sectors_on_device_partition_the_filling(lat, n_up, parity):
    from pyqusolver_amd import get_u1_sym
    total, keys = 0, []
    for key, hs in HilbertSpace.iter_momentum_sectors(lat, include_parity=parity, global_syms=[get_u1_sym(lat, n_up)]):
        keys.append(key)
        sg = [(SG.Translation_x, key[0])] + ([(SG.Translation_y, key[1])] if lat.ly > 1 else [])
        if parity and key[-1] != 0:
            sg.append((SG.Reflection, key[-1]))
        assert hs.nh == _dim(lat, sg, n_up), key
        total += hs.nh
    assert total == comb(lat.ns, n_up)
    assert len(keys) == len(set(keys))


@pytest.mark.gpu
def test_iter_all_sectors_names_and_completeness():
    from pyqusolver_amd import get_u1_sym
    lat = ChainLattice(10)
    total = 0
    for sec, hs in HilbertSpace.iter_all_sectors(lat, ["translation", "reflection", "parity_x"],
                                                 global_syms=[get_u1_sym(lat, 5)]):
        assert set(sec) == {"kx", "reflection", "parity_x"}
        total += hs.nh
    assert total == comb(10, 5)
