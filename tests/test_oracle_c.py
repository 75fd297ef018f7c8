"""The C restatement (oracle/qes_oracle.c) against the NumPy oracle. CPU only."""
import numpy as np
import pytest

from oracle import qes_oracle as O
from oracle import qes_oracle_c as OC


@pytest.mark.parametrize("kind,lx,ly,gens,n_up,model", [
    ("chain", 14, 1, [("tx", 0), ("Px", 1)], 7, "xxz"),
    ("chain", 14, 1, [("tx", 3)], 7, "xxz"),
    ("chain", 14, 1, [("tx", 7), ("R", -1), ("Px", -1)], 7, "j1j2"),
    ("square", 4, 3, [("tx", 1), ("ty", 2)], 6, "j1j2"),
    ("honeycomb", 3, 2, [("tx", 1), ("ty", 0)], None, "hk"),
])
def test_c_oracle_matches_numpy_oracle(kind, lx, ly, gens, n_up, model):
    lat = O.OLattice(kind, lx, ly)
    grp = O.Group(lat, [O.Generator(k, s) for k, s in gens], n_up=n_up)
    b = O.scan_basis(grp)
    t = {"xxz": lambda: O.terms_xxz(lat, 1.0, 0.7),
         "j1j2": lambda: O.terms_j1j2(lat, -1.0, -0.5),
         "hk": lambda: O.terms_heisenberg_kitaev(lat, K=-1.0, J=0.2, hx=0.03, hy=0.02, hz=0.05)}[model]()
    c, s, co = t.arrays()
    rng = np.random.default_rng(3)
    x = rng.standard_normal(len(b.reps)) + 1j * rng.standard_normal(len(b.reps))
    y1 = O.push_matvec(b, lat.ns, c, s, co, t.is_complex, x)
    y2, _ = OC.push_matvec(grp, b.reps, b.norms, c, s, co, t.is_complex, x)
    assert np.linalg.norm(y1 - y2) <= 1e-13 * np.linalg.norm(y1)


@pytest.mark.parametrize("kind,lx,ly,gens,n_up", [
    ("chain", 14, 1, [("tx", 0), ("Px", 1)], 7),
    ("chain", 15, 1, [("tx", 4)], 7),
    ("chain", 12, 1, [("tx", 6), ("R", -1), ("Px", -1)], 6),
    ("square", 4, 3, [("tx", 1), ("ty", 2)], 6),
    ("square", 4, 4, [("tx", 0), ("ty", 0), ("Px", 1)], 8),
])
def test_c_window_scan_matches_numpy_scan(kind, lx, ly, gens, n_up):
    """The C rank-window scan (composite element permutations) equals the
    NumPy orbit scan over the whole filling and over sub-windows."""
    from math import comb
    lat = O.OLattice(kind, lx, ly)
    grp = O.Group(lat, [O.Generator(k, s) for k, s in gens], n_up=n_up)
    b = O.scan_basis(grp)
    total = comb(lat.ns, n_up)
    reps, norms = OC.scan_window(grp, 0, total, threads=3)
    np.testing.assert_array_equal(reps, b.reps)
    np.testing.assert_allclose(norms, b.norms, rtol=1e-15)
    r0, cnt = total // 3, total // 4
    lo, hi = np.searchsorted(b.reps, _unrank(lat.ns, n_up, r0)), np.searchsorted(b.reps, _unrank(lat.ns, n_up, r0 + cnt))
    reps_w, _ = OC.scan_window(grp, r0, cnt)
    np.testing.assert_array_equal(reps_w, b.reps[lo:hi])


def _unrank(ns, k, r):
    """State of colex rank r among the k-subsets (brute force over the sorted filling)."""
    from itertools import combinations
    states = sorted(sum(1 << i for i in c) for c in combinations(range(ns), k))
    return states[r] if r < len(states) else 1 << ns
