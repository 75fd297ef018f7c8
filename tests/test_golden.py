"""Committed golden fixtures (tests/golden/golden.npz, written by
tests/golden/make_golden.py from the pinned CPU oracle): the oracle must keep
reproducing them, and the config-1 energy must equal the independently known
value E0 = -7.14229636061679 (AF Heisenberg ring L=16, Sz=0, k=0)."""
import json
import os

import numpy as np
import pytest

from oracle import qes_oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_golden():
    data = np.load(os.path.join(HERE, "golden.npz"))
    with open(os.path.join(HERE, "golden_cases.json")) as f:
        cases = json.load(f)
    return data, cases


def _terms(model, olat):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.terms(model, olat)


DATA, CASES = load_golden()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_reproduces_golden(case):
    name = case["name"]
    olat = O.OLattice(case["kind"], case["lx"], case["ly"])
    grp = O.Group(olat, [O.Generator(k, s) for k, s in case["gens"]], n_up=case["n_up"])
    b = O.scan_basis(grp)
    np.testing.assert_array_equal(b.reps.astype(np.uint64), DATA[f"{name}/reps"])
    np.testing.assert_array_equal(b.norms, DATA[f"{name}/norms"])
    t = _terms(case["model"], olat)
    c, s, co = t.arrays()
    y = O.push_matvec(b, olat.ns, c, s, co, t.is_complex, DATA[f"{name}/x"])
    yg = DATA[f"{name}/y"]
    assert np.linalg.norm(y - yg) <= 1e-14 * np.linalg.norm(yg)


def test_golden_config1_energy():
    e0 = DATA["chain16_heis_k0/e0"][0]
    assert len(DATA["chain16_heis_k0/reps"]) == 810
    assert abs(e0 - (-7.14229636061679)) <= 1e-12 * 7.2
