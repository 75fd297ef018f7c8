"""Write tests/golden/golden.npz: small fixtures from the CPU oracle.

The reference cannot be imported (its hot-path modules need numba, which is
absent), so the fixtures come from the oracle in oracle/qes_oracle.py after
it has been pinned by the reference's own known answers, closed-form sector
dimensions and energies, and brute-force projections (tests/test_oracle_pins.py).
They freeze those outputs: representatives and normalisations (bit-exact),
H x for a seeded x, and the lowest eigenvalue of the dense reduced matrix.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import qes_oracle as O  # noqa: E402

# (name, lattice kind, lx, ly, generators, n_up, model)
CASES = [
    ("chain16_heis_k0", "chain", 16, 1, [("tx", 0)], 8, "heis"),
    ("chain12_xxz_k3", "chain", 12, 1, [("tx", 3)], 6, "xxz"),
    ("chain14_j1j2_k0_R1_Pxm", "chain", 14, 1, [("tx", 0), ("R", 1), ("Px", -1)], 7, "j1j2"),
    ("chain12_j1j2field_k4_Px", "chain", 12, 1, [("tx", 4), ("Px", 1)], None, "j1j2_field"),
    ("square4x3_j1j2_k1_2", "square", 4, 3, [("tx", 1), ("ty", 2)], 6, "j1j2"),
    ("honeycomb3x2_hk_k0", "honeycomb", 3, 2, [("tx", 0), ("ty", 0)], None, "hk"),
    ("square4x4_j1j2_k0_2_Im", "square", 4, 4, [("tx", 0), ("ty", 2), ("I", -1)], 8, "j1j2"),
    ("honeycomb3x2_hk_k0_Ip", "honeycomb", 3, 2, [("tx", 0), ("ty", 0), ("I", 1)], None, "hk"),
    ("chain14_j1j2_kpi_Im_Px", "chain", 14, 1, [("tx", 7), ("I", -1), ("Px", 1)], 7, "j1j2"),
]


def terms(model, olat):
    if model == "heis":
        return O.terms_j1j2(olat, -1.0, 0.0)
    if model == "xxz":
        return O.terms_xxz(olat, 1.0, 0.7, hz=0.1)
    if model == "j1j2":
        return O.terms_j1j2(olat, -1.0, -0.5)
    if model == "j1j2_field":
        return O.terms_j1j2(olat, -1.0, -0.3, hx=0.2, hz=0.1)
    h = 0.05 / np.sqrt(3)
    return O.terms_heisenberg_kitaev(olat, K=-1.0, J=0.2, hx=h, hy=h, hz=h)


def main():
    out = {}
    meta = []
    for i, (name, kind, lx, ly, gens, n_up, model) in enumerate(CASES):
        olat = O.OLattice(kind, lx, ly)
        grp = O.Group(olat, [O.Generator(k, s) for k, s in gens], n_up=n_up)
        b = O.scan_basis(grp)
        t = terms(model, olat)
        c, s, co = t.arrays()
        rng = np.random.default_rng(100 + i)
        n = len(b.reps)
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        y = O.push_matvec(b, olat.ns, c, s, co, t.is_complex, x)
        e0 = np.linalg.eigvalsh(O.reduced_dense(b, olat.ns, c, s, co, t.is_complex))[0]
        out[f"{name}/reps"] = b.reps.astype(np.uint64)
        out[f"{name}/norms"] = b.norms
        out[f"{name}/x"] = x
        out[f"{name}/y"] = y
        out[f"{name}/e0"] = np.array([e0])
        meta.append({"name": name, "kind": kind, "lx": lx, "ly": ly, "gens": gens, "n_up": n_up, "model": model,
                     "n": int(n), "seed": 100 + i})
        print(f"{name}: N={n} E0={e0:.14f}")
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    with open(os.path.join(HERE, "golden_cases.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main()
