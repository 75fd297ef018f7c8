"""Pins of the CPU oracle (oracle/qes_oracle.py) against the reference's own
known answers and against mathematics that fixes the result independently.
CPU only."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import qes_oracle as O


# ---------------------------------------------------------------- VM KATs
def _one(state, ns, code, s1, s2=-1):
    st, c, diag = O._single(np.array([state], dtype=np.int64), ns, code, s1, s2)
    return int(st[0]), complex(c[0]), diag


def test_reference_kernel_known_answers():
    # tests/algebra/test_custom_hamiltonian_workflows.py:76-92
    st, c, _ = _one(0b0010, 4, O.SIG_X_L, 1)
    assert st == 0b0110 and c == 0.5
    st, c, d = _one(0b0010, 4, O.SIG_Z_L, 1)
    assert st == 0b0010 and c == -0.5 and d


def test_reference_int_vs_array_convention():
    # tests/algebra/test_custom_hamiltonian_workflows.py:101-126: bit 1 <-> +0.5
    for state, arr in ((0, -0.5), (1, 0.5)):
        sx, cx, _ = _one(state, 1, O.SIG_X_L, 0)
        sz, cz, _ = _one(state, 1, O.SIG_Z_L, 0)
        sy, cy, _ = _one(state, 1, O.SIG_Y_L, 0)
        assert sx == 1 - state and cx == 0.5
        assert sz == state and cz == arr
        assert sy == 1 - state and cy == 1j * (2 * state - 1) * 0.5


@pytest.mark.parametrize("code", [O.SIG_X_L, O.SIG_Y_L, O.SIG_Z_L, O.SIG_P_L, O.SIG_M_L, O.SIG_X_C, O.SIG_Y_C,
                                  O.SIG_Z_C, O.SIG_P_C, O.SIG_M_C, O.SIG_XY_C, O.SIG_YX_C, O.SIG_XZ_C,
                                  O.SIG_ZX_C, O.SIG_YZ_C, O.SIG_ZY_C, O.SIG_PM_C, O.SIG_MP_C])
@pytest.mark.parametrize("is_complex", [True, False])
def test_vm_matches_kronecker_brute_force(code, is_complex):
    ns = 4
    sites = [[1, 3]] if code >= 20 else [[2, -1]]
    coeffs = np.array([0.7 - 0.3j if is_complex else 0.7])
    codes = np.array([code])
    Hf = O.full_dense_hamiltonian(ns, codes, sites, coeffs, is_complex)
    # apply the VM to every basis state and assemble columns
    states = np.arange(1 << ns, dtype=np.int64)
    H = np.zeros((1 << ns, 1 << ns), dtype=np.complex128)
    for st, val, keep in O.vm_apply(states, ns, codes, sites, coeffs, is_complex):
        for k in np.nonzero(keep)[0]:
            H[st[k], states[k]] += val[k]
    np.testing.assert_allclose(H, Hf, atol=1e-15)


# ---------------------------------------------------------- group / orbits
def test_reference_orbit_known_answers():
    # tests/algebra/test_momentum_sectors.py:38-62: L=4, rep 3 has orbit {3,6,9,12}
    grp = O.Group(O.OLattice("chain", 4), [O.Generator("tx", 0)])
    orbit = {int(grp.apply(np.array([3]), g)[0][0]) for g in range(grp.n_group)}
    assert orbit == {3, 6, 9, 12}
    orbit1 = {int(grp.apply(np.array([1]), g)[0][0]) for g in range(grp.n_group)}
    assert len(orbit1) == 4
    # allowed k of a period-4 orbit: all four momenta have norm > 0
    for k in range(4):
        g = O.Group(O.OLattice("chain", 4), [O.Generator("tx", k)])
        assert O.normalization(g, np.array([3]))[0] > 0


def test_reference_2d_periods():
    # tests/algebra/test_momentum_sectors.py:64-80: 2x2, rep 1 has periods (2, 2)
    lat = O.OLattice("square", 2, 2)
    for kx in range(2):
        for ky in range(2):
            g = O.Group(lat, [O.Generator("tx", kx), O.Generator("ty", ky)])
            assert O.normalization(g, np.array([1]))[0] > 0


CASES = [
    ("chain", 10, 1, [("tx", 0)], 5),
    ("chain", 10, 1, [("tx", 3)], 5),
    ("chain", 10, 1, [("tx", 5), ("R", -1)], 5),
    ("chain", 10, 1, [("tx", 0), ("R", 1), ("Px", -1)], 5),
    ("chain", 9, 1, [("tx", 2)], 4),
    ("chain", 8, 1, [("tx", 0), ("Px", 1)], None),
    ("chain", 8, 1, [("tx", 4), ("R", 1), ("Px", 1)], None),
    ("square", 4, 2, [("tx", 1), ("ty", 1), ("Px", -1)], 4),
    ("square", 3, 3, [("tx", 1), ("ty", 2)], None),
    ("honeycomb", 2, 2, [("tx", 1), ("ty", 0)], None),
    ("honeycomb", 3, 2, [("tx", 0), ("ty", 1)], 6),
    ("chain", 10, 1, [("tx", 5), ("I", -1), ("Px", 1)], 5),
    ("square", 4, 2, [("tx", 0), ("ty", 1), ("I", 1)], 4),
    ("square", 3, 3, [("tx", 0), ("ty", 0), ("I", -1)], None),
    ("honeycomb", 2, 2, [("tx", 0), ("ty", 0), ("I", -1)], None),
    ("honeycomb", 3, 2, [("tx", 0), ("ty", 0), ("I", 1)], 6),
]


@pytest.mark.parametrize("kind,lx,ly,gens,z2", [("chain", 10, 1, [("tx", 0), ("R", 1)], 1),
                                                ("chain", 10, 1, [("tx", 3)], -1),
                                                ("chain", 9, 1, [("tx", 2)], 1),
                                                ("chain", 8, 1, [("tx", 0), ("Px", 1)], 1),
                                                ("square", 4, 2, [("tx", 1), ("ty", 0)], -1),
                                                ("honeycomb", 2, 2, [("tx", 0), ("ty", 1)], 1)])
def test_z2_scan_dimension_matches_character_formula(kind, lx, ly, gens, z2):
    """Z2 popcount-parity sectors without U(1): the orbit scan's size equals
    the character formula restricted to fixed states of that popcount parity,
    the two parities add up to the unfiltered sector, and every
    representative has the requested parity."""
    lat = O.OLattice(kind, lx, ly)
    gs = [O.Generator(k, s) for k, s in gens]
    grp = O.Group(lat, gs, z2=z2)
    b = O.scan_basis(grp)
    assert len(b.reps) == O.sector_dimension(grp)
    par = np.array([bin(int(r)).count("1") % 2 for r in b.reps])
    assert np.all(np.where(par == 0, 1, -1) == z2)
    other = O.sector_dimension(O.Group(lat, gs, z2=-z2))
    assert len(b.reps) + other == O.sector_dimension(O.Group(lat, gs))
    assert O.sector_dimension(O.Group(lat, gs, n_up=3, z2=1)) == 0          # odd popcount, even parity
    assert O.sector_dimension(O.Group(lat, gs, n_up=3, z2=-1)) == O.sector_dimension(O.Group(lat, gs, n_up=3))


@pytest.mark.parametrize("kind,lx,ly,gens,n_up", CASES)
def test_scan_dimension_matches_character_formula(kind, lx, ly, gens, n_up):
    lat = O.OLattice(kind, lx, ly)
    grp = O.Group(lat, [O.Generator(k, s) for k, s in gens], n_up=n_up)
    b = O.scan_basis(grp)
    assert len(b.reps) == O.sector_dimension(grp)
    assert np.all(np.diff(b.reps) > 0)
    # real sectors: N_r^2 = |Stab(r)| (orbit-stabiliser)
    if np.all(np.abs(grp.chi.imag) < 1e-12):
        _, _, cnt = O.canonicalize(grp, b.reps)
        np.testing.assert_allclose(b.norms ** 2, cnt, rtol=1e-13)


@pytest.mark.parametrize("kind,lx,ly,gens,n_up", CASES)
def test_push_matvec_equals_projected_brute_force(kind, lx, ly, gens, n_up):
    lat = O.OLattice(kind, lx, ly)
    grp = O.Group(lat, [O.Generator(k, s) for k, s in gens], n_up=n_up)
    b = O.scan_basis(grp)
    if kind == "honeycomb":
        t = O.terms_heisenberg_kitaev(lat, K=-1.0, J=0.2, hx=0.03, hy=0.02, hz=0.05)
    elif kind == "square":
        t = O.terms_j1j2(lat, -1.0, -0.5)
    else:
        t = O.terms_xxz(lat, 1.0, 0.7, hz=0.0)
    c, s, co = t.arrays()
    H = O.reduced_dense(b, lat.ns, c, s, co, t.is_complex)
    U = O.rotation_matrix(grp, b)
    Hf = O.full_dense_hamiltonian(lat.ns, c, s, co, t.is_complex)
    np.testing.assert_allclose(H, U.conj().T @ Hf @ U, atol=1e-13)
    np.testing.assert_allclose(H, H.conj().T, atol=1e-13)


def test_union_of_momentum_sectors_is_full_spectrum():
    lat = O.OLattice("chain", 10)
    t = O.terms_xxz(lat, 1.0, 0.7)
    c, s, co = t.arrays()
    Hf = O.full_dense_hamiltonian(10, c, s, co, True)
    mask = np.array([bin(i).count("1") == 5 for i in range(1 << 10)])
    Ef = np.linalg.eigvalsh(Hf[np.ix_(mask, mask)])
    E = []
    for k in range(10):
        grp = O.Group(lat, [O.Generator("tx", k)], n_up=5)
        b = O.scan_basis(grp)
        E += list(np.linalg.eigvalsh(O.reduced_dense(b, 10, c, s, co, True)))
    np.testing.assert_allclose(np.sort(E), Ef, atol=1e-12)


# ------------------------------------------------------------- models / E0
def test_heisenberg_ring_n4_e0_is_minus_two():
    # examples/algebra/example_exact_diagonalization.py:87-97 (H = sum S.S, N=4)
    lat = O.OLattice("chain", 4)
    t = O.terms_j1j2(lat, J1=-1.0, J2=0.0)
    c, s, co = t.arrays()
    Hf = O.full_dense_hamiltonian(4, c, s, co, False)
    assert abs(np.linalg.eigvalsh(Hf)[0] - (-2.0)) < 1e-13


def test_ferromagnet_closed_form():
    # literal J1=+1 through J1J2Model is a ferromagnet: E0 = -n_bonds/4
    lat = O.OLattice("chain", 10)
    grp = O.Group(lat, [O.Generator("tx", 0)], n_up=5)
    b = O.scan_basis(grp)
    t = O.terms_j1j2(lat, J1=1.0, J2=0.0)
    c, s, co = t.arrays()
    e0 = np.linalg.eigvalsh(O.reduced_dense(b, 10, c, s, co, False))[0]
    assert abs(e0 - (-10 / 4)) < 1e-12


def test_chain16_e0_against_full_space_sparse():
    """AF Heisenberg chain L=16, Sz=0, k=0: the sector E0 equals the full
    Sz=0 space E0 (independent sparse Kronecker construction)."""
    lat = O.OLattice("chain", 16)
    t = O.terms_j1j2(lat, J1=-1.0, J2=0.0)
    c, s, co = t.arrays()
    grp = O.Group(lat, [O.Generator("tx", 0)], n_up=8)
    b = O.scan_basis(grp)
    assert len(b.reps) == 810
    e_sector = np.linalg.eigvalsh(O.reduced_dense(b, 16, c, s, co, False))[0]
    Hs = O.full_sparse_hamiltonian(16, c, s, co, False).real
    mask = np.array([bin(i).count("1") == 8 for i in range(1 << 16)])
    idx = np.nonzero(mask)[0]
    Hsz = Hs[idx][:, idx]
    e_full = spla.eigsh(Hsz, k=1, which="SA", tol=1e-14)[0][0]
    assert abs(e_sector - e_full) < 1e-11
    assert abs(e_sector - (-7.14229636061679)) < 1e-11


def test_workload_sector_dimensions_closed_form():
    """Closed-form sector dimensions of the benchmark configurations."""
    def dim(kind, lx, ly, gens, n_up):
        return O.sector_dimension(O.Group(O.OLattice(kind, lx, ly), [O.Generator(k, s) for k, s in gens], n_up=n_up))
    assert dim("chain", 16, 1, [("tx", 0)], 8) == 810
    assert dim("chain", 32, 1, [("tx", 0), ("Px", 1)], 16) == 9_393_119
    assert dim("chain", 32, 1, [("tx", 0), ("Px", -1)], 16) == 9_391_051
    assert dim("chain", 32, 1, [("tx", 0)], 16) == 18_784_170
    assert dim("chain", 24, 1, [("tx", 0), ("R", 1)], 12) == 56_822
    assert dim("chain", 24, 1, [("tx", 0), ("R", -1)], 12) == 55_898
    assert dim("chain", 24, 1, [("tx", 12), ("R", 1)], 12) == 56_360
    tot = sum(dim("chain", 24, 1, [("tx", k)], 12) for k in range(24))
    assert tot == 2_704_156
    assert dim("honeycomb", 3, 4, [("tx", 0), ("ty", 0)], None) == 1_398_500
    assert dim("square", 6, 6, [("tx", 0), ("ty", 0), ("Px", 1)], 18) == 126_056_625


def test_expand_and_transform_to_reduced_closed_forms():
    """Oracle expand_state is an isometry whose columns are the rotation
    matrix (pinned by U^dagger H U above); transform_to_reduced composed with
    it is the diagonal sqrt|G| / N_k^2 (the reference's reduction divides by
    N_k once, with no 1/sqrt|G|)."""
    lat = O.OLattice("chain", 10)
    grp = O.Group(lat, [O.Generator("tx", 3)], n_up=5)
    b = O.scan_basis(grp)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(len(b.reps)) + 1j * rng.standard_normal(len(b.reps))
    full = O.expand_state(grp, b, x)
    np.testing.assert_allclose(np.linalg.norm(full), np.linalg.norm(x), rtol=1e-13)
    np.testing.assert_allclose(full, O.rotation_matrix(grp, b) @ x, rtol=0, atol=1e-14)
    back = O.transform_to_reduced(b, 10, full)
    np.testing.assert_allclose(back, x * np.sqrt(grp.n_group) / b.norms ** 2, rtol=1e-12)


def test_projected_matvec_is_projected_operator():
    """The restated projected kernel equals U_out^dagger O U_in built from the
    pinned rotation matrices and the Kronecker full-space operator, and in
    its own sector it is the ordinary push-form matvec."""
    L = 8
    lat = O.OLattice("chain", L)
    for kin, kout, nin, nout, code in ((0, 3, 4, 4, 13), (1, 2, 4, 5, 14)):
        gi = O.Group(lat, [O.Generator("tx", kin)], n_up=nin)
        go = O.Group(lat, [O.Generator("tx", kout)], n_up=nout)
        bi, bo = O.scan_basis(gi), O.scan_basis(go)
        q = 2 * np.pi * (kout - kin) / L
        codes, sites = [code] * L, [[j, -1] for j in range(L)]
        coeffs = [np.exp(1j * q * j) for j in range(L)]
        x = np.random.default_rng(kin).standard_normal(len(bi.reps)) + 0.5j
        y = O.projected_matvec(gi, bi, go, bo, L, codes, sites, coeffs, True, x)
        Of = O.full_dense_hamiltonian(L, codes, sites, coeffs, True)
        yb = O.rotation_matrix(go, bo).conj().T @ Of @ O.rotation_matrix(gi, bi) @ x
        np.testing.assert_allclose(y, yb, rtol=0, atol=1e-13 * max(1.0, np.abs(yb).max()))
    gi = O.Group(lat, [O.Generator("tx", 2)], n_up=4)
    bi = O.scan_basis(gi)
    c, s, co = O.terms_xxz(lat, 1.0, 0.7).arrays()
    x = np.random.default_rng(1).standard_normal(len(bi.reps)) + 0j
    np.testing.assert_allclose(O.projected_matvec(gi, bi, gi, bi, L, c, s, co, True, x),
                               O.push_matvec(bi, L, c, s, co, True, x), rtol=0, atol=1e-13)


# ------------------------------------------------------------- inversion

@pytest.mark.parametrize("kind,lx,ly", [("chain", 9, 1), ("chain", 10, 1), ("square", 4, 3), ("square", 3, 3),
                                        ("honeycomb", 2, 3), ("honeycomb", 3, 3)])
def test_inversion_perm_is_a_point_inversion_and_lattice_automorphism(kind, lx, ly):
    """inversion_perm (inversion.py:135-230): an involution; r_i + r_perm[i]
    equals twice the bounding-box centre up to lattice periods; it maps the
    bond set onto itself with each honeycomb bond keeping its x/y/z type;
    a chain reduces to the closed form i -> L-1-i."""
    lat = O.OLattice(kind, lx, ly)
    p = O.inversion_perm(lat)
    ns = lat.ns
    np.testing.assert_array_equal(p[p], np.arange(ns))
    a1, a2, a3, basis, fracs, subs, r = lat.geometry()
    r0 = 0.5 * (r.min(axis=0) + r.max(axis=0))
    A = np.column_stack([lx * a1, ly * a2, a3])
    for i in range(ns):
        d = np.linalg.solve(A, r[i] + r[p[i]] - 2 * r0)
        np.testing.assert_allclose(d[:2], np.rint(d[:2]), atol=1e-9)
    orders = (1,) if kind == "honeycomb" else (1, 2)
    for order in orders:
        bonds = O.oracle_bonds(lat, order)
        typed = {(min(i, j), max(i, j)): n for i, j, n in bonds}
        for i, j, n in bonds:
            key = (min(p[i], p[j]), max(p[i], p[j]))
            assert key in typed
            if kind == "honeycomb":
                assert typed[key] == n
    if kind == "chain":
        np.testing.assert_array_equal(p, np.arange(ns)[::-1])


@pytest.mark.parametrize("kind,lx,ly,n_up", [("square", 4, 2, 4), ("honeycomb", 2, 2, None), ("chain", 10, 1, 5)])
def test_inversion_sectors_split_the_momentum_sector(kind, lx, ly, n_up):
    """The k = 0 block's spectrum is the union of the I = +1 and I = -1
    blocks (dimensions add up, eigenvalues match)."""
    lat = O.OLattice(kind, lx, ly)
    if kind == "honeycomb":
        t = O.terms_heisenberg_kitaev(lat, K=-1.0, J=0.2, hx=0.03, hy=0.02, hz=0.05)
    elif kind == "square":
        t = O.terms_j1j2(lat, -1.0, -0.4)
    else:
        t = O.terms_xxz(lat, 1.0, 0.7, hz=0.0)
    c, s, co = t.arrays()
    tr = [("tx", 0)] + ([("ty", 0)] if kind != "chain" else [])

    def block(gens):
        grp = O.Group(lat, [O.Generator(k, v) for k, v in gens], n_up=n_up)
        b = O.scan_basis(grp)
        return np.linalg.eigvalsh(O.reduced_dense(b, lat.ns, c, s, co, t.is_complex)) if len(b.reps) else np.zeros(0)
    e_all = block(tr)
    e_pm = np.concatenate([block(tr + [("I", 1)]), block(tr + [("I", -1)])])
    assert len(e_pm) == len(e_all)
    np.testing.assert_allclose(np.sort(e_pm), e_all, atol=1e-12)



def perm_symmetric_bonds(ns, perm, rng, nb=8):
    """Random bonds (VM sites) closed under the bit-level action of an
    involution perm: the reference moves bit i to bit perm[i]
    (inversion.py:53-86) and the spin VM reads site s on bit ns-1-s, so a
    bond (s, t) maps to (ns-1-perm[ns-1-s], ns-1-perm[ns-1-t]).  Each bond
    and its image carry one coupling (ZZ, XX, YY)."""
    img = lambda s: ns - 1 - int(perm[ns - 1 - s])
    out = {}
    while len(out) < nb:
        s, t = sorted(int(v) for v in rng.choice(ns, 2, replace=False))
        if (s, t) in out:
            continue
        jz, jxy = rng.uniform(0.3, 1.2), rng.uniform(0.3, 1.2)
        out[(s, t)] = (jz, jxy)
        out[tuple(sorted((img(s), img(t))))] = (jz, jxy)
    codes, sites, coeffs = [], [], []
    for (s, t), (jz, jxy) in sorted(out.items()):
        for code, v in ((22, jz), (20, jxy), (21, jxy)):
            codes.append(code)
            sites.append([s, t])
            coeffs.append(v)
    return np.asarray(codes, np.int32), np.asarray(sites, np.int32), np.asarray(coeffs, np.complex128)


def general_inversion_perm(ns, rng):
    """An involution of the sites that is not the full reversal: sites
    numbered in a shuffled order along a ring of ns positions, inverted
    through the ring's centre (position p -> ns-1-p)."""
    order = rng.permutation(ns)              # site s sits at position order[s]
    at = np.argsort(order)                   # the site at each position
    perm = at[ns - 1 - order]
    assert not np.array_equal(perm, np.arange(ns)[::-1])
    return order, perm


@pytest.mark.parametrize("flip", [False, True])
def test_general_inversion_sectors_split_the_u1_sector(flip):
    """An inversion whose permutation is not the full site reversal (applied
    to bits as the reference does): the I = +1 and I = -1 blocks (with the
    spin flip: all four (I, P) blocks) of a Hamiltonian that commutes with it
    are Hermitian, their dimensions add up to C(10, 5), and their spectra to
    the U(1) sector's spectrum of the full-space matrix."""
    ns, n_up = 10, 5
    rng = np.random.default_rng(13 + flip)
    _, perm = general_inversion_perm(ns, rng)
    c, s, co = perm_symmetric_bonds(ns, perm, rng)
    Hf = O.full_dense_hamiltonian(ns, c, s, co, False)
    mask = np.array([bin(i).count("1") == n_up for i in range(1 << ns)])
    e_full = np.linalg.eigvalsh(Hf[np.ix_(mask, mask)])
    lat = O.OLattice("chain", ns)
    es = []
    for iv in (1, -1):
        for pv in ((1, -1) if flip else (None,)):
            gens = [O.Generator("I", iv, perm=perm)] + ([O.Generator("Px", pv)] if pv is not None else [])
            b = O.scan_basis(O.Group(lat, gens, n_up=n_up))
            if len(b.reps) == 0:
                continue
            Hr = O.reduced_dense(b, ns, c, s, co, False)
            assert np.allclose(Hr, Hr.conj().T, atol=1e-12)
            es.append(np.linalg.eigvalsh(Hr))
    e = np.sort(np.concatenate(es))
    assert len(e) == len(e_full)
    np.testing.assert_allclose(e, e_full, atol=1e-12)


# ------------------------------------------------- XXZ term list (xxz.py)
def _pauli_string(ns, ops):
    """Sparse 2^ns matrix of a product of Pauli matrices {site: 'x'|'y'|'z'}
    with site i on bit ns-1-i (leftmost Kronecker factor = site 0)."""
    import scipy.sparse as sp
    P = {"x": np.array([[0, 1], [1, 0]], dtype=np.complex128),
         "y": np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
         "z": np.array([[1, 0], [0, -1]], dtype=np.complex128)}
    M = sp.identity(1, dtype=np.complex128, format="csr")
    for site in range(ns):
        f = sp.csr_matrix(P[ops[site]]) if site in ops else sp.identity(2, dtype=np.complex128, format="csr")
        M = sp.kron(M, f, format="csr")
    return M


def test_xxz_two_site_closed_form_fixes_signs():
    """Chain L=2 (PBC: the bond is listed twice, 2 H_bond).  With
    H_bond = -jxy (SxSx + SySy) - jz SzSz (S = sigma/2) the spectrum is fixed
    in closed form: triplet m = +-1 at -jz/4, the symmetric m = 0 triplet at
    -jxy/2 + jz/4 and the singlet at jxy/2 + jz/4 -- so the sign of jxy
    decides which of the two m = 0 states is lower."""
    rng = np.random.default_rng(7)
    for _ in range(4):
        jxy, jz = rng.uniform(-2, 2, 2)
        c, s, co = O.terms_xxz(O.OLattice("chain", 2), jxy=jxy, jz=jz).arrays()
        H = O.full_dense_hamiltonian(2, c, s, co, True)
        up, dn = 3, 0                       # bit 1 = spin up (+1/2)
        trip0 = np.zeros(4, complex)
        trip0[[1, 2]] = 1 / np.sqrt(2)
        sing = np.zeros(4, complex)
        sing[1], sing[2] = 1 / np.sqrt(2), -1 / np.sqrt(2)
        for vec, e in ((np.eye(4)[up], -jz / 4), (np.eye(4)[dn], -jz / 4),
                       (trip0, -jxy / 2 + jz / 4), (sing, jxy / 2 + jz / 4)):
            np.testing.assert_allclose(H @ vec, 2 * e * vec, atol=1e-14)


def test_xxz_fields_fix_signs():
    """-hz Sz_i and -hx Sx_i per site: the saturated state (all bits 1, Sz =
    +1/2 each) has energy -jz n_bonds/4 - sum_i hz_i/2, and the product of
    +x eigenstates has <H> = -sum hx/2 at jxy = jz = 0; site i is bit ns-1-i,
    checked with a random field on one saturated-but-one state."""
    L = 6
    lat = O.OLattice("chain", L)
    rng = np.random.default_rng(3)
    hz = rng.uniform(-5, 5, L)
    c, s, co = O.terms_xxz(lat, jxy=0.6, jz=0.9, hz=hz).arrays()
    H = O.full_dense_hamiltonian(L, c, s, co, True)
    full = (1 << L) - 1
    assert abs(H[full, full] - (-0.9 * L / 4 - hz.sum() / 2)) < 1e-13
    for site in range(L):        # spin of `site` down: its field term flips sign
        st = full ^ (1 << (L - 1 - site))
        ez = -0.9 * (L - 2) / 4 + 0.9 * 2 / 4     # two of the L bonds now antiparallel
        assert abs(H[st, st].real - (ez - hz.sum() / 2 + hz[site])) < 1e-13
    hx = rng.uniform(-2, 2, L)
    c, s, co = O.terms_xxz(lat, jxy=0.0, jz=0.0, hx=hx).arrays()
    H = O.full_dense_hamiltonian(L, c, s, co, True)
    plus = np.ones(1 << L) / np.sqrt(1 << L)
    assert abs(plus @ H @ plus - (-hx.sum() / 2)) < 1e-13


def test_xxz_at_isotropy_equals_heisenberg_and_known_e0():
    """XXZ(jxy = jz = J) is -J sum S.S, the same operator as J1J2Model(J1=J,
    J2=0) (random bond-resolved J, L=10); XXZ(jxy=-1, jz=-1) on the L=16
    ring at Sz=0, k=0 has the antiferromagnet's E0 = -7.14229636061679.  The
    sign of jxy alone is fixed by the L=14 ring, where the AF ground state
    sits at k = pi: per-momentum spectra of XXZ(-1, -1) and of the
    Heisenberg form must agree sector by sector (a flipped jxy moves the
    ground state to k = 0 when the number of down spins is odd)."""
    lat = O.OLattice("chain", 10)
    J = np.random.default_rng(11).uniform(-1, 1, 10)
    c1, s1, co1 = O.terms_xxz(lat, jxy=J, jz=J).arrays()
    c2, s2, co2 = O.terms_j1j2(lat, J1=J, J2=0.0).arrays()
    H1 = O.full_dense_hamiltonian(10, c1, s1, co1, True)
    H2 = O.full_dense_hamiltonian(10, c2, s2, co2, False)
    np.testing.assert_allclose(H1, H2, atol=1e-14)
    lat16 = O.OLattice("chain", 16)
    b = O.scan_basis(O.Group(lat16, [O.Generator("tx", 0)], n_up=8))
    c, s, co = O.terms_xxz(lat16, jxy=-1.0, jz=-1.0).arrays()
    e0 = np.linalg.eigvalsh(O.reduced_dense(b, 16, c, s, co, True))[0]
    assert abs(e0 - (-7.14229636061679)) < 1e-11
    lat14 = O.OLattice("chain", 14)
    cx, sx, cox = O.terms_xxz(lat14, jxy=-1.0, jz=-1.0).arrays()
    cf, sf, cof = O.terms_xxz(lat14, jxy=1.0, jz=-1.0).arrays()
    e_af, e_flip = [], []
    for k in range(14):
        bk = O.scan_basis(O.Group(lat14, [O.Generator("tx", k)], n_up=7))
        e_af.append(np.linalg.eigvalsh(O.reduced_dense(bk, 14, cx, sx, cox, True))[0])
        e_flip.append(np.linalg.eigvalsh(O.reduced_dense(bk, 14, cf, sf, cof, True))[0])
    assert int(np.argmin(e_af)) == 7                   # AF ring, L/2 odd: k = pi
    np.testing.assert_allclose(e_flip, np.roll(e_af, 7), atol=1e-11)   # jxy -> -jxy: k -> k + pi
    assert int(np.argmin(e_flip)) == 0


# --------------------------- Kitaev-Heisenberg bond labels (heisenberg_kitaev.py)
def _hk_bond_labels(lat):
    """Bond -> Kitaev label read back from the term list at J = h = 0 with
    distinct K per direction."""
    c, s, co = O.terms_heisenberg_kitaev(lat, K=(1.0, 2.0, 3.0), J=0.0).arrays()
    lab = {O.SIG_X_C: "x", O.SIG_Y_C: "y", O.SIG_Z_C: "z"}
    out = {}
    for code, (i, j), v in zip(c, s, co):
        out.setdefault((int(i), int(j)), []).append((lab[int(code)], float(np.real(v))))
    return out


@pytest.mark.parametrize("lx,ly", [(2, 3), (3, 2), (3, 4)])
def test_hk_each_site_has_one_bond_of_each_label(lx, ly):
    lat = O.OLattice("honeycomb", lx, ly)
    labels = _hk_bond_labels(lat)
    per_site = {i: [] for i in range(lat.ns)}
    kval = {"x": -4.0, "y": -8.0, "z": -12.0}    # -4 K_alpha (CORR_TERM_MULT = 4, K subtracted)
    for (i, j), terms in labels.items():
        assert len(terms) == 1
        a, v = terms[0]
        assert abs(v - kval[a]) < 1e-14
        per_site[i].append(a)
        per_site[j].append(a)
    for i, labs in per_site.items():
        assert sorted(labs) == ["x", "y", "z"], (i, labs)


def _six_cycles(ns, bonds):
    adj = {i: set() for i in range(ns)}
    for i, j in bonds:
        adj[i].add(j)
        adj[j].add(i)
    cycles = set()
    for s0 in range(ns):
        stack = [(s0, [s0])]
        while stack:
            v, path = stack.pop()
            if len(path) == 6:
                if s0 in adj[v]:
                    cycles.add(frozenset(path))
                continue
            for w in adj[v]:
                if w not in path and w > s0:
                    stack.append((w, path + [w]))
    return [sorted(c) for c in cycles]


@pytest.mark.parametrize("lx,ly", [(2, 3), (3, 2)])
def test_hk_plaquette_operators_commute_with_kitaev_hamiltonian(lx, ly):
    """At J = h = 0 every closed loop operator W = prod_{j in loop}
    sigma_j^{alpha_j} (alpha_j: label of the bond at j that leaves the loop)
    commutes with H (Kitaev's conserved fluxes).  A mislabelled bond breaks
    it on the loops through that bond."""
    lat = O.OLattice("honeycomb", lx, ly)
    labels = _hk_bond_labels(lat)
    label_of = {}
    for (i, j), terms in labels.items():
        label_of[(i, j)] = label_of[(j, i)] = terms[0][0]
    rng = np.random.default_rng(5)
    K = tuple(rng.uniform(-1.5, 1.5, 3))
    c, s, co = O.terms_heisenberg_kitaev(lat, K=K, J=0.0).arrays()
    H = O.full_sparse_hamiltonian(lat.ns, c, s, co, False)
    loops = _six_cycles(lat.ns, list(labels))
    assert len(loops) >= lx * ly
    nb = {i: [j for (a, j) in label_of if a == i] for i in range(lat.ns)}
    checked = 0
    for loop in loops:
        ops = {}
        for v in loop:
            out = [w for w in nb[v] if w not in loop]
            if len(out) != 1:
                break
            ops[v] = label_of[(v, out[0])]
        else:
            W = _pauli_string(lat.ns, ops)
            comm = (W @ H - H @ W).tocsr()
            assert (np.abs(comm.data).max() if comm.nnz else 0.0) < 1e-12, loop
            assert len(ops) == 6 and W.diagonal().sum() == 0   # a non-trivial Pauli string
            checked += 1
    assert checked >= lx * ly


def test_hk_at_zero_kitaev_is_scaled_heisenberg():
    """K = 0: HeisenbergKitaev(J, h) is 4J sum_<ij> S.S - 2 h.S per site
    (CORR_TERM_MULT = 4, SINGLE_TERM_MULT = 2, heisenberg_kitaev.py:72-76),
    i.e. J1J2Model(J1=-4J) with fields 2h on the same bonds."""
    lat = O.OLattice("honeycomb", 2, 3)
    J, hx, hy, hz = 0.37, 0.11, -0.07, 0.23
    c1, s1, co1 = O.terms_heisenberg_kitaev(lat, K=0.0, J=J, hx=hx, hy=hy, hz=hz).arrays()
    c2, s2, co2 = O.terms_j1j2(lat, J1=-4 * J, J2=None, hx=2 * hx, hy=2 * hy, hz=2 * hz).arrays()
    H1 = O.full_sparse_hamiltonian(lat.ns, c1, s1, co1, True)
    H2 = O.full_sparse_hamiltonian(lat.ns, c2, s2, co2, True)
    d = (H1 - H2).tocsr()
    assert H1.nnz > 0 and (np.abs(d.data).max() if d.nnz else 0.0) < 1e-13


# -------------------- twisted translations and intrinsic phases in elements
def _sector_checks(lat, gens, n_up, terms, bound, check_literal):
    """For every momentum sector: the oracle's reduced matrix with element
    phases equals U^dagger H U of the reference's basis vectors (rotation
    matrix with phases, expand_to_full_state_jit) and is Hermitian; the
    sectors' spectra together are the full spectrum of the filling."""
    c, s, co = terms.arrays()
    Hf = O.full_dense_hamiltonian(lat.ns, c, s, co, True)
    fill = np.array([bin(i).count("1") == n_up for i in range(1 << lat.ns)])
    E_full = np.linalg.eigvalsh(Hf[np.ix_(fill, fill)])
    E = []
    literal_nonherm = 0.0
    for gl in gens:
        grp = O.Group(lat, [O.Generator(k, v) for k, v in gl], n_up=n_up, bound_phase=bound)
        b = O.scan_basis(grp)
        if len(b.reps) == 0:
            continue
        U = O.rotation_matrix(grp, b)
        np.testing.assert_allclose(U.conj().T @ U, np.eye(len(b.reps)), atol=1e-12)
        Hr = O.reduced_dense(b, lat.ns, c, s, co, True, group=grp)
        np.testing.assert_allclose(Hr, U.conj().T @ Hf @ U, atol=1e-12)
        np.testing.assert_allclose(Hr, Hr.conj().T, atol=1e-12)
        E += list(np.linalg.eigvalsh(Hr))
        if check_literal:
            Hl = O.reduced_dense(b, lat.ns, c, s, co, True)
            literal_nonherm = max(literal_nonherm, np.abs(Hl - Hl.conj().T).max())
    np.testing.assert_allclose(np.sort(E), E_full, atol=1e-11)
    return literal_nonherm


@pytest.mark.parametrize("kind,lx,ly,n_up,flux", [("chain", 8, 1, 4, (2 * np.pi / 4, 0.0)),
                                                  ("chain", 9, 1, 3, (4 * np.pi / 3, 0.0)),
                                                  ("square", 3, 2, 3, (2 * np.pi / 3, 0.0)),
                                                  ("square", 2, 3, 3, (0.0, 2 * np.pi / 3))])
def test_twisted_translations_decompose_the_twisted_model(kind, lx, ly, n_up, flux):
    """Boundary phases in the group action (symmetry_container_jit.py:91-98:
    exp(i phi) per up spin crossing, per translation step) make the twisted
    translations commute with the twisted XXZ model; its momentum sectors
    (phi N_up a multiple of 2 pi) are Hermitian blocks whose spectra add up
    to the filling's.  Without the element phase (the reference's kernel
    takes only the character) the blocks are not Hermitian."""
    lat = O.OLattice(kind, lx, ly)
    bound = O.boundary_phase_table(lat.ns, flux)
    terms = O.terms_xxz_twisted(lat, 1.0, 0.6, flux)
    if kind == "chain":
        gens = [[("tx", k)] for k in range(lx)]
    elif flux[1] == 0.0:
        gens = [[("tx", k)] for k in range(lx)]
    else:
        gens = [[("ty", k)] for k in range(ly)]
    literal = _sector_checks(lat, gens, n_up, terms, bound, check_literal=True)
    assert literal > 1e-3
