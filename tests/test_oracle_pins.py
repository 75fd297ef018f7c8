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

