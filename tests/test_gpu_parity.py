"""GPU parity of the HIP path (through the C ABI) against the CPU oracle.

Bars: representatives / index maps bit-exact; matvec <= 1e-12 relative L2;
E0 <= 1e-10 relative.  Sizes the oracle finishes in seconds; larger sizes
through size-independent properties (closed-form dimensions and energies,
hermiticity, determinism, sampled rows)."""
import numpy as np
import pytest

from oracle import qes_oracle as O

pytestmark = pytest.mark.gpu

MATVEC_TOL = 1e-12
E0_TOL = 1e-10


def _lat(kind, lx, ly):
    from pyqusolver_amd import ChainLattice, HoneycombLattice, SquareLattice
    return {"chain": lambda: ChainLattice(lx), "square": lambda: SquareLattice(lx, ly),
            "honeycomb": lambda: HoneycombLattice(lx, ly)}[kind]()


def _sym(gens):
    sym = {}
    tr = {}
    for k, s in gens:
        if k == "tx":
            tr["kx"] = s
        elif k == "ty":
            tr["ky"] = s
        elif k == "R":
            sym["reflection"] = s
        elif k == "I":
            sym["inversion"] = s
        elif k in ("Px", "Py", "Pz"):
            sym.setdefault("parity", {})[k[1]] = s
    if tr:
        sym = {"translation": tr, **sym}
    return sym


def build_pair(kind, lx, ly, gens, n_up, model, z2=None):
    from pyqusolver_amd import get_u1_sym, get_z2_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import HeisenbergKitaev, J1J2Model, XXZ
    lat = _lat(kind, lx, ly)
    olat = O.OLattice(kind, lx, ly)
    gl = []
    if n_up is not None:
        gl.append(get_u1_sym(lat, n_up))
    if z2 is not None:
        gl.append(get_z2_sym(lat, z2))
    hs = HilbertSpace(lattice=lat, sym_gen=_sym(gens), global_syms=gl)
    grp = O.Group(olat, [O.Generator(k, s) for k, s in gens], n_up=n_up, z2=z2)
    b = O.scan_basis(grp)
    if model == "xxz":
        H = XXZ(lat, hilbert_space=hs, jxy=1.0, jz=0.7, hz=0.1)
        t = O.terms_xxz(olat, 1.0, 0.7, hz=0.1)
    elif model == "j1j2":
        H = J1J2Model(lat, J1=-1.0, J2=-0.5, hilbert_space=hs)
        t = O.terms_j1j2(olat, -1.0, -0.5)
    elif model == "heis":
        H = J1J2Model(lat, J1=-1.0, J2=0.0, hilbert_space=hs)
        t = O.terms_j1j2(olat, -1.0, 0.0)
    elif model == "j1j2_field":
        H = J1J2Model(lat, J1=-1.0, J2=-0.3, hilbert_space=hs, hx=0.2, hz=0.1)
        t = O.terms_j1j2(olat, -1.0, -0.3, hx=0.2, hz=0.1)
    else:
        h = 0.05 / np.sqrt(3)
        H = HeisenbergKitaev(lat, K=-1.0, J=0.2, hilbert_space=hs, hx=h, hy=h, hz=h)
        t = O.terms_heisenberg_kitaev(olat, K=-1.0, J=0.2, hx=h, hy=h, hz=h)
    return hs, H, grp, b, t


CASES = [
    ("chain", 16, 1, [("tx", 0)], 8, "j1j2"),
    ("chain", 14, 1, [("tx", 3)], 7, "xxz"),
    ("chain", 14, 1, [("tx", 7), ("R", -1)], 7, "xxz"),
    ("chain", 14, 1, [("tx", 0), ("R", 1), ("Px", -1)], 7, "j1j2"),
    ("chain", 14, 1, [("tx", 7), ("R", -1), ("Px", 1)], 7, "j1j2"),
    ("chain", 13, 1, [("tx", 5)], 6, "xxz"),
    ("chain", 12, 1, [("tx", 0), ("Py", 1)], 6, "j1j2"),
    ("chain", 12, 1, [("tx", 2), ("Py", -1)], 6, "j1j2"),
    ("chain", 12, 1, [("tx", 0), ("R", 1)], None, "j1j2_field"),
    ("chain", 12, 1, [("tx", 4), ("Px", 1)], None, "j1j2_field"),
    ("square", 4, 4, [("tx", 0), ("ty", 0), ("Px", 1)], 8, "j1j2"),
    ("square", 4, 3, [("tx", 1), ("ty", 2)], 6, "j1j2"),
    ("square", 4, 3, [("tx", 2), ("ty", 0), ("R", -1)], 6, "j1j2"),
    ("square", 3, 4, [("ty", 1)], 6, "j1j2"),
    ("honeycomb", 3, 2, [("tx", 0), ("ty", 0)], None, "hk"),
    ("honeycomb", 2, 3, [("tx", 1), ("ty", 2)], None, "hk"),
    ("chain", 14, 1, [("tx", 7), ("I", -1), ("Px", 1)], 7, "j1j2"),
    ("square", 4, 4, [("tx", 0), ("ty", 2), ("I", -1)], 8, "j1j2"),
    ("square", 4, 3, [("tx", 2), ("ty", 0), ("I", 1)], None, "j1j2_field"),
    ("honeycomb", 3, 2, [("tx", 0), ("ty", 0), ("I", 1)], None, "hk"),
    ("honeycomb", 2, 3, [("tx", 1), ("ty", 0), ("I", -1)], 6, "hk"),
]


def _ids(c):
    return f"{c[0]}{c[1]}x{c[2]}-" + "-".join(f"{k}{s}" for k, s in c[3]) + f"-n{c[4]}-{c[5]}"


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_basis_bit_exact_and_matvec(case):
    kind, lx, ly, gens, n_up, model = case
    hs, H, grp, b, t = build_pair(kind, lx, ly, gens, n_up, model)
    assert hs.nh == len(b.reps)
    if not any(k == "Py" for k, _ in gens):
        assert hs.nh == O.sector_dimension(grp)
    np.testing.assert_array_equal(hs.representative_list, b.reps)
    np.testing.assert_allclose(hs.normalization, b.norms, rtol=1e-14, atol=0)
    idx, pidx = hs.map_states(b.states)
    np.testing.assert_array_equal(idx, b.state_rep)
    np.testing.assert_array_equal(pidx, b.state_pidx)
    np.testing.assert_allclose(hs.group.phase_table, b.phase_table, rtol=0, atol=0)
    if any(k == "Py" for k, _ in gens):
        # the reference block is not Hermitian here: basis kept, matvec refused
        from pyqusolver_amd._lib import QesError
        with pytest.raises(QesError):
            H.matvec_fun(np.ones(hs.nh, dtype=np.complex128))
        return
    c, s, co = t.arrays()
    rng = np.random.default_rng(7)
    n = hs.nh
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = H.matvec_fun(x)
    yo = O.push_matvec(b, lx * ly * (2 if kind == "honeycomb" else 1), c, s, co, t.is_complex, x)
    assert np.linalg.norm(y - yo) <= MATVEC_TOL * np.linalg.norm(yo)
    if not H.needs_complex:
        xr = rng.standard_normal(n)
        yr = H.matvec_fun(xr)
        assert yr.dtype == np.float64
        yor = O.push_matvec(b, lx * ly * (2 if kind == "honeycomb" else 1), c, s, co, t.is_complex, xr)
        assert np.linalg.norm(yr - yor) <= MATVEC_TOL * np.linalg.norm(yor)


@pytest.mark.parametrize("case", [("chain", 12, 1, [("tx", 0), ("R", 1)], 1, "xxz"),
                                  ("chain", 12, 1, [("tx", 3)], -1, "xxz"),
                                  ("chain", 12, 1, [("tx", 0), ("Px", 1)], 1, "j1j2"),
                                  ("square", 4, 3, [("tx", 1), ("ty", 0)], -1, "xxz")],
                         ids=lambda c: f"{c[0]}{c[1]}x{c[2]}-" + "-".join(f"{k}{s}" for k, s in c[3]) + f"-z2{c[4]}")
def test_z2_parity_sectors_without_u1(case):
    """Z2 popcount-parity global symmetry without U(1) (globals.py:336):
    representatives, norms and index maps bit-exact against the oracle,
    sector dimension from Burnside's count, matvec and E0 against the oracle's
    dense block."""
    kind, lx, ly, gens, z2, model = case
    hs, H, grp, b, t = build_pair(kind, lx, ly, gens, None, model, z2=z2)
    ns = t.ns
    assert hs.nh == len(b.reps) == O.sector_dimension(grp)
    np.testing.assert_array_equal(hs.representative_list, b.reps)
    np.testing.assert_allclose(hs.normalization, b.norms, rtol=1e-14, atol=0)
    par = np.array([bin(int(r)).count("1") % 2 for r in b.reps])
    assert np.all(np.where(par == 0, 1, -1) == z2)
    idx, pidx = hs.map_states(b.states)
    np.testing.assert_array_equal(idx, b.state_rep)
    c, s, co = t.arrays()
    rng = np.random.default_rng(11)
    x = rng.standard_normal(hs.nh) + 1j * rng.standard_normal(hs.nh)
    yo = O.push_matvec(b, ns, c, s, co, t.is_complex, x)
    assert np.linalg.norm(H.matvec_fun(x) - yo) <= MATVEC_TOL * np.linalg.norm(yo)
    ev = np.linalg.eigvalsh(O.reduced_dense(b, ns, c, s, co, t.is_complex))
    e0 = H.diagonalize(method="lanczos", k=1, tol=1e-13, maxiter=min(300, hs.nh), seed=0)[0]
    assert abs(e0 - ev[0]) <= 1e-10 * abs(ev[0])


@pytest.mark.parametrize("case", [c for c in CASES if c[4] is not None and c[1] * c[2] * (2 if c[0] == "honeycomb" else 1) <= 32
                                  and all(k != "Py" for k, _ in c[3])], ids=_ids)
def test_rank_map_lookup_matches_sorted_keys(case):
    """The rank-map lookup (bitmap over the colex ranks of the U(1) filling)
    finds the same representatives as the sorted keys: index maps and E_off
    equal, products bitwise equal with the same visiting order (the default
    direct pass of the rank map sums in another order: equal to rounding);
    out-of-sector states miss."""
    import torch
    kind, lx, ly, gens, n_up, model = case
    hs, H, grp, b, t = build_pair(kind, lx, ly, gens, n_up, model)
    dt = torch.complex128 if H.needs_complex else torch.float64
    x = torch.randn(hs.nh, dtype=dt, device="cuda", generator=torch.Generator(device="cuda").manual_seed(5))
    y0 = H.matvec_device(x)
    idx0, g0 = hs.map_states(b.states)
    e0 = H.count_offdiag()
    hs.set_lookup("rank")
    assert torch.equal(H.matvec_device(x), y0)
    idx1, g1 = hs.map_states(b.states)
    np.testing.assert_array_equal(idx1, idx0)
    np.testing.assert_array_equal(g1, g0)
    assert H.count_offdiag() == e0
    ns = t.ns
    other = np.array([s for s in range(min(1 << ns, 1 << 16)) if bin(s).count("1") != n_up][:64], dtype=np.uint64)
    if len(other):
        assert np.all(hs.map_states(other)[0] < 0)
    hs.set_lookup("sorted")   # no direct pass there: another summation order
    ys = H.matvec_device(x)
    assert (torch.linalg.vector_norm(ys - y0) / torch.linalg.vector_norm(y0)).item() <= 1e-14
    H.set_option("direct_lookup", 0)   # the same order as the sorted keys
    hs.set_lookup("rank")

    def same(a, b):   # complex multiply-adds may contract differently in the batched-lookup path
        if a.is_complex():
            return (torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)).item() <= 1e-15
        return torch.equal(a, b)
    assert same(H.matvec_device(x), ys)
    H.set_option("direct_lookup", -1)
    assert same(H.matvec_device(x), y0)


@pytest.mark.parametrize("case", [("chain", 36, 1, [("tx", 0), ("R", 1)], 3, "xxz"),
                                  ("chain", 34, 1, [("tx", 5)], 4, "xxz"),
                                  ("chain", 40, 1, [("tx", 0)], 3, "j1j2"),
                                  ("square", 6, 6, [("tx", 1), ("ty", 0)], 3, "j1j2")], ids=_ids)
def test_rank_map_beyond_32_sites(case):
    """33..48 sites: the rank map's three-chunk colex rank (top 16 bits, then
    a row per popcount of the low 32 bits, then the low 16 bits) finds the
    oracle's representatives: products against the oracle and equal to the
    sorted 64-bit keys to rounding (the rank map's direct pass sums in
    another order), index maps equal."""
    import torch
    kind, lx, ly, gens, n_up, model = case
    hs, H, grp, b, t = build_pair(kind, lx, ly, gens, n_up, model)
    np.testing.assert_array_equal(hs.representative_list, b.reps)
    c, s, co = t.arrays()
    rng = np.random.default_rng(3)
    x = rng.standard_normal(hs.nh) + 1j * rng.standard_normal(hs.nh)
    yo = O.push_matvec(b, t.ns, c, s, co, t.is_complex, x)
    y = H.matvec_fun(x)
    assert np.linalg.norm(y - yo) <= MATVEC_TOL * np.linalg.norm(yo)
    idx0, g0 = hs.map_states(b.states)
    np.testing.assert_array_equal(idx0, b.state_rep)
    hs.set_lookup("sorted")
    ys = H.matvec_fun(x)
    assert np.linalg.norm(ys - y) <= 1e-14 * np.linalg.norm(y)
    np.testing.assert_array_equal(hs.map_states(b.states)[0], idx0)
    hs.set_lookup("rank")
    np.testing.assert_array_equal(H.matvec_fun(x), y)


def test_rank_map_refused_without_u1():
    from pyqusolver_amd._lib import QesError
    hs, H, grp, b, t = build_pair("chain", 12, 1, [("tx", 0), ("R", 1)], None, "j1j2_field")
    with pytest.raises(QesError, match="rank map"):
        hs.set_lookup("rank")


@pytest.mark.parametrize("nb", [2, 3, 5])
def test_multi_rhs_matches_single(nb):
    hs, H, grp, b, t = build_pair("chain", 14, 1, [("tx", 3)], 7, "xxz")
    rng = np.random.default_rng(nb)
    X = rng.standard_normal((hs.nh, nb)) + 1j * rng.standard_normal((hs.nh, nb))
    Y = H.matvec_fun(X)
    for j in range(nb):
        yj = H.matvec_fun(np.ascontiguousarray(X[:, j]))
        # one column gathers N_b x_b from a pre-scaled copy; several columns
        # scale each element by N_b: equal to rounding
        assert np.linalg.norm(Y[:, j] - yj) <= 1e-14 * np.linalg.norm(yj)


def test_matvec_is_deterministic_and_linear():
    import torch
    hs, H, grp, b, t = build_pair("square", 4, 4, [("tx", 0), ("ty", 0), ("Px", 1)], 8, "j1j2")
    x = torch.randn(hs.nh, dtype=torch.float64, device="cuda")
    z = torch.randn(hs.nh, dtype=torch.float64, device="cuda")
    y1 = H.matvec_device(x)
    y2 = H.matvec_device(x)
    assert torch.equal(y1, y2)
    lin = H.matvec_device(2.0 * x + z) - (2.0 * y1 + H.matvec_device(z))
    assert lin.norm().item() <= 1e-13 * y1.norm().item()


def test_accumulate_flag():
    import torch
    hs, H, grp, b, t = build_pair("chain", 14, 1, [("tx", 0), ("R", 1), ("Px", -1)], 7, "j1j2")
    x = torch.randn(hs.nh, dtype=torch.float64, device="cuda")
    y0 = torch.randn(hs.nh, dtype=torch.float64, device="cuda")
    y = y0.clone()
    H.matvec_device(x, out=y, accumulate=True)
    assert torch.allclose(y, y0 + H.matvec_device(x), rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("spatial", ["R", "I"])
def test_spatial_symmetry_at_generic_momentum_keeps_basis_refuses_matvec(spatial):
    """Reflection / inversion with a translation at k not in {0, pi}: the
    reference accepts the sector (translation commutes with both by its
    compatibility table, translation.py:137-148) but exp(ikn) p is not a
    representation there (P T P^-1 = T^-1), so its block is not Hermitian:
    the basis is built bit-exactly, the matvec refused."""
    from pyqusolver_amd._lib import QesError
    hs, H, grp, b, t = build_pair("chain", 12, 1, [("tx", 2), (spatial, 1)], 6, "xxz")
    np.testing.assert_array_equal(hs.representative_list, b.reps)
    np.testing.assert_allclose(hs.normalization, b.norms, rtol=1e-14, atol=0)
    with pytest.raises(QesError):
        H.matvec_fun(np.ones(hs.nh, dtype=np.complex128))


def test_real_vector_with_complex_sector_raises():
    from pyqusolver_amd._lib import QesError
    hs, H, grp, b, t = build_pair("chain", 14, 1, [("tx", 3)], 7, "xxz")
    with pytest.raises(QesError):
        H.matvec_fun(np.ones(hs.nh))


def test_linear_operator_drop_in():
    import scipy.sparse.linalg as spla
    hs, H, grp, b, t = build_pair("chain", 16, 1, [("tx", 0)], 8, "j1j2")
    A = H.as_linear_operator()
    assert A.shape == (810, 810)
    w = spla.eigsh(A, k=1, which="SA", tol=1e-12)[0][0]
    c, s, co = t.arrays()
    e_ref = np.linalg.eigvalsh(O.reduced_dense(b, 16, c, s, co, t.is_complex))[0]
    assert abs(w - e_ref) <= E0_TOL * abs(e_ref)


def test_chain16_ground_state_energy():
    hs, H, grp, b, t = build_pair("chain", 16, 1, [("tx", 0)], 8, "j1j2")
    H2 = __import__("pyqusolver_amd.models", fromlist=["J1J2Model"]).J1J2Model(hs.lattice, J1=-1.0, J2=0.0,
                                                                                hilbert_space=hs)
    e0 = H2.diagonalize(method="lanczos", k=1, tol=1e-13, maxiter=200, seed=0)[0]
    c, s, co = O.terms_j1j2(O.OLattice("chain", 16), -1.0, 0.0).arrays()
    e_ref = np.linalg.eigvalsh(O.reduced_dense(b, 16, c, s, co, False))[0]
    assert abs(e0 - e_ref) <= E0_TOL * abs(e_ref)
    assert abs(e0 - (-7.14229636061679)) <= E0_TOL * 7.15


@pytest.mark.parametrize("case", [c for c in CASES if c[1] * c[2] <= 16 and all(k != "Py" for k, _ in c[3])],
                         ids=_ids)
def test_lanczos_e0_against_dense(case):
    kind, lx, ly, gens, n_up, model = case
    hs, H, grp, b, t = build_pair(kind, lx, ly, gens, n_up, model)
    if hs.nh < 2:
        pytest.skip("trivial sector")
    c, s, co = t.arrays()
    ns = lx * ly * (2 if kind == "honeycomb" else 1)
    e_ref = np.linalg.eigvalsh(O.reduced_dense(b, ns, c, s, co, t.is_complex))[0]
    e0 = H.diagonalize(method="lanczos", k=1, tol=1e-13, maxiter=min(300, hs.nh), compute_eigvec=True)[0]
    assert abs(e0 - e_ref) <= E0_TOL * max(1.0, abs(e_ref))
    psi = H.eig_vec
    r = H.matvec_device(psi) - e0 * psi
    assert r.norm().item() <= 1e-6


@pytest.mark.parametrize("reorth", [True, "twice"])
def test_full_reorth_keeps_spectrum_ghost_free(reorth):
    """Many more Lanczos steps than needed for E0 (N = 805 states, 400
    steps): with full reorthogonalisation (DGKS second pass, or always
    twice) the lowest Ritz values are the lowest eigenvalues of the dense
    reduced matrix, with no spurious copies."""
    from pyqusolver_amd.lanczos import lanczos_ground_state
    hs, H, grp, b, t = build_pair("chain", 16, 1, [("tx", 2)], 8, "xxz")
    c, s, co = t.arrays()
    ev = np.linalg.eigvalsh(O.reduced_dense(b, 16, c, s, co, True))
    res = lanczos_ground_state(H, k=4, tol=0.0, maxiter=400, seed=3, reorth=reorth, check_every=400)
    assert res.steps == 400
    w = np.linalg.eigvalsh(np.diag(res.alpha) + np.diag(res.beta[:-1], 1) + np.diag(res.beta[:-1], -1))
    np.testing.assert_allclose(w[:4], ev[:4], rtol=0, atol=1e-9 * abs(ev[0]))
    if reorth == "twice":
        assert res.second_passes == 0     # counted only by the DGKS variant
    else:
        assert 0 <= res.second_passes <= 400


def test_majumdar_ghosh_chain24_exact_energy():
    """J1-J2 chain at J2 = J1/2: E0 = -3L/8 exactly (Majumdar-Ghosh)."""
    from pyqusolver_amd import ChainLattice, get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import J1J2Model
    lat = ChainLattice(24)
    hs = HilbertSpace(lattice=lat, sym_gen={"translation": 0, "parity": {"x": 1}}, global_syms=[get_u1_sym(lat, 12)])
    H = J1J2Model(lat, J1=-1.0, J2=-0.5, hilbert_space=hs)
    e0 = H.diagonalize(method="lanczos", k=1, tol=1e-13, maxiter=300)[0]
    assert abs(e0 - (-9.0)) <= E0_TOL * 9.0


def test_chain32_scale_properties():
    """L=32 J1-J2, Sz=0, k=0, flip+ (9,393,119 states): closed-form dimension,
    sorted reps, sampled rows against the oracle, hermiticity, determinism."""
    import torch
    from pyqusolver_amd import ChainLattice, get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import J1J2Model
    lat = ChainLattice(32)
    hs = HilbertSpace(lattice=lat, sym_gen={"translation": 0, "parity": {"x": 1}}, global_syms=[get_u1_sym(lat, 16)])
    assert hs.nh == 9_393_119
    reps = hs.representative_list
    assert np.all(np.diff(reps) > 0)
    H = J1J2Model(lat, J1=-1.0, J2=-0.5, hilbert_space=hs)
    n = hs.nh
    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    z = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    y = H.matvec_device(x)
    assert torch.equal(y, H.matvec_device(x))
    lhs = torch.dot(z, y).item()
    rhs = torch.dot(H.matvec_device(z), x).item()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # sampled rows against the oracle (group action by permutation tables)
    grp = O.Group(O.OLattice("chain", 32), [O.Generator("tx", 0), O.Generator("Px", 1)], n_up=16)
    t = O.terms_j1j2(O.OLattice("chain", 32), -1.0, -0.5)
    c, s, co = t.arrays()
    rng = np.random.default_rng(1)
    sample = np.sort(rng.choice(n, 300, replace=False))
    xs = x.cpu().numpy()
    ys = O.sampled_rows(grp, reps, hs.normalization, sample, 32, c, s, co, False, xs)
    yg = y.cpu().numpy()[sample]
    assert np.linalg.norm(yg - ys.real) <= MATVEC_TOL * np.linalg.norm(ys)
    # norms: N_r^2 = |Stab(r)| on the sample
    _, _, cnt = O.canonicalize(grp, reps[sample])
    np.testing.assert_allclose(hs.normalization[sample] ** 2, cnt, rtol=1e-14)


def test_square6x6_scale_properties():
    """6x6 J1-J2, Sz=0, k=(0,0), flip+ (126,056,625 states, |G| = 72, 64-bit
    states): closed-form dimension, sorted reps, rows sampled across the
    basis against the oracle's row formula, hermiticity, and the assembled
    table's product equal to the matrix-free one to rounding."""
    import torch
    from pyqusolver_amd import SquareLattice, get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import J1J2Model
    lat = SquareLattice(6, 6)
    hs = HilbertSpace(lattice=lat, sym_gen={"translation": {"kx": 0, "ky": 0}, "parity": {"x": 1}},
                      global_syms=[get_u1_sym(lat, 18)])
    n = hs.nh
    assert n == 126_056_625
    reps = hs.representative_list
    assert np.all(reps[1:] > reps[:-1])
    H = J1J2Model(lat, J1=-1.0, J2=-0.5, hilbert_space=hs)
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    y = H.matvec_device(x)
    z = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    lhs = torch.dot(z, y).item()
    rhs = torch.dot(H.matvec_device(z), x).item()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    del z
    olat = O.OLattice("square", 6, 6)
    grp = O.Group(olat, [O.Generator("tx", 0), O.Generator("ty", 0), O.Generator("Px", 1)], n_up=18)
    c, s, co = O.terms_j1j2(olat, -1.0, -0.5).arrays()
    sample = np.sort(np.random.default_rng(2).choice(n, 200, replace=False))
    xs = x.cpu().numpy()
    ys = O.sampled_rows(grp, reps, hs.normalization, sample, 36, c, s, co, False, xs)
    yg = y.cpu().numpy()[sample]
    assert np.linalg.norm(yg - ys.real) <= MATVEC_TOL * np.linalg.norm(ys)
    _, _, cnt = O.canonicalize(grp, reps[sample])
    np.testing.assert_allclose(hs.normalization[sample] ** 2, cnt, rtol=1e-14)
    H.assemble(True, max_gb=64)
    ya = H.matvec_device(x)
    assert (torch.linalg.vector_norm(ya - y) / torch.linalg.vector_norm(y)).item() <= 1e-14
    del ya
    # Ritz vector without a stored basis at full size (the accumulation must
    # wait for the zero fill of its output, queued on another stream): equal
    # to the stored-basis Ritz vector, residual equal to |beta_m s_m|
    from pyqusolver_amd.lanczos import lanczos_ground_state
    del x, y
    r2 = lanczos_ground_state(H, k=1, tol=1e-10, maxiter=20, seed=1, reorth=False, store_basis=False,
                              compute_eigvec=True)
    r1 = lanczos_ground_state(H, k=1, tol=1e-10, maxiter=20, seed=1, reorth=False, store_basis=True,
                              compute_eigvec=True)
    v1, v2 = r1.eigenvectors, r2.eigenvectors
    assert (torch.linalg.vector_norm(v1 - v2) / torch.linalg.vector_norm(v1)).item() <= 1e-12
    res = torch.linalg.vector_norm(H.matvec_device(v2) - float(r2.eigenvalues[0]) * v2) / torch.linalg.vector_norm(v2)
    assert abs(res.item() - r2.residual_bound) <= 1e-6 * r2.residual_bound + 1e-12
    del v1, v2, r1, r2
    x = torch.randn(n, dtype=torch.float64, device="cuda", generator=torch.Generator(device="cuda").manual_seed(3))
    y = H.matvec_device(x)
    H.assemble(False)
    # a stored basis of 1024 vectors (1 TB) cannot be allocated: the create
    # fails with a message naming it, frees what it had taken, and leaves no
    # error behind for the next call
    from pyqusolver_amd import _lib
    from pyqusolver_amd.lanczos import DeviceLanczos
    free0 = torch.cuda.mem_get_info()[0]
    with pytest.raises(_lib.QesError, match="stored Krylov basis"):
        DeviceLanczos(H, 1023, True, False)
    assert torch.cuda.mem_get_info()[0] >= free0 - (64 << 20)
    # y came from the table; the matrix-free product equals it to rounding
    assert (torch.linalg.vector_norm(H.matvec_device(x) - y) / torch.linalg.vector_norm(y)).item() <= 1e-14


@pytest.mark.parametrize("which", ["honeycomb3x4_hk", "chain24_xxz_k5"])
def test_workload_scale_sampled_rows(which):
    """Full-size sectors of the Kitaev-Heisenberg workload (honeycomb 3x4
    cells, k=0, no U(1), fields along [111]: N = 1,398,500, complex) and of
    the batched sweep (XXZ L=24, Sz=0, k=5: complex characters): closed-form
    dimension, 300 sampled rows against the oracle's row formula, and
    hermiticity <z, Hx> = <Hz, x>."""
    import torch
    from pyqusolver_amd import ChainLattice, HoneycombLattice, get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import HeisenbergKitaev, XXZ
    if which == "honeycomb3x4_hk":
        lat, olat = HoneycombLattice(3, 4), O.OLattice("honeycomb", 3, 4)
        hs = HilbertSpace(lattice=lat, sym_gen={"translation": {"kx": 0, "ky": 0}})
        h = 0.05 / np.sqrt(3.0)
        H = HeisenbergKitaev(lat, K=-1.0, J=0.2, hilbert_space=hs, hx=h, hy=h, hz=h)
        t = O.terms_heisenberg_kitaev(olat, K=-1.0, J=0.2, hx=h, hy=h, hz=h)
        grp = O.Group(olat, [O.Generator("tx", 0), O.Generator("ty", 0)])
        assert hs.nh == 1_398_500
    else:
        lat, olat = ChainLattice(24), O.OLattice("chain", 24)
        hs = HilbertSpace(lattice=lat, sym_gen={"translation": 5}, global_syms=[get_u1_sym(lat, 12)])
        H = XXZ(lat, hilbert_space=hs, jxy=1.0, jz=0.7)
        t = O.terms_xxz(olat, 1.0, 0.7)
        grp = O.Group(olat, [O.Generator("tx", 5)], n_up=12)
    n = hs.nh
    assert n == O.sector_dimension(grp)
    c, s, co = t.arrays()
    g = torch.Generator(device="cuda").manual_seed(4)
    x = torch.complex(torch.randn(n, dtype=torch.float64, device="cuda", generator=g),
                      torch.randn(n, dtype=torch.float64, device="cuda", generator=g))
    z = torch.complex(torch.randn(n, dtype=torch.float64, device="cuda", generator=g),
                      torch.randn(n, dtype=torch.float64, device="cuda", generator=g))
    y = H.matvec_device(x)
    lhs = torch.vdot(z, y).item()
    rhs = torch.vdot(H.matvec_device(z), x).item()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    reps = hs.representative_list
    sample = np.sort(np.random.default_rng(6).choice(n, 300, replace=False))
    ys = O.sampled_rows(grp, reps, hs.normalization, sample, lat.ns, c, s, co, True, x.cpu().numpy())
    yg = y.cpu().numpy()[sample]
    assert np.linalg.norm(yg - ys) <= MATVEC_TOL * np.linalg.norm(ys)


def test_index_map_on_random_states_chain32():
    from pyqusolver_amd import ChainLattice, get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    lat = ChainLattice(32)
    hs = HilbertSpace(lattice=lat, sym_gen={"translation": 0, "parity": {"x": 1}}, global_syms=[get_u1_sym(lat, 16)])
    reps = hs.representative_list
    grp = O.Group(O.OLattice("chain", 32), [O.Generator("tx", 0), O.Generator("Px", 1)], n_up=16)
    rng = np.random.default_rng(5)
    states = []
    while len(states) < 400:
        bits = rng.choice(32, 16, replace=False)
        states.append(int(np.sum(1 << bits.astype(np.int64))))
    states = np.array(states, dtype=np.int64)
    idx, pidx = hs.map_states(states)
    best, bg, _ = O.canonicalize(grp, states)
    pos = np.searchsorted(reps, best)
    ok = (pos < len(reps)) & (reps[np.minimum(pos, len(reps) - 1)] == best)
    np.testing.assert_array_equal(idx, np.where(ok, pos, -1))


def test_golden_fixtures():
    """The HIP path against the committed fixtures: representatives bit-exact,
    normalisations to 1e-14, H x to 1e-12 relative L2, E0 (device Lanczos with
    full reorthogonalisation) to 1e-10 relative."""
    from test_golden import CASES as GCASES, DATA
    for case in GCASES:
        name = case["name"]
        hs, H, grp, b, t = build_pair(case["kind"], case["lx"], case["ly"], [tuple(g) for g in case["gens"]],
                                      case["n_up"], {"heis": "heis"}.get(case["model"], case["model"]))
        np.testing.assert_array_equal(hs.representative_list.astype(np.uint64), DATA[f"{name}/reps"], err_msg=name)
        np.testing.assert_allclose(hs.normalization, DATA[f"{name}/norms"], rtol=1e-14, atol=0, err_msg=name)
        yg = DATA[f"{name}/y"]
        y = H.matvec_fun(DATA[f"{name}/x"])
        assert np.linalg.norm(y - yg) <= MATVEC_TOL * np.linalg.norm(yg), name
        e_ref = DATA[f"{name}/e0"][0]
        e0 = H.diagonalize(method="lanczos", k=1, tol=1e-14, maxiter=min(300, hs.nh), seed=0, reorth=True)[0]
        assert abs(e0 - e_ref) <= E0_TOL * abs(e_ref), (name, e0, e_ref)


FULL_CASES = [
    ("chain", 12, 1, [("tx", 3)], 6, "xxz"),
    ("chain", 12, 1, [("tx", 0), ("R", 1), ("Px", 1)], 6, "j1j2"),
    ("chain", 10, 1, [("tx", 5), ("Px", -1)], None, "j1j2_field"),
    ("square", 4, 3, [("tx", 1), ("ty", 2)], 6, "j1j2"),
    ("honeycomb", 3, 2, [("tx", 0), ("ty", 0)], None, "hk"),
]


@pytest.mark.parametrize("case", FULL_CASES, ids=_ids)
def test_expand_and_reduce_against_oracle(case):
    """expand_state / transform_to_reduced_space on the device against the
    oracle restatements; pins independent of them: reduce(expand(x)) =
    sqrt|G| / N_k^2 x_k (each orbit state carries chi(g*) N_k / sqrt|G|), and
    the expanded ground state is an eigenvector of the full-space H."""
    kind, lx, ly, gens, n_up, model = case
    hs, H, grp, b, t = build_pair(kind, lx, ly, gens, n_up, model)
    ns = lx * ly * (2 if kind == "honeycomb" else 1)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(hs.nh) + 1j * rng.standard_normal(hs.nh)
    full = hs.expand_state(x)
    ref = O.expand_state(grp, b, x)
    assert np.linalg.norm(full - ref) <= 1e-13 * np.linalg.norm(ref)
    psi = rng.standard_normal(1 << ns) + 1j * rng.standard_normal(1 << ns)
    red = hs.transform_to_reduced_space(psi)
    red_ref = O.transform_to_reduced(b, ns, psi)
    assert np.linalg.norm(red - red_ref) <= 1e-13 * np.linalg.norm(red_ref)
    back = hs.transform_to_reduced_space(full)
    np.testing.assert_allclose(back, x * np.sqrt(grp.n_group) / b.norms ** 2, rtol=1e-12, atol=0)
    # eigenvector of the full-space Hamiltonian
    c, s, co = t.arrays()
    e0 = H.diagonalize(method="lanczos", k=1, tol=1e-13, maxiter=min(300, hs.nh), compute_eigvec=True)[0]
    v = H.eig_vec.cpu().numpy()
    vf = hs.expand_state(v)
    Hf = O.full_sparse_hamiltonian(ns, c, s, co, t.is_complex)
    r = Hf @ vf - e0 * vf
    assert np.linalg.norm(r) <= 1e-6 * np.linalg.norm(vf)


@pytest.mark.parametrize("case", [c for c in CASES if all(k != "Py" for k, _ in c[3]) and c[1] * c[2] <= 14],
                         ids=_ids)
def test_explicit_sparse_build_matches_oracle_matrix(case):
    """Hamiltonian.build(): the device table as a scipy CSR matrix equals the
    oracle's reduced matrix (push-form columns) and reproduces the matvec."""
    kind, lx, ly, gens, n_up, model = case
    hs, H, grp, b, t = build_pair(kind, lx, ly, gens, n_up, model)
    c, s, co = t.arrays()
    ns = lx * ly * (2 if kind == "honeycomb" else 1)
    M = O.reduced_dense(b, ns, c, s, co, t.is_complex)
    A = H.build()
    assert A.shape == (hs.nh, hs.nh)
    np.testing.assert_allclose(A.toarray(), M, rtol=0, atol=1e-13 * np.abs(M).max())
    x = np.random.default_rng(2).standard_normal(hs.nh) + 0j
    np.testing.assert_allclose(A @ x, H.matvec_fun(x), rtol=0, atol=1e-12 * np.abs(A @ x).max())
    assert H.assembly_bytes == 0     # the table built for the export is freed


def _sz_q(hs, L, q, code=13):
    from pyqusolver_amd.operator import Operator, SpinOperator
    op = Operator(hs)
    for j in range(L):
        op.add(SpinOperator("Sq", code, "local", code != 13), np.exp(1j * q * j), sites=[j])
    codes = [code] * L
    sites = [[j, -1] for j in range(L)]
    coeffs = [np.exp(1j * q * j) for j in range(L)]
    return op, codes, sites, coeffs


@pytest.mark.parametrize("kin,kout,nin,nout,code", [(0, 3, 6, 6, 13), (2, 5, 6, 6, 13), (1, 4, 6, 7, 14),
                                                     (4, 1, 6, 5, 15)])
def test_projected_matvec_between_momentum_sectors(kin, kout, nin, nout, code):
    """Operator.matvec(x, hilbert_out=...) for S^z_q, S^+_q, S^-_q between
    momentum / filling sectors: against the oracle restatement of the
    reference's projected kernel and against U_out^dagger O U_in."""
    from pyqusolver_amd import ChainLattice, get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    L = 12
    lat = ChainLattice(L)
    hs_in = HilbertSpace(lattice=lat, sym_gen={"translation": kin}, global_syms=[get_u1_sym(lat, nin)])
    hs_out = HilbertSpace(lattice=lat, sym_gen={"translation": kout}, global_syms=[get_u1_sym(lat, nout)])
    q = 2 * np.pi * (kout - kin) / L
    op, codes, sites, coeffs = _sz_q(hs_in, L, q, code)
    x = np.random.default_rng(kin).standard_normal(hs_in.nh) + 1j * np.random.default_rng(9).standard_normal(hs_in.nh)
    y = op.matvec(x, hilbert_out=hs_out)
    olat = O.OLattice("chain", L)
    gi = O.Group(olat, [O.Generator("tx", kin)], n_up=nin)
    go = O.Group(olat, [O.Generator("tx", kout)], n_up=nout)
    bi, bo = O.scan_basis(gi), O.scan_basis(go)
    yo = O.projected_matvec(gi, bi, go, bo, L, codes, sites, coeffs, True, x)
    assert np.linalg.norm(y - yo) <= MATVEC_TOL * max(np.linalg.norm(yo), 1e-300)
    Of = O.full_sparse_hamiltonian(L, codes, sites, coeffs, True)
    yb = O.rotation_matrix(go, bo).conj().T @ (Of @ (O.rotation_matrix(gi, bi) @ x))
    assert np.linalg.norm(y - yb) <= 1e-11 * max(np.linalg.norm(yb), 1e-300)


def test_projected_matvec_same_sector_and_2d():
    """Into its own sector the projected product is the ordinary matvec; a 2D
    case (4x3 square, kx 1 -> 2) against the oracle."""
    from pyqusolver_amd import SquareLattice, get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    hs, H, grp, b, t = build_pair("chain", 12, 1, [("tx", 2)], 6, "xxz")
    x = np.random.default_rng(4).standard_normal(hs.nh) + 0j
    hs2 = hs.__class__(lattice=hs.lattice, sym_gen={"translation": 2}, global_syms=hs.global_syms)
    y_proj = H.projected_matvec(x, hs2)
    y = H.matvec_fun(x)
    assert np.linalg.norm(y_proj - y) <= MATVEC_TOL * np.linalg.norm(y)
    lat = SquareLattice(4, 3)
    hs_in = HilbertSpace(lattice=lat, sym_gen={"translation": {"kx": 1, "ky": 0}}, global_syms=[get_u1_sym(lat, 6)])
    hs_out = HilbertSpace(lattice=lat, sym_gen={"translation": {"kx": 2, "ky": 0}}, global_syms=[get_u1_sym(lat, 6)])
    q = 2 * np.pi / 4
    from pyqusolver_amd.operator import Operator, SpinOperator
    op = Operator(hs_in)
    codes, sites, coeffs = [], [], []
    for j in range(12):
        c = np.exp(1j * q * (j % 4))
        op.add(SpinOperator("Sz", 13, "local", False), c, sites=[j])
        codes.append(13)
        sites.append([j, -1])
        coeffs.append(c)
    xin = np.random.default_rng(5).standard_normal(hs_in.nh) + 0j
    y = op.matvec(xin, hilbert_out=hs_out)
    olat = O.OLattice("square", 4, 3)
    gi = O.Group(olat, [O.Generator("tx", 1), O.Generator("ty", 0)], n_up=6)
    go = O.Group(olat, [O.Generator("tx", 2), O.Generator("ty", 0)], n_up=6)
    bi, bo = O.scan_basis(gi), O.scan_basis(go)
    yo = O.projected_matvec(gi, bi, go, bo, 12, codes, sites, coeffs, True, xin)
    assert np.linalg.norm(y - yo) <= MATVEC_TOL * np.linalg.norm(yo)


@pytest.mark.parametrize("gens", [[], [("tx", 1)], [("tx", 0), ("I", 1)]], ids=["none", "tx1", "tx0-I+"])
def test_boundary_flux_kitaev_sector_parity(gens):
    """Kitaev-Heisenberg honeycomb 3x2 with a flux 2 pi/3 along y (bonds
    carry the Hermitian part of the boundary phase, heisenberg_kitaev.py:
    487-571): reduced matvec and E0 against the oracle in sectors that keep
    the couplings invariant (translation along x, inversion), the explicit
    build Hermitian, and translation along y refused."""
    from pyqusolver_amd import HoneycombLattice
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import HeisenbergKitaev
    phi = 2 * np.pi / 3
    lat = HoneycombLattice(3, 2, flux={"y": phi})
    olat = O.OLattice("honeycomb", 3, 2)
    h = (0.03, 0.02, 0.05)
    hs = HilbertSpace(lattice=lat, sym_gen=_sym(gens))
    H = HeisenbergKitaev(lat, K=-1.0, J=0.3, hilbert_space=hs, hx=h[0], hy=h[1], hz=h[2])
    grp = O.Group(olat, [O.Generator(k, s) for k, s in gens])
    b = O.scan_basis(grp)
    np.testing.assert_array_equal(hs.representative_list, b.reps)
    c, s, co = O.terms_heisenberg_kitaev(olat, K=-1.0, J=0.3, hx=h[0], hy=h[1], hz=h[2], flux=(0.0, phi)).arrays()
    rng = np.random.default_rng(5)
    x = rng.standard_normal(hs.nh) + 1j * rng.standard_normal(hs.nh)
    y = H.matvec_fun(x)
    yo = O.push_matvec(b, lat.ns, c, s, co, True, x)
    assert np.linalg.norm(y - yo) <= MATVEC_TOL * np.linalg.norm(yo)
    A = H.build(use_numpy=True)
    assert np.abs(A - A.conj().T).max() <= 1e-12
    Hd = O.reduced_dense(b, lat.ns, c, s, co, True)
    e_ref = np.linalg.eigvalsh(Hd)[0]
    e0 = H.diagonalize(method="lanczos", k=1, tol=1e-12, maxiter=400)[0]
    assert abs(e0 - e_ref) <= E0_TOL * abs(e_ref)
    with pytest.raises(NotImplementedError):
        HilbertSpace(lattice=lat, sym_gen={"translation": {"ky": 0}})


def test_reference_workflow_checks_tfim_build_and_reduction():
    """The reference's own workflow checks (tests/algebra/
    test_custom_hamiltonian_workflows.py): a translation sector is smaller
    than the full space (L=6, :37-50) and the sparse build equals the dense
    one for a transverse-field Ising chain (L=4, :53-60; here XXZ with
    Jxy=0, Jz=1, hx=0.5), also equal to the oracle's full-space matrix."""
    from pyqusolver_amd import ChainLattice
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import XXZ
    lat6 = ChainLattice(6)
    hs6 = HilbertSpace(lattice=lat6, sym_gen={"translation": 0})
    assert hs6.nh < hs6.nhfull == 64
    lat = ChainLattice(4)
    hs = HilbertSpace(lattice=lat)
    H = XXZ(lat, hilbert_space=hs, jxy=0.0, jz=1.0, hx=0.5)
    A = H.build()
    D = H.build(use_numpy=True)
    np.testing.assert_allclose(A.toarray(), D, atol=1e-10)
    c, s, co = O.terms_xxz(O.OLattice("chain", 4), 0.0, 1.0, hx=0.5).arrays()
    ref = O.full_dense_hamiltonian(4, c, s, co, True)
    np.testing.assert_allclose(D, ref, atol=1e-14)


def test_expectation_value_and_overlap():
    """Operator.expectation_value / overlap (operator.py:926-1075): <x|H|y>
    against the oracle's reduced matrix (numpy and device tensors), pairwise
    and full column blocks, E0 as the expectation in the Ritz vector, and a
    sector-changing matrix element <out|S^z_q|in> against the oracle's
    projected product."""
    import torch
    hs, H, grp, b, t = build_pair("chain", 14, 1, [("tx", 3)], 7, "xxz")
    c, s, co = t.arrays()
    Hd = O.reduced_dense(b, 14, c, s, co, True)
    rng = np.random.default_rng(12)
    n = hs.nh
    X = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    Y = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    ev = H.expectation_value(X[:, 0], Y[:, 0])
    ref = np.vdot(X[:, 0], Hd @ Y[:, 0])
    assert abs(ev - ref) <= 1e-12 * abs(ref)
    np.testing.assert_allclose(H.expectation_value(X, Y), np.einsum("ij,ij->j", X.conj(), Hd @ Y), rtol=1e-12)
    np.testing.assert_allclose(H.expectation_value(X, Y, pairwise=False), X.conj().T @ Hd @ Y, rtol=1e-12)
    Xt, Yt = torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda()
    evt = H.expectation_value(Xt[:, 0], Yt[:, 0])
    assert isinstance(evt, torch.Tensor) and abs(evt.item() - ref) <= 1e-12 * abs(ref)
    np.testing.assert_allclose(H.overlap(X, Y), np.einsum("ij,ij->j", X.conj(), Y), rtol=1e-13)
    H.diagonalize(method="lanczos", k=1, tol=1e-12, maxiter=300, compute_eigvec=True)
    psi = H.eig_vec
    e = (H.expectation_value(psi) / H.overlap(psi)).real.item()
    e_ref = np.linalg.eigvalsh(Hd)[0]
    assert abs(e - e_ref) <= 1e-10 * abs(e_ref)
    from pyqusolver_amd import ChainLattice, get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    L = 12
    lat = ChainLattice(L)
    hs_in = HilbertSpace(lattice=lat, sym_gen={"translation": 1}, global_syms=[get_u1_sym(lat, 6)])
    hs_out = HilbertSpace(lattice=lat, sym_gen={"translation": 4}, global_syms=[get_u1_sym(lat, 6)])
    op, codes, sites, coeffs = _sz_q(hs_in, L, 2 * np.pi * 3 / L, 13)
    xin = rng.standard_normal(hs_in.nh) + 1j * rng.standard_normal(hs_in.nh)
    xout = rng.standard_normal(hs_out.nh) + 1j * rng.standard_normal(hs_out.nh)
    me = op.expectation_value(xout, xin, hilbert_out=hs_out)
    olat = O.OLattice("chain", L)
    gi, go = O.Group(olat, [O.Generator("tx", 1)], n_up=6), O.Group(olat, [O.Generator("tx", 4)], n_up=6)
    yo = O.projected_matvec(gi, O.scan_basis(gi), go, O.scan_basis(go), L, codes, sites, coeffs, True, xin)
    assert abs(me - np.vdot(xout, yo)) <= 1e-12 * abs(np.vdot(xout, yo))


def test_lowest_k_eigenpairs_with_reorthogonalisation():
    """diagonalize(k=4) with the basis kept: the four lowest Ritz values of a
    sector whose lowest levels are non-degenerate match the dense reduced
    matrix to 1e-10, each has a residual below the tolerance, and the (n, 4)
    Ritz vectors satisfy H v = E v."""
    import torch
    hs, H, grp, b, t = build_pair("chain", 14, 1, [("tx", 0), ("R", 1), ("Px", -1)], 7, "j1j2")
    c, s, co = t.arrays()
    ev = np.linalg.eigvalsh(O.reduced_dense(b, 14, c, s, co, True))
    assert np.min(np.diff(ev[:5])) > 1e-3          # lowest levels non-degenerate here
    w = H.diagonalize(method="lanczos", k=4, tol=1e-11, maxiter=min(300, hs.nh), reorth=True, compute_eigvec=True)
    np.testing.assert_allclose(w, ev[:4], rtol=0, atol=1e-10 * abs(ev[0]))
    assert H.last_lanczos.converged and H.last_lanczos.residual_bound <= 1e-11 * max(1.0, abs(ev[3]))
    V = H.eig_vec
    assert tuple(V.shape) == (hs.nh, 4)
    for i in range(4):
        v = V[:, i].contiguous()
        r = H.matvec_device(v) - float(w[i]) * v
        assert torch.linalg.vector_norm(r).item() <= 1e-8 * torch.linalg.vector_norm(v).item()


@pytest.mark.parametrize("case", [("chain", 14, 1, [("tx", 0), ("R", 1)], 7, "j1j2", False),
                                  ("chain", 12, 1, [("tx", 5)], 6, "xxz", True),
                                  ("honeycomb", 2, 2, [("tx", 0), ("ty", 0)], None, "hk", True)])
def test_exact_diagonalisation_on_device(case):
    """method='exact' (the reference's default): the whole spectrum of the
    dense reduced matrix built on the device, against the oracle's dense
    matrix on the CPU to 1e-12 relative; eigenvectors (n, n) with
    H v = E v; 'auto' takes the same path at this size."""
    import torch
    kind, lx, ly, gens, n_up, model, cplx = case
    hs, H, grp, b, t = build_pair(kind, lx, ly, gens, n_up, model)
    c, s, co = t.arrays()
    ev = np.linalg.eigvalsh(O.reduced_dense(b, t.ns, c, s, co, cplx))
    assert bool(H.needs_complex) == cplx
    w = H.diagonalize()
    assert len(w) == hs.nh and H.eig_vec is None
    scale = np.max(np.abs(ev))
    np.testing.assert_allclose(w, ev, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(H.diagonalize(method="auto"), ev, rtol=0, atol=1e-12 * scale)
    w = H.diagonalize(method="exact", compute_eigvec=True)
    V = H.eig_vec
    assert tuple(V.shape) == (hs.nh, hs.nh)
    for i in (0, hs.nh // 2, hs.nh - 1):
        v = V[:, i].contiguous()
        r = H.matvec_device(v) - float(w[i]) * v
        assert torch.linalg.vector_norm(r).item() <= 1e-10 * scale


@pytest.mark.parametrize("which", ["largest", "both"])
def test_largest_and_both_ends(which):
    """which='largest' / 'both' (the reference's diagonalize options,
    hamil_diag_engine.py:183): the k highest, or k//2 lowest and the rest
    highest, eigenvalues of the dense reduced matrix to 1e-10, ascending,
    with Ritz vectors that satisfy H v = E v; the default k is 6."""
    import torch
    hs, H, grp, b, t = build_pair("chain", 14, 1, [("tx", 0), ("R", 1), ("Px", -1)], 7, "j1j2")
    c, s, co = t.arrays()
    ev = np.linalg.eigvalsh(O.reduced_dense(b, 14, c, s, co, True))
    want = ev[-6:] if which == "largest" else np.concatenate([ev[:3], ev[-3:]])
    assert np.min(np.diff(want)) > 1e-3             # separated levels at both ends here
    w = H.diagonalize(method="lanczos", which=which, tol=1e-11, maxiter=min(300, hs.nh), compute_eigvec=True)
    assert len(w) == 6
    np.testing.assert_allclose(w, want, rtol=0, atol=1e-10 * np.max(np.abs(ev)))
    assert H.last_lanczos.converged
    V = H.eig_vec
    for i in range(6):
        v = V[:, i].contiguous()
        r = H.matvec_device(v) - float(w[i]) * v
        assert torch.linalg.vector_norm(r).item() <= 1e-8 * torch.linalg.vector_norm(v).item()
    with pytest.raises(ValueError):
        H.diagonalize(method="lanczos", which="LM")
    with pytest.raises(NotImplementedError):
        H.diagonalize(method="lanczos", sigma=0.5)


def test_two_pass_ritz_vector_without_stored_basis():
    """Ritz vectors without keeping the Krylov basis: the recurrence is run
    a second time from the same start vector (bitwise reproducible steps)
    and sum_j s_j v_j accumulated on the device.  Equal to the stored-basis
    Ritz vector of the same recurrence, an eigenvector of H to the residual,
    for k = 1 and k = 2."""
    import torch
    from pyqusolver_amd.lanczos import lanczos_ground_state
    hs, H, grp, b, t = build_pair("chain", 16, 1, [("tx", 0)], 8, "j1j2")
    for k in (1, 2):
        r2 = lanczos_ground_state(H, k=k, tol=1e-12, maxiter=200, seed=4, reorth=False, store_basis=False,
                                  compute_eigvec=True)
        r1 = lanczos_ground_state(H, k=k, tol=1e-12, maxiter=200, seed=4, reorth=False, store_basis=True,
                                  compute_eigvec=True)
        np.testing.assert_array_equal(r1.alpha, r2.alpha)
        v1, v2 = r1.eigenvectors, r2.eigenvectors
        assert tuple(v2.shape) == tuple(v1.shape)
        assert torch.allclose(v2, v1, rtol=0, atol=1e-12)
        psi = v2 if k == 1 else v2[:, 0].contiguous()
        res = H.matvec_device(psi) - float(r2.eigenvalues[0]) * psi
        assert torch.linalg.vector_norm(res).item() <= 1e-9 * torch.linalg.vector_norm(psi).item()


def test_in_place_and_overlapping_products_are_rejected():
    import torch
    from pyqusolver_amd import ChainLattice, get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import J1J2Model
    lat = ChainLattice(12)
    hs = HilbertSpace(lattice=lat, sym_gen={"translation": 0}, global_syms=[get_u1_sym(lat, 6)])
    H = J1J2Model(lat, J1=-1.0, J2=-0.3, hilbert_space=hs)
    x = torch.randn(2 * hs.nh, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        H.matvec_device(x[:hs.nh], out=x[:hs.nh])
    with pytest.raises(ValueError):
        H.matvec_device(x[:hs.nh], out=x[1:hs.nh + 1])
    y = H.matvec_device(x[:hs.nh], out=x[hs.nh:])    # adjacent, not overlapping
    torch.testing.assert_close(y, H.matvec_device(x[:hs.nh].clone()), rtol=0, atol=0)


def test_operator_in_use_refuses_table_and_plan_changes():
    """A live Lanczos handle or batch holds the operator's plan, assembled
    table and lookup tables (its captured graphs keep the pointers): changing
    or freeing them must fail until the handle is destroyed."""
    from pyqusolver_amd import ChainLattice, get_u1_sym
    from pyqusolver_amd._lib import QesError
    from pyqusolver_amd.batched import BatchedLanczos
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.lanczos import DeviceLanczos
    from pyqusolver_amd.models import J1J2Model
    lat = ChainLattice(12)
    hs = HilbertSpace(lattice=lat, sym_gen={"translation": 0}, global_syms=[get_u1_sym(lat, 6)])
    H = J1J2Model(lat, J1=-1.0, J2=-0.3, hilbert_space=hs)
    B = BatchedLanczos([H], max_steps=20)            # assembles H's table for the batch
    assert H.assembly_bytes > 0
    with pytest.raises(QesError, match="in use"):
        H.assemble(False)
    with pytest.raises(QesError, match="in use"):
        H.set_option("direct_lookup", 0)
    with pytest.raises(QesError, match="in use"):
        hs.set_lookup("sorted")
    B.close()
    assert H.assembly_bytes == 0
    L = DeviceLanczos(H, 10, False, False)
    with pytest.raises(QesError, match="in use"):
        H.assemble(True)
    L.close()
    H.assemble(True)
    H.assemble(False)
    hs.set_lookup("rank")


def _colex_rank(state: int) -> int:
    from math import comb
    r, i, s = 0, 0, int(state)
    while s:
        p = (s & -s).bit_length() - 1
        i += 1
        r += comb(p, i)
        s &= s - 1
    return r


def _colex_unrank(ns: int, k: int, r: int) -> int:
    from math import comb
    s, c = 0, ns - 1
    for i in range(k, 0, -1):
        while c >= i - 1 and comb(c, i) > r:
            c -= 1
        s |= 1 << c
        r -= comb(c, i)
        c -= 1
    return s


@pytest.mark.parametrize("kind", ["chain32", "square6x6"])
def test_basis_windows_bit_exact_at_full_size(kind):
    """At full size (L=32: 9,393,119 representatives; 6x6: 126,056,625) the
    representative list equals the C oracle's orbit scan bit for bit on
    three windows of 2^22 consecutive colex ranks of the filling: at the
    first representative, around the median one and ending at the last one
    (representatives carry their longest zero run on top, so they sit at
    low ranks); norms to 1e-14."""
    from math import comb
    from oracle import qes_oracle_c as OC
    from pyqusolver_amd import ChainLattice, SquareLattice, get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    if kind == "chain32":
        lat, ns, n_up = ChainLattice(32), 32, 16
        sym, gens = {"translation": 0, "parity": {"x": 1}}, [("tx", 0), ("Px", 1)]
        olat = O.OLattice("chain", 32)
    else:
        lat, ns, n_up = SquareLattice(6, 6), 36, 18
        sym, gens = {"translation": {"kx": 0, "ky": 0}, "parity": {"x": 1}}, [("tx", 0), ("ty", 0), ("Px", 1)]
        olat = O.OLattice("square", 6, 6)
    hs = HilbertSpace(lattice=lat, sym_gen=sym, global_syms=[get_u1_sym(lat, n_up)])
    reps = hs.representative_list
    norms = hs.normalization
    grp = O.Group(olat, [O.Generator(g, k) for g, k in gens], n_up=n_up)
    cg = OC.CGroup(grp)
    w = 1 << 22
    total = comb(ns, n_up)
    n = len(reps)
    starts = [_colex_rank(reps[0]), max(0, _colex_rank(reps[n // 2]) - w // 2), _colex_rank(reps[-1]) - w + 1]
    checked = 0
    for r0 in starts:
        r0 = min(max(0, r0), total - w)
        ro, no = OC.scan_window(grp, r0, w, cgroup=cg, threads=16)
        lo = np.searchsorted(reps, _colex_unrank(ns, n_up, r0))
        hi = np.searchsorted(reps, _colex_unrank(ns, n_up, r0 + w)) if r0 + w < total else n
        np.testing.assert_array_equal(reps[lo:hi], ro)
        np.testing.assert_allclose(norms[lo:hi], no, rtol=1e-14)
        checked += len(ro)
    assert checked > 100_000


def _all_codes_terms(L, rng, u1):
    """A Hermitian, translation-invariant instruction list that uses every
    code the VM accepts (operators_spin.py:1084-1118): each non-Hermitian
    term comes with its conjugate (S+ / S-, S+S+ / S-S-, S+S- / S-S+, complex
    coefficients conjugated); the mixed Pauli products XY, YX, XZ, ZX, YZ, ZY
    are Hermitian with real coefficients.  u1: only terms that keep the
    popcount (diagonal, XX + YY exchange, S+S- / S-S+)."""
    c = lambda: complex(rng.uniform(-1, 1), rng.uniform(-1, 1))
    r = lambda: float(rng.uniform(-1, 1))
    terms = []   # (code, coefficient, [sites]) as functions of the first site i
    if u1:
        cz, cx, cpm = r(), r(), c()
        terms += [(13, lambda i: r0, lambda i: [i]) for r0 in [r()]]
        terms += [(22, lambda i, v=cz: v, lambda i: [i, (i + 1) % L]),
                  (20, lambda i, v=cx: v, lambda i: [i, (i + 1) % L]),
                  (21, lambda i, v=cx: v, lambda i: [i, (i + 1) % L]),
                  (31, lambda i, v=cpm: v, lambda i: [i, (i + 2) % L]),
                  (32, lambda i, v=cpm: np.conj(v), lambda i: [i, (i + 2) % L])]
        return terms
    cp, cpp, cpm = c(), c(), c()
    terms += [(11, lambda i, v=r(): v, lambda i: [i]), (12, lambda i, v=r(): v, lambda i: [i]),
              (13, lambda i, v=r(): v, lambda i: [i]),
              (14, lambda i, v=cp: v, lambda i: [i]), (15, lambda i, v=cp: np.conj(v), lambda i: [i]),
              (20, lambda i, v=r(): v, lambda i: [i, (i + 1) % L]), (21, lambda i, v=r(): v, lambda i: [i, (i + 1) % L]),
              (22, lambda i, v=r(): v, lambda i: [i, (i + 1) % L]),
              (23, lambda i, v=cpp: v, lambda i: [i, (i + 1) % L]),
              (24, lambda i, v=cpp: np.conj(v), lambda i: [i, (i + 1) % L]),
              (31, lambda i, v=cpm: v, lambda i: [i, (i + 2) % L]),
              (32, lambda i, v=cpm: np.conj(v), lambda i: [i, (i + 2) % L])]
    for code in (25, 26, 27, 28, 29, 30):
        terms.append((code, lambda i, v=r(): v, lambda i: [i, (i + 1) % L]))
    return terms


@pytest.mark.parametrize("L,gens,n_up", [(10, [("tx", 3)], None), (11, [("tx", 0)], None),
                                         (12, [("tx", 5)], 6), (12, [("tx", 6)], 6)])
def test_every_instruction_code_against_the_oracle(L, gens, n_up):
    """Codes 11-15 and 20-32 through the HIP product (qes_op_create ->
    qes_matvec) against the oracle's push-form restatement of the reference
    kernel on the same instruction arrays; complex coefficients."""
    from pyqusolver_amd import ChainLattice, get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.operator import Hamiltonian, SpinOperator
    rng = np.random.default_rng(L + (n_up or 0))
    lat = ChainLattice(L)
    hs = HilbertSpace(lattice=lat, sym_gen=_sym(gens), global_syms=[get_u1_sym(lat, n_up)] if n_up else [])
    H = Hamiltonian(hs, dtype=np.complex128)
    codes, sites, coeffs = [], [], []
    for code, coef, st in _all_codes_terms(L, rng, n_up is not None):
        for i in range(L):
            H.add(SpinOperator(f"c{code}", code, "local" if code < 20 else "correlation", True), coef(i), sites=st(i))
            codes.append(code)
            sites.append(st(i) + [-1] * (2 - len(st(i))))
            coeffs.append(coef(i))
    used = set(codes)
    assert used >= ({11, 12, 13, 14, 15, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32} if n_up is None
                    else {13, 20, 21, 22, 31, 32})
    grp = O.Group(O.OLattice("chain", L), [O.Generator(k, s) for k, s in gens], n_up=n_up)
    b = O.scan_basis(grp)
    np.testing.assert_array_equal(hs.representative_list, b.reps)
    c_ = np.asarray(codes, np.int32)
    s_ = np.asarray(sites, np.int32)
    co_ = np.asarray(coeffs, np.complex128)
    x = rng.standard_normal(hs.nh) + 1j * rng.standard_normal(hs.nh)
    yo = O.push_matvec(b, L, c_, s_, co_, True, x)
    y = H.matvec_fun(x)
    assert np.linalg.norm(y - yo) <= MATVEC_TOL * np.linalg.norm(yo)
    # the reduced operator is Hermitian (the pull form relies on it)
    z = rng.standard_normal(hs.nh) + 1j * rng.standard_normal(hs.nh)
    assert abs(np.vdot(z, y) - np.vdot(H.matvec_fun(z), x)) <= 1e-12 * np.linalg.norm(y) * np.linalg.norm(z)


@pytest.mark.parametrize("kind,lx,ly,n_up,flux,axis", [("chain", 12, 1, 6, np.pi / 3, "x"),
                                                      ("chain", 16, 1, 8, 3 * np.pi / 4, "x"),
                                                      ("square", 4, 3, 6, np.pi / 3, "x"),
                                                      ("square", 3, 4, 6, 2 * np.pi / 3, "y")])
def test_twisted_translation_sectors_against_the_oracle(kind, lx, ly, n_up, flux, axis):
    """Boundary flux phi along a translated direction: the twisted
    translations (exp(i phi) per up spin crossing, per step;
    symmetry_container_jit.py:91-98) give the basis of the oracle's orbit scan
    with the same phases (representatives bit-exact, norms 1e-14); the
    twisted-boundary XXZ model (XXZ(twist=True)) has the oracle's
    instruction list, its products equal the oracle's push form with the
    element phases, every block is Hermitian, and the lowest energies of all
    momentum sectors contain the filling's ground state."""
    from pyqusolver_amd import ChainLattice, SquareLattice, get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import XXZ
    fl = {axis: flux}
    lat = ChainLattice(lx, flux=flux) if kind == "chain" else SquareLattice(lx, ly, flux=fl)
    olat = O.OLattice(kind, lx, ly)
    fxy = (flux, 0.0) if axis == "x" else (0.0, flux)
    bound = O.boundary_phase_table(olat.ns, fxy)
    t = O.terms_xxz_twisted(olat, 1.0, 0.6, fxy)
    c, s, co = t.arrays()
    L = lx if axis == "x" else ly
    gname = "tx" if axis == "x" else "ty"
    e_min = []
    for k in range(L):
        sym = {"translation": {"kx": k} if axis == "x" else {"ky": k}} if kind != "chain" else {"translation": k}
        hs = HilbertSpace(lattice=lat, sym_gen=sym, global_syms=[get_u1_sym(lat, n_up)])
        grp = O.Group(olat, [O.Generator(gname, k)], n_up=n_up, bound_phase=bound)
        b = O.scan_basis(grp)
        np.testing.assert_array_equal(hs.representative_list, b.reps)
        np.testing.assert_allclose(hs.normalization, b.norms, rtol=1e-14)
        if hs.nh == 0:
            continue
        H = XXZ(lat, hilbert_space=hs, jxy=1.0, jz=0.6, twist=True)
        sites_h, co_h, codes_h = H.instruction_arrays()
        np.testing.assert_array_equal(codes_h, c)
        np.testing.assert_array_equal(sites_h, s)
        np.testing.assert_allclose(co_h, co, rtol=0, atol=1e-15)
        rng = np.random.default_rng(k)
        x = rng.standard_normal(hs.nh) + 1j * rng.standard_normal(hs.nh)
        yo = O.push_matvec(b, olat.ns, c, s, co, True, x, group=grp)
        y = H.matvec_fun(x)
        assert np.linalg.norm(y - yo) <= MATVEC_TOL * np.linalg.norm(yo), k
        z = rng.standard_normal(hs.nh) + 1j * rng.standard_normal(hs.nh)
        assert abs(np.vdot(z, y) - np.vdot(H.matvec_fun(z), x)) <= 1e-12 * np.linalg.norm(y) * np.linalg.norm(z)
        Hd = O.reduced_dense(b, olat.ns, c, s, co, True, group=grp)
        e_ref = np.linalg.eigvalsh(Hd)[0]
        e_min.append(e_ref)
        if hs.nh > 20:
            e0 = H.diagonalize(method="lanczos", k=1, tol=1e-12, maxiter=300)[0]
            assert abs(e0 - e_ref) <= 1e-10 * max(1.0, abs(e_ref))
    Hf = O.full_sparse_hamiltonian(olat.ns, c, s, co, True)
    idx = np.array([i for i in range(1 << olat.ns) if bin(i).count("1") == n_up])
    import scipy.sparse.linalg as spla
    e_full = spla.eigsh(Hf[idx][:, idx], k=1, which="SA", tol=1e-13)[0][0]
    assert abs(min(e_min) - e_full) <= 1e-10 * abs(e_full)


@pytest.mark.parametrize("flip", [False, True])
def test_general_inversion_permutation_against_the_oracle(flip):
    """An inversion whose permutation is not the full site reversal (a ring
    whose sites are numbered in a shuffled order) runs as a permutation base
    element on the device: representatives bit-exact and products against
    the oracle's restatement with the same permutation, for a Hamiltonian
    that commutes with it; the (I, P) blocks' spectra add up to the U(1)
    sector's."""
    from pyqusolver_amd import get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.operator import Hamiltonian, SpinOperator
    import test_host_logic as HL
    import test_oracle_pins as OP
    ns, n_up = 10, 5
    rng = np.random.default_rng(13 + flip)
    order, perm = OP.general_inversion_perm(ns, rng)
    c, s, co = OP.perm_symmetric_bonds(ns, perm, rng)
    lat = HL._ShuffledRing(ns, order)
    e_blocks = []
    for iv in (1, -1):
        for pv in ((1, -1) if flip else (None,)):
            sym = {"inversion": iv}
            if pv is not None:
                sym["parity"] = {"x": pv}
            hs = HilbertSpace(lattice=lat, sym_gen=sym, global_syms=[get_u1_sym(lat, n_up)])
            assert hs.group.perm is not None
            gens = [O.Generator("I", iv, perm=perm)] + ([O.Generator("Px", pv)] if pv is not None else [])
            b = O.scan_basis(O.Group(O.OLattice("chain", ns), gens, n_up=n_up))
            np.testing.assert_array_equal(hs.representative_list, b.reps)
            np.testing.assert_allclose(hs.normalization, b.norms, rtol=1e-14)
            if hs.nh == 0:
                continue
            H = Hamiltonian(hs, dtype=np.float64)
            for code, st, v in zip(c, s, co):
                H.add(SpinOperator(f"c{code}", int(code), "correlation", code != 22), float(v.real),
                      sites=[int(st[0]), int(st[1])])
            x = rng.standard_normal(hs.nh)
            yo = O.push_matvec(b, ns, c, s, co, False, x)
            y = H.matvec_fun(x)
            assert np.linalg.norm(y - yo) <= MATVEC_TOL * np.linalg.norm(yo)
            Hr = O.reduced_dense(b, ns, c, s, co, False)
            e_blocks.append(np.linalg.eigvalsh(Hr))
            if hs.nh > 8:
                e0 = H.diagonalize(method="lanczos", k=1, tol=1e-12, maxiter=200)[0]
                assert abs(e0 - e_blocks[-1][0]) <= 1e-10 * abs(e_blocks[-1][0])
    Hf = O.full_dense_hamiltonian(ns, c, s, co, False)
    mask = np.array([bin(i).count("1") == n_up for i in range(1 << ns)])
    np.testing.assert_allclose(np.sort(np.concatenate(e_blocks)), np.linalg.eigvalsh(Hf[np.ix_(mask, mask)]),
                               atol=1e-12)
