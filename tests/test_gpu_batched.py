"""Batched multi-sector / multi-realisation Lanczos against per-problem solves
and the oracle (GPU)."""
import numpy as np
import pytest

from oracle import qes_oracle as O

pytestmark = pytest.mark.gpu


def _xxz_sectors(L, jz=0.7):
    from pyqusolver_amd import ChainLattice, get_u1_sym
    from pyqusolver_amd.batched import momentum_sectors
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import XXZ
    lat = ChainLattice(L)
    specs = momentum_sectors(L, reflection=True)
    spaces = [HilbertSpace(lattice=lat, sym_gen=s, global_syms=[get_u1_sym(lat, L // 2)]) for s in specs]
    hams = [XXZ(lat, hilbert_space=hs, jxy=1.0, jz=jz) for hs in spaces]
    return lat, specs, spaces, hams


def test_batched_matvec_equals_individual():
    import torch
    from pyqusolver_amd.batched import BatchedLanczos
    lat, specs, spaces, hams = _xxz_sectors(12)
    B = BatchedLanczos(hams, max_steps=50)
    rng = np.random.default_rng(0)
    xs = [rng.standard_normal(h.nh) + 1j * rng.standard_normal(h.nh) for h in hams]
    x = B.concat(xs)
    y = B.matvec(x).cpu().numpy()
    ys = B.split(torch.from_numpy(y))
    for p, h in enumerate(hams):
        yp = h.matvec_fun(xs[p])
        np.testing.assert_array_equal(ys[p].numpy(), yp)
    for order in (0, 1):
        for gm in (1, 3):
            for q in (0, 1):   # static grid stride or dynamic tile queue
                B.set_option("tile_order", order)
                B.set_option("grid_mult", gm)
                B.set_option("tile_queue", q)
                np.testing.assert_array_equal(B.matvec(x).cpu().numpy(), y)
                np.testing.assert_array_equal(B.matvec(x).cpu().numpy(), y)   # the counter restarts
    # sector tables (assembled operators) against on-the-fly rows: same
    # elements in the same order; complex multiply-adds may contract
    # differently in the two kernels
    assert all(h.assembly_bytes > 0 for h in hams)
    B.set_option("op_tables", 0)
    np.testing.assert_allclose(B.matvec(x).cpu().numpy(), y, rtol=0, atol=1e-13 * np.abs(y).max())
    B.close()
    assert all(h.assembly_bytes == 0 for h in hams)


def test_batched_momentum_sweep_e0_against_oracle():
    from pyqusolver_amd.batched import BatchedLanczos
    L = 12
    lat, specs, spaces, hams = _xxz_sectors(L)
    B = BatchedLanczos(hams, max_steps=200)
    res = B.ground_states(tol=1e-13)
    resid = B.residuals()
    olat = O.OLattice("chain", L)
    t = O.terms_xxz(olat, 1.0, 0.7)
    c, s, co = t.arrays()
    for p, spec in enumerate(specs):
        gens = [O.Generator("tx", spec["translation"])]
        if "reflection" in spec:
            gens.append(O.Generator("R", spec["reflection"]))
        b = O.scan_basis(O.Group(olat, gens, n_up=L // 2))
        ev = np.linalg.eigvalsh(O.reduced_dense(b, L, c, s, co, True))
        e_ref = ev[0]
        assert abs(res.e0[p] - e_ref) <= 1e-10 * max(1.0, abs(e_ref)), (spec, res.e0[p], e_ref)
        # the stated FP64 error bound: |E0 - lambda| <= residual for an eigenvalue
        # lambda, and every problem stopped under the residual test (or at its dimension)
        assert np.min(np.abs(ev - res.e0[p])) <= resid[p] * (1 + 1e-6) + 1e-14
        assert resid[p] <= 1e-13 * max(1.0, abs(res.e0[p])) or res.steps[p] >= len(b.reps)
    # the lowest sector energy is the Sz=0 ground state of the full space
    Hf = O.full_dense_hamiltonian(L, c, s, co, True)
    mask = np.array([bin(i).count("1") == L // 2 for i in range(1 << L)])
    e_full = np.linalg.eigvalsh(Hf[np.ix_(mask, mask)])[0]
    assert abs(res.e0.min() - e_full) <= 1e-10 * abs(e_full)


def test_batched_disorder_realisations():
    from pyqusolver_amd import ChainLattice, get_u1_sym
    from pyqusolver_amd.batched import BatchedLanczos, sector_by_sector
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import XXZ
    L = 10
    lat = ChainLattice(L)
    hs = HilbertSpace(lattice=lat, global_syms=[get_u1_sym(lat, L // 2)])
    fields = [np.random.default_rng(1000 + r).uniform(-5, 5, L) for r in range(12)]
    hams = [XXZ(lat, hilbert_space=hs, jxy=1.0, jz=0.7, hz=h) for h in fields]
    B = BatchedLanczos(hams, max_steps=252)
    res = B.ground_states(tol=1e-13)
    ref = sector_by_sector(hams, tol=1e-13, max_steps=252)
    olat = O.OLattice("chain", L)
    b = O.scan_basis(O.Group(olat, [], n_up=L // 2))
    for r, h in enumerate(fields):
        t = O.terms_xxz(olat, 1.0, 0.7, hz=h)
        c, s, co = t.arrays()
        e_ref = np.linalg.eigvalsh(O.reduced_dense(b, L, c, s, co, True))[0]
        assert abs(res.e0[r] - e_ref) <= 1e-10 * abs(e_ref)
        assert abs(ref.e0[r] - e_ref) <= 1e-10 * abs(e_ref)


def test_batched_step_budget_is_respected():
    from pyqusolver_amd.batched import BatchedLanczos
    lat, specs, spaces, hams = _xxz_sectors(10)
    B = BatchedLanczos(hams, max_steps=5)
    B.start()
    B._h and __import__("pyqusolver_amd._lib", fromlist=["lib"]).lib().qes_batch_run(B._h, 9, 1)
    a, b, k = B.coefficients()
    assert np.all(k <= 5)


def _realisations(L, R, seed=1000, jz=0.7, complex_=False):
    from pyqusolver_amd import ChainLattice, get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import XXZ
    lat = ChainLattice(L)
    hs = HilbertSpace(lattice=lat, global_syms=[get_u1_sym(lat, L // 2)])
    fields = [np.random.default_rng(seed + r).uniform(-5, 5, L) for r in range(R)]
    hams = [XXZ(lat, hilbert_space=hs, jxy=1.0, jz=jz, hz=h) for h in fields]
    return lat, hs, fields, hams


@pytest.mark.parametrize("R", [2, 3, 64, 100, 256, 300])
def test_block_matvec_equals_individual(R):
    """Realisations sharing the off-diagonal part run as (n, R) blocks; every
    column must equal the single-operator matvec to round-off (the diagonal
    enters from qes_op_diagonal instead of the term groups)."""
    import torch
    from pyqusolver_amd.batched import BatchedLanczos
    lat, hs, fields, hams = _realisations(8, R)
    B = BatchedLanczos(hams, max_steps=20)
    assert len(B.blocks) == (R + 255) // 256 and not B.regular
    rng = np.random.default_rng(R)
    xs = [rng.standard_normal(hs.nh) for _ in hams]
    x = B.concat(xs)
    yt = B.matvec(x)
    y = B.split(yt)
    for r, h in enumerate(hams):
        ref = h.matvec_fun(xs[r])
        np.testing.assert_allclose(y[r].cpu().numpy(), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    # the cached neighbour table reproduces the on-the-fly rows bit for bit
    for rows in (1, 16, 100):
        B.set_option("block_rows", rows)
        B.set_option("block_table", 0)
        assert torch.equal(B.matvec(x), yt)
        B.set_option("block_table", 1)
        assert torch.equal(B.matvec(x), yt)


def test_mixed_batch_sectors_and_blocks():
    """Momentum sectors (ordinary problems) and a block of realisations in one
    batch: results come back in the order of the input list."""
    import torch
    from pyqusolver_amd.batched import BatchedLanczos
    _, specs, spaces, sect = _xxz_sectors(10)
    _, hs, fields, real = _realisations(10, 5)
    hams = real[:2] + sect + real[2:]
    B = BatchedLanczos(hams, max_steps=252)
    assert len(B.blocks) == 1 and len(B.blocks[0]) == 5
    res = B.ground_states(tol=1e-13)
    Bu = BatchedLanczos(hams, max_steps=252, blocks=False)
    assert not Bu.blocks
    ru = Bu.ground_states(tol=1e-13)
    np.testing.assert_allclose(res.e0, ru.e0, rtol=1e-10, atol=0)
    olat = O.OLattice("chain", 10)
    b = O.scan_basis(O.Group(olat, [], n_up=5))
    for pos, r in ((0, 0), (1, 1), (len(sect) + 2, 2), (len(sect) + 4, 4)):
        c, s, co = O.terms_xxz(olat, 1.0, 0.7, hz=fields[r]).arrays()
        e_ref = np.linalg.eigvalsh(O.reduced_dense(b, 10, c, s, co, True))[0]
        assert abs(res.e0[pos] - e_ref) <= 1e-10 * abs(e_ref)


def test_block_complex_vectors():
    """complex128 vectors on a block of real realisations (complex_=True)."""
    from pyqusolver_amd.batched import BatchedLanczos
    lat, hs, fields, hams = _realisations(8, 7)
    B = BatchedLanczos(hams, max_steps=20, complex_=True)
    assert B.dtype == 1 and len(B.blocks) == 1
    rng = np.random.default_rng(5)
    xs = [rng.standard_normal(hs.nh) + 1j * rng.standard_normal(hs.nh) for _ in hams]
    y = B.split(B.matvec(B.concat(xs)))
    for r, h in enumerate(hams):
        ref = h.matvec_fun(xs[r].real) + 1j * h.matvec_fun(xs[r].imag)
        np.testing.assert_allclose(y[r].cpu().numpy(), ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def _j1j2_chain_sectors(L, J2=0.5):
    from pyqusolver_amd import ChainLattice, get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import J1J2Model
    lat = ChainLattice(L)
    spaces = [HilbertSpace(lattice=lat, sym_gen={"translation": k}, global_syms=[get_u1_sym(lat, L // 2)])
              for k in range(L)]
    hams = [J1J2Model(lat, J1=-1.0, J2=-J2, hilbert_space=hs) for hs in spaces]
    return lat, spaces, hams


def test_coupling_sweep_matvec_equals_scaled_products():
    """Column r of a coupling-scaled block is H_r x_r with H_r = sum_c s_rc H_c:
    equal to the single-operator scaled product over the same tagged table."""
    import torch
    from pyqusolver_amd.batched import CouplingSweep
    lat, spaces, hams = _j1j2_chain_sectors(12)
    scales = np.array([[1.0, 0.8], [1.0, 1.0], [1.0, 1.2], [0.5, 0.0]])
    S = CouplingSweep(hams[:5], scales, max_steps=40)
    rng = np.random.default_rng(2)
    xs = []
    for h in S.hams:
        v = rng.standard_normal(h.nh) + (1j * rng.standard_normal(h.nh) if S.dtype == 1 else 0.0)
        xs.append(v)
    y = S.split(S.matvec(S.concat(xs)))
    dt = torch.complex128 if S.dtype == 1 else torch.float64
    for p, h in enumerate(S.sectors):
        for r in range(len(scales)):
            v = p * len(scales) + r
            ref = h.matvec_scaled_device(scales[r], torch.as_tensor(xs[v], dtype=dt, device="cuda"))
            assert torch.allclose(y[v], ref, rtol=1e-13, atol=1e-13 * ref.abs().max().item()), (p, r)
    S.close()
    assert not any(h.assembly["couplings"] for h in S.sectors)


def test_coupling_sweep_ground_states_against_oracle():
    """All 12 momentum sectors of the J1-J2 chain L=12 (Sz=0) at J2 in
    {0.3, 0.5, 0.7} in one sweep (complex sectors included): E0 of every
    (k, J2) within 1e-10 of the dense reduced matrix of the oracle."""
    from pyqusolver_amd.batched import CouplingSweep
    L = 12
    lat, spaces, hams = _j1j2_chain_sectors(L, J2=0.5)
    j2s = [0.3, 0.5, 0.7]
    scales = np.array([[1.0, j / 0.5] for j in j2s])
    S = CouplingSweep(hams, scales, max_steps=200)
    res = S.ground_states(tol=1e-12)
    e = res.e0.reshape(L, len(j2s))
    olat = O.OLattice("chain", L)
    for k in range(L):
        grp = O.Group(olat, [O.Generator("tx", k)], n_up=L // 2)
        b = O.scan_basis(grp)
        for r, j2 in enumerate(j2s):
            c, s, co = O.terms_j1j2(olat, -1.0, -j2).arrays()
            ref = np.linalg.eigvalsh(O.reduced_dense(b, L, c, s, co, True))[0]
            assert abs(e[k, r] - ref) <= 1e-10 * abs(ref), (k, j2, e[k, r], ref)
    # k and -k sectors agree (time reversal)
    for k in range(1, L // 2):
        np.testing.assert_allclose(e[k], e[L - k], rtol=1e-11)
    S.close()


def test_coupling_sweep_square_against_separate_solves():
    """4x4 J1-J2 square, four momentum sectors x three J2 values in one
    sweep, against a separate Lanczos solve of a model built at each J2."""
    from pyqusolver_amd import SquareLattice, get_u1_sym
    from pyqusolver_amd.batched import CouplingSweep
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import J1J2Model
    lat = SquareLattice(4, 4)
    ks = [(0, 0), (1, 0), (2, 2), (1, 3)]
    spaces = [HilbertSpace(lattice=lat, sym_gen={"translation": {"kx": kx, "ky": ky}},
                           global_syms=[get_u1_sym(lat, 8)]) for kx, ky in ks]
    hams = [J1J2Model(lat, J1=-1.0, J2=-0.5, hilbert_space=hs) for hs in spaces]
    j2s = [0.4, 0.5, 0.6]
    S = CouplingSweep(hams, [[1.0, j / 0.5] for j in j2s], max_steps=300)
    e = S.ground_states(tol=1e-11).e0.reshape(len(ks), len(j2s))
    S.close()
    for p, hs in enumerate(spaces):
        for r, j2 in enumerate(j2s):
            ref = J1J2Model(lat, J1=-1.0, J2=-j2, hilbert_space=hs).diagonalize(method="lanczos", k=1, tol=1e-12, maxiter=300)[0]
            assert abs(e[p, r] - ref) <= 1e-10 * abs(ref), (ks[p], j2)


def test_coupling_sweep_field_scan_against_separate_models():
    """A field scan through the coupling sweep: XXZ L=10 (no U(1)) with
    Jxy, Jz and a uniform field hz tagged separately, four hz values per
    sector and Jz rescaled in one of them, against models built at each value."""
    from pyqusolver_amd import ChainLattice
    from pyqusolver_amd.batched import CouplingSweep
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import XXZ
    lat = ChainLattice(10)
    spaces = [HilbertSpace(lattice=lat, sym_gen={"translation": k}) for k in (0, 3, 5)]
    hams = [XXZ(lat, hilbert_space=hs, jxy=1.0, jz=0.7, hz=0.3) for hs in spaces]
    assert hams[0].n_couplings == 4          # jxy, jz, hx (absent: no instructions), hz
    sets = [(0.7, 0.0), (0.7, 0.3), (0.7, 1.1), (1.2, 0.5)]
    scales = [[1.0, jz / 0.7, 1.0, hz / 0.3] for jz, hz in sets]
    S = CouplingSweep(hams, scales, max_steps=300)
    e = S.ground_states(tol=1e-11).e0.reshape(len(hams), len(sets))
    S.close()
    for p, hs in enumerate(spaces):
        for r, (jz, hz) in enumerate(sets):
            ref = XXZ(lat, hilbert_space=hs, jxy=1.0, jz=jz, hz=hz).diagonalize(method="lanczos", k=1, tol=1e-12, maxiter=300)[0]
            assert abs(e[p, r] - ref) <= 1e-10 * max(1.0, abs(ref)), (p, jz, hz)


def test_autotuned_schedule_changes_speed_only():
    """The schedule autotuner picks tile order / blocks per CU / block rows
    by timing the batch's own product; products and ground states are the
    same under the tuned schedule as under the default one."""
    import torch
    from pyqusolver_amd import ChainLattice, get_u1_sym
    from pyqusolver_amd.batched import BatchedLanczos
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import XXZ
    lat, specs, spaces, hams = _xxz_sectors(12)
    real = [h for h in hams if not h.needs_complex]
    lat10 = ChainLattice(10)
    hs10 = HilbertSpace(lattice=lat10, global_syms=[get_u1_sym(lat10, 5)])
    real += [XXZ(lat10, hilbert_space=hs10, jxy=1.0, jz=0.7, hz=np.random.default_rng(r).uniform(-2, 2, 10))
             for r in range(8)]
    B = BatchedLanczos(real, max_steps=120)
    x = B.concat([np.random.default_rng(i).standard_normal(h.nh) for i, h in enumerate(real)])
    y0 = B.matvec(x)
    r0 = B.ground_states(tol=1e-10)
    sched = B.autotune(reps=2)
    assert set(sched) >= {"tile_order", "grid_mult", "block_rows", "ms"} and sched["ms"] > 0
    assert torch.equal(B.matvec(x), y0)
    r1 = B.ground_states(tol=1e-10)
    np.testing.assert_array_equal(r1.e0, r0.e0)
    np.testing.assert_array_equal(r1.steps, r0.steps)
    B.close()


def test_batched_ground_states_split_batches_with_error_bounds():
    """batched_ground_states: real and complex problems in two batches on two
    streams (threads), realisation blocks inside the real one; E0 equal to
    separate solves to 1e-10, and each result carries its residual bound
    (|E0 - lambda| <= r), all under the stopping tolerance."""
    from pyqusolver_amd import ChainLattice, get_u1_sym
    from pyqusolver_amd.batched import batched_ground_states, sector_by_sector
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import XXZ
    lat, specs, spaces, hams = _xxz_sectors(12)
    lat10 = ChainLattice(10)
    hs10 = HilbertSpace(lattice=lat10, global_syms=[get_u1_sym(lat10, 5)])
    hams = list(hams) + [XXZ(lat10, hilbert_space=hs10, jxy=1.0, jz=0.7,
                             hz=np.random.default_rng(50 + r).uniform(-3, 3, 10)) for r in range(6)]
    assert any(h.needs_complex for h in hams) and not all(h.needs_complex for h in hams)
    res, t_create = batched_ground_states(hams, tol=1e-11, max_steps=250, autotune=False)
    ref = sector_by_sector(hams, tol=1e-11, max_steps=250)
    np.testing.assert_allclose(res.e0, ref.e0, rtol=0, atol=1e-10 * np.max(np.abs(ref.e0)))
    assert res.residuals is not None and res.residuals.shape == (len(hams),)
    assert np.all(res.residuals <= 1e-11 * np.maximum(1.0, np.abs(res.e0)) * (1 + 1e-9))
    assert t_create >= 0.0 and res.seconds > 0.0


def test_one_state_sector_of_positive_energy_after_breakdown():
    """The saturated sector of an antiferromagnetic XXZ ring (n_up = L, one
    state, E = L/4 > 0) breaks down after one step; the batched and the plain
    C-ABI ground-state solves must not report the spurious 0 of the zero rows
    that follow (reference: E0 = -jz L/4 with H = -jz sum SzSz)."""
    import ctypes as C
    import torch
    from pyqusolver_amd import ChainLattice, _lib, get_u1_sym
    from pyqusolver_amd.batched import BatchedLanczos
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import XXZ
    L = 8
    lat = ChainLattice(L)
    hs1 = HilbertSpace(lattice=lat, sym_gen={"translation": 0}, global_syms=[get_u1_sym(lat, L)])
    assert hs1.nh == 1
    h1 = XXZ(lat, hilbert_space=hs1, jxy=1.0, jz=-1.0)
    hs4 = HilbertSpace(lattice=lat, sym_gen={"translation": 0}, global_syms=[get_u1_sym(lat, L // 2)])
    h4 = XXZ(lat, hilbert_space=hs4, jxy=1.0, jz=-1.0)
    B = BatchedLanczos([h1, h4], max_steps=60)
    res = B.ground_states(tol=1e-12, check_every=10)
    assert abs(res.e0[0] - L / 4) < 1e-13
    assert res.steps[0] == 1
    a, b, k = B.coefficients()
    assert k[0] == 1
    e_ref = np.linalg.eigvalsh(_dense(h4))[0]
    assert abs(res.e0[1] - e_ref) <= 1e-10 * abs(e_ref)
    B.close()
    # plain device solve through the C ABI
    v0 = torch.ones(1, dtype=torch.complex128, device="cuda")
    al = np.zeros(40)
    be = np.zeros(40)
    steps = C.c_int32()
    e0 = C.c_double()
    _lib.check(_lib.lib().qes_lanczos_ground_state(h1._device_op(), _lib.dptr(v0), 1, 40, 0, 1e-12, 10,
                                                   al.ctypes.data_as(_lib._PD), be.ctypes.data_as(_lib._PD),
                                                   C.byref(steps), C.byref(e0), _lib.stream_ptr()),
               "qes_lanczos_ground_state")
    assert steps.value == 1 and abs(e0.value - L / 4) < 1e-13
    # and the Python solver
    assert abs(h1.diagonalize(method="lanczos", k=1, tol=1e-12, maxiter=40)[0] - L / 4) < 1e-13


def _dense(h):
    n = h.nh
    cols = [h.matvec_fun(np.eye(n, dtype=np.complex128)[:, j]) for j in range(n)]
    return np.array(cols).T


def test_config3_batch_at_size():
    """The batched workload at full size: XXZ L=24 (Jxy=1, Jz=0.7, Sz=0), all
    24 momenta with reflection parities at k = 0 and 12 (26 sectors, complex
    at generic k), plus 256 random-field L=20 realisations (h_i ~ U[-5, 5]
    from seeds 1000 + r) as one (n, 256) block; every ground state by batched
    Lanczos against one device Lanczos per problem (1e-10 relative), E(k) =
    E(-k), every residual bound within the 1e-10 test, and one block column's
    product against the oracle's row formula on sampled rows."""
    import torch
    from pyqusolver_amd import ChainLattice, get_u1_sym
    from pyqusolver_amd.batched import BatchedLanczos, batched_ground_states, momentum_sectors, sector_by_sector
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import XXZ
    lat24 = ChainLattice(24)
    specs = momentum_sectors(24, reflection=True)
    assert len(specs) == 26
    spaces = [HilbertSpace(lattice=lat24, sym_gen=s, global_syms=[get_u1_sym(lat24, 12)]) for s in specs]
    assert sum(hs.nh for hs in spaces) == 2_704_156          # C(24, 12)
    hams = [XXZ(lat24, hilbert_space=h, jxy=1.0, jz=0.7) for h in spaces]
    lat20 = ChainLattice(20)
    hs20 = HilbertSpace(lattice=lat20, global_syms=[get_u1_sym(lat20, 10)])
    assert hs20.nh == 184_756
    fields = [np.random.default_rng(1000 + r).uniform(-5.0, 5.0, 20) for r in range(256)]
    hams20 = [XXZ(lat20, hilbert_space=hs20, jxy=1.0, jz=0.7, hz=h) for h in fields]
    res, _ = batched_ground_states(hams + hams20, tol=1e-10, max_steps=300)
    assert np.all(res.residuals <= 1e-10 * np.maximum(1.0, np.abs(res.e0)) * 1.0000001)
    seq = sector_by_sector(hams + hams20, tol=1e-10, max_steps=300)
    dev = np.abs(res.e0 - seq.e0) / np.maximum(1.0, np.abs(seq.e0))
    assert dev.max() <= 1e-10, dev.max()
    ek = {s["translation"]: e for s, e in zip(specs, res.e0[:26]) if "reflection" not in s}
    for k in range(1, 12):
        assert abs(ek[k] - ek[24 - k]) <= 1e-10 * abs(ek[k])
    # one realisation column of the block product against the oracle's rows
    B = BatchedLanczos(hams20, max_steps=10)
    assert len(B.blocks) == 1 and not B.regular
    rng = np.random.default_rng(5)
    xs = [rng.standard_normal(hs20.nh) for _ in hams20]
    ys = B.split(B.matvec(B.concat(xs)))
    B.close()
    olat = O.OLattice("chain", 20)
    grp = O.Group(olat, [], n_up=10)
    reps, norms = hs20.representative_list, hs20.normalization
    sample = np.sort(rng.choice(hs20.nh, 300, replace=False))
    for r in (0, 131, 255):
        c, s, co = O.terms_xxz(olat, 1.0, 0.7, hz=fields[r]).arrays()
        yo = O.sampled_rows(grp, reps, norms, sample, 20, c, s, co, True, xs[r])
        yg = ys[r].cpu().numpy()[sample]
        assert np.linalg.norm(yg - yo.real) <= 1e-12 * np.linalg.norm(yo), r


def test_square6x6_coupling_sweep_sector():
    """The 6x6 sweep at full size on two sectors: a complex one, k = (1, 2)
    (N = 126.0 M, complex128), whose three coupling columns (J2 = 0.4, 0.5,
    0.6 by scaling the J2 coupling) must equal the scaled products of its
    tagged table and, on sampled rows, the oracle's rows of models built at
    each J2; and k = (0, 0), whose J2 = 0.5 column's E0 from the sweep equals
    the plain solve's (1e-10 relative)."""
    import torch
    from pyqusolver_amd import SquareLattice, get_u1_sym
    from pyqusolver_amd.batched import CouplingSweep
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import J1J2Model
    lat = SquareLattice(6, 6)
    j2s = [0.4, 0.5, 0.6]
    scales = [[1.0, j / 0.5] for j in j2s]
    olat = O.OLattice("square", 6, 6)
    hs = HilbertSpace(lattice=lat, sym_gen={"translation": {"kx": 1, "ky": 2}, "parity": {"x": 1}},
                      global_syms=[get_u1_sym(lat, 18)])
    H = J1J2Model(lat, J1=-1.0, J2=-0.5, hilbert_space=hs)
    assert H.needs_complex
    S = CouplingSweep([H], scales, max_steps=20, max_gb=160.0)
    g = torch.Generator(device="cuda").manual_seed(9)
    x = torch.randn(3 * hs.nh, dtype=torch.complex128, device="cuda", generator=g)
    xs = [x[i * hs.nh:(i + 1) * hs.nh] for i in range(3)]
    y = S.split(S.matvec(S.concat(xs)))
    grp = O.Group(olat, [O.Generator("tx", 1), O.Generator("ty", 2), O.Generator("Px", 1)], n_up=18)
    reps, norms = hs.representative_list, hs.normalization
    sample = np.sort(np.random.default_rng(4).choice(hs.nh, 100, replace=False))
    for r, j2 in enumerate(j2s):
        ref = H.matvec_scaled_device(scales[r], xs[r].contiguous())
        assert torch.allclose(y[r], ref, rtol=1e-13, atol=1e-13 * ref.abs().max().item()), j2
        c, s, co = O.terms_j1j2(olat, -1.0, -j2).arrays()
        yo = O.sampled_rows(grp, reps, norms, sample, 36, c, s, co, True, xs[r].cpu().numpy())
        yg = y[r].cpu().numpy()[sample]
        assert np.linalg.norm(yg - yo) <= 1e-12 * np.linalg.norm(yo), j2
    S.close()
    del S, H, hs, x, xs, y, ref
    torch.cuda.empty_cache()
    hs0 = HilbertSpace(lattice=lat, sym_gen={"translation": {"kx": 0, "ky": 0}, "parity": {"x": 1}},
                       global_syms=[get_u1_sym(lat, 18)])
    H0 = J1J2Model(lat, J1=-1.0, J2=-0.5, hilbert_space=hs0)
    S0 = CouplingSweep([H0], scales, max_steps=300, max_gb=160.0)
    e = S0.ground_states(tol=1e-10).e0
    resid = S0.residuals()
    S0.close()
    assert np.all(resid <= 1e-10 * np.abs(e) * 1.0000001)
    e_plain = H0.diagonalize(method="lanczos", k=1, tol=1e-10, maxiter=300, seed=0, reorth=False,
                             store_basis=False, assemble=True)[0]
    assert abs(e[1] - e_plain) <= 1e-10 * abs(e_plain)
    assert e[0] < e[1] < e[2]          # E0 rises with J2 at k = 0 (frustration)
