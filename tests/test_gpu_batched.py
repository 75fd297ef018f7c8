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
            B.set_option("tile_order", order)
            B.set_option("grid_mult", gm)
            np.testing.assert_array_equal(B.matvec(x).cpu().numpy(), y)
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
    olat = O.OLattice("chain", L)
    t = O.terms_xxz(olat, 1.0, 0.7)
    c, s, co = t.arrays()
    for p, spec in enumerate(specs):
        gens = [O.Generator("tx", spec["translation"])]
        if "reflection" in spec:
            gens.append(O.Generator("R", spec["reflection"]))
        b = O.scan_basis(O.Group(olat, gens, n_up=L // 2))
        e_ref = np.linalg.eigvalsh(O.reduced_dense(b, L, c, s, co, True))[0]
        assert abs(res.e0[p] - e_ref) <= 1e-10 * max(1.0, abs(e_ref)), (spec, res.e0[p], e_ref)
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
