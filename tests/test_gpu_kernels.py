"""Kernel-variant cross-checks on the GPU: fast canonicalisers against the
brute-force group scan, matvec plans against each other, per-coupling
outputs against the full matvec."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _space(kind, lx, ly, sym, n_up):
    from pyqusolver_amd import ChainLattice, HoneycombLattice, SquareLattice, get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    lat = {"chain": lambda: ChainLattice(lx), "square": lambda: SquareLattice(lx, ly),
           "honeycomb": lambda: HoneycombLattice(lx, ly)}[kind]()
    gl = [get_u1_sym(lat, n_up)] if n_up is not None else []
    return lat, HilbertSpace(lattice=lat, sym_gen=sym, global_syms=gl)


def _random_states(ns, n, rng, n_up=None):
    if n_up is None:
        return rng.integers(0, 1 << ns, size=n, dtype=np.uint64) if ns < 64 else rng.integers(0, 2**63, size=n).astype(np.uint64)
    out = np.zeros(n, dtype=np.uint64)
    for i in range(n):
        bits = rng.choice(ns, n_up, replace=False)
        out[i] = np.uint64(np.sum(np.left_shift(np.uint64(1), bits.astype(np.uint64))))
    return out


SPACES = [
    ("chain", 32, 1, {"translation": 0, "parity": {"x": 1}}, 16, "cyclic"),
    ("chain", 32, 1, {"translation": 5}, 16, "cyclic"),
    ("chain", 24, 1, {"translation": 12, "reflection": 1, "parity": {"x": -1}}, 12, "cyclic"),
    ("chain", 20, 1, {"translation": 3}, None, "cyclic"),
    ("chain", 40, 1, {"translation": 0}, 2, "cyclic"),
    ("chain", 64, 1, {"translation": 7}, 2, "cyclic"),
    ("square", 6, 6, {"translation": {"kx": 0, "ky": 0}, "parity": {"x": 1}}, 18, "rows"),
    ("square", 6, 6, {"translation": {"kx": 2, "ky": 3}}, 18, "rows"),
    ("square", 4, 4, {"translation": {"kx": 0, "ky": 2}, "reflection": 1, "parity": {"x": -1}}, 8, "rows"),
    ("square", 4, 3, {"translation": {"ky": 1}}, 6, "rows"),
    ("square", 4, 3, {"translation": {"kx": 1}}, 6, "rows"),
    ("honeycomb", 3, 4, {"translation": {"kx": 0, "ky": 0}}, None, "rows"),
    ("honeycomb", 2, 6, {"translation": {"kx": 1, "ky": 3}}, None, "rows"),
]


@pytest.mark.parametrize("case", SPACES, ids=[f"{c[0]}{c[1]}x{c[2]}-{c[5]}-{i}" for i, c in enumerate(SPACES)])
def test_fast_canonicaliser_matches_brute_force(case):
    kind, lx, ly, sym, n_up, ck = case
    # long chains: a tiny U(1) sector carries the group; random test states
    # are drawn at half filling regardless
    lat, hs = _space(kind, lx, ly, sym, n_up)
    rng = np.random.default_rng(11)
    st = _random_states(lat.ns, 20000, rng, None if n_up is None else lat.ns // 2)
    assert hs.canon_check(st, ck) == 0
    # all single-bond flips of representatives (the states the matvec sees)
    reps = hs.representative_list[:5000].astype(np.uint64)
    flips = [reps ^ np.uint64((1 << i) | (1 << ((i + 1) % lat.ns))) for i in range(lat.ns)]
    assert hs.canon_check(np.concatenate(flips), ck) == 0


@pytest.mark.parametrize("case", [SPACES[0], SPACES[2], SPACES[8], SPACES[11]], ids=lambda c: f"{c[0]}{c[1]}x{c[2]}")
def test_matvec_plan_variants_agree(case):
    import torch
    from pyqusolver_amd.models import HeisenbergKitaev, J1J2Model
    kind, lx, ly, sym, n_up, ck = case
    lat, hs = _space(kind, lx, ly, sym, n_up)
    if kind == "honeycomb":
        H = HeisenbergKitaev(lat, K=-1.0, J=0.2, hilbert_space=hs, hx=0.03, hy=0.02, hz=0.05)
    else:
        H = J1J2Model(lat, J1=-1.0, J2=-0.4, hilbert_space=hs)
    cplx = H.needs_complex
    dt = torch.complex128 if cplx else torch.float64
    x = torch.randn(hs.nh, dtype=dt, device="cuda")
    assert H.plan["canonicalizer"] == ck
    y_fast = H.matvec_device(x)
    H.set_canonicalizer("brute")
    y_brute = H.matvec_device(x)
    # same elements; the row lattices' top-row test moves some lookups to the
    # canonicalising pass, so the sums run in another order: equal to rounding
    assert (torch.linalg.vector_norm(y_fast - y_brute) / torch.linalg.vector_norm(y_brute)).item() <= 1e-14
    H.set_canonicalizer("auto")
    # every representative lookup finds the same index
    from pyqusolver_amd._lib import QesError
    for mode in ("sorted", "hash", "packed"):
        for rows_log2 in (-4, 0, 3, 6):
            try:
                hs.set_lookup(mode, bucket_rows_log2=rows_log2)
            except QesError:
                continue   # mode not available for this basis
            # the rank map's direct pass sums in another order: equal to rounding
            y = H.matvec_device(x)
            rel = (torch.linalg.vector_norm(y - y_fast) / torch.linalg.vector_norm(y_fast)).item()
            assert rel <= 1e-14, (mode, rows_log2, rel)


def test_per_coupling_outputs_sum_to_matvec():
    import torch
    from pyqusolver_amd.models import J1J2Model
    lat, hs = _space("chain", 24, 1, {"translation": 0, "parity": {"x": 1}}, 12)
    H = J1J2Model(lat, J1=-1.0, J2=-0.45, hilbert_space=hs, hz=None)
    x = torch.randn(hs.nh, dtype=torch.float64, device="cuda")
    Y = H.matvec_couplings_device(x)
    assert Y.shape[0] == H.n_couplings
    y = H.matvec_device(x)
    assert torch.allclose(Y.sum(0), y, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("case", [SPACES[0], SPACES[2], SPACES[8], SPACES[11], SPACES[9]],
                         ids=lambda c: f"{c[0]}{c[1]}x{c[2]}")
def test_assembled_operator_matches_matrix_free(case):
    """The sliced-ELLPACK table holds the same rows in the same order as the
    matrix-free visitor: products equal to rounding (the table holds each
    element with N_b / N_a applied; the matrix-free product gathers N_b x_b)."""
    import torch
    from pyqusolver_amd.models import HeisenbergKitaev, J1J2Model
    kind, lx, ly, sym, n_up, ck = case
    lat, hs = _space(kind, lx, ly, sym, n_up)
    if kind == "honeycomb":
        H = HeisenbergKitaev(lat, K=-1.0, J=0.2, hilbert_space=hs, hx=0.03, hy=0.02, hz=0.05)
    else:
        H = J1J2Model(lat, J1=-1.0, J2=-0.4, hilbert_space=hs)
    dt = torch.complex128 if H.needs_complex else torch.float64
    x = torch.randn(hs.nh, dtype=dt, device="cuda")
    y_free = H.matvec_device(x)
    nbytes = H.assemble()
    assert nbytes > 0
    y_asm = H.matvec_device(x)
    # the table stores conj(val) chi N_b / N_a per element; the matrix-free
    # product sums conj(val) chi (N_b x_b) and divides by N_a once: equal to rounding
    assert torch.allclose(y_asm, y_free, rtol=1e-14, atol=1e-14 * y_free.abs().max().item())
    # accumulate and multi-RHS (matrix-free) still agree
    y2 = y_asm.clone()
    H.matvec_device(x, out=y2, accumulate=True)
    assert torch.allclose(y2, 2 * y_free, rtol=1e-14, atol=1e-14 * y_free.abs().max().item())
    X = torch.stack([x, 2 * x], dim=1).contiguous()
    Y = H.matvec_device(X)
    assert torch.allclose(Y[:, 0], y_free, rtol=1e-14, atol=1e-14 * y_free.abs().max().item())
    # a plan change drops the table
    H.set_option("xcd_remap", 0)
    assert H.assembly_bytes == 0
    H.assemble(False)


@pytest.mark.parametrize("case", [SPACES[0], SPACES[1], SPACES[9], SPACES[11], SPACES[12]],
                         ids=lambda c: f"{c[0]}{c[1]}x{c[2]}-{c[3]}")
def test_coupling_tagged_table(case):
    """A coupling-tagged table (values keyed by coupling id, one diagonal per
    coupling): its plain product is the matrix-free one, the scaled product
    is sum_c s_c H_c x with H_c from the matrix-free per-coupling kernel, and
    the coupling dots are <g, H_c v>.  Complex operators and complex
    characters take the complex value dictionary."""
    import torch
    from pyqusolver_amd.models import HeisenbergKitaev, J1J2Model
    kind, lx, ly, sym, n_up, ck = case
    lat, hs = _space(kind, lx, ly, sym, n_up)
    if kind == "honeycomb":
        H = HeisenbergKitaev(lat, K=-1.0, J=0.2, hilbert_space=hs, hx=0.03, hy=0.02, hz=0.05, split_couplings=True)
    else:
        H = J1J2Model(lat, J1=-1.0, J2=-0.4, hilbert_space=hs)
    nc = H.n_couplings
    assert nc >= 2
    dt = torch.complex128 if H.needs_complex else torch.float64
    g = torch.Generator(device="cuda").manual_seed(21)
    x = torch.randn(hs.nh, dtype=dt, device="cuda", generator=g)
    w = torch.randn(hs.nh, dtype=dt, device="cuda", generator=g)
    y_free = H.matvec_device(x)
    Y = H.matvec_couplings_device(x)
    H.assemble(True, couplings=True)
    info = H.assembly
    assert info["couplings"] and info["format"] in ("packed32", "index16"), info
    scale_tol = 1e-13 * y_free.abs().max().item()
    y_asm = H.matvec_device(x)
    # the table stores conj(val) chi N_b / N_a per element; the matrix-free
    # product sums conj(val) chi (N_b x_b) and divides by N_a once: equal to rounding
    assert torch.allclose(y_asm, y_free, rtol=1e-14, atol=1e-14 * y_free.abs().max().item())
    for s in (np.ones(nc), np.linspace(0.5, 1.7, nc), np.eye(nc)[nc - 1]):
        ys = H.matvec_scaled_device(s, x)
        ref = torch.tensordot(torch.as_tensor(s, device="cuda").to(dt), Y, dims=([0], [0]))
        assert torch.allclose(ys, ref, rtol=1e-13, atol=scale_tol), s
    # unchanged scales reuse the uploaded dictionary; the next call still sees new ones
    s = np.linspace(0.5, 1.7, nc)
    assert torch.equal(H.matvec_scaled_device(s, x), H.matvec_scaled_device(s.copy(), x))
    d = H.coupling_dots_device(x, w)
    ref = np.array([torch.vdot(w, Y[c]).item() if dt == torch.complex128 else torch.dot(w, Y[c]).item()
                    for c in range(nc)])
    assert np.allclose(d, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    # the plain table product is unchanged by a scaled call in between
    assert torch.equal(H.matvec_device(x), y_asm)
    H.assemble(False)
    assert not H.assembly["couplings"]


@pytest.mark.parametrize("case", [SPACES[0], SPACES[1], SPACES[11]], ids=lambda c: f"{c[0]}{c[1]}x{c[2]}-{c[3]}")
@pytest.mark.parametrize("fmt", ["auto", "plain"])
def test_assembled_multi_rhs_equals_single_columns(case, fmt):
    """Several right-hand sides through the assembled table (four columns per
    table pass): every column equals its single-vector table product bit for
    bit, for widths that leave a remainder group; accumulate adds."""
    import torch
    from pyqusolver_amd.models import HeisenbergKitaev, J1J2Model
    kind, lx, ly, sym, n_up, ck = case
    lat, hs = _space(kind, lx, ly, sym, n_up)
    if kind == "honeycomb":
        H = HeisenbergKitaev(lat, K=-1.0, J=0.2, hilbert_space=hs, hx=0.03, hy=0.02, hz=0.05)
    else:
        H = J1J2Model(lat, J1=-1.0, J2=-0.4, hilbert_space=hs)
    # one-pass table: the single-vector product then sums in the multi-vector kernel's order
    H.set_option("asm_panel_cols", 0)
    H.assemble(True, dictionary=(fmt == "auto"))
    dt = torch.complex128 if H.needs_complex else torch.float64
    g = torch.Generator(device="cuda").manual_seed(8)
    for nb in (2, 4, 7):
        X = torch.randn(hs.nh, nb, dtype=dt, device="cuda", generator=g)
        Y = H.matvec_device(X)
        for c in range(nb):
            assert torch.equal(Y[:, c], H.matvec_device(X[:, c].contiguous())), (nb, c)
        Y2 = Y.clone()
        H.matvec_device(X, out=Y2, accumulate=True)
        assert torch.allclose(Y2, 2 * Y, rtol=1e-14, atol=1e-14 * Y.abs().max().item())
    H.assemble(False)


@pytest.mark.parametrize("case", [SPACES[2], SPACES[8], SPACES[11], SPACES[12]],
                         ids=lambda c: f"{c[0]}{c[1]}x{c[2]}-{c[3]}")
@pytest.mark.parametrize("fmt", ["auto", "plain", "couplings"])
def test_column_panels_match_one_pass_table(case, fmt):
    """Column panels (forced narrow, so that small sectors take five passes):
    the panelled single-vector product equals the one-pass table's to
    rounding (each row's sum is split by column range), accumulate adds,
    the scaled product of a coupling-tagged table agrees, a Lanczos solve
    through the panels gives the same ground state, and more than 32 panels
    are refused."""
    import torch
    from pyqusolver_amd import _lib
    from pyqusolver_amd.lanczos import lanczos_ground_state
    from pyqusolver_amd.models import HeisenbergKitaev, J1J2Model
    kind, lx, ly, sym, n_up, ck = case
    lat, hs = _space(kind, lx, ly, sym, n_up)
    if kind == "honeycomb":
        H = HeisenbergKitaev(lat, K=-1.0, J=0.2, hilbert_space=hs, hx=0.03, hy=0.02, hz=0.05, split_couplings=True)
    else:
        H = J1J2Model(lat, J1=-1.0, J2=-0.4, hilbert_space=hs)
    dt = torch.complex128 if H.needs_complex else torch.float64
    g = torch.Generator(device="cuda").manual_seed(5)
    x = torch.randn(hs.nh, dtype=dt, device="cuda", generator=g)
    kw = {"dictionary": fmt != "plain", "couplings": fmt == "couplings"}
    nc = H.n_couplings
    scales = np.linspace(0.6, 1.4, nc)
    H.set_option("asm_panel_cols", 0)
    H.assemble(True, **kw)
    assert H.assembly["panels"] == 0
    y1 = H.matvec_device(x)
    ys1 = H.matvec_scaled_device(scales, x) if fmt == "couplings" else None
    e1 = lanczos_ground_state(H, k=1, tol=1e-11, maxiter=200, seed=2).eigenvalues[0]
    H.assemble(False)
    cols = hs.nh // 5 + 1
    H.set_option("asm_panel_cols", cols)
    H.assemble(True, **kw)
    info = H.assembly
    assert info["panels"] == 5 and info["panel_cols"] == cols and info["panel_entries"] > 0, info
    y2 = H.matvec_device(x)
    tol = 1e-14 * y1.abs().max().item()
    assert torch.allclose(y2, y1, rtol=1e-14, atol=tol)
    assert torch.equal(H.matvec_device(x), y2)   # deterministic
    y3 = y2.clone()
    H.matvec_device(x, out=y3, accumulate=True)
    assert torch.allclose(y3, 2 * y1, rtol=1e-14, atol=2 * tol)
    if fmt == "couplings":
        ys2 = H.matvec_scaled_device(scales, x)
        assert torch.allclose(ys2, ys1, rtol=1e-14, atol=1e-14 * ys1.abs().max().item())
    e2 = lanczos_ground_state(H, k=1, tol=1e-11, maxiter=200, seed=2).eigenvalues[0]
    assert abs(e2 - e1) <= 1e-10 * abs(e1)
    H.assemble(False)
    assert H.assembly["panels"] == 0
    assert hs.nh >= 40
    H.set_option("asm_panel_cols", hs.nh // 40)
    with pytest.raises(_lib.QesError, match="panels"):
        H.assemble(True, **kw)
    assert H.assembly_bytes == 0
    H.set_option("asm_panel_cols", -1)


@pytest.mark.parametrize("case", [SPACES[0], SPACES[1], SPACES[2], SPACES[8], SPACES[9]],
                         ids=lambda c: f"{c[0]}{c[1]}x{c[2]}-{c[3]}")
def test_rank_gather_equals_dense_gather(case):
    """Matrix-free products that gather from the basis's rank-indexed vector
    (N_b x_b at the colex rank of each representative, a NaN marker
    elsewhere) equal the dense pre-scaled gathers bit for bit (same
    elements, same order); an x element whose bits equal the marker is
    still a NaN of the output, not a hole; accumulate adds; the option is
    refused where the basis has no rank map."""
    import torch
    from pyqusolver_amd import _lib
    from pyqusolver_amd.models import J1J2Model
    kind, lx, ly, sym, n_up, ck = case
    lat, hs = _space(kind, lx, ly, sym, n_up)
    H = J1J2Model(lat, J1=-1.0, J2=-0.4, hilbert_space=hs, hx=0.0)
    dt = torch.complex128 if H.needs_complex else torch.float64
    g = torch.Generator(device="cuda").manual_seed(17)
    x = torch.randn(hs.nh, dtype=dt, device="cuda", generator=g)
    y0 = H.matvec_device(x)
    H.set_option("rank_gather", 1)
    y1 = H.matvec_device(x)
    assert torch.equal(y1, y0)
    x2 = torch.randn(hs.nh, dtype=dt, device="cuda", generator=g)   # slots rewritten, no stale values
    y2 = H.matvec_device(x2)
    H.set_option("rank_gather", 0)
    assert torch.equal(y2, H.matvec_device(x2))
    H.set_option("rank_gather", 1)
    y3 = y1.clone()
    H.matvec_device(x, out=y3, accumulate=True)
    assert torch.allclose(y3, 2 * y0, rtol=1e-14, atol=1e-14 * y0.abs().max().item())
    # an element carrying the marker's bits: the products still see a NaN there
    xm = x.clone()
    xv = xm.view(torch.float64) if dt == torch.float64 else torch.view_as_real(xm).reshape(-1)
    xv.view(torch.int64)[0] = 0x7FFC5E17A11C0DE5
    ym = H.matvec_device(xm)
    H.set_option("rank_gather", 0)
    yd = H.matvec_device(xm)
    assert torch.equal(torch.isnan(ym), torch.isnan(yd)) and bool(torch.isnan(yd).any())
    fin = ~torch.isnan(yd)
    assert torch.equal(ym[fin], yd[fin])
    # with the rank map switched off the option falls back to the dense gathers
    H.set_option("rank_gather", 1)
    hs.set_lookup("sorted")
    y4 = H.matvec_device(x)
    assert torch.allclose(y4, y0, rtol=1e-14, atol=1e-14 * y0.abs().max().item())
    hs.set_lookup("rank")
    assert torch.equal(H.matvec_device(x), y0)
    H.set_option("rank_gather", 0)


def test_rank_gather_refused_without_rank_map():
    from pyqusolver_amd import _lib
    from pyqusolver_amd.models import J1J2Model
    lat, hs = _space("chain", 20, 1, {"translation": 3}, None)
    H = J1J2Model(lat, J1=-1.0, J2=-0.4, hilbert_space=hs)
    with pytest.raises(_lib.QesError, match="rank map"):
        H.set_option("rank_gather", 1)


def test_count_outside_sector():
    """Debug count of nonzero elements whose target leaves the sector (the
    reference skips them silently): zero for a U(1)-conserving model at k=0;
    for a transverse field in a U(1) sector every sigma^x_i leaves it, so the
    count is N * ns exactly; the gathers inside are unchanged by the field."""
    from pyqusolver_amd.models import J1J2Model
    lat, hs = _space("chain", 12, 1, {"translation": 0}, 6)
    H0 = J1J2Model(lat, J1=-1.0, J2=-0.4, hilbert_space=hs)
    assert H0.count_outside() == 0
    Hx = J1J2Model(lat, J1=-1.0, J2=-0.4, hilbert_space=hs, hx=0.2)
    assert Hx.count_outside() == hs.nh * 12
    assert Hx.count_offdiag() == H0.count_offdiag()


@pytest.mark.parametrize("case", [SPACES[0], SPACES[1], SPACES[2], SPACES[4], SPACES[7], SPACES[8], SPACES[10]],
                         ids=lambda c: f"{c[0]}{c[1]}x{c[2]}-{c[3]}")
def test_prescaled_product_lookup_batches_agree(case):
    """qes_matvec_ws gathers N_b x_b from a pre-scaled copy, with 1, 2 or 4
    rank-map lookups in flight per lane: the same sums in the same order for
    every batch size (bitwise equal for real vectors), and equal to the unscaled product of
    qes_matvec to rounding; the oracle checks the product itself elsewhere."""
    import torch
    from pyqusolver_amd import _lib
    from pyqusolver_amd.models import J1J2Model, XXZ
    kind, lx, ly, sym, n_up, ck = case
    lat, hs = _space(kind, lx, ly, sym, n_up)
    H = XXZ(lat, hilbert_space=hs, jxy=1.0, jz=0.7) if kind == "chain" else J1J2Model(lat, J1=-1.0, J2=-0.4, hilbert_space=hs)
    cplx = H.needs_complex
    dt = torch.complex128 if cplx else torch.float64
    x = torch.randn(hs.nh, dtype=dt, device="cuda", generator=torch.Generator(device="cuda").manual_seed(8))
    y_plain = torch.empty_like(x)
    _lib.check(_lib.lib().qes_matvec(H._device_op(), _lib.dptr(x), _lib.dptr(y_plain), 1, 1 if cplx else 0, 0,
                                     _lib.stream_ptr()), "qes_matvec")
    outs = []
    for kb in (1, 2, 4, -1):
        H.set_option("direct_batch", kb)
        outs.append(H.matvec_device(x))
    for y in outs[1:]:
        if cplx:   # the complex multiply-adds may contract differently in the two code paths
            assert (torch.linalg.vector_norm(y - outs[0]) / torch.linalg.vector_norm(outs[0])).item() <= 1e-15
        else:
            assert torch.equal(y, outs[0])
    rel = (torch.linalg.vector_norm(outs[0] - y_plain) / torch.linalg.vector_norm(y_plain)).item()
    assert rel <= 1e-14
    ws = torch.empty(hs.nh // 2, dtype=dt, device="cuda")     # too small: refused
    with pytest.raises(_lib.QesError, match="workspace"):
        _lib.check(_lib.lib().qes_matvec_ws(H._device_op(), _lib.dptr(x), _lib.dptr(y_plain), 1, 1 if cplx else 0, 0,
                                            _lib.dptr(ws), ws.numel() * ws.element_size(), _lib.stream_ptr()),
                   "qes_matvec_ws")


@pytest.mark.parametrize("case", [("chain", 24, 1, {"translation": 0, "parity": {"x": 1}}, 12),
                                  ("chain", 34, 1, {"translation": 0, "parity": {"x": 1}}, 17),
                                  ("square", 4, 4, {"translation": {"kx": 0, "ky": 0}, "parity": {"x": 1}}, 8),
                                  SPACES[0][:5], SPACES[6][:5], SPACES[2][:5]],
                         ids=lambda c: f"{c[0]}{c[1]}x{c[2]}-{c[4]}")
def test_lean_kernels_equal_the_general_ones(case):
    """Real pre-scaled products of groups whose base elements are the
    identity and one spin flip run kernels compiled for that case (option
    lean): bitwise the products of the general kernels, with the dense and
    the rank gathers, and of the one-lookup path; a group with a reflection
    (not lean) is unaffected."""
    import torch
    from pyqusolver_amd.models import J1J2Model
    kind, lx, ly, sym, n_up = case
    lat, hs = _space(kind, lx, ly, sym, n_up)
    H = J1J2Model(lat, J1=-1.0, J2=-0.45, hilbert_space=hs)
    assert not H.needs_complex
    x = torch.randn(hs.nh, dtype=torch.float64, device="cuda", generator=torch.Generator(device="cuda").manual_seed(5))
    H.set_option("lean", 0)
    y0 = H.matvec_device(x)
    H.set_option("lean", 1)
    assert torch.equal(H.matvec_device(x), y0)
    H.set_option("direct_batch", 1)
    assert torch.equal(H.matvec_device(x), y0)
    H.set_option("direct_batch", -1)
    if hs.nh > 10**6:
        H.set_option("rank_gather", 1)
        assert torch.equal(H.matvec_device(x), y0)
        H.set_option("lean", 0)
        assert torch.equal(H.matvec_device(x), y0)
        H.set_option("rank_gather", 0)
    H.set_option("lean", 1)
    # the table built by the lean passes holds the same elements in the same order
    if hs.nh < 10**8:
        H.set_option("lean", 0)
        H.assemble(True)
        ya = H.matvec_device(x)
        H.assemble(False)
        H.set_option("lean", 1)
        H.assemble(True)
        assert torch.equal(H.matvec_device(x), ya)
        H.assemble(False)


@pytest.mark.parametrize("case", [("chain", 20, 1, {"translation": 3, "parity": {"x": 1}}, 10),
                                  ("square", 6, 6, {"translation": {"kx": 1, "ky": 2}, "parity": {"x": 1}}, 18)],
                         ids=lambda c: f"{c[0]}{c[1]}x{c[2]}")
def test_lean_kernels_complex(case):
    """Complex momentum sectors with a spin flip run the lean kernels too:
    the same products as the general kernels (the same sums in the same
    order; complex multiply-adds may contract differently in the two
    compilations, so to 1e-15 relative), with the dense and the rank
    gathers."""
    import torch
    from pyqusolver_amd.models import J1J2Model
    kind, lx, ly, sym, n_up = case
    lat, hs = _space(kind, lx, ly, sym, n_up)
    H = J1J2Model(lat, J1=-1.0, J2=-0.45, hilbert_space=hs)
    assert H.needs_complex
    g = torch.Generator(device="cuda").manual_seed(6)
    x = torch.complex(torch.randn(hs.nh, dtype=torch.float64, device="cuda", generator=g),
                      torch.randn(hs.nh, dtype=torch.float64, device="cuda", generator=g))

    def close(a, b):
        return (torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)).item() <= 1e-15
    H.set_option("lean", 0)
    y0 = H.matvec_device(x)
    H.set_option("lean", 1)
    assert close(H.matvec_device(x), y0)
    if kind == "chain":
        H.set_option("rank_gather", 1)
        y1 = H.matvec_device(x)
        H.set_option("lean", 0)
        assert close(H.matvec_device(x), y1) and close(y1, y0)
        H.set_option("rank_gather", 0)
        H.set_option("lean", 1)
        H.set_option("lean", 0)
        H.assemble(True)
        ya = H.matvec_device(x)
        H.assemble(False)
        H.set_option("lean", 1)
        H.assemble(True)
        assert torch.equal(H.matvec_device(x), ya)
        H.assemble(False)
