"""Block Lanczos (Hamiltonian.diagonalize(method='block_lanczos'),
hamil_diag_engine.py:59): the FP64-MFMA contractions against torch FP64
matrix products, and the k extreme eigenvalues against dense
eigensolves of the oracle's reduced matrix (GPU)."""
import numpy as np
import pytest

from oracle import qes_oracle as O

pytestmark = pytest.mark.gpu


def _blocks(nv, ncol, length, cplx, seed):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    dlen = 2 * length if cplx else length
    ld = dlen + (dlen & 1)
    V = torch.randn((nv, ld), dtype=torch.float64, device="cuda", generator=g)
    W = torch.randn((ncol, ld), dtype=torch.float64, device="cuda", generator=g)
    return V, W, dlen, ld


def _as_c(x, dlen, cplx):
    x = x[:, :dlen]
    import torch
    return x.contiguous().view(torch.complex128) if cplx else x


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("nv,ncol,length", [(1, 1, 1), (5, 4, 15), (16, 16, 16), (37, 3, 1000), (70, 8, 12345),
                                            (130, 4, 4099)])
def test_block_contractions_match_torch(cplx, nv, ncol, length):
    import torch
    from pyqusolver_amd import _lib
    if cplx and ncol > 8:
        ncol = 8
    V, W, dlen, ld = _blocks(nv, ncol, length, cplx, nv * 131 + ncol)
    nb = __import__("ctypes").c_int64()
    _lib.check(_lib.lib().qes_block_dot_ws_bytes(nv, dlen, __import__("ctypes").byref(nb)))
    ws = torch.empty(max(1, nb.value // 8), dtype=torch.float64, device="cuda")
    C = torch.empty(nv * ncol * (2 if cplx else 1), dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().qes_block_dot(_lib.dptr(V), ld, nv, _lib.dptr(W), ld, ncol, dlen, 1 if cplx else 0,
                                        _lib.dptr(C), _lib.dptr(ws), nb.value, _lib.stream_ptr()), "qes_block_dot")
    Vc, Wc = _as_c(V, dlen, cplx), _as_c(W, dlen, cplx)
    ref = Vc.conj() @ Wc.transpose(0, 1)
    got = torch.view_as_complex(C.reshape(nv, ncol, 2)) if cplx else C.reshape(nv, ncol)
    tol = 1e-13 * (Vc.abs().max() * Wc.abs().max() * length).item()
    assert (got - ref).abs().max().item() <= tol
    # Y = beta Y + V coef (coef nv x ncol)
    coef = torch.randn((nv, ncol, 2) if cplx else (nv, ncol), dtype=torch.float64, device="cuda")
    Y = torch.randn((ncol, ld), dtype=torch.float64, device="cuda")
    Yc0 = _as_c(Y, dlen, cplx).clone()
    cm = torch.view_as_complex(coef) if cplx else coef
    for beta in (0.0, 1.0, -0.5):
        Yw = Y.clone()
        _lib.check(_lib.lib().qes_block_combine(_lib.dptr(V), ld, nv, _lib.dptr(coef.reshape(-1)), beta, _lib.dptr(Yw),
                                                ld, ncol, dlen, 1 if cplx else 0, _lib.stream_ptr()),
                   "qes_block_combine")
        refy = (beta * Yc0 if beta != 0.0 else 0.0) + cm.transpose(0, 1) @ Vc
        goty = _as_c(Yw, dlen, cplx)
        assert (goty - refy).abs().max().item() <= 1e-13 * max(1.0, refy.abs().max().item()) * nv


CASES = [
    ("chain", 14, 1, [("tx", 0)], 7, "j1j2", 6, 4),
    ("chain", 14, 1, [("tx", 3)], 7, "xxz", 5, 4),
    ("square", 4, 4, [("tx", 0), ("ty", 0)], 8, "j1j2", 8, 8),
    ("chain", 18, 1, [("tx", 0), ("Px", 1)], 9, "j1j2", 4, 6),
    ("honeycomb", 2, 3, [("tx", 0), ("ty", 0)], None, "hk", 6, 6),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}{c[1]}x{c[2]}-{c[5]}-k{c[6]}-p{c[7]}")
def test_block_lanczos_lowest_k_against_dense_oracle(case):
    import torch
    from test_gpu_parity import build_pair
    kind, lx, ly, gens, n_up, model, k, p = case
    hs, H, grp, b, t = build_pair(kind, lx, ly, gens, n_up, model)
    c, s, co = t.arrays()
    ref = np.linalg.eigvalsh(O.reduced_dense(b, t.ns, c, s, co, t.is_complex))
    e = H.diagonalize(method="block_lanczos", k=k, block_size=p, tol=1e-12, maxiter=200, compute_eigvec=True)
    r = H.last_lanczos
    assert r.converged and r.block_size == p
    np.testing.assert_allclose(e, ref[:k], rtol=0, atol=1e-10 * max(1.0, abs(ref[0])))
    V = H._eig_vec
    for i in range(k):
        v = V[:, i].contiguous()
        res = torch.linalg.vector_norm(H.matvec_device(v) - float(e[i]) * v) / torch.linalg.vector_norm(v)
        assert res.item() <= 1e-9
    # the highest end too
    e_hi = H.diagonalize(method="block_lanczos", k=3, block_size=4, tol=1e-12, maxiter=200, which="largest")
    np.testing.assert_allclose(e_hi, ref[-3:], rtol=0, atol=1e-10 * max(1.0, abs(ref[-1])))


def test_block_lanczos_resolves_degenerate_levels():
    """The 4x4 Heisenberg square at k = (0, 0) without spin flip holds
    degenerate multiplets in its low spectrum; a block of 8 start vectors
    finds every copy (one start vector finds one)."""
    from test_gpu_parity import build_pair
    hs, H, grp, b, t = build_pair("square", 4, 4, [("tx", 0), ("ty", 0)], 8, "heis")
    c, s, co = t.arrays()
    ref = np.linalg.eigvalsh(O.reduced_dense(b, t.ns, c, s, co, t.is_complex))
    k = 10
    gaps = np.diff(ref[:k])
    assert np.any(gaps < 1e-9)            # the test needs a degeneracy in range
    e = H.diagonalize(method="block_lanczos", k=k, block_size=8, tol=1e-12, maxiter=200)
    np.testing.assert_allclose(e, ref[:k], rtol=0, atol=1e-10 * abs(ref[0]))
