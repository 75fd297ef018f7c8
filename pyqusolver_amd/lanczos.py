"""Ground-state Lanczos on the device.

The recurrence, dots, axpys and full reorthogonalisation run in libqes_mi
(one hipGraph-captured step replayed per iteration); the host only sees
alpha/beta and solves the FP64 tridiagonal problem with
scipy.linalg.eigh_tridiagonal.  This replaces the eigensolver the
reference dispatches from DiagonalizationEngine
(Algebra/Hamil/hamil_diag_engine.py:176-300, called at Algebra/hamil.py:2063-2081).
"""
from __future__ import annotations

import ctypes as C
import math
import warnings
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch
from scipy.linalg import eigh_tridiagonal

from . import _lib


@dataclass
class LanczosResult:
    eigenvalues: np.ndarray
    alpha: np.ndarray
    beta: np.ndarray
    steps: int
    converged: bool
    history: list = field(default_factory=list)     # (steps, E0) at each check
    eigenvectors: Optional[torch.Tensor] = None
    seconds: float = 0.0
    residual_bound: float = float("nan")
    assembled: bool = False                          # the solve streamed an assembled table
    second_passes: int = 0                           # Gram-Schmidt second passes (DGKS) run


def start_vector(n: int, complex_: bool, seed: int = 0) -> torch.Tensor:
    """Seeded N(0,1) start vector generated on the device (complex: re + 1j*im)."""
    g = torch.Generator(device="cuda")
    g.manual_seed(int(seed))
    re = torch.randn(n, generator=g, dtype=torch.float64, device="cuda")
    if complex_:
        return torch.complex(re, torch.randn(n, generator=g, dtype=torch.float64, device="cuda"))
    return re


class DeviceLanczos:
    """Thin owner of a qes_lanczos handle."""

    def __init__(self, ham, max_steps: int, store_basis: bool, complex_: bool):
        self.ham = ham
        self.max_steps = int(max_steps)
        self.dtype = 1 if complex_ else 0
        h = C.c_void_p()
        _lib.check(_lib.lib().qes_lanczos_create(ham._device_op(), self.max_steps, 1 if store_basis else 0,
                                                 self.dtype, C.byref(h)), "qes_lanczos_create")
        self._h = h

    def start(self, v0: torch.Tensor):
        _lib.check(_lib.lib().qes_lanczos_start(self._h, _lib.dptr(v0), _lib.stream_ptr()), "qes_lanczos_start")

    def run(self, nsteps: int, reorth, use_graph: bool = True):
        """reorth: False/0 none, True/1 Gram-Schmidt with the DGKS second pass,
        2 or 'twice': always two passes."""
        mode = 2 if reorth in (2, "twice") else (1 if reorth else 0)
        _lib.check(_lib.lib().qes_lanczos_run(self._h, int(nsteps), mode, 1 if use_graph else 0),
                   "qes_lanczos_run")

    @property
    def second_passes(self) -> int:
        c = C.c_int32()
        _lib.check(_lib.lib().qes_lanczos_stats(self._h, C.byref(c)), "qes_lanczos_stats")
        return int(c.value)

    def coefficients(self):
        a = np.zeros(self.max_steps)
        b = np.zeros(self.max_steps)
        k = C.c_int32()
        _lib.check(_lib.lib().qes_lanczos_get(self._h, a.ctypes.data_as(_lib._PD), b.ctypes.data_as(_lib._PD),
                                              C.byref(k)), "qes_lanczos_get")
        return a[:k.value].copy(), b[:k.value].copy()

    def ritz(self, coeffs: np.ndarray, out: torch.Tensor):
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        _lib.check(_lib.lib().qes_lanczos_ritz(self._h, c.ctypes.data_as(_lib._PD), len(c), _lib.dptr(out),
                                               _lib.stream_ptr()), "qes_lanczos_ritz")

    def accumulate(self, coeffs: np.ndarray, out: torch.Tensor):
        """out[i] += coeffs[i] * (current Lanczos vector) for the rows of out."""
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        _lib.check(_lib.lib().qes_lanczos_accumulate(self._h, c.ctypes.data_as(_lib._PD), len(c), _lib.dptr(out),
                                                     _lib.stream_ptr()), "qes_lanczos_accumulate")

    def close(self):
        if self._h is not None and self._h.value:
            _lib.lib().qes_lanczos_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tridiag_eigs(alpha, beta, k=1):
    m = len(alpha)
    if m == 1:
        return np.array([alpha[0]]), np.ones((1, 1))
    w, v = eigh_tridiagonal(alpha, beta[: m - 1])
    return w, v


def effective_steps(alpha, beta) -> int:
    """Rows of (alpha, beta) that span the Krylov space: i + 1 for the first
    |beta_i| <= 1e-14 max(1, |theta|) (an invariant subspace: later rows
    belong to a zero vector and add a spurious eigenvalue 0), else all."""
    m = len(alpha)
    if m <= 1:
        return m
    scale = max(1.0, abs(float(tridiag_eigs(alpha, beta)[0][0])))
    small = np.nonzero(~(np.abs(np.asarray(beta[:m])) > 1e-14 * scale))[0]
    return int(small[0]) + 1 if small.size else m


def ritz_residual(alpha, beta):
    """(r, quad) for the lowest Ritz value after m = len(alpha) steps
    (beta[m-1] = beta_m, the coefficient of the next basis vector):
    r = |beta_m s_m| the residual norm ||H y - theta y|| of the Ritz pair
    (s: unit eigenvector of T_m; |theta - lambda| <= r for some eigenvalue
    lambda of H) -- the stopping test, as in eigsh/ARPACK; quad =
    min(r, r^2 / gap), gap = second Ritz value - theta, the usual estimate of
    |theta - E0| once theta is separated from the next eigenvalue (an
    unresolved near-degenerate pair makes it too small)."""
    m = len(alpha)
    if m == 0:
        return math.inf, math.inf
    wv, V = tridiag_eigs(alpha, beta)
    r = abs(float(beta[m - 1]) * float(V[-1, 0]))
    if m < 2:
        return r, r
    gap = float(wv[1] - wv[0])
    return r, (min(r, r * r / gap) if gap > 0 else r)


_WHICH = {"smallest": "SA", "sa": "SA", "largest": "LA", "la": "LA", "both": "BE", "be": "BE"}


def which_code(which: str) -> str:
    """Normalised spectrum end: 'SA' (smallest algebraic, the default), 'LA'
    (largest) or 'BE' (both ends) -- the reference's which='smallest' /
    'largest' / 'both' (hamil_diag_engine.py:183) and eigsh's codes."""
    code = _WHICH.get(str(which).lower())
    if code is None:
        raise ValueError(f"which={which!r}: the Lanczos path computes 'smallest', 'largest' or 'both'")
    return code


def select_ritz(m: int, k: int, which: str = "SA") -> np.ndarray:
    """Indices (ascending) of the k wanted Ritz values among m sorted ones:
    the k lowest (SA), the k highest (LA), or k//2 lowest and the rest
    highest (BE, one more from the top when k is odd, as eigsh does)."""
    k = max(0, min(int(k), m))
    if which == "SA":
        return np.arange(k)
    if which == "LA":
        return np.arange(m - k, m)
    lo = k // 2
    return np.concatenate([np.arange(lo), np.arange(m - (k - lo), m)])


def _pair_residuals(alpha, beta, idx):
    """Ritz values and residual norms |beta_m S[m-1, i]| of the pairs idx."""
    w, S = tridiag_eigs(alpha, beta)
    m = len(alpha)
    return w[idx], np.abs(float(beta[m - 1]) * S[-1, idx])


def next_check(m0: int, r0: float, m1: int, r1: float, target: float, lo: int, hi: int) -> int:
    """Steps to run before the next stopping test: where the geometric decay
    of the residual between the tests after m0 < m1 steps (r0, r1) reaches
    target, clamped to [lo, hi]; hi while r is not decreasing.  Once the
    lowest Ritz pair is isolated its residual falls about geometrically, so
    the test lands near the step where it first passes instead of up to
    hi - 1 steps after it."""
    lo = max(1, int(lo))
    hi = max(lo, int(hi))
    if not (m1 > m0 and r0 > 0.0 and r1 > 0.0 and r1 < r0 and target > 0.0 and r1 > target):
        return hi
    rate = (math.log(r1) - math.log(r0)) / (m1 - m0)
    need = math.ceil((math.log(target) - math.log(r1)) / rate)
    return int(min(hi, max(lo, need)))


def lanczos_ground_state(ham, k: int = 1, tol: float = 1e-10, maxiter: int = 300, v0=None, seed: int = 0,
                         reorth: bool = True, store_basis: Optional[bool] = None, compute_eigvec: bool = False,
                         check_every: int = 10, use_graph: bool = True,
                         adaptive: Optional[bool] = None, which: str = "smallest") -> LanczosResult:
    """Extreme eigenvalue(s) by device Lanczos, the lowest by default.  Stops
    when the Ritz residual |beta_m s_m| of the extreme pair is <=
    tol*max(1,|theta|); with k > 1 and the basis kept (reorthogonalised), all
    k wanted pairs must pass.  ``which``: 'smallest' (k lowest), 'largest'
    (k highest) or 'both' (k//2 lowest, the rest highest); eigenvalues come
    back in ascending order.  A single start vector finds one copy of a
    degenerate eigenvalue; without reorthogonalisation only the extreme
    value(s) at the wanted end(s) are reliable.  Tests are at
    most ``check_every`` steps apart; the next one goes where the residual's
    decay between the last two predicts the crossing (``next_check``;
    ``adaptive=False``: every ``check_every`` steps; None: adaptive for
    N >= 2^20, where a step outlasts a test's host round trip).  With
    ``compute_eigvec`` the Ritz vectors are formed on the device: shape (n,)
    for k = 1, (n, k) otherwise; without a stored basis (store_basis=False)
    by a second run of the recurrence from the same start vector."""
    n = ham.nh
    complex_ = bool(ham.needs_complex)
    code = which_code(which)
    if v0 is None:
        v0 = start_vector(n, complex_, seed)
    if isinstance(v0, torch.Tensor):
        v0t = v0.to("cuda")
    else:
        v0t = torch.from_numpy(np.ascontiguousarray(v0)).to("cuda")
    v0t = v0t.to(torch.complex128 if complex_ else torch.float64).contiguous()
    maxiter = int(min(maxiter, max(n, 1)))
    if adaptive is None:
        adaptive = n >= (1 << 20)
    if store_basis is None:
        store_basis = bool(reorth or compute_eigvec)
    reorth = (2 if reorth in (2, "twice") else bool(reorth)) if store_basis else False
    if k > (2 if code == "BE" else 1) and not reorth:
        warnings.warn("without reorthogonalisation only the extreme Ritz value(s) are converged; "
                      "the other values returned are not", stacklevel=2)
    t0 = time.perf_counter()
    L = DeviceLanczos(ham, maxiter, store_basis, complex_)
    L.start(v0t)
    hist = []
    converged = False
    a = b = np.zeros(0)
    interval, m_prev, r_prev = check_every, 0, 0.0
    while True:
        a_done = len(a)
        nsteps = min(interval, maxiter - a_done)
        if nsteps <= 0:
            break
        L.run(nsteps, reorth, use_graph)
        a, b = L.coefficients()
        w, _ = tridiag_eigs(a, b)
        e0 = float(w[0])
        hist.append((len(a), e0))
        scale = max(1.0, abs(e0))
        if np.any(np.abs(b) <= 1e-14 * scale):          # invariant subspace found
            cut = effective_steps(a, b)
            a, b = a[:cut], b[:cut]
            w, _ = tridiag_eigs(a, b)
            converged = True
            break
        if code == "SA" and not (k > 1 and reorth):
            r_now, target = ritz_residual(a, b)[0], tol * scale
            if r_now <= target:
                converged = True
                break
        else:
            m = len(a)
            if k > 1 and reorth:
                if m < k:
                    continue
                idx = select_ritz(m, k, code)
            else:   # no stored basis: the extreme pair(s) of the wanted end(s)
                idx = select_ritz(m, 1 if code != "BE" else 2, code)
            wk, rk = _pair_residuals(a, b, idx)
            tgt = tol * np.maximum(1.0, np.abs(wk))
            if np.all(rk <= tgt):
                converged = True
                break
            with np.errstate(divide="ignore", invalid="ignore"):   # tol = 0: no pair ever passes
                worst = int(np.argmax(np.where(tgt > 0, rk / tgt, np.inf)))
            r_now, target = float(rk[worst]), float(tgt[worst])
        if adaptive:
            interval = next_check(m_prev, r_prev, len(a), r_now, target, 1, check_every)
        m_prev, r_prev = len(a), r_now
    torch.cuda.synchronize()
    w, V = tridiag_eigs(a, b)
    sel = select_ritz(len(a), k, code)     # wanted pairs, ascending
    kk = len(sel)
    res_bound = float(np.max(np.abs(b[len(a) - 1] * V[-1, sel]))) if kk else float("nan")
    vecs = None
    if compute_eigvec and not store_basis:
        # second pass from the same start vector: the steps are bitwise
        # reproducible, so v_j is regenerated and sum_j S[j, i] v_j formed
        # without keeping the basis (memory: four vectors + the k outputs)
        m = len(a)
        if kk > 16:
            raise ValueError("Ritz vectors without a stored basis: at most 16 (keep the basis for more)")
        second = L.second_passes if reorth else 0
        L.close()
        L = DeviceLanczos(ham, max(m, 1), False, complex_)
        L.start(v0t)
        cols = torch.zeros((kk, n), dtype=v0t.dtype, device="cuda")
        L.accumulate(V[0, sel], cols)
        for j in range(1, m):
            L.run(1, False, use_graph)
            L.accumulate(V[j, sel], cols)
        vecs = cols[0].clone() if k == 1 else cols.transpose(0, 1).contiguous()
        L.close()
        return LanczosResult(eigenvalues=w[sel].copy(), alpha=a, beta=b, steps=len(a), converged=converged,
                             history=hist, eigenvectors=vecs, seconds=time.perf_counter() - t0,
                             residual_bound=res_bound, second_passes=second)
    if compute_eigvec:
        if k == 1:
            vecs = torch.empty((n,), dtype=v0t.dtype, device="cuda")
            L.ritz(V[:, sel[0]], vecs)
        else:
            cols = torch.empty((kk, n), dtype=v0t.dtype, device="cuda")
            for i, s in enumerate(sel):
                L.ritz(V[:, s], cols[i])
            vecs = cols.transpose(0, 1).contiguous()
    second = L.second_passes if reorth else 0
    L.close()
    return LanczosResult(eigenvalues=w[sel].copy(), alpha=a, beta=b, steps=len(a), converged=converged,
                         history=hist, eigenvectors=vecs, seconds=time.perf_counter() - t0,
                         residual_bound=res_bound, second_passes=second)


# ---------------------------------------------------------------------------
# Block Lanczos (Hamiltonian.diagonalize(method='block_lanczos'),
# Algebra/Hamil/hamil_diag_engine.py:59, block_size :219)
# ---------------------------------------------------------------------------

class _BlockStore:
    """Stored Krylov blocks as consecutive vectors of doubles (complex:
    interleaved), with the FP64-MFMA contractions of libqes_mi."""

    def __init__(self, n: int, cplx: bool, nvec: int, p: int):
        self.n, self.cplx, self.p = n, cplx, p
        self.len = 2 * n if cplx else n
        self.ld = self.len + (self.len & 1)
        self.V = torch.zeros((nvec, self.ld), dtype=torch.float64, device="cuda")
        self.W = torch.zeros((p, self.ld), dtype=torch.float64, device="cuda")
        nb = C.c_int64()
        _lib.check(_lib.lib().qes_block_dot_ws_bytes(nvec, self.len, C.byref(nb)), "qes_block_dot_ws_bytes")
        self.ws = torch.empty(max(1, int(nb.value) // 8), dtype=torch.float64, device="cuda")

    def cview(self, t: torch.Tensor) -> torch.Tensor:
        """(rows, n) complex or real view of stored rows."""
        x = t[:, :self.len]
        return x.view(torch.complex128) if self.cplx else x

    def dot(self, V: torch.Tensor, W: torch.Tensor) -> np.ndarray:
        """V^H W for row blocks V (nv rows) and W (ncol rows)."""
        nv, ncol = V.shape[0], W.shape[0]
        C_ = torch.empty(nv * ncol * (2 if self.cplx else 1), dtype=torch.float64, device="cuda")
        _lib.check(_lib.lib().qes_block_dot(_lib.dptr(V), self.ld, nv, _lib.dptr(W), self.ld, ncol, self.len,
                                            1 if self.cplx else 0, _lib.dptr(C_), _lib.dptr(self.ws),
                                            self.ws.numel() * 8, _lib.stream_ptr()), "qes_block_dot")
        c = C_.cpu().numpy()
        return (c[0::2] + 1j * c[1::2]).reshape(nv, ncol) if self.cplx else c.reshape(nv, ncol)

    def combine(self, V: torch.Tensor, coef: np.ndarray, beta: float, Y: torch.Tensor):
        """Y_q = beta Y_q + sum_i V_i coef[i, q]."""
        nv, ncol = coef.shape
        if self.cplx:
            c = np.empty((nv, ncol, 2))
            c[..., 0], c[..., 1] = coef.real, coef.imag
        else:
            c = np.ascontiguousarray(coef.real)
        ct = torch.from_numpy(np.ascontiguousarray(c).reshape(-1)).to("cuda")
        _lib.check(_lib.lib().qes_block_combine(_lib.dptr(V), self.ld, nv, _lib.dptr(ct), float(beta), _lib.dptr(Y),
                                                self.ld, ncol, self.len, 1 if self.cplx else 0, _lib.stream_ptr()),
                   "qes_block_combine")


def _chol_qr(st: _BlockStore, W: torch.Tensor, out: torch.Tensor):
    """out B = W with out orthonormal: shifted Cholesky QR, then Cholesky QR
    twice (stable for condition numbers up to ~1/eps, Fukaya et al. 2020);
    returns B (p x p upper triangular) or None when W vanishes (an
    invariant subspace).  Near-dependent directions of W come out as
    orthonormal vectors with tiny couplings in B, which the full
    reorthogonalisation keeps orthogonal to the basis."""
    p = W.shape[0]
    B = np.eye(p, dtype=np.complex128 if st.cplx else np.float64)
    src = W
    for it in range(3):
        G = st.dot(src, src)
        G = 0.5 * (G + G.conj().T)
        gmax = float(np.max(np.abs(np.diag(G).real)))
        if it == 0 and not gmax > 1e-300:
            return None
        shift = 11.0 * (st.len * p + p * (p + 1)) * np.finfo(np.float64).eps * gmax if it == 0 else 0.0
        try:
            R = np.linalg.cholesky(G + shift * np.eye(p)).conj().T           # G = R^H R
        except np.linalg.LinAlgError:
            return None
        st.combine(src, np.linalg.inv(R), 0.0, out)
        B = R @ B
        src = out
    return B


def block_lanczos(ham, k: int = 6, block_size: Optional[int] = None, tol: float = 1e-10, maxiter: int = 100,
                  v0=None, seed: int = 0, which: str = "smallest", compute_eigvec: bool = False) -> LanczosResult:
    """k extreme eigenpairs by block Lanczos with full reorthogonalisation.

    Each step applies H to a block of p vectors (one product of the
    operator's multi-column path: an assembled table streams once per four
    columns), projects out the previous two blocks and then the whole stored
    basis twice (classical Gram-Schmidt with a second pass), with the
    projections C = V^H W and the updates W -= V C on FP64 matrix cores, and
    orthonormalises the block by Cholesky QR (twice).  The block tridiagonal
    T is solved on the host in FP64; a Ritz pair (theta_i, s_i) is converged
    when ||B_m s_i[last block]|| <= tol max(1, |theta_i|), the residual norm
    of its Ritz vector.  ``maxiter``: block steps; the stored basis holds
    (maxiter + 1) p vectors.  ``block_size`` defaults to max(4, k) (at most 16
    for real, 8 for complex vectors).  Several vectors per step find
    degenerate levels that a single start vector cannot."""
    code = which_code(which)
    n = ham.nh
    cplx = bool(ham.needs_complex)
    pmax = 8 if cplx else 16
    p = int(block_size) if block_size else max(4, int(k))
    if not 1 <= p <= pmax:
        raise ValueError(f"block_size must be 1..{pmax} for {'complex' if cplx else 'real'} vectors")
    k = int(k)
    if k < 1 or k > n:
        raise ValueError(f"k must be in 1..{n}")
    p = min(p, n)
    m_max = max(1, min(int(maxiter), (n + p - 1) // p))
    t0 = time.perf_counter()
    st = _BlockStore(n, cplx, (m_max + 1) * p, p)
    dt = torch.complex128 if cplx else torch.float64
    if v0 is None:
        g = torch.Generator(device="cuda").manual_seed(int(seed))
        X0 = torch.randn((p, n), dtype=dt, device="cuda", generator=g)
    else:
        X0 = torch.as_tensor(v0).to(device="cuda", dtype=dt).reshape(-1, n)[:p].contiguous()
        if X0.shape[0] < p:
            g = torch.Generator(device="cuda").manual_seed(int(seed))
            X0 = torch.cat([X0, torch.randn((p - X0.shape[0], n), dtype=dt, device="cuda", generator=g)])
    st.cview(st.W).copy_(X0)
    if _chol_qr(st, st.W, st.V[0:p]) is None:
        raise _lib.QesError("block Lanczos: the start block is rank deficient")
    As, Bs = [], []
    hist = []
    converged = False
    theta = S = None
    m = 0
    for j in range(m_max):
        Xj = st.V[j * p:(j + 1) * p]
        # W = H X_j through the multi-column product ((n, p) row-major)
        Y = ham.matvec_device(st.cview(Xj).transpose(0, 1).contiguous())
        st.cview(st.W).copy_(Y.transpose(0, 1))
        A = st.dot(Xj, st.W)
        lo = max(0, j - 1) * p
        coef = -A if j == 0 else -np.vstack([Bs[-1].conj().T, A])
        st.combine(st.V[lo:(j + 1) * p], coef, 1.0, st.W)
        for _ in range(2):                           # full reorthogonalisation, twice
            C1 = st.dot(st.V[:(j + 1) * p], st.W)
            st.combine(st.V[:(j + 1) * p], -C1, 1.0, st.W)
            A = A + C1[j * p:(j + 1) * p]
        A = 0.5 * (A + A.conj().T)
        As.append(A)
        m = j + 1
        B = _chol_qr(st, st.W, st.V[(j + 1) * p:(j + 2) * p]) if m < m_max else None
        # block tridiagonal T_m and its Ritz pairs
        T = np.zeros((m * p, m * p), dtype=np.complex128 if cplx else np.float64)
        for i in range(m):
            T[i * p:(i + 1) * p, i * p:(i + 1) * p] = As[i]
            if i + 1 < m:
                T[(i + 1) * p:(i + 2) * p, i * p:(i + 1) * p] = Bs[i]
                T[i * p:(i + 1) * p, (i + 1) * p:(i + 2) * p] = Bs[i].conj().T
        theta, S = np.linalg.eigh(T)
        idx = select_ritz(m * p, min(k, m * p), code)
        if B is None:        # invariant subspace (or step budget): the Ritz pairs are final
            res = np.zeros(len(idx)) if m < m_max else np.full(len(idx), np.inf)
            hist.append((m * p, float(theta[idx[0]])))
            converged = m < m_max and m * p >= k
            break
        res = np.linalg.norm(B @ S[(m - 1) * p:m * p, idx], axis=0)
        Bs.append(B)
        hist.append((m * p, float(theta[idx[0]])))
        if m * p >= k and np.all(res <= tol * np.maximum(1.0, np.abs(theta[idx]))):
            converged = True
            break
    idx = select_ritz(m * p, min(k, m * p), code)
    vals = theta[idx]
    vecs = None
    if compute_eigvec:
        Sk = S[:, idx]
        out = torch.zeros((len(idx), st.ld), dtype=torch.float64, device="cuda")
        step = pmax
        for c0 in range(0, len(idx), step):
            st.combine(st.V[:m * p], Sk[:, c0:c0 + step], 0.0, out[c0:c0 + step])
        vecs = st.cview(out).transpose(0, 1).contiguous()
        if k == 1:
            vecs = vecs[:, 0]
    r = LanczosResult(eigenvalues=np.asarray(vals, dtype=np.float64), alpha=np.zeros(0), beta=np.zeros(0),
                      steps=int(m * p), converged=bool(converged), history=hist, eigenvectors=vecs,
                      seconds=time.perf_counter() - t0,
                      residual_bound=float(np.max(res)) if len(idx) else 0.0)
    r.block_size = p
    r.block_steps = m
    return r
