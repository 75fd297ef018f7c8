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
