"""Differentiable matvec y = H(J) v and the gradients built on it.

The reference has no differentiable Hamiltonian; this module is the
MI355X build's addition on top of the same operator.  H is parametrised by
its coupling ids (the ``coupling`` tag every instruction carries, see
models.py): with the operator compiled at couplings J0,

    H(J) = sum_c (J_c / J0_c) H_c ,   H_c = instructions tagged c at J0,

so dH/dJ_c = H_c / J0_c.  Forward and backward are

    forward   y       = sum_c s_c H_c v                  (s = J / J0)
    backward  grad_v  = H(J)^dagger grad_y = H(J) grad_y (H Hermitian)
              grad_s_c = Re <grad_y, H_c v>

with torch's convention grad = dL/dRe + i dL/dIm for complex tensors.  Two
device paths compute them:

  * a coupling-tagged assembled table (``Hamiltonian.assemble(couplings=
    True)``: every tabulated value keyed by its coupling id, one diagonal
    per coupling): y is one pass with a scaled value dictionary
    (``qes_op_matvec_scaled``) and all Re <grad_y, H_c v> are one more pass
    (``qes_op_coupling_dots``);
  * matrix-free: ``qes_matvec_couplings`` returns every H_c v in a single
    pass over the rows (one canonicalisation per neighbour for all
    couplings) and the sums are taken with torch.

On top of it:
  * ``lanczos_energy``: E0(J) from an m-step Lanczos recurrence written with
    the differentiable matvec (plain three-term recurrence, segments
    recomputed in backward via torch.utils.checkpoint, so the Krylov basis is
    not kept), differentiated through torch.linalg.eigvalsh of T_m;
  * ``hellmann_feynman``: dE0/dJ_c = <psi0|H_c|psi0> / J0_c from the device
    Lanczos Ritz vector;
  * ``krylov_overlap``: |<psi0| exp(-i H t) |phi>|^2 with the m-step Krylov
    propagator V_m exp(-i T_m t) e_1 |phi|, differentiable in J with the same
    segment recomputation.
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint


def _couplings_of(ham) -> np.ndarray:
    """J0 per coupling id, from the model's instruction coefficients: the
    reference value every H_c is compiled at.  Models record them; a plain
    Hamiltonian is taken at unit scale."""
    j0 = getattr(ham, "coupling_values", None)
    n = ham.n_couplings
    if j0 is None:
        return np.ones(n)
    j0 = np.asarray(j0, dtype=np.float64)
    if j0.shape != (n,):
        raise ValueError(f"coupling_values has shape {j0.shape}, operator has {n} coupling ids")
    if np.any(j0 == 0.0):
        raise ValueError("a coupling compiled at 0 has no instructions and cannot be differentiated")
    return j0


class HamiltonianMatvec(torch.autograd.Function):
    """y = sum_c s_c H_c v on the device (s real, shape (n_couplings,))."""

    @staticmethod
    def forward(ctx, s: torch.Tensor, v: torch.Tensor, ham):
        if v.device.type != "cuda":
            raise ValueError("HamiltonianMatvec works on device tensors")
        ctx.ham = ham
        ctx.tagged = ham.assembly["couplings"]
        ctx.save_for_backward(s, v)
        if ctx.tagged:
            return ham.matvec_scaled_device(s.detach().double().cpu().numpy(), v)
        Y = ham.matvec_couplings_device(v.contiguous())          # (n_c, N)
        sc = s.detach().to(device=v.device, dtype=torch.float64)
        return torch.tensordot(sc.to(Y.dtype), Y, dims=([0], [0]))

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        s, v = ctx.saved_tensors
        ham = ctx.ham
        grad_s = grad_v = None
        g = g.contiguous()
        if ctx.tagged:
            if ctx.needs_input_grad[1]:
                grad_v = ham.matvec_scaled_device(s.detach().double().cpu().numpy(), g)
            if ctx.needs_input_grad[0]:
                d = ham.coupling_dots_device(v.contiguous(), g)
                grad_s = torch.as_tensor(d.real, dtype=s.dtype, device=s.device)
            return grad_s, grad_v, None
        if ctx.needs_input_grad[1]:
            Yg = ham.matvec_couplings_device(g)
            sc = s.detach().to(device=g.device, dtype=torch.float64)
            grad_v = torch.tensordot(sc.to(Yg.dtype), Yg, dims=([0], [0]))
        if ctx.needs_input_grad[0]:
            # recompute H_c v rather than keeping n_c vectors per call alive
            Yv = ham.matvec_couplings_device(v.contiguous())
            if Yv.is_complex():
                grad_s = (g.conj().unsqueeze(0) * Yv).real.sum(dim=1)
            else:
                grad_s = (g.unsqueeze(0) * Yv).sum(dim=1)
            grad_s = grad_s.to(s.dtype).to(s.device)
        return grad_s, grad_v, None


class DifferentiableHamiltonian:
    """H(J) as a differentiable linear map: ``dh(J, v)`` = H(J) v.

    ``assemble``: True builds a coupling-tagged table for the operator (an
    error if its values do not fit one), False keeps whatever the operator
    has, None (default) tries the table for spaces of 65536 states or more
    and keeps the matrix-free path when the values do not fit."""

    def __init__(self, ham, coupling_values: Optional[Sequence[float]] = None, assemble: Optional[bool] = None,
                 max_gb: float = 0.0):
        from . import _lib
        self.ham = ham
        if coupling_values is not None:
            ham.coupling_values = np.asarray(coupling_values, dtype=np.float64)
        self.j0 = _couplings_of(ham)
        self.n_couplings = len(self.j0)
        if assemble or (assemble is None and ham.nh >= 65536 and not ham.assembly["couplings"]):
            try:
                ham.assemble(True, max_gb=max_gb, couplings=True)
            except _lib.QesError:
                if assemble:
                    raise

    @property
    def J0(self) -> torch.Tensor:
        return torch.tensor(self.j0, dtype=torch.float64)

    def scale(self, J: torch.Tensor) -> torch.Tensor:
        return J / torch.as_tensor(self.j0, dtype=J.dtype, device=J.device)

    def __call__(self, J: torch.Tensor, v: torch.Tensor) -> torch.Tensor:
        return HamiltonianMatvec.apply(self.scale(J), v, self.ham)


def _dot(a, b):
    """<a, b> = sum conj(a) b (real for real tensors)."""
    return torch.vdot(a, b) if a.is_complex() else torch.dot(a, b)


def _lanczos_segment(H: DifferentiableHamiltonian, J, v, v_prev, beta_prev, nsteps: int):
    """nsteps of the three-term recurrence from (v, v_prev, beta_prev); returns
    the new (v, v_prev, beta_prev) and the alphas/betas of these steps."""
    alphas, betas = [], []
    for _ in range(nsteps):
        w = H(J, v)
        a = _dot(v, w).real
        w = w - a.to(w.dtype) * v - beta_prev.to(w.dtype) * v_prev
        b = torch.linalg.vector_norm(w)
        alphas.append(a)
        betas.append(b)
        v_prev, v = v, w / b.to(w.dtype)
        beta_prev = b
    return v, v_prev, beta_prev, torch.stack(alphas), torch.stack(betas)


def lanczos_tridiagonal(H: DifferentiableHamiltonian, J: torch.Tensor, v0: torch.Tensor, m: int,
                        segment: Optional[int] = None):
    """(alpha[m], beta[m-1]) of an m-step Lanczos run from v0, differentiable
    in J (and v0).  Segments of ``segment`` steps are recomputed in backward
    (default ceil(sqrt(m))), so memory holds O(m/segment + segment) vectors
    instead of the whole basis."""
    seg = int(segment) if segment else max(1, int(math.ceil(math.sqrt(m))))
    v = v0 / torch.linalg.vector_norm(v0).to(v0.dtype)
    v_prev = torch.zeros_like(v)
    beta_prev = torch.zeros((), dtype=torch.float64, device=v.device)
    A, B = [], []
    done = 0
    while done < m:
        k = min(seg, m - done)
        v, v_prev, beta_prev, a, b = checkpoint(_lanczos_segment, H, J, v, v_prev, beta_prev, k,
                                                use_reentrant=False)
        A.append(a)
        B.append(b)
        done += k
    alpha = torch.cat(A)
    beta = torch.cat(B)
    return alpha, beta[: m - 1]


def _tridiag(alpha, beta):
    T = torch.diag(alpha)
    if len(beta):
        T = T + torch.diag(beta, 1) + torch.diag(beta, -1)
    return T


def lanczos_energy(H: DifferentiableHamiltonian, J: torch.Tensor, v0: torch.Tensor, m: int,
                   segment: Optional[int] = None) -> torch.Tensor:
    """Lowest Ritz value of an m-step Lanczos run, differentiable in J.  At
    convergence its gradient is the Hellmann-Feynman one."""
    alpha, beta = lanczos_tridiagonal(H, J, v0, m, segment)
    return torch.linalg.eigvalsh(_tridiag(alpha, beta))[0]


def hellmann_feynman(H: DifferentiableHamiltonian, psi: torch.Tensor) -> torch.Tensor:
    """dE0/dJ_c = <psi|H_c|psi> / (<psi|psi> J0_c) for an eigenvector psi."""
    nrm = _dot(psi, psi).real
    if H.ham.assembly["couplings"]:
        vals = torch.as_tensor(H.ham.coupling_dots_device(psi, psi).real) / nrm.cpu()
    else:
        Y = H.ham.matvec_couplings_device(psi.contiguous())
        vals = torch.stack([_dot(psi, Y[c]).real for c in range(Y.shape[0])]).cpu() / nrm.cpu()
    return vals / torch.as_tensor(H.j0, dtype=torch.float64)


def krylov_propagate(H: DifferentiableHamiltonian, J: torch.Tensor, phi: torch.Tensor, t: float, m: int,
                     segment: Optional[int] = None) -> torch.Tensor:
    """exp(-i H(J) t) phi by the m-step Krylov (Lanczos) propagator
    |phi| V_m exp(-i T_m t) e_1.  The basis is regenerated in a second pass
    (same recurrence, same J) to form the combination, and every pass
    recomputes its segments in backward."""
    phi = phi.to(torch.complex128)
    nrm = torch.linalg.vector_norm(phi)
    alpha, beta = lanczos_tridiagonal(H, J, phi, m, segment)
    T = _tridiag(alpha, beta).to(torch.complex128)
    c = torch.linalg.matrix_exp(-1j * t * T)[:, 0] * nrm.to(torch.complex128)
    # second pass: accumulate sum_j c_j v_j without storing V_m
    seg = int(segment) if segment else max(1, int(math.ceil(math.sqrt(m))))

    def run(J_, c_, v, v_prev, beta_prev, acc, j0, k):
        for j in range(j0, j0 + k):
            acc = acc + c_[j] * v
            if j == m - 1:
                break
            w = H(J_, v)
            w = w - alpha[j].to(w.dtype) * v - beta_prev.to(w.dtype) * v_prev
            b = beta[j]
            v_prev, v = v, w / b.to(w.dtype)
            beta_prev = b
        return v, v_prev, beta_prev, acc

    v = phi / nrm.to(phi.dtype)
    v_prev = torch.zeros_like(v)
    beta_prev = torch.zeros((), dtype=torch.float64, device=v.device)
    acc = torch.zeros_like(v)
    j = 0
    while j < m:
        k = min(seg, m - j)
        v, v_prev, beta_prev, acc = checkpoint(run, J, c, v, v_prev, beta_prev, acc, j, k, use_reentrant=False)
        j += k
    return acc


def krylov_overlap(H: DifferentiableHamiltonian, J: torch.Tensor, psi0: torch.Tensor, phi: torch.Tensor,
                   t: float = 1.0, m: int = 30, segment: Optional[int] = None) -> torch.Tensor:
    """|<psi0| exp(-i H(J) t) |phi>|^2 (psi0 held fixed)."""
    psi_t = krylov_propagate(H, J, phi, t, m, segment)
    ov = torch.vdot(psi0.to(torch.complex128).detach(), psi_t)
    return (ov.conj() * ov).real


def central_difference(f, J: torch.Tensor, h: float = 1e-5) -> torch.Tensor:
    """(f(J + h e_c) - f(J - h e_c)) / 2h for every coupling c."""
    out = torch.zeros_like(J)
    with torch.no_grad():
        for c in range(len(J)):
            e = torch.zeros_like(J)
            e[c] = h
            out[c] = (float(f(J + e)) - float(f(J - e))) / (2.0 * h)
    return out
