"""Dynamical correlations from Lanczos continued fractions.

G(z) = <phi| (z - H)^(-1) |phi>, z = omega + E0 + i eta, with
phi = O |psi0> taken into the sector of H by the sector-changing product
(Operator.matvec(..., hilbert_out=...)).  The device Lanczos run from phi
gives alpha, beta; G(z) = |phi|^2 / (z - a0 - b0^2 / (z - a1 - b1^2 / ...)).
S(q, omega) = -Im G / pi.  The reference reaches the same quantity through
its projected matvec (Algebra/Symmetries/jit/matrix_builder_jit.py:51-245)
and a Lanczos run; the continued fraction itself is the textbook one.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .lanczos import DeviceLanczos


@dataclass
class ContinuedFraction:
    alpha: np.ndarray
    beta: np.ndarray
    weight: float              # <phi|phi>

    def green(self, z: np.ndarray) -> np.ndarray:
        """G(z) for complex z (vectorised), evaluated bottom-up."""
        z = np.asarray(z, dtype=np.complex128)
        m = len(self.alpha)
        g = np.zeros_like(z)
        for j in range(m - 1, -1, -1):
            b2 = self.beta[j] ** 2 if j < m - 1 else 0.0
            g = 1.0 / (z - self.alpha[j] - b2 * g)
        return self.weight * g


def continued_fraction(ham, phi, m: int = 200, reorth: bool = True) -> ContinuedFraction:
    """Lanczos from phi (device or numpy vector in the sector of ``ham``)."""
    complex_ = True
    phi_t = phi if isinstance(phi, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(phi))
    phi_t = phi_t.to(device="cuda", dtype=torch.complex128).contiguous()
    w = float(torch.linalg.vector_norm(phi_t).item() ** 2)
    m = int(min(m, ham.nh))
    L = DeviceLanczos(ham, m, store_basis=bool(reorth), complex_=complex_)
    L.start(phi_t)
    L.run(m, reorth)
    a, b = L.coefficients()
    L.close()
    # stop at an invariant subspace (beta ~ 0): the fraction terminates there
    cut = len(a)
    for j in range(len(b)):
        if abs(b[j]) <= 1e-12 * max(1.0, abs(a[j])):
            cut = j + 1
            break
    return ContinuedFraction(alpha=a[:cut], beta=b[:cut], weight=w)


def structure_factor(ham_out, op, psi0, e0: float, omegas, eta: float = 0.05, m: int = 200,
                     hilbert_out=None) -> np.ndarray:
    """S(omega) = -Im <psi0| O^dagger (omega + E0 + i eta - H)^(-1) O |psi0> / pi
    with O |psi0> carried into ``hilbert_out`` (the sector of ``ham_out``)."""
    hs_out = hilbert_out if hilbert_out is not None else ham_out.hilbert_space
    phi = op.matvec(psi0, hilbert_out=hs_out)
    cf = continued_fraction(ham_out, phi, m=m)
    z = np.asarray(omegas, dtype=np.float64) + e0 + 1j * eta
    return -cf.green(z).imag / np.pi
