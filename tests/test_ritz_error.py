"""Stopping test of the Lanczos ground-state solvers (CPU only).

The solvers stop when the residual norm of the lowest Ritz pair,
r = |beta_m s_m| (s: unit eigenvector of T_m), falls below tol*max(1,|theta|)
-- the eigsh/ARPACK test; |theta - lambda| <= r for some eigenvalue lambda.
Pins: r equals ||H y - theta y|| of the Ritz vector built from an explicit
Krylov basis (the identity the test rests on); the library's host routine
(qes_tridiag_ritz, inverse iteration on T_m) agrees with the dense
eigendecomposition of T_m that the Python solver uses; r bounds the true
error of a converging run; with an unresolved near-degenerate pair, where a
test on successive Ritz values stops early, r stays above tol.
"""
import ctypes as C

import numpy as np
import pytest


def _lanczos(H, v0, m):
    """m steps of Lanczos with full reorthogonalisation (numpy)."""
    n = H.shape[0]
    V = np.zeros((n, m + 1))
    a, b = np.zeros(m), np.zeros(m)
    V[:, 0] = v0 / np.linalg.norm(v0)
    for j in range(m):
        w = H @ V[:, j]
        a[j] = V[:, j] @ w
        w -= V[:, : j + 1] @ (V[:, : j + 1].T @ w)
        w -= V[:, : j + 1] @ (V[:, : j + 1].T @ w)
        b[j] = np.linalg.norm(w)
        V[:, j + 1] = w / b[j]
    return a, b, V[:, :m]


def _lib_ritz(a, b):
    from pyqusolver_amd import _lib
    th, r, q = C.c_double(), C.c_double(), C.c_double()
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    _lib.check(_lib.lib().qes_tridiag_ritz(a.ctypes.data_as(_lib._PD), b.ctypes.data_as(_lib._PD), len(a),
                                           C.byref(th), C.byref(r), C.byref(q)), "qes_tridiag_ritz")
    return th.value, r.value, q.value


def _random_h(n, spectrum, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (Q * spectrum) @ Q.T, rng.standard_normal(n)


@pytest.mark.parametrize("m", [1, 2, 5, 12, 25, 40])
def test_residual_identity_and_library_agreement(m):
    from pyqusolver_amd.lanczos import ritz_residual, tridiag_eigs
    n = 120
    spec = np.concatenate([[-3.0, -2.6, -2.55], np.linspace(-2.0, 2.0, n - 3)])
    H, v0 = _random_h(n, spec, 1)
    a, b, V = _lanczos(H, v0, m)
    w, S = tridiag_eigs(a, b)
    y = V @ S[:, 0]
    r_true = np.linalg.norm(H @ y - w[0] * y)
    r, quad = ritz_residual(a, b)
    assert abs(r - r_true) <= 1e-10 * max(1.0, r_true)
    expect_quad = r_true if m == 1 else min(r_true, r_true ** 2 / (w[1] - w[0]))
    assert abs(quad - expect_quad) <= 1e-8 * expect_quad + 1e-14
    th, r_lib, q_lib = _lib_ritz(a, b)
    assert abs(th - w[0]) <= 1e-13 * max(1.0, abs(w[0]))
    assert abs(r_lib - r) <= 1e-6 * r + 1e-15
    assert abs(q_lib - quad) <= 1e-6 * quad + 1e-15


def test_residual_bounds_true_error_of_converging_run():
    from pyqusolver_amd.lanczos import ritz_residual, tridiag_eigs
    n = 300
    spec = np.concatenate([[-4.0, -3.5], np.linspace(-3.0, 3.0, n - 2)])
    H, v0 = _random_h(n, spec, 2)
    a, b, _ = _lanczos(H, v0, 60)
    for m in range(2, 61, 2):
        w, _ = tridiag_eigs(a[:m], b[:m])
        err_true = w[0] - (-4.0)
        r, quad = ritz_residual(a[:m], b[:m])
        assert err_true >= -1e-12
        if err_true < 0.25:      # -4 is then the eigenvalue nearest theta, within r of it
            assert err_true <= r * (1 + 1e-9) + 1e-14, (m, err_true, r)
        if quad < 1e-4:          # separated: the quadratic estimate is tight
            assert err_true <= 1.5 * quad + 1e-13, (m, err_true, quad)


def test_unresolved_near_degenerate_pair_is_not_declared_converged():
    """Lowest two eigenvalues 1e-7 apart, 0.2 below the rest of [-1, 1]: the
    Krylov space captures the pair long before it resolves it; meanwhile the
    Ritz value changes by less than tol over 10 steps while still off by
    about the splitting (a test on successive Ritz values stops there), and r
    (the splitting times |sin(2 phi)|/2 for the Ritz vector's mixing angle
    phi, plus the rest) stays above tol, so the residual test does not."""
    from pyqusolver_amd.lanczos import ritz_residual, tridiag_eigs
    n = 400
    delta = 1e-7
    spec = np.concatenate([[-1.2 - delta, -1.2], np.linspace(-1.0, 1.0, n - 2)])
    H, v0 = _random_h(n, spec, 3)
    a, b, _ = _lanczos(H, v0, 80)
    tol = 1e-10
    stagnation_fooled = False
    prev = None
    for m in range(10, 81, 10):
        w, _ = tridiag_eigs(a[:m], b[:m])
        theta = w[0]
        err_true = theta - spec[0]
        r, _ = ritz_residual(a[:m], b[:m])
        _, r_lib, _ = _lib_ritz(a[:m], b[:m])
        scale = max(1.0, abs(theta))
        if r <= tol * scale or r_lib <= tol * scale:
            assert err_true <= tol * scale, (m, err_true, r)
        if prev is not None and abs(theta - prev) <= tol * scale and err_true > 100 * tol * scale:
            stagnation_fooled = True
        prev = theta
    assert stagnation_fooled


def test_breakdown_and_one_step_cases():
    from pyqusolver_amd.lanczos import ritz_residual
    th, r, q = _lib_ritz([0.5], [0.0])
    assert th == pytest.approx(0.5) and r == 0.0
    th, r, q = _lib_ritz([0.5], [0.25])
    assert r == pytest.approx(0.25)
    assert ritz_residual(np.array([0.5]), np.array([0.25]))[0] == pytest.approx(0.25)
    # exactly decoupled T (interior zero coupling): finite, and the lowest
    # eigenvector lives in the trailing block
    th, r, q = _lib_ritz([1.0, -1.0, 0.0], [0.0, 0.5, 1e-3])
    T = np.array([[1.0, 0, 0], [0, -1.0, 0.5], [0, 0.5, 0.0]])
    w, S = np.linalg.eigh(T)
    assert th == pytest.approx(w[0])
    assert r == pytest.approx(1e-3 * abs(S[-1, 0]), rel=1e-8)


def test_next_check_interval():
    """The next stopping test goes where the geometric decay between the last
    two tests reaches the target, within [lo, hi]."""
    from pyqusolver_amd.lanczos import next_check
    # r falls by 10x per 10 steps: from 1e-8 at step 90, 1e-10 is 20 steps on
    assert next_check(80, 1e-7, 90, 1e-8, 1e-10, 1, 10) == 10
    assert next_check(80, 1e-7, 90, 1e-8, 1e-10, 1, 30) == 20
    # 1e-9.5 is 5 steps on (ceil of the exact 5.0 up to rounding)
    assert next_check(80, 1e-7, 90, 1e-8, 10 ** -8.5, 1, 10) in (5, 6)
    assert next_check(80, 1e-7, 90, 1e-8, 9.99e-9, 1, 10) == 1
    assert next_check(80, 1e-7, 90, 1e-8, 9.99e-9, 3, 10) == 3
    # not decreasing, no history, or already below target: the longest interval
    assert next_check(80, 1e-8, 90, 1e-8, 1e-10, 1, 10) == 10
    assert next_check(80, 1e-9, 90, 1e-8, 1e-10, 1, 10) == 10
    assert next_check(0, 0.0, 10, 1e-3, 1e-10, 1, 10) == 10
    assert next_check(80, 1e-7, 90, 1e-11, 1e-10, 1, 10) == 10
    # along an exact geometric sequence the predicted step is the first that passes
    rho, target = 0.7, 1e-10
    r = lambda m: 3.0 * rho ** m
    first = next(m for m in range(1000) if r(m) <= target)
    m0, m1 = first - 25, first - 15
    assert m1 + next_check(m0, r(m0), m1, r(m1), target, 1, 100) == first


def test_ritz_selection_by_spectrum_end():
    """which='smallest'/'largest'/'both' select the k lowest, k highest, or
    k//2 lowest and the rest highest Ritz values (eigsh's SA/LA/BE), in
    ascending order; unknown ends raise."""
    from pyqusolver_amd.lanczos import select_ritz, which_code
    assert [which_code(w) for w in ("smallest", "SA", "largest", "la", "both", "BE")] == \
        ["SA", "SA", "LA", "LA", "BE", "BE"]
    with pytest.raises(ValueError):
        which_code("LM")
    assert select_ritz(10, 3, "SA").tolist() == [0, 1, 2]
    assert select_ritz(10, 3, "LA").tolist() == [7, 8, 9]
    assert select_ritz(10, 3, "BE").tolist() == [0, 8, 9]
    assert select_ritz(10, 4, "BE").tolist() == [0, 1, 8, 9]
    assert select_ritz(2, 6, "SA").tolist() == [0, 1]        # k clipped to the Krylov dimension
    assert select_ritz(3, 6, "BE").tolist() == [0, 1, 2]
    assert select_ritz(5, 1, "BE").tolist() == [4]


def test_effective_steps_drops_rows_after_a_breakdown():
    """A problem that broke down (beta = 0) and kept stepping on a zero
    vector appends rows (0, 0): its T would carry a spurious eigenvalue 0.
    Only the rows up to the breakdown span the Krylov space."""
    from pyqusolver_amd.lanczos import effective_steps, tridiag_eigs
    a = np.array([2.0, 0.0, 0.0])
    b = np.array([0.0, 0.0, 0.0])
    assert tridiag_eigs(a, b)[0][0] == 0.0          # the spurious value
    k = effective_steps(a, b)
    assert k == 1 and tridiag_eigs(a[:k], b[:k])[0][0] == 2.0
    a = np.array([1.0, 3.0, 0.0, 0.0])
    b = np.array([0.5, 0.0, 0.0, 0.0])
    assert effective_steps(a, b) == 2
    a = np.array([1.0, 3.0, 2.0])
    b = np.array([0.5, 0.4, 0.1])
    assert effective_steps(a, b) == 3
