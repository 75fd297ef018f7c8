"""The allocation-free coupling dots (qes_op_coupling_dots_noalloc: partial
sums in a static device array, no hipMalloc) against the allocating entry
point qes_op_coupling_dots on the same operator and vectors, twice in a row,
the second call on a smaller operator; GPU."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dots(H, fn, v, g):
    from pyqusolver_amd import _lib
    ncpl = H.n_couplings
    out = np.full(2 * ncpl, np.nan)
    dtype = 1 if v.is_complex() else 0
    _lib.check(getattr(_lib.lib(), fn)(H._device_op(), _lib.dptr(v), _lib.dptr(g), dtype,
                                       out.ctypes.data_as(_lib._PD), _lib.stream_ptr()), fn)
    return out


def test_noalloc_coupling_dots_match_allocating_entry_point():
    import torch
    from pyqusolver_amd import ChainLattice, get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import XXZ
    # first call: 2,704,156 rows (more than the 2048 blocks x 256 rows of one
    # sweep of the fixed grid: every block strides); second: a small complex sector
    cases = [(24, {}, 12, 0.3), (14, {"translation": 3}, 7, 0.2)]
    for L, sym, n_up, hz in cases:
        lat = ChainLattice(L)
        hs = HilbertSpace(lattice=lat, sym_gen=sym, global_syms=[get_u1_sym(lat, n_up)])
        fields = np.random.default_rng(L).uniform(-1, 1, L) * hz
        H = XXZ(lat, hilbert_space=hs, jxy=1.0, jz=0.7, hz=fields)
        H.assemble(True, couplings=True)
        dt = torch.complex128 if H.needs_complex else torch.float64
        gen = torch.Generator(device="cuda").manual_seed(L)
        v = torch.randn(hs.nh, dtype=dt, device="cuda", generator=gen)
        g = torch.randn(hs.nh, dtype=dt, device="cuda", generator=gen)
        ref = _dots(H, "qes_op_coupling_dots", v, g)
        got = _dots(H, "qes_op_coupling_dots_noalloc", v, g)
        scale = np.abs(ref).max()
        assert scale > 0
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * scale)
        # and against the dot products of the per-coupling products
        Y = H.matvec_couplings_device(v)
        direct = np.array([torch.vdot(g, Y[c]).item() for c in range(H.n_couplings)])
        np.testing.assert_allclose(got[0::2] + 1j * got[1::2], direct, rtol=0, atol=1e-12 * scale)
        # a second call on the same operator returns the same bits
        assert np.array_equal(_dots(H, "qes_op_coupling_dots_noalloc", v, g), got)
        H.assemble(False)
