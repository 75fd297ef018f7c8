"""pyqusolver_amd -- MI355X-native symmetric-sector ED matvec and Lanczos.

Drop-in for the reference's (QES) spin-1/2 Hamiltonian matvec on a
symmetry-reduced Hilbert space (SpecialOperator.matvec_fun /
Operator.matvec / scipy LinearOperator) and the Lanczos ground state it
drives.  All compute runs in libqes_mi.so (hand-written HIP for gfx950).
"""
from .lattice import ChainLattice, HoneycombLattice, SquareLattice  # noqa: F401
from .symmetry import GlobalSymmetries, SymmetryGenerators, get_u1_sym, get_z2_sym  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    # GPU-facing classes import torch/ctypes lazily
    if name == "HilbertSpace":
        from .hilbert import HilbertSpace
        return HilbertSpace
    if name in ("Hamiltonian", "spin_op"):
        from . import operator
        return getattr(operator, name)
    if name in ("J1J2Model", "XXZ", "HeisenbergKitaev"):
        from . import models
        return getattr(models, name)
    if name in ("DifferentiableHamiltonian", "HamiltonianMatvec"):
        from . import autograd
        return getattr(autograd, name)
    if name in ("BatchedLanczos", "CouplingSweep"):
        from . import batched
        return getattr(batched, name)
    raise AttributeError(name)
