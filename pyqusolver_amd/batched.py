"""Batched ground states of many sectors / disorder realisations in one go.

The reference sweeps symmetry sectors sequentially
(HilbertSpace.iter_momentum_sectors / iter_symmetry_sectors,
Algebra/hilbert.py:412-652) and diagonalises each one separately.  Here all
problems advance together on the device (libqes_mi batched Lanczos): one
persistent matvec launch over a tile queue of every problem and one launch per
vector step, replayed from a captured hipGraph, with converged problems masked
out.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .lanczos import start_vector
from .symmetry import SymmetryGenerators


@dataclass
class BatchResult:
    e0: np.ndarray
    steps: np.ndarray
    seconds: float
    schedule: Optional[list] = None     # autotuned schedule of each batch (batched_ground_states)
    # residual norm |beta_m s_m| of each lowest Ritz pair: |E0 - lambda| <= residual
    # for an eigenvalue lambda (the FP64 error bound of each result)
    residuals: Optional[np.ndarray] = None


_DIAGONAL_CODES = (13, 22)   # sigma_z local / sigma_z correlation (operators_spin.py:1084-1118)


def _offdiag_signature(h):
    """Key of everything but the diagonal: the basis object, the build type
    and the off-diagonal instructions with their effective coefficients."""
    sites, coeffs, codes = h.instruction_arrays()
    coeffs = np.asarray(coeffs, dtype=np.complex128)
    scale = getattr(h, "_coupling_scale", None)
    if scale is not None and len(codes):
        coeffs = coeffs * scale[np.asarray(h._instr_coupling, dtype=np.int64)]
    off = [(int(c), int(s[0]), int(s[1]), complex(v)) for c, s, v in zip(codes, sites, coeffs)
           if int(c) not in _DIAGONAL_CODES]
    return (id(h.hilbert_space), h._dtype_is_complex(), tuple(off))


class BatchedLanczos:
    """Owner of a qes_batch handle for a list of Hamiltonians.

    Hamiltonians on the same HilbertSpace object whose off-diagonal
    instructions coincide (disorder realisations of a random field) are
    advanced as multi-RHS blocks (up to 256 columns each) unless
    ``blocks=False``; the others are ordinary problems.  Results are always
    reported in the order of ``hams``.
    """

    MAX_RHS = 256

    def __init__(self, hams: Sequence, max_steps: int = 300, complex_: Optional[bool] = None,
                 blocks: bool = True, min_block: int = 2, assemble: bool = True):
        if not hams:
            raise ValueError("empty batch")
        self.hams = list(hams)
        cplx = any(h.needs_complex for h in self.hams) if complex_ is None else bool(complex_)
        self.dtype = 1 if cplx else 0
        self.max_steps = int(max_steps)
        groups = {}
        if blocks:
            for i, h in enumerate(self.hams):
                groups.setdefault(_offdiag_signature(h), []).append(i)
        regular, blocks_ = [], []
        for idx in (groups.values() if blocks else [[i] for i in range(len(self.hams))]):
            if len(idx) < max(2, min_block):
                regular += idx
                continue
            for s in range(0, len(idx), self.MAX_RHS):
                blocks_.append(idx[s:s + self.MAX_RHS])
        regular.sort()
        # ordinary problems stream an assembled table of their operator (built
        # here when it fits a quarter of the free memory, freed by close())
        self._assembled_here = []
        if assemble and regular:
            free, _ = torch.cuda.mem_get_info()
            budget = 0.25 * free
            for i in regular:
                h = self.hams[i]
                if h.assembly_bytes:
                    continue
                try:
                    # the batch streams the one-pass table; column panels
                    # serve single-vector products only (restored by close())
                    h.set_option("asm_panel_cols", 0)
                except _lib.QesError:   # in use by another batch or handle: its table cannot change
                    continue
                try:
                    nb = h.assemble(True, max_gb=budget / 1e9)
                except _lib.QesError:
                    h.set_option("asm_panel_cols", -1)
                    continue
                budget -= nb
                self._assembled_here.append(h)
        self._diag = []
        block_ops, block_r, block_d = [], [], []
        for members in blocks_:
            d = self._block_diagonal(members)
            if d is None:                      # complex diagonal: ordinary problems
                regular += members
                continue
            self._diag.append(d)
            block_ops.append(self.hams[members[0]]._device_op().value)
            block_r.append(len(members))
            block_d.append(d.data_ptr())
        self.regular = regular
        self.blocks = [b for b in blocks_ if not set(b) & set(regular)]
        # virtual problem number of every Hamiltonian
        order = list(regular) + [i for b in self.blocks for i in b]
        self.vp_of = np.empty(len(self.hams), dtype=np.int64)
        self.vp_of[np.asarray(order, dtype=np.int64)] = np.arange(len(order))
        ops = (C.c_void_p * max(len(regular), 1))(*[self.hams[i]._device_op().value for i in regular])
        nb = len(self.blocks)
        bops = (C.c_void_p * max(nb, 1))(*block_ops)
        br = (C.c_int32 * max(nb, 1))(*block_r)
        bd = (C.c_void_p * max(nb, 1))(*block_d)
        h = C.c_void_p()
        _lib.check(_lib.lib().qes_batch_create_blocks(ops, len(regular), bops, br, bd, nb, self.max_steps, self.dtype,
                                                      C.byref(h)), "qes_batch_create_blocks")
        self._h = h
        offs = np.zeros(len(regular) + nb + 1, dtype=np.int64)
        _lib.check(_lib.lib().qes_batch_offsets(self._h, offs.ctypes.data_as(_lib._PI64)))
        self.offsets = offs

    def _block_diagonal(self, members):
        """(n, R) float64 device array of H_aa of every member, or None when a
        diagonal has an imaginary part."""
        n = self.hams[members[0]].nh
        re = torch.empty((len(members), n), dtype=torch.float64, device="cuda")
        im = torch.empty((len(members), n), dtype=torch.float64, device="cuda")
        for r, i in enumerate(members):
            _lib.check(_lib.lib().qes_op_diagonal(self.hams[i]._device_op(), _lib.dptr(re[r]), _lib.dptr(im[r]),
                                                  _lib.stream_ptr()), "qes_op_diagonal")
        if bool((im != 0).any()):
            return None
        return re.t().contiguous()

    @property
    def torch_dtype(self):
        return torch.complex128 if self.dtype == 1 else torch.float64

    # candidate values of the persistent matvec schedule, tried in this order
    SCHEDULE_SPACE = (("tile_queue", (0, 1)), ("tile_order", (1, 0)), ("grid_mult", (8, 4, 16, 2)),
                      ("block_rows", (16, 8, 32)))

    def autotune(self, reps: int = 3, space=None) -> dict:
        """Choose the schedule of the persistent batched matvec (static or
        dynamic tile queue, tile order, blocks per CU, rows per block tile)
        for this batch by timing it:
        coordinate search over ``SCHEDULE_SPACE``, each candidate timed as the
        median of ``reps`` products on a fixed vector with device events; the
        fastest value of each knob is kept before the next knob is tried.
        The products are identical under every schedule, so tuning changes
        only speed.  Returns {option: value, ..., "ms": best time}."""
        space = self.SCHEDULE_SPACE if space is None else space
        total = int(self.offsets[-1])
        g = torch.Generator(device="cuda").manual_seed(0)
        x = torch.randn(total, dtype=torch.float64, device="cuda", generator=g)
        if self.dtype == 1:
            x = torch.complex(x, torch.randn(total, dtype=torch.float64, device="cuda", generator=g))
        y = torch.empty_like(x)
        stream = torch.cuda.current_stream()
        L = _lib.lib()

        def timed() -> float:
            L.qes_batch_matvec(self._h, _lib.dptr(x), _lib.dptr(y), _lib.stream_ptr())    # warm
            ts = []
            for _ in range(max(1, reps)):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                _lib.check(L.qes_batch_matvec(self._h, _lib.dptr(x), _lib.dptr(y), _lib.stream_ptr()),
                           "qes_batch_matvec")
                b.record(stream)
                b.synchronize()
                ts.append(a.elapsed_time(b))
            return float(np.median(ts))
        has_blocks = bool(self.blocks)
        chosen, best_ms = {}, None
        for name, values in space:
            if name == "block_rows" and not has_blocks:
                continue
            times = {}
            for v in values:
                self.set_option(name, v)
                times[v] = timed()
            v_best = min(times, key=times.get)
            self.set_option(name, v_best)
            chosen[name] = v_best
            best_ms = times[v_best]
        chosen["ms"] = best_ms
        self.schedule = chosen
        return chosen

    def set_option(self, name: str, value: int):
        """Schedule knob: tile_order (0/1), grid_mult (blocks per CU),
        block_rows (rows per tile of a block problem), block_table (0:
        matrix-free block rows, 1: cached neighbour table) or op_tables (0:
        matrix-free ordinary problems, 1: their assembled tables); check_min
        (ground states: 0 fixed test interval, n >= 1 adaptive, see
        ``ground_states``)."""
        _lib.check(_lib.lib().qes_batch_set_option(self._h, name.encode(), int(value)), "qes_batch_set_option")

    def concat(self, vecs: Sequence) -> torch.Tensor:
        """Device vector in the batch layout from one vector (numpy or torch)
        per Hamiltonian: ordinary problems concatenated, then each block stored
        (n, R) row-major."""
        dt = self.torch_dtype

        def dev(v):
            t = v if isinstance(v, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(v))
            return t.to(device="cuda", dtype=dt)
        parts = [dev(vecs[i]) for i in self.regular]
        for b in self.blocks:
            parts.append(torch.stack([dev(vecs[i]) for i in b], dim=1).reshape(-1))
        return torch.cat(parts).contiguous()

    def split(self, x: torch.Tensor) -> List[torch.Tensor]:
        """Inverse of concat: one device vector per Hamiltonian."""
        out = [None] * len(self.hams)
        k = 0
        for i in self.regular:
            out[i] = x[int(self.offsets[k]):int(self.offsets[k + 1])]
            k += 1
        for b in self.blocks:
            m = x[int(self.offsets[k]):int(self.offsets[k + 1])].reshape(-1, len(b))
            for r, i in enumerate(b):
                out[i] = m[:, r]
            k += 1
        return out

    def matvec(self, x: torch.Tensor) -> torch.Tensor:
        """y = H_p x_p for every problem, on vectors in the batch layout."""
        y = torch.empty_like(x)
        _lib.check(_lib.lib().qes_batch_matvec(self._h, _lib.dptr(x), _lib.dptr(y), _lib.stream_ptr()),
                   "qes_batch_matvec")
        return y

    def start(self, v0: Optional[torch.Tensor] = None, seeds: Optional[Sequence[int]] = None):
        if v0 is None:
            seeds = list(range(len(self.hams))) if seeds is None else list(seeds)
            v0 = self.concat([start_vector(h.nh, self.dtype == 1, s) for h, s in zip(self.hams, seeds)])
        v0 = v0.to(self.torch_dtype).contiguous()
        _lib.check(_lib.lib().qes_batch_start(self._h, _lib.dptr(v0), _lib.stream_ptr()), "qes_batch_start")

    def capture(self):
        """Capture the step graph now (no steps run)."""
        _lib.check(_lib.lib().qes_batch_run(self._h, 0, 1), "qes_batch_run")

    def ground_states(self, tol: float = 1e-10, check_every: int = 10, v0=None, seeds=None,
                      sync_device: bool = True, check_min: Optional[int] = None) -> BatchResult:
        """Lanczos on every problem to the residual test (qes_batch_ground_states).
        ``check_min`` (the batch option of that name): 0 tests every
        ``check_every`` steps; n >= 1 places each test where the active
        problems' residual decay predicts the first crossing, n..check_every
        steps apart; -1 (the batch's default): 1 for at most 4 problems of
        >= 2^20 rows (a coupling sweep of one large sector), else 0 -- with
        many problems the earliest predicted crossing keeps the interval
        short and every test is a host round trip; None keeps the batch's
        setting."""
        if check_min is not None:
            self.set_option("check_min", int(check_min))
        if sync_device:
            torch.cuda.synchronize()
        else:
            torch.cuda.current_stream().synchronize()
        t0 = time.perf_counter()
        self.start(v0, seeds)
        V = len(self.hams)
        e0 = np.zeros(V)
        steps = np.zeros(V, dtype=np.int32)
        _lib.check(_lib.lib().qes_batch_ground_states(self._h, float(tol), int(check_every),
                                                      e0.ctypes.data_as(_lib._PD), steps.ctypes.data_as(_lib._PI32)),
                   "qes_batch_ground_states")
        return BatchResult(e0=e0[self.vp_of], steps=steps[self.vp_of], seconds=time.perf_counter() - t0)

    def residuals(self) -> np.ndarray:
        """Residual norm |beta_m s_m| of the lowest Ritz pair of every problem
        after the last run (order of hams), from the recurrence coefficients;
        0 after a breakdown (the Krylov space is invariant)."""
        from .lanczos import ritz_residual
        a, b, k = self.coefficients()
        out = np.zeros(len(k))
        for i, m in enumerate(k):
            m = int(m)
            if m > 0:
                out[i] = ritz_residual(a[i, :m], b[i, :m])[0]
        return out

    def coefficients(self):
        """alpha, beta (one row per Hamiltonian, in the order of hams) and steps
        (cut after a breakdown: rows past it belong to a zero vector)."""
        from .lanczos import effective_steps
        P, M = len(self.hams), self.max_steps
        a = np.zeros(P * M)
        b = np.zeros(P * M)
        k = np.zeros(P, dtype=np.int32)
        _lib.check(_lib.lib().qes_batch_get(self._h, a.ctypes.data_as(_lib._PD), b.ctypes.data_as(_lib._PD),
                                            k.ctypes.data_as(_lib._PI32)), "qes_batch_get")
        a, b, k = a.reshape(P, M)[self.vp_of], b.reshape(P, M)[self.vp_of], k[self.vp_of]
        for i in range(len(k)):
            k[i] = effective_steps(a[i, :k[i]], b[i, :k[i]])
        return a, b, k

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().qes_batch_destroy(self._h)
            self._h = None
        for h in getattr(self, "_assembled_here", []):
            try:
                h.assemble(False)
                h.set_option("asm_panel_cols", -1)
            except Exception:
                pass
        self._assembled_here = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CouplingSweep(BatchedLanczos):
    """Ground states of every sector in ``hams`` at several coupling sets at
    once: column r of sector p is its own Lanczos run on
    H_r = sum_c scales[r][c] * H_c (H_c = the instructions tagged with
    coupling id c, at the values the model was built with).  Each sector's
    coupling-tagged table is streamed once per step for all columns.

    ``scales``: (ncol, n_couplings) shared by all sectors, or (nprob, ncol,
    n_couplings), ncol <= 4.  Results (``ground_states``) are in the order
    sector 0 columns 0..ncol-1, sector 1, ... (``e0.reshape(nprob, ncol)``).
    Tagged tables are built here when missing (replacing an untagged one)
    and freed by ``close``."""

    MAX_CPL = 8

    def __init__(self, hams: Sequence, scales, max_steps: int = 300, complex_: Optional[bool] = None,
                 max_gb: float = 0.0):
        if not hams:
            raise ValueError("empty sweep")
        hams = list(hams)
        sc = np.asarray(scales, dtype=np.float64)
        if sc.ndim == 2:
            sc = np.broadcast_to(sc, (len(hams),) + sc.shape)
        if sc.ndim != 3 or sc.shape[0] != len(hams):
            raise ValueError("scales must be (ncol, n_couplings) or (nprob, ncol, n_couplings)")
        ncol = sc.shape[1]
        if not 1 <= ncol <= 4:
            raise ValueError("a coupling sweep runs 1..4 coupling sets per sector")
        cplx = any(h.needs_complex for h in hams) if complex_ is None else bool(complex_)
        self.dtype = 1 if cplx else 0
        self.max_steps = int(max_steps)
        self.ncol = ncol
        self.sectors = hams
        self._assembled_here = []
        packed = np.zeros((len(hams), ncol, self.MAX_CPL))
        for p, h in enumerate(hams):
            nc = h.n_couplings
            if sc.shape[2] < nc:
                raise ValueError(f"sector {p} has {nc} coupling ids, scales give {sc.shape[2]}")
            packed[p, :, :nc] = sc[p, :, :nc]
            if not h.assembly["couplings"]:
                # the sweep streams the one-pass table; column panels serve
                # single-vector products only (restored by close())
                h.set_option("asm_panel_cols", 0)
                h.assemble(True, max_gb=max_gb, couplings=True)
                self._assembled_here.append(h)
        self.scales = packed
        self.hams = [h for h in hams for _ in range(ncol)]          # one entry per virtual problem
        self.regular = []
        self.blocks = [[p * ncol + r for r in range(ncol)] for p in range(len(hams))]
        self.vp_of = np.arange(len(self.hams), dtype=np.int64)
        ops = (C.c_void_p * len(hams))(*[h._device_op().value for h in hams])
        flat = np.ascontiguousarray(packed.reshape(-1))
        h_ = C.c_void_p()
        _lib.check(_lib.lib().qes_batch_create_scaled(ops, len(hams), ncol, flat.ctypes.data_as(_lib._PD),
                                                      self.max_steps, self.dtype, C.byref(h_)),
                   "qes_batch_create_scaled")
        self._h = h_
        offs = np.zeros(len(hams) + 1, dtype=np.int64)
        _lib.check(_lib.lib().qes_batch_offsets(self._h, offs.ctypes.data_as(_lib._PI64)))
        self.offsets = offs

    def ground_states(self, tol: float = 1e-10, check_every: int = 10, v0=None, seeds=None,
                      sync_device: bool = True, check_min: Optional[int] = None) -> BatchResult:
        """Seeds default to sector index (the same start vector for every
        column of a sector, as separate solves with seed = sector would use)."""
        if seeds is None and v0 is None:
            seeds = [p for p in range(len(self.sectors)) for _ in range(self.ncol)]
        return super().ground_states(tol=tol, check_every=check_every, v0=v0, seeds=seeds, sync_device=sync_device,
                                     check_min=check_min)


def momentum_sectors(L: int, reflection: bool = True):
    """sym_gen dicts of a chain's momentum sectors in the order of
    HilbertSpace.iter_momentum_sectors (hilbert.py:546-652), with reflection
    parities only where they form a symmetry sector (k = 0 and k = L/2)."""
    from .hilbert import momentum_symmetry_types, sector_specs
    from .lattice import ChainLattice
    lat = ChainLattice(L)
    out = []
    for sector, _ in sector_specs(momentum_symmetry_types(lat, reflection), lat):
        d = {"translation": sector[SymmetryGenerators.Translation_x]}
        if sector.get(SymmetryGenerators.Reflection, 0):
            d["reflection"] = sector[SymmetryGenerators.Reflection]
        out.append(d)
    return out


def sector_by_sector(hams: Sequence, tol: float = 1e-10, max_steps: int = 300, check_every: int = 10, seeds=None,
                     assemble: bool = True):
    """Reference-style sweep on the same GPU: one device Lanczos per problem,
    same recurrence (no reorthogonalisation) and convergence rule; with
    ``assemble`` each operator is tabulated for its own solve (and freed
    after), as a sequential caller of this library would do."""
    from .lanczos import lanczos_ground_state
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0 = []
    steps = []
    for i, h in enumerate(hams):
        built = False
        if assemble and h.assembly_bytes == 0:
            try:
                h.assemble(True)
                built = True
            except _lib.QesError:
                pass
        r = lanczos_ground_state(h, k=1, tol=tol, maxiter=max_steps, seed=(seeds[i] if seeds else i), reorth=False,
                                 store_basis=False, check_every=check_every)
        if built:
            h.assemble(False)
        e0.append(r.eigenvalues[0])
        steps.append(r.steps)
    return BatchResult(e0=np.asarray(e0), steps=np.asarray(steps), seconds=time.perf_counter() - t0)


def batched_ground_states(hams: Sequence, tol: float = 1e-10, max_steps: int = 300, check_every: int = 10,
                          seeds=None, split_dtypes: bool = True, concurrent: bool = True, warmup: bool = False,
                          autotune: bool = True, check_min: Optional[int] = None):
    """Ground states of many problems with batched Lanczos.

    Problems that need complex128 vectors (generic momenta, complex models)
    and the real ones go to two batches, so the real problems (and their
    realisation blocks) move half the bytes; the two batches run on their
    own streams from two host threads and overlap on the device.  Seeds
    default to the problem's position, as in ``sector_by_sector``.  With
    ``autotune`` each batch's matvec schedule is chosen by timing
    (``BatchedLanczos.autotune``).  Returns (BatchResult in the order of
    ``hams``, seconds spent creating the batches: tables, diagonals,
    tuning).  ``check_min``: see ``BatchedLanczos.ground_states``."""
    import threading
    seeds = list(range(len(hams))) if seeds is None else list(seeds)
    groups = {}
    for i, h in enumerate(hams):
        groups.setdefault(bool(h.needs_complex) if split_dtypes else True, []).append(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batches = [(idx, BatchedLanczos([hams[i] for i in idx], max_steps=max_steps, complex_=c))
               for c, idx in sorted(groups.items(), key=lambda kv: -len(kv[1]))]
    schedules = [B.autotune() if autotune else None for _, B in batches]
    torch.cuda.synchronize()
    t_create = time.perf_counter() - t0
    results = [None] * len(batches)
    errors = []

    streams = [torch.cuda.Stream() for _ in batches]
    for _, B in batches:          # graphs captured up front, one batch at a time
        B.capture()

    def run(k):
        idx, B = batches[k]
        try:
            # each thread feeds its batch from its own stream: no legacy-stream
            # work while another batch's graph runs
            with torch.cuda.stream(streams[k]):
                results[k] = B.ground_states(tol=tol, check_every=check_every, seeds=[seeds[i] for i in idx],
                                             sync_device=False, check_min=check_min)
        except Exception as ex:   # re-raised on the calling thread
            errors.append(ex)

    def once():
        if concurrent and len(batches) > 1:
            th = [threading.Thread(target=run, args=(k,)) for k in range(len(batches))]
            for t in th:
                t.start()
            for t in th:
                t.join()
        else:
            for k in range(len(batches)):
                run(k)
        if errors:
            raise errors[0]

    if warmup:
        once()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    once()
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    e0 = np.zeros(len(hams))
    steps = np.zeros(len(hams), dtype=np.int32)
    resid = np.zeros(len(hams))
    for (idx, B), r in zip(batches, results):
        e0[idx] = r.e0
        steps[idx] = r.steps
        resid[idx] = B.residuals()          # after the timed region
        B.close()
    return BatchResult(e0=e0, steps=steps, seconds=seconds, schedule=schedules, residuals=resid), t_create


# ---------------------------------------------------------------------------
# Momentum classes of a sweep: sectors related by a point-group symmetry of
# the lattice that the Hamiltonian's term list also has hold unitarily
# equivalent blocks (the same spectra), so a sweep may solve one sector per
# class.  No reference counterpart: Algebra/hilbert.py:546-652 iterates every
# momentum; this is an option of the sweep, never implied.
# ---------------------------------------------------------------------------

def _point_group(lattice):
    """(matrix, site permutation) of the lattice's point-group elements that
    map the Bravais lattice onto itself: x -> -x on chains; the 8 elements of
    D4 on square lattices with lx == ly, the 4 reflections otherwise.  Other
    lattices: none (a class is a single sector)."""
    kind = getattr(lattice, "kind", "")
    lx, ly = int(getattr(lattice, "lx", 0)), int(getattr(lattice, "ly", 1))
    if getattr(lattice, "n_basis", 1) != 1 or kind not in ("chain", "square"):
        return []
    if kind == "chain" or ly == 1:
        mats = [np.array([[1, 0], [0, 1]]), np.array([[-1, 0], [0, 1]])]
    else:
        mats = [np.array([[a, 0], [0, b]]) for a in (1, -1) for b in (1, -1)]
        if lx == ly:
            mats += [np.array([[0, a], [b, 0]]) for a in (1, -1) for b in (1, -1)]
    out = []
    for M in mats:
        perm = np.empty(lattice.ns, dtype=np.int64)
        for s in range(lattice.ns):
            x, y, _ = lattice.get_coordinates(s)
            xx, yy = M @ np.array([x, y])
            perm[s] = lattice.site_index(int(xx) % lx, int(yy) % ly)
        out.append((M, perm))
    return out


def _term_multiset(codes, sites, coeffs, perm=None):
    """Canonical multiset of a term list, optionally with permuted sites:
    two-site XX / YY / ZZ terms (codes 20-22) take an unordered pair, and
    S-_i S+_j (code 32) is written as S+_j S-_i (code 31)."""
    out = {}
    for c, st, v in zip(codes, sites, coeffs):
        c = int(c)
        a, b = int(st[0]), int(st[1])
        if perm is not None:
            a = int(perm[a]) if a >= 0 else a
            b = int(perm[b]) if b >= 0 else b
        if c in (20, 21, 22) and b >= 0:
            a, b = min(a, b), max(a, b)
        elif c == 32 and b >= 0:
            c, a, b = 31, b, a
        v = complex(v)
        key = (c, a, b, round(v.real, 12), round(v.imag, 12))
        out[key] = out.get(key, 0) + 1
    return out


def momentum_classes(lattice, ham, keys):
    """Group the momentum keys ``keys`` ((kx,) or (kx, ky)) of a sweep into
    classes of sectors whose blocks are unitarily equivalent: k ~ M k for
    every point-group element M of the lattice (``_point_group``) under
    which the Hamiltonian's instruction list is invariant.  Returns
    [(representative key, [member keys])] in the order of first appearance."""
    sites, coeffs, codes = ham.instruction_arrays()
    ref = _term_multiset(codes, sites, coeffs)
    mats = [M for M, perm in _point_group(lattice) if _term_multiset(codes, sites, coeffs, perm) == ref]
    lx, ly = int(getattr(lattice, "lx", 0)), int(getattr(lattice, "ly", 1))
    keys = [tuple(k) for k in keys]
    index = {k: i for i, k in enumerate(keys)}
    cls = [-1] * len(keys)
    out = []
    for i, k in enumerate(keys):
        if cls[i] >= 0:
            continue
        members = []
        for M in mats or [np.eye(2, dtype=int)]:
            kk = M @ np.array([k[0], k[1] if len(k) > 1 else 0])
            img = (int(kk[0]) % lx,) + ((int(kk[1]) % ly,) if len(k) > 1 else ())
            j = index.get(img)
            if j is not None and cls[j] < 0:
                cls[j] = len(out)
                members.append(img)
        out.append((k, sorted(members, key=lambda m: index[m])))
    return out
