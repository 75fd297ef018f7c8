#!/usr/bin/env python
"""Benchmark: symmetric-sector H.v matvec throughput and time-to-E0 on one MI355X.

Workload (default, --config square6x6): J1-J2 Heisenberg 6x6 square
lattice, PBC, J1=1, J2=0.5 (reference constructor J1J2Model(J1=-1,
J2=-0.5)), Sz=0, 2D translations k=(0,0), spin flip +1: N = 126,056,625
reduced states, float64.  A "step" is one matrix-free H.v on a seeded N(0,1)
vector already resident in HBM.  After the timed steps: the assembled
operator's product (same H.v through a table), a Lanczos ground-state solve
(three-term recurrence: one stored vector is 1 GB) to the 1e-10 residual
test with its time-to-E0, the CPU baseline (the oracle's C restatement of
the reference push-form kernel on column windows strided across the basis),
the J1-J2 chain L=32 (k=0, flip+, N = 9,393,119) as a secondary field, the
batched workload (XXZ L=24 momentum sectors + 256 random-field L=20
realisations: batched against sector by sector) and a one-sector 6x6
coupling sweep (k=(0,0) at J2 in {0.4, 0.5, 0.6} against three solves).

Other workloads: --config chain32 (the L=32 chain alone, Lanczos with full
reorthogonalisation against the exact Majumdar-Ghosh E0 = -3L/8), --config
batched (XXZ L=24 all momentum sectors + 256 random-field L=20
realisations, one batched Lanczos against a sector-by-sector sweep),
--config grad (Kitaev-Heisenberg honeycomb 24 sites: dE0/dJ and Krylov
gradients), --config sweep (6x6 momentum sectors x J2 coupling sweeps).

Prints one JSON line.  Only one GPU is used; --gpus N > 1 exits.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0

# HBM bytes per matvec dispatch from rocprofv3 FETCH_SIZE + WRITE_SIZE
# (KiB per dispatch, FETCH scaled x2.0 and WRITE x1.0 by a 1 GiB device copy
# measured in the same runs); filled from the committed profiles of the
# current kernels (profiles/matvec_kernel_square6x6_j1j2.md,
# profiles/matvec_kernel_chain32_j1j2.md)
TRAFFIC = {
    # matvec_kernel_wide8<false,2,1,true,4,true,true> (lean, rank gather): 244,586,596 KiB x 2.0 + 984,817 KiB x 1.0
    "square6x6_rank_gather": 244586596 * 1024 * 2 + 984817 * 1024,
    # matvec_kernel_wide8<false,2,1,true,4,false,true> (lean): 335,259,233 KiB x 2.0 + 4,924,087 KiB x 1.0
    "square6x6": 335259233 * 1024 * 2 + 4924087 * 1024,
    # matvec_kernel<false,1,1,true,4,true,true> (lean, rank gather): 5,143,062 KiB x 2.0 + 73,384 KiB x 1.0
    "chain32_rank_gather": 5143062 * 1024 * 2 + 73384 * 1024,
    # matvec_kernel<false,1,1,true,4,false,true> (lean): 4,242,720 KiB x 2.0 + 220,151 KiB x 1.0
    "chain32": 4242720 * 1024 * 2 + 220151 * 1024,
}

def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--L", type=int, default=32)
    p.add_argument("--J2", type=float, default=0.5)
    p.add_argument("--lanczos-steps", type=int, default=300)
    p.add_argument("--no-lanczos", action="store_true")
    p.add_argument("--no-assembled", action="store_true", help="skip the assembled-operator timing")
    p.add_argument("--cpu-sample", type=int, default=2000,
                   help="columns timed on one thread for the CPU baseline, in windows strided across the basis "
                        "(0: skip)")
    p.add_argument("--no-secondary", action="store_true", help="square6x6: skip the L=32 chain field")
    p.add_argument("--no-rank-gather", action="store_true",
                   help="time only the default dense gathers (no rank-gather variant for the value)")
    p.add_argument("--no-batched-field", action="store_true",
                   help="square6x6: skip the batched field (XXZ L=24 sectors + L=20 realisations)")
    p.add_argument("--no-sweep-field", action="store_true",
                   help="square6x6: skip the coupling-sweep field (6x6 k=(0,0) at three J2 values)")
    p.add_argument("--no-ritz", action="store_true", help="skip the Ritz-vector residual check (three-term runs)")
    p.add_argument("--quiet", action="store_true")
    p.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                   help="also run a plain device copy of MIB MiB (profiler counter calibration)")
    p.add_argument("--config", default="square6x6",
                   choices=["square6x6", "chain32", "batched", "grad", "sweep", "block", "large", "plain"],
                   help="square6x6 (default; alias large): J1-J2 6x6 square k=(0,0) flip+ matvec + Lanczos (no "
                        "reorthogonalisation) + the L=32 chain as a secondary field; chain32 (alias plain): L=32 "
                        "J1-J2 chain matvec + Lanczos with full reorthogonalisation; batched: XXZ L=24 all momentum "
                        "sectors + random-field L=20 realisations, batched vs sector-by-sector; grad: "
                        "Kitaev-Heisenberg honeycomb 24 sites, dE0/dJ and Krylov time-evolution gradients; sweep: "
                        "J1-J2 6x6 square, every momentum sector (flip+) at J2 in {0.4, 0.5, 0.6}, one coupling "
                        "sweep per sector (several minutes)")
    p.add_argument("--realisations", type=int, default=256)
    p.add_argument("--check-min", type=int, default=None,
                   help="batched/sweep: 0 = stopping test every 10 steps; n >= 1 = adaptive, >= n steps apart; "
                        "default: the library's choice (adaptive only for a few large problems)")
    p.add_argument("--sectors", type=int, default=36, help="sweep: momentum sectors to run (of 36)")
    p.add_argument("--classes", action="store_true",
                   help="sweep: solve one sector per point-group class of momenta (classes checked on the "
                        "Hamiltonian's term list); --sectors and --sector-start then count classes")
    p.add_argument("--sector-start", type=int, default=0, help="sweep: first sector (row-major (kx, ky) order)")
    p.add_argument("--baseline-sectors", type=int, default=1,
                   help="sweep: sectors also solved once per J2 value (the sequential baseline)")
    a = p.parse_args()
    a.config = {"large": "square6x6", "plain": "chain32"}.get(a.config, a.config)
    return a


def bench_batched(a):
    print(json.dumps(run_batched(a)), flush=True)


def run_batched(a):
    """XXZ chain L=24 (Jxy=1, Jz=0.7, Sz=0): all 24 momenta, reflection parity
    at k=0 and k=12 (26 sectors), plus random-field XXZ L=20 Sz=0 realisations
    with h_i ~ U[-5, 5] (seed 1000+r), all ground states by batched Lanczos
    (a real and a complex batch, overlapped), timed against a
    sector-by-sector sweep on the same GPU."""
    import torch
    from pyqusolver_amd import ChainLattice, get_u1_sym
    from pyqusolver_amd.batched import batched_ground_states, momentum_sectors, sector_by_sector
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import XXZ
    lat24 = ChainLattice(24)
    specs = momentum_sectors(24, reflection=True)
    t0 = time.perf_counter()
    spaces = [HilbertSpace(lattice=lat24, sym_gen=s, global_syms=[get_u1_sym(lat24, 12)]) for s in specs]
    hams = [XXZ(lat24, hilbert_space=h, jxy=1.0, jz=0.7) for h in spaces]
    lat20 = ChainLattice(20)
    hs20 = HilbertSpace(lattice=lat20, global_syms=[get_u1_sym(lat20, 10)])
    for r in range(a.realisations):
        h = np.random.default_rng(1000 + r).uniform(-5.0, 5.0, 20)
        hams.append(XXZ(lat20, hilbert_space=hs20, jxy=1.0, jz=0.7, hz=h))
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    rows = int(sum(h.nh for h in hams))
    log(f"{len(hams)} problems, {rows} rows, setup {t_setup:.2f}s", a.quiet)
    # real and complex problems in two batches that overlap on the device;
    # one warm-up sweep (graph capture, code objects) before the timed one
    res, t_create = batched_ground_states(hams, tol=1e-10, max_steps=a.lanczos_steps, warmup=True,
                                          check_min=a.check_min)
    log(f"batched: {res.seconds:.3f}s (+{t_create:.3f}s batch creation: tables, diagonals), "
        f"steps max {res.steps.max()}", a.quiet)
    seq = sector_by_sector(hams, tol=1e-10, max_steps=a.lanczos_steps)
    log(f"sector-by-sector: {seq.seconds:.3f}s", a.quiet)
    dev = np.abs(res.e0 - seq.e0) / np.maximum(1.0, np.abs(seq.e0))
    # accuracy witness independent of the solver: sectors k and L-k hold
    # complex-conjugate blocks with equal spectra (time reversal)
    ek = {s["translation"]: float(e) for s, e in zip(specs, res.e0[:len(specs)]) if "reflection" not in s}
    pair_dev = max(abs(ek[k] - ek[24 - k]) / abs(ek[k]) for k in range(1, 12))
    out = {
        "metric": "H·v matvecs/s + HBM GB/s frac; time-to-E0 at 1e-10 (batched/plain)",
        "value": float(res.steps.sum() / res.seconds),
        "unit": "matvec/s (problem-steps per second, batched)",
        "n_gpus": 1, "steps": int(res.steps.max()), "warmup": 1,
        "ms_per_step": float(1000.0 * res.seconds / max(1, res.steps.max())),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+c128", "data": "synthetic",
        "config": {"workload": "XXZ chain L=24 Jxy=1 Jz=0.7 Sz=0, 26 (k,R) sectors + "
                               f"{a.realisations} random-field L=20 realisations, one batched Lanczos",
                   "problems": len(hams), "rows": rows},
        "time_to_e0_batched_s": res.seconds, "time_to_e0_sector_by_sector_s": seq.seconds,
        "batch_creation_s": t_create,
        "speedup": seq.seconds / res.seconds,
        "speedup_including_creation": seq.seconds / (res.seconds + t_create),
        "sector_by_sector": "one device Lanczos per problem, each operator assembled for its own solve",
        "e0_sectors_L24": [float(v) for v in res.e0[:len(specs)]],
        "residuals_sectors_L24": [float(v) for v in res.residuals[:len(specs)]],
        "residual_bound_max": float(res.residuals.max()),
        "rel_error_bound_max": float(np.max(res.residuals / np.maximum(1.0, np.abs(res.e0)))),
        "max_rel_dev_batched_vs_sequential": float(dev.max()),
        "schedule": res.schedule,
        "max_rel_dev_E0_k_vs_minus_k": pair_dev,
        "stopping": "per problem: Ritz residual |beta_m s_m| <= 1e-10 max(1,|E0|), checked every 10 steps",
    }
    del hams, spaces
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out


def bench_grad(a):
    """Kitaev-Heisenberg honeycomb, 3x4 unit cells (24 sites, PBC), K=-1,
    J=0.2, h=0.05 along [111], k=(0,0), no U(1): N = 1,398,500, complex128.
    dE0/d(K,J,h) by backprop through the Lanczos recurrence, by
    Hellmann-Feynman and by central differences (step 1e-5); gradient of
    |<psi0|exp(-iHt)|phi>|^2 (t=1, 30-step Krylov, phi seeded) against
    central differences.  value = Krylov-objective gradients per second."""
    import torch
    from pyqusolver_amd import HoneycombLattice
    from pyqusolver_amd.autograd import (DifferentiableHamiltonian, central_difference, hellmann_feynman,
                                         krylov_overlap, lanczos_energy)
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.lanczos import lanczos_ground_state, start_vector
    from pyqusolver_amd.models import HeisenbergKitaev
    lat = HoneycombLattice(3, 4)
    hs = HilbertSpace(lattice=lat, sym_gen={"translation": {"kx": 0, "ky": 0}})
    h = 0.05 / np.sqrt(3.0)
    ham = HeisenbergKitaev(lat, K=-1.0, J=0.2, hilbert_space=hs, hx=h, hy=h, hz=h, split_couplings=True)
    H = DifferentiableHamiltonian(ham)
    n = hs.nh
    log(f"honeycomb 3x4 N={n}", a.quiet)

    def sync_time(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    res, t_gs = sync_time(lambda: lanczos_ground_state(ham, tol=1e-12, maxiter=a.lanczos_steps, seed=0, reorth=True,
                                                       compute_eigvec=True))
    e0 = float(res.eigenvalues[0])
    psi0 = res.eigenvectors / torch.linalg.vector_norm(res.eigenvectors)
    log(f"E0={e0:.12f} steps={res.steps} {t_gs:.2f}s", a.quiet)
    g_hf, t_hf = sync_time(lambda: hellmann_feynman(H, psi0))
    m_e = min(int(res.steps) + 40, a.lanczos_steps)
    v0 = start_vector(n, True, seed=0)
    J = H.J0.clone().requires_grad_(True)

    def lanczos_grad():
        J.grad = None
        e = lanczos_energy(H, J, v0, m_e)
        e.backward()
        return e.item(), J.grad.clone()
    (e_l, g_l), t_l = sync_time(lanczos_grad)
    g_fd_e = central_difference(lambda Jp: lanczos_energy(H, Jp, v0, m_e), J.detach(), h=1e-5)
    log(f"dE0/dJ lanczos-AD {g_l.numpy()} HF {g_hf.numpy()} FD {g_fd_e.numpy()}", a.quiet)

    g = torch.Generator(device="cuda").manual_seed(1234)
    phi = torch.randn(n, dtype=torch.complex128, device="cuda", generator=g)
    phi = phi / torch.linalg.vector_norm(phi)

    def krylov_grad():
        J.grad = None
        f = krylov_overlap(H, J, psi0, phi, t=1.0, m=30)
        f.backward()
        return f.item(), J.grad.clone()
    krylov_grad()                      # warm-up
    reps = max(1, a.steps)
    (f_val, g_k), t_k = sync_time(lambda: [krylov_grad() for _ in range(reps)][-1])
    g_fd_k = central_difference(lambda Jp: krylov_overlap(H, Jp, psi0, phi, t=1.0, m=30), J.detach(), h=1e-5)
    log(f"Krylov f={f_val:.6e} AD {g_k.numpy()} FD {g_fd_k.numpy()} {t_k / reps * 1e3:.1f} ms/grad", a.quiet)

    def rel(x, y):
        return float(torch.max(torch.abs(x - y)) / torch.max(torch.abs(y)))
    out = {
        "metric": "H·v matvecs/s + HBM GB/s frac; time-to-E0 at 1e-10 (batched/plain)",
        "value": reps / t_k, "unit": "Krylov-objective gradients/s (30-step Krylov, forward+backward)",
        "n_gpus": 1, "steps": reps, "warmup": 1, "ms_per_step": 1000.0 * t_k / reps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128", "data": "synthetic",
        "config": {"workload": "Kitaev-Heisenberg honeycomb 3x4 cells (24 sites) PBC K=-1 J=0.2 h=0.05 [111], "
                               "k=(0,0), no U(1)", "N": n},
        "E0": e0, "lanczos_steps": int(res.steps), "time_to_E0_s": t_gs,
        "dE0_dJ": {"couplings": ["K", "J", "h"], "lanczos_autograd": g_l.tolist(), "hellmann_feynman": g_hf.tolist(),
                   "central_difference_1e-5": g_fd_e.tolist(), "lanczos_steps": m_e,
                   "max_rel_dev_ad_vs_hf": rel(g_l, g_hf), "max_rel_dev_ad_vs_fd": rel(g_l, g_fd_e),
                   "seconds_autograd": t_l, "seconds_hf": t_hf},
        "krylov_overlap": {"t": 1.0, "m": 30, "value": f_val, "autograd": g_k.tolist(),
                           "central_difference_1e-5": g_fd_k.tolist(), "max_rel_dev": rel(g_k, g_fd_k)},
    }
    print(json.dumps(out), flush=True)


def bench_sweep(a):
    print(json.dumps(run_sweep(a)), flush=True)


def run_sweep(a):
    """J1-J2 6x6 square, Sz=0, spin flip +1, momentum sectors (kx, ky) in
    row-major order, each at J2 in {0.4, 0.5, 0.6}: one coupling sweep per
    sector (its coupling-tagged table built once, streamed once per step for
    the three J2 columns; E0 of each column to the 1e-10 residual test).
    The sequential baseline solves the first ``--baseline-sectors`` sectors
    once per J2 value with an assembled operator each, the way a caller of
    diagonalize would.  value = (sector, J2) ground states per second.

    ``--classes``: sectors whose momenta are related by a point-group
    element of the lattice that the Hamiltonian's term list also has (D4 for
    the 6x6 J1-J2 model: 10 classes of the 36 momenta) hold blocks with the
    same spectrum; only the first sector of each class is solved and its
    results stand for the class (the sectors listed are then the classes'
    first members)."""
    import torch
    from pyqusolver_amd import SquareLattice, get_u1_sym
    from pyqusolver_amd.batched import CouplingSweep, momentum_classes
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import J1J2Model
    lat = SquareLattice(6, 6)
    j2s = [0.4, 0.5, 0.6]
    scales = [[1.0, j / 0.5] for j in j2s]
    from pyqusolver_amd.hilbert import momentum_key, momentum_symmetry_types, sector_specs
    ks = [momentum_key(lat, sec) for sec, _ in sector_specs(momentum_symmetry_types(lat), lat)]
    classes = None
    if a.classes:
        hs0 = HilbertSpace(lattice=lat, sym_gen={"translation": {"kx": 0, "ky": 0}, "parity": {"x": 1}},
                           global_syms=[get_u1_sym(lat, 18)])
        classes = momentum_classes(lat, J1J2Model(lat, J1=-1.0, J2=-0.5, hilbert_space=hs0), ks)
        del hs0
        members_of = dict(classes)
        ks = [rep for rep, _ in classes]
    ks = ks[a.sector_start: a.sector_start + max(1, a.sectors)]
    results, t_total, t_tables, steps_max = [], 0.0, 0.0, 0
    for kx, ky in ks:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        hs = HilbertSpace(lattice=lat, sym_gen={"translation": {"kx": kx, "ky": ky}, "parity": {"x": 1}},
                          global_syms=[get_u1_sym(lat, 18)])
        H = J1J2Model(lat, J1=-1.0, J2=-0.5, hilbert_space=hs)
        n_sec = hs.nh
        t1 = time.perf_counter()
        S = CouplingSweep([H], scales, max_steps=a.lanczos_steps, max_gb=160.0)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        res = S.ground_states(tol=1e-10, check_min=a.check_min)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        resid = S.residuals()
        S.close()
        del H, hs
        t_total += t3 - t0
        t_tables += t2 - t1
        steps_max = max(steps_max, int(res.steps.max()))
        results.append({"k": [kx, ky], "N": int(n_sec), "e0": [float(v) for v in res.e0],
                        "steps": [int(v) for v in res.steps], "residuals": [float(v) for v in resid],
                        "seconds": round(t3 - t0, 3)})
        if classes is not None:
            results[-1]["class_members"] = [list(m) for m in members_of[(kx, ky)]]
        log(f"sector ({kx},{ky}): E0 {' '.join(f'{v:.12f}' for v in res.e0)} steps {list(res.steps)} "
            f"{t3 - t0:.2f}s (table {t2 - t1:.2f}s, solve {t3 - t2:.2f}s)", a.quiet)
    base = None
    nb = min(a.baseline_sectors, len(ks))
    if nb > 0:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dev = 0.0
        for i, (kx, ky) in enumerate(ks[:nb]):
            hs = HilbertSpace(lattice=lat, sym_gen={"translation": {"kx": kx, "ky": ky}, "parity": {"x": 1}},
                              global_syms=[get_u1_sym(lat, 18)])
            for r, j2 in enumerate(j2s):
                H = J1J2Model(lat, J1=-1.0, J2=-j2, hilbert_space=hs)
                e = H.diagonalize(method="lanczos", k=1, tol=1e-10, maxiter=a.lanczos_steps, seed=i, reorth=False,
                                  store_basis=False, assemble=True)[0]
                dev = max(dev, abs(e - results[i]["e0"][r]) / abs(e))
                del H
            del hs
        torch.cuda.synchronize()
        t_seq = time.perf_counter() - t0
        t_sweep = sum(r["seconds"] for r in results[:nb])
        base = {"sectors": nb, "seconds_sequential": round(t_seq, 3), "seconds_sweep": round(t_sweep, 3),
                "speedup": t_seq / t_sweep, "max_rel_dev_sweep_vs_sequential": dev}
        log(f"sequential ({nb} sector(s) x 3 J2): {t_seq:.2f}s vs sweep {t_sweep:.2f}s", a.quiet)
    n_gs = (sum(len(members_of[tuple(k)]) for k in ks) if classes is not None else len(ks)) * len(j2s)
    out = {
        "metric": "H·v matvecs/s + HBM GB/s frac; time-to-E0 at 1e-10 (batched/plain)",
        "value": n_gs / t_total, "unit": "ground states/s ((sector, J2) pairs, 6x6 J1-J2, E0 at 1e-10)",
        "n_gpus": 1, "steps": steps_max, "warmup": 0, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+c128", "data": "synthetic",
        "config": {"workload": "J1-J2 square 6x6 PBC J1=1 Sz=0 flip+, momentum sectors x J2 in {0.4, 0.5, 0.6}, "
                               "one coupling sweep per sector" +
                               (", one sector per point-group class of momenta" if classes is not None else ""),
                   "sectors": len(ks), "first_sector": a.sector_start,
                   "sectors_covered": (sum(len(members_of[tuple(k)]) for k in ks) if classes is not None
                                       else len(ks))},
        "seconds_total": t_total, "seconds_tables": t_tables, "sectors": results, "sequential_baseline": base,
        "stopping": "per (sector, J2): Ritz residual |beta_m s_m| <= 1e-10 max(1,|E0|), tested at most 10 steps "
                    "apart, where the residual's decay predicts the crossing",
    }
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out


def bench_block(a):
    """Block Lanczos on the J1-J2 chain L=32 (Sz=0, k=0, flip+, N =
    9,393,119): the 4 lowest eigenvalues with blocks of 4 and of 8 vectors
    (time to the 1e-10 residual test on every wanted pair), and the
    FP64-MFMA contractions at a full-reorthogonalisation size of that run
    (256 stored vectors against a block of 8): C = V^H W and W -= V C timed
    with HIP events, FLOP and byte rates against the MI355X peaks."""
    import ctypes as C
    import torch
    from pyqusolver_amd import ChainLattice, _lib, get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.lanczos import block_lanczos, lanczos_ground_state
    from pyqusolver_amd.models import J1J2Model
    lat = ChainLattice(32)
    hs = HilbertSpace(lattice=lat, sym_gen={"translation": 0, "parity": {"x": 1}}, global_syms=[get_u1_sym(lat, 16)])
    H = J1J2Model(lat, J1=-1.0, J2=-0.5, hilbert_space=hs)
    n = hs.nh
    H.assemble(True)
    runs = {}
    for p in (4, 8):
        block_lanczos(H, k=4, block_size=p, tol=1e-10, maxiter=3)        # warm-up (code objects)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = block_lanczos(H, k=4, block_size=p, tol=1e-10, maxiter=100)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        runs[f"block_{p}"] = {"eigenvalues": [float(v) for v in r.eigenvalues], "converged": r.converged,
                              "block_steps": r.block_steps, "products": r.steps, "residual_bound": r.residual_bound,
                              "seconds": round(dt, 3)}
        log(f"block {p}: E {r.eigenvalues} steps {r.block_steps} ({r.steps} products) {dt:.2f}s", a.quiet)
    # single-vector reference: lowest value, full reorthogonalisation
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r1 = lanczos_ground_state(H, k=4, tol=1e-10, maxiter=300, seed=0, reorth=True, store_basis=True)
    torch.cuda.synchronize()
    runs["single_vector_k4"] = {"eigenvalues": [float(v) for v in r1.eigenvalues], "steps": int(r1.steps),
                                "seconds": round(time.perf_counter() - t0, 3)}
    H.assemble(False)
    # contraction kernels: 256 stored vectors, a block of 8
    nv, ncol, length = 256, 8, n
    ld = length + (length & 1)
    V = torch.randn((nv, ld), dtype=torch.float64, device="cuda")
    W = torch.randn((ncol, ld), dtype=torch.float64, device="cuda")
    nb = C.c_int64()
    _lib.check(_lib.lib().qes_block_dot_ws_bytes(nv, length, C.byref(nb)))
    ws = torch.empty(nb.value // 8, dtype=torch.float64, device="cuda")
    Cm = torch.empty(nv * ncol, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream()
    L_ = _lib.lib()

    def dot():
        _lib.check(L_.qes_block_dot(_lib.dptr(V), ld, nv, _lib.dptr(W), ld, ncol, length, 0, _lib.dptr(Cm),
                                    _lib.dptr(ws), nb.value, C.c_void_p(st.cuda_stream)), "qes_block_dot")

    def comb():
        _lib.check(L_.qes_block_combine(_lib.dptr(V), ld, nv, _lib.dptr(Cm), 1.0, _lib.dptr(W), ld, ncol, length, 0,
                                        C.c_void_p(st.cuda_stream)), "qes_block_combine")
    Cm.normal_()
    Cm.mul_(1e-3)
    kern = {}
    for name, fn in (("block_dot", dot), ("block_combine", comb)):
        fn()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(max(3, a.steps))]
        torch.cuda.synchronize()
        for e0, e1 in ev:
            e0.record(st)
            fn()
            e1.record(st)
        torch.cuda.synchronize()
        ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in ev]))
        flops = 2.0 * nv * ncol * length
        byts = (nv + ncol) * length * 8 + (ncol * length * 8 if name == "block_combine" else 0)
        kern[name] = {"ms": ms, "tflops": flops / ms / 1e9, "frac_fp64_mfma_peak": flops / ms / 1e9 / 78.6,
                      "GBs": byts / ms / 1e6, "frac_hbm": byts / ms / 1e6 / HBM_PEAK_GBS, "flops": flops, "bytes": byts}
        log(f"{name}: {ms:.3f} ms, {flops / ms / 1e9:.2f} TFLOP/s, {byts / ms / 1e6:.0f} GB/s", a.quiet)
    out = {
        "metric": "H·v matvecs/s + HBM GB/s frac; time-to-E0 at 1e-10 (batched/plain)",
        "value": runs["block_8"]["products"] / runs["block_8"]["seconds"],
        "unit": "matvec/s (block Lanczos products per second, 4 lowest eigenvalues, blocks of 8)",
        "n_gpus": 1, "steps": runs["block_8"]["block_steps"], "warmup": 1,
        "ms_per_step": 1000.0 * runs["block_8"]["seconds"] / max(1, runs["block_8"]["block_steps"]),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "block Lanczos, J1-J2 chain L=32 PBC J1=1 J2=0.5 Sz=0 k=0 flip+, 4 lowest", "N": n},
        "runs": runs,
        "contractions": {"shape": {"stored_vectors": nv, "block": ncol, "length": length},
                         "kernels": kern,
                         "roofline": {"bound": "hbm", "peak_hbm_GBs": HBM_PEAK_GBS, "peak_fp64_mfma_TFLOPs": 78.6,
                                      "intensity_flop_per_byte": 2.0 * ncol / 8.0}},
    }
    print(json.dumps(out), flush=True)


def log(msg, quiet=False):
    if not quiet:
        print(msg, file=sys.stderr, flush=True)


def _cpu_model() -> str:
    """Host CPU model name (the lscpu "Model name" line, from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.lower().startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, J2, reps, norms, sample, n):
    """CPU port of the reference push-form kernel (oracle/qes_oracle.c),
    extrapolated to matvecs/s from bounded samples: column windows at
    positions i*N/W (i = 0..W-1, strided across the whole sorted basis, so
    low and high representatives both count), first ``sample`` columns in
    all on one thread (the reference matvec_fun is single-threaded,
    special_operator.py:656-663), then the windows that follow them on every
    host core the box gives this process (at most 16), each thread with its
    own output buffer like the reference's multithreaded kernel
    (matrix_builder.py:1380-1470).  The multi-core rate is the reported
    value."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    from oracle import qes_oracle as O
    from oracle import qes_oracle_c as OC
    lat = O.OLattice(wl["kind"], wl["lx"], wl["ly"])
    grp = O.Group(lat, [O.Generator(g, k) for g, k in wl["oracle_gens"]], n_up=wl["n_up"])
    t = O.terms_j1j2(lat, -1.0, -J2)
    c, s, co = t.arrays()
    x = np.random.default_rng(0).standard_normal(len(reps))
    cg = OC.CGroup(grp)
    nwin = 32
    starts = [(i * n) // nwin for i in range(nwin)]
    w1 = max(1, sample // nwin)
    r1 = [(st, min(n, st + w1)) for st in starts]
    y = np.zeros_like(x)
    t0 = time.perf_counter()
    for r in r1:
        OC.push_matvec(grp, reps, norms, c, s, co, False, x, k_range=r, cgroup=cg, y=y)
    dt = time.perf_counter() - t0
    cols_1 = sum(b - a for a, b in r1) / dt
    nth = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1), os.cpu_count() or 1))
    wm = max(1, (sample // 2) * nth // nwin)
    rm = [(min(n, st + w1), min(n, st + w1 + wm)) for st in starts]
    per_thread = [rm[i::nth] for i in range(nth)]
    per_thread = [p for p in per_thread if p]
    ys = [np.zeros_like(x) for _ in per_thread]

    def work(i):
        for r in per_thread[i]:
            OC.push_matvec(grp, reps, norms, c, s, co, False, x, k_range=r, cgroup=cg, y=ys[i])
    t0 = time.perf_counter()
    # ctypes releases the GIL inside the C kernel, so the threads run in parallel
    with ThreadPoolExecutor(max_workers=len(per_thread)) as ex:
        list(ex.map(work, range(len(per_thread))))
    dtm = time.perf_counter() - t0
    cols_m = sum(b - a for a, b in rm) / dtm
    return {"value": cols_m / n, "unit": "matvec/s", "cores": len(per_thread), "kind": "port",
            "single_thread_value": cols_1 / n, "cpu_model": _cpu_model(), "host_cpus": os.cpu_count(),
            "sample_positions": starts,
            "sample": f"push-form columns of the same {wl['name']} sector in {nwin} windows at i*N/{nwin} "
                      f"(i = 0..{nwin - 1}): {w1} columns per window on 1 thread ({dt:.1f} s, {cols_1:.0f} "
                      f"columns/s), then the next {wm} columns per window on {len(per_thread)} threads "
                      f"({dtm:.1f} s wall, {cols_m:.0f} columns/s), extrapolated to N={n}; reference group action "
                      f"per element and binary search in place of the 2^{wl['ns']}-entry dense map"}


WORKLOADS = {
    # J1-J2 chain, k=0, flip +1 (the plain-at-scale workload)
    "chain32": lambda a: {"kind": "chain", "lx": a.L, "ly": 1, "ns": a.L, "n_up": a.L // 2,
                        "sym": {"translation": 0, "parity": {"x": 1}},
                        "oracle_gens": [("tx", 0), ("Px", 1)],
                        "name": f"J1-J2 chain L={a.L} PBC J1=1 J2={a.J2} Sz=0 k=0 flip+",
                        "e_exact": (-3.0 * a.L / 8.0) if abs(a.J2 - 0.5) < 1e-15 else None,
                        "reorth": True},
    # J1-J2 6x6 square, k=(0,0), flip +1: no full reorthogonalisation (one
    # basis vector is 1 GB; 300 of them do not fit next to the operator)
    "square6x6": lambda a: {"kind": "square", "lx": 6, "ly": 6, "ns": 36, "n_up": 18,
                        "sym": {"translation": {"kx": 0, "ky": 0}, "parity": {"x": 1}},
                        "oracle_gens": [("tx", 0), ("ty", 0), ("Px", 1)],
                        "name": f"J1-J2 square 6x6 PBC J1=1 J2={a.J2} Sz=0 k=(0,0) flip+",
                        "e_exact": None, "reorth": False},
}


def measure_sector(a, config, steps, warmup, primary=True):
    """One sector workload: basis, E_off, timed matrix-free products (HIP
    events on the launch stream), the assembled product, host-buffer
    products (N < 5e7), the Lanczos solve and, for the primary workload,
    the CPU baseline.  Returns the fields of the JSON line."""
    import torch
    from pyqusolver_amd import ChainLattice, SquareLattice, get_u1_sym
    from pyqusolver_amd.hilbert import HilbertSpace
    from pyqusolver_amd.models import J1J2Model
    # the solver module (with scipy.linalg, ~0.12 s) is imported here, outside
    # the timed time-to-E0
    from pyqusolver_amd import lanczos as _lanczos  # noqa: F401

    wl = WORKLOADS[config](a)
    J2 = a.J2
    lat = ChainLattice(wl["lx"]) if wl["kind"] == "chain" else SquareLattice(wl["lx"], wl["ly"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hs = HilbertSpace(lattice=lat, sym_gen=wl["sym"], global_syms=[get_u1_sym(lat, wl["n_up"])])
    torch.cuda.synchronize()
    t_basis = time.perf_counter() - t0
    n = hs.nh
    H = J1J2Model(lat, J1=-1.0, J2=-J2, hilbert_space=hs)
    e_off = H.count_offdiag()
    log(f"{wl['name']}: basis N={n} built in {t_basis:.3f}s, E_off={e_off}", a.quiet)

    # seeded N(0,1) inputs resident on the device before the timed region
    nx = min(max(steps, 1), 10 if n < 5e7 else 2)
    xs = [torch.from_numpy(np.random.default_rng(i).standard_normal(n)).to("cuda") for i in range(nx)]
    y = torch.empty(n, dtype=torch.float64, device="cuda")
    for i in range(warmup):
        H.matvec_device(xs[i % nx], out=y)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for i in range(steps):
        evs[i][0].record(stream)
        H.matvec_device(xs[i % nx], out=y)
        evs[i][1].record(stream)
    stop.record(stream)
    torch.cuda.synchronize()
    total_ms = start.elapsed_time(stop)
    step_ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in evs]))
    ms_per_step = total_ms / steps
    import ctypes as C
    from pyqusolver_amd import _lib
    # the rank-gather variant (the library option rank_gather: gathers from a
    # vector indexed by the colex rank of every state of the filling, built
    # once per basis); used for the value where it builds, the default gathers
    # reported beside it
    default_gather = {"ms_per_step": ms_per_step, "matvec_per_s": 1000.0 / ms_per_step}
    rank_gather = None
    if primary and not a.no_rank_gather:
        try:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            H.set_option("rank_gather", 1)
            torch.cuda.synchronize()
            t_rg = time.perf_counter() - t0
        except _lib.QesError as ex:   # no rank map, or the vector does not fit: the default gathers stay
            rank_gather = {"error": str(ex)}
        else:
            for i in range(warmup):
                H.matvec_device(xs[i % nx], out=y)
            torch.cuda.synchronize()
            start.record(stream)
            for i in range(steps):
                evs[i][0].record(stream)
                H.matvec_device(xs[i % nx], out=y)
                evs[i][1].record(stream)
            stop.record(stream)
            torch.cuda.synchronize()
            ms_per_step = start.elapsed_time(stop) / steps
            step_ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in evs]))
            import math
            words = -(-math.comb(wl["ns"], wl["n_up"]) // 64)
            rank_gather = {"ms_per_step": ms_per_step, "build_s": round(t_rg, 4),
                           "vector_bytes": words * 64 * 8}
    # the dominant kernel alone (the product is a short prescale pass, then
    # the row kernel): its launch duration from HIP events on the same stream
    kev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    pev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    ws = torch.empty_like(y)
    op = H._device_op()
    for i in range(steps):
        x = xs[i % nx]
        pev[i][0].record(stream)
        _lib.check(_lib.lib().qes_matvec_ws(op, _lib.dptr(x), _lib.dptr(y), 1, 0, 0, _lib.dptr(ws), n * 8,
                                            C.c_void_p(stream.cuda_stream)), "qes_matvec_ws")
        pev[i][1].record(stream)
    torch.cuda.synchronize()
    del ws
    rg_on = rank_gather is not None and "error" not in rank_gather
    if rg_on:
        H.set_option("rank_gather", 0)   # frees the basis's vector before the table is built
    b_alg = n * (8 + 8 + 8 + 8) + e_off * 8
    prod_ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in pev]))
    log(f"matvec {ms_per_step:.3f} ms/step (events {step_ms:.3f} ms, product {prod_ms:.3f} ms), B_alg "
        f"{b_alg/1e9:.3f} GB, {b_alg / (prod_ms * 1e-3) / 1e9:.1f} GB/s", a.quiet)
    # assembled operator (table built once, streamed per matvec): same H.v,
    # reported beside the matrix-free value
    assembled = None
    if not a.no_assembled:
        free, _ = torch.cuda.mem_get_info()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        try:
            tb = H.assemble(True, max_gb=min(128.0, 0.4 * free / 1e9))
        except Exception as ex:   # table larger than the budget: reported, not required
            assembled = {"error": str(ex)}
        else:
            torch.cuda.synchronize()
            t_build = time.perf_counter() - t0
            for i in range(warmup):
                H.matvec_device(xs[i % nx], out=y)
            ev2 = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
            torch.cuda.synchronize()
            for i in range(steps):
                ev2[i][0].record(stream)
                H.matvec_device(xs[i % nx], out=y)
                ev2[i][1].record(stream)
            torch.cuda.synchronize()
            asm_ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in ev2]))
            info = H.assembly
            H.assemble(False)
            if info["panels"]:
                # column panels: the panel tables, the diagonal, x_a, y_a, the
                # gathers, and one more read and write of y per later pass
                b_asm = info["panel_bytes"] + n * 8 + e_off * 8 + n * 16 + (info["panels"] - 1) * n * 16
            else:
                b_asm = tb + e_off * 8 + n * 16
            assembled = {"ms_per_matvec": asm_ms, "matvec_per_s": 1000.0 / asm_ms, "table_bytes": tb,
                         "panels": info["panels"], "panel_cols": info["panel_cols"],
                         "panel_bytes": info["panel_bytes"],
                         "build_s": round(t_build, 4),
                         "achieved_alg_GBs": b_alg / (asm_ms * 1e-3) / 1e9,
                         "frac_alg": b_alg / (asm_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "bytes_streamed": b_asm,
                         "frac_streamed": b_asm / (asm_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "kernel": "sell_matvec_kernel"}
            log(f"assembled: table {tb/1e9:.2f} GB built in {t_build:.3f}s, {asm_ms:.3f} ms/matvec", a.quiet)

    # host-buffer path (matvec_fun on numpy arrays): PCIe copies in and out
    host_ms = None
    if n < 5e7:
        xh = xs[0].cpu().numpy()
        H.matvec_fun(xh)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            H.matvec_fun(xh)
        host_ms = (time.perf_counter() - t0) / 3 * 1e3
        log(f"matvec_fun on host numpy arrays: {host_ms:.2f} ms/call (PCIe-inclusive)", a.quiet)
    del xs

    if primary and a.calibrate_copy > 0:
        # plain copy of a known size: its FETCH_SIZE / WRITE_SIZE per dispatch
        # calibrate the profiler's byte counters
        nel = a.calibrate_copy * (1 << 20) // 8
        src = torch.ones(nel, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    lanczos = None
    if not a.no_lanczos:
        e_exact = wl["e_exact"]
        reorth = wl["reorth"]
        # time-to-E0 = tabulating H (when it fits) + the Lanczos solve
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        free, _ = torch.cuda.mem_get_info()
        try:
            H.assemble(True, max_gb=min(128.0, 0.4 * free / 1e9))
        except Exception:
            pass    # too large: the solve runs matrix-free
        torch.cuda.synchronize()
        t_asm = time.perf_counter() - t0
        e0 = H.diagonalize(method="lanczos", k=1, tol=1e-10, maxiter=a.lanczos_steps, seed=0, reorth=reorth,
                           store_basis=reorth, assemble=False)[0]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        H.assemble(False)
        r = H.last_lanczos
        target = e_exact if e_exact is not None else float(e0)
        hit = None
        for st_, e in r.history:
            if abs(e - target) <= 1e-10 * abs(target):
                hit = st_
                break
        lanczos = {"E0": float(e0), "E0_exact": e_exact,
                   "rel_err": (abs(e0 - e_exact) / abs(e_exact)) if e_exact else None,
                   "steps": int(r.steps), "steps_to_1e-10": hit,
                   "steps_to_1e-10_reference": "exact E0" if e_exact is not None else "final E0 (self)",
                   "residual_bound": r.residual_bound,
                   "rel_error_bound": r.residual_bound / max(1.0, abs(float(e0))),
                   "converged": bool(r.converged),
                   "stopping": "Ritz residual |beta_m s_m| <= 1e-10 max(1,|E0|), tested at most 10 steps apart, "
                               "where the residual's decay predicts the crossing; |E0 - lambda| <= residual_bound",
                   "time_to_e0_s": round(dt, 4),
                   "seconds_assembly": round(t_asm, 4), "seconds_solve": round(dt - t_asm, 4),
                   "assembled_operator": bool(r.assembled),
                   "reorth": (f"full FP64 classical Gram-Schmidt against all stored vectors, DGKS second pass "
                              f"({r.second_passes} of {r.steps} steps)") if reorth else
                             "none (three-term recurrence)"}
        log(f"lanczos E0={e0:.14f} steps={r.steps} {dt:.2f}s residual<={r.residual_bound:.2e}", a.quiet)
        if not reorth and not a.no_ritz:
            # accuracy witness outside the timed solve: the ground-state Ritz
            # vector (second run of the recurrence, no stored basis) and its
            # true residual ||H psi - E0 psi|| / ||psi||
            from pyqusolver_amd.lanczos import lanczos_ground_state
            H.assemble(True, max_gb=min(128.0, 0.4 * torch.cuda.mem_get_info()[0] / 1e9))
            rv = lanczos_ground_state(H, k=1, tol=1e-10, maxiter=a.lanczos_steps, seed=0, reorth=False,
                                      store_basis=False, compute_eigvec=True)
            psi = rv.eigenvectors
            res = torch.linalg.vector_norm(H.matvec_device(psi) - float(rv.eigenvalues[0]) * psi)
            lanczos["ritz_vector_residual"] = float(res / torch.linalg.vector_norm(psi))
            H.assemble(False)
            del psi
            log(f"ritz vector residual {lanczos['ritz_vector_residual']:.2e}", a.quiet)

    cpu = None
    if primary and a.cpu_sample > 0:
        try:
            cpu = cpu_baseline(wl, J2, hs.representative_list, hs.normalization, a.cpu_sample, n)
        except Exception as ex:  # baseline is reported, never required
            cpu = {"value": None, "error": str(ex)}
    achieved = b_alg / (prod_ms * 1e-3) / 1e9
    out = {
        "value": 1000.0 / ms_per_step, "ms_per_step": ms_per_step,
        "config": {"workload": wl["name"], "N": n, "E_off": e_off, "group_order": hs.group.order,
                   "basis_build_s": round(t_basis, 4)},
        "host_buffer_ms_per_matvec": host_ms,
        "assembled": assembled,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": TRAFFIC.get(config + ("_rank_gather" if rg_on else "")),
                     "kernel": ("matvec_kernel_wide8" if wl["ns"] > 32 else "matvec_kernel") +
                               # the profiled workloads run the lean kernels (spin flip only)
                               ((" (lean, rank gather)" if rg_on else " (lean)") if config in ("square6x6", "chain32")
                                else (" (rank gather)" if rg_on else "")),
                     "kernel_ms": prod_ms, "bytes_alg": b_alg,
                     "bytes_alg_formula": "N*(8 rep + 8 norm + 8 x_a + 8 y_a) + E_off*8 x gathers",
                     "launches": ("rank_scatter_kernel (N*24 B read, N 8-byte stores at the representatives' ranks)"
                                  if rg_on else "prescale_kernel (N*24 B, ~0.4% of the time)") +
                                 " + the row kernel; kernel_ms is the product's HIP-event time on its stream"},
        "matvec_variant": "rank gather (option rank_gather=1)" if rg_on else "dense pre-scaled gathers (default)",
        "default_gather": default_gather,
        "rank_gather": rank_gather,
        "cpu_baseline": cpu,
        "lanczos": lanczos,
    }
    del H, hs
    torch.cuda.synchronize()
    return out


def bench_plain(a):
    import torch
    r = measure_sector(a, a.config, a.steps, a.warmup, primary=True)
    out = {
        "metric": "H·v matvecs/s + HBM GB/s frac; time-to-E0 at 1e-10 (batched/plain)",
        "value": r["value"],
        "unit": "matvec/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": r["config"],
        "roofline": r["roofline"],
        "matvec_variant": r["matvec_variant"],
        "default_gather": r["default_gather"],
        "rank_gather": r["rank_gather"],
        "cpu_baseline": r["cpu_baseline"],
        "assembled": r["assembled"],
        "lanczos": r["lanczos"],
        "time_to_e0_s": (r["lanczos"] or {}).get("time_to_e0_s"),
        "host_buffer_ms_per_matvec": r["host_buffer_ms_per_matvec"],
    }
    if a.config == "square6x6" and not a.no_secondary:
        torch.cuda.empty_cache()
        sec = measure_sector(a, "chain32", 10, 2, primary=False)
        out["secondary"] = {"workload": sec["config"]["workload"], "N": sec["config"]["N"],
                            "E_off": sec["config"]["E_off"], "matvec_per_s": sec["value"],
                            "ms_per_matvec": sec["ms_per_step"], "roofline": sec["roofline"],
                            "assembled": sec["assembled"], "host_buffer_ms_per_matvec": sec["host_buffer_ms_per_matvec"],
                            "lanczos": sec["lanczos"]}
    if a.config == "square6x6" and not a.no_batched_field:
        # the batched half of the metric: XXZ L=24 sectors + 256 random-field
        # L=20 realisations, batched against sector by sector
        torch.cuda.empty_cache()
        b = run_batched(a)
        out["batched"] = {k: b[k] for k in ("config", "time_to_e0_batched_s", "time_to_e0_sector_by_sector_s",
                                            "batch_creation_s", "speedup", "speedup_including_creation",
                                            "residual_bound_max", "rel_error_bound_max",
                                            "max_rel_dev_batched_vs_sequential", "max_rel_dev_E0_k_vs_minus_k")}
        out["batched"]["matvec_per_s"] = b["value"]
    if a.config == "square6x6" and not a.no_sweep_field:
        # the 6x6 coupling sweep on its first momentum sector, k = (0,0), at
        # J2 in {0.4, 0.5, 0.6}, against three separate solves of that sector
        torch.cuda.empty_cache()
        a.sector_start, a.sectors, a.baseline_sectors = 0, 1, 1
        w = run_sweep(a)
        out["sweep"] = {"workload": "J1-J2 square 6x6 k=(0,0) flip+ at J2 in {0.4, 0.5, 0.6}, one coupling sweep",
                        "sectors": w["sectors"], "sequential_baseline": w["sequential_baseline"],
                        "seconds_tables": w["seconds_tables"]}
    print(json.dumps(out), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); the library has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)
    if a.config == "batched":
        return bench_batched(a)
    if a.config == "grad":
        return bench_grad(a)
    if a.config == "sweep":
        return bench_sweep(a)
    if a.config == "block":
        return bench_block(a)
    return bench_plain(a)


if __name__ == "__main__":
    main()
