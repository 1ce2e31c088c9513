#!/usr/bin/env python3
"""Electron PIC step benchmark on one MI355X.

Workload (default): the plain GPU step at scale -- axisymmetric r-z grid
256 x 512 (dr = dz = 10 um, PHIp = 0.1), 4e6 electrons from a Gaussian X-ray
spot (sigma_r = 50 um) with the 70/30 400 +- 20 eV / 30 +- 10 eV energy mix,
uniform N2 at 2.5e25 m^-3, one collision check per push (dt = 1e-14 s,
10 field substeps), ionization secondaries on (capacity 2x), cell sort every
10 steps, FP32 particles, FP64 field solver.

A step = one push of every electron + its collision check + the field cycle.
value = particle-steps/s over the K timed steps (inputs resident in HBM).
roofline: the step against HBM, 48 algorithmic bytes per particle-step
(FP32 x, y, z, vx, vy, vz read + write) at the measured particle-steps/s;
peak 8 TB/s.  kernel_frac: the same bytes over the dominant kernel's
average launch duration (HIP events on the library's stream, every launch
of the timed region).  traffic: HBM bytes per step from the committed
rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this same command.
cpu_baseline: the FP64 C oracle (OpenMP) on this host, one thread and all
the threads OMP_NUM_THREADS allows, on a slice of this workload and on the
plain CPU-runnable configuration (64 x 128, 1e5 electrons, 200 steps).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "gas-atenuator_amd"))
sys.path.insert(0, str(REPO / "oracle"))

METRIC = "particle-steps/s + %HBM roofline; batched vs plain; adjoint/forward time"
HBM_PEAK_GBS = 8000.0
BYTES_PER_PARTICLE_STEP = 48
DEPOSIT_BYTES_PER_PARTICLE_STEP = 24
# The timed region times every other k_particles launch (0, 2, 4, ...): a
# launch timed by its own HIP events idles the GPU for about 4 us before and
# after it (rocprofv3 kernel trace), so timing all of them would add some 9 us
# to every step of the headline.  The even launches of the 20-step default
# sample the population's growth evenly (the mean of 0, 2, .., 18 equals that
# of 0, 1, .., 18, the launches of the pushing kernel variant).
TIME_EVERY = 2


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="plain_gpu_256x512")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-source", action="store_true", help="skip the photoionization-source leg")
    ap.add_argument("--no-dsmc", action="store_true", help="skip the DSMC neutral-gas leg")
    ap.add_argument("--cpu-particles", type=int, default=1_000_000)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-tuned-sort", action="store_true",
                    help="skip the extra run at the tuned cell-sort interval")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (5 times) and print its size: "
                         "the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def copy_calibration(mib: int):
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 5}}))


def cpu_threads() -> int:
    """Threads of the all-core CPU leg: OMP_NUM_THREADS when set (the host's
    share for one GPU), else the CPUs this process may run on."""
    import os
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else len(os.sched_getaffinity(0))


def cpu_model() -> str:
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_leg(w, host, n_particles: int, n_steps: int, seed: int, threads: int, secondaries: bool):
    """The FP64 oracle loop on the first n_particles of host for n_steps with
    `threads` OpenMP threads; returns (particle-steps/s, wall s, particle-steps)."""
    import numpy as np

    import gasx
    import oracle
    from gasx.workloads import density_grid

    p = gasx.make_params(w.nr, w.nz, w.dr, w.dz, w.phip, w.ddt, w.n_sub, fn=1.0)
    op = oracle.params_from(p)
    ot = oracle.build_tables(gasx.raw_tables())
    n = min(n_particles, host[0].size)
    st = np.stack([a[:n] for a in host]).astype(np.float64)
    o = oracle.State(st, cap=2 * n if secondaries else n)
    dens, gdr, gdz = density_grid(w)
    og = oracle.make_gas(dens.astype(np.float64), gdr, gdz)
    er = np.zeros((w.nz + 1) * w.nr)
    ez = np.zeros(w.nz * (w.nr + 1))
    h = np.zeros(w.nz * w.nr)
    oracle.set_threads(threads)
    try:
        t0 = time.perf_counter()
        oracle.run(op, ot, og, o, er, ez, h, n_steps, collide=True, secondaries=secondaries, seed=seed)
        dt = time.perf_counter() - t0
    finally:
        oracle.set_threads(1)
    # pushes performed: live particles each step (absorption is rare; the
    # population grows about linearly with secondaries)
    live_end = int(np.sum(o.dele[: o.n] == 0))
    psteps = 0.5 * (n + live_end) * n_steps
    return psteps / dt, dt, psteps


def cpu_baseline(w, host, n_particles: int, n_steps: int, seed: int):
    """The FP64 oracle on this host: a slice of the benchmarked workload (its
    first n_particles electrons for n_steps) and the plain CPU-runnable
    configuration exactly, each with one thread and with all threads."""
    from gasx.workloads import WORKLOADS, seed_electrons

    T = cpu_threads()
    vN, tN, _ = oracle_leg(w, host, n_particles, n_steps, seed, T, w.secondaries)
    v1, t1, _ = oracle_leg(w, host, n_particles, n_steps, seed, 1, w.secondaries)
    c1 = WORKLOADS["plain_cpu_64x128"]
    h1 = seed_electrons(c1.n_electrons, c1.rp, c1.lp, c1.phip, c1.sigma_r, seed=seed)
    c1N, c1tN, c1ps = oracle_leg(c1, h1, c1.n_electrons, c1.n_steps, seed, T, False)
    c11, c1t1, _ = oracle_leg(c1, h1, c1.n_electrons, c1.n_steps, seed, 1, False)
    n = min(n_particles, host[0].size)
    return {
        "value": vN,
        "unit": "particle-steps/s",
        "cores": T,
        "kind": "port",
        "sample": (f"FP64 C oracle, {T} OpenMP threads: first {n} of the seeded electrons on the "
                   f"{w.nr}x{w.nz} grid, {n_steps} steps with collisions"
                   f"{' and secondaries' if w.secondaries else ''} ({tN:.1f} s)"),
        "single_thread": {"value": v1, "cores": 1, "wall_s": t1},
        "nproc": len(__import__("os").sched_getaffinity(0)),
        "cpu_model": cpu_model(),
        "plain_cpu_64x128": {
            "sample": (f"the whole configuration: {c1.nr}x{c1.nz}, {c1.n_electrons} electrons, "
                       f"{c1.n_steps} steps, collisions on, secondaries off"),
            "particle_steps": c1ps,
            "value": c1N, "cores": T, "wall_s": c1tN,
            "single_thread": {"value": c11, "cores": 1, "wall_s": c1t1},
        },
    }


def cell_sorted(host, w):
    """Initial electrons ordered by their (iz, ir) cell (both legs below use it)."""
    import numpy as np
    x, y, z = (a.astype(np.float64) for a in host[:3])
    ir = np.minimum((np.hypot(x, y) / w.dr).astype(np.int64), w.nr - 1)
    iz = np.minimum((z / w.dz).astype(np.int64), w.nz - 1)
    order = np.argsort(iz * w.nr + ir, kind="stable")
    return tuple(a[order] for a in host)


def stencil_sorted(host, w):
    """Initial electrons ordered by the first cell of their deposit stencil
    (the cell sort's key: floor(r/dr - 1/2), floor(z/dz - 1/2) in FP32,
    clamped into the grid), as a source that emits cell by cell leaves them
    (INCIDENCE1.m:96-158 loops over rings and slabs)."""
    import numpy as np
    x, y, z = (a.astype(np.float32) for a in host[:3])
    r = np.sqrt(x * x + y * y)
    sr = r * np.float32(1.0 / w.dr) - np.float32(0.5)
    sz = z * np.float32(1.0 / w.dz) - np.float32(0.5)
    ir = np.clip(np.floor(sr).astype(np.int64), 0, w.nr - 1)
    iz = np.clip(np.floor(sz).astype(np.int64), 0, w.nz - 1)
    order = np.argsort(iz * w.nr + ir, kind="stable")
    return tuple(a[order] for a in host)


def bench_batched(seed: int, steps: int | None = None, sequential: bool = True, secondaries: bool = True):
    """Batched ensemble (64 runs of 128x256, log-uniform density and electron
    count, collisions and secondaries on with capacity 2x) in one captured graph with the
    autotuned launch shapes, against the same runs one after another with
    gasx_run.  Both legs start from the same cell-ordered electrons and produce
    identical events run by run (checked from the per-run counters)."""
    import numpy as np
    import torch

    import gasx
    from gasx.workloads import WORKLOADS, batched_runs, seed_electrons

    w = WORKLOADS["batched_128x256"]
    n_steps = steps or w.n_steps
    counts, dens = batched_runs(w, seed)
    caps = [(2 if secondaries else 1) * int(n) for n in counts]   # growth capped at 2x
    p = gasx.make_params(w.nr, w.nz, w.dr, w.dz, w.phip, w.ddt, w.n_sub, fn=1.0)
    ctx = gasx.Context(p, device=0)
    runs = [cell_sorted(seed_electrons(int(n), w.rp, w.lp, w.phip, w.sigma_r, seed=seed + 100 + r), w)
            for r, n in enumerate(counts)]
    ens = gasx.Ensemble(ctx, caps, dens, w.rp, w.lp)
    for r, h in enumerate(runs):
        ens.load_run(r, *h)
    init = (ens.pool.clone(), ens.counts.clone())

    def reset():
        ens.pool.copy_(init[0])
        ens.counts.copy_(init[1])
        for t in (ens.er, ens.ez, ens.hphi):
            t.zero_()
        ens.counters(reset=True)
        torch.cuda.synchronize()

    kw = dict(collide=True, secondaries=secondaries, seed=seed + 7, run_id=0)
    tuning = ens.autotune(trial_steps=5, **kw)
    reset()
    ens.run(n_steps, use_graph=True, **kw)            # capture + instantiate + first replay
    reset()
    t0 = time.perf_counter()
    ens.run(n_steps, use_graph=True, **kw)
    torch.cuda.synchronize()
    wall_b = time.perf_counter() - t0
    per_run = ens.counters()
    psteps_b = sum(c.n_pushed for c in per_run)
    out = {"value": psteps_b / wall_b, "unit": "particle-steps/s", "runs": len(counts),
           "steps": n_steps, "electrons": int(counts.sum()), "capacity_factor": 2.0 if secondaries else 1.0,
           "secondaries": secondaries, "wall_s": wall_b, "tuning": tuning,
           "graph": True, "grid": f"{w.nr}x{w.nz}"}
    del ens, init
    torch.cuda.empty_cache()
    if not sequential:
        return out
    wall_s, psteps_s, same = 0.0, 0, True
    gas_grids = [gasx.GasGrid(d, w.rp, w.lp, temperature=300.0) for d in dens]
    for r, h in enumerate(runs):
        e = gasx.ParticleSet(caps[r])
        e.load(*h)
        f = gasx.FieldState(w.nr, w.nz)
        ctx.counters(reset=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.run(e, f, gas_grids[r], n_steps, sort_every=0, **{**kw, "run_id": r})
        torch.cuda.synchronize()
        wall_s += time.perf_counter() - t0
        c = ctx.counters()
        psteps_s += c.n_pushed
        b = per_run[r]
        same &= (c.n_elastic, c.n_ionize, c.n_excite, c.n_pushed) == \
                (b.n_elastic, b.n_ionize, b.n_excite, b.n_pushed)
    out.update({"sequential_value": psteps_s / wall_s, "sequential_wall_s": wall_s,
                "speedup_vs_sequential": (psteps_b / wall_b) / (psteps_s / wall_s),
                "identical_events_to_sequential": bool(same)})
    return out


def bench_adjoint(seed: int, steps: int | None = None):
    """Differentiable run of the adjoint workload (128x256, 5e5 electrons,
    200 pushes, 32-node radial profile with nodes 2.5e25 (1 + 0.1 U(-1, 1)),
    checkpoints every 20 steps, objective W_Er over r >= 0.25 Rp): the time of
    one replayed forward and of one adjoint (backward, recomputation
    included) pass, reported as (forward + adjoint) / forward, and against
    the production step (gasx_run with collisions, same electrons, profile
    and steps)."""
    import numpy as np
    import torch

    import gasx
    from gasx.workloads import WORKLOADS, seed_electrons

    w = WORKLOADS["adjoint_128x256"]
    n_steps = steps or w.n_steps
    nn, every = w.extra["nodes"], w.extra["checkpoint"]
    p = gasx.make_params(w.nr, w.nz, w.dr, w.dz, w.phip, w.ddt, w.n_sub, fn=1.0)
    ctx = gasx.Context(p, device=0)
    host = cell_sorted(seed_electrons(w.n_electrons, w.rp, w.lp, w.phip, w.sigma_r, seed=seed), w)
    e = gasx.ParticleSet(w.n_electrons)
    e.load(*host)
    nodes0 = w.density * (1 + 0.1 * np.random.default_rng(seed).uniform(-1.0, 1.0, nn))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run = gasx.DiffRun(ctx, e, nodes0, n_steps=n_steps, checkpoint_every=every, seed=seed + 3,
                       r_frac=w.extra["shell"])
    t_rec = time.perf_counter() - t0
    run.forward(nodes0, 1.0)                      # warm
    run.backward()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        wv = run.forward(nodes0, 1.0)
    t_fwd = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        g, ga = run.backward()
    t_bwd = (time.perf_counter() - t0) / reps
    run.release()
    # production forward of the same electrons and steps (radial profile as a gas grid)
    rc = (np.arange(w.nr) + 0.5) * w.dr
    dens = np.interp(rc, (np.arange(nn) + 0.5) * w.rp / nn, nodes0).astype(np.float32).reshape(1, -1)
    gas = gasx.GasGrid(dens, w.dr, w.lp, temperature=300.0)
    f = gasx.FieldState(w.nr, w.nz)
    ctx.run(e, f, gas, 5, collide=True, seed=seed + 3)
    e.load(*host)
    for t in (f.er, f.ez, f.hphi):
        t.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.run(e, f, gas, n_steps, collide=True, seed=seed + 3)
    torch.cuda.synchronize()
    t_prod = time.perf_counter() - t0
    return {"workload": w.name, "grid": f"{w.nr}x{w.nz}", "electrons": w.n_electrons, "steps": n_steps,
            "nodes": nn, "checkpoint_every": every, "shell_r_over_rp": w.extra["shell"], "record_s": t_rec, "forward_s": t_fwd,
            "backward_s": t_bwd, "backward_over_forward": t_bwd / t_fwd,
            "forward_plus_adjoint_over_forward": (t_fwd + t_bwd) / t_fwd,
            "production_forward_s": t_prod, "backward_over_production_forward": t_bwd / t_prod,
            "forward_plus_adjoint_over_production_forward": (t_fwd + t_bwd) / t_prod,
            "objective_J": wv, "grad_amplitude": ga, "grad_nodes_norm": float(np.linalg.norm(g))}


def bench_photoionization(seed: int):
    """The photoionization source (gasx_photoionize) on the 256x512 workload's
    grid: a 1 keV pulse, Gaussian beam of sigma 50 um (beam radius 150 um),
    uniform N2 at 2.5e25 m^-3 and 300 K, FN = 1, gas depletion on; the photon
    number gives about 2e6 photoionizations (a photoelectron and an Auger
    electron each, and an N2++ ion).  Timed over 5 pulses into fresh
    populations."""
    import numpy as np
    import torch

    import gasx
    from gasx.core import M_ION
    from gasx.tables import source_tables
    from gasx.workloads import WORKLOADS

    w = WORKLOADS["plain_gpu_256x512"]
    p = gasx.make_params(w.nr, w.nz, w.dr, w.dz, w.phip, w.ddt, w.n_sub, fn=1.0)
    ctx = gasx.Context(p, device=0)
    tabs = source_tables()
    n_photons = 2.45e6
    e = gasx.ParticleSet(5_000_000)
    ions = gasx.ParticleSet(2_500_000, qnum=2.0, mass=M_ION)
    dens0 = np.full((w.nz, w.nr), w.density, np.float32)
    gas = gasx.GasGrid(dens0, w.dr, w.dz, temperature=np.full_like(dens0, 300.0))
    kw = dict(ions=ions, deplete=True, tables=tabs)
    ctx.photoionize(e, gas, 1.0, n_photons, 150e-6, seed=seed, **kw)       # warm
    walls, res = [], None
    for k in range(5):
        e.count_t.zero_()
        ions.count_t.zero_()
        gas.density.copy_(torch.from_numpy(dens0))
        gas.temperature.fill_(300.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = ctx.photoionize(e, gas, 1.0, n_photons, 150e-6, seed=seed, pulse=k, **kw)
        walls.append(time.perf_counter() - t0)
    wall = float(np.median(walls))
    return {"events": res["n_events"], "electrons": res["n_electrons"], "ions": res["n_ions"],
            "photons_left_fraction": res["photons_left"] / n_photons, "sigma_pe_m2": res["sigma_pe"],
            "wall_s": wall, "events_per_s": res["n_events"] / wall,
            "note": "median of 5 pulses, host call to return (result read back included)"}


DSMC_BYTES_PER_MOLECULE_STEP = 56      # FP32 x, y, z, vx, vy, vz, ei read + written


def dsmc_main2_case():
    """main2.m:5-47: N2 at 10 Pa and 300 K in a wedge of Rp = 1 cm, Lp = 3 cm,
    phimax = 0.5; cells of the mean free path (15 rings x 45 layers, 2880
    cells), 5000 molecules per cell, dt = 0.2 us."""
    import math

    from gasx.dsmc import CellMap, K_B_DSMC, N2_DIAMETER

    n_gas = 10.0 / K_B_DSMC / 300.0
    lam = 1 / (math.sqrt(2) * math.pi * n_gas * N2_DIAMETER ** 2)
    cm = CellMap(int(1e-2 / lam) + 1, int(0.03 / lam) + 1, 1e-2, 0.03, 0.5, lam)
    return cm, n_gas


def dsmc_cpu_baseline(cm, n_gas, seed: int):
    """The oracle's DSMC step (move, cells, collisions; OpenMP over molecules
    and cells) on the whole workload (2880 cells, 1.44e7 molecules) for 2
    steps."""
    import oracle as O
    from gasx.dsmc import CellMap, seed_molecules

    import numpy as np

    layers = cm.n_layers
    sl = CellMap(cm.n_rings, layers, cm.rp, cm.lp * layers / cm.n_layers, cm.phip, cm.delta)
    om = O.cell_map(sl.n_rings, sl.n_layers, sl.rp, sl.lp, sl.phip, sl.delta)
    nxy = O.cell_map_build(om)[4]
    n = 5000 * nxy * layers
    fn = n_gas * 0.5 * sl.rp ** 2 * sl.phip * sl.lp / n
    from gasx.dsmc import dsmc_params
    p = dsmc_params(2e-7, fn)
    op = O.dsmc_params(p.dt, p.fn, p.mass, p.diameter, p.sc_max, p.omega, p.zeta, p.alpha, p.min_attempts)
    st0 = np.ascontiguousarray(seed_molecules(n, sl, seed=seed).astype(np.float64))
    out = {}
    for threads in (cpu_threads(), 1):
        O.set_threads(threads)
        st = st0.copy()
        t0 = time.perf_counter()
        for k in range(2):
            O.dsmc_move(om, p.dt, st, round32=True)
            O.dsmc_collide(om, op, st, seed=seed, step=k, round32=True)
        wall = time.perf_counter() - t0
        out[threads] = (2 * n / wall, wall)
    O.set_threads(1)
    t_all, t_one = out[cpu_threads()], out[1]
    return {"value": t_all[0], "unit": "molecule-steps/s", "cores": cpu_threads(), "kind": "port",
            "sample": f"FP64 C oracle: the whole workload ({nxy * layers} cells, {n} molecules), "
                      f"2 steps ({t_all[1]:.1f} s on {cpu_threads()} threads)",
            "single_thread": {"value": t_one[0], "cores": 1, "wall_s": t_one[1]}}


def bench_dsmc(seed: int, steps: int = 20, cpu: bool = True):
    """The DSMC neutral-gas step (gasx_dsmc_step) on main2.m's case: 1.44e7
    molecules, one move + cell sort + collision step per 0.2 us, molecules
    permuted into cell order every step.  value = molecule-steps/s over the
    timed steps (molecules resident in HBM)."""
    import torch

    import gasx
    from gasx.dsmc import MoleculeSet, dsmc_params, seed_molecules

    cm, n_gas = dsmc_main2_case()
    ctx = gasx.Context(gasx.make_params(16, 16), device=0)
    _, nxy = cm.info(ctx.lib)
    ncell = nxy * cm.n_layers
    n = 5000 * ncell
    fn = n_gas * 0.5 * cm.rp ** 2 * cm.phip * cm.lp / n
    mol = MoleculeSet(n).load(seed_molecules(n, cm, seed=seed))
    p = dsmc_params(2e-7, fn, reorder=True)
    for k in range(3):
        ctx.dsmc.step(mol, cm, p, seed=seed, step=k)
    torch.cuda.synchronize()
    ctx.dsmc.counters(reset=True)
    t0 = time.perf_counter()
    for k in range(steps):
        ctx.dsmc.step(mol, cm, p, seed=seed, step=3 + k)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    c = ctx.dsmc.counters(reset=True)
    value = n * steps / wall
    achieved = DSMC_BYTES_PER_MOLECULE_STEP * value / 1e9
    out = {"workload": "dsmc_main2", "molecules": n, "cells": ncell, "rings": cm.n_rings, "layers": cm.n_layers,
           "dt_s": p.dt, "steps": steps, "warmup": 3, "ms_per_step": wall * 1e3 / steps, "value": value,
           "unit": "molecule-steps/s",
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "bytes_per_molecule_step": DSMC_BYTES_PER_MOLECULE_STEP,
                        "scope": "whole step (move, sort, reorder, collisions)"},
           "attempts_per_step": c.n_attempts / steps, "collisions_per_step": c.n_collisions / steps,
           "sc_exceed": c.n_sc_exceed, "wall_events_per_step": c.n_boundary / steps}
    del mol
    torch.cuda.empty_cache()
    if cpu:
        out["cpu_baseline"] = dsmc_cpu_baseline(cm, n_gas, seed)
    return out


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "the multi-GPU sharding is a design note, not a measured configuration.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return

    import numpy as np
    import torch

    import gasx
    from gasx.workloads import WORKLOADS, density_grid, seed_electrons

    if args.workload == "dsmc_main2":
        d = bench_dsmc(args.seed, steps=args.steps if args.steps != 1000 else 20, cpu=not args.no_cpu_baseline)
        print(json.dumps({
            "metric": METRIC, "value": d["value"], "unit": d["unit"], "n_gpus": 1, "steps": d["steps"],
            "warmup": d["warmup"], "ms_per_step": d["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp64 arithmetic, fp32 state", "data": "synthetic",
            "config": {"workload": d["workload"], "molecules": d["molecules"], "cells": d["cells"]},
            "roofline": d["roofline"], "cpu_baseline": d.get("cpu_baseline"), "dsmc": d}))
        return
    w = WORKLOADS[args.workload]
    if w.name == "batched_128x256":
        b = bench_batched(args.seed, steps=args.steps if args.steps != 1000 else None)
        print(json.dumps({
            "metric": METRIC, "value": b["value"], "unit": "particle-steps/s", "n_gpus": 1,
            "steps": b["steps"], "warmup": 1, "ms_per_step": b["wall_s"] * 1e3 / b["steps"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
            "data": "synthetic", "config": {"workload": w.name, "runs": b["runs"], "grid": b["grid"],
                                            "electrons": b["electrons"]},
            "batched": b}))
        return
    if w.name == "adjoint_128x256":
        a = bench_adjoint(args.seed, steps=args.steps if args.steps != 1000 else None)
        print(json.dumps({
            "metric": METRIC, "value": a["forward_plus_adjoint_over_forward"],
            "unit": "(forward + adjoint) time / forward time",
            "n_gpus": 1, "steps": a["steps"], "warmup": 1, "ms_per_step": a["backward_s"] * 1e3 / a["steps"],
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
            "data": "synthetic", "config": {"workload": w.name, "grid": a["grid"], "electrons": a["electrons"]},
            "adjoint": a}))
        return
    if w.n_electrons <= 0:
        raise SystemExit(f"workload {w.name} is not a single-run bench workload")
    p = gasx.make_params(w.nr, w.nz, w.dr, w.dz, w.phip, w.ddt, w.n_sub, fn=1.0)
    ctx = gasx.Context(p, device=0)
    host = seed_electrons(w.n_electrons, w.rp, w.lp, w.phip, w.sigma_r, seed=args.seed)
    # the GPU run starts from the electrons in cell order (the CPU baseline's
    # sample stays the seeded order: a prefix of it covers the whole spot)
    host_t = [torch.from_numpy(np.ascontiguousarray(a)).pin_memory() for a in stencil_sorted(host, w)]
    e = gasx.ParticleSet(w.capacity)
    f = gasx.FieldState(w.nr, w.nz)
    dens, gdr, gdz = density_grid(w)
    gas = gasx.GasGrid(dens, gdr, gdz, temperature=300.0)

    def reset():
        for row, t in enumerate(host_t):
            e.data[row, : w.n_electrons].copy_(t, non_blocking=True)
        e.count_t.fill_(w.n_electrons)
        for t in (f.er, f.ez, f.hphi):
            t.zero_()

    run_kw = dict(collide=True, secondaries=w.secondaries, sort_every=w.sort_every, seed=args.seed + 1)
    reset()
    ctx.run(e, f, gas, args.warmup, time_kernels=2, time_every=TIME_EVERY, **run_kw)
    torch.cuda.synchronize()
    reset()
    ctx.counters(reset=True)
    torch.cuda.synchronize()
    # timed region: every TIME_EVERY-th k_particles launch (the roofline
    # kernel) timed by HIP events of its own dispatch
    t0 = time.perf_counter()
    ctx.run(e, f, gas, args.steps, time_kernels=2, time_every=TIME_EVERY, **run_kw)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    c = ctx.counters()
    tm = ctx.timing()
    psteps = c.n_pushed
    value = psteps / wall
    # per-kernel breakdown: the next steps of the same run with every launch timed
    n_bd = min(args.steps, 200)
    ctx.run(e, f, gas, n_bd, time_kernels=1, step0=args.steps, **run_kw)
    torch.cuda.synchronize()
    tb = ctx.timing()
    psteps_bd = ctx.counters().n_pushed - psteps

    # The particle pass of a step is two kernels: k_particles (push, both half
    # drifts, collisions; the dominant one, reported as the roofline kernel),
    # which reads and writes the 24-byte state, and k_deposit (the current
    # deposit), which reads it.
    # one k_particles launch per step (one species); launches 0, k, 2k, ...
    # timed, so the average duration is over a uniform sample of the steps
    n_timed = tm["n_particle_launches"]
    assert n_timed == (args.steps + TIME_EVERY - 1) // TIME_EVERY, n_timed
    avg_ms = tm["particle_ms"] / n_timed
    n_launch = args.steps
    bytes_per_launch = BYTES_PER_PARTICLE_STEP * psteps / n_launch
    kernel_achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    # the step: the same 48 bytes per particle-step at the measured rate
    step_achieved = BYTES_PER_PARTICLE_STEP * value / 1e9
    # the drift/deposit kernel and the whole particle pass from the breakdown steps
    n_dd = max(1, tb["n_drift_launches"])
    dd_ms = tb["drift_ms"] / n_dd
    pass_ms = (tb["particle_ms"] + tb["drift_ms"]) / n_bd
    # HBM bytes per step from the committed FETCH_SIZE + WRITE_SIZE passes of
    # this same command (calibrated against a plain copy); none for other
    # step counts
    traffic, traffic_src = None, None
    tfile = REPO / "profiles" / f"step_{w.name}_steps{args.steps}_warmup{args.warmup}.traffic.txt"
    if tfile.exists():
        for ln in tfile.read_text().splitlines():
            if ln.startswith("hbm_bytes_per_step:"):
                traffic = float(ln.split(":", 1)[1])
                traffic_src = f"profiles/{tfile.name} (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, of this command)"
    tuned = None
    ts = w.extra.get("sort_every_tuned")
    if ts and not args.no_tuned_sort:
        # the same timed steps from the same start with the tuned sort interval
        reset()
        ctx.counters(reset=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.run(e, f, gas, args.steps, **{**run_kw, "sort_every": ts})
        torch.cuda.synchronize()
        wt = time.perf_counter() - t0
        tuned = {"sort_every": ts, "value": ctx.counters().n_pushed / wt, "ms_per_step": wt * 1e3 / args.steps}
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": w.name, "grid": f"{w.nr}x{w.nz}", "electrons": w.n_electrons,
            "capacity": w.capacity, "density_m3": w.density, "dt_s": w.ddt, "field_substeps": w.n_sub,
            "secondaries": w.secondaries, "sort_every": w.sort_every,
            "field_solver": "fp64", "particles": "fp32 SoA", "initial_order": "deposit-stencil cell",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": step_achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": step_achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_per": "step",
            "traffic_source": traffic_src,
            "bytes_per_particle_step": BYTES_PER_PARTICLE_STEP,
            "scope": "whole step (48 B per particle-step at the measured particle-steps/s)",
            "kernel": "k_particles (gather + push + collision Monte Carlo + secondary staging)",
            "kernel_achieved": kernel_achieved,
            "kernel_frac": kernel_achieved / HBM_PEAK_GBS,
            "kernel_avg_launch_ms": avg_ms,
        },
        "deposit_kernel": {   # breakdown steps; 24 algorithmic bytes (the state, read)
            "avg_launch_ms": dd_ms,
            "achieved_GBs": DEPOSIT_BYTES_PER_PARTICLE_STEP * psteps_bd / n_dd / (dd_ms * 1e-3) / 1e9,
        },
        "particle_pass": {   # both kernels of a step against the step's 48 algorithmic bytes per particle
            "ms_per_step": pass_ms,
            "achieved_GBs": BYTES_PER_PARTICLE_STEP * psteps_bd / n_bd / (pass_ms * 1e-3) / 1e9,
        },
        "breakdown_us_per_step": {k[:-3]: tb[k] * 1e3 / n_bd
                                  for k in ("particle_ms", "drift_ms", "fdtd_ms", "append_ms", "sort_ms")},
        "breakdown_note": f"every launch timed, {n_bd} further steps of the same run",
        "events": {"elastic": c.n_elastic, "ionize": c.n_ionize, "excite": c.n_excite,
                   "absorbed": c.n_absorbed, "dropped_secondaries": c.n_dropped_sec,
                   "bound_violations": c.n_bound_violations, "final_slots": e.n},
        "fdtd_cell_updates_per_s": w.nr * w.nz * w.n_sub * n_bd / (tb["fdtd_ms"] * 1e-3)
        if tb["fdtd_ms"] > 0 else None,
    }
    if tuned:
        line["tuned_sort"] = tuned
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(w, host, args.cpu_particles, args.cpu_steps, args.seed + 1)
    if not args.no_batched:
        del e, f, host_t
        torch.cuda.empty_cache()
        line["batched"] = bench_batched(args.seed)
        torch.cuda.empty_cache()
        nb = bench_batched(args.seed, secondaries=False)
        line["batched"]["without_secondaries"] = {
            k: nb[k] for k in ("value", "wall_s", "sequential_value", "sequential_wall_s", "speedup_vs_sequential",
                               "identical_events_to_sequential", "tuning")}
    if not args.no_adjoint:
        torch.cuda.empty_cache()
        line["adjoint"] = bench_adjoint(args.seed)
    if not args.no_source:
        torch.cuda.empty_cache()
        line["photoionization"] = bench_photoionization(args.seed)
    if not args.no_dsmc:
        torch.cuda.empty_cache()
        line["dsmc"] = bench_dsmc(args.seed, cpu=not args.no_cpu_baseline)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
