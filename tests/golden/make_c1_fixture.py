"""Writes tests/golden/c1_oracle_stats.npz: the FP64 oracle's statistics of the
plain CPU-runnable configuration (64 x 128 grid, dr = dz = 10 um, uniform N2
at 2.5e25 m^-3, 1e5 Gaussian-spot electrons with the 70/30 400 +- 20 eV /
30 +- 10 eV mix, collisions on, secondaries off, 200 steps of 1e-14 s),
over 20 seeds.

Per seed it keeps the elastic / ionization / excitation counts and 1000
final electron energies drawn uniformly from the survivors.  The GPU test
(tests/test_gpu_configs.py) runs the HIP path on 20 independent seeds of the
same configuration and compares with two-sample KS.  Run in this repository:

    python tests/golden/make_c1_fixture.py
"""
from __future__ import annotations

import sys
import time
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(REPO / "gas-atenuator_amd"))
sys.path.insert(0, str(REPO / "oracle"))

NR, NZ, N, STEPS, DENS = 64, 128, 100_000, 200, 2.5e25
SEEDS = range(20)
ELECTRON_SEED0, PHILOX_SEED0 = 2000, 5000      # the GPU side uses other seeds
KEEP = 1000
OUT = Path(__file__).resolve().parent / "c1_oracle_stats.npz"


def energies_ev(v, ok):
    b2 = (v[3:6, ok] ** 2).sum(0) / 9e16
    return 9.1096e-31 * 9e16 / 1.6e-19 * (1 / np.sqrt(1 - b2) - 1)


def one_seed(s: int):
    import gasx
    import oracle
    from gasx.tables import raw_tables
    from gasx.workloads import seed_electrons
    p = gasx.make_params(NR, NZ, fn=1.0)
    op = oracle.params_from(p)
    ot = oracle.build_tables(raw_tables())
    og = oracle.make_gas(np.full((1, 1), DENS), NR * p.dr, NZ * p.dz)
    st = np.stack(seed_electrons(N, NR * p.dr, NZ * p.dz, p.phip, seed=ELECTRON_SEED0 + s))
    o = oracle.State(st, cap=N)
    z = np.zeros
    t0 = time.perf_counter()
    _, _, _, c = oracle.run(op, ot, og, o, z((NZ + 1) * NR), z(NZ * (NR + 1)), z(NZ * NR), STEPS,
                            collide=True, secondaries=False, seed=PHILOX_SEED0 + s)
    ok = o.dele[:o.n] == 0
    e = energies_ev(o.v[:, :o.n], ok)
    pick = np.random.default_rng(s).choice(e.size, KEEP, replace=False)
    return s, (c.n_elastic, c.n_ionize, c.n_excite), np.sort(e[pick]), time.perf_counter() - t0


def main():
    with ProcessPoolExecutor(max_workers=8) as ex:
        res = sorted(ex.map(one_seed, SEEDS))
    counts = np.array([r[1] for r in res], dtype=np.int64)
    energies = np.stack([r[2] for r in res]).astype(np.float32)
    np.savez_compressed(OUT, counts=counts, energies_ev=energies,
                        config=np.array([NR, NZ, N, STEPS], dtype=np.int64), density=np.array(DENS),
                        electron_seed0=np.array(ELECTRON_SEED0), philox_seed0=np.array(PHILOX_SEED0))
    print(f"wrote {OUT}: counts mean {counts.mean(0)}, {sum(r[3] for r in res):.0f} s of oracle time")


if __name__ == "__main__":
    main()
