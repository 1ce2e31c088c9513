"""Collision statistics of the production loop (gasx_run: push, collision
Monte Carlo, secondaries, deposit, field cycle) against the FP64 oracle loop
(electron_transport.cu / electron_collision.m restated), over 20 seeds.

The two sides use independent Philox keys (GPU seeds 0-19, oracle seeds
100-119), so they are independent samples of the same process: two-sample
Kolmogorov-Smirnov p > 0.01 on the per-seed elastic / ionization / excitation
counts and on the pooled energy spectra of the surviving electrons."""
import numpy as np
import pytest
from scipy import stats

import oracle

pytestmark = pytest.mark.gpu

NR, NZ, N, STEPS, DENS = 32, 64, 4000, 16, 2.5e25
SEEDS = range(20)
ME_C2_EV = 9.1096e-31 * 9e16 / 1.6e-19


def energies_ev(st):
    v = st[3:6].astype(np.float64)
    ok = ~np.isnan(st[2])
    b2 = (v[:, ok] ** 2).sum(0) / 9e16
    g = 1 / np.sqrt(1 - b2)
    return ME_C2_EV * (g - 1)


@pytest.fixture(scope="module")
def samples(gpu, otab):
    import gasx
    from gasx.workloads import seed_electrons
    p = gasx.make_params(NR, NZ, fn=1.0)
    op = oracle.params_from(p)
    ctx = gasx.Context(p)
    gas = gasx.GasGrid.uniform(DENS, NR * p.dr, NZ * p.dz)
    og = oracle.make_gas(np.full((1, 1), DENS), NR * p.dr, NZ * p.dz)
    z = np.zeros
    out = {"g_counts": [], "o_counts": [], "g_e": [], "o_e": []}
    for s in SEEDS:
        st = np.stack(seed_electrons(N, NR * p.dr, NZ * p.dz, p.phip, seed=1000 + s))
        e = gasx.ParticleSet(3 * N).load(*st)
        f = gasx.FieldState(NR, NZ)
        ctx.counters(reset=True)
        ctx.run(e, f, gas, STEPS, collide=True, secondaries=True, seed=s)
        c = ctx.counters()
        assert c.n_bound_violations == 0 and c.n_dropped_sec == 0
        out["g_counts"].append((c.n_elastic, c.n_ionize, c.n_excite))
        out["g_e"].append(energies_ev(e.state()))
        o = oracle.State(st, cap=3 * N)
        _, _, _, oc = oracle.run(op, otab, og, o, z((NZ + 1) * NR), z(NZ * (NR + 1)), z(NZ * NR), STEPS,
                                 collide=True, secondaries=True, seed=100 + s)
        out["o_counts"].append((oc.n_elastic, oc.n_ionize, oc.n_excite))
        ov = o.v[:, :o.n].copy()
        ov[2, o.dele[:o.n] == 1] = np.nan
        out["o_e"].append(energies_ev(ov))
    return {k: (np.array(v) if k.endswith("counts") else np.concatenate(v)) for k, v in out.items()}


@pytest.mark.parametrize("kind,col", [("elastic", 0), ("ionization", 1), ("excitation", 2)])
def test_event_counts_ks(samples, kind, col):
    g, o = samples["g_counts"][:, col], samples["o_counts"][:, col]
    assert g.min() > 100, f"too few {kind} events for the test"
    p = stats.ks_2samp(g, o).pvalue
    assert p > 0.01, (kind, g, o, p)


def test_energy_spectrum_ks(samples):
    g, o = samples["g_e"], samples["o_e"]
    assert g.size > 0.9 * N * len(SEEDS) and o.size > 0.9 * N * len(SEEDS)
    p = stats.ks_2samp(g, o).pvalue
    assert p > 0.01, p


def test_low_energy_tail_ks(samples):
    """Below 100 eV: the secondaries and the degraded primaries."""
    g, o = samples["g_e"], samples["o_e"]
    g, o = g[g < 100], o[o < 100]
    assert g.size > 1000
    p = stats.ks_2samp(g, o).pvalue
    assert p > 0.01, p
