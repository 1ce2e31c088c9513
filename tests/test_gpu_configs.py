"""The five measured configurations, each run at its own size on the GPU.

* plain_cpu_64x128 (64 x 128, 1e5 electrons, 200 steps, secondaries off):
  20 GPU seeds against the FP64 oracle's 20 independent seeds stored in
  tests/golden/c1_oracle_stats.npz (tests/golden/make_c1_fixture.py):
  two-sample KS on the per-seed event counts and on the final spectrum.
* plain_gpu_256x512 (256 x 512, 4e6 electrons, capacity 8e6, secondaries,
  cell sort every 10 steps): 100 steps, across the capacity.
* batched_128x256 (64 runs, log-uniform density and count): batch ==
  sequential bit for bit for every run, and one run against the oracle.
* adjoint_128x256 (5e5 electrons, 200 steps, 32 nodes, checkpoints every
  20): gradient against central differences (h = 1e-3).
* avalanche_1024x2048 (3e7 electrons, capacity 9e7, density gradient, sort
  every 5 steps): 10 steps with the same property checks as the 256 x 512
  run and the cell order of a final sort.

Property checks at full size (the oracle would need hours there): no
collision-bound violation, live = seeded + appended - absorbed, finite fields,
run-to-run determinism of particles and fields, and the fixed-point current
sum rule sum_cells Q = sum_particles kq v (EM7.cu:174-202 weights sum to one)
on the final state.
"""
from pathlib import Path

import numpy as np
import pytest
from scipy import stats

import oracle
from conftest import log

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
ME_C2_EV = 9.1096e-31 * 9e16 / 1.6e-19


def energies_ev(st):
    v = st[3:6].astype(np.float64)
    ok = ~np.isnan(st[2])
    b2 = (v[:, ok] ** 2).sum(0) / 9e16
    return ME_C2_EV * (1 / np.sqrt(1 - b2) - 1)


def workload(name):
    from gasx.workloads import WORKLOADS
    return WORKLOADS[name]


def setup_run(w, seed, capacity=None):
    import gasx
    from gasx.workloads import density_grid, seed_electrons
    p = gasx.make_params(w.nr, w.nz, w.dr, w.dz, w.phip, w.ddt, w.n_sub, fn=1.0)
    ctx = gasx.Context(p)
    host = seed_electrons(w.n_electrons, w.rp, w.lp, w.phip, w.sigma_r, seed=seed)
    e = gasx.ParticleSet(capacity or w.capacity)
    e.load(*host)
    dens, gdr, gdz = density_grid(w)
    gas = gasx.GasGrid(dens, gdr, gdz, temperature=300.0)
    return p, ctx, host, e, gas


def same_bits(a, b):
    """Bitwise equality of two float tensors on the device (NaN slots included)."""
    import torch
    return a.shape == b.shape and torch.equal(a.contiguous().view(torch.int32), b.contiguous().view(torch.int32))


def sum_rule(p, ctx, e):
    """Fixed-point current of the state against kq * sum of v (FP64 on the
    host), within the rounding of each contribution: 1/2 per cell and four
    cells per particle, plus FP32 rounding of the weights and products."""
    st = e.state().astype(np.float64)
    ok = ~np.isnan(st[2])
    x, y, vx, vy, vz = st[0, ok], st[1, ok], st[3, ok], st[4, ok], st[5, ok]
    r = np.hypot(x, y)
    vr = np.where(r > 0, (x * vx + y * vy) / np.where(r > 0, r, 1.0), 0.0)
    kq = -(2.0 ** 32) / p.c
    qr, qz = ctx.deposit(e)
    sr, sz = int(qr.sum().item()), int(qz.sum().item())
    tol_r = np.sum(2.0 + 64 * 2.0 ** -24 * np.abs(kq * np.sqrt(vx ** 2 + vy ** 2)))
    tol_z = np.sum(2.0 + 64 * 2.0 ** -24 * np.abs(kq * vz))
    assert abs(sr - kq * vr.sum()) <= tol_r, (sr, kq * vr.sum(), tol_r)
    assert abs(sz - kq * vz.sum()) <= tol_z, (sz, kq * vz.sum(), tol_z)


def bookkeeping(e, n0, c):
    st = e.state()
    live = int(np.count_nonzero(~np.isnan(st[2])))
    assert live == n0 + c.n_ionize - c.n_dropped_sec - c.n_absorbed, (live, n0, c)
    assert c.n_bound_violations == 0
    return st


# ---- plain_cpu_64x128 ----------------------------------------------------------------
def test_plain_cpu_64x128_statistics_match_oracle_fixture(gpu):
    import gasx
    from gasx.workloads import seed_electrons
    fx = np.load(GOLDEN / "c1_oracle_stats.npz")
    nr, nz, n, steps = (int(v) for v in fx["config"])
    w = workload("plain_cpu_64x128")
    assert (nr, nz, n, steps) == (w.nr, w.nz, w.n_electrons, w.n_steps)
    assert float(fx["density"]) == w.density
    log(f"plain_cpu_64x128: 20 seeds x {n} electrons x {steps} steps")
    p = gasx.make_params(nr, nz, w.dr, w.dz, w.phip, w.ddt, w.n_sub, fn=1.0)
    ctx = gasx.Context(p)
    gas = gasx.GasGrid.uniform(w.density, w.rp, w.lp)
    counts, spec = [], []
    for s in range(20):
        st = np.stack(seed_electrons(n, w.rp, w.lp, w.phip, w.sigma_r, seed=3000 + s))
        e = gasx.ParticleSet(n).load(*st)
        f = gasx.FieldState(nr, nz)
        ctx.counters(reset=True)
        ctx.run(e, f, gas, steps, collide=True, secondaries=False, seed=s)
        c = ctx.counters()
        assert c.n_bound_violations == 0 and c.n_pushed > 0.99 * n * steps
        counts.append((c.n_elastic, c.n_ionize, c.n_excite))
        spec.append(energies_ev(e.state()))
    counts = np.array(counts)
    for col, kind in enumerate(("elastic", "ionization", "excitation")):
        pv = stats.ks_2samp(counts[:, col], fx["counts"][:, col]).pvalue
        print(kind, counts[:, col].mean(), fx["counts"][:, col].mean(), pv)
        assert pv > 0.01, (kind, counts[:, col], fx["counts"][:, col])
    g = np.concatenate(spec)
    o = fx["energies_ev"].ravel().astype(np.float64)
    pv = stats.ks_2samp(g, o).pvalue
    print("spectrum", np.median(g), np.median(o), pv)
    assert pv > 0.01


# ---- plain_gpu_256x512 ---------------------------------------------------------------
def test_plain_gpu_256x512_across_capacity(gpu):
    import torch
    w = workload("plain_gpu_256x512")
    assert w.sort_every == 10 and w.capacity == 2 * w.n_electrons
    steps = 100
    log(f"plain_gpu_256x512: {w.n_electrons} electrons, cap {w.capacity}, {steps} steps, sort every 10")
    out = []
    for rep in range(2):
        p, ctx, host, e, gas = setup_run(w, seed=0)
        import gasx
        f = gasx.FieldState(w.nr, w.nz)
        ctx.counters(reset=True)
        ctx.run(e, f, gas, steps, collide=True, secondaries=True, sort_every=w.sort_every, seed=1)
        torch.cuda.synchronize()
        out.append((p, ctx, e, f, ctx.counters()))
    (p, ctx, e, f, c), (_, _, e2, f2, c2) = out
    print(c)
    assert c == c2
    assert same_bits(e.data, e2.data) and same_bits(e.count_t, e2.count_t)
    for a, b in ((f.er, f2.er), (f.ez, f2.ez), (f.hphi, f2.hphi)):
        assert torch.equal(a, b) and bool(torch.isfinite(a).all())
    assert c.n_dropped_sec > 0 and e.n == w.capacity      # crossed the capacity
    bookkeeping(e, w.n_electrons, c)
    sum_rule(p, ctx, e)


# ---- batched_128x256 -----------------------------------------------------------------
def test_batched_128x256_equals_sequential_and_tracks_oracle(gpu, otab):
    import torch
    import gasx
    from gasx.workloads import batched_runs, seed_electrons
    w = workload("batched_128x256")
    steps = 6
    counts, dens = batched_runs(w, 0)
    caps = [2 * int(n) for n in counts]                  # secondaries on, growth capped at 2x
    log(f"batched_128x256: {len(counts)} runs, {int(counts.sum())} electrons, {steps} steps")
    p = gasx.make_params(w.nr, w.nz, w.dr, w.dz, w.phip, w.ddt, w.n_sub, fn=1.0)
    ctx = gasx.Context(p)
    runs = [seed_electrons(int(n), w.rp, w.lp, w.phip, w.sigma_r, seed=100 + r) for r, n in enumerate(counts)]
    ens = gasx.Ensemble(ctx, caps, dens, w.rp, w.lp)
    for r, h in enumerate(runs):
        ens.load_run(r, *h)
    kw = dict(collide=True, secondaries=True, seed=7)
    ens.run(steps, run_id=0, use_graph=True, **kw)
    per_run = ens.counters()
    torch.cuda.synchronize()
    for r, h in enumerate(runs):
        e = gasx.ParticleSet(caps[r]).load(*h)
        f = gasx.FieldState(w.nr, w.nz)
        gas = gasx.GasGrid(dens[r], w.rp, w.lp, temperature=300.0)
        ctx.counters(reset=True)
        ctx.run(e, f, gas, steps, sort_every=0, run_id=r, **kw)
        c = ctx.counters()
        b = int(ens.base[r])
        n = int(ens.counts[r].item())
        assert n == e.n, r
        assert same_bits(ens.pool[:, b:b + n], e.data[:, :n]), r
        assert torch.equal(ens.er[r], f.er) and torch.equal(ens.ez[r], f.ez) and torch.equal(ens.hphi[r], f.hphi), r
        g = per_run[r]
        assert (g.n_elastic, g.n_ionize, g.n_excite, g.n_pushed, g.n_dropped_sec, g.n_bound_violations) == \
               (c.n_elastic, c.n_ionize, c.n_excite, c.n_pushed, c.n_dropped_sec, 0), r
    # one run of the ensemble against the FP64 oracle loop on an independent seed
    r = int(np.argmin(np.abs(counts - 60_000)))
    gst = ens.run_state(r)
    op = oracle.params_from(p)
    og = oracle.make_gas(dens[r].astype(np.float64), w.rp, w.lp)
    o = oracle.State(np.stack(runs[r]), cap=caps[r])
    z = np.zeros
    _, _, _, oc = oracle.run(op, otab, og, o, z((w.nz + 1) * w.nr), z(w.nz * (w.nr + 1)), z(w.nz * w.nr), steps,
                             collide=True, secondaries=True, seed=9001)
    g = per_run[r]
    for kind, a, b in (("elastic", g.n_elastic, oc.n_elastic), ("ionization", g.n_ionize, oc.n_ionize),
                       ("excitation", g.n_excite, oc.n_excite)):
        assert abs(a - b) <= 5 * np.sqrt(a + b) + 5, (kind, a, b)   # two Poisson-like counts
    ov = o.v[:, :o.n].copy()
    ov[2, o.dele[:o.n] == 1] = np.nan
    pv = stats.ks_2samp(energies_ev(gst), energies_ev(ov)).pvalue
    print("batched run", r, "counts", g, oc.n_elastic, oc.n_ionize, oc.n_excite, "spectrum KS p", pv)
    assert pv > 0.01


# ---- adjoint_128x256 -----------------------------------------------------------------
def test_adjoint_128x256_gradient_matches_central_differences(gpu):
    import gasx
    from gasx.workloads import seed_electrons
    w = workload("adjoint_128x256")
    nn, every, shell = w.extra["nodes"], w.extra["checkpoint"], w.extra["shell"]
    log(f"adjoint_128x256: {w.n_electrons} electrons, {w.n_steps} steps, {nn} nodes, checkpoints every {every}")
    p = gasx.make_params(w.nr, w.nz, w.dr, w.dz, w.phip, w.ddt, w.n_sub, fn=1.0)
    ctx = gasx.Context(p)
    host = seed_electrons(w.n_electrons, w.rp, w.lp, w.phip, w.sigma_r, seed=0)
    e = gasx.ParticleSet(w.n_electrons).load(*host)
    nodes0 = w.density * (1 + 0.1 * np.random.default_rng(0).uniform(-1.0, 1.0, nn))
    run = gasx.DiffRun(ctx, e, nodes0, n_steps=w.n_steps, checkpoint_every=every, seed=3, r_frac=shell)
    w0 = run.forward(nodes0, 1.0)
    g, ga = run.backward()
    assert w0 > 0 and np.all(np.isfinite(g)) and np.linalg.norm(g) > 0
    h = 1e-3
    fda = (run.forward(nodes0, 1 + h) - run.forward(nodes0, 1 - h)) / (2 * h)
    d = g * nodes0 / np.linalg.norm(g * nodes0) * nodes0        # the gradient's direction, relative steps
    fd = (run.forward(nodes0 + h * d, 1.0) - run.forward(nodes0 - h * d, 1.0)) / (2 * h)
    print("W", w0, "dW/dA", ga, "fd", fda, "directional", float(g @ d), "fd", fd)
    assert fda == pytest.approx(ga, rel=1e-3, abs=0)
    assert fd == pytest.approx(float(g @ d), rel=2e-2, abs=0)
    run.release()


# ---- avalanche_1024x2048 -------------------------------------------------------------
def test_avalanche_1024x2048_ten_steps(gpu):
    import torch
    import gasx
    w = workload("avalanche_1024x2048")
    steps = 10
    log(f"avalanche_1024x2048: {w.n_electrons} electrons, cap {w.capacity}, {steps} steps, sort every {w.sort_every}")
    p, ctx, host, e, gas = setup_run(w, seed=0)
    del host
    init = e.data[:, : w.n_electrons].clone()
    res = []
    for rep in range(2):
        if rep:
            e.data[:, : w.n_electrons].copy_(init)
            e.count_t.fill_(w.n_electrons)
        f = gasx.FieldState(w.nr, w.nz)
        ctx.counters(reset=True)
        ctx.run(e, f, gas, steps, collide=True, secondaries=True, sort_every=w.sort_every, seed=1)
        torch.cuda.synchronize()
        c = ctx.counters()
        n = e.n
        res.append((e.data[:, :n].clone(), f, c))
    (d1, f1, c1), (d2, f2, c2) = res
    print(c1)
    assert c1 == c2 and same_bits(d1, d2)
    del d1, d2, init
    for a, b in ((f1.er, f2.er), (f1.ez, f2.ez), (f1.hphi, f2.hphi)):
        assert torch.equal(a, b) and bool(torch.isfinite(a).all())
    assert c1.n_ionize > 0.1 * w.n_electrons
    bookkeeping(e, w.n_electrons, c1)
    sum_rule(p, ctx, e)
    # the cell order of a sort: 2^21 cells + the absorbed key need 22 bits
    ctx.sort(e)
    st = e.state()
    assert not np.isnan(st[2]).any()
    x, y, z = st[0], st[1], st[2]
    r2 = x * x
    r2 = r2 + y * y
    r = np.sqrt(r2)
    kr = np.clip(np.floor(r * np.float32(1 / p.dr) - np.float32(0.5)), 0, w.nr - 1).astype(np.int64)
    kz = np.clip(np.floor(z * np.float32(1 / p.dz) - np.float32(0.5)), 0, w.nz - 1).astype(np.int64)
    key = kz * w.nr + kr
    assert key.max() >= 1 << 20
    assert np.all(np.diff(key) >= 0)
