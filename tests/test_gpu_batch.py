"""Batched ensemble (gasx_batch_run) against sequential single runs (GPU).

Every run of an ensemble must be bit-identical to gasx_run of the same
particles, gas, seed and run id: same Philox counters, same fixed-point
current, same field arithmetic, same append order.
"""
import numpy as np
import pytest

from conftest import log
from test_gpu_parity import wedge_particles

pytestmark = pytest.mark.gpu

NR, NZ = 24, 40


def ensemble_case(seed=0):
    rng = np.random.default_rng(seed)
    sizes = [3000, 0, 17, 1500, 4096, 800]                # uneven, one empty, one tiny
    dens = [np.full((NZ // 4, NR // 4), d, np.float32) for d in (2.5e25, 1e26, 2.5e24, 2.5e26, 5e25, 8e25)]
    dens[3][:, : NR // 8] *= 0.5                          # a non-uniform one
    parts = [wedge_particles(n, NR, NZ, 1e-5, 1e-5, 0.1, 100 + k, vmax=1.5e7, edge_frac=0.05)
             if n else np.zeros((6, 0), np.float32) for k, n in enumerate(sizes)]
    caps = [max(1, 2 * n) for n in sizes]
    return sizes, dens, parts, caps, rng


def run_single(params, part, cap, dens, n_steps, seed, run_id, secondaries):
    import gasx
    ctx = gasx.Context(params)
    e = gasx.ParticleSet(cap)
    if part.shape[1]:
        e.load(*part)
    f = gasx.FieldState(NR, NZ)
    gas = gasx.GasGrid(dens, 4e-5, 4e-5, temperature=300.0)
    ctx.run(e, f, gas, n_steps, collide=True, secondaries=secondaries, sort_every=0, seed=seed,
            run_id=run_id)
    c = ctx.counters()
    return e.state(), f.numpy(), c


def make_ensemble(params, dens, parts, caps):
    import gasx
    ctx = gasx.Context(params)
    ens = gasx.Ensemble(ctx, caps, dens, 4e-5, 4e-5)
    for r, p in enumerate(parts):
        if p.shape[1]:
            ens.load_run(r, *p)
    return ctx, ens


def assert_same(ens, r, ref_state, ref_fields):
    got = ens.run_state(r)
    assert got.shape == ref_state.shape, (r, got.shape, ref_state.shape)
    assert np.array_equal(got, ref_state, equal_nan=True), r
    er, ez, h = ref_fields
    assert np.array_equal(ens.er[r].cpu().numpy(), er), r
    assert np.array_equal(ens.ez[r].cpu().numpy(), ez), r
    assert np.array_equal(ens.hphi[r].cpu().numpy(), h), r


@pytest.mark.parametrize("secondaries,n_steps", [(False, 5), (True, 6)])
def test_batch_equals_sequential_runs(gpu, secondaries, n_steps):
    import gasx
    log(f"batch vs sequential secondaries={secondaries} n_steps={n_steps}")
    params = gasx.make_params(NR, NZ, fn=2e3)
    sizes, dens, parts, caps, _ = ensemble_case()
    ctx, ens = make_ensemble(params, dens, parts, caps)
    ens.run(n_steps, collide=True, secondaries=secondaries, seed=11, run_id=40)
    per_run = ens.counters()
    for r in range(len(sizes)):
        st, fl, c = run_single(params, parts[r], caps[r], dens[r], n_steps, 11, 40 + r, secondaries)
        assert_same(ens, r, st, fl)
        got = per_run[r]
        assert (got.n_elastic, got.n_ionize, got.n_excite, got.n_pushed, got.n_dropped_sec) == \
               (c.n_elastic, c.n_ionize, c.n_excite, c.n_pushed, c.n_dropped_sec), r
    assert sum(c.n_ionize for c in per_run) > 0


@pytest.mark.parametrize("secondaries,use_graph", [(False, False), (True, False), (True, True)])
def test_batch_with_ions_equals_sequential_runs(gpu, secondaries, use_graph):
    """Runs with an ion pool: each run's ions (the N2+ of its ionizations,
    appended in parent slot order) and electrons equal gasx_run's with an ion
    species; one run's ion capacity is too small, so its drops are counted."""
    import torch
    import gasx
    log(f"batch with ions secondaries={secondaries} graph={use_graph}")
    params = gasx.make_params(NR, NZ, fn=2e3)
    sizes, dens, parts, caps, _ = ensemble_case(4)
    icaps = [max(1, 3 * n) for n in sizes]
    icaps[3] = 5                                          # fills within the first step
    ctx = gasx.Context(params)
    ens = gasx.Ensemble(ctx, caps, dens, 4e-5, 4e-5, ion_capacities=icaps)
    for r, p in enumerate(parts):
        if p.shape[1]:
            ens.load_run(r, *p)
    ens.run(6, collide=True, secondaries=secondaries, seed=21, run_id=7, use_graph=use_graph)
    torch.cuda.synchronize()
    per_run = ens.counters()
    total_ions = 0
    for r in range(len(sizes)):
        c1 = gasx.Context(params)
        e = gasx.ParticleSet(caps[r])
        if parts[r].shape[1]:
            e.load(*parts[r])
        ions = gasx.ParticleSet(icaps[r])
        f = gasx.FieldState(NR, NZ)
        gas = gasx.GasGrid(dens[r], 4e-5, 4e-5, temperature=300.0)
        c1.run(e, f, gas, 6, collide=True, secondaries=secondaries, sort_every=0, seed=21, run_id=7 + r,
               ions=ions)
        c = c1.counters()
        assert_same(ens, r, e.state(), f.numpy())
        gi = ens.run_ions(r)
        ri = ions.state()
        assert gi.shape == ri.shape, (r, gi.shape, ri.shape)
        assert np.array_equal(gi, ri, equal_nan=True), r
        got = per_run[r]
        assert (got.n_ionize, got.n_dropped_ion, got.n_dropped_sec) == \
               (c.n_ionize, c.n_dropped_ion, c.n_dropped_sec), r
        total_ions += gi.shape[1]
    assert total_ions > 0
    assert per_run[3].n_dropped_ion > 0 and ens.run_ions(3).shape[1] == 5


def test_batch_graph_replays_match_eager(gpu):
    """One captured graph of the whole step sequence, replayed twice with
    eager work and synchronizations between the replays."""
    import torch
    import gasx
    log("batch graph replay")
    params = gasx.make_params(NR, NZ, fn=2e3)
    sizes, dens, parts, caps, _ = ensemble_case(1)
    ctx, eager = make_ensemble(params, dens, parts, caps)
    eager.run(7, collide=True, secondaries=True, seed=3, run_id=5)
    ref = [eager.run_state(r) for r in range(len(sizes))]
    ref_f = [(eager.er[r].cpu().numpy(), eager.ez[r].cpu().numpy(), eager.hphi[r].cpu().numpy())
             for r in range(len(sizes))]
    ctx2, g = make_ensemble(params, dens, parts, caps)
    init = (g.pool.clone(), g.counts.clone())
    for rep in range(2):
        g.pool.copy_(init[0]); g.counts.copy_(init[1])
        for t in (g.er, g.ez, g.hphi):
            t.zero_()
        torch.cuda.synchronize()
        g.run(7, collide=True, secondaries=True, seed=3, run_id=5, use_graph=True)
        torch.cuda.synchronize()
        _ = (g.pool.sum() + 1).item()                    # unrelated eager work between replays
        for r in range(len(sizes)):
            assert_same(g, r, ref[r], ref_f[r])
    assert ctx2.timing()["n_graph_launches"] == 0        # timing was not requested


def test_graph_replay_after_eager_run_uses_its_own_descriptors(gpu):
    """A captured graph, then an eager run of another description (other run
    ids and density bounds) on the same context, then a replay of the first:
    the replay must compute what an all-eager sequence computes, bit for bit
    (the eager run overwrites the device descriptors the graph reads)."""
    import torch
    import gasx
    log("batch graph -> eager (other run ids, n_max) -> graph")
    params = gasx.make_params(NR, NZ, fn=2e3)
    sizes, dens, parts, caps, _ = ensemble_case(3)
    dens_b = [d * 0.5 for d in dens]                      # other n_max per run
    kw_a = dict(collide=True, secondaries=True, seed=5, run_id=10)
    kw_b = dict(collide=True, secondaries=True, seed=6, run_id=500)

    def sequence(use_graph):
        ctx = gasx.Context(params)
        a = gasx.Ensemble(ctx, caps, dens, 4e-5, 4e-5)
        b = gasx.Ensemble(ctx, caps, dens_b, 4e-5, 4e-5)
        for r, p in enumerate(parts):
            if p.shape[1]:
                a.load_run(r, *p)
                b.load_run(r, *p)
        a.run(3, use_graph=use_graph, **kw_a)             # capture (or eager)
        b.run(3, use_graph=False, **kw_b)                 # eager: uploads b's descriptors
        a.run(3, use_graph=use_graph, **kw_a)             # replay of a's graph
        torch.cuda.synchronize()
        return ([a.run_state(r) for r in range(len(sizes))],
                [a.er[r].cpu().numpy() for r in range(len(sizes))],
                [b.run_state(r) for r in range(len(sizes))])

    ref_a, ref_er, ref_b = sequence(False)
    got_a, got_er, got_b = sequence(True)
    for r in range(len(sizes)):
        assert np.array_equal(got_a[r], ref_a[r], equal_nan=True), r
        assert np.array_equal(got_er[r], ref_er[r]), r
        assert np.array_equal(got_b[r], ref_b[r], equal_nan=True), r


def test_autotune_restores_state_and_picks_a_valid_shape(gpu):
    import gasx
    log("batch autotune")
    params = gasx.make_params(NR, NZ, fn=2e3)
    sizes, dens, parts, caps, _ = ensemble_case(2)
    ctx, ens = make_ensemble(params, dens, parts, caps)
    before = [ens.run_state(r) for r in range(len(sizes))]
    t = ens.autotune(trial_steps=2, collide=True, secondaries=True, seed=1)
    # particle grid: one or two generations of resident workgroups, or 4..32 per CU
    assert t["fdtd_ti"] >= 1 and t["fdtd_tk"] >= 1 and 1 <= t["particle_blocks_per_cu"] <= 64
    assert ens.tuning() == t
    for r in range(len(sizes)):
        assert np.array_equal(ens.run_state(r), before[r], equal_nan=True)
    # the tuned shape computes the same thing
    ens.run(4, collide=True, secondaries=True, seed=9, run_id=0)
    _, ref = make_ensemble(params, dens, parts, caps)
    ref.set_tuning(16, 32, 16)
    ref.run(4, collide=True, secondaries=True, seed=9, run_id=0)
    for r in range(len(sizes)):
        assert np.array_equal(ens.run_state(r), ref.run_state(r), equal_nan=True)
        assert np.array_equal(ens.er[r].cpu().numpy(), ref.er[r].cpu().numpy())


def test_batch_rejects_bad_descriptions(gpu):
    import ctypes as C
    import gasx
    from gasx import _abi
    params = gasx.make_params(NR, NZ)
    sizes, dens, parts, caps, _ = ensemble_case()
    ctx, ens = make_ensemble(params, dens, parts, caps)
    b = ens.struct()
    o = _abi.RunOpts(2, 1, 0, 5, 0, 0, 0, 0, 0)          # sort_every != 0
    assert ctx.lib.gasx_batch_run(ctx.ctx, C.byref(b), C.byref(o)) == _abi.E_ARG
    bad_base = np.array(ens.base, dtype=np.int64)
    bad_base[1] = bad_base[0] + 1                         # misaligned and overlapping
    b.base = bad_base.ctypes.data_as(C.POINTER(C.c_int64))
    o = _abi.RunOpts(2, 1, 0, 0, 0, 0, 0, 0, 0)
    assert ctx.lib.gasx_batch_run(ctx.ctx, C.byref(b), C.byref(o)) == _abi.E_ARG
    # ion pool ranges that overlap
    ens_i = gasx.Ensemble(ctx, caps, dens, 4e-5, 4e-5, ion_capacities=[64] * len(caps))
    b = ens_i.struct()
    bad_ib = np.zeros(len(caps), dtype=np.int64)
    b.ion_base = bad_ib.ctypes.data_as(C.POINTER(C.c_int64))
    assert ctx.lib.gasx_batch_run(ctx.ctx, C.byref(b), C.byref(o)) == _abi.E_ARG
