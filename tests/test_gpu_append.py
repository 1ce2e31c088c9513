"""Secondary append at production-like capacities (GPU).

Up to 32768 staging segments (8.4e6 slots) the append is one launch whose
workgroups sum the counts of all earlier segments; above, a scan launch in
groups of 1024 segments (2^18 slots) and a copy launch.  The tests here use
populations above 3 * 2^18 slots, and the order test runs both paths:

* gasx_collide: every secondary is placed at its parent's position with the
  velocity the collision gave it, and the appended block must list the
  parents in increasing slot order (electron_collision.m:68-75 appends the
  secondaries of the logical mask in index order);
* gasx_run with and without an ion species: the ions' append runs first and
  shares the scan buffers with the electrons' own append, which must give
  the same electrons and fields bit for bit.
"""
import numpy as np
import pytest

from conftest import log
from test_gpu_parity import wedge_particles

pytestmark = pytest.mark.gpu

NR, NZ = 32, 64


@pytest.mark.parametrize("cap", [1_200_000, 9_000_000])   # one launch; scan + copy (35157 segments)
def test_collide_appends_in_parent_slot_order_across_scan_groups(gpu, cap):
    import gasx
    n = 900_000                                             # 3516 staging segments: 4 groups
    log(f"append order n={n} cap={cap}")
    params = gasx.make_params(NR, NZ, fn=1e3)
    ctx = gasx.Context(params)
    part = wedge_particles(n, NR, NZ, 1e-5, 1e-5, 0.1, 77, vmax=2.5e7, edge_frac=0.0)
    e = gasx.ParticleSet(cap).load(*part)
    gas = gasx.GasGrid.uniform(1e26, NR * 1e-5, NZ * 1e-5)
    ctx.counters(reset=True)
    ctx.collide(e, gas, 1e-14, seed=5, step=2, run=0)
    c = ctx.counters()
    st = e.state()
    n_sec = st.shape[1] - n
    assert n_sec == c.n_ionize - c.n_dropped_sec and n_sec > 20_000
    # collisions change velocities only: a parent keeps its position
    assert np.array_equal(st[:3, :n], part[:3])
    key = {tuple(v): i for i, v in enumerate(part[:3].T.view(np.uint32).reshape(n, 3).tolist())}
    sec_pos = st[:3, n:].T.copy().view(np.uint32).reshape(n_sec, 3)
    parents = np.array([key[tuple(v)] for v in sec_pos.tolist()])
    assert np.all(np.diff(parents) > 0)                   # strictly increasing parent slots
    # every parent's speed dropped by at least the ionization energy (15.6 eV)
    v0 = np.linalg.norm(part[3:, parents].astype(np.float64), axis=0)
    v1 = np.linalg.norm(st[3:, parents].astype(np.float64), axis=0)
    assert np.all(v1 < v0)
    # and staged segments beyond the first scan group hold parents
    assert parents.max() > 3 * 1024 * 256


def test_ion_append_leaves_the_electron_append_unchanged(gpu):
    import gasx
    from gasx.core import M_ION
    n, cap = 800_000, 1_000_000                           # 3907 segments: 4 scan groups
    log(f"run with and without ions n={n} cap={cap}")
    params = gasx.make_params(NR, NZ, fn=1e2)
    part = wedge_particles(n, NR, NZ, 1e-5, 1e-5, 0.1, 78, vmax=2.5e7, edge_frac=0.05)
    gas_d = np.full((1, 1), 1e26, np.float32)
    out = []
    for with_ions in (False, True):
        ctx = gasx.Context(params)
        e = gasx.ParticleSet(cap).load(*part)
        f = gasx.FieldState(NR, NZ)
        gas = gasx.GasGrid(gas_d, NR * 1e-5, NZ * 1e-5)
        ions = gasx.ParticleSet(cap, qnum=1.0, mass=M_ION) if with_ions else None
        ctx.run(e, f, gas, 4, collide=True, secondaries=True, sort_every=0, seed=8, ions=ions)
        c = ctx.counters()
        out.append((e.state(), f.numpy(), c, ions.n if ions is not None else None))
    (s0, f0, c0, _), (s1, f1, c1, n_ions) = out
    assert c0.n_ionize == c1.n_ionize > 20_000
    assert n_ions == c1.n_ionize - c1.n_dropped_ion
    assert s0.shape == s1.shape and s0.shape[1] > n
    assert np.array_equal(s0, s1, equal_nan=True)
    for a, b in zip(f0, f1):
        assert np.array_equal(a, b)


def test_capacity_crossing_keeps_the_first_secondaries(gpu):
    """With room for k secondaries, gasx_collide keeps exactly the first k of
    what a run with room for all appends (parent slot order) and counts the
    rest as dropped."""
    import gasx
    n = 700_000
    part = wedge_particles(n, NR, NZ, 1e-5, 1e-5, 0.1, 79, vmax=2.5e7, edge_frac=0.0)
    gas = gasx.GasGrid.uniform(1e26, NR * 1e-5, NZ * 1e-5)
    params = gasx.make_params(NR, NZ, fn=1e3)
    outs = []
    for cap in (2 * n, n + 30_000):
        log(f"capacity crossing n={n} cap={cap}")
        ctx = gasx.Context(params)
        e = gasx.ParticleSet(cap).load(*part)
        ctx.collide(e, gas, 1e-14, seed=6, step=0)
        outs.append((e.state(), ctx.counters()))
    (big, cb), (small, cs) = outs
    assert cb.n_ionize == cs.n_ionize > 30_000 and cb.n_dropped_sec == 0
    assert small.shape[1] == n + 30_000 and cs.n_dropped_sec == cs.n_ionize - 30_000
    assert np.array_equal(small, big[:, : n + 30_000])
