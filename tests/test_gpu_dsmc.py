"""GPU parity of the DSMC neutral-gas step (gasx_dsmc_*) against the oracle's
restatement of main2.m:247-264 on the same FP32-stored inputs and Philox
streams.  The two compute in FP64 with the same operations; they can differ
only where atan2, sin, cos or pow round their last FP64 bit differently, so
the comparisons are bit for bit with a small allowance for such cases.
"""
import math

import numpy as np
import pytest
import torch

import gasx
import oracle as O
from gasx.dsmc import CellMap, MoleculeSet, dsmc_params, seed_molecules

pytestmark = pytest.mark.gpu

KB = 1.38064853e-23
M = 4.65e-26
D = 3.7e-10


def main2_map(lp=0.03):
    n = 10.0 / KB / 300.0                                         # main2.m:12,18-19
    lam = 1 / (math.sqrt(2) * math.pi * n * D ** 2)               # main2.m:20
    return CellMap(int(1e-2 / lam) + 1, int(lp / lam) + 1, 1e-2, lp, 0.5, lam), n


def omap(cm: CellMap):
    return O.cell_map(cm.n_rings, cm.n_layers, cm.rp, cm.lp, cm.phip, cm.delta)


@pytest.fixture(scope="module")
def ctx(gpu):
    return gasx.Context(gasx.make_params(16, 16), device=0)


def to64(st32):
    return np.ascontiguousarray(st32.astype(np.float64))


def near_cell_edge(cm: CellMap, st, tol=1e-9):
    pd, s0, dp, vc, nxy = O.cell_map_build(omap(cm))
    r = np.hypot(st[0], st[1])
    ph = np.arctan2(st[1], st[0])
    ir = np.clip(np.ceil(r / cm.dr).astype(int), 1, cm.n_rings) - 1
    def frac(a):
        return np.abs(a - np.round(a))
    return (frac(r / cm.dr) < tol) | (frac(st[2] / cm.dz) < tol) | (frac(ph / dp[ir]) < tol)


def test_cells_match_oracle(ctx):
    cm, _ = main2_map()
    st = seed_molecules(300_000, cm, seed=1)
    mol = MoleculeSet(st.shape[1]).load(st)
    g = ctx.dsmc.cells(mol, cm).cpu().numpy()
    o = O.dsmc_cells(omap(cm), to64(st))
    bad = g != o
    assert bad.sum() <= 3
    assert np.all(near_cell_edge(cm, to64(st)[:, bad]))


def wall_state(cm: CellMap, n, seed):
    """Molecules near every wall and face, moving into them."""
    rng = np.random.default_rng(seed)
    st = seed_molecules(n, cm, seed=seed).astype(np.float64)
    k = n // 5
    r = cm.rp - rng.uniform(0, 1e-4, k)
    ph = rng.uniform(0, cm.phip, k)
    st[0, :k], st[1, :k] = r * np.cos(ph), r * np.sin(ph)
    st[3, :k] = np.abs(st[3, :k]) * 3 * np.cos(ph)
    st[4, :k] = np.abs(st[4, :k]) * 3 * np.sin(ph)
    st[2, k:2 * k] = cm.lp - rng.uniform(0, 1e-4, k)
    st[5, k:2 * k] = np.abs(st[5, k:2 * k]) * 3
    st[2, 2 * k:3 * k] = rng.uniform(0, 1e-4, k)
    st[5, 2 * k:3 * k] = -np.abs(st[5, 2 * k:3 * k]) * 3
    return st.astype(np.float32)


def test_move_matches_oracle(ctx):
    cm, _ = main2_map()
    st = wall_state(cm, 200_000, 2)
    mol = MoleculeSet(st.shape[1]).load(st)
    ctx.dsmc.counters(reset=True)
    ctx.dsmc.move(mol, cm, 2e-7)
    g = mol.state()
    o = to64(st)
    oc = O.dsmc_move(omap(cm), 2e-7, o, round32=True)
    c = ctx.dsmc.counters(reset=True)
    assert [c.n_boundary, c.n_wall_r, c.n_wall_z, c.n_wrap] == list(oc)
    assert c.n_wall_r > 10_000 and c.n_wall_z > 20_000 and c.n_wrap > 1000
    diff = g[:6].astype(np.float64) != o[:6]
    rows = diff.any(0)
    assert rows.sum() <= 20
    # the rest differ by the last FP32 bit (a libm-rounded atan2/cos/sin)
    gi, oi = g[:6][:, rows].view(np.int32), o[:6][:, rows].astype(np.float32).view(np.int32)
    assert np.all(np.abs(gi.astype(np.int64) - oi) <= 2)
    assert np.array_equal(g[6], st[6])


def collide_case(cm: CellMap, n, fn, seed):
    st = seed_molecules(n, cm, seed=seed)
    p = dsmc_params(2e-7, fn)
    op = O.dsmc_params(p.dt, p.fn, p.mass, p.diameter, p.sc_max, p.omega, p.zeta, p.alpha, p.min_attempts)
    return st, p, op


def test_collide_matches_oracle(ctx):
    cm, _ = main2_map(lp=0.004)
    # about 3000 molecules per cell and main2's collision rate per molecule
    st, p, op = collide_case(cm, 400_000, 3.4e8, 5)
    mol = MoleculeSet(st.shape[1]).load(st)
    ctx.dsmc.counters(reset=True)
    ctx.dsmc.collide(mol, cm, p, seed=77, step=3)
    c = ctx.dsmc.counters(reset=True)
    g = mol.state()
    o = to64(st)
    cell = O.dsmc_cells(omap(cm), o)
    oc = O.dsmc_collide(omap(cm), op, o, seed=77, step=3, cell=cell, round32=True)
    assert c.n_attempts == oc[0] and c.n_attempts > 100_000
    assert abs(c.n_collisions - oc[1]) <= 3 and c.n_collisions > 10_000
    assert abs(c.n_sc_exceed - oc[2]) <= 3
    same = (g[3:].astype(np.float64) == o[3:]).all(0)
    # cells whose molecules all agree bit for bit
    ncell = cm.n_layers * cm.info(ctx.lib)[1]
    bad_cells = np.unique(cell[~same])
    assert bad_cells.size <= max(2, ncell // 100), (bad_cells.size, ncell)
    assert np.array_equal(g[:3], st[:3])


def test_step_equals_move_then_collide_and_matches_oracle_with_reorder(ctx):
    cm, _ = main2_map(lp=0.004)
    st, p, op = collide_case(cm, 200_000, 6.8e8, 9)
    a = MoleculeSet(st.shape[1]).load(st)
    b = MoleculeSet(st.shape[1]).load(st)
    ctx.dsmc.step(a, cm, p, seed=5, step=0)
    ctx.dsmc.move(b, cm, p.dt)
    ctx.dsmc.collide(b, cm, p, seed=5, step=0)
    assert np.array_equal(a.state(), b.state())
    # two steps with reorder against the oracle's move, collide, stable cell sort
    p.reorder = 1
    g = MoleculeSet(st.shape[1]).load(st)
    o = to64(st)
    om = omap(cm)
    for k in range(2):
        ctx.dsmc.step(g, cm, p, seed=5, step=k)
        O.dsmc_move(om, p.dt, o, round32=True)
        O.dsmc_collide(om, op, o, seed=5, step=k, round32=True)
        o = O.dsmc_reorder(om, o)
    gs = g.state().astype(np.float64)
    assert gs.shape == o.shape
    gc = O.dsmc_cells(om, gs)
    assert np.all(np.diff(gc) >= 0)                       # cell order
    same = (gs == o).all(0)
    assert same.mean() > 0.98, same.mean()


def test_sample_matches_oracle(ctx):
    cm, _ = main2_map(lp=0.004)
    st = seed_molecules(300_000, cm, seed=11)
    mol = MoleculeSet(st.shape[1]).load(st)
    fn = 1.7e8
    s = ctx.dsmc.sample(mol, cm, fn, M)
    o = O.dsmc_sample(omap(cm), to64(st), fn, M)
    for k in ("cell_density", "cell_temperature", "rz_density", "rz_temperature"):
        gv = s[k].cpu().numpy().reshape(o[k].shape)
        assert np.allclose(gv, o[k], rtol=2e-7, atol=0), k


def test_dsmc_rejects_bad_arguments(ctx):
    cm, _ = main2_map()
    mol = MoleculeSet(16).load(seed_molecules(16, cm, seed=0))
    p = dsmc_params(2e-7, 1e8)
    p.zeta = 1.0
    with pytest.raises(gasx._abi.GasxError):
        ctx.dsmc.step(mol, cm, p, seed=0, step=0)
    bad = CellMap(0, 4, 1e-2, 0.03, 0.5, 1e-3)
    with pytest.raises(gasx._abi.GasxError):
        ctx.dsmc.move(mol, bad, 1e-7)


def test_main2_scale_conserves_energy_and_keeps_the_gas(ctx):
    """main2.m's case: 5000 molecules per cell (1.44e7), 10 Pa, 300 K, three
    steps of 0.2 us with the cell reorder."""
    cm, n_gas = main2_map()
    _, nxy = cm.info(ctx.lib)
    ncell = nxy * cm.n_layers
    n = 5000 * ncell                                             # main2.m:45
    v = 0.5 * cm.rp ** 2 * cm.phip * cm.lp
    fn = n_gas * v / n                                           # main2.m:47
    st = seed_molecules(n, cm, seed=3)
    mol = MoleculeSet(n).load(st)
    p = dsmc_params(2e-7, fn, reorder=True)
    d = mol.data[:, :n].double()
    e0 = (0.5 * M * (d[3:6] ** 2).sum() + d[6].sum()).item()
    ctx.dsmc.counters(reset=True)
    for k in range(3):
        ctx.dsmc.step(mol, cm, p, seed=1, step=k)
    torch.cuda.synchronize()
    c = ctx.dsmc.counters(reset=True)
    assert mol.n == n
    d = mol.data[:, :n].double()
    e1 = (0.5 * M * (d[3:6] ** 2).sum() + d[6].sum()).item()
    assert e1 == pytest.approx(e0, rel=2e-6)
    # attempts: floor(SCmax FN N (N-1) / VC / 2 dt) per cell; accepted ~ pi d^2 <cr> / SCmax
    cr = math.sqrt(16 * KB * 300 / (math.pi * M))
    assert c.n_collisions / c.n_attempts == pytest.approx(math.pi * D ** 2 * cr / p.sc_max, rel=0.03)
    nu = n_gas * math.sqrt(2) * math.sqrt(8 * KB * 300 / (math.pi * M)) * math.pi * D ** 2   # main2.m:51
    assert 2 * c.n_collisions / (3 * n) == pytest.approx(nu * p.dt, rel=0.05)
    s = ctx.dsmc.sample(mol, cm, fn, M)
    assert s["rz_density"].mean().item() == pytest.approx(n_gas, rel=0.01)
    assert s["rz_temperature"].mean().item() == pytest.approx(300.0, rel=0.05)
