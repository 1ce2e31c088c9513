"""Photoionization source on the GPU (gasx_photoionize) against the FP64
oracle (orc_photoionize) on the same Philox streams.

The attenuation runs in FP64 with the oracle's operation order, so the event
counts per cell are equal exactly; events are emitted in the same order, so
every photoelectron, Auger electron and ion is compared one to one (the GPU
stores FP32: positions and velocities agree to FP32 rounding).  The depleted
gas agrees to FP32 rounding of the FP64 update.
"""
import numpy as np
import pytest

import oracle
from conftest import log

pytestmark = pytest.mark.gpu

GNR, GNZ, D = 64, 128, 1e-5


@pytest.fixture(scope="module")
def tabs():
    from gasx.tables import source_tables
    return source_tables()


def gas_arrays(seed=0):
    rng = np.random.default_rng(seed)
    dens = (2.5e25 * (1 + 0.2 * rng.standard_normal((GNZ, GNR)))).astype(np.float32)
    temp = (300.0 + 50.0 * rng.random((GNZ, GNR))).astype(np.float32)
    return dens, temp


def close(a, b, scale, rel):
    return np.all(np.abs(a.astype(np.float64) - b) <= rel * scale)


def test_photoionize_matches_oracle(gpu, tabs):
    import torch
    import gasx
    from gasx.core import M_ION
    log("photoionize vs oracle: 64x128 gas, FN=3, 1 keV, depletion")
    p = gasx.make_params(GNR, GNZ, fn=3.0)
    ctx = gasx.Context(p)
    dens, temp = gas_arrays()
    gas = gasx.GasGrid(dens, D, D, temperature=temp)
    n0 = 1000
    e = gasx.ParticleSet(400_000).load(*np.zeros((6, n0), np.float32))
    ions = gasx.ParticleSet(200_000, qnum=2.0, mass=M_ION)
    res = ctx.photoionize(e, gas, 1.0, 1.5e6, 150e-6, ions=ions, seed=9, pulse=4, run=1, deplete=True, tables=tabs)
    torch.cuda.synchronize()
    op = oracle.params_from(p)
    o = oracle.photoionize(op, dens.astype(np.float64), D, D, 1.0, 1.5e6, 150e-6, tabs,
                           temp=temp.astype(np.float64), deplete=True, seed=9, pulse=4, run=1)
    info = o["info"]
    n = info["n_events"]
    print(res, info)
    assert res["n_events"] == n > 50_000
    assert res["n_rings"] == info["n_rings"] and res["auger"] == info["auger"] == 1
    assert res["sigma_pe"] == info["sigma_pe"]
    assert res["photons_left"] == pytest.approx(info["photons_left"], rel=1e-12, abs=0)
    assert res["n_electrons"] == 2 * n and res["n_ions"] == n and res["n_dropped_electrons"] == 0
    st = e.state()
    assert st.shape[1] == n0 + 2 * n
    pe, ae = st[:, n0:n0 + n], st[:, n0 + n:]
    ist = ions.state()
    for got, ref in ((pe, o["pe"]), (ae, o["ae"]), (ist, o["ion"])):
        rscale = np.abs(ref[:3]).max()
        vscale = np.linalg.norm(ref[3:], axis=0)
        assert close(got[:3], ref[:3], rscale, 2e-7)
        assert close(got[3:], ref[3:], vscale, 4e-7)
    # depleted gas, in place on the device grids
    gd, gt = gas.density.cpu().numpy(), gas.temperature.cpu().numpy()
    assert not np.array_equal(gd, dens)
    assert np.allclose(gd, o["dens"], rtol=2e-7, atol=0)
    assert np.allclose(gt, o["temp"], rtol=1e-6, atol=0)


def test_photoionize_capacity_and_valence_shell(gpu, tabs):
    import gasx
    log("photoionize capacity crossing; valence photon without ions")
    p = gasx.make_params(GNR, GNZ, fn=3.0)
    ctx = gasx.Context(p)
    dens, _ = gas_arrays(1)
    gas = gasx.GasGrid(dens, D, D, temperature=300.0)
    op = oracle.params_from(p)
    o = oracle.photoionize(op, dens.astype(np.float64), D, D, 1.0, 1.5e6, 150e-6, tabs, seed=2)
    n = o["info"]["n_events"]
    cap = n + n // 2                                          # every photoelectron, half the Auger electrons
    e = gasx.ParticleSet(cap)
    res = ctx.photoionize(e, gas, 1.0, 1.5e6, 150e-6, seed=2, tables=tabs)
    assert res["n_events"] == n and res["n_electrons"] == cap and res["n_dropped_electrons"] == 2 * n - cap
    st = e.state()
    assert close(st[3:, :n], o["pe"][3:], np.linalg.norm(o["pe"][3:], axis=0), 4e-7)
    assert close(st[3:, n:], o["ae"][3:, : cap - n], np.linalg.norm(o["ae"][3:, : cap - n], axis=0), 4e-7)
    # below the K edge: one electron per event, no Auger electron
    e2 = gasx.ParticleSet(2 * n)
    res2 = ctx.photoionize(e2, gas, 0.3, 1.5e6, 150e-6, seed=3, tables=tabs)
    o2 = oracle.photoionize(op, dens.astype(np.float64), D, D, 0.3, 1.5e6, 150e-6, tabs, seed=3)
    assert res2["auger"] == 0 and res2["n_events"] == o2["info"]["n_events"] == res2["n_electrons"]
    assert close(e2.state()[3:], o2["pe"][3:], np.linalg.norm(o2["pe"][3:], axis=0), 4e-7)
    from gasx._abi import GasxError
    with pytest.raises(GasxError):
        ctx.photoionize(e2, gas, 0.01, 1.5e6, 150e-6, tables=tabs)   # below every binding energy


def test_photoelectrons_feed_the_transport_loop(gpu, tabs):
    """The source's electrons are the production loop's input (main3.m:134
    then :154): a few steps with collisions and secondaries from them."""
    import gasx
    log("photoionize then gasx_run")
    nr, nz = 64, 128
    p = gasx.make_params(nr, nz, fn=3.0)
    ctx = gasx.Context(p)
    gas = gasx.GasGrid.uniform(2.5e25, nr * D, nz * D, temperature=300.0)
    e = gasx.ParticleSet(1_000_000)
    res = ctx.photoionize(e, gas, 1.0, 1.5e6, 150e-6, seed=4, tables=tabs)
    assert res["n_electrons"] > 100_000
    f = gasx.FieldState(nr, nz)
    ctx.counters(reset=True)
    ctx.run(e, f, gas, 10, collide=True, secondaries=True, sort_every=5, seed=5)
    c = ctx.counters()
    assert c.n_bound_violations == 0 and c.n_ionize > 0 and c.n_elastic > 0
    er, ez, h = f.numpy()
    assert np.isfinite(er).all() and np.abs(er).max() > 0
