"""The reference-named entry points (gasx.reference_api) against the oracle (GPU)."""
import numpy as np
import pytest

import oracle
from conftest import log
from test_gpu_parity import rel, wedge_particles

pytestmark = pytest.mark.gpu


def test_electricmagneticfield6_matches_oracle(gpu):
    """electricmagneticfield6.m:1-190 in the reference's argument convention:
    MATLAB-shaped fields, N x 3 particles, delta/dt/Rp/Lp/PHIp/FN/cycle_time."""
    import gasx
    from gasx.reference_api import electricmagneticfield6
    log("reference_api.electricmagneticfield6")
    nr, nz, delta, dt, phip, fn = 24, 40, 1e-5, 2e-15, 0.1, 4e3
    se = wedge_particles(6000, nr, nz, delta, delta, phip, 21, edge_frac=0.0).astype(np.float64)
    si = wedge_particles(1500, nr, nz, delta, delta, phip, 22, vmax=1e3, edge_frac=0.0).astype(np.float64)
    # away from r = Rp, so absorption cannot hinge on rounding
    se = se[:, np.hypot(se[0], se[1]) < 0.9 * nr * delta]
    si = si[:, np.hypot(si[0], si[1]) < 0.9 * nr * delta]
    rng = np.random.default_rng(5)
    er0 = rng.standard_normal((nr, nz + 1)) * 1e3      # MATLAB shape NR x (NZ+1)
    ez0 = rng.standard_normal((nr + 1, nz)) * 1e3
    h0 = rng.standard_normal((nr, nz)) * 1.0
    out = electricmagneticfield6(se[:3].T, se[3:].T, si[:3].T, si[3:].T, np.zeros((0, 3)), np.zeros((0, 3)),
                                 er0, ez0, h0, delta, dt, nr * delta, nz * delta, phip, fn, 15)
    epos, evel, ipos, ivel, _, _, er, ez, h, ke = out
    assert er.shape == er0.shape and ez.shape == ez0.shape and h.shape == h0.shape
    p = gasx.make_params(nr, nz, dr=delta, dz=delta, phip=phip, ddt=5 * dt, n_sub=10, fn=fn)
    op = oracle.params_from(p)
    both = np.concatenate([se.astype(np.float32), si.astype(np.float32)], axis=1)
    q = np.concatenate([np.full(se.shape[1], -gasx.E_EM), np.full(si.shape[1], gasx.E_EM)])
    m = np.concatenate([np.full(se.shape[1], gasx.ME_EM), np.full(si.shape[1], gasx.M_ION)])
    o = oracle.State(both, q=q, m=m)
    oer, oez, oh, oke = oracle.em_step(op, o, er0.T.ravel(), ez0.T.ravel(), h0.T.ravel(), 3)
    e_er = rel(er.T.ravel(), oer)
    e_ez = rel(ez.T.ravel(), oez)
    e_h = rel(h.T.ravel(), oh)
    print(f"electricmagneticfield6 field rel err: Er {e_er:.2e} Ez {e_ez:.2e} H {e_h:.2e}")
    assert e_er < 1e-5 and e_ez < 1e-5 and e_h < 1e-5
    alive_e = o.dele[: se.shape[1]] == 0
    assert epos.shape == (int(alive_e.sum()), 3) and evel.shape == epos.shape
    o_r = np.hypot(o.v[0][: se.shape[1]][alive_e], o.v[1][: se.shape[1]][alive_e])
    assert np.max(np.abs(np.hypot(epos[:, 0], epos[:, 1]) - o_r)) < 1e-6 * nr * delta
    # kinetic energy in keV with the constants of electricmagneticfield6.m
    o_ke = oke[: se.shape[1]][alive_e] / (1000 * gasx.E_EM)
    assert np.allclose(ke, o_ke, rtol=2e-5, atol=1e-9)
    assert ipos.shape[0] == int((o.dele[se.shape[1]:] == 0).sum())
    with pytest.raises(ValueError):
        electricmagneticfield6(se[:3].T, se[3:].T, si[:3].T, si[3:].T, np.zeros((0, 3)), np.zeros((0, 3)),
                               er0[:, :-1], ez0, h0, delta, dt, nr * delta, nz * delta, phip, fn, 5)


def test_find_CS0_matches_oracle(gpu, raw, otab):
    from gasx.reference_api import find_CS0
    log("reference_api.find_CS0")
    e_kev = np.array([[0.0150, 0.0157, 0.03, 0.4], [1.0, 5.0, 20.0, 20.1]])
    se, si, sx = find_CS0(raw["t_el"], raw["t_sel"], raw["t_ion"], raw["t_exc"], e_kev)
    assert se.shape == e_kev.shape
    ose, osi, osx = oracle.cross_sections(otab, e_kev.ravel())
    for g, o in ((se, ose), (si, osi), (sx, osx)):
        g = g.ravel()
        nan = np.isnan(o)
        assert np.array_equal(np.isnan(g), nan)
        assert np.allclose(g[~nan], o[~nan], rtol=2e-5)
    assert si.ravel()[3] == pytest.approx(1.632e-20, rel=1e-5, abs=0)      # TotalIonCS_Nitrogen.csv:12


def test_electron_collision_matches_direct_call(gpu, raw):
    """electron_collision.m with the reference's phi-resolved cell arrays (one
    density per ring here, so the ring mean is exact) equals gasx_collide on
    the same Philox stream; secondaries and ions are appended."""
    import gasx
    from gasx.reference_api import electron_collision
    log("reference_api.electron_collision")
    nr, nz, dr, dz, phip = 20, 30, 1e-5, 1e-5, 0.1
    rmax = nr
    phi_div = np.ceil(np.arange(1, rmax + 1) / rmax * nr * dr * phip / dr).astype(np.int64)
    sum_cell = np.cumsum(phi_div)
    cellnumxy = int(sum_cell[-1])
    ring_n = np.linspace(2.5e26, 1e26, rmax)
    dens_cells = np.tile(np.repeat(ring_n, phi_div), nz)
    temp_cells = np.full(nz * cellnumxy, 300.0)
    st = wedge_particles(4000, nr, nz, dr, dz, phip, 31, vmax=1.5e7, edge_frac=0.0).astype(np.float64)
    pos, vel = st[:3].T.copy(), st[3:].T.copy()
    n = pos.shape[0]
    dummy = np.zeros(n)
    ep, ev, ip, iv = electron_collision(pos, vel, np.zeros((0, 3)), np.zeros((0, 3)), raw["t_el"], raw["t_sel1"],
                                        raw["t_del"], raw["t_sdel"], raw["t_exc"], dummy, sum_cell, cellnumxy,
                                        dens_cells, temp_cells, phip, phi_div, dr, dz, dummy, dummy, dummy,
                                        1e-13, seed=77, step=3)
    n_sec = ep.shape[0] - n
    assert n_sec > 0 and ip.shape[0] == n_sec and iv.shape == ip.shape
    # the same call through the direct API
    p = gasx.make_params(2, 2, phip=phip)
    ctx = gasx.Context(p)
    e = gasx.ParticleSet(2 * n).load(*st.astype(np.float32))
    ions = gasx.ParticleSet(2 * n, qnum=1.0, mass=gasx.M_ION)
    gas = gasx.GasGrid(np.tile(ring_n, (nz, 1)).astype(np.float32), dr, dz, temperature=np.full((nz, rmax), 300.0, np.float32))
    ctx.collide(e, gas, 1e-13, 77, 3, ions=ions)
    d = e.state()
    assert d.shape[1] == ep.shape[0]
    assert np.array_equal(d[:3].T, ep.astype(np.float32)) and np.array_equal(d[3:].T, ev.astype(np.float32))
    # secondaries sit at their parents' positions (electron_collision.m:68-70)
    parents = {tuple(r) for r in pos.astype(np.float32)}
    assert all(tuple(r) in parents for r in ep[n:].astype(np.float32))
    assert np.array_equal(ions.state()[:3].T, ip.astype(np.float32))
