"""Pins of the oracle's photoionization source (orc_photoionize, restating
INCIDENCE1.m:27-158 with GUASS.m, findEShell.m, findSigmaPE.m,
findPEangle.m and findV.m) by what the mathematics and the data fix:

* the beam's ring weights against a numerical quadrature of the Gaussian
  beam's radial density (GUASS.m integrates x exp(-x^2 / 2 s^2) per ring);
* sigma_pe against the 14_MAC.csv node at 1 keV;
* Beer-Lambert: photons conserved (absorbed = emitted x FN within the
  rounding) and the transmitted fraction against exp(-n sigma L);
* the photoelectron's polar angle against the Sauter cross section
  d sigma / d Omega ~ sin^2 t / (1 - b cos t)^4 [1 + g (g-1)(g-2)(1 - b cos t) / 2];
* speeds against the relativistic closed form, the Auger angle law;
* momentum: the ion's velocity minus the recoil is a Maxwellian of the cell
  temperature; the depleted gas loses exactly FN molecules per event and
  their thermal energy.
"""
import math

import numpy as np
import pytest
from scipy import integrate, stats

import oracle

NR, NZ = 64, 128
DR = DZ = 1e-5


@pytest.fixture(scope="module")
def src_tables():
    from gasx.tables import source_tables
    return source_tables()


def params(fn=1.0):
    import gasx
    return oracle.params_from(gasx.make_params(NR, NZ, fn=fn))


@pytest.fixture(scope="module")
def pulse(src_tables):
    p = params(fn=3.0)
    dens = np.full((NZ, NR), 2.5e25)
    temp = np.full((NZ, NR), 300.0)
    temp[:, :5] = 450.0                                   # a hotter core
    r = oracle.photoionize(p, dens, DR, DZ, 1.0, 1.2e6, 150e-6, src_tables, temp=temp, deplete=True, seed=5,
                           pulse=2)
    return p, dens, temp, r


def test_beam_ring_weights_match_quadrature(src_tables):
    """In an optically thin slab every ring keeps its photons, so the events
    per ring are round(n N_i sigma dz / FN) with N_i the ring's share of the
    beam: the shares against a quadrature of the beam's radial density
    x exp(-x^2 / 2 s^2), s = laser_r / 3 (GUASS.m:3-15)."""
    p = params(fn=1.0)
    n, n0 = 1e10, 1e19                                     # n sigma dz ~ 8e-14: no attenuation
    r = oracle.photoionize(p, np.full((1, NR), n), DR, 1.0, 1.0, n0, 150e-6, src_tables, max_events=1 << 20)
    nr = r["info"]["n_rings"]
    assert nr == round(150e-6 / DR)
    cnt = np.bincount(r["cell"], minlength=nr).astype(float)
    s = 150e-6 / 3
    w = np.array([integrate.quad(lambda x: x * math.exp(-x * x / (2 * s * s)), i * DR, (i + 1) * DR,
                                 epsabs=0, epsrel=1e-13)[0] for i in range(nr)])
    expect = n * (n0 * w / w.sum()) * r["info"]["sigma_pe"] * 1.0 / p.fn
    assert cnt.min() > 1000
    assert np.all(np.abs(cnt - expect) <= 0.5 + 1e-6 * expect)


def test_sigma_pe_at_table_node(pulse):
    _, _, _, r = pulse
    # 14_MAC.csv: 1.00E-03 MeV -> 3.31E+03 cm^2/g; sigma = MAC m_a / N_A * 1e-4
    assert r["info"]["sigma_pe"] == pytest.approx(3.31e3 * 14 / 6.02e23 * 1e-4, rel=1e-12, abs=0)
    assert r["info"]["electron_kev"] == pytest.approx(1.0 - 0.4099, rel=1e-12, abs=0)
    assert r["info"]["auger"] == 1


def test_beer_lambert_attenuation(src_tables):
    p = params(fn=1e3)
    n0, L = 2.5e25, NZ * DZ
    r = oracle.photoionize(p, np.full((NZ, NR), n0), DR, DZ, 1.0, 1e9, 150e-6, src_tables, max_events=1 << 22)
    sig = r["info"]["sigma_pe"]
    left = r["info"]["photons_left"]
    k = n0 * sig * DZ                                      # per-slab absorption, first order
    # (1 - k)^NZ against exp(-n sigma L): differ by about NZ k^2 / 2
    assert left / 1e9 == pytest.approx(math.exp(-n0 * sig * L), rel=NZ * k * k, abs=0)
    assert abs(left / 1e9 - math.exp(-n0 * sig * L)) > 0          # the explicit slab update, not exp
    # photons are conserved: each cell rounds dN / FN to the nearest integer
    ncell = r["info"]["n_rings"] * NZ
    absorbed = 1e9 - left
    assert abs(r["info"]["n_events"] * 1e3 - absorbed) <= 0.5 * 1e3 * ncell


@pytest.mark.parametrize("kev", [0.5901, 5.0])
def test_sauter_angle_distribution(kev):
    """cos(theta) of the sampler against the Sauter cross section (chi-square
    over 40 bins of cos theta)."""
    dE = kev * 1.6e-16
    me_c2 = 9.11e-31 * 9e16
    g = 1 + dE / me_c2
    b = math.sqrt(dE * (dE + 2 * me_c2)) / (dE + me_c2)
    L = oracle.lib()
    n = 60000
    ct = np.array([1.0 - L.orc_sauter_v(dE, 3, e, 1, 0) for e in range(n)])
    assert ct.min() >= -1 and ct.max() <= 1
    dens = lambda c: (1 - c * c) / (1 - b * c) ** 4 * (1 + 0.5 * g * (g - 1) * (g - 2) * (1 - b * c))
    edges = np.linspace(-1, 1, 41)
    mass = np.array([integrate.quad(dens, edges[i], edges[i + 1])[0] for i in range(40)])
    exp_ = n * mass / mass.sum()
    obs = np.histogram(ct, edges)[0]
    keep = exp_ >= 20
    o, e = obs[keep], exp_[keep] * obs[keep].sum() / exp_[keep].sum()
    assert stats.chisquare(o, e).pvalue > 1e-3


def test_speeds_and_auger_angles(pulse):
    _, _, _, r = pulse
    c, m = 3e8, 9.109e-31

    def find_v(ke):
        gm = 1 + ke / (m * c * c)
        return c * math.sqrt(1 - 1 / gm ** 2)

    vpe = np.linalg.norm(r["pe"][3:], axis=0)
    vae = np.linalg.norm(r["ae"][3:], axis=0)
    assert np.allclose(vpe, find_v(0.5901 * 1.6e-16), rtol=1e-12)
    assert np.allclose(vae, find_v(0.383 * 1.6e-16), rtol=1e-12)
    # Auger polar angle uniform in [0, pi] (INCIDENCE1.m:138), azimuth uniform
    th = np.arccos(np.clip(r["ae"][5] / vae, -1, 1))
    assert stats.kstest(th / math.pi, "uniform").pvalue > 1e-3
    ch = np.mod(np.arctan2(r["ae"][4], r["ae"][3]), 2 * math.pi)
    assert stats.kstest(ch / (2 * math.pi), "uniform").pvalue > 1e-3


def test_positions_lie_in_their_cells(pulse):
    _, _, _, r = pulse
    nr = r["info"]["n_rings"]
    ring, slab = r["cell"] % nr, r["cell"] // nr
    x, y, z = r["pe"][:3]
    rr = np.hypot(x, y)
    assert np.all(rr >= ring * DR - 1e-15) and np.all(rr <= (ring + 1) * DR + 1e-15)
    assert np.all(z >= slab * DZ) and np.all(z <= (slab + 1) * DZ)
    ph = np.arctan2(y, x)
    assert np.all(ph >= 0) and np.all(ph <= 0.1 + 1e-15)
    # uniform in r inside the ring (INCIDENCE1.m:67), in z inside the slab, in phi over the wedge
    assert stats.kstest(rr / DR - ring, "uniform").pvalue > 1e-3
    assert stats.kstest(z / DZ - slab, "uniform").pvalue > 1e-3
    assert stats.kstest(ph / 0.1, "uniform").pvalue > 1e-3
    assert np.array_equal(r["ae"][:3], r["pe"][:3]) and np.array_equal(r["ion"][:3], r["pe"][:3])
    # slab by slab, rings outward
    assert np.all(np.diff(r["cell"]) >= 0)


def test_recoil_leaves_a_maxwellian(pulse):
    """m_ion v_ion = p_photon z^ + m_ion v_th - m_e (v_pe + v_ae): the
    residual v_th is N(0, kB T / m_ion) per component at the cell's T."""
    p, dens, temp, r = pulse
    m_mol, m_e, kb = 2 * 14 * 1.6726e-27, 9.11e-31, 1.38064853e-23
    pgam = 1.0 * 1.6e-16 / 3e8
    vth = r["ion"][3:] - (np.array([[0], [0], [pgam]]) - m_e * (r["pe"][3:] + r["ae"][3:])) / m_mol
    assert np.allclose(vth, r["vth"], rtol=1e-6, atol=1e-9)
    nr = r["info"]["n_rings"]
    T = np.where(r["cell"] % nr < 5, 450.0, 300.0)
    u = vth / np.sqrt(kb * T / m_mol)
    n = u.shape[1]
    for comp in u:
        assert abs(comp.mean()) < 5 / math.sqrt(n)
        assert stats.kstest(comp, "norm").pvalue > 1e-3
    assert stats.kstest(np.sum(u ** 2, axis=0), stats.chi2(3).cdf).pvalue > 1e-3


def test_depletion_removes_molecules_and_their_thermal_energy(pulse, src_tables):
    p, dens, temp, r = pulse
    nr = r["info"]["n_rings"]
    cnt = np.bincount(r["cell"], minlength=nr * NZ).reshape(NZ, nr)
    i = np.arange(nr)
    V = p.phip / 2 * ((i + 1) ** 2 - i ** 2) * DR * DR * DZ
    n1 = r["dens"][:, :nr]
    # molecules: n' V = n V - FN count, exactly up to rounding
    assert np.allclose(n1 * V, dens[:, :nr] * V - p.fn * cnt, rtol=1e-13, atol=0)
    assert np.array_equal(r["dens"][:, nr:], dens[:, nr:])
    # energy: 3/2 kB (n V T - n' V T') = sum 1/2 FN m v_th^2 over the cell's ions
    m_mol, kb = 2 * 14 * 1.6726e-27, 1.38064853e-23
    ke = np.zeros(nr * NZ)
    np.add.at(ke, r["cell"], 0.5 * p.fn * m_mol * np.sum(r["vth"] ** 2, axis=0))
    e0 = 1.5 * kb * dens[:, :nr] * V * temp[:, :nr]
    de = e0 - 1.5 * kb * n1 * V * r["temp"][:, :nr]
    # the difference of two cell energies ~1e8 times larger than it: a few
    # ulps of the cell energy
    assert np.all(np.abs(de - ke.reshape(NZ, nr)) <= 16 * 2.0 ** -52 * e0)
    assert ke.max() > 1e3 * 16 * 2.0 ** -52 * e0.max()


def test_valence_photon_has_no_auger(src_tables):
    p = params(fn=10.0)
    r = oracle.photoionize(p, np.full((NZ, NR), 2.5e25), DR, DZ, 0.3, 1e7, 150e-6, src_tables, max_events=1 << 22)
    assert r["info"]["auger"] == 0 and r["ae"] is None
    assert r["info"]["electron_kev"] == pytest.approx(0.3 - 0.0175, rel=1e-12, abs=0)
    assert r["info"]["n_events"] > 1000


def test_photon_below_every_shell_is_rejected(src_tables):
    with pytest.raises(ValueError):
        oracle.photoionize(params(), np.full((NZ, NR), 2.5e25), DR, DZ, 0.01, 1e7, 150e-6, src_tables)
