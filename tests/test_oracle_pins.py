"""Pins of the FP64 oracle against known answers, closed forms and invariants.

The reference ships no tests or golden vectors, so each oracle function is
pinned here by something it does not share code with: published test
vectors (Philox), the reference's own data tables (cross sections), closed
forms (Green-Sawada CDF, binary-encounter limit, uniform-field momentum gain,
Bessel cavity frequency, trapezoid integral of a linear function) and
invariants (|v| under rotation, current sum rule, wedge-wrap invariance).
"""
import math

import numpy as np
import pytest
from scipy import integrate, special

import oracle

A0SQ = 2.8002852e-21


def params(nr=16, nz=16, **kw):
    import gasx
    return oracle.params_from(gasx.make_params(nr, nz, **kw))


# ---- Philox4x32-10: Random123 known-answer vectors ----------------------------
@pytest.mark.parametrize("ctr,key,out", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_kat(ctr, key, out):
    assert oracle.philox(ctr, key) == out


def test_u01_open_interval_and_fp32_exact():
    L = oracle.lib()
    for x in (0, 1, 255, 256, 0xffffffff, 0x80000000):
        u = L.orc_u01(x)
        assert 0.0 < u < 1.0
        assert float(np.float32(u)) == u


# ---- cross sections at table nodes (data files of the reference) ------------
def test_cross_section_known_answers(otab):
    e = np.array([0.4, 0.0156, 0.03, 0.0155, 20.1, 0.0009, 0.001, 20.0])
    se, si, sx = oracle.cross_sections(otab, e)
    assert si[0] == pytest.approx(1.632e-20, rel=1e-12, abs=0)       # TotalIonCS_Nitrogen.csv:12
    assert si[1] == pytest.approx(1e-23, rel=1e-12, abs=0)           # :1
    assert se[0] == pytest.approx(3.935 * 2 * A0SQ, rel=1e-12, abs=0)  # TotalElasticCS_Nitrogen.csv
    assert se[2] == pytest.approx(10.67e-20, rel=1e-12, abs=0)       # TotalElasticCS_smalleV_Nitrogen.csv:25
    assert np.isnan(si[3]) and not np.isnan(se[3])            # below the ionization threshold
    assert np.isnan(se[4]) and np.isnan(si[4])                # above 20 keV
    assert np.isnan(se[5]) and se[6] == pytest.approx(9.74e-20, rel=1e-6, abs=0)   # small-energy table starts at 1 eV
    assert se[7] == pytest.approx(0.125 * 2 * A0SQ, rel=1e-6, abs=0) and si[7] == pytest.approx(0.073e-20, rel=1e-6, abs=0)


def test_interp1_linear_between_nodes(otab):
    # 450 eV lies halfway between the 400 and 500 eV ionization nodes
    _, si, _ = oracle.cross_sections(otab, np.array([0.45]))
    assert si[0] == pytest.approx(0.5 * (1.632 + 1.428) * 1e-20, rel=1e-12, abs=0)


def test_table_merge(raw, otab):
    """electron_collision.m:15-33: 7 small-energy + 29 large-energy columns."""
    assert otab.n_dcs == 36
    e = np.array(otab.dcs_e[:36])
    assert np.allclose(e[:7], [0.005, 0.01, 0.015, 0.02, 0.025, 0.03, 0.04])
    assert np.allclose(e[7:], raw["t_el"][0])
    # small-energy TCS converted to a0^2 per atom; the 50 eV row is dropped
    assert otab.dcs_tcs[0] == pytest.approx(12.74e-20 / A0SQ / 2, rel=1e-6, abs=0)
    assert otab.dcs_tcs[6] == pytest.approx(9.68e-20 / A0SQ / 2, rel=1e-6, abs=0)
    assert otab.dcs_tcs[7] == pytest.approx(15.5, rel=1e-6, abs=0)
    lv = np.array(otab.dcs_logv[:36 * 181]).reshape(36, 181)
    assert np.allclose(lv[0], np.log(raw["t_sdel"][1:, 1]))
    assert np.allclose(lv[35], np.log(raw["t_del"][1:, 29]))
    assert np.allclose(lv[7], np.log(raw["t_del"][1:, 1]))


# ---- excitation --------------------------------------------------------------
def test_excitation_loss_limits(otab, raw):
    L = oracle.lib()
    W = raw["t_exc"][:, 0]
    for T in (15.6, 50.0, 400.0, 5000.0, 20000.0):
        ed = L.orc_exc_edep(otab, T, 0)
        assert W.min() <= ed <= W.max()
    # T -> infinity: only the omega = 0.75 terms survive and (1-(W/T)^g)^nu -> 1,
    # so E_dep -> sum A W^-0.25 / sum A W^-1.25 over those levels
    t = raw["t_exc"]
    m = t[:, 2] == 0.75
    lim = np.sum(t[m, 1] * t[m, 0] ** -0.25) / np.sum(t[m, 1] * t[m, 0] ** -1.25)
    assert L.orc_exc_edep(otab, 1e40, 0) == pytest.approx(lim, rel=1e-8, abs=0)


def test_excitation_cross_section_high_energy_slope(otab):
    _, _, sx = oracle.cross_sections(otab, np.array([10.0, 20.0]))
    slope = math.log(sx[1] / sx[0]) / math.log(2.0)
    assert -0.9 < slope < -0.7   # dominated by the omega = 0.75 levels


# ---- elastic angle -----------------------------------------------------------
def _mean_cos_from_table(raw, col_e_kev):
    d = raw["t_del"]
    j = int(np.nonzero(np.isclose(d[0, 1:], col_e_kev))[0][0]) + 1
    th = np.deg2rad(d[1:, 0])
    dcs = d[1:, j]
    f = lambda t: np.interp(t, th, dcs) * np.sin(t)
    num = integrate.quad(lambda t: f(t) * np.cos(t), 0, np.pi, limit=400, points=th[1:-1][::5])[0]
    den = integrate.quad(f, 0, np.pi, limit=400, points=th[1:-1][::5])[0]
    return num / den


@pytest.mark.parametrize("e_kev", [0.1, 0.4, 2.0])
def test_elastic_angle_matches_table_distribution(otab, raw, e_kev):
    L = oracle.lib()
    u = (np.arange(20000) + 0.5) / 20000
    th = np.array([L.orc_elastic_theta(otab, e_kev, x) for x in u])
    assert np.all(np.diff(th) >= 0) and th[0] >= 0 and th[-1] <= np.pi
    assert np.mean(np.cos(th)) == pytest.approx(_mean_cos_from_table(raw, e_kev), abs=2e-3)


def test_elastic_angle_between_nodes(otab, raw):
    L = oracle.lib()
    u = (np.arange(4000) + 0.5) / 4000
    m = lambda e: np.mean(np.cos([L.orc_elastic_theta(otab, e, x) for x in u]))
    lo, mid, hi = m(0.3), m(0.35), m(0.4)
    assert min(lo, hi) <= mid <= max(lo, hi)


@pytest.mark.parametrize("e_kev", [0.4, 0.35, 0.02, 3.3])
def test_elastic_rejection_sampler_matches_inverse_cdf(otab, e_kev):
    """The kernel's rejection sampler draws the reference (inverse-CDF)
    distribution: chi-square over the 1-degree intervals and uniformity
    inside them."""
    from scipy import stats
    L = oracle.lib()
    nu = 60000
    u = (np.arange(nu) + 0.5) / nu
    th_ref = np.array([L.orc_elastic_theta(otab, e_kev, x) for x in u])
    p_ref = np.bincount(np.minimum((th_ref * 180 / np.pi).astype(int), 179), minlength=180) / nu
    ns = 40000
    th = np.array([L.orc_elastic_theta_rej(otab, e_kev, s, 5, 1, 12345) for s in range(ns)])
    deg = th * 180 / np.pi
    k = np.minimum(deg.astype(int), 179)
    obs = np.bincount(k, minlength=180)
    exp_ = p_ref * ns
    # merge sparse backward bins so that every expected count is >= 20
    order = np.argsort(exp_)
    groups, cur, acc = [], [], 0.0
    for i in order:
        cur.append(i); acc += exp_[i]
        if acc >= 20:
            groups.append(cur); cur, acc = [], 0.0
    if cur:
        groups[-1].extend(cur)
    o = np.array([obs[g].sum() for g in groups], float)
    e = np.array([exp_[g].sum() for g in groups])
    e *= o.sum() / e.sum()
    assert stats.chisquare(o, e).pvalue > 1e-3
    assert stats.kstest(deg - k, "uniform").pvalue > 1e-3


# ---- ionization -----------------------------------------------------------------
@pytest.mark.parametrize("T", [20.0, 100.0, 400.0, 5000.0])
def test_green_sawada_inverse_cdf(T):
    L = oracle.lib()
    I, T0 = 15.6, 4.17 - 1000.0 / (T + 31.2)
    G = 13.8 * T / (T + 15.6)
    Tm = (T - I) / 2
    cdf = lambda E: (math.atan((E - T0) / G) + math.atan(T0 / G)) / (math.atan((Tm - T0) / G) + math.atan(T0 / G))
    import ctypes as C
    out = [C.c_double() for _ in range(5)]
    ee = 9.109e-31 * 9e16 / 1.6022e-19
    for u in (0.01, 0.3, 0.5, 0.77, 0.99):
        L.orc_ionize(T, u, ee, *[C.byref(o) for o in out])
        k0, ep, es, thp, ths = (o.value for o in out)
        assert cdf(k0) == pytest.approx(u, abs=1e-12)
        assert ep + es + I == pytest.approx(T, rel=1e-12, abs=0)      # energy balance
        assert es == pytest.approx(k0, rel=1e-6, abs=0) and ep >= es
        # binary encounter: close to 90 degrees between the outgoing electrons
        assert thp + ths == pytest.approx(math.pi / 2, abs=0.02 * math.sqrt(T / 1000) + 1e-3)


def test_rotate_preserves_speed_and_angle():
    import ctypes as C
    L = oracle.lib()
    rng = np.random.default_rng(0)
    for _ in range(200):
        v = rng.standard_normal(3) * 1e7
        th, ph = rng.random() * np.pi, rng.random() * 2 * np.pi
        a = [C.c_double(x) for x in v]
        L.orc_rotate(*[C.byref(x) for x in a], th, ph, 1e-5)
        w = np.array([x.value for x in a])
        assert np.linalg.norm(w) == pytest.approx(np.linalg.norm(v), rel=1e-12, abs=0)
        assert np.dot(v, w) / np.dot(v, v) == pytest.approx(math.cos(th), abs=1e-12)
    # degenerate branch: velocity along z
    a = [C.c_double(0.0), C.c_double(0.0), C.c_double(5.0)]
    L.orc_rotate(*[C.byref(x) for x in a], 0.3, 1.0, 1e-5)
    assert a[2].value == pytest.approx(5 * math.cos(0.3), rel=1e-6, abs=0)


# ---- FDTD ----------------------------------------------------------------------
def _cavity_frequency(ez_metric):
    nr, nz = 64, 4
    p = params(nr, nz, dr=1e-5, dz=1e-5, ddt=1e-14, n_sub=10, ez_metric=ez_metric)
    rp = nr * p.dr
    r_nodes = np.arange(nr + 1) * p.dr
    ez = np.tile(special.j0(2.404825557695773 * r_nodes / rp), (nz, 1))
    ez[:, -1] = 0.0
    er = np.zeros((nz + 1, nr))
    h = np.zeros((nz, nr))
    j0 = np.zeros(nr * nz)
    samples = []
    for _ in range(800):
        er, ez, h = oracle.fdtd_push(p, er, ez, h, j0, j0, 1)
        ez = ez.reshape(nz, nr + 1)
        er = er.reshape(nz + 1, nr)
        h = h.reshape(nz, nr)
        samples.append(ez[1, 0])
    s = np.array(samples)
    zc = np.nonzero(np.diff(np.sign(s)) != 0)[0]
    # linear interpolation of the crossings
    tz = zc + s[zc] / (s[zc] - s[zc + 1])
    period = 2 * np.mean(np.diff(tz)) * p.ddt
    return 2 * math.pi / period, 2.404825557695773 * 3e8 / rp


def test_fdtd_tm010_cavity_frequency():
    """PEC cylinder: omega = c * j01 / Rp (z-uniform TM010 mode)."""
    w, w_exact = _cavity_frequency(0)
    assert w == pytest.approx(w_exact, rel=2e-3, abs=0)


def test_fdtd_ez_metric_variant_is_off_resonance():
    """The EM3.cu:39 metric r=(i+1)dr detunes the mode (documented divergence D1)."""
    w, w_exact = _cavity_frequency(1)
    assert abs(w / w_exact - 1) > 5e-3


def test_fdtd_vacuum_energy_bounded():
    nr, nz = 24, 32
    p = params(nr, nz)
    rng = np.random.default_rng(0)
    er = rng.standard_normal((nz + 1, nr))
    ez = rng.standard_normal((nz, nr + 1))
    ez[:, -1] = 0
    h = rng.standard_normal((nz, nr)) / 377.0
    st = oracle.State(np.zeros((6, 1)))
    st.dele[:] = 1
    e0 = oracle.tallies(p, er, ez, h, st)[3]
    er2, ez2, h2 = oracle.fdtd_push(p, er, ez, h, np.zeros(nr * nz), np.zeros(nr * nz), 200)
    e1 = oracle.tallies(p, er2, ez2, h2, st)[3]
    assert np.all(np.isfinite(er2)) and e1 < 3 * e0


# ---- push ---------------------------------------------------------------------------
def _uniform_fields(nr, nz, Er=0.0, Ez=0.0, H=0.0):
    return (np.full((nz + 1) * nr, Er), np.full(nz * (nr + 1), Ez), np.full(nz * nr, H))


def test_push_uniform_ez_momentum_gain():
    """dp/dt = qE: gamma*v_z grows by qE dt/m per push, exactly in exact arithmetic."""
    nr, nz = 16, 4000
    p = params(nr, nz, dz=1e-5)
    er, ez, h = _uniform_fields(nr, nz, Ez=-2e7)
    st = oracle.State(np.array([[3e-5], [1e-6], [0.02], [1e5], [2e5], [1e6]]))
    u0 = st.v[3:6, 0] / math.sqrt(1 - np.sum(st.v[3:6, 0] ** 2) / 9e16)
    n = 50
    for _ in range(n):
        r, phi, vr = oracle.em2(p, st)
        oracle.em4(p, st, r, phi, vr, er, ez, h)
    v = st.v[3:6, 0]
    u = v / math.sqrt(1 - np.sum(v ** 2) / 9e16)
    du = -1.6e-19 * -2e7 * p.ddt / 9.1096e-31 * n
    assert u[2] - u0[2] == pytest.approx(du, rel=1e-9, abs=0)
    assert np.hypot(u[0], u[1]) == pytest.approx(np.hypot(u0[0], u0[1]), rel=1e-9, abs=0)


def test_push_free_flight_and_reflection():
    nr, nz = 16, 16
    p = params(nr, nz)
    er, ez, h = _uniform_fields(nr, nz)
    lp = nz * p.dz
    st = oracle.State(np.array([[5e-5], [1e-6], [lp - 1e-9], [0.0], [0.0], [1e7]]))
    r, phi, vr = oracle.em2(p, st)
    oracle.em4(p, st, r, phi, vr, er, ez, h)
    assert st.v[5, 0] == pytest.approx(-1e7, rel=1e-6, abs=0)                 # reflected at z = Lp
    assert st.v[2, 0] == pytest.approx(lp - 1e-9 - (1e7 * p.ddt - 1e-9) + 0, abs=1e-12) or st.v[2, 0] < lp
    assert np.linalg.norm(st.v[3:6, 0]) == pytest.approx(1e7, rel=1e-12, abs=0)


def test_wedge_wrap_is_rotation_invariant():
    """A particle and its copy rotated by PHIp end with the same r, v_r, v_phi, v_z."""
    nr, nz = 16, 16
    p = params(nr, nz)
    rot = lambda a, t: np.array([a[0] * math.cos(t) - a[1] * math.sin(t), a[0] * math.sin(t) + a[1] * math.cos(t)])
    x = np.array([6e-5, 0.2e-6])
    v = np.array([1e6, -3e7])
    z, vz = 5e-5, 2e6
    outs = []
    for t in (0.0, p.phip):
        xr, vrr = rot(x, t), rot(v, t)
        st = oracle.State(np.array([[xr[0]], [xr[1]], [z], [vrr[0]], [vrr[1]], [vz]]))
        oracle.em2(p, st)
        X, V = st.v[0:2, 0], st.v[3:5, 0]
        rr = np.linalg.norm(X)
        outs.append((rr, (X @ V) / rr, (X[0] * V[1] - X[1] * V[0]) / rr, st.v[5, 0]))
        assert 0 <= math.atan2(X[1], X[0]) <= p.phip + 1e-12
    assert np.allclose(outs[0], outs[1], rtol=1e-9)


# ---- deposit ------------------------------------------------------------------------
def test_em7_current_sum_rule():
    """sum J V = FN * sum q v (the clamped area weights sum to one)."""
    nr, nz = 20, 30
    p = params(nr, nz, fn=2.0)
    rng = np.random.default_rng(1)
    n = 5000
    r = rng.random(n) * nr * p.dr
    ph = rng.random(n) * p.phip
    st = oracle.State(np.array([r * np.cos(ph), r * np.sin(ph), rng.random(n) * nz * p.dz,
                                rng.standard_normal(n) * 1e6, rng.standard_normal(n) * 1e6,
                                rng.standard_normal(n) * 1e6]))
    rr, phi, vr = oracle.em2(p, st)
    jr, jz = oracle.em7(p, st, rr, vr)
    i = np.arange(nr)
    V = np.tile(p.phip / 2 * ((i + 1) ** 2 - i ** 2) * p.dr ** 2 * p.dz, nz)
    live = st.dele[:n] == 0
    assert np.sum(jr * V) == pytest.approx(p.fn * np.sum(st.q[:n][live] * vr[:n][live]), rel=1e-10, abs=0)
    assert np.sum(jz * V) == pytest.approx(p.fn * np.sum(st.q[:n][live] * st.v[5, :n][live]), rel=1e-10, abs=0)


def test_em7_cell_centre_hits_one_cell():
    nr, nz = 10, 10
    p = params(nr, nz)
    r0, z0 = 3.5 * p.dr, 6.5 * p.dz
    st = oracle.State(np.array([[r0], [0.0], [z0], [2e6], [0.0], [1e6]]))
    rr, phi, vr = np.array([r0]), np.array([0.0]), np.array([2e6])
    jr, jz = oracle.em7(p, st, rr, vr)
    assert np.count_nonzero(jr) == 1 and jr.reshape(nz, nr)[6, 3] != 0


def test_fixed_point_deposit_tracks_fp64():
    nr, nz = 20, 30
    p = params(nr, nz, fn=1.0)
    rng = np.random.default_rng(2)
    n = 4000
    r = rng.random(n) * nr * p.dr * 0.999
    ph = rng.random(n) * p.phip
    s32 = np.array([r * np.cos(ph), r * np.sin(ph), rng.random(n) * nz * p.dz,
                    rng.standard_normal(n) * 1e6, rng.standard_normal(n) * 1e6,
                    rng.standard_normal(n) * 1e6]).astype(np.float32)
    qr, qz = oracle.deposit_fx32(p, s32, -1.0)
    jr, jz = oracle.j_from_fixed(p, qr, qz)
    st = oracle.State(s32)
    rr = np.hypot(st.v[0], st.v[1])
    vr = (st.v[3] * st.v[0] + st.v[4] * st.v[1]) / rr
    ojr, ojz = oracle.em7(p, st, rr, vr)
    from conftest import deposit_error_bound
    bound = deposit_error_bound(p, s32, np.sqrt(np.sum(s32[3:6].astype(np.float64) ** 2, axis=0)))
    assert np.all(np.abs(jr - ojr) <= bound) and np.all(np.abs(jz - ojz) <= bound)
    # and the bound is tight enough to catch a plausible mistake: a deposit
    # that drops the z half-cell offset is far outside it
    sh = s32.copy()
    sh[2] += np.float32(0.5 * p.dz)
    qb, _ = oracle.deposit_fx32(p, sh, -1.0)
    jb, _ = oracle.j_from_fixed(p, qb, qb)
    assert not np.all(np.abs(jb - ojr) <= bound)
    # order independence of the integer sum
    perm = rng.permutation(n)
    qr2, qz2 = oracle.deposit_fx32(p, s32[:, perm], -1.0)
    assert np.array_equal(qr, qr2) and np.array_equal(qz, qz2)


# ---- tallies -----------------------------------------------------------------------
def test_tallies_trapz_closed_form():
    nr, nz = 12, 9
    p = params(nr, nz, fn=3.0)
    E0 = 5.0
    er = np.full((nz + 1) * nr, E0)
    ez = np.zeros(nz * (nr + 1))
    h = np.zeros(nz * nr)
    st = oracle.State(np.array([[1e-5], [0.0], [1e-5], [3e6], [0.0], [4e6]]))
    out = oracle.tallies(p, er, ez, h, st)
    r0, r1 = p.dr / 2, nr * p.dr - p.dr / 2
    we = p.phip * 0.5 * p.eps0 * E0 ** 2 * (r1 ** 2 - r0 ** 2) / 2 * (nz * p.dz)
    assert out[0] == pytest.approx(we, rel=1e-12, abs=0)
    g = 1 / math.sqrt(1 - 25e12 / 9e16)
    assert out[2] == pytest.approx(3.0 * (g - 1) * 9.1096e-31 * 9e16, rel=1e-9, abs=0)


# ---- collisions: decision logic against the tables -----------------------------------
def test_collision_branch_frequencies(otab):
    """Event fractions follow sigma_el : sigma_ion : sigma_exc (binomial bounds)."""
    import gasx
    p = oracle.params_from(gasx.make_params(20, 20))
    n = 40000
    from gasx.workloads import speed_from_ev
    v = speed_from_ev(np.full(n, 400.0))
    rng = np.random.default_rng(3)
    d = rng.standard_normal((3, n))
    d /= np.linalg.norm(d, axis=0)
    st = oracle.State(np.vstack([np.full(n, 5e-5), np.full(n, 1e-6), np.full(n, 1e-4), d * v]))
    g = oracle.make_gas(np.full((1, 1), 1e25), 20 * p.dr, 20 * p.dz)
    ev, sv, _ = oracle.collide(p, otab, g, st, 1e-13, seed=3, step=0)
    se, si, sx = oracle.cross_sections(otab, np.array([0.4]))
    st_ = se[0] + si[0] + sx[0]
    pt = v[0] * 1e-13 * 1e25 * st_
    ncol = np.count_nonzero(ev)
    assert abs(ncol - n * pt) < 5 * math.sqrt(n * pt * (1 - pt))
    for code, s in ((1, se[0]), (2, si[0]), (3, sx[0])):
        k = np.count_nonzero(ev == code)
        f = s / st_
        assert abs(k - ncol * f) < 5 * math.sqrt(ncol * f * (1 - f))
    # excitation leaves T - E_dep, ionization conserves T = Ep + Es + I
    from gasx.workloads import speed_from_ev as sp
    ex = np.nonzero(ev == 3)[0][:5]
    edep = oracle.lib().orc_exc_edep(otab, 400.0, 0)
    # speeds after excitation come from the collision constants (electron_transport.cu)
    g1 = (400.0 - edep) * 1.6022e-19 / (9.109e-31 * 9e16)
    vexp = 3e8 * math.sqrt(g1 * (g1 + 2)) / (1 + g1)
    assert np.allclose(np.linalg.norm(st.v[3:6, ex], axis=0), vexp, rtol=1e-12)


# ---- EM4 gathers: fields linear on each staggered grid ---------------------------
# Bilinear interpolation reproduces a field that is linear in r and z exactly at
# interior points; the EM4 edge rules (EM4.cu:48-53, 68-73, 89-116) hold the
# value of the nearest node row / column outside the node range, i.e. the
# linear field evaluated at the position clamped into the node range.  A
# half-cell offset error misses by (slope * cell / 2); a transposed index or a
# wrong edge rule misses by more.
Q_E, M_E = -1.6e-19, 9.1096e-31


def _gather_case(nr=16, nz=20, n=3000, seed=4):
    p = params(nr, nz)
    rng = np.random.default_rng(seed)
    r0 = rng.random(n) * nr * p.dr * 0.999999
    z0 = rng.random(n) * nz * p.dz
    # force a share of the points into the edge bands of every grid
    k = n // 6
    r0[:k] = rng.random(k) * 0.5 * p.dr                         # r < dr/2
    r0[k:2 * k] = nr * p.dr - rng.random(k) * 0.5 * p.dr * 0.999  # r > Rp - dr/2
    z0[2 * k:3 * k] = rng.random(k) * 0.5 * p.dz                # z < dz/2
    z0[3 * k:4 * k] = nz * p.dz - rng.random(k) * 0.5 * p.dz     # z > Lp - dz/2
    phi = p.phip * (0.25 + 0.5 * rng.random(n))
    return p, r0, z0, phi


def _kick(p, r0, z0, phi, v, er, ez, h, vr):
    """One EM4 push of particles at (r0, phi, z0) with velocity v (3 x n);
    returns the momentum change per unit mass u' - u (3 x n)."""
    n = r0.size
    st = oracle.State(np.vstack([r0 * np.cos(phi), r0 * np.sin(phi), z0, v]))
    c = 3e8
    u0 = v / np.sqrt(1 - np.sum(v ** 2, axis=0) / c ** 2)
    oracle.em4(p, st, r0.copy(), phi.copy(), vr.copy(), er, ez, h)
    v1 = st.v[3:6, :n]
    u1 = v1 / np.sqrt(1 - np.sum(v1 ** 2, axis=0) / c ** 2)
    return u1 - u0


@pytest.mark.parametrize("which", ["er", "ez"])
def test_em4_gather_reproduces_linear_electric_fields(which):
    p, r0, z0, phi = _gather_case()
    nr, nz, dr, dz = p.nr, p.nz, p.dr, p.dz
    a0, gr, gz = 700.0, 3.1e8, -1.7e8                            # V/m, V/m^2
    f = lambda r, z: a0 + gr * r + gz * z
    er = np.zeros((nz + 1, nr))
    ez = np.zeros((nz, nr + 1))
    if which == "er":        # Er nodes at r = (i + 1/2) dr, z = k dz (EM3.cu:18)
        ri, zk = (np.arange(nr) + 0.5) * dr, np.arange(nz + 1) * dz
        er[:] = f(ri[None, :], zk[:, None])
        expect = f(np.clip(r0, 0.5 * dr, (nr - 0.5) * dr), z0)
    else:                    # Ez nodes at r = i dr, z = (k + 1/2) dz (EM3.cu:19)
        ri, zk = np.arange(nr + 1) * dr, (np.arange(nz) + 0.5) * dz
        ez[:] = f(ri[None, :], zk[:, None])
        expect = f(r0, np.clip(z0, 0.5 * dz, (nz - 0.5) * dz))
    h = np.zeros((nz, nr))
    du = _kick(p, r0, z0, phi, np.zeros((3, r0.size)), er.ravel(), ez.ravel(), h.ravel(), np.zeros(r0.size))
    k = Q_E * p.ddt / M_E
    got_r = (du[0] * np.cos(phi) + du[1] * np.sin(phi)) / k
    got_z = du[2] / k
    got, other = (got_r, got_z) if which == "er" else (got_z, got_r)
    assert np.allclose(got, expect, rtol=1e-9, atol=1e-9 * a0)
    assert np.allclose(other, 0.0, atol=1e-9 * a0)
    # a half-cell offset would be visible: slope * cell / 2 >> tolerance
    assert min(abs(gr) * dr, abs(gz) * dz) / 2 > 1e3 * 1e-9 * np.abs(expect).max()


@pytest.mark.parametrize("vel", ["axial", "radial"])
def test_em4_gather_and_magnetic_force_of_linear_hphi(vel):
    """Hphi linear on its grid (r = (i+1/2) dr, z = (k+1/2) dz, EM3_1.cu:12):
    with E = 0 the kick is (q dt / m) mu0 H (vr z^ - vz r^) (EM4.cu:126-129),
    H at the clamped position."""
    p, r0, z0, phi = _gather_case(seed=5)
    nr, nz, dr, dz = p.nr, p.nz, p.dr, p.dz
    z0 = np.clip(z0, 1e-7, nz * dz - 1e-7)                        # no z reflection in the drift
    h0, hr, hz = 2.0, 4e4, 6e4                                     # A/m, A/m^2
    f = lambda r, z: h0 + hr * r + hz * z
    ri, zk = (np.arange(nr) + 0.5) * dr, (np.arange(nz) + 0.5) * dz
    h = f(ri[None, :], zk[:, None])
    expect = f(np.clip(r0, 0.5 * dr, (nr - 0.5) * dr), np.clip(z0, 0.5 * dz, (nz - 0.5) * dz))
    n, v0, mu0 = r0.size, 1e6, 4e-7 * math.pi
    if vel == "axial":        # vr = 0, vz = v0: kick along -r^
        v = np.vstack([np.zeros(n), np.zeros(n), np.full(n, v0)])
        vr = np.zeros(n)
    else:                     # vr = v0, vz = 0: kick along +z^
        v = np.vstack([v0 * np.cos(phi), v0 * np.sin(phi), np.zeros(n)])
        vr = np.full(n, v0)
    z = np.zeros((nz + 1) * nr), np.zeros(nz * (nr + 1))
    du = _kick(p, r0, z0, phi, v, z[0], z[1], h.ravel(), vr)
    k = Q_E * p.ddt / M_E * mu0 * v0
    du_r = du[0] * np.cos(phi) + du[1] * np.sin(phi)
    if vel == "axial":
        assert np.allclose(-du_r / k, expect, rtol=1e-8, atol=0)
        assert np.allclose(du[2], 0.0, atol=1e-12 * abs(k) * h0)
    else:
        assert np.allclose(du[2] / k, expect, rtol=1e-8, atol=0)
        assert np.allclose(du_r, 0.0, atol=1e-12 * abs(k) * h0)


def test_em4_uniform_hphi_rotates_momentum_in_the_rz_plane():
    """Uniform Hphi, E = 0: du = (q dt/m) mu0 H (vr z^ - vz r^), which is
    perpendicular to u, so |u'|^2 = |u|^2 + (q dt mu0 H / m)^2 (vr^2 + vz^2)."""
    p, r0, z0, phi = _gather_case(n=500, seed=6)
    # away from the wedge faces and z ends: the drift after the kick must not
    # wrap (which rotates the velocity by PHIp) or reflect
    r0 = np.maximum(r0, 2 * p.dr)
    z0 = np.clip(z0, 1e-6, p.nz * p.dz - 1e-6)
    rng = np.random.default_rng(7)
    n, H, mu0 = r0.size, 30.0, 4e-7 * math.pi
    v = rng.standard_normal((3, n)) * 2e7
    vr = v[0] * np.cos(phi) + v[1] * np.sin(phi)
    er, ez, h = np.zeros((p.nz + 1) * p.nr), np.zeros(p.nz * (p.nr + 1)), np.full(p.nz * p.nr, H)
    st = oracle.State(np.vstack([r0 * np.cos(phi), r0 * np.sin(phi), z0, v]))
    c = 3e8
    u0 = v / np.sqrt(1 - np.sum(v ** 2, axis=0) / c ** 2)
    oracle.em4(p, st, r0.copy(), phi.copy(), vr.copy(), er, ez, h)
    v1 = st.v[3:6, :n]
    u1 = v1 / np.sqrt(1 - np.sum(v1 ** 2, axis=0) / c ** 2)
    k = Q_E * p.ddt / M_E * mu0 * H
    du = u1 - u0
    rhat = np.vstack([np.cos(phi), np.sin(phi), np.zeros(n)])
    expect = k * (vr * np.array([[0], [0], [1.0]]) - v[2] * rhat)
    assert np.allclose(du, expect, rtol=1e-8, atol=1e-8 * np.abs(expect).max())
    lhs = np.sum(u1 ** 2, axis=0)
    rhs = np.sum(u0 ** 2, axis=0) + k ** 2 * (vr ** 2 + v[2] ** 2)
    assert np.allclose(lhs, rhs, rtol=1e-12, atol=0)


# ---- FDTD current coupling (EM3.cu:29, 35, 39) ----------------------------------------
def _e_update(p, er, ez, h, jr, jz, tau, axis_sign=1):
    er, ez = er.copy().ravel(), ez.copy().ravel()
    oracle.lib().orc_fdtd_e(oracle.ptr(er), oracle.ptr(ez), oracle.ptr(oracle.a(h)), oracle.ptr(oracle.a(jr)),
                            oracle.ptr(oracle.a(jz)), p.nr, p.nz, p.eps0, p.dr, p.dz, tau, p.ez_metric,
                            axis_sign)
    return er.reshape(p.nz + 1, p.nr), ez.reshape(p.nz, p.nr + 1)


@pytest.mark.parametrize("axis_sign", [1, -1])
def test_fdtd_uniform_current_follows_ampere(axis_sign):
    """With H = 0 Ampere-Maxwell reads dE/dt = -J/eps0: a uniform J lowers every
    E node the update owns by tau J / eps0 (interior Er rows, Ez columns 1..nr-1).
    On the axis the reference adds +tau Jz / eps0 (EM3.cu:35, divergence D2):
    the axis node follows Ampere only with axis_j_sign = -1."""
    nr, nz = 12, 10
    p = params(nr, nz)
    rng = np.random.default_rng(8)
    er0 = rng.standard_normal((nz + 1, nr)) * 1e3
    ez0 = rng.standard_normal((nz, nr + 1)) * 1e3
    h = np.zeros(nz * nr)
    J, tau = 3e7, 1e-15
    er, ez = _e_update(p, er0, ez0, h, np.full(nr * nz, J), np.full(nr * nz, -2 * J), tau, axis_sign)
    d = tau * J / p.eps0
    assert np.allclose(er[1:nz] - er0[1:nz], -d, rtol=1e-9, atol=0)
    assert np.array_equal(er[0], er0[1]) and np.array_equal(er[nz], er0[nz - 1])   # EM3.cu:22,26
    assert np.allclose(ez[:, 1:nr] - ez0[:, 1:nr], 2 * d, rtol=1e-9, atol=0)
    assert np.array_equal(ez[:, nr], ez0[:, nr])                                     # PEC wall
    ampere_axis = np.allclose(ez[:, 0] - ez0[:, 0], 2 * d, rtol=1e-9, atol=0)
    assert ampere_axis == (axis_sign == -1)


@pytest.mark.parametrize("comp,cell", [("r", (3, 4)), ("r", (9, 7)), ("z", (2, 5)), ("z", (10, 3))])
def test_fdtd_current_energy_theorem(comp, cell):
    """Poynting: the first-order energy change of one E update from a current
    J in one cell is -tau J . E(cell centre) V(i) (EM3.cu:29 / :39 averaging of
    J onto the E nodes).  E is uniform in r and linear in z, so the centre value
    is exact; the energy is the reference's trapezoid tally
    (electricmagneticfield5.m:167-177).  The linear part is isolated with +J and
    -J (the quadratic part cancels); a sign, averaging or index mistake in the
    current terms changes it by 2x, 0.5x or a cell's worth of E."""
    nr, nz = 12, 10
    p = params(nr, nz)
    i0, k0 = cell
    a0, g = 200.0, 4e8
    er0 = np.tile((a0 + g * np.arange(nz + 1) * p.dz)[:, None], (1, nr))
    ez0 = np.tile((a0 + g * (np.arange(nz) + 0.5) * p.dz)[:, None], (1, nr + 1))
    h = np.zeros(nz * nr)
    tau, J = 1e-15, 1e6
    st = oracle.State(np.zeros((6, 1)))
    st.dele[:] = 1

    def energy(sign):
        j = np.zeros(nr * nz)
        j[i0 + k0 * nr] = sign * J
        jr, jz = (j, np.zeros_like(j)) if comp == "r" else (np.zeros_like(j), j)
        er, ez = _e_update(p, er0, ez0, h, jr, jz, tau)
        return oracle.tallies(p, er, ez, h, st)[0]

    lin = (energy(1) - energy(-1)) / 2
    e_centre = a0 + g * (k0 + 0.5) * p.dz
    V = p.phip / 2 * ((i0 + 1) ** 2 - i0 ** 2) * p.dr ** 2 * p.dz
    assert lin == pytest.approx(-tau * J * e_centre * V, rel=1e-7, abs=0)


# ---- excitation exponents (find_CS0.m:47-55, ParameterExcCS_Nitrogen.csv columns) -----
@pytest.mark.parametrize("level", range(7))
def test_excitation_level_exponents(raw, level):
    """One Jackman-Green level, sigma ~ (W/T)^Omega (1 - (W/T)^gamma)^nu:
    near threshold (1 - (W/T)^gamma)^nu ~ (gamma (T - W)/W)^nu, so the log-slope
    in (T - W) is nu (column "v"); at high energy the log-slope in T is -Omega.
    Swapping the "v" and "miu" columns changes the threshold slope of the
    levels where they differ."""
    row = raw["t_exc"][level]
    W, om, nu = row[0], row[2], row[3]
    one = dict(raw)
    one["t_exc"] = raw["t_exc"][level:level + 1].copy()
    t = oracle.build_tables(one)
    eps = np.array([1e-7, 1e-6])
    _, _, sx = oracle.cross_sections(t, W * (1 + eps) / 1000.0)
    slope = math.log(sx[1] / sx[0]) / math.log(eps[1] / eps[0])
    assert slope == pytest.approx(nu, abs=2e-4)
    _, _, sx = oracle.cross_sections(t, np.array([1e6, 2e6]) * W / 1000.0)
    assert math.log(sx[1] / sx[0]) / math.log(2.0) == pytest.approx(-om, abs=1e-4)


def test_excitation_table_has_levels_that_tell_nu_from_gamma(raw):
    t = raw["t_exc"]
    assert np.any(t[:, 3] != t[:, 4])


# ---- ionization side products (electron_transport.cu:167-204) --------------------------
def _ionizations(otab, n=60000, seed=11):
    import gasx
    from gasx.workloads import speed_from_ev
    p = oracle.params_from(gasx.make_params(20, 20))
    rng = np.random.default_rng(seed)
    T = rng.uniform(40.0, 3000.0, n)
    d = rng.standard_normal((3, n))
    d /= np.linalg.norm(d, axis=0)
    v0 = d * speed_from_ev(T)
    st = oracle.State(np.vstack([np.full(n, 5e-5), np.full(n, 1e-6), np.full(n, 1e-4), v0]))
    g = oracle.make_gas(np.full((1, 1), 3e26), 20 * p.dr, 20 * p.dz, t_scalar=450.0)
    ev, sv, iv = oracle.collide(p, otab, g, st, 1e-13, seed=seed, step=1, ions=True)
    m = ev == 2
    return p, v0[:, m], st.v[3:6, :n][:, m], sv[m].T, iv[m].T


def test_ion_velocity_is_maxwellian(otab):
    """Each component ~ N(0, k_B T / m_ion) (electron_transport.cu:202-204)."""
    from scipy import stats
    p, _, _, _, iv = _ionizations(otab)
    sig = math.sqrt(p.k_b * 450.0 / p.m_ion)
    n = iv.shape[1]
    assert n > 5000
    for comp in iv:
        assert abs(comp.mean()) < 5 * sig / math.sqrt(n)
        chi = np.sum(comp ** 2) / sig ** 2                         # chi^2 with n degrees of freedom
        assert stats.chi2.sf(chi, n) > 1e-3 and stats.chi2.cdf(chi, n) > 1e-3
        assert stats.kstest(comp / sig, "norm").pvalue > 1e-3
    # independent components: |v|^2 / sigma^2 follows chi^2 with 3 degrees of
    # freedom (the Maxwell speed distribution), which dependent components with
    # normal marginals would not
    assert abs(np.corrcoef(iv)[0, 1]) < 5 / math.sqrt(n)
    assert stats.kstest(np.sum(iv ** 2, axis=0) / sig ** 2, stats.chi2(3).cdf).pvalue > 1e-3


def test_ionization_kinematics_are_moller_binary_encounters(otab):
    """Free-electron (Moller) kinematics with the binding energy left out of the
    angles: an electron of kinetic energy T giving W to the other leaves it at
    cos^2 = W (T + 2mc^2) / (T (W + 2mc^2)) and turns by
    cos^2 = (T - W)(T + 2mc^2) / (T (T - W + 2mc^2)).  The reference measures the
    secondary's polar angle from +z of the lab (not from the primary's
    direction) and puts it at azimuth phi_p - pi, opposite the primary's
    scattering azimuth about its old direction (electron_transport.cu:167-200)."""
    p, v0, v1, vs, _ = _ionizations(otab, n=20000, seed=12)
    c = p.c
    mc2_em = p.me_em * c * c / p.e_em
    mc2 = p.me_mc * c * c / p.e_mc

    def ev(v, m_c2):
        b2 = np.sum(v ** 2, axis=0) / c ** 2
        return m_c2 * (1 / np.sqrt(1 - b2) - 1)

    T = ev(v0, mc2_em)
    W = ev(vs, mc2)
    cs = vs[2] / np.linalg.norm(vs, axis=0)
    assert np.all(cs >= 0)                                        # lab frame, measured from +z
    assert np.allclose(cs ** 2, W * (T + 2 * mc2) / (T * (W + 2 * mc2)), rtol=1e-9)
    cp = np.sum(v0 * v1, axis=0) / (np.linalg.norm(v0, axis=0) * np.linalg.norm(v1, axis=0))
    assert np.all(cp > 0)
    assert np.allclose(cp ** 2, (T - W) * (T + 2 * mc2) / (T * (T - W + 2 * mc2)), rtol=1e-7)
    # azimuths: the primary's about its old direction (rotate.m frame), the
    # secondary's about +z
    vp2 = v0[0] ** 2 + v0[1] ** 2
    vn = np.linalg.norm(v0, axis=0)
    e1 = np.vstack([v0[0] * v0[2], v0[1] * v0[2], -vp2]) / (vn * np.sqrt(vp2))
    e2 = np.vstack([-v0[1], v0[0], np.zeros_like(vp2)]) / np.sqrt(vp2)
    phi_p = np.arctan2(np.sum(v1 * e2, axis=0), np.sum(v1 * e1, axis=0))
    phi_s = np.arctan2(vs[1], vs[0])
    dphi = np.angle(np.exp(1j * (phi_s - (phi_p - math.pi))))
    ok = vp2 > 1e-3 * vn ** 2
    assert ok.sum() > 1000 and np.allclose(dphi[ok], 0.0, atol=1e-7)


# ---- appends at a capacity crossing (electron_collision.m:68-75) ------------------------
def test_capacity_crossing_keeps_the_first_parents_in_slot_order(otab):
    """The secondaries of one collision step are appended in parent slot order
    (logical-mask order of electron_collision.m:68-75); when the capacity cuts
    the append, the kept ones are the first in that order: a run with room for
    k secondaries holds exactly the first k secondaries of a run with room for
    all, and the dropped count is the rest."""
    import gasx
    from gasx.workloads import seed_electrons
    p = oracle.params_from(gasx.make_params(16, 32))
    n = 6000
    st = np.stack(seed_electrons(n, 16 * p.dr, 32 * p.dz, p.phip, seed=21))
    g = oracle.make_gas(np.full((1, 1), 2e26), 16 * p.dr, 32 * p.dz)
    z = np.zeros
    outs = []
    for cap in (3 * n, n + 200):
        o = oracle.State(st, cap=cap)
        _, _, _, c = oracle.run(p, otab, g, o, z((32 + 1) * 16), z(32 * 17), z(32 * 16), 1,
                                collide=True, secondaries=True, seed=4)
        outs.append((o, c))
    (big, cb), (small, cs) = outs
    assert cb.n_ionize == cs.n_ionize > 400 and cb.n_dropped == 0
    assert small.n == n + 200 and cs.n_dropped == cs.n_ionize - 200
    assert np.array_equal(small.v[:, n:n + 200], big.v[:, n:n + 200])
    # the kept secondaries sit at their parents' positions, parents ascending
    pos = {tuple(v): i for i, v in enumerate(big.v[:3, :n].T.tolist())}
    parents = [pos[tuple(v)] for v in big.v[:3, n:big.n].T.tolist()]
    assert np.all(np.diff(parents) > 0) and len(parents) == cb.n_ionize
