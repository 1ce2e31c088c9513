"""Pins of the oracle's electrostatic ion step (orc_es_poisson, orc_es_field,
orc_es_push; ion_staticelectrofield_push.m:21-156 with Poisson.m:25-68)
against what the mathematics fixes, on the CPU.

* the field is numpy.gradient of PHI (MATLAB's gradient rule: central
  differences inside, one-sided at the edges);
* without electrons the iteration solves the discrete Poisson problem: its
  fixed point is the direct sparse solve of the same system, and as the cells
  shrink it converges to the analytic potential of a chosen charge density
  (a wrong sign or offset of the stencil's r terms does not);
* with Boltzmann electrons a large uniform plasma is quasi-neutral away from
  the walls (rho_e = rho_i) and the electrons carry the ions' charge;
* the deposit conserves charge, and the kick along r-hat of the half-drifted
  position conserves the ion's angular momentum about the axis exactly.
"""
import math

import numpy as np
import pytest

import oracle

EPS0 = 8.854187817e-12
E = 1.6e-19


def analytic_case(nr, nz, delta):
    """phi = (Rp^2 - r^2) sin(pi z / Lp) >= 0 and rho = -eps0 lap(phi)."""
    rp, lp = nr * delta, nz * delta
    r = (np.arange(nr) + 0.5) * delta
    z = (np.arange(nz) + 0.5) * delta
    R, Z = np.meshgrid(r, z)
    phi = (rp ** 2 - R ** 2) * np.sin(np.pi * Z / lp)
    lap = -(4 + (rp ** 2 - R ** 2) * (np.pi / lp) ** 2) * np.sin(np.pi * Z / lp)
    return phi, -EPS0 * lap


def test_field_is_matlab_gradient():
    rng = np.random.default_rng(1)
    p = oracle.es_params(17, 23, 2e-4, 0.5, 1.0, 1e5, 1e-9)
    phi = rng.standard_normal((23, 17))
    er, ez = oracle.es_field(p, phi)
    assert np.allclose(er, -np.gradient(phi, 2e-4, axis=1), rtol=1e-13, atol=1e-9)
    assert np.allclose(ez, -np.gradient(phi, 2e-4, axis=0), rtol=1e-13, atol=1e-9)


def solve_direct(nr, nz, delta, rho):
    """The Jacobi fixed point: the same discrete system solved directly
    (interior stencil of Poisson.m:46, PHI(0,:) = PHI(1,:), zero edges)."""
    from scipy.sparse import lil_matrix
    from scipy.sparse.linalg import spsolve
    d2 = delta ** 2
    n = nr * nz
    A = lil_matrix((n, n))
    b = np.zeros(n)
    idx = lambda i, j: j * nr + i
    for j in range(nz):
        for i in range(nr):
            c = idx(i, j)
            if j == 0 or j == nz - 1 or i == nr - 1:
                A[c, c] = 1.0
            elif i == 0:
                A[c, c] = 1.0
                A[c, idx(1, j)] = -1.0
            else:
                R = (i + 0.5) * delta
                A[c, c] = 4 / d2
                m = 1 / (R + delta) / (2 * delta)
                A[c, idx(i + 1, j)] = -(1 / d2 + m)
                A[c, idx(i - 1, j)] = -(1 / d2 - m)
                A[c, idx(i, j - 1)] = -1 / d2
                A[c, idx(i, j + 1)] = -1 / d2
                b[c] = rho[j, i] / EPS0
    return spsolve(A.tocsr(), b).reshape(nz, nr)


def test_jacobi_fixed_point_is_the_direct_solve():
    nr, nz, d = 20, 30, 1e-4
    _, rho = analytic_case(nr, nz, d)
    p = oracle.es_params(nr, nz, d, 0.5, 1.0, 1e5, 1e-9, tol=1e-13, max_iter=400_000, boltzmann=False)
    phi, rho_e, info = oracle.es_poisson(p, rho)
    assert info["converged"] == 1 and not rho_e.any()
    ref = solve_direct(nr, nz, d, rho)
    assert np.max(np.abs(phi - ref)) < 1e-9 * np.max(np.abs(ref))


def test_poisson_converges_to_the_analytic_potential():
    errs = []
    for nr, nz, d in ((10, 15, 2e-4), (20, 30, 1e-4), (40, 60, 5e-5)):
        exact, rho = analytic_case(nr, nz, d)
        p = oracle.es_params(nr, nz, d, 0.5, 1.0, 1e5, 1e-9, tol=1e-12, max_iter=2_000_000, boltzmann=False)
        phi, _, info = oracle.es_poisson(p, rho)
        assert info["converged"] == 1
        errs.append(np.max(np.abs(phi - exact)) / np.max(exact))
    print("max error / max phi:", errs)
    # first order (the stencil's 1/(R + dr) metric and the edge rows are O(dr))
    assert errs[1] < 0.6 * errs[0] and errs[2] < 0.6 * errs[1]
    assert errs[2] < 0.05


def test_boltzmann_plasma_is_quasi_neutral_inside():
    nr, nz, d = 60, 90, 1e-4
    te = 1e5
    lam = 5 * d                                          # Debye length of five cells
    n_i = EPS0 * 1.38064853e-23 * te / (lam ** 2 * E ** 2)
    rho = np.full((nz, nr), E * n_i)
    p = oracle.es_params(nr, nz, d, 0.5, 1.0, te, 1e-9, tol=1e-7, max_iter=2_000_000)
    phi, rho_e, info = oracle.es_poisson(p, rho)
    assert info["converged"] == 1 and phi.min() >= 0
    # more than six Debye lengths from the wall at r = Rp and the z ends
    r = (np.arange(nr) + 0.5) * d
    z = (np.arange(nz) + 0.5) * d
    # (row 0, the axis row, takes PHI from row 1 and has no electrons: Poisson.m:41,59)
    inner = ((r[None, :] > d) & (r[None, :] < nr * d - 6 * lam) & (z[:, None] > 6 * lam) &
             (z[:, None] < nz * d - 6 * lam))
    ratio = rho_e[inner] / rho[inner]
    print("rho_e / rho_i inside:", ratio.min(), ratio.max(), info)
    # flat inside; the electrons missing from the sheaths sit a few percent
    # above the ions there (their totals are equal)
    assert np.ptp(ratio) < 5e-3 and abs(ratio.mean() - 1) < 0.05
    i = np.arange(nr)
    R = (i + 0.5) * d
    w = np.ones(nr); w[[0, -1]] = 0.5
    wz = np.ones(nz); wz[[0, -1]] = 0.5
    tot_e = np.sum(wz[:, None] * w[None, :] * rho_e * R[None, :])
    tot_i = np.sum(wz[:, None] * w[None, :] * rho * R[None, :])
    assert tot_e * info["n_e0"] / ne0_of(rho_e, phi, te) == pytest.approx(tot_i, rel=1e-9)
    # the potential falls towards the walls
    assert phi[nz // 2, nr // 2] > phi[nz // 2, nr - 2] > phi[nz // 2, nr - 1] == 0
    # Boltzmann relation wherever there are electrons (the sheaths too, where
    # PHI varies): rho_e = e n_e0 exp(e PHI / kT) with one n_e0 (converged:
    # the last sweep's PHI is the final one)
    assert info["max_dphi"] <= 1e-7 * info["mean_phi"]
    beta = E / (1.38064853e-23 * te)
    has = rho_e > 0
    assert np.ptp(phi[has]) > 0.2 * phi.max()
    ne0 = rho_e[has] / (E * np.exp(beta * phi[has]))
    assert np.ptp(ne0) < 1e-9 * ne0.mean()


def ne0_of(rho_e, phi, te):
    """The n_e0 of the last sweep's electrons (they saw the PHI before it;
    converged, that PHI equals the last one to the stopping tolerance)."""
    beta = E / (1.38064853e-23 * te)
    i = np.nonzero(rho_e)
    return float(np.median(rho_e[i] / (E * np.exp(beta * phi[i]))))


def ions(n, nr, nz, d, seed, vmax=3e3):
    rng = np.random.default_rng(seed)
    r = nr * d * np.sqrt(rng.random(n)) * 0.98
    ph = rng.random(n) * 0.5
    z = rng.random(n) * nz * d
    v = rng.standard_normal((3, n)) * vmax
    return np.ascontiguousarray(np.stack([r * np.cos(ph), r * np.sin(ph), z, v[0], v[1], v[2]]))


def test_push_conserves_charge_and_angular_momentum():
    nr, nz, d = 30, 45, 1e-4
    w = 1e4
    a = ions(4000, nr, nz, d, 3)
    b = ions(1500, nr, nz, d, 4)
    a0, b0 = a.copy(), b.copy()
    dt = 2e-9
    p = oracle.es_params(nr, nz, d, 0.5, w, 1e6, dt)
    out = oracle.es_push(p, [(a, 1.0, 4.65e-26), (b, 2.0, 4.65e-26)])
    assert out["info"]["converged"] == 1 and out["phi"].max() > 0
    # deposit: the density carries every ion's charge
    i = np.arange(nr)
    V = 0.5 * 0.5 * ((i + 1) ** 2 - i ** 2) * d ** 3
    assert np.sum(out["rho_i"] * V[None, :]) == pytest.approx(w * E * (4000 + 2 * 1500), rel=1e-12)
    for st, st0 in ((a, a0), (b, b0)):
        # leapfrog: x1 - x0 = dt/2 (v0 + v1)
        assert np.allclose(st[:3] - st0[:3], 0.5 * dt * (st0[3:] + st[3:]), rtol=1e-9, atol=1e-18)
        # the kick is along r-hat: L_z = x vy - y vx is unchanged
        lz0 = st0[0] * st0[4] - st0[1] * st0[3]
        lz1 = st[0] * st[4] - st[1] * st[3]
        assert np.allclose(lz1, lz0, rtol=1e-9, atol=1e-12 * np.abs(lz0).max())
        assert not np.allclose(st[3:], st0[3:])          # the field did push them


def test_ion_in_a_cell_centre_lands_in_that_cell():
    nr, nz, d = 8, 8, 1e-4
    st = np.ascontiguousarray(np.array([[3.5 * d], [0.0], [5.5 * d], [0.0], [0.0], [0.0]]))
    p = oracle.es_params(nr, nz, d, 0.5, 7.0, 1e6, 1e-9)
    out = oracle.es_push(p, [(st, 2.0, 4.65e-26)])
    V = 0.5 * 0.5 * (16 - 9) * d ** 3
    expect = np.zeros((nz, nr))
    expect[5, 3] = 7.0 * E * 2 / V
    assert np.array_equal(out["rho_i"] == 0, expect == 0)
    assert out["rho_i"][5, 3] == pytest.approx(expect[5, 3], rel=1e-14)
