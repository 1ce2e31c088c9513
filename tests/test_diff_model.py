"""Pins of the FP64 torch model of the differentiable run (tests/diff_model.py),
which is the reference gradient of the library's hand-written adjoint.

Its field cycle and its collision-free trajectory are checked against the C
oracle (electricmagneticfield6.m's EM2/EM7/EM3/EM4 loop, em_step), its
amplitude scaling against the linear regime's W(A) = A^2 W(1), and autograd
through it against central differences in FP64 (per node, along a uniform
profile scaling, and in the amplitude).
"""
import numpy as np
import pytest
import torch

import diff_model as dm
import oracle


def grid_and_params(nr=16, nz=24, **kw):
    import gasx
    p = gasx.make_params(nr, nz, **kw)
    return dm.Grid.from_params(p), oracle.params_from(p)


def electrons(n, g, seed=0):
    from gasx.workloads import seed_electrons
    return np.stack(seed_electrons(n, g.rp, g.lp, g.phip, sigma_r=0.2 * g.rp, seed=seed, dtype=np.float64))


def test_field_cycle_matches_oracle():
    g, op = grid_and_params(12, 20, n_sub=4)
    rng = np.random.default_rng(3)
    er = rng.standard_normal((g.nz + 1) * g.nr)
    ez = rng.standard_normal(g.nz * (g.nr + 1))
    h = rng.standard_normal(g.nz * g.nr) / 377.0
    jr = rng.standard_normal(g.nz * g.nr) * 1e3
    jz = rng.standard_normal(g.nz * g.nr) * 1e3
    oer, oez, oh = oracle.fdtd_push(op, er, ez, h, jr, jz, 1)
    te, tz, th = dm.field_cycle(g, *dm.layout_fields(g, er, ez, h),
                                torch.as_tensor(jr).view(g.nz, g.nr), torch.as_tensor(jz).view(g.nz, g.nr))
    for t, o in ((te, oer), (tz, oez), (th, oh)):
        np.testing.assert_allclose(t.numpy().ravel(), o, rtol=1e-12, atol=1e-12 * np.abs(o).max())


@pytest.mark.parametrize("ez_metric,axis", [(0, 1), (1, -1)])
def test_field_cycle_variants_match_oracle(ez_metric, axis):
    g, op = grid_and_params(8, 10, n_sub=3, ez_metric=ez_metric, axis_j_sign=axis)
    rng = np.random.default_rng(4)
    er = rng.standard_normal((g.nz + 1) * g.nr)
    ez = rng.standard_normal(g.nz * (g.nr + 1))
    h = rng.standard_normal(g.nz * g.nr) / 377.0
    jr = rng.standard_normal(g.nz * g.nr)
    jz = rng.standard_normal(g.nz * g.nr)
    oer, oez, oh = oracle.fdtd_push(op, er, ez, h, jr, jz, 1)
    te, tz, th = dm.field_cycle(g, *dm.layout_fields(g, er, ez, h),
                                torch.as_tensor(jr).view(g.nz, g.nr), torch.as_tensor(jz).view(g.nz, g.nr))
    for t, o in ((te, oer), (tz, oez), (th, oh)):
        np.testing.assert_allclose(t.numpy().ravel(), o, rtol=1e-12, atol=1e-12 * np.abs(o).max())


def no_events(n, steps):
    return [{"dv": np.zeros((3, n)), "p0": np.zeros(n), "col": np.zeros(n, np.uint8)} for _ in range(steps)]


def test_collision_free_run_matches_oracle_em_loop():
    """Without collisions the model is the reference's EM2 -> EM7 -> field
    cycle -> EM4 loop: the same final Er, hence the same objective."""
    g, op = grid_and_params(16, 24, fn=2e5)
    st = electrons(300, g, seed=5)
    steps = 6
    nodes0 = np.full(8, 1e25)
    w = dm.run(g, st, no_events(st.shape[1], steps), torch.as_tensor(nodes0), nodes0,
               torch.tensor(1.0, dtype=torch.float64), steps, r_frac=0.0)
    ost = oracle.State(st)
    z = np.zeros
    er, ez, h, _ = oracle.em_step(op, ost, z((g.nz + 1) * g.nr), z(g.nz * (g.nr + 1)), z(g.nz * g.nr), steps)
    w_or = dm.objective(g, torch.as_tensor(er).view(g.nz + 1, g.nr), 0.0).item()
    assert w_or > 0
    assert w.item() == pytest.approx(w_or, rel=1e-9, abs=0)


def test_amplitude_scales_linear_regime():
    """Weak source: at FN = 1 the self-field moves the particles by ~1e-8 of
    their paths, so Er is linear in A to that level and W(A) = A^2 W(1)."""
    g, _ = grid_and_params(16, 24, fn=1.0)
    st = electrons(200, g, seed=6)
    nodes0 = np.full(8, 1e25)
    ev = no_events(st.shape[1], 4)
    w1 = dm.run(g, st, ev, torch.as_tensor(nodes0), nodes0, torch.tensor(1.0, dtype=torch.float64), 4, 0.0)
    w3 = dm.run(g, st, ev, torch.as_tensor(nodes0), nodes0, torch.tensor(3.0, dtype=torch.float64), 4, 0.0)
    assert w3.item() == pytest.approx(9 * w1.item(), rel=1e-6, abs=0)


def synthetic_events(n, steps, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(steps):
        p0 = rng.uniform(0.0, 0.3, n)
        p0[rng.random(n) < 0.1] = 0.0
        col = (rng.random(n) < p0).astype(np.uint8)
        dv = rng.standard_normal((3, n)) * 2e6 * col
        out.append({"dv": dv, "p0": p0, "col": col})
    return out


def test_uniform_scaling_derivative_matches_richardson_differences():
    """Derivative along n = lam n0 (every ratio n/n0 equal to lam, so both
    likelihood-ratio branches move together) against a fourth-order central
    difference in lam."""
    g, _ = grid_and_params(16, 24, fn=1.0)
    st = electrons(150, g, seed=8)
    steps = 3
    ev = synthetic_events(st.shape[1], steps, 9)
    nodes0 = np.linspace(3e25, 1e25, 6)
    amp = torch.tensor(1.0, dtype=torch.float64)
    lam = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    w = dm.run(g, st, ev, lam * torch.as_tensor(nodes0), nodes0, amp, steps, 0.0)
    (dw,) = torch.autograd.grad(w, lam)
    lams = np.array([0.98, 0.99, 1.01, 1.02])
    ws = [dm.run(g, st, ev, float(l) * torch.as_tensor(nodes0), nodes0, amp, steps, 0.0).item() for l in lams]
    fd = (8 * (ws[2] - ws[1]) - (ws[3] - ws[0])) / (12 * 0.01)
    assert dw.item() == pytest.approx(fd, rel=1e-6, abs=0)


def test_autograd_matches_central_differences():
    """Self-consistent field (large FN): autograd through the model against
    central differences in FP64 for two nodes and the amplitude."""
    g, _ = grid_and_params(16, 24, fn=3e5, n_sub=4)
    st = electrons(200, g, seed=10)
    steps = 5
    ev = synthetic_events(st.shape[1], steps, 11)
    nodes0 = np.linspace(2e25, 0.5e25, 6)
    nodes = torch.tensor(nodes0, requires_grad=True)
    amp = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    w = dm.run(g, st, ev, nodes, nodes0, amp, steps, 0.25)
    gn, ga = torch.autograd.grad(w, (nodes, amp))

    def W(nv, a):
        return dm.run(g, st, ev, torch.as_tensor(nv), nodes0, torch.tensor(a, dtype=torch.float64), steps,
                      0.25).item()

    h = 1e-4
    for j in (0, 2):
        e = np.zeros_like(nodes0)
        e[j] = h * nodes0[j]
        fd = (W(nodes0 + e, 1.0) - W(nodes0 - e, 1.0)) / (2 * e[j])
        assert gn[j].item() == pytest.approx(fd, rel=1e-5, abs=1e-6 * np.abs(gn.numpy()).max())
    fda = (W(nodes0, 1 + h) - W(nodes0, 1 - h)) / (2 * h)
    assert ga.item() == pytest.approx(fda, rel=1e-6, abs=0)
