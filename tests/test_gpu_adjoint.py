"""Differentiable run on the GPU (gasx_diff_*): the hand-written adjoint
against autograd through the FP64 torch model of the same frozen-event run
(tests/diff_model.py, pinned to the C oracle in test_diff_model.py), against
central differences of the GPU forward (h = 1e-3), across checkpoint
intervals, and through torch.autograd.Function.

FP32 particles against the FP64 model: the objective agrees to 1e-5
relative and the gradient to 2e-3 of its norm (the adjoint runs its particle
sensitivities in FP32)."""
import numpy as np
import pytest
import torch

import diff_model as dm

pytestmark = pytest.mark.gpu

NR, NZ, N, STEPS = 32, 64, 3000, 24


def setup(fn=3e4, n=N, seed=2):
    import gasx
    from gasx.workloads import seed_electrons
    p = gasx.make_params(NR, NZ, fn=fn)
    ctx = gasx.Context(p, device=0)
    st = np.stack(seed_electrons(n, NR * p.dr, NZ * p.dz, p.phip, sigma_r=60e-6, seed=seed))
    e = gasx.ParticleSet(n).load(*st)
    nodes0 = np.linspace(1.5e26, 3e25, 12)
    return gasx, p, ctx, st, e, nodes0


def record(gasx, ctx, e, nodes0, every=5, steps=STEPS, r_frac=0.25):
    return gasx.DiffRun(ctx, e, nodes0, n_steps=steps, checkpoint_every=every, seed=11, r_frac=r_frac)


@pytest.fixture(scope="module")
def case(gpu):
    gasx, p, ctx, st, e, nodes0 = setup()
    run = record(gasx, ctx, e, nodes0)
    events = [run.events(s) for s in range(STEPS)]
    w = run.forward(nodes0, 1.0)
    g, ga = run.backward()
    return dict(gasx=gasx, p=p, ctx=ctx, st=st, e=e, nodes0=nodes0, run=run, events=events, w=w, g=g, ga=ga)


def reference(case, nodes=None, amp=1.0, steps=STEPS, r_frac=0.25):
    g = dm.Grid.from_params(case["p"])
    nodes0 = case["nodes0"]
    nv = torch.tensor(nodes0 if nodes is None else nodes, requires_grad=True)
    a = torch.tensor(amp, dtype=torch.float64, requires_grad=True)
    w = dm.run(g, case["st"].astype(np.float64), case["events"], nv, nodes0, a, steps, r_frac)
    gn, ga = torch.autograd.grad(w, (nv, a))
    return w.item(), gn.numpy(), ga.item()


def test_recorded_events_are_collisions(case):
    cols = np.array([ev["col"].sum() for ev in case["events"]])
    p0 = np.concatenate([ev["p0"] for ev in case["events"]])
    assert cols.sum() > 0.05 * N                  # both likelihood-ratio branches are exercised
    assert (p0 >= 0).all() and (p0 < 1).all() and p0.max() > 0.01
    kicked = np.concatenate([np.abs(ev["dv"]).sum(0) > 0 for ev in case["events"]])
    collided = np.concatenate([ev["col"] != 0 for ev in case["events"]])
    assert np.array_equal(kicked, collided)


def test_objective_matches_fp64_model(case):
    w_ref, _, _ = reference(case)
    assert case["w"] > 0
    assert case["w"] == pytest.approx(w_ref, rel=1e-5, abs=0)


def test_gradient_matches_fp64_autograd(case):
    _, g_ref, ga_ref = reference(case)
    g = case["g"]
    assert np.linalg.norm(g_ref) > 0
    err = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    assert err < 2e-3, (g, g_ref)
    assert case["ga"] == pytest.approx(ga_ref, rel=2e-3, abs=0)


def test_gradient_matches_central_differences(case):
    """h = 1e-3 relative: along the gradient's own direction in the profile,
    per large node, and in the amplitude."""
    run, nodes0, g = case["run"], case["nodes0"], case["g"]
    h = 1e-3
    d = g * nodes0 / np.linalg.norm(g * nodes0) * nodes0        # relative direction, nodes-weighted
    fd = (run.forward(nodes0 + h * d, 1.0) - run.forward(nodes0 - h * d, 1.0)) / (2 * h)
    assert fd == pytest.approx(float(g @ d), rel=1e-2, abs=0)
    j = int(np.argmax(np.abs(g * nodes0)))
    e = np.zeros_like(nodes0)
    e[j] = h * nodes0[j]
    fdj = (run.forward(nodes0 + e, 1.0) - run.forward(nodes0 - e, 1.0)) / (2 * e[j])
    assert fdj == pytest.approx(g[j], rel=2e-2, abs=0)
    fda = (run.forward(nodes0, 1 + h) - run.forward(nodes0, 1 - h)) / (2 * h)
    assert fda == pytest.approx(case["ga"], rel=1e-3, abs=0)
    run.forward(nodes0, 1.0)


def test_checkpoint_interval_does_not_change_gradient(case):
    gasx, ctx, e, nodes0 = case["gasx"], case["ctx"], case["e"], case["nodes0"]
    outs = []
    for every in (STEPS, 7):
        r = record(gasx, ctx, e, nodes0, every=every)
        outs.append((r.forward(nodes0, 1.0), *r.backward()))
    base = case["run"] = record(gasx, ctx, e, nodes0)   # restore the module's recorded run
    base.forward(nodes0, 1.0)
    for w, g, ga in outs:
        assert w == pytest.approx(case["w"], rel=1e-6, abs=0)
        assert np.linalg.norm(g - case["g"]) <= 1e-4 * np.linalg.norm(case["g"])
        assert ga == pytest.approx(case["ga"], rel=1e-5, abs=0)


def test_gradient_off_reference_profile(case):
    """Away from n0 the weights differ from 1 and the ratio's position
    dependence enters; the adjoint still matches the FP64 model."""
    run, nodes0 = case["run"], case["nodes0"]
    nodes = nodes0 * np.linspace(1.3, 0.8, nodes0.size)
    w = run.forward(nodes, 1.7)
    g, ga = run.backward()
    w_ref, g_ref, ga_ref = reference(case, nodes=nodes, amp=1.7)
    assert w == pytest.approx(w_ref, rel=1e-5, abs=0)
    assert np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref) < 2e-3
    assert ga == pytest.approx(ga_ref, rel=2e-3, abs=0)
    run.forward(nodes0, 1.0)


def test_tiny_objective_gradient(gpu):
    """FN = 1e-16 scales the current (and every field) by ~3e-21 against the
    other cases, so W ~ 1e-41 of theirs: the adjoint's power-of-two seed
    scaling keeps the FP32 sensitivities representable, and objective and
    gradient still match the FP64 model."""
    gasx, p, ctx, st, e, nodes0 = setup(fn=1e-16)
    run = record(gasx, ctx, e, nodes0)
    c = {"p": p, "st": st, "nodes0": nodes0, "events": [run.events(s) for s in range(STEPS)]}
    w = run.forward(nodes0, 1.0)
    g, ga = run.backward()
    w_ref, g_ref, ga_ref = reference(c)
    assert 0 < w_ref < 1e-50
    assert w == pytest.approx(w_ref, rel=1e-5, abs=0)
    assert np.linalg.norm(g_ref) > 0
    assert np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref) < 2e-3
    assert ga == pytest.approx(ga_ref, rel=2e-3, abs=0)


def test_autograd_function(case):
    gasx, run, nodes0 = case["gasx"], case["run"], case["nodes0"]
    nodes = torch.tensor(nodes0, requires_grad=True)
    amp = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    w = gasx.shell_energy(run, nodes, amp)
    # a second forward in between: backward restores the first one's point
    run.forward(nodes0 * 1.1, 1.0)
    (2.0 * w).backward()
    assert w.item() == pytest.approx(case["w"], rel=1e-6, abs=0)
    np.testing.assert_allclose(nodes.grad.numpy(), 2.0 * case["g"], rtol=1e-6,
                               atol=1e-6 * np.abs(case["g"]).max())
    assert amp.grad.item() == pytest.approx(2.0 * case["ga"], rel=1e-6, abs=0)


def test_adjoint_errors(case):
    gasx, ctx = case["gasx"], case["ctx"]
    from gasx._abi import GasxError
    r = case["run"]
    with pytest.raises(ValueError):
        r.forward(np.ones(3), 1.0)
    with pytest.raises(GasxError):
        r.forward(case["nodes0"], -1.0)
    with pytest.raises(GasxError):
        gasx.DiffRun(ctx, case["e"], np.array([1e25]), n_steps=4)          # one node
    with pytest.raises(GasxError):
        gasx.DiffRun(ctx, case["e"], -case["nodes0"], n_steps=4)           # negative profile
    # argument errors leave the recorded run and its last forward untouched
    g, ga = r.backward()
    np.testing.assert_allclose(g, case["g"], rtol=1e-9, atol=1e-9 * np.abs(case["g"]).max())
    # a new recording has no forward yet: backward is a sequence error
    case["run"] = record(gasx, ctx, case["e"], case["nodes0"])
    with pytest.raises(GasxError):
        case["run"].backward()
    case["run"].forward(case["nodes0"], 1.0)


@pytest.mark.parametrize("nr,nz,kw", [
    (40, 70, {}),
    (37, 101, {"ez_metric": 1, "axis_j_sign": -1}),
    (16, 32, {"n_sub": 4}),
    (90, 33, {"n_sub": 1}),
    (24, 2, {"n_sub": 3}),
])
def test_transposed_field_cycle_dot_product(gpu, nr, nz, kw):
    """<lam, M F + B J> = <M^T lam, F> + <B^T lam, J> for random F, J, lam,
    with the forward cycle of gasx_fdtd (pinned to the oracle in
    test_gpu_parity)."""
    import gasx
    p = gasx.make_params(nr, nz, **kw)
    ctx = gasx.Context(p)
    rng = np.random.default_rng(nr * 1000 + nz)
    F = [rng.standard_normal((nz + 1, nr)), rng.standard_normal((nz, nr + 1)), rng.standard_normal((nz, nr))]
    L = [rng.standard_normal((nz + 1, nr)), rng.standard_normal((nz, nr + 1)), rng.standard_normal((nz, nr))]
    jr, jz = rng.standard_normal((nz, nr)) * 1e3, rng.standard_normal((nz, nr)) * 1e3
    fs = gasx.FieldState(nr, nz)
    for t, a in zip((fs.er, fs.ez, fs.hphi), F):
        t.copy_(torch.from_numpy(a))
    ctx.fdtd(fs, torch.from_numpy(jr).cuda(), torch.from_numpy(jz).cuda(), 1)
    MF = fs.numpy()
    lam = gasx.FieldState(nr, nz)
    for t, a in zip((lam.er, lam.ez, lam.hphi), L):
        t.copy_(torch.from_numpy(a))
    ljr = torch.zeros((nz, nr), dtype=torch.float64, device="cuda")
    ljz = torch.zeros_like(ljr)
    ctx.fdtd_transpose(lam, ljr, ljz)
    MtL = lam.numpy()
    lhs = sum(float((l * m).sum()) for l, m in zip(L, MF))
    rhs = sum(float((a * b).sum()) for a, b in zip(MtL, F)) + float((ljr.cpu().numpy() * jr).sum()) + \
        float((ljz.cpu().numpy() * jz).sum())
    scale = sum(float(np.abs(l * m).sum()) for l, m in zip(L, MF))
    assert abs(lhs - rhs) <= 1e-11 * scale
    # and the current part alone (F = 0): the transposed map's B^T is not zero
    assert np.abs(ljr.cpu().numpy()).max() > 0 and np.abs(ljz.cpu().numpy()).max() > 0
