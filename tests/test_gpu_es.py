"""Electrostatic ion step on the GPU (gasx_ion_es_push) against the FP64
oracle (orc_es_push) from the same FP32 ion states.

The charge counts per cell are integers, so rho_i agrees exactly unless an
ion's FP32 half-drifted position falls on the other side of a cell edge than
the oracle's FP64 one; the Poisson iteration runs the oracle's operations in
FP64 (sums in another order, the device's exp), so PHI agrees to rounding
while both stop after the same sweep; the ions agree to FP32 rounding.
"""
import numpy as np
import pytest

import oracle
from conftest import log

pytestmark = pytest.mark.gpu

M_ION = 4.65e-26


def ions(n, nr, nz, d, seed, vmax=3e3, outside=0):
    rng = np.random.default_rng(seed)
    r = nr * d * np.sqrt(rng.random(n)) * 0.98
    ph = rng.random(n) * 0.5
    z = rng.random(n) * nz * d
    if outside:
        r[:outside] = nr * d * (1.0 + rng.random(outside))   # beyond Rp: the last ring's cells
        z[outside:2 * outside] = -d * rng.random(outside)    # below z = 0: the first slab's
        r[2 * outside] = 0.0                                  # on the axis
    v = rng.standard_normal((3, n)) * vmax
    return np.stack([r * np.cos(ph), r * np.sin(ph), z, v[0], v[1], v[2]]).astype(np.float32)


@pytest.mark.parametrize("case", ["two_species", "edges"])
def test_es_push_matches_oracle(gpu, case):
    import gasx
    nr, nz, d, w, te, dt = 30, 45, 1e-4, 1e4, 1e6, 2e-9
    log(f"ion es push {case}")
    p = gasx.make_params(nr, nz, phip=0.5)
    ctx = gasx.Context(p)
    a = ions(4000, nr, nz, d, 3, outside=20 if case == "edges" else 0)
    b = ions(1500, nr, nz, d, 4)
    sa = gasx.ParticleSet(a.shape[1], qnum=1.0, mass=M_ION).load(*a)
    sb = gasx.ParticleSet(b.shape[1], qnum=2.0, mass=M_ION).load(*b)
    res = ctx.ion_es_push([sa, sb], nr, nz, d, w, te, dt)
    oa, ob = a.astype(np.float64), b.astype(np.float64)
    op = oracle.es_params(nr, nz, d, 0.5, w, te, dt)
    ref = oracle.es_push(op, [(oa, 1.0, M_ION), (ob, 2.0, M_ION)])
    info = ref["info"]
    print({k: res[k] for k in ("iterations", "converged", "n_e0", "max_dphi", "mean_phi")}, info)
    assert res["converged"] == info["converged"] == 1
    g = {k: res[k].cpu().numpy() for k in ("phi", "rho_i", "rho_e", "e_r", "e_z")}
    diff_cells = int(np.sum(g["rho_i"] != ref["rho_i"]))
    assert diff_cells <= 4
    scale = np.abs(ref["phi"]).max()
    assert scale > 0
    if diff_cells == 0 and res["iterations"] == info["iterations"]:
        assert res["int_rho_i"] == pytest.approx(info["int_rho_i"], rel=1e-13)
        assert res["n_e0"] == pytest.approx(info["n_e0"], rel=1e-10)
        assert np.max(np.abs(g["phi"] - ref["phi"])) < 1e-10 * scale
        assert np.max(np.abs(g["rho_e"] - ref["rho_e"])) < 1e-10 * np.abs(ref["rho_e"]).max()
        for k in ("e_r", "e_z"):
            assert np.max(np.abs(g[k] - ref[k])) < 1e-8 * np.abs(ref[k]).max()
    else:
        assert abs(res["iterations"] - info["iterations"]) <= 1
        assert np.max(np.abs(g["phi"] - ref["phi"])) < 0.02 * scale
    for st, o, start in ((sa.state(), oa, a), (sb.state(), ob, b)):
        assert np.max(np.abs(st[:3] - o[:3])) < 2e-7 * nr * d
        kick = np.abs(o[3:] - start[3:].astype(np.float64))
        err = np.abs(st[3:] - o[3:])
        assert np.quantile(err.max(axis=0), 0.999) < 1e-6 * np.abs(o[3:]).max() + 1e-3 * kick.max()
    # deterministic: the same call on the same start gives the same bits
    sa2 = gasx.ParticleSet(a.shape[1], qnum=1.0, mass=M_ION).load(*a)
    sb2 = gasx.ParticleSet(b.shape[1], qnum=2.0, mass=M_ION).load(*b)
    res2 = ctx.ion_es_push([sa2, sb2], nr, nz, d, w, te, dt)
    assert res2["iterations"] == res["iterations"]
    assert np.array_equal(res2["phi"].cpu().numpy(), g["phi"])
    assert np.array_equal(sa2.state(), sa.state()) and np.array_equal(sb2.state(), sb.state())


def test_es_push_rejects_bad_arguments(gpu):
    import gasx
    from gasx._abi import GasxError
    ctx = gasx.Context(gasx.make_params(8, 8))
    s = gasx.ParticleSet(10, qnum=1.0, mass=M_ION)
    with pytest.raises(GasxError):
        ctx.ion_es_push([s], 2, 8, 1e-4, 1.0, 1e5, 1e-9)        # nr < 3
    with pytest.raises(GasxError):
        ctx.ion_es_push([s], 8, 8, 1e-4, 1.0, -1.0, 1e-9)       # T_e <= 0
