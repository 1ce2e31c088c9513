"""gasx_collide_noalloc (the allocation-free collision call) against
gasx_collide (GPU, through the C ABI).

Both calls run the same kernels on the same Philox streams; only where the
secondaries, the ions and the append scan are staged differs (the library's
fixed device array instead of context buffers), so electrons, appended
secondaries, ions and event counters must agree bit for bit.  Two calls in a
row, the second with a smaller population, show that the fixed array's
contents from the first call do not leak into the second.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import log
from test_gpu_parity import wedge_particles

pytestmark = pytest.mark.gpu

NR, NZ = 32, 64


def _species_pair(n, cap, seed):
    import gasx
    part = wedge_particles(n, NR, NZ, 1e-5, 1e-5, 0.1, seed, vmax=2.5e7, edge_frac=0.02)
    return gasx.ParticleSet(cap).load(*part), gasx.ParticleSet(cap).load(*part)


def _ions(cap):
    import gasx
    from gasx.core import M_ION
    return gasx.ParticleSet(cap, qnum=1.0, mass=M_ION)


def test_collide_noalloc_matches_collide(gpu):
    import gasx
    from gasx import _abi
    params = gasx.make_params(NR, NZ, fn=1e3)
    ctx = gasx.Context(params)
    lib = ctx.lib
    gas = gasx.GasGrid.uniform(1e26, NR * 1e-5, NZ * 1e-5)
    # first: 7e5 electrons (2735 staging segments: three scan groups) with
    # ions; second: a smaller population without ions
    for call, (n, cap, with_ions, step) in enumerate(((700_000, 1 << 20, True, 4), (20_000, 50_000, False, 9))):
        log(f"collide_noalloc call {call}: n={n} cap={cap} ions={with_ions} step={step}")
        a, b = _species_pair(n, cap, 60 + call)
        ia, ib = (_ions(cap), _ions(cap)) if with_ions else (None, None)
        ctx.counters(reset=True)
        ctx.collide(a, gas, 1e-14, seed=17, step=step, run=3, ions=ia)
        ca = ctx.counters(reset=True)
        sb, gs = b.struct(), gas.struct()
        si = ib.struct() if ib is not None else None
        st = lib.gasx_collide_noalloc(ctx.ctx, C.byref(sb), C.byref(si) if si is not None else None,
                                      C.byref(gs), 1e-14, 17, step, 3)
        assert st == _abi.OK, lib.gasx_last_error(ctx.ctx)
        cb = ctx.counters(reset=True)
        assert ca == cb
        assert ca.n_ionize > 1000 and ca.n_elastic > 1000
        assert a.n == b.n == n + ca.n_ionize - ca.n_dropped_sec
        assert np.array_equal(a.state(), b.state(), equal_nan=True)
        if with_ions:
            assert ia.n == ib.n == ca.n_ionize
            assert np.array_equal(ia.state(), ib.state(), equal_nan=True)


def test_collide_noalloc_rejects_capacity_above_limit(gpu):
    import gasx
    from gasx import _abi
    params = gasx.make_params(NR, NZ, fn=1e3)
    ctx = gasx.Context(params)
    gas = gasx.GasGrid.uniform(1e26, NR * 1e-5, NZ * 1e-5)
    big = gasx.ParticleSet((1 << 20) + 1)
    sb, gs = big.struct(), gas.struct()
    st = ctx.lib.gasx_collide_noalloc(ctx.ctx, C.byref(sb), None, C.byref(gs), 1e-14, 1, 0, 0)
    assert st == _abi.E_ARG
    assert b"GASX_NOALLOC_MAX_SLOTS" in ctx.lib.gasx_last_error(ctx.ctx)
