"""The reference's MATLAB entry points of the electron step, with the same
names, argument order and meaning, running on libgasx.

* ``electricmagneticfield6``  -> gasx_em_step   (electricmagneticfield6.m:1)
* ``find_CS0``                -> gasx_cross_sections (find_CS0.m:1)
* ``electron_collision``      -> gasx_collide   (electron_collision.m:1)

Arrays follow the reference's shapes: particle positions and velocities are
N x 3 (Cartesian x, y, z in the wedge frame, the state electron_collision.m:74
returns), field arrays are MATLAB's NR x (NZ+1), (NR+1) x NZ and NR x NZ
(numpy/torch arrays of that shape; they are transposed to the library's
r-fastest layout on the way in and back on the way out).  numpy arrays and
torch tensors (CPU or GPU) are accepted; results come back as the same kind.
Host arrays cross PCIe on every call, as the reference's implicit gpuArray
uploads did; callers that keep state on the GPU use ``gasx.Context``
directly.

Errors follow the reference's: wrong shapes raise ValueError (MATLAB's
size-mismatch exceptions); there is no silent CPU path.
"""
from __future__ import annotations

import numpy as np
import torch

from .core import (C_LIGHT, E_EM, M_ION, ME_EM, Context, FieldState, GasGrid, ParticleSet,
                   make_params)
from .tables import raw_tables

RATIO = 5            # electricmagneticfield6.m:134: ddt = 5 dt, 2*ratio field substeps of dt/2


class _Sessions:
    """Contexts cached per (grid, tables) so repeated calls reuse tables and buffers."""

    def __init__(self):
        self._cache = {}
        self._step = 0

    def get(self, key, factory):
        ctx = self._cache.get(key)
        if ctx is None:
            ctx = self._cache[key] = factory()
        return ctx

    def next_step(self) -> int:
        self._step += 1
        return self._step


_SESSIONS = _Sessions()


def _as_np(a) -> np.ndarray:
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def _like(ref, a: np.ndarray):
    if isinstance(ref, torch.Tensor):
        return torch.from_numpy(np.ascontiguousarray(a)).to(ref.device)
    return a


def _nx3(name, a) -> np.ndarray:
    a = _as_np(a)
    if a.size == 0:
        return np.zeros((0, 3))
    if a.ndim != 2 or a.shape[1] != 3:
        raise ValueError(f"{name} must be N x 3, got {a.shape}")
    return a


def _species(pos, vel, qnum, mass, extra=0):
    n = pos.shape[0]
    ps = ParticleSet(max(1, n + extra), qnum=qnum, mass=mass)
    if n:
        ps.load(pos[:, 0], pos[:, 1], pos[:, 2], vel[:, 0], vel[:, 1], vel[:, 2])
    return ps


def _kinetic_kev(v: np.ndarray) -> np.ndarray:
    """(gamma - 1) m c^2 in keV with the constants of electricmagneticfield6.m:5,8,11,172."""
    v2 = np.sum(v.astype(np.float64) ** 2, axis=1)
    b2 = v2 / C_LIGHT ** 2
    g = 1.0 / np.sqrt(1.0 - b2)
    return (b2 * g * g / (1.0 + g)) * ME_EM * C_LIGHT ** 2 / (1000 * E_EM)


def electricmagneticfield6(electron_position0, electron_velocity, ion1_position0, ion1_velocity,
                           ion2_position0, ion2_velocity, Er, Ez, Hphi, delta, dt, Rp, Lp, PHIp, FN,
                           cycle_time_for_electromagneticfield, *, device: int = 0):
    """electricmagneticfield6.m:1-190: cycle_time/5 pushes of step 5 dt, each
    EM2 -> EM7 -> field cycle (10 substeps of dt/2) -> EM4, for electrons,
    N2+ and N2++.  Absorbed particles are removed from each species (the
    electron masks of :176-178 applied per species).

    Returns (electron_position, electron_velocity, ion1_position, ion1_velocity,
    ion2_position, ion2_velocity, Er, Ez, Hphi, electron_kinetic [keV])."""
    nr, nz = int(round(Rp / delta)), int(round(Lp / delta))
    er_np, ez_np, h_np = _as_np(Er), _as_np(Ez), _as_np(Hphi)
    for name, a, shp in (("Er", er_np, (nr, nz + 1)), ("Ez", ez_np, (nr + 1, nz)), ("Hphi", h_np, (nr, nz))):
        if a.shape != shp:
            raise ValueError(f"{name} must be {shp} (NR = Rp/delta, NZ = Lp/delta), got {a.shape}")
    n_push = int(cycle_time_for_electromagneticfield) // RATIO
    p = make_params(nr, nz, dr=delta, dz=delta, phip=PHIp, ddt=RATIO * dt, n_sub=2 * RATIO, fn=FN)
    key = ("em", nr, nz, float(delta), float(PHIp), float(dt), float(FN), device)
    ctx = _SESSIONS.get(key, lambda: Context(p, device=device))
    pos = [_nx3("electron_position0", electron_position0), _nx3("ion1_position0", ion1_position0),
           _nx3("ion2_position0", ion2_position0)]
    vel = [_nx3("electron_velocity", electron_velocity), _nx3("ion1_velocity", ion1_velocity),
           _nx3("ion2_velocity", ion2_velocity)]
    for a, b in zip(pos, vel):
        if a.shape[0] != b.shape[0]:
            raise ValueError("position and velocity counts differ")
    sp = [_species(pos[0], vel[0], -1.0, ME_EM), _species(pos[1], vel[1], 1.0, M_ION),
          _species(pos[2], vel[2], 2.0, M_ION)]
    for s, a in zip(sp, pos):
        s.count_t.fill_(a.shape[0])
    f = FieldState(nr, nz)
    f.er.copy_(torch.from_numpy(np.ascontiguousarray(er_np.T)))
    f.ez.copy_(torch.from_numpy(np.ascontiguousarray(ez_np.T)))
    f.hphi.copy_(torch.from_numpy(np.ascontiguousarray(h_np.T)))
    if n_push > 0:
        ctx.em_step(sp, f, n_push)
    out = []
    for s in sp:
        st = s.state(live_only=True)
        out.append((st[:3].T.astype(np.float64), st[3:].T.astype(np.float64)))
    er2, ez2, h2 = (a.T.copy() for a in f.numpy())
    ke = _kinetic_kev(out[0][1])
    return (_like(electron_position0, out[0][0]), _like(electron_velocity, out[0][1]),
            _like(ion1_position0, out[1][0]), _like(ion1_velocity, out[1][1]),
            _like(ion2_position0, out[2][0]), _like(ion2_velocity, out[2][1]),
            _like(Er, er2), _like(Ez, ez2), _like(Hphi, h2), _like(electron_velocity, ke))


def _table_key(*arrs):
    return tuple(hash(np.ascontiguousarray(_as_np(a), dtype=np.float64).tobytes()) for a in arrs)


def find_CS0(t_el, t_sel, t_ion, t_exc, E_PE, *, device: int = 0):
    """find_CS0.m:1-61: sigma_el, sigma_ion, sigma_exc (m^2) of energies E_PE
    (keV); NaN outside the tables as MATLAB interp1."""
    tabs = dict(raw_tables())
    tabs.update(t_el=_as_np(t_el), t_sel=_as_np(t_sel), t_ion=_as_np(t_ion), t_exc=_as_np(t_exc))
    key = ("cs", device) + _table_key(t_el, t_sel, t_ion, t_exc)
    ctx = _SESSIONS.get(key, lambda: Context(make_params(2, 2), device=device, tables=tabs))
    e = torch.as_tensor(_as_np(E_PE).astype(np.float32).ravel()).cuda(device)
    se, si, sx = ctx.cross_sections(e)
    shape = np.shape(_as_np(E_PE))
    return tuple(_like(E_PE, t.cpu().numpy().astype(np.float64).reshape(shape)) for t in (se, si, sx))


def ring_density(cell_values, sum_cell, cellnumxy, nz):
    """Collapses the reference's phi-resolved gas cells (cell_structure.m:9-25:
    plane of cellnumxy cells, ring k holding cells sum_cell0[k] ..
    sum_cell0[k+1]-1, planes stacked in z) to one value per (z, r) ring: the
    mean over the ring's phi cells.  Returns an (nz, n_rings) array."""
    v = _as_np(cell_values).astype(np.float64).ravel()
    sc0 = np.concatenate([[0], np.asarray(_as_np(sum_cell)).astype(np.int64).ravel()])
    cellnumxy = int(cellnumxy)
    if sc0[-1] != cellnumxy:
        raise ValueError("sum_cell(end) must equal cellnumxy")
    if v.size < nz * cellnumxy:
        raise ValueError("fewer cell values than nz * cellnumxy")
    planes = v[: nz * cellnumxy].reshape(nz, cellnumxy)
    sums = np.add.reduceat(planes, sc0[:-1], axis=1)
    return sums / np.diff(sc0)[None, :]


def electron_collision(electron_position, electron_velocity, ion1_position, ion1_velocity, t_el,
                       t_sel_1, t_del, t_sdel, t_exc, electron_kinetic, sum_cell, cellnumxy,
                       particle_num_density_in_every_cell, temperature_in_every_cell, PHIp, phimax,
                       dr, dz, sig_elastic, sig_ionization, sig_excitation, dt_transport, *,
                       seed: int = 0, step: int | None = None, device: int = 0):
    """electron_collision.m:1-75 with electron_transport.cu: one collision check
    of duration dt_transport per electron; ionization appends a secondary
    electron (electron_collision.m:68-75) and an N2+ ion with a Maxwellian
    velocity at the parent's position (:63-66).

    The cross sections and kinetic energy are recomputed from each electron's
    speed inside the kernel, so sig_* and electron_kinetic are accepted for
    signature parity and only checked for length.  The gas is taken as one
    density per (z, r) ring: the mean of the ring's phi cells.  Random numbers
    come from Philox with key ``seed`` and counter word ``step`` (an internal
    call counter when None), where the reference seeded curand from clock()."""
    pos = _nx3("electron_position", electron_position)
    vel = _nx3("electron_velocity", electron_velocity)
    ipos = _nx3("ion1_position", ion1_position)
    ivel = _nx3("ion1_velocity", ion1_velocity)
    n = pos.shape[0]
    for name, a in (("electron_kinetic", electron_kinetic), ("sig_elastic", sig_elastic),
                    ("sig_ionization", sig_ionization), ("sig_excitation", sig_excitation)):
        if np.size(_as_np(a)) != n:
            raise ValueError(f"{name} must have one entry per electron")
    rings = len(np.ravel(_as_np(phimax)))
    cellnumxy = int(cellnumxy)
    nzc = np.size(_as_np(particle_num_density_in_every_cell)) // cellnumxy
    dens = ring_density(particle_num_density_in_every_cell, sum_cell, cellnumxy, nzc)
    temp = ring_density(temperature_in_every_cell, sum_cell, cellnumxy, nzc)
    if dens.shape[1] != rings:
        raise ValueError("phimax and sum_cell disagree on the number of rings")
    tabs = dict(raw_tables())
    tabs.update(t_el=_as_np(t_el), t_sel1=_as_np(t_sel_1), t_del=_as_np(t_del),
                t_sdel=_as_np(t_sdel), t_exc=_as_np(t_exc))
    key = ("coll", float(PHIp), device) + _table_key(t_el, t_sel_1, t_del, t_sdel, t_exc)
    ctx = _SESSIONS.get(key, lambda: Context(make_params(2, 2, phip=PHIp), device=device, tables=tabs))
    e = _species(pos, vel, -1.0, ME_EM, extra=n)
    e.count_t.fill_(n)
    ions = _species(ipos, ivel, 1.0, M_ION, extra=n)
    ions.count_t.fill_(ipos.shape[0])
    gas = GasGrid(dens.astype(np.float32), dr, dz, temperature=temp.astype(np.float32))
    st = _SESSIONS.next_step() if step is None else int(step)
    ctx.collide(e, gas, float(dt_transport), seed, st, ions=ions)
    es, is_ = e.state(), ions.state()
    return (_like(electron_position, es[:3].T.astype(np.float64)),
            _like(electron_velocity, es[3:].T.astype(np.float64)),
            _like(ion1_position, is_[:3].T.astype(np.float64)),
            _like(ion1_velocity, is_[3:].T.astype(np.float64)))


__all__ = ["electricmagneticfield6", "find_CS0", "electron_collision", "ring_density"]
