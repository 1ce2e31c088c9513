"""Host-side wrapper of libgasx.so.

Device memory comes from torch (plumbing only); every computation runs in
the HIP library.  The library is bound to torch's current HIP stream so that
torch copies and library kernels are ordered.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _abi
from ._abi import Counters, Fields, Gas, Params, RawTables, RunOpts, Species, check
from .tables import RawTableHandle, raw_tables

# Physical constants as the reference spells them per file.
C_LIGHT = 3e8                       # electricmagneticfield6.m:8 (everywhere)
EPS0 = 8.854187817e-12              # electricmagneticfield6.m:9
MU0 = 4 * math.pi * 1e-7            # electricmagneticfield6.m:10
E_EM = 1.6e-19                      # electricmagneticfield6.m:11
ME_EM = 9.1096e-31                  # electricmagneticfield6.m:5
E_MC = 1.6022e-19                   # electron_transport.cu:47
ME_MC = 9.109e-31                   # electron_transport.cu:49
M_ION = 4.65e-26                    # electron_transport.cu:50 / electricmagneticfield6.m:6
K_B = 1.38064853e-23                # electron_transport.cu:51


def make_params(nr: int, nz: int, dr: float = 1e-5, dz: float = 1e-5, phip: float = 0.1,
                ddt: float = 1e-14, n_sub: int = 10, fn: float = 1.0, ez_metric: int = 0,
                axis_j_sign: int = 1, exc_form: int = 0, rot_eps: float = 1e-5) -> Params:
    """Parameters of the r-z PIC step.  Where the reference's files disagree
    the defaults are: ez_metric=0 -> 1/(2r) at r = i*dr (electricmagneticfield3.m:187
    rather than EM3.cu:39), axis current sign +1 as EM3.cu:35, the excitation
    form of find_CS0.m:52, and the rotate.m:5 degeneracy threshold 1e-5."""
    return Params(nr, nz, dr, dz, phip, ddt, n_sub, fn, C_LIGHT, EPS0, MU0, E_EM, ME_EM,
                  E_MC, ME_MC, M_ION, K_B, ez_metric, axis_j_sign, exc_form, rot_eps)


def _ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    assert t.is_contiguous()
    return t.data_ptr()


class ParticleSet:
    """FP32 SoA particle population in device memory (rows x, y, z, vx, vy, vz)."""

    def __init__(self, capacity: int, qnum: float = -1.0, mass: float = ME_EM, device="cuda"):
        if capacity < 1:
            raise ValueError("capacity must be positive")
        self.capacity = int(capacity)
        self.qnum = float(qnum)
        self.mass = float(mass)
        self.data = torch.zeros((6, self.capacity), dtype=torch.float32, device=device)
        self.count_t = torch.zeros(1, dtype=torch.int32, device=device)

    @property
    def x(self): return self.data[0]
    @property
    def y(self): return self.data[1]
    @property
    def z(self): return self.data[2]
    @property
    def vx(self): return self.data[3]
    @property
    def vy(self): return self.data[4]
    @property
    def vz(self): return self.data[5]

    def struct(self) -> Species:
        d = self.data
        return Species(d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), d[3].data_ptr(),
                       d[4].data_ptr(), d[5].data_ptr(), self.count_t.data_ptr(), self.capacity,
                       self.qnum, self.mass)

    def load(self, x, y, z, vx, vy, vz) -> "ParticleSet":
        n = len(x)
        if n > self.capacity:
            raise ValueError("more particles than capacity")
        for row, a in enumerate((x, y, z, vx, vy, vz)):
            t = torch.as_tensor(np.asarray(a), dtype=torch.float32)
            self.data[row, :n].copy_(t.to(self.data.device))
        self.count_t.fill_(n)
        return self

    @property
    def n(self) -> int:
        return int(self.count_t.item())

    def state(self, live_only: bool = False) -> np.ndarray:
        """(6, n) float32 host copy; absorbed slots have z = NaN."""
        a = self.data[:, : self.n].cpu().numpy()
        if live_only:
            a = a[:, ~np.isnan(a[2])]
        return a


class FieldState:
    """FP64 field arrays in the reference layout (r fastest):
    er (nz+1, nr), ez (nz, nr+1), hphi (nz, nr) as row-major tensors."""

    def __init__(self, nr: int, nz: int, device="cuda"):
        self.nr, self.nz = nr, nz
        self.er = torch.zeros((nz + 1, nr), dtype=torch.float64, device=device)
        self.ez = torch.zeros((nz, nr + 1), dtype=torch.float64, device=device)
        self.hphi = torch.zeros((nz, nr), dtype=torch.float64, device=device)

    def struct(self) -> Fields:
        return Fields(self.er.data_ptr(), self.ez.data_ptr(), self.hphi.data_ptr())

    def numpy(self):
        return self.er.cpu().numpy(), self.ez.cpu().numpy(), self.hphi.cpu().numpy()


class GasGrid:
    """Neutral density n(r, z) (m^-3) on a gnz x gnr grid of cells gdr x gdz."""

    def __init__(self, density, gdr: float, gdz: float, temperature=300.0, device="cuda"):
        d = torch.as_tensor(np.asarray(density, dtype=np.float32)).to(device)
        if d.dim() == 0:
            d = d.reshape(1, 1)
        self.density = d.contiguous()
        self.gdr, self.gdz = float(gdr), float(gdz)
        if np.ndim(temperature) == 0:
            self.temperature = None
            self.t_scalar = float(temperature)
        else:
            self.temperature = torch.as_tensor(np.asarray(temperature, dtype=np.float32)).to(device).contiguous()
            self.t_scalar = 0.0
        self.n_max = float(self.density.max().item())

    @classmethod
    def uniform(cls, n: float, rp: float, lp: float, temperature=300.0, device="cuda"):
        return cls(np.full((1, 1), n, dtype=np.float32), rp, lp, temperature, device)

    def struct(self) -> Gas:
        gnz, gnr = self.density.shape
        return Gas(self.density.data_ptr(), _ptr(self.temperature), gnr, gnz, self.gdr, self.gdz,
                   self.t_scalar, self.n_max)


@dataclass
class CounterValues:
    n_elastic: int
    n_ionize: int
    n_excite: int
    n_dropped_sec: int
    n_dropped_ion: int
    n_bound_violations: int
    n_absorbed: int
    n_saturated: int
    n_pushed: int


class Context:
    """One libgasx context on one device, bound to torch's current stream."""

    def __init__(self, params: Params, device: int = 0, tables=None):
        self.lib = _abi.load()
        self.params = params
        self.device = device
        self._tabs = RawTableHandle(tables if tables is not None else raw_tables(), RawTables)
        ctx = C.c_void_p()
        st = self.lib.gasx_ctx_create(device, C.byref(params), C.byref(self._tabs.struct), C.byref(ctx))
        check(self.lib, None, st)
        self.ctx = ctx
        with torch.cuda.device(device):
            stream = torch.cuda.current_stream().cuda_stream
        self._check(self.lib.gasx_ctx_set_stream(self.ctx, C.c_void_p(stream)))

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.gasx_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int):
        check(self.lib, self.ctx, st)

    # ---- fields --------------------------------------------------------------
    def fdtd(self, f: FieldState, jr: torch.Tensor, jz: torch.Tensor, n_push: int = 1):
        fs = f.struct()
        self._check(self.lib.gasx_fdtd(self.ctx, C.byref(fs), _ptr(jr), _ptr(jz), n_push))

    def fdtd_transpose(self, lam: FieldState, ljr: torch.Tensor, ljz: torch.Tensor):
        """Transposed field cycle of one push (the adjoint's): lam <- M^T lam,
        (ljr, ljz) <- B^T lam."""
        fs = lam.struct()
        self._check(self.lib.gasx_fdtd_transpose(self.ctx, C.byref(fs), _ptr(ljr), _ptr(ljz)))

    def fields_sync(self, f: FieldState):
        fs = f.struct()
        self._check(self.lib.gasx_fields_sync(self.ctx, C.byref(fs)))

    # ---- particle kernels ---------------------------------------------------------
    def half_drift(self, ps: ParticleSet):
        s = ps.struct()
        self._check(self.lib.gasx_half_drift(self.ctx, C.byref(s)))

    def deposit(self, ps: ParticleSet, qr: torch.Tensor | None = None, qz: torch.Tensor | None = None):
        p = self.params
        if qr is None:
            qr = torch.zeros(p.nr * p.nz, dtype=torch.int64, device=ps.data.device)
            qz = torch.zeros_like(qr)
        s = ps.struct()
        self._check(self.lib.gasx_deposit(self.ctx, C.byref(s), _ptr(qr), _ptr(qz)))
        return qr, qz

    def current_from_fixed(self, qr: torch.Tensor, qz: torch.Tensor):
        jr = torch.empty(qr.shape, dtype=torch.float64, device=qr.device)
        jz = torch.empty_like(jr)
        self._check(self.lib.gasx_current_from_fixed(self.ctx, _ptr(qr), _ptr(qz), _ptr(jr), _ptr(jz)))
        return jr, jz

    def push(self, ps: ParticleSet, f: FieldState | None = None):
        s = ps.struct()
        fs = f.struct() if f is not None else Fields()
        self._check(self.lib.gasx_push(self.ctx, C.byref(s), C.byref(fs)))

    def em_step(self, species: list[ParticleSet], f: FieldState, n_push: int):
        arr = (Species * len(species))(*[s.struct() for s in species])
        fs = f.struct()
        self._check(self.lib.gasx_em_step(self.ctx, arr, len(species), C.byref(fs), n_push))

    # ---- collisions --------------------------------------------------------------
    def cross_sections(self, e_kev: torch.Tensor):
        e_kev = e_kev.to(torch.float32).contiguous()
        se, si, sx = (torch.empty_like(e_kev) for _ in range(3))
        self._check(self.lib.gasx_cross_sections(self.ctx, _ptr(e_kev), e_kev.numel(), _ptr(se),
                                                 _ptr(si), _ptr(sx)))
        return se, si, sx

    def collide(self, e: ParticleSet, gas: GasGrid, dt_coll: float, seed: int, step: int,
                run: int = 0, ions: ParticleSet | None = None):
        s = e.struct()
        g = gas.struct()
        si = ions.struct() if ions is not None else None
        self._check(self.lib.gasx_collide(self.ctx, C.byref(s), C.byref(si) if si is not None else None,
                                          C.byref(g), dt_coll, seed, step, run))

    def collision_bound(self, gas: GasGrid, dt_coll: float) -> float:
        g = gas.struct()
        out = C.c_double()
        self._check(self.lib.gasx_collision_bound(self.ctx, C.byref(g), dt_coll, C.byref(out)))
        return out.value

    # ---- photoionization source -----------------------------------------------------
    def photoionize(self, e: ParticleSet, gas: GasGrid, photon_kev: float, n_photons: float,
                    beam_radius: float, ions: ParticleSet | None = None, seed: int = 0, pulse: int = 0,
                    run: int = 0, deplete: bool = False, auger_kev: float = 0.383,
                    atomic_mass: float = 14.0, tables: dict | None = None) -> dict:
        """INCIDENCE1.m: photoelectrons (+ Auger electrons) appended to e, ions
        to `ions`; with deplete the gas grid loses the ionized molecules and
        their thermal energy (in place; a scalar temperature becomes a grid).
        Returns the gasx_photo_result fields; electrons or ions beyond the
        capacities are dropped and counted (n_dropped_*)."""
        from .tables import source_tables
        t = tables if tables is not None else source_tables()
        be = np.ascontiguousarray(t["binding_ev"], dtype=np.float64)
        mac = np.ascontiguousarray(t["mac"], dtype=np.float64)
        if deplete and gas.temperature is None:
            gas.temperature = torch.full_like(gas.density, gas.t_scalar)
        PD = C.POINTER(C.c_double)
        src = _abi.PhotoSource(photon_kev, n_photons, beam_radius, be.ctypes.data_as(PD), be.size,
                               mac.ctypes.data_as(PD), mac.shape[0], atomic_mass, auger_kev, seed, pulse, run,
                               _ptr(gas.density) if deplete else None,
                               _ptr(gas.temperature) if deplete else None)
        res = _abi.PhotoResult()
        s = e.struct()
        si = ions.struct() if ions is not None else None
        g = gas.struct()
        st = self.lib.gasx_photoionize(self.ctx, C.byref(s), C.byref(si) if si is not None else None,
                                       C.byref(g), C.byref(src), C.byref(res))
        if st not in (_abi.OK, _abi.E_CAPACITY):
            self._check(st)
        return {n: getattr(res, n) for n, _ in _abi.PhotoResult._fields_}

    def ion_es_push(self, ions: list[ParticleSet], nr: int, nz: int, delta: float, weight: float, t_e: float,
                    dt: float, n_e0: float = 1e15, tol: float = 0.01, max_iter: int = 1_000_000,
                    phip: float | None = None) -> dict:
        """ion_staticelectrofield_push.m with Poisson.m: the ion species are
        pushed in place; returns the gasx_es_result fields and the grids phi,
        rho_i, rho_e, e_r, e_z (torch FP64, nz x nr)."""
        dev = ions[0].data.device
        grids = {k: torch.empty((nz, nr), dtype=torch.float64, device=dev)
                 for k in ("phi", "rho_i", "rho_e", "e_r", "e_z")}
        arr = (Species * len(ions))(*[s.struct() for s in ions])
        p = _abi.EsParams(nr, nz, delta, self.params.phip if phip is None else phip, weight, t_e, dt, n_e0, tol,
                          max_iter)
        res = _abi.EsResult()
        self._check(self.lib.gasx_ion_es_push(self.ctx, arr, len(ions), C.byref(p), *[_ptr(grids[k]) for k in
                                              ("phi", "rho_i", "rho_e", "e_r", "e_z")], C.byref(res)))
        out = {n: getattr(res, n) for n, _ in _abi.EsResult._fields_}
        out.update(grids)
        return out

    # ---- production loop ---------------------------------------------------------
    def run(self, e: ParticleSet, f: FieldState, gas: GasGrid | None, n_steps: int,
            collide: bool = True, secondaries: bool = False, sort_every: int = 0, seed: int = 0,
            run_id: int = 0, step0: int = 0, ions: ParticleSet | None = None, use_graph: bool = False,
            time_kernels: bool | int = False, time_every: int = 1):
        s = e.struct()
        fs = f.struct()
        g = gas.struct() if gas is not None else None
        si = ions.struct() if ions is not None else None
        o = RunOpts(n_steps, int(collide), int(secondaries), sort_every, seed, run_id, step0,
                    int(use_graph), int(time_kernels), time_every)
        self._check(self.lib.gasx_run(self.ctx, C.byref(s), C.byref(si) if si is not None else None,
                                      C.byref(fs), C.byref(g) if g is not None else None, C.byref(o)))

    def timing(self) -> dict:
        t = _abi.Timing()
        self._check(self.lib.gasx_get_timing(self.ctx, C.byref(t)))
        return {n: getattr(t, n) for n, _ in _abi.Timing._fields_}

    def sort(self, ps: ParticleSet):
        s = ps.struct()
        self._check(self.lib.gasx_sort(self.ctx, C.byref(s)))

    def tallies(self, f: FieldState, species: list[ParticleSet]):
        arr = (Species * max(1, len(species)))(*[s.struct() for s in species])
        fs = f.struct()
        out = (C.c_double * 4)()
        self._check(self.lib.gasx_tallies(self.ctx, C.byref(fs), arr, len(species), out))
        return tuple(out)

    def counters(self, reset: bool = False) -> CounterValues:
        c = Counters()
        st = self.lib.gasx_get_counters(self.ctx, C.byref(c), int(reset))
        vals = CounterValues(*[getattr(c, n) for n, _ in Counters._fields_])
        if st not in (_abi.OK, _abi.E_CAPACITY):
            self._check(st)
        return vals

    def synchronize(self):
        self._check(self.lib.gasx_synchronize(self.ctx))

    # ---- DSMC neutral gas ------------------------------------------------------------
    @property
    def dsmc(self):
        """The DSMC entry points (gasx.dsmc.Dsmc) of this context."""
        from .dsmc import Dsmc
        return Dsmc(self)


class Ensemble:
    """Independent runs on the context's grid, advanced together by
    gasx_batch_run: one particle launch and one field launch per step cover
    every run.  Run r owns pooled slots [base[r], base[r] + capacity[r]), its
    own fields, gas density (gnz x gnr) and Philox run id (run_id + r).  With
    ion_capacities, run r also owns ion slots [ion_base[r], ion_base[r] +
    ion_capacity[r]) of an ion pool, where the N2+ ions of its ionizations are
    appended as gasx_run appends them to its ion species."""

    def __init__(self, ctx: Context, capacities, densities, gdr: float, gdz: float,
                 t_scalar: float = 300.0, device="cuda", ion_capacities=None):
        self.ctx = ctx
        lib = ctx.lib
        align = lib.gasx_batch_align()
        self.nruns = len(capacities)
        if not 1 <= self.nruns <= 1024:
            raise ValueError("an ensemble holds 1..1024 runs")
        self.capacity = np.asarray([int(c) for c in capacities], dtype=np.int64)
        if np.any(self.capacity < 1):
            raise ValueError("capacities must be positive")
        padded = (self.capacity + align - 1) // align * align
        self.base = np.concatenate([[0], np.cumsum(padded)[:-1]]).astype(np.int64)
        total = int(padded.sum())
        p = ctx.params
        self.nr, self.nz = p.nr, p.nz
        self.pool = torch.zeros((6, total), dtype=torch.float32, device=device)
        self.counts = torch.zeros(self.nruns, dtype=torch.int32, device=device)
        self.er = torch.zeros((self.nruns, p.nz + 1, p.nr), dtype=torch.float64, device=device)
        self.ez = torch.zeros((self.nruns, p.nz, p.nr + 1), dtype=torch.float64, device=device)
        self.hphi = torch.zeros((self.nruns, p.nz, p.nr), dtype=torch.float64, device=device)
        d = [np.atleast_2d(np.asarray(x, dtype=np.float32)) for x in densities]
        if len(d) != self.nruns or any(x.shape != d[0].shape for x in d):
            raise ValueError("one density grid of a common shape per run")
        self.density = torch.from_numpy(np.stack(d)).to(device).contiguous()
        self.n_max = np.asarray([float(x.max()) for x in d], dtype=np.float64)
        self.gnz, self.gnr = d[0].shape
        self.gdr, self.gdz, self.t_scalar = float(gdr), float(gdz), float(t_scalar)
        self.ion_pool = self.ion_counts = None
        if ion_capacities is not None:
            self.ion_capacity = np.asarray([int(c) for c in ion_capacities], dtype=np.int64)
            if self.ion_capacity.shape != (self.nruns,) or np.any(self.ion_capacity < 1):
                raise ValueError("one positive ion capacity per run")
            ipad = (self.ion_capacity + align - 1) // align * align
            self.ion_base = np.concatenate([[0], np.cumsum(ipad)[:-1]]).astype(np.int64)
            self.ion_pool = torch.zeros((6, int(ipad.sum())), dtype=torch.float32, device=device)
            self.ion_counts = torch.zeros(self.nruns, dtype=torch.int32, device=device)

    def load_run(self, r: int, x, y, z, vx, vy, vz):
        n = len(x)
        if n > self.capacity[r]:
            raise ValueError("more particles than the run's capacity")
        b = int(self.base[r])
        for row, a in enumerate((x, y, z, vx, vy, vz)):
            self.pool[row, b: b + n].copy_(torch.as_tensor(np.asarray(a), dtype=torch.float32))
        self.counts[r] = n

    def run_state(self, r: int) -> np.ndarray:
        """(6, n) float32 host copy of run r (absorbed slots have z = NaN)."""
        n = int(self.counts[r].item())
        b = int(self.base[r])
        return self.pool[:, b: b + n].cpu().numpy()

    def run_ions(self, r: int) -> np.ndarray:
        """(6, n) float32 host copy of run r's ions."""
        if self.ion_pool is None:
            raise ValueError("the ensemble has no ion pool")
        n = int(self.ion_counts[r].item())
        b = int(self.ion_base[r])
        return self.ion_pool[:, b: b + n].cpu().numpy()

    def struct(self) -> "_abi.Batch":
        d = self.pool
        ions = self.ion_pool is not None
        self._keep = (np.ascontiguousarray(self.base), np.ascontiguousarray(self.capacity),
                      np.ascontiguousarray(self.n_max),
                      np.ascontiguousarray(self.ion_base) if ions else None,
                      np.ascontiguousarray(self.ion_capacity) if ions else None)
        P64, PD = C.POINTER(C.c_int64), C.POINTER(C.c_double)
        ion_args = ([self.ion_pool[k].data_ptr() for k in range(6)] +
                    [self._keep[3].ctypes.data_as(P64), self._keep[4].ctypes.data_as(P64),
                     self.ion_counts.data_ptr()]) if ions else [None] * 6 + [P64(), P64(), None]
        return _abi.Batch(self.nruns, d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(),
                          d[3].data_ptr(), d[4].data_ptr(), d[5].data_ptr(),
                          self._keep[0].ctypes.data_as(P64), self._keep[1].ctypes.data_as(P64),
                          self.counts.data_ptr(), self.er.data_ptr(), self.ez.data_ptr(),
                          self.hphi.data_ptr(), self.density.data_ptr(),
                          self._keep[2].ctypes.data_as(PD), self.gnr, self.gnz, self.gdr, self.gdz,
                          self.t_scalar, *ion_args)

    @staticmethod
    def _opts(n_steps, collide, secondaries, seed, run_id, step0, use_graph, time_kernels):
        return RunOpts(n_steps, int(collide), int(secondaries), 0, seed, run_id, step0,
                       int(use_graph), int(time_kernels), 1)

    def run(self, n_steps: int, collide: bool = True, secondaries: bool = False, seed: int = 0,
            run_id: int = 0, step0: int = 0, use_graph: bool = False, time_kernels: bool = False):
        b = self.struct()
        o = self._opts(n_steps, collide, secondaries, seed, run_id, step0, use_graph, time_kernels)
        self.ctx._check(self.ctx.lib.gasx_batch_run(self.ctx.ctx, C.byref(b), C.byref(o)))

    def autotune(self, trial_steps: int = 5, collide: bool = True, secondaries: bool = False,
                 seed: int = 0, run_id: int = 0) -> dict:
        b = self.struct()
        o = self._opts(trial_steps, collide, secondaries, seed, run_id, 0, False, False)
        t = _abi.BatchTuning()
        self.ctx._check(self.ctx.lib.gasx_batch_autotune(self.ctx.ctx, C.byref(b), C.byref(o),
                                                         trial_steps, C.byref(t)))
        return {n: getattr(t, n) for n, _ in _abi.BatchTuning._fields_}

    def set_tuning(self, fdtd_ti: int, fdtd_tk: int, particle_blocks_per_cu: int):
        t = _abi.BatchTuning(fdtd_ti, fdtd_tk, particle_blocks_per_cu)
        self.ctx._check(self.ctx.lib.gasx_batch_set_tuning(self.ctx.ctx, C.byref(t)))

    def tuning(self) -> dict:
        t = _abi.BatchTuning()
        self.ctx._check(self.ctx.lib.gasx_batch_get_tuning(self.ctx.ctx, C.byref(t)))
        return {n: getattr(t, n) for n, _ in _abi.BatchTuning._fields_}

    def counters(self, reset: bool = False) -> list[CounterValues]:
        arr = (Counters * self.nruns)()
        self.ctx._check(self.ctx.lib.gasx_batch_counters(self.ctx.ctx, self.nruns, arr, int(reset)))
        return [CounterValues(*[getattr(c, n) for n, _ in Counters._fields_]) for c in arr]
