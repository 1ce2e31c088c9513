"""DSMC neutral gas over libgasx.so: the molecule step of main2.m:247-271
(free_motion.m, boundary_conditions_specular.m, put_into_cells_symmetry.m,
collision_HS_internal.m) on the cell map of cell_structure.m, and the gas
density / temperature it hands to the electron step (particle_density0.m,
count_T_numbensity.m).

Device memory comes from torch; every computation runs in the HIP library.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _abi

# main2.m:5-17,35-36
K_B_DSMC = 1.38064853e-23
N2_MASS = 4.65e-26
N2_DIAMETER = 3.7e-10
VHS_DIAREF = 4.17e-10
OMEGA = 0.74
ZETA = 2.0


@dataclass
class CellMap:
    """cell_structure.m's (r, phi, z) map of the wedge: n_rings rings of width
    Rp / n_rings, ring i cut into ceil((i+1) / n_rings Rp PHIp / delta) phi
    cells, n_layers z layers."""
    n_rings: int
    n_layers: int
    rp: float
    lp: float
    phip: float
    delta: float

    def struct(self) -> _abi.CellMap:
        return _abi.CellMap(self.n_rings, self.n_layers, self.rp, self.lp, self.phip, self.delta)

    def info(self, lib) -> tuple[np.ndarray, int]:
        pd = (C.c_int32 * self.n_rings)()
        nxy = C.c_int64()
        m = self.struct()
        st = lib.gasx_cell_map_info(C.byref(m), pd, C.byref(nxy))
        if st != _abi.OK:
            raise _abi.GasxError(st, "invalid cell map")
        return np.frombuffer(pd, dtype=np.int32).copy(), nxy.value

    @property
    def dr(self) -> float:
        return self.rp / self.n_rings

    @property
    def dz(self) -> float:
        return self.lp / self.n_layers


def dsmc_params(dt: float, fn: float, mass: float = N2_MASS, diameter: float = N2_DIAMETER,
                sc_max: float | None = None, t0: float = 300.0, omega: float = OMEGA, zeta: float = ZETA,
                alpha: float = 1.0, min_attempts: int = 1, reorder: bool = False) -> _abi.DsmcParams:
    """Parameters of collision_HS_internal.m.  sc_max defaults to main2.m:35,
    15 sigma_ref sqrt(kB T0 / m) with sigma_ref = pi * 4.17e-10^2."""
    if sc_max is None:
        sc_max = 15.0 * (math.pi * VHS_DIAREF ** 2) * math.sqrt(K_B_DSMC * t0 / mass)
    return _abi.DsmcParams(dt, fn, mass, diameter, sc_max, omega, zeta, alpha, int(min_attempts), int(reorder))


class MoleculeSet:
    """FP32 SoA molecules in device memory (rows x, y, z, vx, vy, vz, ei)."""

    def __init__(self, capacity: int, device="cuda"):
        if capacity < 1:
            raise ValueError("capacity must be positive")
        self.capacity = int(capacity)
        self.data = torch.zeros((7, self.capacity), dtype=torch.float32, device=device)
        self.count_t = torch.zeros(1, dtype=torch.int32, device=device)

    def struct(self) -> _abi.Molecules:
        d = self.data
        return _abi.Molecules(*[d[k].data_ptr() for k in range(7)], self.count_t.data_ptr(), self.capacity)

    def load(self, state) -> "MoleculeSet":
        """state: (7, n) rows x, y, z, vx, vy, vz, ei."""
        a = np.asarray(state)
        n = a.shape[1]
        if n > self.capacity:
            raise ValueError("more molecules than capacity")
        self.data[:, :n].copy_(torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32)).to(self.data.device))
        self.count_t.fill_(n)
        return self

    @property
    def n(self) -> int:
        return int(self.count_t.item())

    def state(self) -> np.ndarray:
        return self.data[:, : self.n].cpu().numpy()


def seed_molecules(n: int, cmap: CellMap, t0: float = 300.0, mass: float = N2_MASS, seed: int = 0) -> np.ndarray:
    """main2.m:59-106: r = Rp U^0.5, z = Lp U, phi = PHIp U, velocity components
    N(0, sqrt(kB T0 / m)), internal energy -kB T0 log U (synthetic streams:
    numpy in place of the reference's Park-Miller sequence and normrnd).
    Returns (7, n) float32 rows x, y, z, vx, vy, vz, ei."""
    rng = np.random.default_rng(seed)
    r = cmap.rp * rng.random(n) ** 0.5
    z = rng.random(n) * cmap.lp
    phi = rng.random(n) * cmap.phip
    sig = math.sqrt(K_B_DSMC * t0 / mass)
    v = rng.normal(0.0, sig, (3, n))
    ei = -K_B_DSMC * t0 * np.log(1.0 - rng.random(n))
    st = np.stack([r * np.cos(phi), r * np.sin(phi), z, v[0], v[1], v[2], ei]).astype(np.float32)
    # keep every molecule strictly inside the wedge after the FP32 rounding
    st[2] = np.clip(st[2], np.float32(1e-9) * cmap.lp, np.nextafter(np.float32(cmap.lp), np.float32(0)))
    return st


@dataclass
class DsmcCounterValues:
    n_attempts: int
    n_collisions: int
    n_sc_exceed: int
    n_boundary: int
    n_wall_r: int
    n_wall_z: int
    n_wrap: int


def _p(t):
    return None if t is None else t.data_ptr()


class Dsmc:
    """DSMC entry points of one libgasx Context."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.lib = ctx.lib

    def cells(self, mol: MoleculeSet, cmap: CellMap) -> torch.Tensor:
        out = torch.empty(mol.capacity, dtype=torch.int32, device=mol.data.device)
        m, cm = mol.struct(), cmap.struct()
        self.ctx._check(self.lib.gasx_dsmc_cells(self.ctx.ctx, C.byref(m), C.byref(cm), out.data_ptr()))
        return out[: mol.n]

    def move(self, mol: MoleculeSet, cmap: CellMap, dt: float):
        m, cm = mol.struct(), cmap.struct()
        self.ctx._check(self.lib.gasx_dsmc_move(self.ctx.ctx, C.byref(m), C.byref(cm), dt))

    def collide(self, mol: MoleculeSet, cmap: CellMap, p: _abi.DsmcParams, seed: int, step: int):
        m, cm = mol.struct(), cmap.struct()
        self.ctx._check(self.lib.gasx_dsmc_collide(self.ctx.ctx, C.byref(m), C.byref(cm), C.byref(p), seed, step))

    def step(self, mol: MoleculeSet, cmap: CellMap, p: _abi.DsmcParams, seed: int, step: int):
        m, cm = mol.struct(), cmap.struct()
        self.ctx._check(self.lib.gasx_dsmc_step(self.ctx.ctx, C.byref(m), C.byref(cm), C.byref(p), seed, step))

    def sample(self, mol: MoleculeSet, cmap: CellMap, fn: float, mass: float = N2_MASS, cells: bool = True,
               rz: bool = True):
        """Per-cell and per-(ring, layer) density (m^-3) and temperature (K),
        torch float32: cell arrays (n_cells,), rz arrays (n_layers, n_rings)."""
        _, nxy = cmap.info(self.lib)
        dev = mol.data.device
        cd = torch.empty(nxy * cmap.n_layers, dtype=torch.float32, device=dev) if cells else None
        ct = torch.empty_like(cd) if cells else None
        rd = torch.empty((cmap.n_layers, cmap.n_rings), dtype=torch.float32, device=dev) if rz else None
        rt = torch.empty_like(rd) if rz else None
        m, cm = mol.struct(), cmap.struct()
        self.ctx._check(self.lib.gasx_dsmc_sample(self.ctx.ctx, C.byref(m), C.byref(cm), fn, mass, _p(cd), _p(ct),
                                                  _p(rd), _p(rt)))
        return {"cell_density": cd, "cell_temperature": ct, "rz_density": rd, "rz_temperature": rt}

    def counters(self, reset: bool = False) -> DsmcCounterValues:
        c = _abi.DsmcCounters()
        self.ctx._check(self.lib.gasx_dsmc_get_counters(self.ctx.ctx, C.byref(c), int(reset)))
        return DsmcCounterValues(*[getattr(c, n) for n, _ in _abi.DsmcCounters._fields_])
