"""gasx: MI355X-native electron PIC step of the GAS-ATENUATOR X-ray gas
attenuator model (axisymmetric r-z FDTD, relativistic push, N2 Monte Carlo
collisions, fixed-point current deposition) behind the C ABI of
include/gasx.h.

The compute runs in libgasx.so (HIP, gfx950).  Importing this package does
not load the library; the first Context does, and raises if it is missing.
"""
from .core import (C_LIGHT, E_EM, E_MC, EPS0, K_B, M_ION, ME_EM, ME_MC, MU0, Context,
                   CounterValues, Ensemble, FieldState, GasGrid, ParticleSet, make_params)
from .adjoint import DiffRun, shell_energy
from .dsmc import CellMap, MoleculeSet, dsmc_params, seed_molecules
from .tables import raw_tables

__all__ = [
    "Context", "Ensemble", "ParticleSet", "FieldState", "GasGrid", "CounterValues", "make_params",
    "raw_tables", "DiffRun", "shell_energy", "CellMap", "MoleculeSet", "dsmc_params", "seed_molecules", "C_LIGHT", "EPS0", "MU0", "E_EM", "ME_EM", "E_MC", "ME_MC", "M_ION", "K_B",
]
