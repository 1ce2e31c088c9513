"""ctypes mirror of include/gasx.h and the loader of libgasx.so.

The loader raises when the HIP library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libgasx.so"

# gasx_status
OK, E_ARG, E_CAPACITY, E_HIP, E_NOMEM, E_STATE, E_NUMERIC = range(7)


class GasxError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"gasx status {status}: {msg}")
        self.status = status


class CapacityError(GasxError):
    pass


class Params(C.Structure):
    _fields_ = [
        ("nr", C.c_int32), ("nz", C.c_int32),
        ("dr", C.c_double), ("dz", C.c_double),
        ("phip", C.c_double), ("ddt", C.c_double),
        ("n_sub", C.c_int32), ("fn", C.c_double),
        ("c", C.c_double), ("eps0", C.c_double), ("mu0", C.c_double),
        ("e_em", C.c_double), ("me_em", C.c_double),
        ("e_mc", C.c_double), ("me_mc", C.c_double),
        ("m_ion", C.c_double), ("k_b", C.c_double),
        ("ez_metric", C.c_int32), ("axis_j_sign", C.c_int32), ("exc_form", C.c_int32),
        ("rot_eps", C.c_double),
    ]


class RawTables(C.Structure):
    _fields_ = [
        ("t_el", C.POINTER(C.c_double)), ("n_el", C.c_int32),
        ("t_sel", C.POINTER(C.c_double)), ("n_sel", C.c_int32),
        ("t_sel1", C.POINTER(C.c_double)), ("n_sel1", C.c_int32),
        ("t_del", C.POINTER(C.c_double)), ("w_del", C.c_int32),
        ("t_sdel", C.POINTER(C.c_double)), ("w_sdel", C.c_int32),
        ("t_ion", C.POINTER(C.c_double)), ("n_ion", C.c_int32),
        ("t_exc", C.POINTER(C.c_double)), ("n_exc", C.c_int32),
    ]


class Species(C.Structure):
    _fields_ = [
        ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
        ("vx", C.c_void_p), ("vy", C.c_void_p), ("vz", C.c_void_p),
        ("count", C.c_void_p), ("capacity", C.c_int64),
        ("qnum", C.c_float), ("mass", C.c_double),
    ]


class Fields(C.Structure):
    _fields_ = [("er", C.c_void_p), ("ez", C.c_void_p), ("hphi", C.c_void_p)]


class Gas(C.Structure):
    _fields_ = [
        ("density", C.c_void_p), ("temperature", C.c_void_p),
        ("gnr", C.c_int32), ("gnz", C.c_int32),
        ("gdr", C.c_double), ("gdz", C.c_double),
        ("t_scalar", C.c_double), ("n_max", C.c_double),
    ]


class Counters(C.Structure):
    _fields_ = [
        ("n_elastic", C.c_int64), ("n_ionize", C.c_int64), ("n_excite", C.c_int64),
        ("n_dropped_sec", C.c_int64), ("n_dropped_ion", C.c_int64),
        ("n_bound_violations", C.c_int64), ("n_absorbed", C.c_int64), ("n_saturated", C.c_int64),
        ("n_pushed", C.c_int64),
    ]


class Timing(C.Structure):
    _fields_ = [
        ("particle_ms", C.c_double), ("fdtd_ms", C.c_double), ("append_ms", C.c_double),
        ("sort_ms", C.c_double), ("n_particle_launches", C.c_int64), ("n_fdtd_launches", C.c_int64),
        ("n_append_launches", C.c_int64), ("n_sorts", C.c_int64),
        ("graph_ms", C.c_double), ("n_graph_launches", C.c_int64),
        ("drift_ms", C.c_double), ("n_drift_launches", C.c_int64),
    ]


class Batch(C.Structure):
    _fields_ = [
        ("nruns", C.c_int32),
        ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
        ("vx", C.c_void_p), ("vy", C.c_void_p), ("vz", C.c_void_p),
        ("base", C.POINTER(C.c_int64)), ("capacity", C.POINTER(C.c_int64)),
        ("counts", C.c_void_p),
        ("er", C.c_void_p), ("ez", C.c_void_p), ("hphi", C.c_void_p),
        ("density", C.c_void_p), ("n_max", C.POINTER(C.c_double)),
        ("gnr", C.c_int32), ("gnz", C.c_int32),
        ("gdr", C.c_double), ("gdz", C.c_double), ("t_scalar", C.c_double),
        ("ix", C.c_void_p), ("iy", C.c_void_p), ("iz", C.c_void_p),
        ("ivx", C.c_void_p), ("ivy", C.c_void_p), ("ivz", C.c_void_p),
        ("ion_base", C.POINTER(C.c_int64)), ("ion_capacity", C.POINTER(C.c_int64)),
        ("ion_counts", C.c_void_p),
    ]


class BatchTuning(C.Structure):
    _fields_ = [("fdtd_ti", C.c_int32), ("fdtd_tk", C.c_int32), ("particle_blocks_per_cu", C.c_int32)]


class DiffSpec(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32), ("nodes0", C.POINTER(C.c_double)), ("n_steps", C.c_int32),
        ("checkpoint_every", C.c_int32), ("seed", C.c_uint64), ("run_id", C.c_uint32),
        ("r_frac", C.c_double),
    ]


class RunOpts(C.Structure):
    _fields_ = [
        ("n_steps", C.c_int32), ("collide", C.c_int32), ("secondaries", C.c_int32),
        ("sort_every", C.c_int32), ("seed", C.c_uint64), ("run_id", C.c_uint32),
        ("step0", C.c_uint32), ("use_graph", C.c_int32), ("time_kernels", C.c_int32),
        ("time_every", C.c_int32),
    ]


class PhotoSource(C.Structure):
    _fields_ = [
        ("photon_kev", C.c_double), ("n_photons", C.c_double), ("beam_radius", C.c_double),
        ("binding_ev", C.POINTER(C.c_double)), ("n_shells", C.c_int32),
        ("mac", C.POINTER(C.c_double)), ("n_mac", C.c_int32),
        ("atomic_mass", C.c_double), ("auger_kev", C.c_double),
        ("seed", C.c_uint64), ("pulse", C.c_uint32), ("run", C.c_uint32),
        ("density_out", C.c_void_p), ("temperature_out", C.c_void_p),
    ]


class PhotoResult(C.Structure):
    _fields_ = [
        ("photons_left", C.c_double), ("n_events", C.c_int64), ("n_electrons", C.c_int64),
        ("n_ions", C.c_int64), ("n_dropped_electrons", C.c_int64), ("n_dropped_ions", C.c_int64),
        ("sigma_pe", C.c_double), ("electron_kev", C.c_double), ("auger", C.c_int32),
        ("n_rings", C.c_int32),
    ]


class EsParams(C.Structure):
    _fields_ = [
        ("nr", C.c_int32), ("nz", C.c_int32), ("delta", C.c_double), ("phip", C.c_double),
        ("weight", C.c_double), ("t_e", C.c_double), ("dt", C.c_double), ("n_e0", C.c_double),
        ("tol", C.c_double), ("max_iter", C.c_int32),
    ]


class EsResult(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32), ("converged", C.c_int32), ("n_e0", C.c_double),
        ("max_dphi", C.c_double), ("mean_phi", C.c_double), ("int_rho_i", C.c_double),
    ]


class CellMap(C.Structure):
    _fields_ = [("n_rings", C.c_int32), ("n_layers", C.c_int32), ("rp", C.c_double), ("lp", C.c_double),
                ("phip", C.c_double), ("delta", C.c_double)]


class Molecules(C.Structure):
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("vx", C.c_void_p),
                ("vy", C.c_void_p), ("vz", C.c_void_p), ("ei", C.c_void_p), ("count", C.c_void_p),
                ("capacity", C.c_int64)]


class DsmcParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("fn", C.c_double), ("mass", C.c_double), ("diameter", C.c_double),
                ("sc_max", C.c_double), ("omega", C.c_double), ("zeta", C.c_double), ("alpha", C.c_double),
                ("min_attempts", C.c_int32), ("reorder", C.c_int32)]


class DsmcCounters(C.Structure):
    _fields_ = [("n_attempts", C.c_int64), ("n_collisions", C.c_int64), ("n_sc_exceed", C.c_int64),
                ("n_boundary", C.c_int64), ("n_wall_r", C.c_int64), ("n_wall_z", C.c_int64),
                ("n_wrap", C.c_int64)]


# exported symbols and their prototypes (every declaration of include/gasx.h)
_P = C.c_void_p
PROTOTYPES = {
    "gasx_abi_version": (C.c_int, []),
    "gasx_status_string": (C.c_char_p, [C.c_int]),
    "gasx_last_error": (C.c_char_p, [_P]),
    "gasx_ctx_create": (C.c_int, [C.c_int, C.POINTER(Params), C.POINTER(RawTables), C.POINTER(_P)]),
    "gasx_ctx_destroy": (C.c_int, [_P]),
    "gasx_ctx_set_stream": (C.c_int, [_P, _P]),
    "gasx_synchronize": (C.c_int, [_P]),
    "gasx_fdtd": (C.c_int, [_P, C.POINTER(Fields), _P, _P, C.c_int]),
    "gasx_half_drift": (C.c_int, [_P, C.POINTER(Species)]),
    "gasx_deposit": (C.c_int, [_P, C.POINTER(Species), _P, _P]),
    "gasx_current_from_fixed": (C.c_int, [_P, _P, _P, _P, _P]),
    "gasx_push": (C.c_int, [_P, C.POINTER(Species), C.POINTER(Fields)]),
    "gasx_fields_sync": (C.c_int, [_P, C.POINTER(Fields)]),
    "gasx_em_step": (C.c_int, [_P, C.POINTER(Species), C.c_int, C.POINTER(Fields), C.c_int]),
    "gasx_cross_sections": (C.c_int, [_P, _P, C.c_int64, _P, _P, _P]),
    "gasx_collide": (C.c_int, [_P, C.POINTER(Species), C.POINTER(Species), C.POINTER(Gas),
                               C.c_double, C.c_uint64, C.c_uint32, C.c_uint32]),
    "gasx_collide_noalloc": (C.c_int, [_P, C.POINTER(Species), C.POINTER(Species), C.POINTER(Gas),
                                       C.c_double, C.c_uint64, C.c_uint32, C.c_uint32]),
    "gasx_run": (C.c_int, [_P, C.POINTER(Species), C.POINTER(Species), C.POINTER(Fields),
                           C.POINTER(Gas), C.POINTER(RunOpts)]),
    "gasx_sort": (C.c_int, [_P, C.POINTER(Species)]),
    "gasx_tallies": (C.c_int, [_P, C.POINTER(Fields), C.POINTER(Species), C.c_int,
                               C.POINTER(C.c_double)]),
    "gasx_get_counters": (C.c_int, [_P, C.POINTER(Counters), C.c_int]),
    "gasx_get_count": (C.c_int, [_P, C.POINTER(Species), C.POINTER(C.c_int64)]),
    "gasx_collision_bound": (C.c_int, [_P, C.POINTER(Gas), C.c_double, C.POINTER(C.c_double)]),
    "gasx_get_timing": (C.c_int, [_P, C.POINTER(Timing)]),
    "gasx_batch_align": (C.c_int, []),
    "gasx_batch_run": (C.c_int, [_P, C.POINTER(Batch), C.POINTER(RunOpts)]),
    "gasx_batch_autotune": (C.c_int, [_P, C.POINTER(Batch), C.POINTER(RunOpts), C.c_int,
                                      C.POINTER(BatchTuning)]),
    "gasx_batch_set_tuning": (C.c_int, [_P, C.POINTER(BatchTuning)]),
    "gasx_batch_get_tuning": (C.c_int, [_P, C.POINTER(BatchTuning)]),
    "gasx_batch_counters": (C.c_int, [_P, C.c_int, C.POINTER(Counters), C.c_int]),
    "gasx_photoionize": (C.c_int, [_P, C.POINTER(Species), C.POINTER(Species), C.POINTER(Gas),
                                   C.POINTER(PhotoSource), C.POINTER(PhotoResult)]),
    "gasx_ion_es_push": (C.c_int, [_P, C.POINTER(Species), C.c_int, C.POINTER(EsParams), _P, _P, _P, _P, _P,
                                   C.POINTER(EsResult)]),
    "gasx_cell_map_info": (C.c_int, [C.POINTER(CellMap), C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "gasx_dsmc_cells": (C.c_int, [_P, C.POINTER(Molecules), C.POINTER(CellMap), _P]),
    "gasx_dsmc_move": (C.c_int, [_P, C.POINTER(Molecules), C.POINTER(CellMap), C.c_double]),
    "gasx_dsmc_collide": (C.c_int, [_P, C.POINTER(Molecules), C.POINTER(CellMap), C.POINTER(DsmcParams),
                                    C.c_uint64, C.c_uint32]),
    "gasx_dsmc_step": (C.c_int, [_P, C.POINTER(Molecules), C.POINTER(CellMap), C.POINTER(DsmcParams),
                                 C.c_uint64, C.c_uint32]),
    "gasx_dsmc_sample": (C.c_int, [_P, C.POINTER(Molecules), C.POINTER(CellMap), C.c_double, C.c_double,
                                   _P, _P, _P, _P]),
    "gasx_dsmc_get_counters": (C.c_int, [_P, C.POINTER(DsmcCounters), C.c_int]),
    "gasx_diff_record": (C.c_int, [_P, C.POINTER(Species), C.POINTER(DiffSpec)]),
    "gasx_diff_forward": (C.c_int, [_P, C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_double)]),
    "gasx_diff_backward": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "gasx_diff_events": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P, _P]),
    "gasx_diff_release": (C.c_int, [_P]),
    "gasx_fdtd_transpose": (C.c_int, [_P, C.POINTER(Fields), _P, _P]),
}

ABI_VERSION = 8

_lib = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libgasx.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(str(p))
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.gasx_abi_version() != ABI_VERSION:
        raise RuntimeError("libgasx ABI version mismatch")
    if path is None:
        _lib = lib
    return lib


def check(lib, ctx, status: int) -> None:
    if status == OK:
        return
    msg = (lib.gasx_last_error(ctx) or b"").decode() if ctx else ""
    msg = msg or lib.gasx_status_string(status).decode()
    if status == E_CAPACITY:
        raise CapacityError(status, msg)
    raise GasxError(status, msg)
