"""ctypes binding of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker of the HIP product path.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
NANG = 181
MAXT = 64

_P = C.c_void_p
_D = C.POINTER(C.c_double)


class OrcParams(C.Structure):
    _fields_ = [
        ("nr", C.c_int32), ("nz", C.c_int32), ("dr", C.c_double), ("dz", C.c_double),
        ("phip", C.c_double), ("ddt", C.c_double), ("n_sub", C.c_int32), ("fn", C.c_double),
        ("c", C.c_double), ("eps0", C.c_double), ("mu0", C.c_double),
        ("e_em", C.c_double), ("me_em", C.c_double), ("e_mc", C.c_double), ("me_mc", C.c_double),
        ("m_ion", C.c_double), ("k_b", C.c_double),
        ("ez_metric", C.c_int32), ("axis_j_sign", C.c_int32), ("exc_form", C.c_int32),
        ("rot_eps", C.c_double),
    ]


class OrcTables(C.Structure):
    _fields_ = [
        ("n_el", C.c_int32), ("el_e", C.c_double * MAXT), ("el_s", C.c_double * MAXT),
        ("n_sel", C.c_int32), ("sel_e", C.c_double * MAXT), ("sel_s", C.c_double * MAXT),
        ("n_ion", C.c_int32), ("ion_e", C.c_double * MAXT), ("ion_s", C.c_double * MAXT),
        ("n_exc", C.c_int32), ("exc_w", C.c_double * 16), ("exc_a", C.c_double * 16),
        ("exc_om", C.c_double * 16), ("exc_nu", C.c_double * 16), ("exc_ga", C.c_double * 16),
        ("n_dcs", C.c_int32), ("dcs_e", C.c_double * MAXT), ("dcs_loge", C.c_double * MAXT),
        ("dcs_tcs", C.c_double * MAXT), ("dcs_logv", C.c_double * (MAXT * NANG)),
    ]


class OrcGas(C.Structure):
    _fields_ = [("dens", _D), ("temp", _D), ("gnr", C.c_int32), ("gnz", C.c_int32),
                ("gdr", C.c_double), ("gdz", C.c_double), ("t_scalar", C.c_double)]


class OrcCounts(C.Structure):
    _fields_ = [("n_elastic", C.c_int64), ("n_ionize", C.c_int64), ("n_excite", C.c_int64),
                ("n_dropped", C.c_int64)]


class OrcPhotoSource(C.Structure):
    _fields_ = [("photon_kev", C.c_double), ("n_photons", C.c_double), ("beam_radius", C.c_double),
                ("binding_ev", _D), ("n_shells", C.c_int32), ("mac", _D), ("n_mac", C.c_int32),
                ("atomic_mass", C.c_double), ("auger_kev", C.c_double), ("seed", C.c_uint64),
                ("pulse", C.c_uint32), ("run", C.c_uint32)]


class OrcPhotoInfo(C.Structure):
    _fields_ = [("photons_left", C.c_double), ("sigma_pe", C.c_double), ("electron_kev", C.c_double),
                ("auger", C.c_int32), ("n_rings", C.c_int32), ("n_events", C.c_int64)]


class OrcEsParams(C.Structure):
    _fields_ = [("nr", C.c_int32), ("nz", C.c_int32), ("delta", C.c_double), ("phip", C.c_double),
                ("weight", C.c_double), ("t_e", C.c_double), ("dt", C.c_double), ("n_e0", C.c_double),
                ("tol", C.c_double), ("max_iter", C.c_int32), ("boltzmann", C.c_int32)]


class OrcEsInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("n_e0", C.c_double),
                ("max_dphi", C.c_double), ("mean_phi", C.c_double), ("int_rho_i", C.c_double)]


class OrcCellMap(C.Structure):
    _fields_ = [("n_rings", C.c_int32), ("n_layers", C.c_int32), ("rp", C.c_double), ("lp", C.c_double),
                ("phip", C.c_double), ("delta", C.c_double)]


class OrcDsmcParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("fn", C.c_double), ("mass", C.c_double), ("diameter", C.c_double),
                ("sc_max", C.c_double), ("omega", C.c_double), ("zeta", C.c_double), ("alpha", C.c_double),
                ("min_attempts", C.c_int32)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(str(LIB_PATH))
        L.orc_interp1.restype = C.c_double
        L.orc_interp1.argtypes = [_D, _D, C.c_int, C.c_double]
        L.orc_elastic_theta.restype = C.c_double
        L.orc_elastic_theta.argtypes = [C.POINTER(OrcTables), C.c_double, C.c_double]
        L.orc_exc_edep.restype = C.c_double
        L.orc_exc_edep.argtypes = [C.POINTER(OrcTables), C.c_double, C.c_int]
        L.orc_u01.restype = C.c_double
        L.orc_u01.argtypes = [C.c_uint32]
        L.orc_ionize.restype = None
        L.orc_ionize.argtypes = [C.c_double, C.c_double, C.c_double] + [_D] * 5
        L.orc_rotate.restype = None
        L.orc_rotate.argtypes = [_D, _D, _D, C.c_double, C.c_double, C.c_double]
        L.orc_collide.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcTables), C.POINTER(OrcGas),
                                  _P, _P, _P, _P, _P, _P, _P, C.c_int64, C.c_double,
                                  C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, _P, _P, _P]
        L.orc_run.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcTables), C.POINTER(OrcGas),
                              _P, _P, _P, _P, _P, _P, _P, C.POINTER(C.c_int64), C.c_int64,
                              _P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint32,
                              C.c_uint32, C.c_int, C.POINTER(OrcCounts)]
        L.orc_elastic_theta_rej.restype = C.c_double
        L.orc_elastic_theta_rej.argtypes = [C.POINTER(OrcTables), C.c_double, C.c_uint32, C.c_uint32,
                                            C.c_uint32, C.c_uint64]
        L.orc_fdtd_push.argtypes = [C.POINTER(OrcParams), _P, _P, _P, _P, _P, C.c_int]
        L.orc_fdtd_h.argtypes = [_P, _P, _P, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_fdtd_e.argtypes = [_P, _P, _P, _P, _P, C.c_int, C.c_int, C.c_double, C.c_double,
                                 C.c_double, C.c_double, C.c_int, C.c_int]
        L.orc_em2.argtypes = [C.POINTER(OrcParams)] + [_P] * 10 + [C.c_int64]
        L.orc_em4.argtypes = [C.POINTER(OrcParams)] + [_P] * 16 + [C.c_int64]
        L.orc_em7.argtypes = [C.POINTER(OrcParams)] + [_P] * 8 + [C.c_int64]
        L.orc_em_step.argtypes = [C.POINTER(OrcParams)] + [_P] * 10 + [C.c_int64] + [_P] * 5 + [C.c_int]
        L.orc_deposit_fx32.argtypes = [C.POINTER(OrcParams)] + [_P] * 6 + [C.c_int64, C.c_float, _P, _P]
        L.orc_j_from_fixed.argtypes = [C.POINTER(OrcParams), _P, _P, _P, _P]
        L.orc_tallies.argtypes = [C.POINTER(OrcParams)] + [_P] * 8 + [C.c_int64, _P]
        L.orc_cross_sections.argtypes = [C.POINTER(OrcTables), _P, C.c_int64, _P, _P, _P]
        L.orc_tables_build.argtypes = [C.POINTER(OrcTables), _P, C.c_int, _P, C.c_int, _P, C.c_int,
                                       _P, C.c_int, _P, C.c_int, _P, C.c_int, _P, C.c_int]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_photoionize.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcGas), _P, _P, C.c_int,
                                      C.POINTER(OrcPhotoSource), C.c_int64, _P, _P, _P, _P, _P,
                                      C.POINTER(OrcPhotoInfo)]
        L.orc_sauter_v.restype = C.c_double
        L.orc_sauter_v.argtypes = [C.c_double, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_es_poisson.argtypes = [C.POINTER(OrcEsParams), _P, _P, _P, C.POINTER(OrcEsInfo)]
        L.orc_es_field.argtypes = [C.POINTER(OrcEsParams), _P, _P, _P]
        L.orc_es_push.argtypes = [C.POINTER(OrcEsParams), C.c_int] + [_P] * 6 + [_P, _P, _P] + [_P] * 5 + \
            [C.POINTER(OrcEsInfo)]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_cell_map_build.restype = C.c_int64
        L.orc_cell_map_build.argtypes = [C.POINTER(OrcCellMap), _P, _P, _P, _P]
        L.orc_dsmc_cells.argtypes = [C.POINTER(OrcCellMap), C.c_int64, _P, _P, _P, _P]
        L.orc_dsmc_move.argtypes = [C.POINTER(OrcCellMap), C.c_double, C.c_int64] + [_P] * 6 + [C.c_int, _P]
        L.orc_dsmc_u.restype = C.c_double
        L.orc_dsmc_u.argtypes = [C.c_uint32]
        L.orc_dsmc_collide.argtypes = [C.POINTER(OrcCellMap), C.POINTER(OrcDsmcParams), C.c_int64, _P] + \
            [_P] * 4 + [C.c_uint64, C.c_uint32, C.c_int, _P]
        L.orc_dsmc_sample.argtypes = [C.POINTER(OrcCellMap), C.c_int64, _P, _P, _P, _P, C.c_double, C.c_double,
                                      C.c_double, _P, _P, _P, _P]
        L.orc_set_threads.restype = None
        L.orc_get_threads.restype = C.c_int
        _lib = L
    return _lib


def a(x, dtype=np.float64):
    """contiguous array + its pointer"""
    return np.ascontiguousarray(x, dtype=dtype)


def ptr(x: np.ndarray):
    return x.ctypes.data_as(C.c_void_p)


def params_from(p) -> OrcParams:
    """Copy a gasx Params (or any object with the same fields) into OrcParams."""
    return OrcParams(*[getattr(p, n) for n, _ in OrcParams._fields_])


def build_tables(raw: dict) -> OrcTables:
    t = OrcTables()
    r = {k: a(v) for k, v in raw.items()}
    st = lib().orc_tables_build(C.byref(t), ptr(r["t_el"]), r["t_el"].shape[1], ptr(r["t_sel"]),
                                r["t_sel"].shape[0], ptr(r["t_sel1"]), r["t_sel1"].shape[0],
                                ptr(r["t_del"]), r["t_del"].shape[1], ptr(r["t_sdel"]),
                                r["t_sdel"].shape[1], ptr(r["t_ion"]), r["t_ion"].shape[0],
                                ptr(r["t_exc"]), r["t_exc"].shape[0])
    if st != 0:
        raise ValueError(f"orc_tables_build failed ({st})")
    return t


def set_threads(n: int) -> None:
    """Threads of the oracle's loops (1 by default: one fixed FP64 summation order)."""
    lib().orc_set_threads(int(n))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def cross_sections(t: OrcTables, e_kev):
    e = a(np.atleast_1d(e_kev))
    se, si, sx = (np.empty_like(e) for _ in range(3))
    lib().orc_cross_sections(C.byref(t), ptr(e), e.size, ptr(se), ptr(si), ptr(sx))
    return se, si, sx


def fdtd_push(p: OrcParams, er, ez, h, jr, jz, n_push=1):
    er, ez, h = a(er).copy(), a(ez).copy(), a(h).copy()
    jr, jz = a(jr), a(jz)
    lib().orc_fdtd_push(C.byref(p), ptr(er), ptr(ez), ptr(h), ptr(jr), ptr(jz), n_push)
    return er, ez, h


def deposit_fx32(p: OrcParams, state32: np.ndarray, qnum=-1.0):
    s = [a(state32[i], np.float32) for i in range(6)]
    n = s[0].size
    qr = np.zeros(p.nr * p.nz, dtype=np.int64)
    qz = np.zeros_like(qr)
    lib().orc_deposit_fx32(C.byref(p), *[ptr(v) for v in s], n, qnum, ptr(qr), ptr(qz))
    return qr, qz


def j_from_fixed(p: OrcParams, qr, qz):
    qr, qz = a(qr, np.int64), a(qz, np.int64)
    jr = np.empty(qr.size)
    jz = np.empty(qr.size)
    lib().orc_j_from_fixed(C.byref(p), ptr(qr), ptr(qz), ptr(jr), ptr(jz))
    return jr, jz


class State:
    """FP64 particle state in the oracle's (reference) representation."""

    def __init__(self, arr6, cap=None, q=None, m=None):
        arr6 = np.asarray(arr6, dtype=np.float64)
        n = arr6.shape[1]
        cap = cap or n
        self.v = np.zeros((6, cap))
        self.v[:, :n] = arr6
        self.n = n
        self.dele = np.zeros(cap, dtype=np.int32)
        self.dele[:n][np.isnan(arr6[2])] = 1
        self.v[2, :n][np.isnan(arr6[2])] = 0.0
        self.q = np.full(cap, -1.6e-19) if q is None else a(q)
        self.m = np.full(cap, 9.1096e-31) if m is None else a(m)

    def rows(self):
        return [self.v[i] for i in range(6)]


def em2(p, st: State):
    r, phi, vr = (np.zeros(st.v.shape[1]) for _ in range(3))
    lib().orc_em2(C.byref(p), *[ptr(x) for x in st.rows()], ptr(r), ptr(phi), ptr(vr), ptr(st.dele), st.n)
    return r, phi, vr


def em4(p, st: State, r, phi, vr, er, ez, h):
    ke = np.zeros(st.v.shape[1])
    er, ez, h = a(er), a(ez), a(h)
    lib().orc_em4(C.byref(p), *[ptr(x) for x in st.rows()], ptr(r), ptr(phi), ptr(vr), ptr(er),
                  ptr(ez), ptr(h), ptr(st.q), ptr(st.m), ptr(ke), ptr(st.dele), st.n)
    return ke


def em7(p, st: State, r, vr):
    jr = np.zeros(p.nr * p.nz)
    jz = np.zeros(p.nr * p.nz)
    lib().orc_em7(C.byref(p), ptr(r), ptr(st.v[2]), ptr(st.q), ptr(vr), ptr(st.v[5]), ptr(jr),
                  ptr(jz), ptr(st.dele), st.n)
    return jr, jz


def em_step(p, st: State, er, ez, h, n_push):
    er, ez, h = a(er).copy(), a(ez).copy(), a(h).copy()
    jr = np.zeros(p.nr * p.nz)
    jz = np.zeros(p.nr * p.nz)
    ke = np.zeros(st.v.shape[1])
    lib().orc_em_step(C.byref(p), *[ptr(x) for x in st.rows()], ptr(st.q), ptr(st.m), ptr(ke),
                      ptr(st.dele), st.n, ptr(er), ptr(ez), ptr(h), ptr(jr), ptr(jz), n_push)
    return er, ez, h, ke


def make_gas(dens, gdr, gdz, t_scalar=300.0):
    d = a(np.atleast_2d(dens))
    g = OrcGas(d.ctypes.data_as(_D), None, d.shape[1], d.shape[0], gdr, gdz, t_scalar)
    g._keep = d
    return g


ELASTIC_INVERSE_CDF, ELASTIC_REJECTION = 0, 1


def collide(p, t, g, st: State, dt_coll, seed, step, run=0, ions=False, elastic_mode=ELASTIC_REJECTION):
    ev = np.zeros(st.v.shape[1], dtype=np.int8)
    sv = np.zeros(3 * st.v.shape[1])
    iv = np.zeros(3 * st.v.shape[1]) if ions else None
    lib().orc_collide(C.byref(p), C.byref(t), C.byref(g), *[ptr(x) for x in st.rows()],
                      ptr(st.dele), st.n, dt_coll, seed, step, run, elastic_mode, ptr(ev), ptr(sv),
                      ptr(iv) if ions else None)
    return ev[: st.n], sv.reshape(-1, 3)[: st.n], (iv.reshape(-1, 3)[: st.n] if ions else None)


def run(p, t, g, st: State, er, ez, h, n_steps, collide=True, secondaries=False, seed=0, run_id=0,
        step0=0, elastic_mode=ELASTIC_REJECTION):
    er, ez, h = a(er).copy(), a(ez).copy(), a(h).copy()
    n = C.c_int64(st.n)
    cnt = OrcCounts()
    rc = lib().orc_run(C.byref(p), C.byref(t), C.byref(g), *[ptr(x) for x in st.rows()],
                       ptr(st.dele), C.byref(n), st.v.shape[1], ptr(er), ptr(ez), ptr(h), n_steps,
                       int(collide), int(secondaries), seed, run_id, step0, elastic_mode,
                       C.byref(cnt))
    if rc != 0:
        raise RuntimeError("orc_run failed")
    st.n = n.value
    return er, ez, h, cnt


def tallies(p, er, ez, h, st: State):
    out = np.zeros(4)
    lib().orc_tallies(C.byref(p), ptr(a(er)), ptr(a(ez)), ptr(a(h)), ptr(st.v[3]), ptr(st.v[4]),
                      ptr(st.v[5]), ptr(st.m), ptr(st.dele), st.n, ptr(out))
    return out


def photoionize(p, dens, gdr, gdz, photon_kev, n_photons, beam_radius, tables, temp=None, t_scalar=300.0,
                deplete=False, seed=0, pulse=0, run=0, auger_kev=0.383, atomic_mass=14.0, max_events=None):
    """INCIDENCE1.m restated (orc_photoionize).  dens / temp: (gnz, gnr) FP64
    grids (copied; the depleted grids are returned).  Returns a dict with the
    per-event arrays pe, ae, ion (6 x n), vth (3 x n), cell (n) and the info."""
    d = np.array(np.atleast_2d(dens), dtype=np.float64)
    t = np.array(np.atleast_2d(temp), dtype=np.float64) if temp is not None else np.full_like(d, t_scalar)
    g = OrcGas(d.ctypes.data_as(_D), t.ctypes.data_as(_D) if (temp is not None or deplete) else None,
               d.shape[1], d.shape[0], gdr, gdz, t_scalar)
    be = a(tables["binding_ev"]).ravel()
    mac = a(tables["mac"])
    s = OrcPhotoSource(photon_kev, n_photons, beam_radius, be.ctypes.data_as(_D), be.size,
                       mac.ctypes.data_as(_D), mac.shape[0], atomic_mass, auger_kev, seed, pulse, run)
    m = int(max_events or 1 << 20)
    pe, ae, ion = (np.zeros((6, m)) for _ in range(3))
    vth = np.zeros((3, m))
    cell = np.zeros(m, dtype=np.int32)
    info = OrcPhotoInfo()
    rc = lib().orc_photoionize(C.byref(p), C.byref(g), ptr(d), ptr(t), int(deplete), C.byref(s), m, ptr(pe),
                               ptr(ae), ptr(ion), ptr(vth), ptr(cell), C.byref(info))
    if rc != 0:
        raise ValueError(f"orc_photoionize failed ({rc})")
    n = info.n_events
    return {"pe": pe[:, :n], "ae": ae[:, :n] if info.auger else None, "ion": ion[:, :n], "vth": vth[:, :n],
            "cell": cell[:n], "dens": d, "temp": t,
            "info": {k: getattr(info, k) for k, _ in OrcPhotoInfo._fields_}}


def es_params(nr, nz, delta, phip, weight, t_e, dt, n_e0=1e15, tol=0.01, max_iter=1_000_000, boltzmann=True):
    return OrcEsParams(nr, nz, delta, phip, weight, t_e, dt, n_e0, tol, max_iter, int(boltzmann))


def es_poisson(p: OrcEsParams, rho_i):
    """Poisson.m's iteration restated (orc_es_poisson); rho_i (nz, nr)."""
    rho_i = np.ascontiguousarray(rho_i, dtype=np.float64)
    phi = np.zeros_like(rho_i)
    rho_e = np.zeros_like(rho_i)
    info = OrcEsInfo()
    lib().orc_es_poisson(C.byref(p), ptr(rho_i), ptr(phi), ptr(rho_e), C.byref(info))
    return phi, rho_e, {k: getattr(info, k) for k, _ in OrcEsInfo._fields_}


def es_field(p: OrcEsParams, phi):
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    er, ez = np.zeros_like(phi), np.zeros_like(phi)
    lib().orc_es_field(C.byref(p), ptr(phi), ptr(er), ptr(ez))
    return er, ez


def es_push(p: OrcEsParams, species):
    """ion_staticelectrofield_push.m restated (orc_es_push).  species: list
    of (state (6, n) FP64, charge number, mass); states are pushed in place.
    Returns the grids (nz, nr) and the info."""
    nsp = len(species)
    rows = [np.ascontiguousarray(st, dtype=np.float64) for st, _, _ in species]
    for st, r in zip([s for s, _, _ in species], rows):
        if r is not st:
            raise ValueError("species states must be C-contiguous FP64 (6, n) arrays")
    PA = C.c_void_p * nsp
    cols = [PA(*[r[k].ctypes.data for r in rows]) for k in range(6)]
    n = (C.c_int64 * nsp)(*[r.shape[1] for r in rows])
    q = (C.c_double * nsp)(*[qq for _, qq, _ in species])
    m = (C.c_double * nsp)(*[mm for _, _, mm in species])
    shape = (p.nz, p.nr)
    phi, rho_i, rho_e, er, ez = (np.zeros(shape) for _ in range(5))
    info = OrcEsInfo()
    lib().orc_es_push(C.byref(p), nsp, *cols, n, q, m, ptr(phi), ptr(rho_i), ptr(rho_e), ptr(er), ptr(ez),
                      C.byref(info))
    return {"phi": phi, "rho_i": rho_i, "rho_e": rho_e, "e_r": er, "e_z": ez,
            "info": {k: getattr(info, k) for k, _ in OrcEsInfo._fields_}}


# ---- DSMC neutral gas (main2.m:247-264, collision_HS_internal.m) ----------------
def cell_map(n_rings, n_layers, rp, lp, phip, delta) -> OrcCellMap:
    return OrcCellMap(n_rings, n_layers, rp, lp, phip, delta)


def cell_map_build(m: OrcCellMap):
    """cell_structure.m + calculate_VC.m: (phi_divide, sum0, dphi, vc_ring, cellnumxy)."""
    nxy = lib().orc_cell_map_build(C.byref(m), None, None, None, None)
    pd = np.zeros(m.n_rings, dtype=np.int32)
    s0 = np.zeros(m.n_rings + 1, dtype=np.int32)
    dp = np.zeros(m.n_rings)
    vc = np.zeros(m.n_rings)
    lib().orc_cell_map_build(C.byref(m), ptr(pd), ptr(s0), ptr(dp), ptr(vc))
    return pd, s0, dp, vc, nxy


def dsmc_params(dt, fn, mass, diameter, sc_max, omega=0.74, zeta=2.0, alpha=1.0, min_attempts=1) -> OrcDsmcParams:
    return OrcDsmcParams(dt, fn, mass, diameter, sc_max, omega, zeta, alpha, min_attempts)


def _mol(st):
    if st.dtype != np.float64 or not st.flags.c_contiguous or st.shape[0] != 7:
        raise ValueError("molecule state must be a C-contiguous FP64 (7, n) array")
    return [st[k].ctypes.data for k in range(7)]


def dsmc_cells(m: OrcCellMap, st: np.ndarray) -> np.ndarray:
    """put_into_cells_symmetry.m: 0-based cell of each molecule (-1: z NaN)."""
    n = st.shape[1]
    cell = np.zeros(n, dtype=np.int32)
    x, y, z = (np.ascontiguousarray(st[k], dtype=np.float64) for k in range(3))
    lib().orc_dsmc_cells(C.byref(m), n, ptr(x), ptr(y), ptr(z), ptr(cell))
    return cell


def dsmc_move(m: OrcCellMap, dt: float, st: np.ndarray, round32: bool = True) -> np.ndarray:
    """free_motion.m + boundary_conditions_specular.m in place; returns
    [boundary events, r wall, z walls, phi wraps]."""
    c = np.zeros(4, dtype=np.int64)
    a = _mol(st)
    lib().orc_dsmc_move(C.byref(m), dt, st.shape[1], *a[:6], int(round32), ptr(c))
    return c


def dsmc_collide(m: OrcCellMap, p: OrcDsmcParams, st: np.ndarray, seed: int, step: int, cell=None,
                 round32: bool = True) -> np.ndarray:
    """collision_HS_internal.m in place; returns [attempts, collisions, n_SC]."""
    if cell is None:
        cell = dsmc_cells(m, st)
    cell = np.ascontiguousarray(cell, dtype=np.int32)
    c = np.zeros(3, dtype=np.int64)
    a = _mol(st)
    lib().orc_dsmc_collide(C.byref(m), C.byref(p), st.shape[1], ptr(cell), a[3], a[4], a[5], a[6], seed, step,
                           int(round32), ptr(c))
    return c


def dsmc_reorder(m: OrcCellMap, st: np.ndarray) -> np.ndarray:
    """Stable sort of the molecules by cell, unused (z NaN) slots dropped."""
    cell = dsmc_cells(m, st)
    keep = cell >= 0
    order = np.argsort(np.where(keep, cell, np.iinfo(np.int32).max), kind="stable")[: int(keep.sum())]
    return np.ascontiguousarray(st[:, order])


def dsmc_sample(m: OrcCellMap, st: np.ndarray, fn: float, mass: float, kb: float = 1.38064853e-23):
    """particle_density0.m + count_T_numbensity.m per cell and per (ring, layer)."""
    pd, s0, dp, vc, nxy = cell_map_build(m)
    cell = dsmc_cells(m, st)
    n = st.shape[1]
    cd, ct = np.zeros(nxy * m.n_layers), np.zeros(nxy * m.n_layers)
    rd, rt = np.zeros((m.n_layers, m.n_rings)), np.zeros((m.n_layers, m.n_rings))
    vx, vy, vz = (np.ascontiguousarray(st[k], dtype=np.float64) for k in range(3, 6))
    lib().orc_dsmc_sample(C.byref(m), n, ptr(cell), ptr(vx), ptr(vy), ptr(vz), fn, mass, kb, ptr(cd), ptr(ct),
                          ptr(rd), ptr(rt))
    return {"cell_density": cd, "cell_temperature": ct, "rz_density": rd, "rz_temperature": rt}
