"""N2 cross-section tables, read as the reference reads them.

The CSV files under ``data/n2`` are the reference's own data tables
(electron_transport_field3.m:8-37 reads them with readtable).  The files
are used unchanged; ParameterExcCS_Nitrogen.csv has a header row and CR line
endings, which the reader handles.  ``raw_tables()`` returns the matrices in
the row-major layout of ``gasx_raw_tables``; the merges of
electron_collision.m:12-41 happen inside the library (and, separately, in
the oracle).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

DATA_DIR = Path(__file__).resolve().parent.parent / "data" / "n2"

FILES = {
    "t_el": "TotalElasticCS_Nitrogen.csv",          # 2 x 29  (keV ; a0^2 per atom)
    "t_sel": "TotalElasticCS_smalleV_Nitrogen.csv", # 26 x 2  (eV ; 1e-20 m^2)
    "t_sel1": "TotalElasticCS_smalleV_Nitrogen_1.csv",  # 8 x 2
    "t_del": "ElasticCS_Nitrogen.csv",              # 182 x 30 DCS (row 0 energies keV, col 0 angle)
    "t_sdel": "ElasticDCS_smalleV_Nitrogen.csv",    # 182 x 8
    "t_ion": "TotalIonCS_Nitrogen.csv",             # 38 x 2  (eV ; 1e-20 m^2)
    "t_exc": "ParameterExcCS_Nitrogen.csv",         # 7 x 5   (j, A, omega, v, miu)
}


def read_csv(path: Path) -> np.ndarray:
    """Numeric CSV with optional header row and any line-ending convention."""
    text = Path(path).read_bytes().decode("ascii").replace("\r\n", "\n").replace("\r", "\n")
    rows = []
    for line in text.split("\n"):
        line = line.strip()
        if not line:
            continue
        cells = line.split(",")
        try:
            rows.append([float(c) for c in cells])
        except ValueError:
            if rows:
                raise
            continue  # header row (readtable takes it as variable names)
    return np.asarray(rows, dtype=np.float64)


def raw_tables(data_dir: Path | None = None) -> dict[str, np.ndarray]:
    d = Path(data_dir) if data_dir else DATA_DIR
    out = {k: np.ascontiguousarray(read_csv(d / f)) for k, f in FILES.items()}
    assert out["t_el"].shape[0] == 2
    assert out["t_del"].shape[0] == 182 and out["t_sdel"].shape[0] == 182
    assert out["t_exc"].shape[1] == 5
    return out


SOURCE_FILES = {
    "binding_ev": "EB_14.csv",   # N shell binding energies (eV), K then valence (findEShell.m:2)
    "mac": "14_MAC.csv",         # photon energy (MeV) ; mass attenuation (cm^2/g) (findSigmaPE.m:9)
}


def source_tables(data_dir: Path | None = None) -> dict[str, np.ndarray]:
    """The photoionization source's data files of the reference, unchanged."""
    d = Path(data_dir) if data_dir else DATA_DIR
    out = {"binding_ev": np.ascontiguousarray(read_csv(d / SOURCE_FILES["binding_ev"]).ravel()),
           "mac": np.ascontiguousarray(read_csv(d / SOURCE_FILES["mac"]))}
    assert out["mac"].shape[1] == 2
    return out


class RawTableHandle:
    """Keeps the numpy buffers alive while the C struct points at them."""

    def __init__(self, tabs: dict[str, np.ndarray], struct_type):
        self.tabs = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in tabs.items()}
        t = self.tabs
        P = C.POINTER(C.c_double)
        ptr = lambda a: a.ctypes.data_as(P)
        self.struct = struct_type(
            ptr(t["t_el"]), t["t_el"].shape[1],
            ptr(t["t_sel"]), t["t_sel"].shape[0],
            ptr(t["t_sel1"]), t["t_sel1"].shape[0],
            ptr(t["t_del"]), t["t_del"].shape[1],
            ptr(t["t_sdel"]), t["t_sdel"].shape[1],
            ptr(t["t_ion"]), t["t_ion"].shape[0],
            ptr(t["t_exc"]), t["t_exc"].shape[0],
        )
