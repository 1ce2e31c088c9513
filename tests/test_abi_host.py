"""Host-side checks of the C ABI and the table loader (no GPU needed).

The library is loaded and its exports are compared with every function that
include/gasx.h declares; argument validation that returns before any HIP call
is exercised; no compute entry point is called.
"""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import gasx
from gasx import _abi
from gasx.tables import raw_tables, read_csv

HEADER = Path(__file__).resolve().parents[1] / "include" / "gasx.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(gasx_[a-z_0-9]+)\s*\(", text, flags=re.M)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("gasx_ctx_create", "gasx_em_step", "gasx_cross_sections", "gasx_collide",
                     "gasx_tallies", "gasx_run", "gasx_fdtd", "gasx_push", "gasx_deposit"):
        assert required in names
    assert len(names) >= 23


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(_abi.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_prototypes_cover_the_header():
    assert set(declared_functions()) == set(_abi.PROTOTYPES)


def test_abi_version_and_status_strings():
    lib = _abi.load()
    v = re.search(r"#define GASX_ABI_VERSION (\d+)", HEADER.read_text()).group(1)
    assert lib.gasx_abi_version() == int(v)
    seen = set()
    for s in range(7):
        msg = lib.gasx_status_string(s)
        assert msg and msg not in seen
        seen.add(msg)


def test_ctx_create_rejects_bad_arguments_without_touching_a_device():
    lib = _abi.load()
    p = gasx.make_params(16, 16)
    handle = gasx.tables.RawTableHandle(raw_tables(), _abi.RawTables)
    ctx = C.c_void_p(1)
    # no CPU backend behind the ABI: a negative device is an argument error
    assert lib.gasx_ctx_create(-1, C.byref(p), C.byref(handle.struct), C.byref(ctx)) == _abi.E_ARG
    assert ctx.value is None
    assert lib.gasx_ctx_create(0, None, C.byref(handle.struct), C.byref(ctx)) == _abi.E_ARG
    assert lib.gasx_ctx_create(0, C.byref(p), None, C.byref(ctx)) == _abi.E_ARG
    assert lib.gasx_ctx_create(0, C.byref(p), C.byref(handle.struct), None) == _abi.E_ARG
    for field, bad in (("nr", 1), ("dr", 0.0), ("phip", 2.0), ("n_sub", 0), ("ddt", -1.0)):
        q = gasx.make_params(16, 16)
        setattr(q, field, bad)
        assert lib.gasx_ctx_create(0, C.byref(q), C.byref(handle.struct), C.byref(ctx)) == _abi.E_ARG, field


def test_null_context_is_an_argument_error():
    lib = _abi.load()
    assert lib.gasx_ctx_destroy(None) in (_abi.OK, _abi.E_ARG)
    assert lib.gasx_synchronize(None) == _abi.E_ARG
    assert lib.gasx_fdtd(None, None, None, None, 1) == _abi.E_ARG
    assert lib.gasx_run(None, None, None, None, None, None) == _abi.E_ARG


def test_raw_tables_shapes():
    t = raw_tables()
    assert t["t_el"].shape == (2, 29)          # TotalElasticCS_Nitrogen.csv: keV row, a0^2 row
    assert t["t_sel"].shape == (26, 2)         # 1..50 eV
    assert t["t_sel1"].shape[1] == 2
    # DCS tables: energy header row + 181 angles; angle column + one column per energy
    assert t["t_del"].shape == (182, 30)
    assert t["t_sdel"].shape == (182, 8)
    assert t["t_ion"].shape == (38, 2)         # 15.6 eV .. 20 keV
    assert t["t_exc"].shape == (7, 5)          # CR line endings handled
    assert t["t_ion"][0, 0] == pytest.approx(15.6, rel=1e-6, abs=0)
    assert t["t_ion"][-1, 0] == pytest.approx(20000.0, rel=1e-6, abs=0)


def test_read_csv_handles_cr_line_endings(tmp_path):
    f = tmp_path / "cr.csv"
    f.write_bytes(b"1,2,3\r4,5,6\r")
    a = read_csv(f)
    assert a.shape == (2, 3) and a[1, 2] == 6


def test_seeded_electrons_follow_the_energy_mixture():
    from gasx.workloads import seed_electrons
    n = 200_000
    x, y, z, vx, vy, vz = seed_electrons(n, 64 * 1e-5, 128 * 1e-5, 0.1, seed=3)
    r = np.hypot(x, y)
    assert np.all(r < 64e-5) and np.all(np.arctan2(y, x) >= 0) and np.all(np.arctan2(y, x) <= 0.1 + 1e-6)
    v2 = vx.astype(np.float64) ** 2 + vy ** 2 + vz ** 2
    c = 3e8
    g = 1 / np.sqrt(1 - v2 / c ** 2)
    e_ev = (g - 1) * 9.1096e-31 * c ** 2 / 1.6e-19
    hi = e_ev > 200
    assert abs(hi.mean() - 0.7) < 0.01
    assert abs(np.mean(e_ev[hi]) - 400) < 1.0
    assert abs(np.mean(e_ev[~hi]) - 30) < 1.0


def test_ring_density_collapses_phi_cells():
    """cell_structure.m:9-25 layout: ring k holds cells sum_cell0[k]..sum_cell(k)-1 of each z plane."""
    from gasx.reference_api import ring_density
    phi_div = np.array([1, 2, 3])
    sum_cell = np.cumsum(phi_div)
    vals = np.arange(2 * 6, dtype=float)       # two z planes of 6 cells
    out = ring_density(vals, sum_cell, 6, 2)
    assert out.shape == (2, 3)
    assert np.allclose(out, [[0, 1.5, 4], [6, 7.5, 10]])
    with pytest.raises(ValueError):
        ring_density(vals, sum_cell, 7, 2)


DRIVER = Path(__file__).resolve().parents[1] / "gas-atenuator_amd" / "shim" / "transport_driver"


def test_transport_driver_argument_and_table_errors(tmp_path):
    """The C driver's failures that come before any HIP call: usage, a
    missing data directory, a table of the wrong shape."""
    import shutil
    import subprocess
    from gasx.tables import DATA_DIR
    assert DRIVER.exists(), "build() makes the driver"
    r = subprocess.run([str(DRIVER)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    args = ["64", "128", "3", "1e6", "2", "2", "1"]
    r = subprocess.run([str(DRIVER), str(tmp_path / "none"), str(tmp_path / "o.bin")] + args,
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "create: invalid argument" in r.stderr
    bad = tmp_path / "n2"
    shutil.copytree(DATA_DIR, bad)
    lines = (bad / "ElasticCS_Nitrogen.csv").read_text().replace("\r\n", "\n").replace("\r", "\n").splitlines()
    (bad / "ElasticCS_Nitrogen.csv").write_text("\n".join(lines[:-1]) + "\n")   # 181 rows, not 182
    r = subprocess.run([str(DRIVER), str(bad), str(tmp_path / "o.bin")] + args, capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "create: invalid argument" in r.stderr
    assert not (tmp_path / "o.bin").exists()


STRUCTS = {   # ctypes mirror -> C typedef of include/gasx.h
    "Params": "gasx_params", "RawTables": "gasx_raw_tables", "Species": "gasx_species", "Fields": "gasx_fields",
    "Gas": "gasx_gas", "Counters": "gasx_counters", "Timing": "gasx_timing", "Batch": "gasx_batch",
    "BatchTuning": "gasx_batch_tuning", "DiffSpec": "gasx_diff_spec", "RunOpts": "gasx_run_opts",
    "PhotoSource": "gasx_photo_source", "PhotoResult": "gasx_photo_result", "EsParams": "gasx_es_params",
    "EsResult": "gasx_es_result", "CellMap": "gasx_cell_map", "Molecules": "gasx_molecules",
    "DsmcParams": "gasx_dsmc_params", "DsmcCounters": "gasx_dsmc_counters",
}


def test_ctypes_structs_match_the_header_layout(tmp_path):
    """Every ctypes structure has the size, field names and field offsets of
    its C struct (compiled against include/gasx.h with the host compiler)."""
    import shutil
    import subprocess
    from gasx import _abi
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "gasx.h"', "int main(void) {"]
    for py, c in STRUCTS.items():
        cls = getattr(_abi, py)
        lines.append(f'printf("{py} size %zu\\n", sizeof({c}));')
        for name, _ in cls._fields_:
            lines.append(f'printf("{py} {name} %zu\\n", offsetof({c}, {name}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", f"-I{HEADER.parent}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            py, key, val = line.split()
            got[(py, key)] = int(val)
    for py in STRUCTS:
        cls = getattr(_abi, py)
        assert got[(py, "size")] == C.sizeof(cls), py
        for name, _ in cls._fields_:
            assert got[(py, name)] == getattr(cls, name).offset, (py, name)
