"""Build step for libgasx.so (HIP, gfx950) and the CPU oracle.

``build_library()`` always recompiles: each ``csrc/*.hip`` is compiled to an
object with hipcc for gfx950 and linked into ``gasx/libgasx.so``.  All HIP
code is built with packed FP32 instructions disabled; FP contraction is on
except where the source turns it off (the fixed-point deposit weights, which
must be reproducible on the CPU, and the sort keys).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
ROOT = PKG_DIR.parent                     # gas-atenuator_amd/
REPO = ROOT.parent
CSRC = ROOT / "csrc"
INCLUDE = REPO / "include"
LIB_PATH = PKG_DIR / "libgasx.so"
OBJ_DIR = ROOT / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GASX_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    "-ffp-contract=fast-honor-pragmas",   # the deposit arithmetic turns contraction off locally
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wno-unused-value",   # hipFree/hipStreamDestroy results on cleanup paths are ignored on purpose
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}")
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if out:
        sys.stderr.write(out + "\n")


def build_library(jobs: int = 4) -> Path:
    """Compile every csrc/*.hip for gfx950 and link libgasx.so."""
    OBJ_DIR.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    objs = [OBJ_DIR / (s.stem + ".o") for s in sources]

    def compile_one(pair):
        src, obj = pair
        _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])

    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(compile_one, zip(sources, objs)))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


SHIM = ROOT / "shim"
DRIVER_PATH = SHIM / "transport_driver"
DSMC_DRIVER_PATH = SHIM / "dsmc_driver"
ES_DRIVER_PATH = SHIM / "es_driver"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))


def build_shim() -> Path:
    """The C drivers of the ABI (shim/transport_driver.c, shim/dsmc_driver.c and
    shim/es_driver.c over shim/gasx_shim.c), plain C against libgasx.so and the
    HIP runtime.
    The MEX gateway (shim/gasx_mex.c) needs MATLAB's mex and is not built here."""
    for main, out in (("transport_driver.c", DRIVER_PATH), ("dsmc_driver.c", DSMC_DRIVER_PATH),
                      ("es_driver.c", ES_DRIVER_PATH)):
        _run(["gcc", "-O2", "-std=c11", "-Wall", "-Wextra", f"-I{INCLUDE}", f"-I{ROCM / 'include'}",
              str(SHIM / "gasx_shim.c"), str(SHIM / main),
              f"-L{PKG_DIR}", "-lgasx", f"-L{ROCM / 'lib'}", "-lamdhip64", "-lm",
              "-Wl,-rpath,$ORIGIN/../gasx", f"-Wl,-rpath,{ROCM / 'lib'}", "-o", str(out)])
    return DRIVER_PATH


def build_oracle() -> Path:
    """Compile the FP64 CPU oracle (test infrastructure) with its Makefile."""
    oracle_dir = REPO / "oracle"
    _run(["make", "-s", "-B", "-C", str(oracle_dir)])
    return oracle_dir / "liboracle.so"


def clean() -> None:
    shutil.rmtree(OBJ_DIR, ignore_errors=True)


if __name__ == "__main__":
    print(build_library())
    print(build_shim())
    print(build_oracle())
