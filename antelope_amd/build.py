"""Build the HIP library (gfx950) and the C++ driver in-tree with hipcc.

    python -m antelope_amd.build            # library + driver
    python -m antelope_amd.build --lib-only

Outputs land in antelope_amd/lib/ (git-ignored, but shipped with the tree).
Every source is compiled on every call (the driver's "does it build" check).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
OBJDIR = LIBDIR / "obj"
LIBNAME = "libantelope_sbe.so"
ARCH = os.environ.get("ANTELOPE_OFFLOAD_ARCH", "gfx950")

# Library sources (kernels + C ABI). The driver links against the library.
LIB_SOURCES = [
    "host_setup.cpp", "sbe_api.cpp", "nband_host.cpp", "frontend.cpp", "server.cpp", "wannier_tb.cpp", "atl.cpp", "spectrum.cpp", "tuner.cpp", "material.cpp",
    "haldane.hip", "haldane_velocity.hip", "haldane_adaptive.hip", "nband.hip", "nband16.hip", "nband32.hip", "analytic.hip", "series.hip", "noalloc.hip",
]
DRIVER = "antelope_mi355x"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 (v_pk_*_f32) stays off in every device pass
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        ln for ln in r.stdout.splitlines() if "packed-fp32-ops" not in ln and ln.strip()
    )
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if out.strip():
        sys.stderr.write(out + "\n")


def compile_one(src: str) -> Path:
    obj = OBJDIR / (Path(src).name + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    _run([hipcc(), *HIPCC_FLAGS, *lang, "-c", str(CSRC / src), "-o", str(obj)])
    return obj


def build_library(jobs: int = 8) -> Path:
    OBJDIR.mkdir(parents=True, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, LIB_SOURCES))
    out = LIBDIR / LIBNAME
    tmp = LIBDIR / (LIBNAME + ".tmp")
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_driver() -> Path:
    """The reference-compatible executable (reads inputParam.cfg, writes the .dat files)."""
    out = LIBDIR / DRIVER
    _run([hipcc(), "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(CSRC / "driver_main.cpp"),
          f"-L{LIBDIR}", "-lantelope_sbe", "-Wl,-rpath,$ORIGIN", "-o", str(out)])
    return out


def build_oracle() -> Path:
    """The CPU checker (tests only); compiled here so tests can load it."""
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "liboracle.so"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    args = ap.parse_args(argv)
    print(build_library(args.j))
    if not args.lib_only:
        print(build_driver())
        print(build_oracle())
    return 0


if __name__ == "__main__":
    sys.exit(main())
