"""Build script for libwsim (HIP, gfx950), sim_cli, and the CPU oracle.

Products (all in-tree, git-ignored, shipped with the tree):
  water_simulation_amd/lib/libwsim.so   -- the product library (C ABI, include/wsim.h)
  water_simulation_amd/bin/sim_cli      -- the scene driver
  water_simulation_amd/bin/mesh_frame   -- one frame to a PLY surface mesh
  oracle/liboracle.so                   -- CPU oracle (test infrastructure)
  oracle/_ref/libref_host.so            -- reference host sources (only where
                                           /root/reference exists)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
BIN = PKG / "bin"
OBJ = PKG / "_build"
ORACLE = ROOT / "oracle"
REFERENCE = Path("/root/reference")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off and the IEEE divide/sqrt default keep every kernel
# expression bit-identical to its binary32 evaluation on the host; packed
# FP32 instruction selection is disabled.
HIP_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-parameter", "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall"]

HIP_SOURCES = ["radix_sort.hip", "grid.hip", "grid_noalloc.hip", "morton.hip", "store.hip", "solver.hip", "adjoint.hip",
               "copy.hip", "pair_ops.hip", "surface.hip"]
HOST_SOURCES = ["host_surface.cpp", "capi.cpp", "frames.cpp"]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")
    noisy = [l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if noisy:
        sys.stderr.write("\n".join(noisy) + "\n")


def _compile(src: str) -> Path:
    out = OBJ / (src + ".o")
    path = CSRC / src
    if src.endswith(".hip"):
        cmd = [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", str(path), "-o", str(out)]
    else:
        cmd = [HIPCC, *HOST_FLAGS, "-D__HIP_PLATFORM_AMD__", "-c", str(path), "-o", str(out)]
    _run(cmd)
    return out


def build_library(jobs: int = 8) -> Path:
    OBJ.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, HIP_SOURCES + HOST_SOURCES))
    tmp = LIB / "libwsim.so.tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", *map(str, objs), "-o", str(tmp)])
    out = LIB / "libwsim.so"
    os.replace(tmp, out)
    return out


def build_sim_cli() -> Path:
    BIN.mkdir(exist_ok=True)
    out = BIN / "sim_cli"
    _run(["g++", "-std=c++17", "-O2", "-Wall", f"-I{ROOT / 'include'}", f"-I{PKG / 'include'}",
          str(PKG / "apps" / "sim_cli.cpp"), f"-L{LIB}", "-lwsim", "-Wl,-rpath,$ORIGIN/../lib",
          "-o", str(out)])
    return out


def build_mesh_frame() -> Path:
    BIN.mkdir(exist_ok=True)
    out = BIN / "mesh_frame"
    _run(["g++", "-std=c++17", "-O2", "-Wall", f"-I{ROOT / 'include'}", f"-I{PKG / 'include'}",
          str(PKG / "apps" / "mesh_frame.cpp"), f"-L{LIB}", "-lwsim", "-Wl,-rpath,$ORIGIN/../lib",
          "-o", str(out)])
    return out


def build_oracle() -> Path:
    out = ORACLE / "liboracle.so"
    tmp = ORACLE / "liboracle.so.tmp"
    _run(["g++", "-std=c++17", "-O2", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC",
          "-shared", "-Wall", str(ORACLE / "wsim_oracle.cpp"), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_reference_host() -> Path | None:
    """oracle/_ref from the reference's host-only sources, where they exist."""
    if not (REFERENCE / "solvers" / "dfsph" / "src" / "boundary_sampler.cpp").exists():
        return None
    _run(["make", "-B", "-s", "-C", str(ORACLE / "ref")])
    return ORACLE / "_ref" / "libref_host.so"


def build_all(jobs: int = 8) -> None:
    build_oracle()
    build_reference_host()
    build_library(jobs)
    build_sim_cli()
    build_mesh_frame()


if __name__ == "__main__":
    build_all()
    print("built:", LIB / "libwsim.so", BIN / "sim_cli", ORACLE / "liboracle.so")
