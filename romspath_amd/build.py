"""Build librp.so (HIP, gfx950) in-tree with hipcc.

The library is the product: hand-written kernels plus the C-ABI of
include/rp_abi.h.  Built here, it travels with the source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib" / "librp.so"
SOURCES = ["rp_step.hip", "rp_adjoint.hip", "rp_sort.hip", "rp_locate.hip", "rp_abi.hip"]
HEADERS = ["rp_device.h", "rp_kernels.h", "rp_dual.h", "rp_step_dev.h"]
# host-only C++ (no device code): the trajectory file writer and the
# netCDF classic reader of the hydrodynamic input
HOST_SOURCES = ["rp_ncout.cpp", "rp_ncin.cpp", "rp_modal_host.cpp"]
CXX = os.environ.get("CXX", "g++")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # FP64 parity with the reference's IEEE arithmetic: no FMA contraction
    "-ffp-contract=off",
    # no packed FP32 instructions (some forms have given wrong results)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


FORTRAN = PKG / "fortran"
DRIVER = PKG / "lib" / "rp_driver"
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")


def build_fortran(verbose: bool = True) -> Path:
    """The ISO_C_BINDING module and the Fortran driver, linked to librp.so."""
    out = PKG / "lib"
    mods = out / "fmod"
    mods.mkdir(parents=True, exist_ok=True)
    steps = [
        [FC, "-O2", "-J", str(mods), "-c", str(FORTRAN / "romspath_hip.f90"), "-o", str(mods / "romspath_hip.o")],
        [FC, "-O2", "-I", str(mods), "-c", str(FORTRAN / "rp_driver.f90"), "-o", str(mods / "rp_driver.o")],
        [FC, "-o", str(DRIVER), str(mods / "rp_driver.o"), str(mods / "romspath_hip.o"),
         "-L", str(out), "-lrp", "-Wl,-rpath,$ORIGIN"],
    ]
    for cmd in steps:
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"Fortran build failed ({r.returncode})")
    return DRIVER


NC_CHECK = PKG / "lib" / "rp_nc_check_san"


def build_nc_check(verbose: bool = True) -> Path:
    """The netCDF reader and writer (host C++) with a small driver, built with
    AddressSanitizer and UndefinedBehaviorSanitizer for the tests that feed
    them damaged files.  Host code only: g++, no device code."""
    NC_CHECK.parent.mkdir(parents=True, exist_ok=True)
    cmd = [CXX, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-I", str(ROOT / "include"), "-o", str(NC_CHECK),
           str(CSRC / "rp_nc_check.cpp"), str(CSRC / "rp_ncin.cpp"), str(CSRC / "rp_ncout.cpp")]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"sanitizer build failed ({r.returncode})")
    return NC_CHECK


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)


def _check(proc, what, verbose):
    out, err = proc.communicate()
    if proc.returncode != 0:
        sys.stderr.write(out + err)
        raise RuntimeError(f"{what} failed ({proc.returncode})")
    if verbose and err.strip():
        # hipcc's host pass does not know the packed-fp32 feature: harmless
        lines = [l for l in err.splitlines() if "packed-fp32-ops" not in l]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")


def build(verbose: bool = True) -> Path:
    """Compile every source to an object in parallel, then link librp.so."""
    LIB.parent.mkdir(parents=True, exist_ok=True)
    objdir = LIB.parent / "obj"
    objdir.mkdir(exist_ok=True)
    compile_flags = [f for f in FLAGS if f != "-shared"]
    procs = []
    objs = []
    for src in SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        objs.append(str(obj))
        procs.append((src, _run([HIPCC, *compile_flags, "-I", str(ROOT / "include"), "-c",
                                 str(CSRC / src), "-o", str(obj)], verbose)))
    for src in HOST_SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        objs.append(str(obj))
        procs.append((src, _run([CXX, "-O2", "-std=c++17", "-fPIC", "-Wall", "-I", str(ROOT / "include"),
                                 "-c", str(CSRC / src), "-o", str(obj)], verbose)))
    for src, proc in procs:
        _check(proc, f"compile {src}", verbose)
    tmp = LIB.with_suffix(".so.tmp")
    _check(_run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objs],
                verbose), "link", verbose)
    os.replace(tmp, LIB)
    build_fortran(verbose)
    build_nc_check(verbose)
    return LIB


if __name__ == "__main__":
    build()
