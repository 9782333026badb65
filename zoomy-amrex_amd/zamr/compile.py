"""Compile system libraries: the counterpart of the reference's per-system
`make` (zoomy_amrex/run_case.py:390-398, build.py:21-52).

A system is a directory holding the printer's Model.H and Numerics.H.  The
kernel + driver translation unit (csrc/zamr_system.hip) is compiled once per
system for gfx950; the MarchConstants.H of each (mesh dimension, order)
variant is compiled into a tiny object and linked beside it, giving one
shared library per variant: lib/libzamr_<system>_d<dim>o<order>.so.
"""
import os
import pathlib
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = pathlib.Path(__file__).resolve().parent.parent  # zoomy-amrex_amd/
REPO = PKG.parent
CSRC = PKG / "csrc"
COMPAT = PKG / "compat"
SYSTEMS = PKG / "systems"
TEST_SYSTEMS = REPO / "tests" / "systems"  # test-only headers (pathological sources for error paths)
INCLUDE = REPO / "include"
LIBDIR = PKG / "lib"
OBJDIR = PKG / "_build"

ARCH = os.environ.get("ZAMR_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# Shipped variants: (system, mesh dimension, spatial order).
VARIANTS = [("swe1d", 1, 1), ("swe2d", 2, 1), ("swe2d", 2, 2), ("swe2d_nf", 2, 2), ("sme2d", 2, 2),
            ("swe2d_visc", 2, 1), ("swe2d_visc", 2, 2)]
# Test systems (tests/systems/<name>): built beside the shipped ones, with
# their FMA-off twins, for the tests that drive a model's error paths.
TEST_VARIANTS = [("swe2d_fallback", 2, 2)]

# -amdgpu-use-amdgpu-trackers: the scheduler's AMDGPU register-pressure
# trackers; measured 1-2% faster on the order-2 stage kernels, which run at
# their three-waves-per-SIMD register limit.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-mllvm", "-amdgpu-use-amdgpu-trackers"]


def render_march_constants(cfl, dimension, degree):
    """MarchConstants.H in the shape zoomy_amrex/_constants.py:59-88 renders:
    c_cfl (the safety factor), c_cfl_dimension = d, c_cfl_degree_factor = 2k+1."""
    return ("#pragma once\n#include <AMReX_REAL.H>\n"
            "// c_cfl, c_cfl_dimension, c_cfl_degree_factor of the CFL law\n"
            "// dt = c_cfl * dx / (d * (2k+1) * lambda)\n"
            "namespace ZoomyConstants {\n"
            f"    constexpr amrex::Real c_cfl = {float(cfl)!r};\n"
            f"    constexpr int c_cfl_dimension = {int(dimension)};\n"
            f"    constexpr int c_cfl_degree_factor = {2 * int(degree) + 1};\n"
            "}\n")


def lib_name(system, dim, order, contract=True):
    return f"libzamr_{system}_d{dim}o{order}{'' if contract else '_nofma'}.so"


def lib_path(system, dim, order, contract=True):
    return LIBDIR / lib_name(system, dim, order, contract)


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed:\n" + " ".join(map(str, cmd)) + "\n" + r.stdout + r.stderr)
    return r


def compile_system(system_dir, name, variants, contract=True, objdir=OBJDIR, libdir=LIBDIR, cfl=0.9):
    """Compile one system directory (Model.H + Numerics.H) and link one
    library per (mesh dimension, order) in `variants`.  Returns the paths."""
    system_dir = pathlib.Path(system_dir)
    objdir = pathlib.Path(objdir)
    libdir = pathlib.Path(libdir)
    objdir.mkdir(parents=True, exist_ok=True)
    libdir.mkdir(parents=True, exist_ok=True)
    tag = "" if contract else "_nofma"
    obj = objdir / f"{name}{tag}.o"
    fp = ["-ffp-contract=fast"] if contract else ["-ffp-contract=off", "-DZAMR_BITWISE"]
    _run([HIPCC, *HIP_FLAGS, *fp, f"-I{COMPAT}", f"-I{system_dir}", f"-I{INCLUDE}", f"-I{CSRC}",
          "-c", str(CSRC / "zamr_system.hip"), "-o", str(obj)])
    outs = []
    for dim, order in variants:
        vdir = objdir / f"{name}_d{dim}o{order}"
        vdir.mkdir(parents=True, exist_ok=True)
        (vdir / "MarchConstants.H").write_text(render_march_constants(cfl, dim, order - 1))
        mobj = vdir / "march.o"
        _run([HIPCC, "-O2", "-std=c++17", "-fPIC", "-x", "c++", f"-I{COMPAT}", f"-I{vdir}", "-c",
              str(CSRC / "zamr_march.cpp"), "-o", str(mobj)])
        out = libdir / f"libzamr_{name}_d{dim}o{order}{tag}.so"
        _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", f"-Wl,-soname,{out.name}", str(obj), str(mobj),
              "-o", str(out)])
        outs.append(out)
        # the drop-in executable `main_<variant>.ex inputs` (Source/main.cpp counterpart)
        exe = libdir / f"main_{name}_d{dim}o{order}{tag}.ex"
        _run([shutil.which("g++") or "g++", "-O2", "-std=c++17", f"-I{INCLUDE}", str(CSRC / "zamr_main.cpp"),
              str(out), "-Wl,-rpath,$ORIGIN", "-o", str(exe)])
    return outs


def compile_probe(libdir=LIBDIR):
    """The bandwidth-calibration library bench.py times the stage kernels'
    memory access pattern with (csrc/zamr_probe.hip)."""
    libdir = pathlib.Path(libdir)
    libdir.mkdir(parents=True, exist_ok=True)
    out = libdir / "libzamr_probe.so"
    _run([HIPCC, *HIP_FLAGS, "-shared", str(CSRC / "zamr_probe.hip"), "-o", str(out)])
    return out


def main_path(system, dim, order, contract=True):
    return LIBDIR / f"main_{system}_d{dim}o{order}{'' if contract else '_nofma'}.ex"


def build_all(jobs=6, with_nofma=True):
    """Build every shipped variant (and the FMA-off twins used by the bitwise
    parity tests), then the test systems."""
    by_sys = {}
    for s, d, o in VARIANTS:
        by_sys.setdefault(s, []).append((d, o))
    tasks = [(SYSTEMS / s, s, v, True) for s, v in by_sys.items()]
    if with_nofma:
        tasks.append((SYSTEMS / "swe2d", "swe2d", [(2, 2), (2, 1)], False))
        tasks.append((SYSTEMS / "swe2d_nf", "swe2d_nf", [(2, 2)], False))
        tasks.append((SYSTEMS / "sme2d", "sme2d", [(2, 2)], False))
        tasks.append((SYSTEMS / "swe2d_visc", "swe2d_visc", [(2, 2), (2, 1)], False))
    tests = {}
    for s, d, o in TEST_VARIANTS:
        tests.setdefault(s, []).append((d, o))
    for s, v in tests.items():
        tasks.append((TEST_SYSTEMS / s, s, v, True))
        if with_nofma:
            tasks.append((TEST_SYSTEMS / s, s, v, False))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(compile_system, d, s, v, c) for d, s, v, c in tasks]
        return [p for f in futs for p in f.result()]


if __name__ == "__main__":
    for p in build_all():
        print(p)
    sys.exit(0)
