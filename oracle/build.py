"""Builds the CPU oracle libraries (test infrastructure): one
oracle/_build/liboracle_<system>.so per system directory, from
oracle/oracle_capi.cpp with FMA contraction off (as the reference builds with
-fmad=false, build.py:33-41) and OpenMP over boxes."""
import pathlib
import shutil
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
REPO = HERE.parent
SYSTEMS = REPO / "zoomy-amrex_amd" / "systems"
TEST_SYSTEMS = REPO / "tests" / "systems"
COMPAT = REPO / "zoomy-amrex_amd" / "compat"
OUT = HERE / "_build"
CXX = shutil.which("g++") or "g++"


def build(systems=None):
    OUT.mkdir(exist_ok=True)
    dirs = {p.name: p for base in (SYSTEMS, TEST_SYSTEMS) if base.exists() for p in base.iterdir()
            if (p / "Model.H").exists()}
    systems = systems or sorted(dirs)
    outs = []
    procs = []
    for s in systems:
        out = OUT / f"liboracle_{s}.so"
        cmd = [CXX, "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp", "-std=c++17", "-fPIC", "-shared",
               f"-I{COMPAT}", f"-I{dirs[s]}", str(HERE / "oracle_capi.cpp"), "-o", str(out)]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        outs.append(out)
    for cmd, p in procs:
        log, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError("oracle build failed:\n" + " ".join(cmd) + "\n" + log)
    return outs


if __name__ == "__main__":
    for p in build():
        print(p)
