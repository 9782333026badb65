"""Drop-in counterpart of zoomy_amrex.run_case (run_case.py:346-426) and of
the plotfile reader the reference tests use (deliverable.py:33-129).

run_case(system, settings, output_dir, ic=...) follows the reference
sequence: sample the initial condition on the cell centres into one float64
raster per state row (run_case.py:98-115), write the inputs deck with
write_inputs (run_case.py:235-288), build the system library if needed
(`make`, run_case.py:398), run `main_<variant>.ex inputs` as a subprocess in
<output_dir>/_build/Exec with check=True (run_case.py:418-422) and return the
last plotfile.  `system` is a shipped system name or a directory holding a
printer-generated Model.H + Numerics.H.
"""
import os
import pathlib
import re
import shutil
import struct
import subprocess

import numpy as np

from . import compile as zc
from .deck import write_inputs


def _grid(settings):  # run_case.py:75-84
    dim = int(settings.get("dimension", 2))
    dom = settings["domain"]
    nc = settings["n_cells"]
    if dim == 1:
        nx = int(nc[0] if isinstance(nc, (list, tuple)) else nc)
        ny = 1
        x0, x1 = float(dom[0]), float(dom[1])
        y0, y1 = 0.0, (x1 - x0) / nx
    else:
        nx, ny = (int(nc[0]), int(nc[1])) if isinstance(nc, (list, tuple)) else (int(nc), int(nc))
        x0, x1, y0, y1 = [float(v) for v in dom[:4]]
    return dim, {"nx": nx, "ny": ny, "prob_lo": (x0, y0), "prob_hi": (x1, y1)}


def _cell_centers(geom, dim):  # run_case.py:87-95
    x0, x1 = geom["prob_lo"][0], geom["prob_hi"][0]
    nx = geom["nx"]
    dx = (x1 - x0) / nx
    xc = x0 + (np.arange(nx) + 0.5) * dx
    if dim == 1:
        return np.vstack([xc, np.zeros_like(xc)])
    y0, y1 = geom["prob_lo"][1], geom["prob_hi"][1]
    ny = geom["ny"]
    dy = (y1 - y0) / ny
    yc = y0 + (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(xc, yc, indexing="xy")
    return np.vstack([X.ravel(order="C"), Y.ravel(order="C")])


def _variant(system, dim, order, contract=True):
    if isinstance(system, (str, os.PathLike)) and pathlib.Path(system).is_dir():
        d = pathlib.Path(system)
        name = d.name
        exe = zc.main_path(name, dim, order, contract)
        if not exe.exists():
            zc.compile_system(d, name, [(dim, order)], contract=contract)
        return exe
    exe = zc.main_path(system, dim, order, contract)
    if not exe.exists():
        raise FileNotFoundError(f"{exe} is not built; run __graft_entry__.build()")
    return exe


def run_case(system, settings, output_dir, ic=None, bc_sides=None, contract=True):
    """Run a case like zoomy_amrex.run_case and return the last plotfile path.

    ic: callable X (2, ncell) -> (n_q, ncell), the model's initial condition
    sampled on the cell centres.  bc_sides: tag names per side (default the
    West/East/South/North mapping of run_case.py:281-283)."""
    out = pathlib.Path(output_dir)
    dim, geom = _grid(settings)
    bdir = out / "_build"
    ex = bdir / "Exec"
    ex.mkdir(parents=True, exist_ok=True)
    rasters = None
    if ic is not None:
        rdir = bdir / "raster"
        rdir.mkdir(parents=True, exist_ok=True)
        Q = np.asarray(ic(_cell_centers(geom, dim)), dtype=np.float64)
        rasters = []
        for c in range(Q.shape[0]):
            p = rdir / f"ic_{c}.raw"
            Q[c].astype(np.float64).tofile(p)
            rasters.append(str(p.resolve()))
    order = int(settings.get("spatial_order", 1))
    if bc_sides is None:
        bc_sides = ({"x_lo": "West", "x_hi": "East", "y_lo": "South", "y_hi": "North"} if dim == 2 else
                    {"x_lo": "West", "x_hi": "East"})
    write_inputs(ex / "inputs", geom["nx"], 2, tend=settings.get("time_end", 0.1), order=order,
                 plot_dt=settings.get("time_end", 0.1) / max(1, settings.get("output_snapshots", 10)),
                 cfl=settings.get("cfl", 0.45), max_level=int(settings.get("max_level", 0)),
                 dtmax=settings.get("dtmax"), max_step=settings.get("max_step"), dtmin=settings.get("dtmin"),
                 geom=geom, state_rasters=rasters, max_grid_size=int(settings.get("max_grid_size", 64)),
                 bc_sides=bc_sides, well_balanced=bool(settings.get("well_balanced", False)),
                 clamp_positivity=False, positivity=str(settings.get("positivity", "none")),
                 implicit_source=bool(settings.get("implicit_source", False)),
                 friction=settings.get("friction"))
    exe = _variant(system, dim, order, contract)
    for stale in ex.glob("plt_*"):
        shutil.rmtree(stale, ignore_errors=True)
    subprocess.run([str(exe), "inputs"], cwd=ex, check=True, capture_output=True, text=True)
    plts = sorted(ex.glob("plt_*"))
    if not plts:
        raise RuntimeError("run_case: the run produced no plotfile")
    return plts[-1]


_BOX = re.compile(r"\(\(([-\d,]+)\) \(([-\d,]+)\)")


def _read_level(lvl_dir, ncomp):
    """(ncomp, ny, nx) array over the level's box bounding box (NaN where no
    box), its (ilo, jlo, nx, ny) and the boxes (lo0, lo1, hi0, hi1)."""
    txt = (lvl_dir / "Cell_H").read_text()
    boxes = [([int(v) for v in a.split(",")], [int(v) for v in b.split(",")]) for a, b in _BOX.findall(txt)]
    fods = re.findall(r"FabOnDisk:\s+(\S+)\s+(\d+)", txt)
    ilo = min(lo[0] for lo, _ in boxes)
    jlo = min(lo[1] for lo, _ in boxes)
    nx = max(hi[0] for _, hi in boxes) - ilo + 1
    ny = max(hi[1] for _, hi in boxes) - jlo + 1
    out = np.full((ncomp, ny, nx), np.nan)
    for (lo, hi), (fname, off) in zip(boxes, fods):
        bx, by = hi[0] - lo[0] + 1, hi[1] - lo[1] + 1
        with open(lvl_dir / fname, "rb") as fb:
            fb.seek(int(off))
            fb.readline()
            n = bx * by * ncomp
            data = np.array(struct.unpack("<%dd" % n, fb.read(8 * n))).reshape(ncomp, by, bx)
        out[:, lo[1] - jlo:hi[1] - jlo + 1, lo[0] - ilo:hi[0] - ilo + 1] = data
    return out, (ilo, jlo, nx, ny), [(lo[0], lo[1], hi[0], hi[1]) for lo, hi in boxes]


def read_plotfile(path):
    """(ncomp, ny, nx) level-0 array, variable names and time of a plotfile."""
    path = pathlib.Path(path)
    lines = (path / "Header").read_text().splitlines()
    ncomp = int(lines[1])
    names = lines[2:2 + ncomp]
    time = float(lines[2 + ncomp + 1])
    out, _, _ = _read_level(path / "Level_0", ncomp)
    return out, names, time


def read_levels(path):
    """Every level of a plotfile, as deliverable.read_levels (deliverable.py:
    114-129) returns them: a list of dicts arr / ext / boxes / ref
    (cumulative refinement), and the header fields ncomp, names,
    finest_level, prob_lo, prob_hi, refs, time."""
    path = pathlib.Path(path)
    lines = (path / "Header").read_text().splitlines()
    ncomp = int(lines[1])
    i = 2 + ncomp
    H = dict(ncomp=ncomp, names=lines[2:i], dim=int(lines[i]), time=float(lines[i + 1]),
             finest_level=int(lines[i + 2]), prob_lo=[float(v) for v in lines[i + 3].split()],
             prob_hi=[float(v) for v in lines[i + 4].split()],
             refs=[int(v) for v in lines[i + 5].split()])
    levels, cum = [], 1
    for lev in range(H["finest_level"] + 1):
        if lev > 0:
            cum *= H["refs"][lev - 1] if lev - 1 < len(H["refs"]) else 2
        arr, ext, boxes = _read_level(path / f"Level_{lev}", ncomp)
        levels.append(dict(arr=arr, ext=ext, boxes=boxes, ref=cum))
    return levels, H
