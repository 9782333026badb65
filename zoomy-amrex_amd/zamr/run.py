"""Drop-in counterpart of zoomy_amrex.run_case (run_case.py:346-426), and a
reader of the AMReX plotfiles the drop-in writes (the on-disk format the
reference's tests read back through deliverable.py:33-129: a Header, and per
level a Cell_H whose FabOnDisk lines locate self-describing FABs).

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
import subprocess

import numpy as np

from . import compile as zc
from .deck import write_inputs


def _mesh(settings):
    """(dim, geom) of a run_case settings dict (keys dimension, domain,
    n_cells, as run_case.py:75-84 reads them).  A one-dimensional case is an
    nx x 1 strip whose single row of cells is square: y spans one cell width."""
    dim = int(settings.get("dimension", 2))
    counts = [int(v) for v in np.atleast_1d(settings["n_cells"])]
    bounds = [float(v) for v in settings["domain"]]
    if dim == 1:
        nx = counts[0]
        lo = (bounds[0], 0.0)
        hi = (bounds[1], (bounds[1] - bounds[0]) / nx)
        return 1, {"nx": nx, "ny": 1, "prob_lo": lo, "prob_hi": hi}
    nx, ny = counts[0], counts[-1]
    return dim, {"nx": nx, "ny": ny, "prob_lo": (bounds[0], bounds[2]), "prob_hi": (bounds[1], bounds[3])}


def _centres(geom, dim):
    """Cell centres as a (2, ncell) array, x varying fastest (the raster
    order, idx = i + nx j); y is 0 in one dimension."""
    axes = []
    for d, n in ((0, geom["nx"]), (1, geom["ny"])):
        lo, hi = geom["prob_lo"][d], geom["prob_hi"][d]
        h = (hi - lo) / n
        axes.append(lo + h * (0.5 + np.arange(n)))
    if dim == 1:
        return np.stack([axes[0], np.zeros(geom["nx"])])
    yy, xx = np.meshgrid(axes[1], axes[0], indexing="ij")
    return np.stack([xx.reshape(-1), yy.reshape(-1)])


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
    dim, geom = _mesh(settings)
    bdir = out / "_build"
    ex = bdir / "Exec"
    ex.mkdir(parents=True, exist_ok=True)
    rasters = None
    if ic is not None:
        rdir = bdir / "raster"
        rdir.mkdir(parents=True, exist_ok=True)
        Q = np.asarray(ic(_centres(geom, dim)), dtype=np.float64)
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


def _fab(path, offset):
    """One FAB of a Cell_D file: its header line (FAB <format> ((lo) (hi) (type)) ncomp)
    names the box and the component count, little-endian float64 data
    (component-major, i fastest) follows.  Returns (lo, hi, data)."""
    with open(path, "rb") as f:
        f.seek(offset)
        head = f.readline().decode()
        ints = [int(v) for v in re.findall(r"-?\d+", head[head.rindex(")))") + 3:])]
        lo, hi, ncomp = ints[0:2], ints[2:4], ints[-1]
        shape = (ncomp, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
        data = np.fromfile(f, dtype="<f8", count=int(np.prod(shape))).reshape(shape)
    return lo, hi, data


def _level(level_dir, ncomp):
    """A level's FABs (located by the FabOnDisk lines of its Cell_H) pasted
    into one (ncomp, ny, nx) array over their bounding box, NaN where no box
    lies; also that box's (ilo, jlo, nx, ny) and the boxes (lo0, lo1, hi0, hi1)."""
    fabs = []
    for line in (level_dir / "Cell_H").read_text().splitlines():
        if line.startswith("FabOnDisk:"):
            _, name, off = line.split()
            fabs.append(_fab(level_dir / name, int(off)))
    ilo = min(lo[0] for lo, _, _ in fabs)
    jlo = min(lo[1] for lo, _, _ in fabs)
    nx = max(hi[0] for _, hi, _ in fabs) - ilo + 1
    ny = max(hi[1] for _, hi, _ in fabs) - jlo + 1
    arr = np.full((ncomp, ny, nx), np.nan)
    for lo, hi, data in fabs:
        arr[:, lo[1] - jlo:hi[1] - jlo + 1, lo[0] - ilo:hi[0] - ilo + 1] = data[:ncomp]
    return arr, (ilo, jlo, nx, ny), [(lo[0], lo[1], hi[0], hi[1]) for lo, hi, _ in fabs]


def _header(path):
    """The fields of a plotfile's Header used here, read line by line."""
    it = iter((path / "Header").read_text().splitlines())
    next(it)  # format tag
    ncomp = int(next(it))
    names = [next(it) for _ in range(ncomp)]
    dim = int(next(it))
    time = float(next(it))
    finest = int(next(it))
    prob_lo = [float(v) for v in next(it).split()]
    prob_hi = [float(v) for v in next(it).split()]
    refs = [int(v) for v in next(it).split()]
    return dict(ncomp=ncomp, names=names, dim=dim, time=time, finest_level=finest, prob_lo=prob_lo,
                prob_hi=prob_hi, refs=refs)


def read_plotfile(path):
    """(ncomp, ny, nx) level-0 array, variable names and time of a plotfile."""
    path = pathlib.Path(path)
    H = _header(path)
    arr, _, _ = _level(path / "Level_0", H["ncomp"])
    return arr, H["names"], H["time"]


def read_levels(path):
    """Every level of a plotfile: a list of dicts arr (the level's array over
    its boxes' bounding box), ext (ilo, jlo, nx, ny), boxes and ref (the
    refinement relative to level 0), and the Header fields ncomp, names, dim,
    time, finest_level, prob_lo, prob_hi, refs."""
    path = pathlib.Path(path)
    H = _header(path)
    levels = []
    for lev in range(H["finest_level"] + 1):
        arr, ext, boxes = _level(path / f"Level_{lev}", H["ncomp"])
        levels.append(dict(arr=arr, ext=ext, boxes=boxes, ref=int(np.prod(H["refs"][:lev], dtype=int))))
    return levels, H
