"""The `inputs` deck: the reference's runtime contract between build.py and the
driver (build.py:55-120, Source/ZoomyAmr.cpp:29-83).

write_inputs keeps the reference signature and writes the same keys, so a
caller of zoomy_amrex.build.write_inputs gets a deck this library reads.
deck_from_case turns the case dictionaries used by the tests and the bench
into deck text.
"""
import pathlib


def _flag(v):
    return "true" if v else "false"


def _domain(ncell, dim_mesh, geom):
    """(n_cell, prob_lo, prob_hi) as deck vectors: the square unit domain of
    ncell^2 cells, or the given rectangle; a three-dimensional mesh carries one
    cell across [0, 1] in z."""
    if geom is None:
        cells, lo, hi = [ncell, ncell], [0.0, 0.0], [1.0, 1.0]
    else:
        cells = [geom["nx"], geom["ny"]]
        lo, hi = list(geom["prob_lo"]), list(geom["prob_hi"])
    if dim_mesh == 3:
        cells, lo, hi = cells + [1], lo + [0.0], hi + [1.0]
    return cells, lo, hi


def write_inputs(path, ncell, dim_mesh, tend, order, plot_dt, cfl=0.45, bc="extrap",
                 implicit_source=False, implicit_global=False, friction=None, slip=None,
                 max_level=0, ref_ratio=2, geom=None, dem_file=None, release_file=None,
                 well_balanced=False, clamp_positivity=True, tag_b_max=None,
                 bc_sides=None, state_rasters=None, positivity="none", dtmax=None,
                 max_step=None, dtmin=None, max_grid_size=64):
    """Write the inputs deck of a run at `path` and return its text.

    The arguments and the keys they set are those of the reference's deck
    writer (build.py:55-120), so a caller of zoomy_amrex.build.write_inputs
    gets a deck this library reads: the mesh and box keys, geometry (never
    periodic), boundary tags (per side, or `wall` on every side), model
    parameters and rasters, output, time stepping and solver switches, and the
    tagging threshold.  Optional entries are left out when their argument is
    None, and then take the driver's defaults.  A hierarchy (max_level > 0) on
    a two-dimensional mesh uses blocking factor 2, anything else 1."""
    cells, lo, hi = _domain(ncell, dim_mesh, geom)
    vec = lambda v: " ".join(str(x) for x in v)  # noqa: E731
    sides = dict(bc_sides) if bc_sides else ({s: "wall" for s in ("x_lo", "x_hi", "y_lo", "y_hi")}
                                             if bc == "wall" else {})
    entries = [
        ("amr.max_level", max_level), ("amr.n_cell", vec(cells)), ("amr.max_grid_size", max_grid_size),
        ("amr.blocking_factor", 2 if (max_level > 0 and dim_mesh == 2) else 1), ("amr.ref_ratio", ref_ratio),
        ("amr.regrid_int", 2),
        ("geometry.prob_lo", vec(lo)), ("geometry.prob_hi", vec(hi)), ("geometry.is_periodic", vec([0] * len(cells))),
    ]
    entries += [(f"bc.{s}", sides[s]) for s in ("x_lo", "x_hi", "y_lo", "y_hi") if s in sides]
    entries += [("params.n_m", friction), ("params.lambda_s", slip), ("init.dem_file", dem_file),
                ("init.release_file", release_file),
                ("init.state_rasters", vec(state_rasters) if state_rasters else None),
                ("output.identifier", 0), ("output.plot_dt_interval", plot_dt), ("solver.time_end", tend),
                ("solver.max_step", max_step), ("solver.cfl", cfl),
                ("solver.dtmin", "1.e-7" if dtmin is None else dtmin),
                ("solver.dtmax", tend if dtmax is None else dtmax), ("solver.spatial_order", order),
                ("solver.implicit_source", _flag(implicit_source)), ("solver.implicit_global", _flag(implicit_global)),
                ("solver.well_balanced", _flag(well_balanced)), ("solver.clamp_positivity", _flag(clamp_positivity)),
                ("solver.positivity", positivity), ("tagging.threshold", 0.02), ("tagging.b_max", tag_b_max)]
    text = "".join(f"{k} = {v}\n" for k, v in entries if v is not None)
    pathlib.Path(path).write_text(text)
    return text


AMR_KEYS = dict(max_level=1, ref_ratio=2, blocking_factor=2, n_error_buf=1, grid_eff=0.7, threshold=0.01,
                b_max=1.0e30)


def deck_from_case(case, amr=None):
    """Deck text for a case dictionary (keys as in the oracle's Python handle:
    nx, ny, prob_lo, prob_hi, max_grid_size, spatial_order, implicit_source,
    well_balanced, clamp_positivity, positivity, cfl, dtmin, dtmax, time_end,
    max_step, bc, params).  `amr` (a dict over AMR_KEYS) adds the hierarchy:
    amr.max_level, ref_ratio, blocking_factor, n_error_buf, n_proper, grid_eff
    and tagging.threshold / tagging.b_max."""
    c = dict(case)
    bc = c.get("bc", ("", "", "", ""))
    a = None
    if amr is not None:
        a = dict(AMR_KEYS)
        a.update(amr)
    lines = [f"amr.max_level = {a['max_level'] if a else 0}",
             f"amr.n_cell = {c['nx']} {c['ny']}",
             f"amr.max_grid_size = {c.get('max_grid_size', 64)}",
             f"geometry.prob_lo = {c.get('prob_lo', (0.0, 0.0))[0]!r} {c.get('prob_lo', (0.0, 0.0))[1]!r}",
             f"geometry.prob_hi = {c.get('prob_hi', (1.0, 1.0))[0]!r} {c.get('prob_hi', (1.0, 1.0))[1]!r}",
             f"solver.time_end = {c.get('time_end', 1.0)!r}",
             f"solver.max_step = {c.get('max_step', -1)}",
             f"solver.cfl = {c.get('cfl', 0.5)!r}",
             f"solver.dtmin = {c.get('dtmin', 1e-7)!r}",
             f"solver.dtmax = {c.get('dtmax', 1e-2)!r}",
             f"solver.spatial_order = {c.get('spatial_order', 1)}",
             f"solver.implicit_source = {'true' if c.get('implicit_source', True) else 'false'}",
             f"solver.implicit_global = {'true' if c.get('implicit_global', False) else 'false'}",
             f"solver.well_balanced = {'true' if c.get('well_balanced', False) else 'false'}",
             f"solver.clamp_positivity = {'true' if c.get('clamp_positivity', True) else 'false'}",
             f"solver.positivity = {c.get('positivity', 'none')}"]
    for side, name in zip(("x_lo", "x_hi", "y_lo", "y_hi"), bc):
        if name:
            lines.append(f"bc.{side} = {name}")
    for k, v in (c.get("params") or {}).items():
        lines.append(f"params.{k} = {v!r}")
    if a:
        lines += [f"amr.ref_ratio = {a['ref_ratio']}", f"amr.blocking_factor = {a['blocking_factor']}",
                  f"amr.n_error_buf = {a['n_error_buf']}", f"amr.grid_eff = {a['grid_eff']!r}",
                  f"amr.n_proper = {a.get('n_proper', 1)}",
                  f"tagging.threshold = {a['threshold']!r}", f"tagging.b_max = {a['b_max']!r}"]
    return "\n".join(lines) + "\n"
