"""The `inputs` deck: the reference's runtime contract between build.py and the
driver (build.py:55-120, Source/ZoomyAmr.cpp:29-83).

write_inputs keeps the reference signature and writes the same keys, so a
caller of zoomy_amrex.build.write_inputs gets a deck this library reads.
deck_from_case turns the case dictionaries used by the tests and the bench
into deck text.
"""
import pathlib


def write_inputs(path, ncell, dim_mesh, tend, order, plot_dt, cfl=0.45, bc="extrap",
                 implicit_source=False, implicit_global=False, friction=None, slip=None,
                 max_level=0, ref_ratio=2, geom=None, dem_file=None, release_file=None,
                 well_balanced=False, clamp_positivity=True, tag_b_max=None,
                 bc_sides=None, state_rasters=None, positivity="none", dtmax=None,
                 max_step=None, dtmin=None, max_grid_size=64):
    """Write an inputs deck at `path` (same parameters as build.py:55-60)."""
    if geom is not None:
        nx, ny = geom["nx"], geom["ny"]
        gx0, gy0 = geom["prob_lo"]
        gx1, gy1 = geom["prob_hi"]
        ncell_line = f"{nx} {ny}" + (" 1" if dim_mesh == 3 else "")
        prob_lo = f"{gx0} {gy0}" + (" 0.0" if dim_mesh == 3 else "")
        prob_hi = f"{gx1} {gy1}" + (" 1.0" if dim_mesh == 3 else "")
    else:
        ncell_line = " ".join([str(ncell)] * 2 + (["1"] if dim_mesh == 3 else []))
        prob_hi = "1.0 1.0 1.0" if dim_mesh == 3 else "1.0 1.0"
        prob_lo = "0.0 0.0 0.0" if dim_mesh == 3 else "0.0 0.0"
    isper = "0 0 0" if dim_mesh == 3 else "0 0"
    lines = [f"amr.max_level     = {max_level}",
             f"amr.n_cell        = {ncell_line}",
             f"amr.max_grid_size = {max_grid_size}",
             f"amr.blocking_factor = {2 if (max_level > 0 and dim_mesh == 2) else 1}",
             f"amr.ref_ratio     = {ref_ratio}",
             "amr.regrid_int    = 2",
             f"geometry.prob_lo  = {prob_lo}",
             f"geometry.prob_hi  = {prob_hi}",
             f"geometry.is_periodic = {isper}"]
    if bc_sides:
        for side in ("x_lo", "x_hi", "y_lo", "y_hi"):
            if side in bc_sides:
                lines.append(f"bc.{side} = {bc_sides[side]}")
    elif bc == "wall":
        lines += ["bc.x_lo = wall", "bc.x_hi = wall", "bc.y_lo = wall", "bc.y_hi = wall"]
    if friction is not None:
        lines.append(f"params.n_m = {friction}")
    if slip is not None:
        lines.append(f"params.lambda_s = {slip}")
    if dem_file is not None:
        lines.append(f"init.dem_file = {dem_file}")
    if release_file is not None:
        lines.append(f"init.release_file = {release_file}")
    if state_rasters:
        lines.append("init.state_rasters = " + " ".join(map(str, state_rasters)))
    lines += ["output.identifier       = 0",
              f"output.plot_dt_interval = {plot_dt}",
              f"solver.time_end        = {tend}"]
    if max_step is not None:
        lines.append(f"solver.max_step        = {max_step}")
    lines += [f"solver.cfl             = {cfl}",
              f"solver.dtmin           = {'1.e-7' if dtmin is None else dtmin}",
              f"solver.dtmax           = {tend if dtmax is None else dtmax}",
              f"solver.spatial_order   = {order}",
              f"solver.implicit_source = {'true' if implicit_source else 'false'}",
              f"solver.implicit_global = {'true' if implicit_global else 'false'}",
              f"solver.well_balanced   = {'true' if well_balanced else 'false'}",
              f"solver.clamp_positivity = {'true' if clamp_positivity else 'false'}",
              f"solver.positivity      = {positivity}",
              "tagging.threshold      = 0.02"]
    if tag_b_max is not None:
        lines.append(f"tagging.b_max          = {tag_b_max}")
    text = "\n".join(lines) + "\n"
    pathlib.Path(path).write_text(text)
    return text


AMR_KEYS = dict(max_level=1, ref_ratio=2, blocking_factor=2, n_error_buf=1, grid_eff=0.7, threshold=0.01,
                b_max=1.0e30)


def deck_from_case(case, amr=None):
    """Deck text for a case dictionary (keys as in the oracle's Python handle:
    nx, ny, prob_lo, prob_hi, max_grid_size, spatial_order, implicit_source,
    well_balanced, clamp_positivity, positivity, cfl, dtmin, dtmax, time_end,
    max_step, bc, params).  `amr` (a dict over AMR_KEYS) adds the hierarchy:
    amr.max_level, ref_ratio, blocking_factor, n_error_buf, grid_eff and
    tagging.threshold / tagging.b_max."""
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
                  f"tagging.threshold = {a['threshold']!r}", f"tagging.b_max = {a['b_max']!r}"]
    return "\n".join(lines) + "\n"
