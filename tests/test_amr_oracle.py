"""Pins of the oracle's two-level AMR restatement (oracle/zamr_oracle_amr.hpp):
the AMReX algorithms the reference's AMR path calls (ZoomyAmr.cpp:137-288,
733-762) are checked against what mathematics fixes, never against a retyped
formula:
  * cell_cons_interp conserves (the four children average to the parent) and
    reproduces linear data exactly; its limiter zeroes a direction at an
    extremum of any component;
  * average_down of the interpolant returns the coarse data;
  * Berger-Rigoutsos clusters cover every tag, are disjoint and tight, reach
    the efficiency target, and split two separated blobs at the hole;
  * the fine level's decomposition into boxes does not change any result;
  * with no tags the hierarchy is the single-level run, bit for bit;
  * with every cell tagged the fine level is a single-level run at twice the
    resolution, bit for bit, and the composite mass is conserved to round-off
    inside walls.
"""
import numpy as np
import pytest

from oracle.pyoracle import AmrOracle, Oracle, cluster
from zamr import cases


def column_case(n=32, steps=6, cfl=0.4, bc=("wall",) * 4, order=2, **kw):
    c = dict(nx=n, ny=n, prob_lo=(0.0, 0.0), prob_hi=(1.0, 1.0), max_grid_size=16, spatial_order=order,
             implicit_source=False, well_balanced=False, clamp_positivity=False, positivity="none", cfl=cfl,
             dtmin=1e-9, dtmax=1.0, time_end=1e30, max_step=steps, bc=tuple(bc), params={})
    c.update(kw)
    return c


def column_state(n, wet_outside=0.0):
    x, y = cases.cell_centers(n, n)
    X, Y = np.meshgrid(x, y)
    h = np.where((X - 0.5) ** 2 + (Y - 0.5) ** 2 < 0.04, 1.0, wet_outside)
    z = np.zeros_like(h)
    return np.stack([z, h, z, z])


def test_interp_conserves_and_is_exact_on_linear_data():
    n = 16
    o = AmrOracle("swe2d", column_case(n=n))
    j, i = np.mgrid[0:n, 0:n].astype(float)
    Q = np.stack([0.1 + 0.01 * i - 0.02 * j, 1.0 + 0.03 * i + 0.05 * j, 0.2 - 0.01 * i, 0.3 + 0.02 * j])
    o.set_state(Q)
    for jc in range(1, n - 1):
        for ic in range(1, n - 1):
            kids = np.array([o.interp(2 * ic + a, 2 * jc + b) for b in (0, 1) for a in (0, 1)])
            assert np.allclose(kids.mean(axis=0), Q[:, jc, ic], rtol=0, atol=1e-15)
            for b in (0, 1):
                for a in (0, 1):
                    xf, yf = ic + (a - 0.5) * 0.5, jc + (b - 0.5) * 0.5  # child centre in coarse index units
                    exact = [0.1 + 0.01 * xf - 0.02 * yf, 1.0 + 0.03 * xf + 0.05 * yf, 0.2 - 0.01 * xf,
                             0.3 + 0.02 * yf]
                    assert np.allclose(kids[2 * b + a], exact, rtol=0, atol=2e-15)


def test_interp_limiter_is_shared_across_components():
    """An x-extremum in one component (the bed) flattens every component's x
    slope in that coarse cell; y slopes are untouched."""
    n = 16
    o = AmrOracle("swe2d", column_case(n=n))
    j, i = np.mgrid[0:n, 0:n].astype(float)
    b = np.where(i == 8, 0.5, np.where(i == 9, 0.2, 0.0))  # a peak with a non-zero central slope
    Q = np.stack([b, 1.0 + 0.03 * i + 0.05 * j, 0 * i, 0 * i])
    o.set_state(Q)
    k00, k10 = o.interp(16, 10), o.interp(17, 10)
    assert k00[1] == k10[1]  # no x variation of h inside coarse cell (8, 5)
    k01 = o.interp(16, 11)
    assert k01[1] > k00[1]  # y slope kept
    k_other = o.interp(12, 10), o.interp(13, 10)  # coarse cell (6, 5): no extremum, linear h
    assert k_other[1][1] > k_other[0][1]


def test_interp_at_domain_edge_extrapolates_first_order():
    """Beyond the domain the coarse stencil repeats the edge cell, so the
    edge cell's x slope is limited to zero."""
    n = 16
    o = AmrOracle("swe2d", column_case(n=n))
    j, i = np.mgrid[0:n, 0:n].astype(float)
    Q = np.stack([0 * i, 1.0 + 0.03 * i, 0 * i, 0 * i])
    o.set_state(Q)
    assert o.interp(0, 6)[1] == o.interp(1, 6)[1] == Q[1, 3, 0]


def test_average_down_of_interpolant_is_identity():
    n = 32
    o = AmrOracle("swe2d", column_case(n=n))
    j, i = np.mgrid[0:n, 0:n].astype(float)
    Q = np.stack([0 * i, 1.0 + 0.01 * i + 0.02 * j, 0.001 * i, 0.002 * j])
    o.set_state(Q)
    o.init_levels()  # everything is wet: one fine box over the domain
    assert o.finest_level() == 1 and (_cover(o.boxes(1), (64, 64)) == 1).all()
    o.average_down()
    assert np.allclose(o.get_state(0), Q, rtol=0, atol=1e-15)


def _cover(boxes, shape):
    m = np.zeros(shape, dtype=int)
    for (x0, y0, x1, y1) in boxes:
        m[y0:y1 + 1, x0:x1 + 1] += 1
    return m


@pytest.mark.parametrize("seed", range(6))
def test_clusters_cover_tags_disjoint_tight_efficient(seed):
    rng = np.random.default_rng(seed)
    shape = (40, 48)
    t = np.zeros(shape, dtype=bool)
    for _ in range(4):  # a few blobs and a scatter
        cy, cx, r = rng.integers(4, 36), rng.integers(4, 44), rng.integers(2, 7)
        yy, xx = np.mgrid[0:shape[0], 0:shape[1]]
        t |= (yy - cy) ** 2 + (xx - cx) ** 2 < r * r
    t |= rng.random(shape) < 0.01
    pts = np.argwhere(t)[:, ::-1]
    boxes = cluster(pts, eff=0.7)
    cov = _cover(boxes, shape)
    assert cov.max() == 1  # disjoint
    assert np.all(cov[t] == 1)  # every tag covered
    for (x0, y0, x1, y1) in boxes:
        sub = t[y0:y1 + 1, x0:x1 + 1]
        assert sub[0, :].any() and sub[-1, :].any() and sub[:, 0].any() and sub[:, -1].any()  # tight
        assert sub.mean() >= 0.7


def test_cluster_cuts_at_the_hole():
    pts = [(x, y) for x in range(0, 4) for y in range(0, 4)] + [(x, y) for x in range(10, 13) for y in range(0, 4)]
    assert sorted(cluster(pts)) == [(0, 0, 3, 3), (10, 0, 12, 3)]  # 28 of 52 cells: below 0.7
    # at exactly the efficiency target nothing is cut
    pts = [(x, y) for x in range(0, 4) for y in range(0, 4)] + [(x, y) for x in range(7, 10) for y in range(0, 4)]
    assert cluster(pts) == [(0, 0, 9, 3)]  # 28 of 40 cells = 0.7
    # a single full rectangle is already efficient
    assert cluster([(x, y) for x in range(5) for y in range(3)]) == [(0, 0, 4, 2)]


def test_cluster_bisects_a_sparse_diagonal():
    """No hole and no inflection on a diagonal: the bisection in the later
    direction (y wins a tie) and recursion down to single points."""
    pts = [(k, k) for k in range(4)]
    boxes = cluster(pts)
    assert sorted(boxes) == [(k, k, k, k) for k in range(4)]


def _run(amr_kw, case, Q0, steps):
    o = AmrOracle("swe2d", case, amr=amr_kw)
    o.set_state(Q0)
    o.init_levels()
    n, t, dts, ms, nl = o.evolve(step_limit=steps)
    return o, n, t, dts, ms, nl


def test_fine_decomposition_does_not_change_results():
    c = column_case(n=32, steps=8)
    Q0 = column_state(32)
    a, *ra = _run(dict(fine_max_grid_size=64), c, Q0, 8)
    b, *rb = _run(dict(fine_max_grid_size=8), c, Q0, 8)
    assert len(b.boxes(1)) > len(a.boxes(1))
    assert np.array_equal(ra[2], rb[2])  # dts
    assert np.array_equal(a.get_state(0), b.get_state(0))
    assert np.array_equal(a.get_state(1), b.get_state(1), equal_nan=True)


def test_no_tags_is_the_single_level_run():
    c = column_case(n=32, steps=6)
    Q0 = column_state(32)
    o, n, t, dts, ms, nl = _run(dict(b_max=-1.0), c, Q0, 6)
    assert o.finest_level() == 0 and set(nl) == {1}
    s = Oracle("swe2d", c)
    s.set_state(Q0)
    n2, t2, dts2, _ = s.evolve(step_limit=6)
    assert n == n2 and t == t2 and np.array_equal(dts, dts2)
    assert np.array_equal(o.get_state(0), s.get_state())


def test_all_tagged_fine_level_is_the_fine_single_level_run():
    """Every cell wet, so every cell tagged: level 1 covers the domain, has no
    coarse-fine ghosts, and takes two of its own CFL steps per coarse step —
    a single-level run at twice the resolution from the interpolated start."""
    c = column_case(n=16, steps=5, cfl=0.45)
    Q0 = column_state(16, wet_outside=0.5)
    o = AmrOracle("swe2d", c)
    o.set_state(Q0)
    o.init_levels()
    assert o.boxes(1) == [(0, 0, 31, 31)]
    Q1 = o.get_state(1)
    m0 = o.total_mass()
    n, t, dts, ms, nl = o.evolve(step_limit=5)
    f = Oracle("swe2d", dict(c, nx=32, ny=32, max_step=10))
    f.set_state(Q1)
    f.evolve(step_limit=10)
    assert np.array_equal(o.get_state(1), f.get_state())
    assert np.all(np.abs(ms - m0) <= 1e-14 * m0)  # walls, fine level over everything
