"""HIP path vs the CPU oracle and the reference fixtures, through the C-ABI.

Bar: <= 1e-12 relative L-infinity per stage for the FMA-contracting build,
bitwise equality for the FMA-off build of the same kernels, and the
reference's own fixtures under its np.allclose rule.
"""
import sys

import numpy as np
import pytest

from helpers import (FIXTURES_1D, bed_step, bed_step_case, bed_step_expected, case_1d, case_2d_pulse, fixture,
                     rel_linf, rel_linf_rows, smooth_state)
from oracle.pyoracle import Oracle
from zamr import cases
from zamr.deck import deck_from_case
from zamr.lib import ZamrError, ZoomyAmr

pytestmark = pytest.mark.gpu

TOL_STAGE = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu(system, case, Q0, dim=2, contract=True, init=True):
    z = ZoomyAmr(deck_from_case(case), system=system, dim=dim, order=case["spatial_order"], contract=contract)
    if init:
        z.init_data()
    if Q0 is not None:
        z.set_state(Q0)
    return z


def orc(system, case, Q0, threads=1):
    o = Oracle(system, case, threads=threads)
    o.init_solution()
    if Q0 is not None:
        o.set_state(Q0)
    return o


def run_until_done(z, chunk=1000):
    total = 0
    while True:
        n, t = z.run_steps(chunk)
        total += n
        if n < chunk:
            return total, t


@pytest.mark.parametrize("name", sorted(FIXTURES_1D))
def test_reference_fixture_1d_gpu(name):
    say("case", name)
    ic, nx, tend, wb, wall = FIXTURES_1D[name]
    c = case_1d(nx, tend, wb, wall)
    x = (np.arange(nx) + 0.5) * (10.0 / nx)
    Q0 = ic(x)[:, None, :]
    z = gpu("swe1d", c, Q0, dim=1)
    run_until_done(z)
    Q, A = z.get_state(0), z.get_state(1)
    Qr, Ar = fixture(name)
    assert np.allclose(Q, Qr.reshape(Q.shape)) and np.allclose(A, Ar.reshape(A.shape))
    o = orc("swe1d", c, Q0)
    o.evolve()
    assert rel_linf(Q, o.get_state()) <= 1e-12
    assert rel_linf(A, o.get_aux()) <= 1e-12


def test_reference_fixture_2d_pulse_gpu():
    c, Q0 = case_2d_pulse()
    z = gpu("swe2d", c, Q0)
    run_until_done(z)
    Q, A = z.get_state(0), z.get_state(1)
    Qr, Ar = fixture("swe_2d")
    assert np.allclose(Q, Qr) and np.allclose(A, Ar)
    assert np.abs(Q - Qr).max() < 1e-12


def _case(n, ny, order, wb, impl, mgs, bc=("wall", "", "", "wall"), clamp=True):
    return dict(nx=n, ny=ny, prob_lo=(0.0, 0.0), prob_hi=(1.0, ny / n), max_grid_size=mgs, spatial_order=order,
                implicit_source=impl, well_balanced=wb, clamp_positivity=clamp, cfl=0.45, dtmin=1e-7, dtmax=1.0,
                time_end=1e30, max_step=-1, bc=bc, params={"n_m": 0.03 if impl else 0.0})


STAGE_CASES = [(o, wb, impl, mgs) for o in (1, 2) for wb in (False, True) for impl in (False, True)
               for mgs in (300, 24)]


@pytest.mark.parametrize("order,wb,impl,mgs", STAGE_CASES)
def test_stage_parity(order, wb, impl, mgs):
    say("stage", order, wb, impl, mgs)
    n, ny = 70, 54
    Q0 = smooth_state(n, ny, seed=order + 2 * wb + 4 * impl)
    Q0[1, 10:14, 20:30] = 0.0  # a dry patch
    Q0[2:, 10:14, 20:30] = 0.0
    c = _case(n, ny, order, wb, impl, mgs)
    dt, t = 2.5e-3, 0.0
    o = orc("swe2d", c, Q0)
    o.update_state(t)
    o.stage(order, t, dt)
    ref = o.get_state()
    z = gpu("swe2d", c, Q0)
    z.stage(order, t, dt)
    assert rel_linf(z.get_state(), ref) <= TOL_STAGE
    zb = gpu("swe2d", c, Q0, contract=False)
    zb.stage(order, t, dt)
    assert np.array_equal(zb.get_state(), ref), "FMA-off build must match the oracle bitwise"


@pytest.mark.parametrize("order", [1, 2])
def test_advance_and_dt_parity(order):
    n, ny = 64, 48
    Q0 = smooth_state(n, ny, seed=7)
    c = _case(n, ny, order, True, True, 32)
    o = orc("swe2d", c, Q0)
    z = gpu("swe2d", c, Q0)
    zb = gpu("swe2d", c, Q0, contract=False)
    assert z.compute_dt() == o.compute_dt()
    t = 0.0
    for k in range(3):
        dt = o.compute_dt()
        assert zb.compute_dt() == dt
        assert z.compute_dt() == pytest.approx(dt, rel=1e-13)
        o.advance(t, dt)
        z.advance(t, dt)
        zb.advance(t, dt)
        t += dt
        assert rel_linf(z.get_state(), o.get_state()) <= 1e-12
        assert np.array_equal(zb.get_state(), o.get_state())
        assert np.array_equal(zb.get_state(1), o.get_aux())


def test_dam_break_run_bitwise():
    """The reference's default deck (Exec/inputs, 100^2 dam break, 4 boxes,
    order 2, t = 0.1) run to the end: FMA-off build equals the oracle bitwise,
    the FMA build within 1e-11."""
    c, _ = cases.dam_break_2d()
    o = Oracle("swe2d", c)
    o.init_solution()
    n_o, t_o, dts, ms = o.evolve()
    z = gpu("swe2d", c, None)
    n_g, t_g = run_until_done(z)
    zb = gpu("swe2d", c, None, contract=False)
    n_b, t_b = run_until_done(zb)
    assert n_b == n_o and t_b == t_o
    assert np.array_equal(zb.get_state(), o.get_state())
    assert np.array_equal(zb.get_state(1), o.get_aux())
    assert n_g == n_o
    assert rel_linf(z.get_state(), o.get_state()) <= 1e-11


def test_bump_crop_run():
    """A 256^2 instance of the 4096^2 seeded-bump workload (WB, Manning via the
    point-implicit solve, wet/dry), 6 steps on the device clock."""
    c, Q0 = cases.swe_bumps(n=256, seed=3, steps=6)
    c["max_grid_size"] = 64
    o = orc("swe2d", c, Q0)
    o.evolve()
    zb = gpu("swe2d", c, Q0, contract=False)
    run_until_done(zb)
    assert np.array_equal(zb.get_state(), o.get_state())
    z = gpu("swe2d", c, Q0)
    run_until_done(z)
    assert rel_linf(z.get_state(), o.get_state()) <= 1e-11
    assert z.get_time()[1] == 6


def test_graph_steps_equal_host_steps():
    """run_steps (dt on the device, hipGraph replay) equals TimeStep-by-TimeStep
    host driving bitwise."""
    c, Q0 = cases.swe_bumps(n=192, seed=5, steps=20)
    a = gpu("swe2d", c, Q0)
    run_until_done(a)
    b = gpu("swe2d", c, Q0)
    t = 0.0
    for _ in range(20):
        t += b.time_step(t)
    assert np.array_equal(a.get_state(), b.get_state())
    assert a.get_time()[0] == t


def test_decomposition_invariance_wet_dry_gpu():
    """The throughput build's wet fast paths (hinv = 1/h where h >= eps; one
    reciprocal per side in the well-balanced face) are decided per cell and
    per face, so results stay bitwise independent of the box decomposition
    when waves mix wet and dry cells: the surface lowered so that the bumps
    leave a third of the domain dry."""
    c, Q0 = cases.swe_bumps(n=200, seed=9, steps=6)
    Q0 = Q0.copy()
    Q0[1] = np.maximum(0.0, np.quantile(Q0[0], 0.67) - Q0[0])
    dry = float(np.mean(Q0[1] < 1e-4))
    assert 0.2 < dry < 0.5, dry
    outs = []
    for mgs in (200, 64, 37):
        z = gpu("swe2d", dict(c, max_grid_size=mgs), Q0)
        run_until_done(z)
        outs.append(z.get_state())
    assert np.isfinite(outs[0]).all()
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_decomposition_invariance_gpu():
    c, Q0 = cases.swe_bumps(n=200, seed=9, steps=4)
    outs = []
    for mgs in (200, 64, 37):
        c2 = dict(c, max_grid_size=mgs)
        z = gpu("swe2d", c2, Q0)
        run_until_done(z)
        outs.append(z.get_state())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_ragged_boxes_run_bitwise():
    """Boxes that fit no strip or row segment evenly (157 x 83 cells cut at
    37: widths 37 and 9, heights 37 and 9), walls on two sides and
    extrapolation on the others, WB, Manning and clamp: 8 steps on the
    device clock, FMA-off build equal to the oracle bitwise (lanes past a
    box's last ghost column, the last row's repeated prefetch and the
    boundary mirrors all enter)."""
    n, ny = 157, 83
    Q0 = smooth_state(n, ny, seed=11)
    Q0[1, 30:36, 40:52] = 0.0  # a dry patch
    Q0[2:, 30:36, 40:52] = 0.0
    c = dict(_case(n, ny, 2, True, True, 37, bc=("wall", "", "", "wall")), max_step=8)
    say("ragged boxes 157x83 / 37")
    o = orc("swe2d", c, Q0)
    n_o, t_o, dts, ms = o.evolve()
    zb = gpu("swe2d", c, Q0, contract=False)
    n_b, t_b = run_until_done(zb)
    assert n_b == n_o == 8 and t_b == t_o
    assert np.array_equal(zb.get_state(), o.get_state())
    assert np.array_equal(zb.get_state(1), o.get_aux())
    z = gpu("swe2d", c, Q0)
    run_until_done(z)
    assert rel_linf(z.get_state(), o.get_state()) <= 1e-11


def test_lake_at_rest_1024():
    n = 1024
    Q = smooth_state(n, n, bed=True, amp=0.0)
    Q[2:] = 0.0
    c = _case(n, n, 2, True, True, n, bc=("wall", "wall", "", ""))
    c["max_step"] = 10
    z = gpu("swe2d", c, Q)
    run_until_done(z)
    R = z.get_state()
    assert np.abs(R[0] + R[1] - 1.0).max() < 1e-13
    assert np.abs(R[2:]).max() < 1e-13


def test_full_size_one_step_parity():
    """The 4096^2 workload itself: one SSP-RK2 step against the oracle (run on
    16 host threads), plus size-independent properties after 4 more steps:
    the bed row never changes, h stays >= 0, everything finite."""
    c, Q0 = cases.swe_bumps(n=4096, seed=0, steps=1)
    say("full size: generated")
    z = gpu("swe2d", c, Q0)
    run_until_done(z)
    Qg = z.get_state()
    del z
    zb = gpu("swe2d", c, Q0, contract=False)
    run_until_done(zb)
    Qb = zb.get_state()
    del zb
    c_o = dict(c, max_grid_size=256)
    o = orc("swe2d", c_o, Q0, threads=16)
    o.evolve()
    say("full size: oracle done")
    ref = o.get_state()
    del o
    assert np.array_equal(Qb, ref), "FMA-off build must match the oracle bitwise at 4096^2"
    assert rel_linf(Qg, ref) <= 1e-12
    # momentum rows near zero: per-row error stays at FMA round-off of O(1) fluxes
    assert max(rel_linf_rows(Qg, ref)) <= 1e-9
    c5 = dict(c, max_step=5)
    z5 = gpu("swe2d", c5, Q0)
    run_until_done(z5)
    R = z5.get_state()
    assert np.isfinite(R).all()
    assert np.array_equal(R[0], Q0[0])
    assert R[1].min() >= 0.0


def test_bad_tag_is_an_error():
    c = _case(16, 16, 2, False, False, 16, bc=("Nowhere", "", "", ""))
    with pytest.raises(ZamrError) as e:
        ZoomyAmr(deck_from_case(c), system="swe2d", dim=2, order=2)
    assert "Nowhere" in str(e.value)


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("axis", [0, 1])
def test_wb_bed_step_moving_water(order, axis):
    """Well-balanced stage over a bed step with moving water (flat surface,
    one velocity): the momentum rescale of the hydrostatic reconstruction
    makes the step face's two states equal, so only the cells below the step
    change, by a closed form (tests/helpers.py: bed_step); the bitwise build
    equals the oracle, the FMA build the closed form to 1e-14."""
    say("bed step", order, axis)
    nx, ny = (70, 12) if axis == 0 else (12, 70)
    k = 33
    Q, rate = bed_step(nx=nx, ny=ny, axis=axis, k=k)
    c = dict(bed_step_case(nx, ny, order, 2), max_grid_size=24)
    dt = 0.01
    E = bed_step_expected(Q, rate, dt, axis, k)
    o = orc("swe2d", c, Q)
    o.update_state(0.0)
    o.stage(order, 0.0, dt)
    ref = o.get_state()
    assert np.array_equal(ref[:2], E[:2]) and np.abs(ref - E).max() <= 1e-14
    zb = gpu("swe2d", c, Q, contract=False)
    zb.stage(order, 0.0, dt)
    assert np.array_equal(zb.get_state(), ref)
    z = gpu("swe2d", c, Q)
    z.stage(order, 0.0, dt)
    assert np.abs(z.get_state() - E).max() <= 1e-14
