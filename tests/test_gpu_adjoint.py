"""Adjoint of the frozen-dt advance (per-cell-Manning system): the forward it
differentiates against the oracle, and its gradients with respect to the
bathymetry and Manning fields against central finite differences of the GPU
forward and of the CPU oracle's forward, perturbed cell by cell.

The loss is 0.5 * sum over steps and gauges of (h - h_truth)^2; the initial
state is [b, h0, h0 u0, h0 v0] with h0 = max(0, eta0 - b), so d/db includes
the initial depth.  minmod is not differentiable where it switches argument
(ties, and slopes of opposite sign, which is everywhere the fluid rests): the
adjoint takes the derivative of the branch the forward took, and the finite
differences use the same convention by replaying the base forward's limiter
choices (zamr_adj_freeze_limiter, the oracle's limiter tape).  The small
cases add a smooth background current so that no region rests."""
import numpy as np
import pytest
import torch

from oracle.pyoracle import Oracle
from zamr import cases
from zamr.adjoint import GaugeRun, GaugeSeries, frozen_dts
from zamr.lib import ZamrError, ZoomyAmr
from zamr.deck import deck_from_case

pytestmark = pytest.mark.gpu

FD_EPS = 1e-6
# the loss is ~1e-5 as sensitive to the Manning field as to the bathymetry:
# its finite differences take a larger step; over these short runs the
# friction number tau ~ n^2 stays small, so the loss is close to quadratic in
# n and a central difference of step 1e-3 has negligible truncation error
FD_EPS_N = 1e-3


def vortex(n):
    """A smooth background current, tangential at the walls and nonzero
    almost everywhere, so that the flow has no resting regions."""
    x, y = cases.cell_centers(n, n)
    X, Y = np.meshgrid(x, y)
    u = 0.2 * np.sin(np.pi * X) * np.cos(np.pi * Y) + 0.05
    v = -0.2 * np.cos(np.pi * X) * np.sin(np.pi * Y) + 0.03
    return u, v


def island(n):
    """A bed bump that rises 0.25 above the still surface (eta = 1), off the
    domain's centre: its shore dries and wets as the hump's wave runs up."""
    x, y = cases.cell_centers(n, n)
    X, Y = np.meshgrid(x, y)
    return 1.25 * np.exp(-((X - 0.27) ** 2 + (Y - 0.7) ** 2) / (2 * 0.08 ** 2))


def setup(n=64, steps=40, seg=7, seed=1, bc=("", "", "", ""), contract=True, ngauges=24, nfd=10, gauges=None,
          current=True, freeze=True, manning=None, keys=None, bed=None):
    """The inversion case at size n; the truth run, then (freeze) the limiter
    frozen at the base forward's choices for the derivative checks.
    `manning` raises the base Manning coefficient (and shifts the truth's by
    the same amount) so that short runs feel the friction; `keys` overrides
    solver keys of the case (well_balanced, clamp_positivity,
    max_grid_size); `bed` is added to the base and truth bathymetry."""
    c = cases.swe_adjoint(n=n, seed=seed, steps=steps, ngauges=ngauges, nfd=nfd, seg=seg, bc=bc)
    if keys:
        c["case"] = dict(c["case"], **keys)
    if bed is not None:
        c["b0"] = c["b0"] + bed
        c["b_truth"] = c["b_truth"] + bed
    if gauges is not None:
        c["gauges"] = gauges
    if manning is not None:
        c["n_truth"] = c["n_truth"] + (manning - 0.03)
        c["n0"] = np.full((n, n), manning)
    u, v = vortex(n) if current else (np.zeros((n, n)), np.zeros((n, n)))

    def state(b):
        q = cases.initial_state(b, c["eta0"])
        q[2] = q[1] * u
        q[3] = q[1] * v
        return q

    c["state"] = state
    c["uv"] = (u, v)
    dts = frozen_dts(c["case"], state(c["b0"]), c["n0"], steps, contract=contract)
    run = GaugeRun(c["case"], c["gauges"], dts, seg=seg, contract=contract)
    truth = run.forward(state(c["b_truth"]), c["n_truth"])
    if freeze:
        run.freeze_limiter(True)
    return c, run, truth, dts


def misfit_diff(Gp, Gm, truth):
    """L(+) - L(-) for L = 0.5 sum (G - T)^2, without forming L."""
    return 0.5 * np.sum((Gp - Gm) * (Gp + Gm - 2.0 * truth))


def grads(run, c, truth, device="cpu"):
    """(loss, dL/db, dL/dn) through GaugeSeries with tensors on `device`
    (CUDA tensors take the zero-copy path)."""
    b = torch.tensor(c["b0"], device=device, requires_grad=True)
    nf = torch.tensor(c["n0"], device=device, requires_grad=True)
    eta0 = torch.tensor(c["eta0"], device=device)
    u, v = (torch.tensor(a, device=device) for a in c["uv"])
    h = torch.clamp(eta0 - b, min=0.0)
    Q0 = torch.stack([b, h, h * u, h * v])
    G = GaugeSeries.apply(Q0, nf, run)
    assert G.device == Q0.device
    loss = 0.5 * ((G - torch.tensor(truth, device=device)) ** 2).sum()
    loss.backward()
    return float(loss.detach()), b.grad.cpu().numpy().copy(), nf.grad.cpu().numpy().copy()


def fd_noise(Gp, truth, eps):
    """Rounding of a central difference of the misfit: one independent ulp on
    every gauge depth, weighted by its misfit (root-sum-square)."""
    return np.sqrt(np.sum((np.spacing(np.abs(Gp)) * (Gp - truth)) ** 2)) / (2 * eps)


class FD(float):
    """A central difference that carries its rounding estimate."""

    def __new__(cls, value, noise):
        x = float.__new__(cls, value)
        x.noise = noise
        return x


def fd_cell(run, c, truth, field, i, j, eps=None):
    eps = eps or (FD_EPS if field == "b" else FD_EPS_N)
    b, nf = c["b0"].copy(), c["n0"].copy()
    a = b if field == "b" else nf
    a0 = a[j, i]
    a[j, i] = a0 + eps
    Gp = run.forward(c["state"](b), nf)
    a[j, i] = a0 - eps
    Gm = run.forward(c["state"](b), nf)
    return FD(misfit_diff(Gp, Gm, truth) / (2 * eps), fd_noise(Gp, truth, eps))


def rel_err(g, fd, scale, raw=False):
    """|g - fd| relative to |fd| (floored at 1e-3 of the field's largest
    gradient), not counting what the difference's own rounding explains
    (raw=True counts all of it)."""
    err = abs(g - fd) - (0.0 if raw else 4.0 * getattr(fd, "noise", 0.0))
    return max(err, 0.0) / max(abs(fd), 1e-3 * scale)


def test_frozen_forward_matches_oracle_bitwise():
    c, run, truth, dts = setup(contract=False, steps=20, seg=6, freeze=False)
    Q0 = c["state"](c["b0"])
    nf = c["n_truth"]
    G = run.forward(Q0, nf)
    final = run.z.get_state()
    o = Oracle("swe2d_nf", c["case"])
    o.init_solution()
    o.set_state(Q0)
    aux = run.aux0.copy()
    aux[1] = nf
    o.set_aux(aux)
    gi, gj = c["gauges"]
    t = 0.0
    Go = []
    for dt in dts:
        o.advance(t, dt)
        t += dt
        Go.append(o.get_state()[1][gj, gi])
    assert np.array_equal(G, np.array(Go))
    assert np.array_equal(final, o.get_state())


def test_gradient_matches_gpu_finite_differences():
    c, run, truth, _ = setup(manning=0.3)
    L, gb, gn = grads(run, c, truth)
    assert L > 0 and np.isfinite(gb).all() and np.isfinite(gn).all()
    sb, sn = np.abs(gb).max(), np.abs(gn).max()
    worst, where = 0.0, None
    for i, j in zip(*c["fd_cells"]):
        for field, g, sc in (("b", gb, sb), ("n", gn, sn)):
            err = rel_err(g[j, i], fd_cell(run, c, truth, field, i, j), sc)
            if err > worst:
                worst, where = err, (field, i, j, g[j, i])
    assert worst <= 1e-6, (worst, where)


def test_gradient_matches_oracle_finite_differences():
    """The reference gradient: the CPU oracle's forward perturbed cell by
    cell, with its limiter frozen at its own base run's choices.  The GPU
    side is the FMA-free build, whose forward is the oracle's bit for bit, so
    that both freeze the same limiter branches."""
    c, run, truth, dts = setup(steps=16, seg=5, manning=0.3, contract=False)
    L, gb, gn = grads(run, c, truth)
    gi, gj = c["gauges"]
    o = Oracle("swe2d_nf", c["case"])
    o.init_solution()

    def oracle_series(b, nf):
        o.set_state(c["state"](b))
        aux = run.aux0.copy()
        aux[1] = nf
        o.set_aux(aux)
        t, G = 0.0, []
        for dt in dts:
            o.advance(t, dt)
            t += dt
            G.append(o.get_state()[1][gj, gi])
        return np.array(G)

    o.set_tape(1)
    G0 = oracle_series(c["b0"], c["n0"])
    assert np.array_equal(G0, run.forward(c["state"](c["b0"]), c["n0"]))
    sb, sn = np.abs(gb).max(), np.abs(gn).max()
    worst = 0.0
    for i, j in list(zip(*c["fd_cells"]))[:4]:
        for field, g, sc in (("b", gb, sb), ("n", gn, sn)):
            b, nf = c["b0"].copy(), c["n0"].copy()
            a = b if field == "b" else nf
            a0 = a[j, i]
            e = FD_EPS if field == "b" else FD_EPS_N
            a[j, i] = a0 + e
            o.set_tape(2)
            Gp = oracle_series(b, nf)
            a[j, i] = a0 - e
            o.set_tape(2)
            Gm = oracle_series(b, nf)
            err = rel_err(g[j, i], misfit_diff(Gp, Gm, truth) / (2 * e), sc)
            if err > worst:
                worst, where = err, (field, i, j)
    o.close()
    assert worst <= 1e-6, (worst, where)


def test_walls_and_edges():
    """Reflective walls on all sides, gauges and perturbed cells at the walls
    and corners: the ghost-cell adjoints fold back through the transposed
    wall condition in x and in y."""
    n = 48
    gi = np.array([0, 1, n - 1, n - 2, 0, n - 1, 5, 20, 0, n - 1, 17, 30], dtype=np.int32)
    gj = np.array([0, 7, n - 1, 11, n - 1, 0, 0, n - 1, 25, 33, 1, n - 2], dtype=np.int32)
    c, run, truth, _ = setup(n=n, steps=30, seg=8, bc=("wall",) * 4, gauges=(gi, gj), manning=0.3)
    L, gb, gn = grads(run, c, truth)
    sb, sn = np.abs(gb).max(), np.abs(gn).max()
    cells = [(0, 0), (1, 0), (0, 1), (n - 1, n - 1), (n - 2, n - 1), (n - 1, 3), (2, n - 1), (n - 1, 20), (1, 1)]
    worst, where = 0.0, None
    for i, j in cells:
        for field, g, sc in (("b", gb, sb), ("n", gn, sn)):
            err = rel_err(g[j, i], fd_cell(run, c, truth, field, i, j), sc)
            if err > worst:
                worst, where = err, (field, i, j, g[j, i])
    assert worst <= 1e-6, (worst, where)


def test_directional_derivative():
    """<grad, d> against the derivative of the loss along a smooth direction d
    touching every cell (bathymetry and Manning together), with the limiter
    frozen and, separately, free."""
    for freeze in (True, False):
        c, run, truth, _ = setup(n=64, steps=30, seg=9, freeze=freeze)
        L, gb, gn = grads(run, c, truth)
        x, y = cases.cell_centers(64, 64)
        X, Y = np.meshgrid(x, y)
        db = np.sin(3 * X + 1) * np.cos(2 * Y)
        dn = 0.01 * np.cos(5 * X) * np.sin(4 * Y + 0.5)
        an = np.sum(gb * db) + np.sum(gn * dn)
        e = 1e-6
        Gp = run.forward(c["state"](c["b0"] + e * db), c["n0"] + e * dn)
        Gm = run.forward(c["state"](c["b0"] - e * db), c["n0"] - e * dn)
        fd = misfit_diff(Gp, Gm, truth) / (2 * e)
        assert abs(an - fd) <= 1e-6 * abs(fd), (freeze, an, fd)


def test_zero_copy_path_matches_host_path():
    """CUDA tensors go through the library's device arrays on the caller's
    stream (zamr_device_ptr, zamr_set_stream, zamr_adj_device_io): the same
    loss and gradients, bit for bit, as the host-memory path, on torch's
    default stream and on a side stream, with the host entry points never
    called, and the host path still right after it."""
    c, run, truth, _ = setup(n=64, steps=30, seg=7, manning=0.3, freeze=False)
    ref = grads(run, c, truth)
    host_calls = ("set_state", "get_state", "adj_forward", "adj_backward", "adj_gradient")
    saved = {k: getattr(run.z, k) for k in host_calls}

    def refuse(*a, **k):
        raise AssertionError("host-memory entry point called on the zero-copy path")

    for k in host_calls:
        setattr(run.z, k, refuse)
    try:
        got = grads(run, c, truth, device="cuda")
        side = torch.cuda.Stream()
        with torch.cuda.stream(side):
            got_side = grads(run, c, truth, device="cuda")
        torch.cuda.synchronize()
    finally:
        for k, f in saved.items():
            setattr(run.z, k, f)
    for g in (got, got_side):
        assert abs(g[0] - ref[0]) <= 1e-14 * abs(ref[0])  # torch's device sum of the loss
        assert np.array_equal(g[1], ref[1]) and np.array_equal(g[2], ref[2])
    again = grads(run, c, truth)
    assert again[0] == ref[0] and np.array_equal(again[1], ref[1]) and np.array_equal(again[2], ref[2])
    # in device I/O mode a host array is refused before anything is queued
    host = np.zeros((len(run.dts), len(run.gauges[0])))
    with pytest.raises(ZamrError) as e:
        run.z.adj_forward_device(host.ctypes.data)
    assert e.value.code == 1
    run.close()


WB_CLAMP = dict(well_balanced=True, clamp_positivity=True)
# Near a shore the friction (~ n^2 / h^(4/3)) makes the loss strongly curved
# in n, and a central difference of step 1e-3 carries truncation errors of a
# few 1e-6 there (measured: 2.4e-6 at 1e-3, 2.5e-7 at 3e-4, 2.3e-8 at 1e-4 on
# a shore cell, falling as eps^2); these cases take 3e-4.
FD_EPS_SHORE = 3e-4


def min_depth_unclamped(c, dts, contract=True):
    """The least depth after any step of the base forward run with the
    positivity clamp off (negative where the clamp has work to do)."""
    z = ZoomyAmr(deck_from_case(dict(c["case"], clamp_positivity=False, max_step=10 ** 9)), system="swe2d_nf",
                 dim=2, order=2, contract=contract)
    z.init_data()
    aux = z.get_state(field=1)
    aux[1] = c["n0"]
    z.set_state(c["state"](c["b0"]))
    z.set_state(aux, field=1)
    lo, t = np.inf, 0.0
    for dt in dts:
        z.advance(t, dt)
        t += dt
        lo = min(lo, float(z.get_state()[1].min()))
    z.close()
    return lo


@pytest.mark.parametrize("bed,wb", [("bumps", True), ("island", True), ("island", False)])
def test_well_balanced_clamped_gradient_matches_finite_differences(bed, wb):
    """The C2-style forward (hydrostatic reconstruction with its balancing bed
    source, UpdateState's positivity clamp): gradients against central
    differences of the GPU forward, cell by cell and along a smooth
    whole-field direction.  `island`: a bump above the surface with a dry
    top and a shore; without the hydrostatic reconstruction the run there
    makes negative depths that the clamp cuts (checked: the same run without
    the clamp goes negative)."""
    extra = island(64) if bed == "island" else None
    c, run, truth, dts = setup(n=64, steps=40, seg=7, manning=0.3, keys=dict(WB_CLAMP, well_balanced=wb), bed=extra)
    if not wb:
        assert min_depth_unclamped(c, dts) < 0.0
    L, gb, gn = grads(run, c, truth)
    assert np.isfinite(gb).all() and np.isfinite(gn).all()
    sb, sn = np.abs(gb).max(), np.abs(gn).max()
    worst, where = 0.0, None
    for i, j in zip(*c["fd_cells"]):
        for field, g, sc in (("b", gb, sb), ("n", gn, sn)):
            err = rel_err(g[j, i], fd_cell(run, c, truth, field, i, j, eps=FD_EPS if field == "b" else FD_EPS_SHORE),
                          sc)
            if err > worst:
                worst, where = err, (field, i, j, g[j, i])
    assert worst <= 1e-6, (worst, where)
    x, y = cases.cell_centers(64, 64)
    X, Y = np.meshgrid(x, y)
    db = np.sin(3 * X + 1) * np.cos(2 * Y)
    dn = 0.01 * np.cos(5 * X) * np.sin(4 * Y + 0.5)
    an = np.sum(gb * db) + np.sum(gn * dn)
    e = 1e-6
    Gp = run.forward(c["state"](c["b0"] + e * db), c["n0"] + e * dn)
    Gm = run.forward(c["state"](c["b0"] - e * db), c["n0"] - e * dn)
    fd = misfit_diff(Gp, Gm, truth) / (2 * e)
    assert abs(an - fd) <= 1e-6 * abs(fd), (an, fd)


def test_well_balanced_clamped_gradient_matches_oracle_finite_differences():
    """The same forward in the FMA-free build against central differences of
    the CPU oracle's well-balanced, clamped forward (limiter frozen at its own
    base run's choices), over the drying island."""
    c, run, truth, dts = setup(steps=16, seg=5, manning=0.3, contract=False, keys=WB_CLAMP, bed=island(64))
    L, gb, gn = grads(run, c, truth)
    gi, gj = c["gauges"]
    o = Oracle("swe2d_nf", c["case"])
    o.init_solution()

    def oracle_series(b, nf):
        o.set_state(c["state"](b))
        aux = run.aux0.copy()
        aux[1] = nf
        o.set_aux(aux)
        t, G = 0.0, []
        for dt in dts:
            o.advance(t, dt)
            t += dt
            G.append(o.get_state()[1][gj, gi])
        return np.array(G)

    o.set_tape(1)
    G0 = oracle_series(c["b0"], c["n0"])
    assert np.array_equal(G0, run.forward(c["state"](c["b0"]), c["n0"]))
    sb, sn = np.abs(gb).max(), np.abs(gn).max()
    worst, where = 0.0, None
    for i, j in list(zip(*c["fd_cells"]))[:4]:
        for field, g, sc in (("b", gb, sb), ("n", gn, sn)):
            b, nf = c["b0"].copy(), c["n0"].copy()
            a = b if field == "b" else nf
            a0 = a[j, i]
            e = FD_EPS if field == "b" else FD_EPS_SHORE
            a[j, i] = a0 + e
            o.set_tape(2)
            Gp = oracle_series(b, nf)
            a[j, i] = a0 - e
            o.set_tape(2)
            Gm = oracle_series(b, nf)
            err = rel_err(g[j, i], FD(misfit_diff(Gp, Gm, truth) / (2 * e), fd_noise(Gp, truth, e)), sc)
            if err > worst:
                worst, where = err, (field, i, j)
    o.close()
    assert worst <= 1e-6, (worst, where)


def test_multi_box_forward_and_gradient_equal_single_box():
    """A deck chopped into 16 boxes (max_grid_size 16): the gauge series, the
    end state the forward leaves in the boxes and the gradients equal those
    of the one-box deck bit for bit, through the host path and the zero-copy
    path."""
    one = setup(n=64, steps=24, seg=7, manning=0.3, keys=WB_CLAMP, bed=island(64), freeze=False)
    many = setup(n=64, steps=24, seg=7, manning=0.3, keys=dict(WB_CLAMP, max_grid_size=16), bed=island(64),
                 freeze=False)
    (c1, r1, t1, d1), (c2, r2, t2, d2) = one, many
    assert r2.z.shape()[2] == 16 and np.array_equal(d1, d2) and np.array_equal(t1, t2)
    Q0 = c1["state"](c1["b0"])
    assert np.array_equal(r1.forward(Q0, c1["n0"]), r2.forward(Q0, c2["n0"]))
    assert np.array_equal(r1.z.get_state(), r2.z.get_state())
    g1 = grads(r1, c1, t1)
    for dev in ("cpu", "cuda"):
        g2 = grads(r2, c2, t2, device=dev)
        # (torch sums the loss in another order on the device)
        assert abs(g1[0] - g2[0]) <= 1e-14 * abs(g1[0]), dev
        assert np.array_equal(g1[1], g2[1]) and np.array_equal(g1[2], g2[2]), dev


def test_checkpoint_interval_does_not_change_the_gradient():
    """Recomputation from checkpoints reproduces the stored states bit for bit."""
    c, run, truth, dts = setup(steps=24, seg=24, freeze=False)
    _, gb, gn = grads(run, c, truth)
    for seg in (1, 5):
        r = GaugeRun(c["case"], c["gauges"], dts, seg=seg)
        _, b2, n2 = grads(r, c, truth)
        assert np.array_equal(gb, b2) and np.array_equal(gn, n2), seg


def test_full_size_gradient_matches_finite_differences():
    """The 1024^2, 200-step inversion case (fluid at rest at t = 0, 256
    gauges, a checkpoint every 20 steps): central FD at the 32 seeded cells,
    eps = 1e-6 for the bathymetry, raw.  Over these 200 short steps (t ~ 4e-3)
    the friction barely acts: the Manning gradients are ~1e-7 of the
    bathymetry ones, and a single-cell difference resolves them only to the
    rounding of the misfit sums, so the per-cell Manning check takes eps =
    1e-3 net of that rounding; the Manning gradient is checked raw along the
    direction of its largest derivative."""
    c, run, truth, _ = setup(n=1024, steps=200, seg=20, seed=0, ngauges=256, nfd=32, current=False)
    L, gb, gn = grads(run, c, truth)
    sb, sn = np.abs(gb).max(), np.abs(gn).max()
    eb, en, rb, rn = [], [], [], []
    for i, j in zip(*c["fd_cells"]):
        fb, fn = fd_cell(run, c, truth, "b", i, j), fd_cell(run, c, truth, "n", i, j)
        eb.append(rel_err(gb[j, i], fb, sb))
        en.append(rel_err(gn[j, i], fn, sn))
        rb.append(rel_err(gb[j, i], fb, sb, raw=True))
        rn.append(rel_err(gn[j, i], fn, sn, raw=True))
    print("full-size FD relative errors (raw): b max %.3e median %.3e; n max %.3e median %.3e"
          % (max(rb), float(np.median(rb)), max(rn), float(np.median(rn))), flush=True)
    print("beyond the differences' rounding: b max %.3e; n max %.3e" % (max(eb), max(en)), flush=True)
    assert max(rb) <= 1e-6, max(rb)
    assert max(en) <= 1e-6, max(en)
    # The Manning gradient raw, along d = sign(dL/dn): of all perturbations
    # of unit amplitude per cell this has the largest derivative, sum |dL/dn|
    # (~6e-5 here).  Two FP64 forwards differ by rounding of ~2e-15 in the
    # misfit whatever the step, so a central difference carries ~1e-15/eps of
    # absolute rounding: ~4e-5 of the derivative at eps = 1e-6 (printed), and
    # below 1e-6 of it from eps = 1e-4 on, where the truncation (the loss is
    # close to quadratic in n) is still far smaller.
    d = np.sign(gn)
    an = float(np.sum(gn * d))
    Q0 = c["state"](c["b0"])
    for e in (1e-6, 1e-4, 1e-3):
        Gp = run.forward(Q0, c["n0"] + e * d)
        Gm = run.forward(Q0, c["n0"] - e * d)
        fd = misfit_diff(Gp, Gm, truth) / (2 * e)
        print("Manning along sign(dL/dn), eps %.0e: adjoint %.12e FD %.12e raw rel %.3e"
              % (e, an, fd, abs(an - fd) / abs(fd)), flush=True)
        if e >= 1e-4:
            assert abs(an - fd) <= 1e-6 * abs(fd), (e, an, fd)


def test_unsupported_configurations():
    case = cases.swe_adjoint(n=32, steps=4)["case"]
    z = ZoomyAmr(deck_from_case(case), system="swe2d", dim=2, order=2)
    z.init_data()
    with pytest.raises(ZamrError) as e:
        z.adj_setup([1e-3] * 4, [1], [1], 2)
    assert e.value.code == 5
    for keys in (dict(positivity="mood"), dict(implicit_source=False), dict(implicit_global=True)):
        z = ZoomyAmr(deck_from_case(dict(case, **keys)), system="swe2d_nf", dim=2, order=2)
        z.init_data()
        with pytest.raises(ZamrError) as e:
            z.adj_setup([1e-3] * 4, [1], [1], 2)
        assert e.value.code == 5, keys
    z = ZoomyAmr(deck_from_case(case), system="swe2d_nf", dim=2, order=2)
    z.init_data()
    with pytest.raises(ZamrError):
        z.adj_setup([1e-3] * 4, [1, 1], [2, 2], 2)  # repeated gauge cell
    with pytest.raises(ZamrError):
        z.adj_forward()  # the failed setup left nothing to run
