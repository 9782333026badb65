"""Seeded synthetic workloads.

Random numbers come from SplitMix64 (Steele, Lea & Flood 2014): a 64-bit
counter advanced by the golden-ratio increment and finalised by two
xor-shift-multiply rounds; U[0,1) takes the top 53 bits.  The same arrays
feed the HIP library and the CPU oracle.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    def __init__(self, seed=0):
        self.state = np.uint64(seed)

    def next_u64(self, n):
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + k * _GOLDEN
            self.state = z[-1] if n else self.state
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            return z ^ (z >> np.uint64(31))

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u


def cell_centers(nx, ny, plo=(0.0, 0.0), phi=(1.0, 1.0)):
    dx = (phi[0] - plo[0]) / nx
    dy = (phi[1] - plo[1]) / ny
    x = plo[0] + (np.arange(nx) + 0.5) * dx
    y = plo[1] + (np.arange(ny) + 0.5) * dy
    return x, y


def gaussian_bumps(x, y, rng, nbumps, amp=(0.0, 0.8), sigma=(0.02, 0.1)):
    """Sum of nbumps Gaussians A exp(-|X-c|^2 / (2 s^2)), separable in x and y."""
    A = rng.uniform(nbumps, *amp)
    s = rng.uniform(nbumps, *sigma)
    cx = rng.uniform(nbumps)
    cy = rng.uniform(nbumps)
    b = np.zeros((len(y), len(x)))
    for k in range(nbumps):
        gx = np.exp(-((x - cx[k]) ** 2) / (2 * s[k] ** 2))
        gy = np.exp(-((y - cy[k]) ** 2) / (2 * s[k] ** 2))
        b += A[k] * np.outer(gy, gx)
    return b


def surface_perturbation(x, y, rng, nmodes=8, kmax=8, amplitude=1e-2):
    """Random-phase sum of nmodes plane waves with integer wave vectors
    |k_x|, |k_y| <= kmax, scaled so that its maximum magnitude is amplitude."""
    kx = np.floor(rng.uniform(nmodes, -kmax, kmax + 1)).astype(int)
    ky = np.floor(rng.uniform(nmodes, -kmax, kmax + 1)).astype(int)
    ph = rng.uniform(nmodes, 0.0, 2 * np.pi)
    a = rng.uniform(nmodes, 0.5, 1.0)
    eta = np.zeros((len(y), len(x)))
    for m in range(nmodes):
        # sin(a+b) expanded so the 2-D field is built from 1-D factors
        sx, cx = np.sin(2 * np.pi * kx[m] * x), np.cos(2 * np.pi * kx[m] * x)
        sy, cy = np.sin(2 * np.pi * ky[m] * y + ph[m]), np.cos(2 * np.pi * ky[m] * y + ph[m])
        eta += a[m] * (np.outer(cy, sx) + np.outer(sy, cx))
    mx = np.abs(eta).max()
    return eta * (amplitude / mx) if mx > 0 else eta


def dam_break_2d(n=100, time_end=0.1):
    """The reference's default deck (Exec/inputs): 2-D dam break on [0,1]^2,
    h = 2 left of x = 0.5 and 1 right, flat bed, order 2, CFL 0.45,
    extrapolation, explicit source, t_end = 0.1.  The state is init_solution's."""
    case = dict(nx=n, ny=n, prob_lo=(0.0, 0.0), prob_hi=(1.0, 1.0), max_grid_size=64, spatial_order=2,
                implicit_source=False, well_balanced=False, clamp_positivity=True, cfl=0.45, dtmin=1e-7,
                dtmax=0.1, time_end=time_end, max_step=-1, bc=("", "", "", ""), params={})
    return case, None


def box_list(seed=0, n_amr=600, n_members=64, steps=200, member_range=(256, 1024)):
    """A batched box list: n_amr patches of a 3-level hierarchy (refinement
    ratio 2, side lengths per axis drawn from {16, 32, 64, 128} with weights
    0.4/0.3/0.2/0.1, level l with dx = 2^-l / 1024, smooth seeded wet states),
    plus n_members dam-break ensemble members of sizes uniform in
    member_range^2 (h_left ~ U[1.5, 2.5], h_right = 1, dx = 1/nx).

    Every box is its own physical domain with extrapolation boundaries.
    Time-step groups: one per AMR level (boxes of a level share dt, the
    segmented max of their CFL speeds) and one per ensemble member.
    Returns (boxes, states, solver_keys): boxes are dicts with lo, hi, dx,
    prob_lo, group; states are (4, ny, nx) arrays."""
    rng = SplitMix64(seed)
    sides = np.array([16, 32, 64, 128])
    cum = np.cumsum([0.4, 0.3, 0.2, 0.1])
    boxes, states = [], []
    u = rng.uniform(3 * n_amr)
    for k in range(n_amr):
        nx = int(sides[np.searchsorted(cum, u[3 * k], side="right")])
        ny = int(sides[np.searchsorted(cum, u[3 * k + 1], side="right")])
        lev = min(int(u[3 * k + 2] * 3), 2)
        dx = 2.0 ** (-lev) / 1024
        ph = rng.uniform(4, 0, 2 * np.pi)
        x = (np.arange(nx) + 0.5) * dx
        y = (np.arange(ny) + 0.5) * dx
        X, Y = np.meshgrid(x, y, indexing="xy")
        kx, ky = 2 * np.pi / (nx * dx), 2 * np.pi / (ny * dx)
        h = 1.0 + 0.05 * np.sin(kx * X + ph[0]) * np.cos(ky * Y + ph[1])
        hu = 0.1 * h * np.sin(ky * Y + ph[2])
        hv = 0.1 * h * np.cos(kx * X + ph[3])
        boxes.append(dict(lo=(0, 0), hi=(nx - 1, ny - 1), dx=(dx, dx), prob_lo=(0.0, 0.0), group=lev))
        states.append(np.stack([np.zeros_like(h), h, hu, hv]))
    lo_m, hi_m = member_range
    dims = np.floor(rng.uniform(2 * n_members, lo_m, hi_m + 1)).astype(int)
    hl = rng.uniform(n_members, 1.5, 2.5)
    for m in range(n_members):
        nx, ny = int(dims[2 * m]), int(dims[2 * m + 1])
        x = (np.arange(nx) + 0.5) / nx
        h = np.where(x < 0.5, hl[m], 1.0)[None, :].repeat(ny, axis=0)
        z = np.zeros_like(h)
        boxes.append(dict(lo=(0, 0), hi=(nx - 1, ny - 1), dx=(1.0 / nx, 1.0 / ny), prob_lo=(0.0, 0.0), group=3 + m))
        states.append(np.stack([z, h, z, z]))
    keys = dict(spatial_order=2, implicit_source=False, well_balanced=False, clamp_positivity=True, cfl=0.45,
                dtmin=1e-12, dtmax=1.0, time_end=1e30, max_step=steps)
    return boxes, states, keys


def swe_bumps(n=4096, seed=0, nbumps=32, steps=100, manning=0.03, max_grid_size=None):
    """Single-level 2-D shallow water over nbumps seeded Gaussian bumps:
    eta = 1 + a 1e-2 random-phase perturbation, h = max(0, eta - b) (dry
    where b > eta), zero velocity, Manning friction through the point-implicit
    source, well-balanced reconstruction, order 2, `steps` SSP-RK2 steps."""
    rng = SplitMix64(seed)
    x, y = cell_centers(n, n)
    b = gaussian_bumps(x, y, rng, nbumps)
    eta = 1.0 + surface_perturbation(x, y, rng)
    h = np.maximum(0.0, eta - b)
    z = np.zeros_like(h)
    Q = np.stack([b, h, z, z])
    case = dict(nx=n, ny=n, prob_lo=(0.0, 0.0), prob_hi=(1.0, 1.0),
                max_grid_size=max_grid_size or n, spatial_order=2, implicit_source=True,
                well_balanced=True, clamp_positivity=True, cfl=0.45, dtmin=1e-7, dtmax=1.0,
                time_end=1.0e30, max_step=steps, bc=("", "", "", ""), params={"n_m": manning})
    return case, Q


def dry_columns(n=2048, seed=0, ncols=24, radius=(0.02, 0.05), steps=20, max_grid_size=256, positivity="clamp"):
    """Water columns (depth 1) at seeded positions over a dry flat bed, walls:
    the AMR workload (amr.max_level = 1, ErrorEst's wet/front tags, so level 1
    follows the spreading water).  Order 2, Manning 0.03 point-implicit, no
    well-balancing; positivity "clamp" (the reference's default,
    solver.clamp_positivity) or "mood"."""
    rng = SplitMix64(seed)
    x, y = cell_centers(n, n)
    X, Y = np.meshgrid(x, y)
    h = np.zeros((n, n))
    u = rng.uniform(3 * ncols)
    for k in range(ncols):
        cx, cy = 0.1 + 0.8 * u[3 * k], 0.1 + 0.8 * u[3 * k + 1]
        r = radius[0] + (radius[1] - radius[0]) * u[3 * k + 2]
        h = np.where((X - cx) ** 2 + (Y - cy) ** 2 < r * r, 1.0, h)
    z = np.zeros_like(h)
    case = dict(nx=n, ny=n, prob_lo=(0.0, 0.0), prob_hi=(1.0, 1.0), max_grid_size=max_grid_size,
                spatial_order=2, implicit_source=True, well_balanced=False,
                clamp_positivity=positivity == "clamp", positivity="mood" if positivity == "mood" else "none",
                cfl=0.45, dtmin=1e-9, dtmax=1.0, time_end=1.0e30, max_step=steps, bc=("wall",) * 4,
                params={"n_m": 0.03})
    return case, np.stack([z, h, z, z])


def box_filter(a, passes=3, width=5):
    """`passes` passes of a width x width moving average, edges replicated."""
    r = width // 2
    for _ in range(passes):
        for ax in (0, 1):
            p = np.pad(a, [(r, r) if k == ax else (0, 0) for k in range(2)], mode="edge")
            c = np.cumsum(p, axis=ax)
            c = np.concatenate([np.zeros_like(np.take(c, [0], axis=ax)), c], axis=ax)
            hi = np.take(c, np.arange(width, width + a.shape[ax]), axis=ax)
            lo = np.take(c, np.arange(0, a.shape[ax]), axis=ax)
            a = (hi - lo) / width
    return a


def swe_adjoint(n=1024, seed=0, steps=200, ngauges=256, nfd=32, seg=20, bc=("", "", "", "")):
    """Gauge-misfit inversion on the per-cell-Manning shallow-water system.

    Base: bathymetry b0 = 16 seeded Gaussian bumps (A ~ U[0, 0.25], sigma ~
    U[0.05, 0.15]), Manning n0 = 0.03 everywhere.  Truth: b = b0 + 8 bumps of
    amplitude 0.1 (sigma ~ U[0.02, 0.06]); n = 0.03 + U[-0.01, 0.01] per
    cell, smoothed by 3 passes of a 5x5 box filter.  Initial free surface
    eta0 = 1 + 0.1 exp(-|x - (1/2, 1/2)|^2 / (2 0.05^2)), h0 = max(0, eta0 - b),
    zero velocity.  Order 2, Rusanov, point-implicit Manning, no
    well-balancing and no positivity clamp (everything stays wet), `steps`
    SSP-RK2 steps with the dt sequence of the base run frozen; `ngauges`
    distinct seeded gauge cells; `nfd` seeded finite-difference cells, each a
    gauge shifted by up to 3 cells.  Returns a dict."""
    rng = SplitMix64(seed)
    x, y = cell_centers(n, n)
    b0 = gaussian_bumps(x, y, rng, 16, amp=(0.0, 0.25), sigma=(0.05, 0.15))
    bt = b0 + gaussian_bumps(x, y, rng, 8, amp=(0.1, 0.1), sigma=(0.02, 0.06))
    nt = box_filter(0.03 + rng.uniform(n * n, -0.01, 0.01).reshape(n, n), passes=3, width=5)
    X, Y = np.meshgrid(x, y, indexing="xy")
    eta0 = 1.0 + 0.1 * np.exp(-((X - 0.5) ** 2 + (Y - 0.5) ** 2) / (2 * 0.05 ** 2))
    cells = []
    seen = set()
    while len(cells) < ngauges:
        u = rng.uniform(2)
        c = (int(u[0] * n), int(u[1] * n))
        if c not in seen:
            seen.add(c)
            cells.append(c)
    gi = np.array([c[0] for c in cells], dtype=np.int32)
    gj = np.array([c[1] for c in cells], dtype=np.int32)
    fi, fj = [], []
    for _ in range(nfd):
        u = rng.uniform(3)
        g = min(int(u[0] * ngauges), ngauges - 1)
        off = np.floor(rng.uniform(2, -3, 4)).astype(int)
        fi.append(int(np.clip(gi[g] + off[0], 0, n - 1)))
        fj.append(int(np.clip(gj[g] + off[1], 0, n - 1)))
    case = dict(nx=n, ny=n, prob_lo=(0.0, 0.0), prob_hi=(1.0, 1.0), max_grid_size=n, spatial_order=2,
                implicit_source=True, well_balanced=False, clamp_positivity=False, cfl=0.45, dtmin=1e-12,
                dtmax=1.0, time_end=1.0e30, max_step=steps, bc=tuple(bc), params={})
    return dict(case=case, b0=b0, b_truth=bt, n0=np.full((n, n), 0.03), n_truth=nt, eta0=eta0,
                gauges=(gi, gj), fd_cells=(np.array(fi), np.array(fj)), steps=steps, seg=seg)


def initial_state(b, eta0):
    """[b, max(0, eta0 - b), 0, 0]."""
    z = np.zeros_like(b)
    return np.stack([b, np.maximum(0.0, eta0 - b), z, z])


def sme_bumps(n=2048, seed=0, steps=50, nu=0.01, lambda_s=0.1, max_grid_size=None, bc=("", "", "", "")):
    """Level-2 shallow-moment workload: bathymetry = 16 seeded Gaussian bumps
    (A ~ U[0, 0.5], sigma ~ U[0.05, 0.15]), h0 = 1 - b (all wet: b < 0.87 for
    seed 0), every velocity moment hu_k, hv_k (k = 0, 1, 2) = U[-0.05, 0.05] h0
    per cell, slip (lambda_s) + Newtonian (nu) friction through the
    point-implicit source, order-2 MUSCL / path-conservative Rusanov, no
    well-balancing (its row scaling is written for the shallow-water layout),
    `steps` SSP-RK2 steps.  State rows [b, h, hu0, hu1, hu2, hv0, hv1, hv2]."""
    rng = SplitMix64(seed)
    x, y = cell_centers(n, n)
    b = gaussian_bumps(x, y, rng, 16, amp=(0.0, 0.5), sigma=(0.05, 0.15))
    h = 1.0 - b
    Q = np.zeros((8, n, n))
    Q[0], Q[1] = b, h
    for k in (2, 3, 4, 5, 6, 7):
        Q[k] = rng.uniform(n * n, -0.05, 0.05).reshape(n, n) * h
    case = dict(nx=n, ny=n, prob_lo=(0.0, 0.0), prob_hi=(1.0, 1.0), max_grid_size=max_grid_size or n,
                spatial_order=2, implicit_source=True, well_balanced=False, clamp_positivity=True, cfl=0.45,
                dtmin=1e-9, dtmax=1.0, time_end=1.0e30, max_step=steps, bc=tuple(bc),
                params={"nu": nu, "lambda_s": lambda_s})
    return case, Q
