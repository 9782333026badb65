"""Gradients through the frozen-dt SSP-RK2 advance of the per-cell-Manning
shallow-water system: the library's adjoint (include/zamr.h, zamr_adj_*)
behind a torch.autograd.Function.

    run = GaugeRun(case, gauges, dts)            # one library context
    Q0 = torch.stack([b, (eta0 - b).clamp(min=0), 0*b, 0*b])
    G = GaugeSeries.apply(Q0, n_field, run)      # (nsteps, ngauges) depths
    loss = 0.5 * ((G - G_truth) ** 2).sum()
    loss.backward()                               # b.grad, n_field.grad

The library does the work on the GPU; torch only carries the tensors (the
wrapper moves them through host memory)."""
import numpy as np
import torch

from .deck import deck_from_case
from .lib import ZoomyAmr


class GaugeRun:
    """A library context of the per-cell-Manning system with a frozen dt
    sequence, a checkpoint every `seg` steps and a set of gauge cells."""

    def __init__(self, case, gauges, dts, seg=20, contract=True, system="swe2d_nf"):
        self.z = ZoomyAmr(deck_from_case(case), system=system, dim=2, order=2, contract=contract)
        self.z.init_data()
        self.aux0 = self.z.get_state(field=1)
        self.dts = np.asarray(dts, dtype=np.float64)
        self.gauges = gauges
        self.z.adj_setup(self.dts, gauges[0], gauges[1], seg)
        self.generation = 0

    def set_inputs(self, Q0, nfield):
        self.z.set_state(np.asarray(Q0, dtype=np.float64))
        aux = self.aux0.copy()
        aux[1] = nfield
        self.z.set_state(aux, field=1)

    def forward(self, Q0, nfield):
        """Gauge depths after every step, shape (nsteps, ngauges)."""
        self.set_inputs(Q0, nfield)
        gh = self.z.adj_forward()
        self.generation += 1
        return gh

    def backward(self, gbar):
        """Adjoint of the initial state (4, ny, nx) and of the Manning field
        (ny, nx) for the gauge-depth adjoint gbar of the last forward."""
        self.z.adj_backward(gbar)
        return self.z.adj_gradient()

    def freeze_limiter(self, on=True):
        """Record the limiter choices on the next forward and replay them on
        later ones (finite-difference checks through minmod switch points)."""
        self.z.adj_freeze_limiter(on)

    def close(self):
        self.z.close()


def frozen_dts(case, Q0, nfield, steps, contract=True, system="swe2d_nf"):
    """The dt sequence of a CFL-driven run of `steps` steps from (Q0, nfield)."""
    z = ZoomyAmr(deck_from_case(dict(case, max_step=steps)), system=system, dim=2, order=2, contract=contract)
    z.init_data()
    aux = z.get_state(field=1)
    aux[1] = nfield
    z.set_state(np.asarray(Q0, dtype=np.float64))
    z.set_state(aux, field=1)
    done, _ = z.run_steps(steps)
    dts, _ = z.dt_history()
    z.close()
    if done != steps or len(dts) != steps:
        raise RuntimeError(f"base run took {done} of {steps} steps")
    return dts


class GaugeSeries(torch.autograd.Function):
    """(Q0, n_field) -> gauge depths after every step, for a GaugeRun.  The
    backward must follow its own forward (the run keeps one set of
    checkpoints)."""

    @staticmethod
    def forward(ctx, Q0, nfield, run):
        gh = run.forward(Q0.detach().cpu().numpy(), nfield.detach().cpu().numpy())
        ctx.run = run
        ctx.generation = run.generation
        return torch.from_numpy(gh).to(device=Q0.device, dtype=Q0.dtype)

    @staticmethod
    def backward(ctx, gbar):
        run = ctx.run
        if run.generation != ctx.generation:
            raise RuntimeError("GaugeRun was run forward again before this backward; its checkpoints are gone")
        qb, nb = run.backward(gbar.detach().cpu().numpy())
        return (torch.from_numpy(qb).to(device=gbar.device, dtype=gbar.dtype),
                torch.from_numpy(nb).to(device=gbar.device, dtype=gbar.dtype), None)
