"""Gradients through the frozen-dt SSP-RK2 advance of the per-cell-Manning
shallow-water system: the library's adjoint (include/zamr.h, zamr_adj_*)
behind a torch.autograd.Function.

    run = GaugeRun(case, gauges, dts)            # one library context
    Q0 = torch.stack([b, (eta0 - b).clamp(min=0), 0*b, 0*b])
    G = GaugeSeries.apply(Q0, n_field, run)      # (nsteps, ngauges) depths
    loss = 0.5 * ((G - G_truth) ** 2).sum()
    loss.backward()                               # b.grad, n_field.grad

The library does the work on the GPU.  CUDA tensors stay on the device: the
initial state and the Manning field are copied in place into the library's
state arrays (zamr_device_ptr), the gauge series and the gradients are
written into torch-allocated device tensors, and everything is queued on the
caller's current torch stream (zamr_set_stream), with no host round trip and
no host wait.  CPU tensors go through host memory."""
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
        self._stream = None  # the stream handed to the library (None: its own)
        self._device = None
        self._boxes = None
        self._aux_views = None

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

    # ---- zero-copy path: CUDA tensors on the caller's current stream ----
    def _bind_stream(self, device):
        st = torch.cuda.current_stream(device).cuda_stream
        if st != self._stream:
            self.z.set_stream(st)
            self._stream = st

    def _valid_view(self, loc, box, row0, nrows):
        """Strided torch view of rows [row0, row0 + nrows) of the valid cells
        of a box's library array located by zamr_device_ptr, shape (nrows,
        ny, nx) of the box (lo0, lo1, hi0, hi1)."""
        ptr, rs, cs, ngrow, _ = loc
        lo0, lo1, hi0, hi1 = box
        start = ptr + 8 * (row0 * cs + ngrow * rs + ngrow)
        return _device_array(start, (nrows, hi1 - lo1 + 1, hi0 - lo0 + 1), (8 * cs, 8 * rs, 8), self._device)

    def forward_device(self, Q0, nfield):
        """Gauge depths (nsteps, ngauges) as a device tensor; Q0 (n_q, ny, nx)
        and nfield (ny, nx) are CUDA float64 tensors."""
        self._device = Q0.device
        self._bind_stream(Q0.device)
        if self._boxes is None:
            # the boxes' Qaux arrays stay where they are for the context's
            # life; only their stored input row (Manning) is written
            self._boxes = self.z.level_boxes(0)
            self._aux_views = [self._valid_view(self.z.device_ptr(field=1, box=k), bx, 1, 1)[0]
                               for k, bx in enumerate(self._boxes)]
        for k, (lo0, lo1, hi0, hi1) in enumerate(self._boxes):
            q = self._valid_view(self.z.device_ptr(field=0, box=k), self._boxes[k], 0, self.z.nq)
            q.copy_(Q0[:, lo1:hi1 + 1, lo0:hi0 + 1])
            self._aux_views[k].copy_(nfield[lo1:hi1 + 1, lo0:hi0 + 1])
        gh = torch.empty((len(self.dts), len(self.gauges[0])), dtype=torch.float64, device=Q0.device)
        self.z.adj_forward_device(gh.data_ptr())
        self.generation += 1
        return gh

    def backward_device(self, gbar):
        """(adjoint of Q0, adjoint of the Manning field) as device tensors for
        the device gauge adjoint gbar (nsteps, ngauges)."""
        self._bind_stream(gbar.device)
        gbar = gbar.to(dtype=torch.float64).contiguous()
        self.z.adj_backward_device(gbar.data_ptr())
        nx, ny, _ = self.z.shape()
        qb = torch.empty((self.z.nq, ny, nx), dtype=torch.float64, device=gbar.device)
        nb = torch.empty((ny, nx), dtype=torch.float64, device=gbar.device)
        self.z.adj_gradient_device(qb.data_ptr(), nb.data_ptr())
        return qb, nb

    def close(self):
        self.z.close()


class _CudaArray:
    """A device address with a shape and byte strides, exported through the
    CUDA array interface so that torch.as_tensor views it without a copy."""

    def __init__(self, ptr, shape, strides):
        self.__cuda_array_interface__ = {"shape": tuple(shape), "strides": tuple(strides), "typestr": "<f8",
                                         "data": (int(ptr), False), "version": 2}


def _device_array(ptr, shape, strides, device):
    t = torch.as_tensor(_CudaArray(ptr, shape, strides), device=device)
    if t.data_ptr() != ptr or t.dtype != torch.float64:
        raise RuntimeError("torch did not view the library's device array in place")
    return t


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
        ctx.run = run
        if Q0.is_cuda:
            if not nfield.is_cuda or Q0.device != nfield.device:
                raise ValueError("Q0 and nfield must be on the same device")
            gh = run.forward_device(Q0.detach().to(torch.float64), nfield.detach().to(torch.float64))
            ctx.generation = run.generation
            return gh.to(dtype=Q0.dtype)
        gh = run.forward(Q0.detach().numpy(), nfield.detach().numpy())
        ctx.generation = run.generation
        return torch.from_numpy(gh).to(dtype=Q0.dtype)

    @staticmethod
    def backward(ctx, gbar):
        run = ctx.run
        if run.generation != ctx.generation:
            raise RuntimeError("GaugeRun was run forward again before this backward; its checkpoints are gone")
        if gbar.is_cuda:
            qb, nb = run.backward_device(gbar.detach())
            return qb.to(dtype=gbar.dtype), nb.to(dtype=gbar.dtype), None
        qb, nb = run.backward(gbar.detach().numpy())
        return torch.from_numpy(qb).to(dtype=gbar.dtype), torch.from_numpy(nb).to(dtype=gbar.dtype), None
