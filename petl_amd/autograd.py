"""torch.autograd bindings of the planogram projector.

`project(server, f)` is differentiable in `f`: its forward is the HIP forward projector (the fast
separable kernels where the geometry allows), its backward the matched back projector, i.e. the exact
transpose of that forward, so gradients of any loss of the planograms are exact up to fp32 rounding.
Both run on torch's current HIP stream on device-resident tensors; nothing is copied to the host, and a
batch costs one forward and one back projector launch per planogram.

Shapes: `f` is (nz, ny, nx) or a batch (B, nz, ny, nx) of float32 volumes on the GPU; the result is a
tuple with one tensor per planogram of shape (N_v1, N_v0, N_u1, N_u0) or (B, N_v1, N_v0, N_u1, N_u0).

There is no counterpart in the reference; the loss helpers below restate the Poisson log-likelihood
that its MLEM maximises (src/petl_server.py:1534-1606).
"""
import torch

from .petl_server import PETLserver

__all__ = ["PlanogramProjection", "project", "backproject_matched", "poisson_nll"]


def _shapes(server):
    return [tuple(int(x) for x in server.get_planogramSize(i)) for i in range(server.get_numPlanograms())]


def _vol_shape(server):
    return (server.numZ(), server.numY(), server.numX())


def _check_volume(server, f):
    if not isinstance(f, torch.Tensor) or f.device.type != "cuda":
        raise TypeError("f must be a torch tensor on the GPU")
    if f.dtype != torch.float32:
        raise TypeError("f must be float32")
    if tuple(f.shape[-3:]) != _vol_shape(server) or f.dim() not in (3, 4):
        raise ValueError(f"f must have shape {_vol_shape(server)} or (B, *{_vol_shape(server)}), got {tuple(f.shape)}")


def _stream(device):
    return torch.cuda.current_stream(device).cuda_stream


def _forward(server, f):
    """f: (B, nz, ny, nx) contiguous -> list of (B, *shape_i) planogram tensors, one launch per planogram
    for the whole batch on the fast kernels."""
    shapes = _shapes(server)
    g = [torch.empty((f.shape[0],) + sh, dtype=torch.float32, device=f.device) for sh in shapes]
    server.project_gpu_batch(g, f, stream=_stream(f.device), sync=False)
    return g


def backproject_matched(server, g, batch):
    """Sum over planograms of A_i^T g_i with the matched back projector, per batch element (the mode is an
    argument of the call; the server's own mode is left alone).  g: list of (B, *shape_i) device tensors.
    Returns (B, nz, ny, nx)."""
    dev = g[0].device
    out = torch.empty((batch,) + _vol_shape(server), dtype=torch.float32, device=dev)
    server.backproject_gpu_batch(g, out, mode="matched", stream=_stream(dev), sync=False)
    return out


class PlanogramProjection(torch.autograd.Function):
    """y_i = A_i f for every planogram i of `server`; backward: df = sum_i A_i^T dy_i (matched)."""

    @staticmethod
    def forward(ctx, server, f):
        ctx.server = server
        ctx.batch = f.shape[0]
        return tuple(_forward(server, f))

    @staticmethod
    def backward(ctx, *grads):
        server = ctx.server
        shapes = _shapes(server)
        dev = next(gr.device for gr in grads if gr is not None)
        gl = []
        for sh, gr in zip(shapes, grads):
            if gr is None:
                gr = torch.zeros((ctx.batch,) + sh, dtype=torch.float32, device=dev)
            gl.append(gr.contiguous())
        return None, backproject_matched(server, gl, ctx.batch)


def project(server, f):
    """Differentiable forward projection.  Returns a tuple of planogram tensors (see module doc)."""
    if not isinstance(server, PETLserver):
        raise TypeError("server must be a PETLserver")
    _check_volume(server, f)
    single = f.dim() == 3
    fb = (f.unsqueeze(0) if single else f).contiguous()
    out = PlanogramProjection.apply(server, fb)
    return tuple(o.squeeze(0) for o in out) if single else out


def poisson_nll(q, y, eps=0.0):
    """sum(q - y log q) over bins with q > eps (the negative Poisson log-likelihood up to a constant),
    accumulated in float64.  q, y: matching tensors."""
    qd = q.double()
    yd = y.double()
    pos = qd > eps
    safe = torch.where(pos, qd, torch.ones_like(qd))
    return torch.where(pos, qd - yd * torch.log(safe), torch.zeros_like(qd)).sum()
