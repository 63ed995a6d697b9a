"""torch.autograd.Function over the hand-written HIP adjoint (no reference
counterpart: voftrial1.cu has no differentiation).

``centroid_height(F0, u0, v0, w0, p0, rho_l, rho_g, sigma, spec)`` returns the
liquid-centroid height J = sum(F z) / sum(F) after ``spec.nsteps`` steps of the
VOF timestep from the given state.  Its forward runs the solver with
checkpoints every ``spec.checkpoint_every`` steps (vof_adj_forward); its
backward runs the reverse sweep (vof_adj_backward) and returns dJ/d(state)
and dJ/d(rho_l, rho_g, sigma).  sigma does not enter the reference physics, so
its gradient is zero.  State tensors use the reference layout
(nz+2, ny+2, nx+2); they may live on any device and are copied through the
host.  There is no CPU fallback: the computation runs in lib/libvof_hip.so.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

import vof


@dataclass
class AdjointSpec:
    params: vof.VofParams
    first_istep: int = 1
    nsteps: int = 1
    checkpoint_every: int = 20
    last: dict = field(default_factory=dict)  # timings and scalar gradients of the last call


def _params_with(p: vof.VofParams, **kw) -> vof.VofParams:
    q = vof.VofParams.from_buffer_copy(p)
    for k, v in kw.items():
        setattr(q, k, float(v))
    return q


class CentroidHeight(torch.autograd.Function):
    @staticmethod
    def forward(ctx, F0, u0, v0, w0, p0, rho_l, rho_g, sigma, spec: AdjointSpec):
        prm = _params_with(spec.params, rho_l=rho_l.item(), rho_g=rho_g.item(), sigma=sigma.item())
        s = vof.Solver(prm)
        for name, t in zip("FUVWP", (F0, u0, v0, w0, p0)):
            s.upload(name, t.detach().cpu().numpy())
        J = s.adj_forward(spec.first_istep, spec.nsteps, spec.checkpoint_every)
        ctx.solver = s
        ctx.spec = spec
        ctx.devices = [t.device for t in (F0, u0, v0, w0, p0)]
        return torch.tensor(J, dtype=torch.float64, device=F0.device)

    @staticmethod
    def backward(ctx, gJ):
        s = ctx.solver
        sc = s.adj_backward(float(gJ))
        grads = [torch.from_numpy(s.adj_download(n)).to(d) for n, d in zip("FUVWP", ctx.devices)]
        fwd_ms, bwd_ms = s.adj_times()
        ctx.spec.last = {"scalars": sc, "fwd_ms": fwd_ms, "bwd_ms": bwd_ms}
        s.close()
        ctx.solver = None
        dev = ctx.devices[0]
        sc_t = lambda x: torch.tensor(x, dtype=torch.float64, device=dev)
        return (*grads, sc_t(sc["rho_l"]), sc_t(sc["rho_g"]), sc_t(sc["sigma"]), None)


def centroid_height(F0, u0, v0, w0, p0, rho_l, rho_g, sigma, spec: AdjointSpec) -> torch.Tensor:
    return CentroidHeight.apply(F0, u0, v0, w0, p0, rho_l, rho_g, sigma, spec)


def smoothed_noise(shape, std: float, width: float, seed: int) -> np.ndarray:
    """Seeded Gaussian noise smoothed by a Gaussian of `width` cells (reflecting
    edges), rescaled to standard deviation `std`."""
    from scipy.ndimage import gaussian_filter

    a = gaussian_filter(np.random.default_rng(seed).normal(0.0, 1.0, shape), width, mode="reflect")
    return a * (std / a.std())
