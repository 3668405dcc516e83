"""PyTorch autograd binding of the stepper: records as a differentiable function
of the medium.

``SeismogramFunction`` runs the HIP forward (checkpointed) in ``forward`` and the
discrete adjoint in ``backward`` through the C ABI (libmsfdm.so, ctypes).  Inputs
are the seven medium arrays in the model-file layout (fluid K, rho_f as
[nz_f][nx]; solid C11, C13, C33, C55, rho_s as [nz_s][nx]); outputs are the three
receiver records (solid-row vx, vz and fluid interface-row p, [nrec][nx]).
``media_from_velocity`` maps (vp, rho) of the water and (vp, vs, rho) of the
isotropic solid to those arrays, so gradients with respect to velocities and
density follow by the chain rule.
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import GRADS, Solver

_ORDER = GRADS   # K, rho_f, C11, C13, C33, C55, rho_s


class SeismogramFunction(torch.autograd.Function):
    """forward: checkpointed run of shot isrc in the given medium -> records;
    backward: the reverse sweep.  One solver may serve several nodes (shots,
    media) of a graph: each node keeps its shot, step count and medium, and its
    backward first redoes its own checkpointed forward when the solver's
    checkpoints or medium have been replaced since (Solver.adjoint_token)."""

    @staticmethod
    def forward(ctx, solver: Solver, isrc: int, nsteps: int, every: int, *media):
        arrays = {n: m.detach().cpu().numpy().astype(np.float32) for n, m in zip(_ORDER, media)}
        solver.update_media(**arrays)
        solver.adjoint_forward(isrc, nsteps, every)
        recs = solver.records()
        ctx.solver = solver
        ctx.args = (int(isrc), int(nsteps), int(every))
        ctx.arrays = arrays
        ctx.token = solver.adjoint_token
        ctx.dtypes = [m.dtype for m in media]
        ctx.devices = [m.device for m in media]
        dev = media[0].device
        return tuple(torch.from_numpy(r).to(dev) for r in recs)

    @staticmethod
    def backward(ctx, g_vx, g_vz, g_p):
        res = [g.detach().cpu().numpy().astype(np.float32) for g in (g_vx, g_vz, g_p)]
        solver = ctx.solver
        if solver.adjoint_token != ctx.token:   # another node (or call) replaced this node's forward
            solver.update_media(**ctx.arrays)
            solver.adjoint_forward(*ctx.args)
            ctx.token = solver.adjoint_token
        grads = solver.adjoint_backward(res)
        out = [torch.from_numpy(grads[n]).to(device=d, dtype=t) for n, d, t in zip(_ORDER, ctx.devices, ctx.dtypes)]
        return (None, None, None, None, *out)


def records(solver: Solver, media: dict, isrc: int = 0, nsteps: int | None = None, every: int = 0):
    """Differentiable receiver records for the medium tensors in ``media``."""
    nsteps = solver.nt if nsteps is None else nsteps
    return SeismogramFunction.apply(solver, isrc, nsteps, every, *[media[n] for n in _ORDER])


def media_from_velocity(vp_f, rho_f, vp, vs, rho):
    """Water (vp_f, rho_f) and isotropic solid (vp, vs, rho) -> the seven arrays
    (K = rho vp^2; C11 = C33 = rho vp^2, C13 = rho (vp^2 - 2 vs^2), C55 = rho vs^2)."""
    m = rho * vp * vp
    mu = rho * vs * vs
    return {"K": rho_f * vp_f * vp_f, "rho_f": rho_f, "C11": m, "C13": m - 2 * mu, "C33": m, "C55": mu,
            "rho_s": rho}


def velocity_from_model(model):
    """(vp_f, rho_f, vp, vs, rho) float64 arrays of a modelgen.Model."""
    a = {k: v.astype(np.float64) for k, v in model.arrays.items()}
    rho_f, rho = a["rho_f"], a["rho_s"]
    return (np.sqrt(a["K"] / rho_f), rho_f, np.sqrt(a["C11"] / rho), np.sqrt(a["C55"] / rho), rho)
