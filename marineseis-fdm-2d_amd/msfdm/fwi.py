"""L2-misfit gradients of the receiver records with respect to vp, vs and rho.

The workload of the adjoint: observed records come from a true model, the
starting model is that model smoothed by a Gaussian on the global grid (water
rows carry vs = 0), and the gradient of 0.5 * ||rec - obs||^2 with respect to the
water (vp, rho) and solid (vp, vs, rho) grids comes from the discrete adjoint
through ``msfdm.autograd``.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from . import autograd as ag
from ._lib import Solver


def gaussian_smooth(a: np.ndarray, sigma: float) -> np.ndarray:
    """Separable Gaussian smoothing (std ``sigma`` cells, edge-replicated)."""
    if sigma <= 0:
        return a.copy()
    r = int(np.ceil(4 * sigma))
    x = np.arange(-r, r + 1, dtype=np.float64)
    w = np.exp(-0.5 * (x / sigma) ** 2)
    w /= w.sum()
    out = a.astype(np.float64)
    for axis in (0, 1):
        pad = [(0, 0), (0, 0)]
        pad[axis] = (r, r)
        p = np.pad(out, pad, mode="edge")
        out = sum(w[k] * np.take(p, np.arange(k, k + a.shape[axis]), axis=axis) for k in range(2 * r + 1))
    return out


def smoothed_model(model, sigma_m: float, h: float):
    """Copy of ``model`` whose vp, vs, rho are smoothed on the global grid."""
    vp_f, rho_f, vp, vs, rho = ag.velocity_from_model(model)
    nzf = model.nz_fluid
    def glob(fluid, solid):
        g = np.zeros((model.nz, model.nx))
        g[:nzf] = fluid
        g[nzf - 1:] = solid
        return g
    sig = sigma_m / h
    gvp = gaussian_smooth(glob(vp_f, vp), sig)
    gvs = gaussian_smooth(glob(np.zeros_like(vp_f), vs), sig)
    grho = gaussian_smooth(glob(rho_f, rho), sig)
    vp_f2, rho_f2 = gvp[:nzf], grho[:nzf]
    vp2, vs2, rho2 = gvp[nzf - 1:], gvs[nzf - 1:], grho[nzf - 1:]
    media = ag.media_from_velocity(vp_f2, rho_f2, vp2, vs2, rho2)
    arrays = dict(model.arrays)
    for k, v in media.items():
        arrays[k] = np.ascontiguousarray(v, dtype="<f4")
    return dataclasses.replace(model, arrays=arrays)


def misfit_and_gradient(solver: Solver, model, obs, isrc: int = 0, nsteps: int | None = None, every: int = 0):
    """0.5 ||rec - obs||^2 and its gradient w.r.t. (vp_f, rho_f, vp, vs, rho) at ``model``."""
    vel = [torch.tensor(v, dtype=torch.float64, requires_grad=True) for v in ag.velocity_from_model(model)]
    media = ag.media_from_velocity(*vel)
    recs = ag.records(solver, media, isrc=isrc, nsteps=nsteps, every=every)
    loss = sum(0.5 * torch.sum((r.to(torch.float64) - torch.as_tensor(o, dtype=torch.float64)) ** 2)
               for r, o in zip(recs, obs))
    loss.backward()
    names = ("vp_f", "rho_f", "vp", "vs", "rho")
    return float(loss.detach()), {n: v.grad.numpy() for n, v in zip(names, vel)}
