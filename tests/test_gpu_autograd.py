"""GPU: the torch.autograd binding.  The backward of the L2 misfit through
SeismogramFunction equals the C-ABI gradient with the residual as adjoint
source, and the velocity/density gradient matches central differences of the
misfit (GPU forward, eps = 1e-3 relative, 2 % tolerance for FP32 noise)."""
import numpy as np
import pytest
import torch

import msfdm
from msfdm import autograd as ag
from msfdm import fwi, modelgen

pytestmark = pytest.mark.gpu


def _small(**kw):
    base = dict(nx=72, nz=56, h=5.0, water_depth=100.0, npd=8, fpeak=25.0, dt=4e-4, nt=60,
                src_x=[36], src_height=10.0, seafloor=[(6.0, 150.0)], random_std=0.05, corr_len=20.0, seed=3)
    base.update(kw)
    return modelgen.build_model(**base)


def test_autograd_backward_equals_c_abi_gradient():
    true = _small()
    init = fwi.smoothed_model(true, 25.0, 5.0)
    s = msfdm.Solver(true)
    s.reset(0); s.run(true.nt); s.sync()
    obs = s.records()
    media = {n: torch.tensor(init.arrays[n].astype(np.float64), requires_grad=True) for n in msfdm._lib.GRADS}
    recs = ag.records(s, media, nsteps=true.nt)
    loss = sum(0.5 * torch.sum((r.double() - torch.from_numpy(o).double()) ** 2) for r, o in zip(recs, obs))
    loss.backward()
    s2 = msfdm.Solver(init)
    s2.reset(0); s2.run(init.nt); s2.sync()
    res = [a - b for a, b in zip(s2.records(), obs)]
    g = s2.gradient(0, init.nt, res)
    for n in msfdm._lib.GRADS:
        a, b = media[n].grad.numpy(), g[n]
        assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b), n
        assert np.linalg.norm(b) > 0, n


def test_velocity_gradient_matches_central_differences():
    true = _small()
    init = fwi.smoothed_model(true, 25.0, 5.0)
    s = msfdm.Solver(true)
    s.reset(0); s.run(true.nt); s.sync()
    obs = s.records()
    solver = msfdm.Solver(init)
    loss0, g = fwi.misfit_and_gradient(solver, init, obs, nsteps=init.nt)
    assert loss0 > 0
    vel0 = ag.velocity_from_model(init)
    names = ("vp_f", "rho_f", "vp", "vs", "rho")
    rng = np.random.default_rng(16)

    def loss_at(vel):
        media = ag.media_from_velocity(*vel)
        arrays = dict(init.arrays)
        for k, v in media.items():
            arrays[k] = np.ascontiguousarray(v, dtype="<f4")
        t = msfdm.Solver(modelgen.Model(**{**init.__dict__, "arrays": arrays}))
        t.reset(0); t.run(init.nt)
        return sum(0.5 * float(np.sum((a.astype(np.float64) - b.astype(np.float64)) ** 2))
                   for a, b in zip(t.records(), obs))

    eps = 1e-3
    for _ in range(3):
        d = [rng.standard_normal(v.shape) * v for v in vel0]
        lp = loss_at([v + eps * dv for v, dv in zip(vel0, d)])
        lm = loss_at([v - eps * dv for v, dv in zip(vel0, d)])
        fd = (lp - lm) / (2 * eps)
        an = sum(float(np.sum(g[n] * dv)) for n, dv in zip(names, d))
        print(f"  fd {fd:.6e} adjoint {an:.6e}", flush=True)
        assert an == pytest.approx(fd, rel=2e-2)


def test_two_shot_loss_through_one_solver_equals_sum_of_single_shot_gradients():
    """Two nodes of one graph on one solver: the second forward replaces the
    first one's checkpoints, so the first node's backward must redo its own
    forward (shot 0) before its reverse sweep."""
    m = _small(src_x=[24, 48])
    s = msfdm.Solver(m)
    rng = np.random.default_rng(12)
    nrec = msfdm._lib.num_records(m.nt, m.dt)
    res = [[rng.standard_normal((nrec, m.nx)) for _ in range(3)] for _ in range(2)]
    media = {n: torch.tensor(m.arrays[n].astype(np.float64), requires_grad=True) for n in msfdm._lib.GRADS}
    rec0 = ag.records(s, media, isrc=0, nsteps=m.nt)
    rec1 = ag.records(s, media, isrc=1, nsteps=m.nt)
    loss = sum(torch.sum(r.double() * torch.from_numpy(q)) for r, q in zip(rec0, res[0])) + \
        sum(torch.sum(r.double() * torch.from_numpy(q)) for r, q in zip(rec1, res[1]))
    loss.backward()
    s2 = msfdm.Solver(m)
    g0 = s2.gradient(0, m.nt, [q.astype(np.float32) for q in res[0]])
    g1 = s2.gradient(1, m.nt, [q.astype(np.float32) for q in res[1]])
    for n in msfdm._lib.GRADS:
        want = g0[n] + g1[n]
        got = media[n].grad.numpy()
        assert np.linalg.norm(got - want) <= 1e-6 * np.linalg.norm(want), n
        assert np.linalg.norm(g0[n]) > 0 and np.linalg.norm(g1[n]) > 0, n
