"""GPU: the benchmark configurations at their own shapes.

* Batched forward: the 64-shot family (nx 800-2000 x nz 400-1000 windows of the
  flat 5 m water-over-solid medium, seeded), one batched run of 200 steps:
  every shot's records and final fields equal the shot stepped alone, bit for
  bit; two shots against the FP64 oracle (records within 1e-4).
* Adjoint gradient: the 401x201 flat-seafloor model over 3000 steps, observed
  records from the true model, the initial model smoothed by a 100 m Gaussian:
  the (vp_f, rho_f, vp, vs, rho) gradient against central differences along 16
  seeded relative directions (eps 1e-3); on the 201x101 cut, the gradient
  against the FP64 oracle adjoint (1e-4 relative L2 per array).
* Multi-shot gradient: 32 shots on the 4001x1501 two-scale seafloor grid over a
  60-step record (the 8000-step record's per-step work, at a fraction of its
  length): the batched gradient equals the in-order sum of the 32 single-shot
  gradients bit for bit, and a repeated batched gradient gives the same bits; a
  401x421 window of the same geometry against the FP64 oracle.
"""
import numpy as np
import pytest

import msfdm
from msfdm import autograd as ag
from msfdm import fwi, modelgen
import pyoracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

RECORD_TOL = 1e-4
NB = 200


@pytest.fixture(scope="module")
def batch64():
    models = modelgen.batch_shots(n=64, nt=4000)
    print(f"[batch64] {len(models)} shots, nx {min(m.nx for m in models)}-{max(m.nx for m in models)}, "
          f"nz {min(m.nz for m in models)}-{max(m.nz for m in models)}, {NB} steps", flush=True)
    b = msfdm.Batch(models)
    b.reset()
    b.run(NB)
    b.sync()
    yield models, b
    b.close()


def test_batch64_family_matches_single_shots_bitwise(batch64):
    models, b = batch64
    for i, m in enumerate(models):
        s = msfdm.Solver(m)
        s.tile_rows = b.tile_rows
        s.reset(0)
        s.run(NB)
        sb = b.shot(i)
        for a, c in zip(sb.records(), s.records()):
            assert a.shape == (NB // 2, m.nx) and np.array_equal(a, c), i
        for name in ("FP", "FVZ", "SVX", "SVZ", "SSXZ", "SXEX", "SZVZ"):
            assert np.array_equal(sb.get(name), s.get(name)), (i, name)
        s.close()
        if i % 16 == 15:
            print(f"  shots 0..{i} equal", flush=True)


def test_batch64_shots_match_oracle(batch64):
    models, b = batch64
    small = int(np.argmin([m.nx * m.nz for m in models]))
    for i in sorted({small, 5}):
        m = models[i]
        print(f"[batch64 vs oracle] shot {i}: {m.nx}x{m.nz}", flush=True)
        o = pyoracle.Oracle(m)
        o.reset(0)
        o.run(NB)
        for name, a, c in zip(("solid_vx", "solid_vz", "fluid_p"), b.shot(i).records(), o.records()):
            e = rel_l2(a, c)
            print(f"  {name}: rel L2 {e:.2e}", flush=True)
            assert e <= RECORD_TOL, (i, name, e)


def _loss_at(solver, init, obs, vel):
    media = ag.media_from_velocity(*vel)
    solver.update_media(**{k: np.ascontiguousarray(v, dtype="<f4") for k, v in media.items()})
    solver.reset(0)
    solver.run(init.nt)
    solver.sync()
    return sum(0.5 * float(np.sum((a.astype(np.float64) - c.astype(np.float64)) ** 2))
               for a, c in zip(solver.records(), obs))


def test_gradient_flat401_3000_steps_central_differences():
    true = modelgen.preset("flat401", nt=3000)
    init = fwi.smoothed_model(true, 100.0, 5.0)
    print("[gradient] flat401, 3000 steps, 16 directions", flush=True)
    s = msfdm.Solver(true)
    s.reset(0)
    s.run(true.nt)
    s.sync()
    obs = s.records()
    solver = msfdm.Solver(init)
    loss, g = fwi.misfit_and_gradient(solver, init, obs, nsteps=init.nt)
    assert loss > 0
    vel0 = ag.velocity_from_model(init)
    names = ("vp_f", "rho_f", "vp", "vs", "rho")
    rng = np.random.default_rng(4)
    eps = 1e-3
    errs = []
    for k in range(16):
        d = [rng.standard_normal(v.shape) * v for v in vel0]
        fd = (_loss_at(solver, init, obs, [v + eps * dv for v, dv in zip(vel0, d)])
              - _loss_at(solver, init, obs, [v - eps * dv for v, dv in zip(vel0, d)])) / (2 * eps)
        an = sum(float(np.sum(g[n] * dv)) for n, dv in zip(names, d))
        errs.append(abs(an - fd) / abs(fd))
        print(f"  direction {k}: adjoint {an:.6e} central difference {fd:.6e} rel {errs[-1]:.2e}", flush=True)
    assert max(errs) <= 1e-3 and float(np.median(errs)) <= 2e-4, errs


def test_gradient_cut201_matches_fp64_oracle():
    cut_true = modelgen.preset("cut201")
    cut = fwi.smoothed_model(cut_true, 100.0, 5.0)
    print(f"[gradient] cut201 {cut.nx}x{cut.nz}, {cut.nt} steps vs FP64 oracle", flush=True)
    o = pyoracle.Oracle(cut_true)
    o.reset(0)
    o.run(cut.nt)
    obs = o.records()
    oi = pyoracle.Oracle(cut)
    oi.reset(0)
    oi.run(cut.nt)
    res = [a - c for a, c in zip(oi.records(), obs)]
    go = oi.gradient(0, cut.nt, res, every=24)
    gs = msfdm.Solver(cut).gradient(0, cut.nt, [r.astype(np.float32) for r in res])
    for n in msfdm._lib.GRADS:
        e = rel_l2(gs[n], go[n])
        print(f"  {n}: rel L2 {e:.2e}", flush=True)
        assert e <= 1e-4, (n, e)


def _residuals(m, nsteps, nshot, seed):
    rng = np.random.default_rng(seed)
    nrec = msfdm._lib.num_records(nsteps, m.dt)
    return [[rng.standard_normal((nrec, m.nx)).astype(np.float32) for _ in range(3)] for _ in range(nshot)]


def test_gradient_shots_twoscale4001_equal_sum_of_single_shots():
    nsteps, n = 60, 32
    nx = modelgen.preset("twoscale4001", nt=1).nx
    src = [int(round(x)) for x in np.linspace(0.1 * (nx - 1), 0.9 * (nx - 1), n)]
    m = modelgen.preset("twoscale4001", src_x=src, nt=nsteps)
    print(f"[shots] twoscale4001 {m.nx}x{m.nz}, {n} shots, {nsteps} steps", flush=True)
    s = msfdm.Solver(m)
    res = _residuals(m, nsteps, n, seed=8)
    every = 6
    g, wave = s.gradient_shots(list(range(n)), nsteps, res, every=every)
    print(f"  batched: waves of {wave} shots", flush=True)
    g2, _ = s.gradient_shots(list(range(n)), nsteps, res, every=every)
    ref = None
    for i in range(n):
        gi = s.gradient(i, nsteps, res[i], every=every)
        ref = gi if ref is None else {k: ref[k] + gi[k] for k in ref}
    for k in msfdm._lib.GRADS:
        print(f"  {k}: |g| {np.linalg.norm(g[k]):.4e} rel L2 vs in-order sum of single shots "
              f"{rel_l2(g[k], ref[k]):.2e}", flush=True)
        assert np.array_equal(g[k], ref[k]), k
        assert np.array_equal(g2[k], g[k]), k
        assert np.linalg.norm(ref[k]) > 0, k


def test_twoscale_window_gradient_matches_fp64_oracle():
    nsteps = 100
    m = modelgen.preset("twoscale4001", nx=401, nz=421, nt=nsteps, src_x=[120, 280])
    print(f"[shots] twoscale window {m.nx}x{m.nz}, 2 shots, {nsteps} steps vs FP64 oracle", flush=True)
    res = _residuals(m, nsteps, 2, seed=9)
    g, _ = msfdm.Solver(m).gradient_shots([0, 1], nsteps, res)
    o = pyoracle.Oracle(m)
    go = None
    for i in range(2):
        gi = o.gradient(i, nsteps, [r.astype(np.float64) for r in res[i]], every=10)
        go = gi if go is None else {k: go[k] + gi[k] for k in go}
    for k in msfdm._lib.GRADS:
        e = rel_l2(g[k], go[k])
        print(f"  {k}: rel L2 vs FP64 oracle {e:.2e}", flush=True)
        assert e <= 1e-4, (k, e)
