"""GPU discrete adjoint against the FP64 oracle's adjoint and against finite differences.

* L^T B^T and Filter^T applied on random inputs match the oracle's transposes
  (each field within 2e-5 of its max |value|);
* the FP32 transposition identity <B L x, y> = <x, L^T B^T y> holds to 1e-5;
* the checkpointed gradient of a record functional w.r.t. the seven model
  arrays matches the oracle's FP64 gradient (relative L2 per array <= 1e-4 on a
  curvilinear random-medium model, <= 2e-4 on a flat model);
* central differences of the GPU forward along random relative directions
  (eps = 1e-3) agree with <g, dm> within 2 % (FP32 forward noise).
"""
import numpy as np
import pytest

import msfdm
from msfdm import modelgen
import pyoracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _small(**kw):
    base = dict(nx=72, nz=56, h=5.0, water_depth=100.0, npd=8, fpeak=25.0, dt=4e-4, nt=60,
                src_x=[36], src_height=10.0, seafloor=[(6.0, 150.0)], random_std=0.05, corr_len=20.0, seed=3)
    base.update(kw)
    return modelgen.build_model(**base)


def _rand(shape_of, rng, names):
    return {n: rng.standard_normal(shape_of(msfdm.FID[n])).astype(np.float32) for n in names}


@pytest.mark.parametrize("direction", [0, 1])
def test_adjoint_operators_match_oracle(direction):
    print(f"[adj-op] dir={direction}", flush=True)
    m = _small()
    s = msfdm.Solver(m)
    o = pyoracle.Oracle(m)
    rng = np.random.default_rng(30 + direction)
    state = _rand(s.shape, rng, msfdm.FIELDS)
    y = _rand(s.shape, rng, msfdm.FIELDS)
    for n in msfdm.FIELDS:
        s.set(n, state[n]); o.set(n, state[n].astype(np.float64))
        s.adjoint_set_input(n, y[n]); o.set_adj_in(n, y[n].astype(np.float64))
    # L^T B^T
    o.adjoint_op(1, direction)
    qb = {n: o.get_adj_out(n) for n in msfdm.FIELDS}
    for n in msfdm.FIELDS:
        o.set_adj_in(n, qb[n])
    o.adjoint_op(0, direction)
    s.adjoint_apply(0, direction)
    for n in msfdm.FIELDS:
        ref, got = o.get_adj_out(n), s.adjoint_get_output(n).astype(np.float64)
        scale = np.abs(ref).max()
        if n in ("FXVX", "FXVZ", "FXEX", "FXEZ", "SXVX", "SXVZ", "SXEX", "SXEZ"):
            cols = list(range(3, 3 + m.npd)) + list(range(m.nx - 3 - m.npd, m.nx - 3))
            ref, got = ref[:, cols], got[:, cols]
        elif n.startswith("SZ"):
            z0 = m.nz_solid - 3 - m.npd
            ref, got = ref[z0:z0 + m.npd], got[z0:z0 + m.npd]
        assert np.abs(got - ref).max() <= 2e-5 * scale, n
    # Filter^T
    for n in msfdm.FIELDS:
        o.set_adj_in(n, y[n].astype(np.float64))
    o.adjoint_op(2, 0)
    s.adjoint_apply(1, 0)
    for n in ("FVX", "FVZ", "FP", "SVX", "SVZ", "SSX", "SSZ", "SSXZ"):
        ref, got = o.get_adj_out(n), s.adjoint_get_output(n).astype(np.float64)
        assert np.abs(got - ref).max() <= 2e-5 * np.abs(ref).max(), n


def test_fp32_transposition_identity():
    m = _small()
    s = msfdm.Solver(m)
    rng = np.random.default_rng(7)
    x = _rand(s.shape, rng, msfdm.FIELDS)
    y = _rand(s.shape, rng, msfdm.FIELDS)
    for n in msfdm.FIELDS:
        s.set(n, x[n])
        s.adjoint_set_input(n, y[n])
    s.eval_rhs(1, 0.0)
    s.adjoint_apply(0, 1)
    lhs = sum(float(np.sum(s.get_rhs(n).astype(np.float64) * _mask(s, m, n) * y[n])) for n in msfdm.FIELDS)
    rhs = sum(float(np.sum(_mask(s, m, n) * x[n].astype(np.float64) * s.adjoint_get_output(n))) for n in msfdm.FIELDS)
    assert lhs == pytest.approx(rhs, rel=1e-5)


def _mask(s, m, n):
    """Cells a field actually holds (PML memory only inside its strip)."""
    mk = np.ones(s.shape(msfdm.FID[n]))
    if n in ("FXVX", "FXVZ", "FXEX", "FXEZ", "SXVX", "SXVZ", "SXEX", "SXEZ"):
        mk[:] = 0
        mk[:, 3:3 + m.npd] = 1
        mk[:, m.nx - 3 - m.npd:m.nx - 3] = 1
    elif n.startswith("SZ"):
        mk[:] = 0
        z0 = m.nz_solid - 3 - m.npd
        mk[z0:z0 + m.npd] = 1
    return mk


@pytest.mark.parametrize("curvi", [True, False])
def test_gradient_matches_oracle(curvi):
    print(f"[adj-grad] curvi={curvi}", flush=True)
    m = _small() if curvi else _small(seafloor=(), random_std=0.0)
    nsteps = 60
    o = pyoracle.Oracle(m)
    nrec = (nsteps - 1) // o.gap + 1
    rng = np.random.default_rng(5)
    res = [rng.standard_normal((nrec, m.nx)).astype(np.float32) for _ in range(3)]
    go = o.gradient(0, nsteps, [r.astype(np.float64) for r in res], every=8)
    s = msfdm.Solver(m)
    gs = s.gradient(0, nsteps, res, every=8)
    tol = 1e-4 if curvi else 2e-4
    for n in msfdm._lib.GRADS:
        e = rel_l2(gs[n], go[n])
        print(f"  {n}: rel L2 {e:.2e}", flush=True)
        assert e <= tol, (n, e)
    # the checkpoint interval does not change the result beyond rounding
    gs2 = s.gradient(0, nsteps, res, every=0)
    for n in msfdm._lib.GRADS:
        assert rel_l2(gs2[n], gs[n]) <= 1e-6, n
    # and a repeated gradient gives the same bits
    gs3 = s.gradient(0, nsteps, res, every=8)
    for n in msfdm._lib.GRADS:
        assert np.array_equal(gs3[n], gs[n]), n
    # checkpoint buffers kept from the longer gradient serve a shorter one: same
    # bits as a solver that never ran the longer one
    short = 30
    rs = [r[:(short - 1) // o.gap + 1] for r in res]
    gshort = s.gradient(0, short, rs, every=8)
    fresh = msfdm.Solver(m)
    gfresh = fresh.gradient(0, short, rs, every=8)
    fresh.close()
    for n in msfdm._lib.GRADS:
        assert np.array_equal(gshort[n], gfresh[n]), n


def test_gradient_matches_gpu_central_differences():
    m = _small()
    nsteps = 60
    s = msfdm.Solver(m)
    s.reset(0); s.run(nsteps); s.sync()
    nrec = s.num_records()
    rng = np.random.default_rng(11)
    res = [rng.standard_normal((nrec, m.nx)).astype(np.float32) for _ in range(3)]
    g = s.gradient(0, nsteps, res)

    def J(model):
        t = msfdm.Solver(model)
        t.reset(0); t.run(nsteps)
        return sum(float(np.sum(r.astype(np.float64) * a)) for r, a in zip(res, t.records()))

    eps = 1e-3
    for aname in ("K", "C11", "C55", "rho_s", "rho_f"):
        base = m.arrays[aname].astype(np.float64)
        d = rng.standard_normal(base.shape) * base
        mp = modelgen.Model(**{**m.__dict__}); mm = modelgen.Model(**{**m.__dict__})
        mp.arrays = dict(m.arrays); mm.arrays = dict(m.arrays)
        mp.arrays[aname] = (base + eps * d).astype(np.float32)
        mm.arrays[aname] = (base - eps * d).astype(np.float32)
        fd = J(mp) - J(mm)
        an = float(np.sum(g[aname] * (mp.arrays[aname].astype(np.float64) - mm.arrays[aname].astype(np.float64))))
        print(f"  {aname}: fd {fd:.6e} adjoint {an:.6e}", flush=True)
        assert an == pytest.approx(fd, rel=2e-2), aname
