"""The oracle's discrete adjoint, pinned by exact transposition identities
(<A x, y> = <x, A^T y> for the RHS with its boundary condition, the interface
filter and a whole RK4 step in both F/B orders) and by central finite
differences of a record functional with respect to every model array."""
import numpy as np
import pytest

import pyoracle
from msfdm import modelgen


def _small(**kw):
    base = dict(nx=72, nz=56, h=5.0, water_depth=100.0, npd=8, fpeak=25.0, dt=4e-4, nt=60,
                src_x=[36], src_height=10.0, seafloor=[(6.0, 150.0)], random_std=0.05, corr_len=20.0, seed=3)
    base.update(kw)
    return modelgen.build_model(**base)


def _rand_state(o, rng):
    return {n: rng.standard_normal(o.shape(i)) for i, n in enumerate(pyoracle.FIELDS)}


def _dot(a, b):
    return sum(float(np.sum(a[n] * b[n])) for n in pyoracle.FIELDS)


@pytest.mark.parametrize("direction", [0, 1])
def test_rhs_with_bc_transpose_identity(direction):
    m = _small()
    o = pyoracle.Oracle(m)
    rng = np.random.default_rng(10 + direction)
    x, y = _rand_state(o, rng), _rand_state(o, rng)
    for n, a in x.items():
        o.set(n, a)
    o.rhs(direction, 0.0)
    Ax = {n: o.get_k(n) for n in pyoracle.FIELDS}
    for n, a in y.items():
        o.set_adj_in(n, a)
    o.adjoint_op(1, direction)                 # B^T
    qb = {n: o.get_adj_out(n) for n in pyoracle.FIELDS}
    for n, a in qb.items():
        o.set_adj_in(n, a)
    o.adjoint_op(0, direction)                 # L^T (forward state irrelevant for the transpose)
    ATy = {n: o.get_adj_out(n) for n in pyoracle.FIELDS}
    lhs, rhs = _dot(Ax, y), _dot(x, ATy)
    assert lhs == pytest.approx(rhs, rel=1e-11)


def test_filter_transpose_identity():
    m = _small()
    o = pyoracle.Oracle(m)
    rng = np.random.default_rng(4)
    x, y = _rand_state(o, rng), _rand_state(o, rng)
    for n, a in x.items():
        o.set(n, a)
    o.filter()
    Fx = {n: o.get(n) for n in pyoracle.FIELDS}
    for n, a in y.items():
        o.set_adj_in(n, a)
    o.adjoint_op(2, 0)
    FTy = {n: o.get_adj_out(n) for n in pyoracle.FIELDS}
    assert _dot(Fx, y) == pytest.approx(_dot(x, FTy), rel=1e-12)


@pytest.mark.parametrize("it", [6, 7])
def test_whole_step_transpose_identity(it):
    m = _small()
    o = pyoracle.Oracle(m)
    rng = np.random.default_rng(20 + it)
    x, y = _rand_state(o, rng), _rand_state(o, rng)
    for n, a in x.items():
        o.set(n, a)
    o.step_nosrc(it)
    Sx = {n: o.get(n) for n in pyoracle.FIELDS}
    for n, a in x.items():
        o.set(n, a)
    for n, a in y.items():
        o.set_adj_in(n, a)
    o.adjoint_op(3, it)
    STy = {n: o.get_adj_out(n) for n in pyoracle.FIELDS}
    assert _dot(Sx, y) == pytest.approx(_dot(x, STy), rel=1e-10)


def _functional(m, res, nsteps):
    o = pyoracle.Oracle(m)
    o.reset(0)
    o.run(nsteps)
    return sum(float(np.sum(r * a)) for r, a in zip(res, o.records()))


def test_gradient_matches_central_differences():
    m = _small()
    nsteps = 60
    o = pyoracle.Oracle(m)
    gap = o.gap
    nrec = (nsteps - 1) // gap + 1
    rng = np.random.default_rng(99)
    res = [rng.standard_normal((nrec, m.nx)) for _ in range(3)]
    g = o.gradient(0, nsteps, res, every=7)
    names = {"K": "K", "rho_f": "rho_f", "C11": "C11", "C13": "C13", "C33": "C33", "C55": "C55", "rho_s": "rho_s"}
    eps = 1e-3
    for gname, aname in names.items():
        base = m.arrays[aname].astype(np.float64)
        d = rng.standard_normal(base.shape) * base          # relative direction
        mp, mm = modelgen.Model(**{**m.__dict__}), modelgen.Model(**{**m.__dict__})
        mp.arrays = dict(m.arrays); mm.arrays = dict(m.arrays)
        ap = (base + eps * d).astype(np.float32)
        am = (base - eps * d).astype(np.float32)
        mp.arrays[aname] = ap; mm.arrays[aname] = am
        fd = (_functional(mp, res, nsteps) - _functional(mm, res, nsteps))
        an = float(np.sum(g[gname] * (ap.astype(np.float64) - am.astype(np.float64))))
        assert an == pytest.approx(fd, rel=2e-3), gname
