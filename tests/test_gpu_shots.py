"""Multi-shot gradient (msfdm_gradient_shots): all shots of a wave stepped together.

* equals the in-order sum of the single-shot gradients (msfdm_gradient) bit for
  bit at the same checkpoint interval (every adjoint output cell has one writer
  that sums its terms in a fixed order, and both paths fold the same segment
  sums), with the memory budget forcing waves of 2 + 1 shots and with all
  shots in one wave;
* the adjoint products as one launch or as three give the same bits, single
  shot and multi-shot;
* matches the sum of the FP64 oracle's per-shot gradients to 1e-4 relative L2;
* records of 1-3 steps and an interval longer than the record: single-shot =
  multi-shot bit for bit, both against the oracle;
* it releases the single-shot checkpoints (a pending adjoint forward must be
  repeated; the next single-shot gradient gives the same bits);
* a budget below one shot is refused with MSFDM_EDEVICE, a bad shot index
  with MSFDM_EINVAL.
"""
import numpy as np
import pytest

import msfdm
from msfdm import modelgen
import pyoracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

NSTEPS = 60


def _model():
    return modelgen.build_model(nx=72, nz=56, h=5.0, water_depth=100.0, npd=8, fpeak=25.0, dt=4e-4, nt=NSTEPS,
                                src_x=[20, 36, 52], src_height=10.0, seafloor=[(6.0, 150.0)], random_std=0.05,
                                corr_len=20.0, seed=3)


def _residuals(m, nrec, seed=21):
    rng = np.random.default_rng(seed)
    return [[rng.standard_normal((nrec, m.nx)).astype(np.float32) for _ in range(3)] for _ in m.src_x]


def test_shots_equal_sum_of_single_shot_gradients():
    m = _model()
    s = msfdm.Solver(m)
    s.reset(0); s.run(NSTEPS); s.sync()
    nrec = s.num_records()
    res = _residuals(m, nrec)
    single = [s.gradient(i, NSTEPS, res[i], every=8) for i in range(3)]
    ref = {n: single[0][n] + single[1][n] + single[2][n] for n in msfdm._lib.GRADS}
    before = [a.copy() for a in s.records()]
    per_shot, every = s.gradient_shots_plan(NSTEPS, 8)
    assert every == 8 and per_shot > 0
    for budget, want in ((64 * 2 ** 20 + 2.9 * per_shot, 2), (0.0, 3)):
        print(f"[shots] budget {budget:.3e} B", flush=True)
        g, wave = s.gradient_shots([0, 1, 2], NSTEPS, res, every=8, mem_bytes=budget)
        assert wave == want
        for n in msfdm._lib.GRADS:
            assert np.array_equal(g[n], ref[n]), n
    # the solver's own records (of its last forward run) are untouched
    for a, b in zip(before, s.records()):
        assert np.array_equal(a, b)


def test_shots_match_oracle():
    m = _model()
    o = pyoracle.Oracle(m)
    nrec = (NSTEPS - 1) // o.gap + 1
    res = _residuals(m, nrec, seed=5)
    s = msfdm.Solver(m)
    g, wave = s.gradient_shots([2, 0, 1], NSTEPS, [res[2], res[0], res[1]])
    assert wave == 3
    go = None
    for i in (2, 0, 1):
        gi = o.gradient(i, NSTEPS, [r.astype(np.float64) for r in res[i]], every=10)
        go = gi if go is None else {n: go[n] + gi[n] for n in go}
    for n in msfdm._lib.GRADS:
        e = rel_l2(g[n], go[n])
        print(f"  {n}: rel L2 vs FP64 oracle {e:.2e}", flush=True)
        assert e <= 1e-4, (n, e)


def test_shots_errors():
    m = _model()
    s = msfdm.Solver(m)
    res = _residuals(m, msfdm._lib.num_records(NSTEPS, m.dt))[:1]   # not read: both calls fail before the run
    with pytest.raises(msfdm.MsfdmError) as e:
        s.gradient_shots([0], NSTEPS, res, mem_bytes=1e6)
    assert e.value.code == 4   # MSFDM_EDEVICE
    with pytest.raises(msfdm.MsfdmError) as e:
        s.gradient_shots([7], NSTEPS, res)
    assert e.value.code == 1   # MSFDM_EINVAL
    # residuals of the wrong record count are refused before any device call
    with pytest.raises(ValueError):
        s.gradient_shots([0], NSTEPS, _residuals(m, 30)[:1])
    with pytest.raises(ValueError):
        s.gradient(0, NSTEPS, _residuals(m, 30)[0])


def test_adjoint_launch_modes_agree_bitwise():
    print("[shots] adjoint products: one launch against three", flush=True)
    m = _model()
    s = msfdm.Solver(m)
    s.reset(0); s.run(NSTEPS); s.sync()
    res = _residuals(m, s.num_records())
    assert s.adjoint_launch == -1
    out = {}
    for mode in (0, 1):
        s.adjoint_launch = mode
        assert s.adjoint_launch == mode
        g1 = s.gradient(1, NSTEPS, res[1], every=8)
        gs, _ = s.gradient_shots([0, 1, 2], NSTEPS, res, every=8)
        out[mode] = (g1, gs)
    for k in range(2):
        for n in msfdm._lib.GRADS:
            assert np.array_equal(out[0][k][n], out[1][k][n]), (k, n)
            assert np.any(out[0][k][n] != 0), (k, n)
    with pytest.raises(msfdm.MsfdmError):
        s.adjoint_launch = 2


def test_shots_release_the_single_shot_buffers():
    print("[shots] the multi-shot gradient releases the single-shot checkpoints", flush=True)
    m = _model()
    s = msfdm.Solver(m)
    s.reset(0); s.run(NSTEPS); s.sync()
    res = _residuals(m, s.num_records())
    g1 = s.gradient(1, NSTEPS, res[1], every=8)
    s.adjoint_forward(1, NSTEPS, every=8)
    s.gradient_shots([0], NSTEPS, res[:1], every=8)
    with pytest.raises(msfdm.MsfdmError):   # the pending forward's checkpoints are gone
        s.adjoint_backward(res[1])
    g2 = s.gradient(1, NSTEPS, res[1], every=8)   # allocates them again
    for n in msfdm._lib.GRADS:
        assert np.array_equal(g1[n], g2[n]), n


@pytest.mark.parametrize("nsteps,every", [(1, 2), (2, 2), (3, 2), (12, 20)])
def test_short_records_and_long_intervals(nsteps, every):
    """Records of 1-3 steps (a single, partial segment) and an interval longer
    than the record: the segment recomputation writes each step's U^{n+1}
    into the next slot, the last step's into the state buffer.  Single-shot and
    multi-shot gradients agree bit for bit and match the FP64 oracle."""
    print(f"[shots] nsteps={nsteps} every={every}", flush=True)
    m = _model()
    o = pyoracle.Oracle(m)
    nrec = (nsteps - 1) // o.gap + 1
    res = _residuals(m, nrec, seed=9)
    s = msfdm.Solver(m)
    g1 = s.gradient(1, nsteps, res[1], every=every)
    gs, _ = s.gradient_shots([1], nsteps, res[1:2], every=every)
    go = o.gradient(1, nsteps, [r.astype(np.float64) for r in res[1]], every=every)
    for n in msfdm._lib.GRADS:
        assert np.array_equal(g1[n], gs[n]), n
        assert np.all(np.isfinite(g1[n])), n
        if np.linalg.norm(go[n]) > 0:
            e = rel_l2(g1[n], go[n])
            print(f"  {n}: rel L2 {e:.2e}", flush=True)
            assert e <= 1e-4, (n, e)
        else:
            assert not np.any(g1[n]), n
