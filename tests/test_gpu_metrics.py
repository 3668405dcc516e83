"""FP64 GPU metric kernel (msfdm_grid_metrics) against the FP64 oracle restatement
of the reference's numpy.gradient metric block.

The kernel evaluates the same explicitly rounded FP64 operations in numpy's
order, so every output is within 1 ulp of the oracle (in practice identical),
on the body-fitted seafloor grid (vertical columns) and on a sheared grid
whose columns are not vertical (all five metric terms non-trivial).  The
float32 arrays written as model files are then identical.
"""
import numpy as np
import pytest

import metrics
import msfdm
from msfdm import modelgen

pytestmark = pytest.mark.gpu

KEYS = ("J", "xi_x", "xi_z", "ze_x", "ze_z")


def _grids():
    X, Z = modelgen.grid_coordinates(2001, 1001, 2.5, 500.0, npd=30, seafloor=[(40.0, 800.0)])
    yield "seafloor2001", X, Z
    ze, xi = np.meshgrid(np.arange(257.0), np.arange(333.0), indexing="ij")
    Xs = 5.0 * xi + 0.8 * ze + 3.0 * np.sin(2 * np.pi * ze / 97.0)
    Zs = 5.0 * ze + 1.5 * xi * 0.01 * ze + 7.0 * np.sin(2 * np.pi * xi / 61.0)
    yield "sheared", Xs, Zs


@pytest.mark.parametrize("case", ["seafloor2001", "sheared"])
def test_gpu_metrics_match_oracle(case):
    name, X, Z = next(g for g in _grids() if g[0] == case)
    print(f"[metrics] {name} {X.shape}", flush=True)
    got = msfdm.grid_metrics(X, Z)
    ref = metrics.grid_metrics(X, Z)
    for k in KEYS:
        np.testing.assert_array_max_ulp(got[k], ref[k], maxulp=1)
        assert np.array_equal(got[k].astype("<f4"), ref[k].astype("<f4")), k
        same = np.mean(got[k] == ref[k])
        print(f"  {k}: identical fraction {same:.6f}", flush=True)


def test_model_with_gpu_metrics_runs():
    m = modelgen.build_model(120, 90, 5.0, 150.0, npd=10, fpeak=20.0, dt=4e-4, nt=40, src_x=[60],
                             seafloor=[(8.0, 200.0)], metrics="gpu")
    a = modelgen.build_model(120, 90, 5.0, 150.0, npd=10, fpeak=20.0, dt=4e-4, nt=40, src_x=[60],
                             seafloor=[(8.0, 200.0)])
    # FD truncation of the reference's stencil at 40 points per seafloor wavelength:
    # (2 pi h / lambda)^2 / 6 = 4e-3 inside, first order at the blend's kinks
    for k in KEYS:
        rel = np.abs(m.arrays[k].astype(np.float64) - a.arrays[k]).max() / max(np.abs(a.arrays[k]).max(), 1e-30)
        assert rel < 2e-2, (k, rel)
    s = msfdm.Solver(m)
    s.reset(0)
    s.run(40)
    vx, vz, p = s.records()
    assert np.isfinite(p).all() and np.abs(p).max() > 0
