"""Pins of the FP64 metric oracle (oracle/metrics.py) and host checks of
msfdm_grid_metrics' argument handling.

* the written-out stencil is numpy.gradient(edge_order=2) bit for bit (the
  library routine the reference's generators call);
* it is exact on quadratic coordinate maps, so every metric term matches the
  closed form of such a map (a sign or operand swap in the inverse fails here);
* flat grid: J = h^2, xi_x = ze_z = 1/h, xi_z = ze_x = 0;
* vertical columns x = h xi: the reference's J = h z_ze, xi_x = 1/h, xi_z = +0;
* J (xi_x ze_z - xi_z ze_x) = 1 (inverse of the coordinate Jacobian);
* on the sinusoidal body-fitted grid the FD metrics converge to the closed-form
  metrics of the same geometry at second order (away from the interface row and
  the start of the PML taper, where the blend's second derivative jumps and a
  central difference is first order).
"""
import ctypes as C

import numpy as np
import pytest

import metrics
import msfdm
from msfdm import modelgen


@pytest.mark.parametrize("shape", [(7, 5), (40, 61), (3, 3)])
def test_stencil_is_numpy_gradient(shape):
    rng = np.random.default_rng(sum(shape))
    f = rng.standard_normal(shape) * 1e3
    for ax in (0, 1):
        assert np.array_equal(metrics.gradient2(f, ax), np.gradient(f, axis=ax, edge_order=2))


def test_exact_on_quadratic_map():
    nz, nx = 23, 31
    ze, xi = np.meshgrid(np.arange(nz, dtype=np.float64), np.arange(nx, dtype=np.float64), indexing="ij")
    x = 4.0 * xi + 0.7 * ze + 0.01 * xi * xi - 0.02 * xi * ze
    z = 0.3 * xi + 2.5 * ze + 0.015 * ze * ze + 0.01 * xi * ze
    x_xi, x_ze = 4.0 + 0.02 * xi - 0.02 * ze, 0.7 - 0.02 * xi
    z_xi, z_ze = 0.3 + 0.01 * ze, 2.5 + 0.03 * ze + 0.01 * xi
    J = x_xi * z_ze - x_ze * z_xi
    ref = {"J": J, "xi_x": z_ze / J, "xi_z": -x_ze / J, "ze_x": -z_xi / J, "ze_z": x_xi / J}
    m = metrics.grid_metrics(x, z)
    for k in ref:
        assert np.abs(m[k] - ref[k]).max() <= 1e-12 * np.abs(ref[k]).max(), k


def test_flat_grid():
    h = 2.5
    ze, xi = np.meshgrid(np.arange(9.0), np.arange(12.0), indexing="ij")
    m = metrics.grid_metrics(h * xi, h * (ze - 3.0))
    assert np.all(m["J"] == h * h)
    assert np.all(m["xi_x"] == 1.0 / h) and np.all(m["ze_z"] == 1.0 / h)
    assert np.all(m["xi_z"] == 0.0) and np.all(m["ze_x"] == 0.0)


def test_vertical_columns_reduce_to_reference_form():
    h = 5.0
    X, Z = modelgen.grid_coordinates(64, 48, h, 100.0, npd=8, seafloor=[(6.0, 150.0)])
    m = metrics.grid_metrics(X, Z)
    z_ze = np.gradient(Z, axis=0, edge_order=2)
    z_xi = np.gradient(Z, axis=1, edge_order=2)
    assert np.array_equal(m["J"], h * z_ze)
    assert np.array_equal(m["ze_z"], h / (h * z_ze))
    assert np.array_equal(m["ze_x"], -z_xi / (h * z_ze))
    assert np.all(np.abs(m["xi_x"] - 1.0 / h) <= 2 * np.spacing(1.0 / h))
    assert not np.any(np.signbit(m["xi_z"])) and np.all(m["xi_z"] == 0.0)
    assert np.array_equal(m["xi_x"].astype("<f4"), np.full(m["xi_x"].shape, 1.0 / h, dtype="<f4"))


def test_inverse_identity():
    rng = np.random.default_rng(3)
    ze, xi = np.meshgrid(np.arange(30.0), np.arange(40.0), indexing="ij")
    x = 3.0 * xi + 0.4 * np.sin(0.3 * ze + 0.2 * xi) + 0.1 * rng.standard_normal()
    z = 2.0 * ze + 0.5 * np.cos(0.25 * xi) * np.sin(0.1 * ze)
    m = metrics.grid_metrics(x, z)
    one = m["J"] * (m["xi_x"] * m["ze_z"] - m["xi_z"] * m["ze_x"])
    assert np.abs(one - 1.0).max() < 1e-13


def test_second_order_convergence_to_closed_form():
    # same physical geometry (L x D, seafloor 500 + 40 sin(2 pi x / 800)) at h and h/2
    errs = []
    for h in (10.0, 5.0):
        nx, nz = int(2000 / h) + 1, int(1000 / h) + 1
        npd = int(100 / h)
        ana = modelgen.build_model(nx, nz, h, 500.0, npd=npd, fpeak=10.0, dt=1e-3, nt=1, src_x=[nx // 2],
                                   seafloor=[(40.0, 800.0)])
        X, Z = modelgen.grid_coordinates(nx, nz, h, 500.0, npd=npd, seafloor=[(40.0, 800.0)])
        m = metrics.grid_metrics(X, Z)
        zeta0, b_pml = ana.nz_fluid - 1, nz - 4 - npd
        rows = np.array([k for k in range(5, nz - 5) if abs(k - zeta0) > 2 and abs(k - b_pml) > 2])
        sl = (rows[:, None], np.arange(4, nx - 4)[None, :])
        errs.append(max(np.abs(m[k][sl] - ana.arrays[k][sl].astype(np.float64)).max() / np.abs(m[k][sl]).max()
                        for k in ("J", "ze_x", "ze_z")))
    ratio = errs[0] / errs[1]
    assert errs[1] < 2e-3
    assert 3.0 < ratio < 5.0, (errs, ratio)


def test_grid_metrics_argument_checks():
    L = msfdm.lib()
    dp = C.POINTER(C.c_double)
    a = np.zeros((2, 5))
    outs = [np.zeros((2, 5)) for _ in range(5)]
    ptr = lambda v: v.ctypes.data_as(dp)
    assert L.msfdm_grid_metrics(ptr(a), ptr(a), 2, 5, 0, *[ptr(o) for o in outs]) == 1   # MSFDM_EINVAL
    assert "3 x 3" in msfdm.last_error()
    assert L.msfdm_grid_metrics(None, ptr(a), 2, 5, 0, *[ptr(o) for o in outs]) == 1
    with pytest.raises(ValueError):
        msfdm.grid_metrics(np.zeros((4, 4)), np.zeros((4, 5)))
