"""Pins of the solid-only CPML (boundary_type 1, include/cpml_kernels.cuh:8-233)
and Dirichlet (boundary_type 0, include/dirichlet_kernels.cuh:7-107) steppers of
the FP64 oracle, against what the mathematics fixes:

* the interior velocity-stress operator on a curvilinear grid reproduces the
  analytic time derivatives of a plane wave (rho dv/dt = div sigma,
  dsigma/dt = C : grad v) to the DRP truncation error, and the F/B pair to its
  centred error: a swapped stiffness, metric or sign fails by O(1);
* the CPML zeta profile of the top strip is the mirror image of the bottom one
  (src/elements.cu:832-869, lo and hi strips of one formula);
* the CPML absorbs what the Dirichlet box does not (all four sides damp);
* records: row 30 vz for CPML, none for Dirichlet (src/main.cu:199-209, 297-322).
"""
import dataclasses

import numpy as np
import pytest

import pyoracle
from msfdm import modelgen

F, B = 0, 1


def _curvi(boundary_type, nx=96, nz=90, h=5.0, amp=6.0, lam=160.0):
    """Solid-only model on the map x = h xi, z = h zeta + amp sin(2 pi x / lam):
    J = h z_zeta = h^2, ze_x = -z_xi / J, ze_z = h / J (closed form)."""
    m = modelgen.build_solid_model(nx=nx, nz=nz, h=h, npd=10, fpeak=20.0, dt=4e-4, nt=10, src_x=[nx // 2],
                                   src_z=[nz // 2], boundary_type=boundary_type,
                                   solid=(3100.0, 1750.0, 2150.0), random_std=0.05, corr_len=30.0, seed=3)
    x = h * np.arange(nx, dtype=np.float64)
    zxi = amp * 2 * np.pi * h / lam * np.cos(2 * np.pi * x / lam)
    J = np.full((nz, nx), h * h)
    a = dict(m.arrays)
    a["J"] = J.astype(np.float32)
    a["xi_x"] = np.full((nz, nx), 1.0 / h, np.float32)
    a["xi_z"] = np.zeros((nz, nx), np.float32)
    a["ze_x"] = np.broadcast_to((-zxi / (h * h))[None, :], (nz, nx)).astype(np.float32)
    a["ze_z"] = np.full((nz, nx), 1.0 / h, np.float32)
    X = np.broadcast_to(x[None, :], (nz, nx))
    Z = h * np.arange(nz, dtype=np.float64)[:, None] + amp * np.sin(2 * np.pi * X / lam)
    return dataclasses.replace(m, arrays=a), X, Z


@pytest.mark.parametrize("boundary_type", [0, 1])
def test_interior_operator_matches_plane_wave_derivatives(boundary_type):
    m, X, Z = _curvi(boundary_type)
    o = pyoracle.Oracle(m)
    kx, kz = 2 * np.pi / 120.0, 2 * np.pi / 150.0          # 24 and 30 points per wavelength
    ph = kx * X + kz * Z
    amp = {"SVX": 1.0, "SVZ": -0.7, "SSX": 2.0e6, "SSZ": -1.3e6, "SSXZ": 0.9e6}
    off = {"SVX": 0.1, "SVZ": 0.7, "SSX": 1.3, "SSZ": 2.1, "SSXZ": 2.9}
    f = {k: amp[k] * np.sin(ph + off[k]) for k in amp}
    df = {k: (amp[k] * kx * np.cos(ph + off[k]), amp[k] * kz * np.cos(ph + off[k])) for k in amp}   # d/dx, d/dz
    for k, v in f.items():
        o.set(k, v)
    C11, C13, C33, C55, rho = (m.arrays[k].astype(np.float64) for k in ("C11", "C13", "C33", "C55", "rho_s"))
    exact = {
        "SVX": (df["SSX"][0] + df["SSXZ"][1]) / rho,
        "SVZ": (df["SSXZ"][0] + df["SSZ"][1]) / rho,
        "SSX": C11 * df["SVX"][0] + C13 * df["SVZ"][1],
        "SSZ": C13 * df["SVX"][0] + C33 * df["SVZ"][1],
        "SSXZ": C55 * (df["SVX"][1] + df["SVZ"][0]),
    }
    ks = {}
    for d in (F, B):
        o.rhs(d, 0.0)
        ks[d] = {k: o.get_k(k) for k in exact}
    # interior rows/columns whose stencils do not reach the frozen bands or the PML
    r = slice(6 + m.npd, m.nz_solid - 7 - m.npd)
    c = slice(6 + m.npd, m.nx - 7 - m.npd)
    for k, e in exact.items():
        scale = np.abs(e[r, c]).max()
        for d in (F, B):
            assert np.abs(ks[d][k][r, c] - e[r, c]).max() < 1e-2 * scale, (k, d)
        avg = 0.5 * (ks[F][k] + ks[B][k])
        assert np.abs(avg[r, c] - e[r, c]).max() < 2e-3 * scale, k


def test_cpml_zeta_profile_mirrors_top_and_bottom():
    m = modelgen.preset("solid_cpml_tiny")
    o = pyoracle.Oracle(m)
    al, be, dp = o.pml(1)
    nz, npd = m.nz_solid, m.npd
    top = slice(3, 3 + npd)
    bot = slice(nz - 4, nz - 4 - npd, -1)
    for a in (al, be, dp):
        assert np.array_equal(a[top], a[bot])
    assert np.all(dp[top] > 0) and np.all(dp[3 + npd:nz - 3 - npd] == 0)
    assert dp[3] == dp.max() and be[3] == pytest.approx(m.beta0, rel=1e-6)


def _energy(o, m):
    rho, C11, C13, C33, C55, J = (m.arrays[k].astype(np.float64) for k in ("rho_s", "C11", "C13", "C33", "C55", "J"))
    vx, vz, sx, sz, sxz = (o.get(f) for f in ("SVX", "SVZ", "SSX", "SSZ", "SSXZ"))
    det = C11 * C33 - C13 ** 2
    es = 0.5 * ((C33 * sx * sx - 2 * C13 * sx * sz + C11 * sz * sz) / det + sxz ** 2 / C55)
    return float(np.sum((es + 0.5 * rho * (vx ** 2 + vz ** 2)) * J))


def test_cpml_absorbs_more_than_the_dirichlet_box():
    """Homogeneous 121 x 121 grid, central 25 Hz source: once the wave has
    crossed the boundaries the CPML grid keeps < 1e-6 of the peak energy and
    at least 30x less than the Dirichlet grid."""
    out = {}
    for bt in (0, 1):
        m = modelgen.build_solid_model(nx=121, nz=121, h=5.0, npd=15, fpeak=25.0, dt=4e-4, nt=1600,
                                       src_x=[60], src_z=[60], boundary_type=bt, solid=(3000.0, 1700.0, 2200.0))
        o = pyoracle.Oracle(m)
        o.reset(0)
        es = []
        for _ in range(16):
            o.run(100)
            es.append(_energy(o, m))
        out[bt] = es
        assert o.records()[0].shape[0] == (0 if bt == 0 else (1600 - 1) // o.gap + 1)
    peak = max(out[1])
    assert out[1][-1] < 1e-6 * peak
    assert out[1][-1] * 30 < out[0][-1]
    assert out[0][:3] == pytest.approx(out[1][:3], rel=1e-12)   # identical until the wave reaches the PML


def test_cpml_records_row_30():
    m = modelgen.preset("solid_cpml_tiny")
    o = pyoracle.Oracle(m)
    o.reset(1)
    assert o.gap == 3                                  # (int)(1e-3 / 4e-4 + 0.5)
    o.run(40)
    (rec,) = o.records()
    assert rec.shape == (14, m.nx)                     # steps 0, 3, .., 39
    assert np.array_equal(rec[-1], o.get("SVZ")[30])
    assert np.abs(rec).max() > 0
