"""Pins of the solid-only FP64 oracle (grid.calc_type = 0, traction-free top).

The reference holds no tests or outputs for PmlFs::SolidOnly either, so the
oracle's restatement (include/pml_fs_kernels.cuh:8-307, src/pml_fs_kernels.cu:1-381,
src/elements.cu:907-977) is checked against what the mathematics fixes:

* on a flat grid the ALM coefficients reduce to the textbook free-surface
  elimination dvz/dz = -(C13/C33) dvx/dx, dvx/dz = -dvz/dx;
* on curvilinear grids the free-surface stress rates carry zero traction
  n . dsigma/dt = 0 along the surface normal n ~ (zeta_x, zeta_z), PML columns
  included (an exchanged, mis-signed or mis-indexed ALM term breaks this);
* the three conservative rows below the surface see the odd image of the
  traction flux about row 3 (src/pml_fs_kernels.cu:7-19): on a flat grid they
  equal the DRP of the odd extension of sigma_xz / sigma_zz;
* the Rayleigh wave travels along the free surface at the root of the
  half-space secular equation (the ALM closure, its sign and the ghost rows
  together fix it);
* frozen rows and columns stay exactly zero.
"""
import math

import numpy as np
import pytest

import pyoracle
from msfdm import modelgen

F, B = 0, 1
DRP = np.array([-0.30874, -0.63260, 1.23300, -0.33340, 0.04168])   # offsets -1..3 (forward)


def _small_flat(**over):
    kw = dict(nx=64, nz=48, h=5.0, npd=8, fpeak=20.0, dt=4e-4, nt=100, src_x=[32], src_z=[10],
              solid=(3000.0, 1600.0, 2100.0))
    kw.update(over)
    return modelgen.build_solid_model(**kw)


def test_alm_flat_grid_closed_form():
    """Flat grid (xi_x = zeta_z = 1/h): al1_xx = 1/h, al1_zz = -C13/(C33 h) and
    every other coefficient 0, i.e. sigma_zz and sigma_xz rates vanish on the
    surface and sigma_xx' = (C11 - C13^2/C33) dvx/dx."""
    m = _small_flat()
    o = pyoracle.Oracle(m)
    a = o.alm()
    h = 5.0
    C13 = m.arrays["C13"][3].astype(np.float64)
    C33 = m.arrays["C33"][3].astype(np.float64)
    xix = m.arrays["xi_x"][3].astype(np.float64)
    assert np.allclose(a[0], xix, rtol=1e-14)
    assert np.allclose(a[0], 1.0 / h, rtol=1e-7)
    assert np.allclose(a[1], -(C13 / C33) * m.arrays["ze_z"][3], rtol=1e-13)
    for k in (2, 3, 4, 5):
        assert np.abs(a[k]).max() < 1e-14 / h, k


@pytest.mark.parametrize("direction", [F, B])
def test_free_surface_rates_carry_zero_traction(direction):
    """Random state on a curvilinear random-medium grid: the surface-row stress
    rates satisfy sigma' . n = 0 for n ~ (zeta_x, zeta_z) at every updated
    column, including the xi-PML columns (their memory terms are again ALM
    combinations), to rounding."""
    m = modelgen.preset("solid_tiny")
    o = pyoracle.Oracle(m)
    rng = np.random.default_rng(3 + direction)
    for f in ("SVX", "SVZ", "SSX", "SSZ", "SSXZ", "SXEX", "SXEZ"):
        shape = o.shape(pyoracle.FID[f])
        a = np.zeros(shape)
        a[3:-3, 3:-3] = rng.standard_normal((shape[0] - 6, shape[1] - 6))
        if f in ("SXEX", "SXEZ"):   # memory only lives in the side strips
            mask = np.zeros(shape[1], bool)
            mask[3:3 + m.npd] = True
            mask[m.nx - 3 - m.npd:m.nx - 3] = True
            a[:, ~mask] = 0.0
        o.set(f, a)
    o.rhs(direction, 0.0)
    ksx, ksz, ksxz = (o.get_k(f)[3] for f in ("SSX", "SSZ", "SSXZ"))
    zx = m.arrays["ze_x"][3].astype(np.float64)
    zz = m.arrays["ze_z"][3].astype(np.float64)
    cols = slice(3, m.nx - 3)
    tx = (ksx * zx + ksxz * zz)[cols]
    tz = (ksxz * zx + ksz * zz)[cols]
    scale = np.abs(np.stack([ksx, ksz, ksxz]))[:, cols].max() * np.hypot(zx, zz)[cols].max()
    assert scale > 0
    assert np.abs(tx).max() < 1e-12 * scale
    assert np.abs(tz).max() < 1e-12 * scale
    # the curvilinear case is not the flat one: the tangential rate is non-zero
    assert np.abs(ksxz[cols]).max() > 1e-3 * scale


@pytest.mark.parametrize("direction", [F, B])
def test_conservative_rows_see_odd_image_about_the_surface(direction):
    """Flat grid, rows 3..5, interior columns (no PML): k_vx = (xi_x Dx sxx +
    zeta_z Dz sxz~) / rho and k_vz = (xi_x Dx sxz + zeta_z Dz szz~) / rho, with
    s~ the odd extension of s about row 3 (s~(3 - n) = -s(3 + n))."""
    m = _small_flat()
    o = pyoracle.Oracle(m)
    rng = np.random.default_rng(11)
    st = {}
    for f in ("SSX", "SSZ", "SSXZ"):
        a = np.zeros((m.nz_solid, m.nx))
        a[3:-3, 3:-3] = rng.standard_normal((m.nz_solid - 6, m.nx - 6))
        o.set(f, a)
        st[f] = a
    o.rhs(direction, 0.0)
    kvx, kvz = o.get_k("SVX"), o.get_k("SVZ")
    h = 5.0
    rho = m.arrays["rho_s"].astype(np.float64)

    def odd(a):
        b = a.copy()
        for n in (1, 2, 3):
            b[3 - n] = -a[3 + n]
        return b

    def d1(line, i):   # DRP of a 1-D line at index i
        if direction == F:
            return sum(DRP[q] * line[i - 1 + q] for q in range(5))
        return -sum(DRP[q] * line[i + 1 - q] for q in range(5))

    lo, hi = 3 + m.npd, m.nx - 3 - m.npd
    for iz in (3, 4, 5):
        for ix in range(lo, hi):
            ex = (1.0 / h) * d1(st["SSX"][iz], ix) + (1.0 / h) * d1(odd(st["SSXZ"])[:, ix], iz)
            ez = (1.0 / h) * d1(st["SSXZ"][iz], ix) + (1.0 / h) * d1(odd(st["SSZ"])[:, ix], iz)
            # metrics are float32 values of 1/h; compare at float32 relative precision
            assert kvx[iz, ix] == pytest.approx(ex / rho[iz, ix], rel=2e-6, abs=1e-9), (iz, ix)
            assert kvz[iz, ix] == pytest.approx(ez / rho[iz, ix], rel=2e-6, abs=1e-9), (iz, ix)


def _rayleigh_ratio(vp, vs):
    """Root xi = c_R / vs of (2 - xi^2)^2 = 4 sqrt(1 - xi^2) sqrt(1 - xi^2 vs^2/vp^2)."""
    k2 = (vs / vp) ** 2
    f = lambda x: (2 - x * x) ** 2 - 4 * math.sqrt(1 - x * x) * math.sqrt(1 - x * x * k2)
    lo, hi = 0.5, 0.999999
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(lo) * f(mid) <= 0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def test_rayleigh_wave_speed_along_the_free_surface():
    """Homogeneous Poisson solid (vp = sqrt(3) vs), shallow source: the vz
    records of the surface row at two offsets, cross-correlated, move at the
    Rayleigh speed 0.9194 vs within 1.5 % (S is 8.8 % faster, P 88 %)."""
    vs = 1800.0
    vp = math.sqrt(3.0) * vs
    m = modelgen.build_solid_model(nx=321, nz=101, h=5.0, npd=20, fpeak=15.0, dt=5e-4, nt=1200,
                                   src_x=[60], src_z=[5], solid=(vp, vs, 2000.0))
    o = pyoracle.Oracle(m)
    o.reset(0)
    o.run(1200)
    (vz,) = o.records()
    dt_rec = o.gap * m.dt
    x1, x2 = 60 + 40, 60 + 120          # 200 m and 600 m from the source, outside the PML
    a, b = vz[:, x1], vz[:, x2]
    cc = np.correlate(b, a, mode="full")
    k = int(np.argmax(cc))
    # parabolic refinement of the peak
    y0, y1, y2 = cc[k - 1], cc[k], cc[k + 1]
    frac = 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)
    lag = (k - (len(a) - 1) + frac) * dt_rec
    c = (x2 - x1) * 5.0 / lag
    cr = _rayleigh_ratio(vp, vs) * vs
    assert _rayleigh_ratio(vp, vs) == pytest.approx(0.9194, abs=2e-4)
    assert c == pytest.approx(cr, rel=0.015), (c, cr)


def test_frozen_bands_stay_zero_and_run_is_deterministic():
    m = modelgen.preset("solid_tiny")
    runs = []
    for _ in range(2):
        o = pyoracle.Oracle(m)
        o.reset(1)
        o.run(60)
        runs.append({f: o.get(f) for f in ("SVX", "SVZ", "SSX", "SSZ", "SSXZ")})
    for f, a in runs[0].items():
        assert np.array_equal(a, runs[1][f])
        assert not a[:3].any() and not a[-3:].any() and not a[:, :3].any() and not a[:, -3:].any(), f
        assert np.abs(a).max() > 0, f
