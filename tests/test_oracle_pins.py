"""Pins of the FP64 oracle against mathematics that does not come from the oracle.

The reference repository holds no tests, golden vectors or output data, so each
part of the oracle is checked against closed forms, exact polynomial identities,
independent constructions and physics (reflection coefficients):

* one-sided DRP / 4-2 / 2-2 operators vs analytic derivatives and their F/B mirror
  identity (include/tools_func.cuh:51-157);
* the distance-dependent zeta operator choice (include/tools_func.cuh:131-140);
* ghost fluxes = odd (linear) extensions about the interface and the surface
  (src/pml_fs_kernels.cu:385-426), seen through the conservative-form RHS rows;
* the CFS-PML profile vs the design reflection coefficient Rc
  (src/elements.cu:832-854, docs/test.md:863-880);
* the Ricker wavelet's peak, zeros and cut-off (src/elements.cu:385-392);
* normal-incidence plane-wave reflection at the seafloor (Z2-Z1)/(Z2+Z1) and at
  the pressure-release surface (-1), which fixes the sign and scaling of the
  interface conditions and the filter (include/pml_fs_kernels.cuh:754-827,
  src/pml_fs_kernels.cu:428-525).
"""
import math

import numpy as np
import pytest

import pyoracle
from msfdm import modelgen

F, B = 0, 1


def _op(kind, d, f, i, dist=3, solid=False):
    return pyoracle.op1d(kind, d, f, i, dist, solid)


def test_drp_matches_analytic_derivative_of_smooth_waves():
    """Each one-sided DRP operator is within 0.5 % of k cos(kx) from 10 to 80
    points per wavelength; the F/B pair (the MacCormack predictor-corrector
    average) is a centred operator with errors below 1e-3 at 10 and 1e-4 at
    20+ points per wavelength.  A mistyped weight breaks both bounds."""
    n = 400
    x = np.arange(n, dtype=np.float64)
    for ppw in (10.0, 20.0, 40.0, 80.0):
        k = 2 * np.pi / ppw
        f = np.sin(k * x + 0.3)
        df = k * np.cos(k * x + 0.3)[10:n - 10]
        one = {d: np.array([_op(0, d, f, i) for i in range(10, n - 10)]) for d in (F, B)}
        for d in (F, B):
            assert np.abs(one[d] - df).max() / k < 5e-3, (ppw, d)
        avg = 0.5 * (one[F] + one[B])
        assert np.abs(avg - df).max() / k < (1e-3 if ppw == 10.0 else 1e-4), ppw


def test_drp_backward_is_mirrored_forward():
    rng = np.random.default_rng(0)
    f = rng.standard_normal(64)
    g = f[::-1].copy()
    for i in range(8, 56):
        assert _op(0, B, f, i) == pytest.approx(-_op(0, F, g, 63 - i), rel=1e-13, abs=1e-13)
        assert _op(1, B, f, i) == pytest.approx(-_op(1, F, g, 63 - i), rel=1e-13, abs=1e-13)
        assert _op(2, B, f, i) == pytest.approx(-_op(2, F, g, 63 - i), rel=1e-13, abs=1e-13)


def test_closures_exact_on_lines_and_fb_average_on_quadratics():
    x = np.arange(20, dtype=np.float64)
    lin = 3.0 - 1.7 * x
    quad = 0.4 * x * x - 2.0 * x + 1.0
    for kind in (1, 2):                       # M42, M22
        for d in (F, B):
            assert _op(kind, d, lin, 9) == pytest.approx(-1.7, abs=1e-12)
        avg = 0.5 * (_op(kind, F, quad, 9) + _op(kind, B, quad, 9))
        assert avg == pytest.approx(0.8 * 9 - 2.0, abs=1e-12)
    # the DRP pair's average is a centred 7-point operator: near-exact on quadratics
    avg = 0.5 * (_op(0, F, quad, 9) + _op(0, B, quad, 9))
    assert avg == pytest.approx(0.8 * 9 - 2.0, rel=1e-3)


def test_distance_dependent_operator_choice():
    rng = np.random.default_rng(1)
    f = rng.standard_normal(32)
    i = 16
    for d in (F, B):
        assert _op(3, d, f, i, dist=5) == _op(0, d, f, i)
        assert _op(3, d, f, i, dist=3) == _op(0, d, f, i)
        assert _op(3, d, f, i, dist=2) == _op(1, d, f, i)
        assert _op(3, d, f, i, dist=1) == _op(2, d, f, i)
        # on the interface row the one-sided 2-2 difference always points into the medium
        assert _op(3, d, f, i, dist=0, solid=True) == pytest.approx(f[i + 1] - f[i])
        assert _op(3, d, f, i, dist=0, solid=False) == pytest.approx(f[i] - f[i - 1])


def _flat_model(**kw):
    base = dict(nx=96, nz=70, h=5.0, water_depth=120.0, npd=10, fpeak=20.0, dt=4e-4, nt=10,
                src_x=[48], src_height=10.0)
    base.update(kw)
    return modelgen.build_model(**base)


def _drp_line(d, vals, i):
    return _op(0, d, np.asarray(vals, dtype=np.float64), i)


@pytest.mark.parametrize("d", [F, B])
def test_conservative_interface_rows_equal_drp_of_linear_extension(d):
    """Flat grid, p linear in depth near the seafloor: the conservative rows with
    ghost fluxes must equal -grad p / rho built from the DRP of the linearly
    extended column (the ghost 2F(0)-F(-n) is exact for lines)."""
    m = _flat_model()
    o = pyoracle.Oracle(m)
    nzf, nx = m.nz_fluid, m.nx
    rows = np.arange(nzf, dtype=np.float64)
    cols = np.arange(nx, dtype=np.float64)
    a, b = 2.0, -0.35
    p = (a + b * rows)[:, None] * (1.0 + 0.1 * np.cos(cols / 7.0))[None, :]
    o.set("FP", p)
    o.rhs(d, 0.0)
    kvz = o.get_k("FVZ")
    rho = m.arrays["rho_f"][0, 0].astype(np.float64)
    zz = float(m.arrays["ze_z"][0, 0])
    for iz in (nzf - 3, nzf - 2):   # rows whose stencil reaches the ghosts
        for ix in (30, 47, 61):
            col = (a + b * np.arange(nzf + 3)) * (1.0 + 0.1 * np.cos(ix / 7.0))
            expect = -zz * _drp_line(d, col, iz) / rho
            assert kvz[iz, ix] == pytest.approx(expect, rel=1e-10), (iz, ix)


@pytest.mark.parametrize("d", [F, B])
def test_conservative_surface_rows_equal_drp_of_odd_extension(d):
    """Near the pressure-release row 3 the ghost -F(3+n) is the odd image: for p
    odd about row 3 the conservative rows equal the DRP of the odd extension."""
    m = _flat_model()
    o = pyoracle.Oracle(m)
    nzf, nx = m.nz_fluid, m.nx
    rows = np.arange(nzf, dtype=np.float64)
    prof = np.sin(0.21 * (rows - 3.0))
    p = prof[:, None] * np.ones(nx)[None, :]
    p[:3] = 0.0   # frozen rows hold nothing; the ghosts must supply the image
    o.set("FP", p)
    o.rhs(d, 0.0)
    kvz = o.get_k("FVZ")
    rho = float(m.arrays["rho_f"][0, 0])
    zz = float(m.arrays["ze_z"][0, 0])
    ext = np.sin(0.21 * (np.arange(nzf, dtype=np.float64) - 3.0))   # odd about row 3
    for iz in (3, 4, 5):
        expect = -zz * _drp_line(d, ext, iz) / rho
        assert kvz[iz, 48] == pytest.approx(expect, rel=1e-10, abs=1e-14), iz


def test_pml_profile_meets_design_reflection_coefficient():
    m = _flat_model(npd=20, nx=120, nz=90)
    o = pyoracle.Oracle(m)
    al, be, dp = o.pml(0)
    npd, nx = m.npd, m.nx
    inside = np.r_[np.arange(0, 3), np.arange(3 + npd, nx - 3 - npd), np.arange(nx - 3, nx)]
    assert np.all(dp[inside] == 0.0) and np.all(be[inside] == 1.0)
    strip = dp[3:3 + npd][::-1]          # from the inner edge outwards
    assert np.all(np.diff(strip) > 0)
    assert np.allclose(dp[3:3 + npd], dp[nx - 4:nx - 4 - npd:-1])   # left/right symmetric
    # integral over the strip thickness L of d(r) gives R = exp(-2/c * int d)
    r = np.r_[0.0, np.arange(1, npd + 1) / npd]
    d = np.r_[0.0, strip]
    integral = m.thickness * np.trapezoid(d, r)
    R = math.exp(-2.0 * integral / m.vp_max)
    assert abs(math.log(R) / math.log(m.rc) - 1.0) < 0.01
    assert be[3] == pytest.approx(m.beta0, rel=1e-6)          # outermost cell: beta = beta0
    assert al[3] == pytest.approx(0.0, abs=1e-9)               # alpha -> 0 at the edge
    assert al[nx // 2] == pytest.approx(2 * math.pi * m.fpeak, rel=1e-6)   # alpha0 = 2 pi fpeak


def test_ricker_closed_form():
    m = _flat_model(fpeak=25.0, dt=1e-4)
    o = pyoracle.Oracle(m)
    f = 25.0
    T = 1.2 / f
    it_peak = int(round(T / 1e-4))
    assert o.stage_amp(it_peak, 0) == pytest.approx(1.0, abs=2e-3)
    # zeros of (1 - 2 a) exp(-a) at |t - T| = 1/(pi f sqrt 2)
    t0 = T + 1.0 / (math.pi * f * math.sqrt(2.0))
    it0 = int(t0 / 1e-4)
    assert o.stage_amp(it0, 0) * o.stage_amp(it0 + 1, 0) <= 0.0
    assert o.stage_amp(0, 0) == pytest.approx((1 - 2 * (1.2 * math.pi) ** 2) * math.exp(-(1.2 * math.pi) ** 2), rel=1e-5)
    assert o.stage_amp(int(2.2 * T / 1e-4), 0) == 0.0        # cut off once t - T > T
    # stages 2 and 3 both sit at t + dt/2
    assert o.stage_amp(17, 1) == o.stage_amp(17, 2)


def _plane_wave_run(water_depth, z_pulse, nsteps, probe_rows, pulse_down=True, boundary_type=2):
    """x-invariant P pulse in the water column of a flat model; records p at probe rows."""
    h = 2.5
    m = modelgen.build_model(nx=240, nz=130, h=h, water_depth=water_depth, npd=10, fpeak=40.0,
                             dt=2e-4, nt=nsteps, src_x=[], src_z=[],
                             solid=(3000.0, 1700.0, 2200.0), vp_max=3000.0, boundary_type=boundary_type)
    o = pyoracle.Oracle(m)
    nzf, nx = m.nz_fluid, m.nx
    rows = np.arange(nzf, dtype=np.float64)
    z = (rows - 3.0) * h
    width = 6.0 * h
    p = np.exp(-((z - z_pulse) / width) ** 2)
    rho, c = 1000.0, 1500.0
    vz = (1.0 if pulse_down else -1.0) * p / (rho * c)
    P = np.zeros((nzf, nx)); VZ = np.zeros((nzf, nx))
    P[3:, 3:nx - 3] = p[3:, None]
    VZ[3:, 3:nx - 3] = vz[3:, None]
    o.set("FP", P)
    o.set("FVZ", VZ)
    o.step = 0
    out = {r: [] for r in probe_rows}
    for _ in range(nsteps):
        o.run(1)
        pp = o.get("FP")
        for r in probe_rows:
            out[r].append(pp[r, nx // 2])
    return m, {r: np.array(v) for r, v in out.items()}


def test_seafloor_normal_incidence_reflection_coefficient():
    """A downgoing pulse reflects off the flat seafloor with R = (Z2 - Z1)/(Z2 + Z1)."""
    m, tr = _plane_wave_run(water_depth=200.0, z_pulse=120.0, nsteps=520, probe_rows=[3 + 60])
    sig = tr[63]
    nt = sig.size
    # incident pulse passes the probe (z = 150 m) at ~ (150-120)/1500 s, the
    # reflection at ~ (200-120 + 200-150)/1500 s
    dt = m.dt
    i_inc = int(30 / 1500 / dt); i_ref = int(130 / 1500 / dt)
    w = int(0.012 / dt)
    a_inc = sig[max(0, i_inc - w):i_inc + w].max()
    a_ref = sig[i_ref - w:min(nt, i_ref + w)].max()
    z1, z2 = 1000.0 * 1500.0, 2200.0 * 3000.0
    R = (z2 - z1) / (z2 + z1)
    assert a_ref / a_inc == pytest.approx(R, rel=0.03)


def test_pressure_release_surface_reflection_is_minus_one():
    """An upgoing pulse returns from the p = 0 sea surface with reflection -1."""
    m, tr = _plane_wave_run(water_depth=250.0, z_pulse=100.0, nsteps=620, probe_rows=[3 + 24],
                            pulse_down=False)
    sig = tr[27]
    dt = m.dt
    i_inc = int(40 / 1500 / dt); i_ref = int(160 / 1500 / dt)
    w = int(0.012 / dt)
    a_inc = sig[max(0, i_inc - w):i_inc + w].max()
    a_ref = sig[i_ref - w:i_ref + w].min()
    assert a_ref / a_inc == pytest.approx(-1.0, rel=0.03)


def test_cpml_sea_surface_absorbs_the_upgoing_pulse():
    """With the top PML of the fluid-solid CPML grid (include/cpml_kernels.cuh:239-415)
    the same upgoing pulse that the pressure-release surface returns with -1
    comes back with less than 2 % of its amplitude."""
    m, tr = _plane_wave_run(water_depth=250.0, z_pulse=100.0, nsteps=620, probe_rows=[3 + 24],
                            pulse_down=False, boundary_type=1)
    sig = tr[27]
    dt = m.dt
    i_inc = int(40 / 1500 / dt); i_ref = int(160 / 1500 / dt)
    w = int(0.012 / dt)
    a_inc = sig[max(0, i_inc - w):i_inc + w].max()
    a_ref = np.abs(sig[i_ref - w:i_ref + w]).max()
    assert a_inc > 0.5
    assert a_ref / a_inc < 0.02


def test_cpml_fluid_zeta_profile_is_the_top_strip_formula():
    """The fluid zeta profile of the CPML grid is the lo strip of the profile
    formula (src/elements.cu:832-854, 870-873): equal to the left xi strip, zero
    damping below it; the PANDF grid has none."""
    m = modelgen.preset("cpml_fs_tiny")
    o = pyoracle.Oracle(m)
    af, bf, df = o.pml(2)
    ax, bx, dx = o.pml(0)
    npd = m.npd
    for a, b in ((af, ax), (bf, bx), (df, dx)):
        assert np.array_equal(a[3:3 + npd], b[3:3 + npd])
    assert np.all(df[3 + npd:] == 0.0) and np.all(df[3:3 + npd] > 0)
    _, _, d0 = pyoracle.Oracle(modelgen.preset("tiny_curvi")).pml(2)
    assert not d0.any()


def test_oracle_step_invariants():
    m = modelgen.preset("tiny_curvi")
    o = pyoracle.Oracle(m)
    o.reset(0)
    o.run(80)
    p = o.get("FP")
    assert np.all(p[3] == 0.0)
    for name in pyoracle.FIELDS:
        a = o.get(name)
        assert np.all(a[:, :3] == 0.0) and np.all(a[:, -3:] == 0.0), name
    nzf = m.nz_fluid
    n1, n2 = o.normals()
    vnf = n1 * o.get("FVX")[nzf - 1] + n2 * o.get("FVZ")[nzf - 1]
    vns = n1 * o.get("SVX")[0] + n2 * o.get("SVZ")[0]
    c = slice(3, m.nx - 3)
    assert np.abs(vnf - vns)[c].max() <= 1e-12 * np.abs(vns).max()
    sxx, szz, sxz = o.get("SSX")[0], o.get("SSZ")[0], o.get("SSXZ")[0]
    snn = n1 * n1 * sxx + 2 * n1 * n2 * sxz + n2 * n2 * szz
    # p = -sigma_nn holds up to the float normal's |n|^2 - 1 (about 1e-7): the
    # stresses are rebuilt from sigma_nn along n, so their sigma_nn carries |n|^4
    dn = np.abs(n1 * n1 + n2 * n2 - 1.0)[c].max()
    assert np.abs(p[nzf - 1] + snn)[c].max() <= (4 * dn + 1e-12) * np.abs(snn).max()
    # the normal is the reference's float computation (src/elements.cu:892-905):
    # float components, a float square root and float quotients
    zx, zz = m.arrays["ze_x"][nzf - 1], m.arrays["ze_z"][nzf - 1]
    s32 = np.sqrt(zx * zx + zz * zz, dtype=np.float32)
    assert np.array_equal(n1, (zx / s32).astype(np.float64)) and np.array_equal(n2, (zz / s32).astype(np.float64))
    assert np.abs(n1 * n1 + n2 * n2 - 1.0).max() > 1e-9      # not renormalised in double
    # flat models have the vertical normal exactly
    mf = _flat_model()
    n1f, n2f = pyoracle.Oracle(mf).normals()
    assert np.all(n1f == 0.0) and np.all(n2f == 1.0)


def test_filter_divides_by_the_float_tap_sum():
    """The reference sums the seven float filter taps in float and divides by
    that sum (src/elements.cu:30-33, src/pml_fs_kernels.cu:474-487).  The float
    sum rounds 1.8e-8 below the exact sum of the taps, so the filter has a DC
    gain of 1 + 1.8e-8 per step: 1.1e-4 over a 6000-step run.  On a flat
    interface with a state that already meets the interface conditions and is
    constant along x, the filter can only apply that gain.  Dividing by the
    exact sum (gain 1) fails the test."""
    taps = np.array([1.1e-13, 1.678e-06, 0.03419, 0.93163, 0.03419, 1.678e-06, 1.1e-13], dtype=np.float32)
    fsum = np.float32(0.0)
    for w in taps:                      # the reference's left-to-right float accumulation
        fsum = np.float32(fsum + w)
    gain = math.fsum(float(w) for w in taps) / float(fsum)
    assert 1.5e-8 < gain - 1.0 < 2.5e-8

    m = modelgen.build_model(nx=72, nz=56, h=5.0, water_depth=100.0, npd=8, fpeak=25.0, dt=4e-4, nt=60,
                             src_x=[36], src_height=10.0)   # flat seafloor: normal (0, +-1)
    o = pyoracle.Oracle(m)
    rf = m.nz_fluid - 1
    # v_n equal on both sides, p = -sigma_zz, sigma_xz = 0; the tangential parts are free
    iface = {"FVX": 0.3, "FVZ": 0.7, "FP": 2.0, "SVX": -0.4, "SVZ": 0.7, "SSX": 1.5, "SSZ": -2.0, "SSXZ": 0.0}
    for i, n in enumerate(pyoracle.FIELDS):
        a = np.zeros(o.shape(i))
        if n in iface:
            a[rf if n.startswith("F") else 0, :] = iface[n]
        o.set(n, a)
    o.filter()
    inner = slice(6, m.nx - 6)          # all seven taps inside columns 3..nx-4
    for n, v in iface.items():
        row = o.get(n)[rf if n.startswith("F") else 0, inner]
        if v == 0.0:
            assert np.abs(row).max() < 1e-15, n
        else:
            assert np.allclose(row / v, gain, rtol=0, atol=1e-13), (n, row[0] / v - 1.0)
