"""GPU dispersion image (msfdm_dispersion_image, the phase-shift method of
tools/MAS.py:50-98 in FP64) against the numpy restatement oracle/mas.py.

Tolerance: 1e-9 of the image maximum on a noise gather (every frequency has a
well-defined phase, so both sides normalise the same spectrum), and the plane
wave's peak velocity exactly."""
import math

import numpy as np
import pytest

import mas
import msfdm
from msfdm import modelgen, records

pytestmark = pytest.mark.gpu


def test_dispersion_image_matches_oracle_on_noise():
    rng = np.random.default_rng(9)
    nt, ntr, fs = 700, 173, 1000.0
    g = rng.standard_normal((nt, ntr)).astype(np.float32)
    x = (np.arange(ntr) - 40) * 2.5
    f, vel, amp = records.dispersion_image(g, x, fs, 200.0, 4000.0, 12.5)
    fo, ao = mas.dispersion_image(g.astype(np.float64), x, fs, vel)
    assert np.array_equal(f, fo)
    assert np.abs(amp - ao).max() <= 1e-9 * ao.max()


def test_dispersion_image_of_a_plane_wave_peaks_at_its_velocity():
    fs, nt, c0, f0 = 1000.0, 1024, 1350.0, 20.0
    x = np.arange(0, 200) * 5.0
    t = np.arange(nt) / fs
    a = (math.pi * f0 * (t[:, None] - 0.2 - x[None, :] / c0)) ** 2
    g = ((1 - 2 * a) * np.exp(-a)).astype(np.float32)
    f, vel, amp = records.dispersion_image(g, x, fs, 800.0, 2500.0, 5.0)
    band = (f > 8.0) & (f < 45.0)
    assert np.all(np.abs(vel[np.argmax(amp[band], axis=1)] - c0) <= 5.0)


def test_dispersion_image_of_solver_records():
    """End to end: a Scholte-regime record of the flat seafloor model through the
    reader and the GPU image, against the oracle on the same gather."""
    m = modelgen.preset("cut201", nt=1200)
    s = msfdm.Solver(m)
    s.reset(0)
    s.run(m.nt)
    vx, vz, p = s.records()
    g, off = records.gather(vz, m.npd + 3, 5.0, m.src_x[0])
    fs = 1.0 / (s.dt * 2)          # records every gap = 2 steps at dt = 0.5 ms
    f, vel, amp = records.dispersion_image(g, off, fs, 300.0, 3500.0, 10.0)
    fo, ao = mas.dispersion_image(g.astype(np.float64), off, fs, vel)
    band = (f > 3.0) & (f < 40.0)
    assert np.abs(amp[band] - ao[band]).max() <= 1e-8 * ao[band].max()


def test_dispersion_image_errors():
    g = np.zeros((10, 3), np.float32)
    with pytest.raises(ValueError):
        records.dispersion_image(g, [0.0, 1.0], 100.0, 100.0, 200.0, 10.0)
    with pytest.raises(msfdm.MsfdmError):
        records.dispersion_image(g, [0.0, 1.0, np.inf], 100.0, 100.0, 200.0, 10.0)
