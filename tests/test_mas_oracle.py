"""Pins of the phase-shift dispersion-image oracle (oracle/mas.py, restating
tools/MAS.py:50-98) and of the record readers, on the CPU.

* a plane wave g(t, x) = s(t - x / c0) stacks coherently at c = c0 only: in the
  source band the image peaks at c0 (a closed form, independent of the code);
* the stack at any single frequency equals the brute-force double sum over
  time and offset written from the definitions (no FFT);
* the readers return what the writer put in the reference's layout.
"""
import math

import numpy as np
import pytest

import mas
from msfdm import records


def _ricker(t, f):
    a = (math.pi * f * t) ** 2
    return (1 - 2 * a) * np.exp(-a)


def test_plane_wave_peaks_at_its_velocity():
    fs, nt, c0, f0 = 1000.0, 1024, 1350.0, 20.0
    x = np.arange(0, 60) * 10.0 - 100.0            # signed offsets
    t = np.arange(nt) / fs
    g = _ricker(t[:, None] - 0.2 - x[None, :] / c0, f0)
    vel = mas.velocity_grid(800.0, 2500.0, 5.0)
    freqs, amp = mas.dispersion_image(g, x, fs, vel)
    band = (freqs > 8.0) & (freqs < 45.0)
    peaks = vel[np.argmax(amp[band], axis=1)]
    assert np.all(np.abs(peaks - c0) <= 5.0)
    assert np.all(amp[band].max(axis=1) > 0.999)    # phase-only stack of a pure delay is exact


def test_single_frequency_equals_brute_force_sums():
    rng = np.random.default_rng(2)
    fs, nt, ntr = 500.0, 64, 9
    g = rng.standard_normal((nt, ntr))
    x = rng.uniform(-300.0, 300.0, ntr)
    vel = np.array([700.0, 1500.0, 3100.0])
    freqs, amp = mas.dispersion_image(g, x, fs, vel)
    for k in (1, 7, nt // 2):
        S = np.array([sum(g[n, j] * np.exp(-2j * np.pi * k * n / nt) for n in range(nt)) for j in range(ntr)])
        P = np.conj(S) / np.abs(S)
        w = 2 * np.pi * k * fs / nt
        ref = [abs(sum(P[j] * np.exp(-1j * w * x[j] / c) for j in range(ntr))) / ntr for c in vel]
        assert freqs[k] == pytest.approx(k * fs / nt, rel=1e-15)
        assert np.allclose(amp[k], ref, rtol=1e-10)


def test_readers_and_gather(tmp_path):
    nx, nrec = 50, 12
    rec = np.arange(nrec * nx, dtype=np.float32).reshape(nrec, nx)
    d = tmp_path / "out" / "0" / "record"
    d.mkdir(parents=True)
    rec.tofile(d / "solid_vz.bin")
    got = records.read_record(str(tmp_path / "out"), 0, "solid_vz", nx)
    assert np.array_equal(got, rec)
    g, off = records.gather(got, npd=5, h=2.5, src_x=20)
    assert g.shape == (nrec, nx - 10) and np.array_equal(g[:, 0], rec[:, 5])
    assert off[0] == (5 - 20) * 2.5 and off[-1] == (nx - 6 - 20) * 2.5
    with pytest.raises(ValueError):
        records.read_record(str(tmp_path / "out"), 0, "solid_vz", 7)
    snap = tmp_path / "out" / "0" / "snapshot" / "solid" / "vz"
    snap.mkdir(parents=True)
    rec.tofile(snap / "vz_000040.bin")
    assert np.array_equal(records.read_snapshot(str(tmp_path / "out"), 0, "solid", "vz", 40, nrec, nx), rec)
    a = np.sin(np.arange(100) / 5.0)
    assert records.ncc(a, 3 * a + 1) == pytest.approx(1.0)
    assert records.ncc(a, -a) == pytest.approx(-1.0)
