"""Phase-shift dispersion image, restated in numpy (float64).

TEST INFRASTRUCTURE ONLY: the checker for the HIP dispersion-image kernel
(msfdm_dispersion_image); imported by tests/ only, never by the product path.

Restates the reference's ``dispersion_imaging`` (tools/MAS.py:50-98, the same
function in tools/test.py:50-99): the real FFT of every trace over time, the
phase-only spectrum conj(S) / |S| (0 where |S| = 0), and for each frequency
f and test velocity c the stack |sum_x P(f, x) exp(-i 2 pi f x / c)| / ntr.
"""
from __future__ import annotations

import numpy as np


def dispersion_image(gather, offsets, fs, velocities):
    """gather [nt][ntr] (time-major), offsets [ntr] (m), fs (Hz), velocities [nc]
    (m/s) -> (frequencies [nf], amplitude [nf][nc]) with nf = nt // 2 + 1."""
    g = np.asarray(gather, dtype=np.float64)
    x = np.asarray(offsets, dtype=np.float64)
    c = np.asarray(velocities, dtype=np.float64)
    spectrum = np.fft.rfft(g, axis=0)
    freqs = np.fft.rfftfreq(g.shape[0], d=1.0 / fs)
    mag = np.abs(spectrum)
    phase = np.divide(spectrum.conj(), mag, out=np.zeros_like(spectrum), where=mag > 0)
    omega = 2.0 * np.pi * freqs
    amp = np.empty((freqs.size, c.size))
    for i, w in enumerate(omega):
        shifts = np.exp(-1j * w * x[:, None] / c[None, :])
        amp[i] = np.abs(np.sum(phase[i, :, None] * shifts, axis=0)) / g.shape[1]
    return freqs, amp


def velocity_grid(c_min, c_max, dc):
    """Test velocities c_min, c_min + dc, ... up to c_max (tools/MAS.py:82)."""
    return np.arange(c_min, c_max + dc / 2.0, dc)
