"""Consumer side of the solver output: readers for the record and snapshot files,
the phase-shift dispersion image on the GPU, and the trace comparison measure.

The file layout is the one bin/main_debug writes (src/main.cu:218-339):
``<out_root>/<shot>/record/{solid_vx,solid_vz,fluid_p}.bin`` time-major
``[nrec][nx]`` little-endian float32, and
``<out_root>/<shot>/snapshot/{solid,fluid}/{vz,vx}/{vz,vx}_%06d.bin`` full
``[rows][nx]`` arrays.  The readers follow tools/MAS.py:9-47 (PML columns cut,
signed offsets of the kept traces) and tools/plot_wavefield.py.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import check, lib

RECORDS = ("solid_vx", "solid_vz", "fluid_p")


def record_path(out_root: str, shot: int, name: str) -> str:
    if name not in RECORDS:
        raise ValueError(f"record must be one of {RECORDS}, not {name!r}")
    return os.path.join(out_root, str(shot), "record", name + ".bin")


def read_record(out_root: str, shot: int, name: str, nx: int) -> np.ndarray:
    """The [nrec][nx] float32 record of one component."""
    a = np.fromfile(record_path(out_root, shot, name), dtype="<f4")
    if a.size % nx:
        raise ValueError(f"{name}: {a.size} floats is not a whole number of {nx}-trace samples")
    return a.reshape(-1, nx)


def read_snapshot(out_root: str, shot: int, medium: str, comp: str, it: int, rows: int, nx: int) -> np.ndarray:
    """One snapshot array ([rows][nx]) of medium 'solid' / 'fluid', component 'vx' / 'vz'."""
    if medium not in ("solid", "fluid") or comp not in ("vx", "vz"):
        raise ValueError("medium must be 'solid' or 'fluid' and comp 'vx' or 'vz'")
    path = os.path.join(out_root, str(shot), "snapshot", medium, comp, f"{comp}_{it:06d}.bin")
    a = np.fromfile(path, dtype="<f4")
    if a.size != rows * nx:
        raise ValueError(f"{path}: expected {rows * nx} floats, found {a.size}")
    return a.reshape(rows, nx)


def gather(record: np.ndarray, npd: int, h: float, src_x: int):
    """Cut the npd PML columns (and the frozen ones inside them, as
    tools/MAS.py:37-44 does) off a [nt][nx] record; returns the gather and the
    signed offsets (m) of the kept columns from the source column."""
    nt, nx = record.shape
    if npd < 0 or 2 * npd >= nx:
        raise ValueError("npd must leave at least one trace")
    g = np.ascontiguousarray(record[:, npd:nx - npd])
    offsets = (np.arange(npd, nx - npd) - src_x) * float(h)
    return g, offsets


def dispersion_image(gather_tx, offsets, fs: float, c_min: float, c_max: float, dc: float, device: int = 0):
    """Phase-shift dispersion image (tools/MAS.py:50-98) computed on the GPU in
    FP64.  Returns (frequencies [nf], velocities [nc], amplitude [nf][nc])."""
    g = np.ascontiguousarray(gather_tx, dtype=np.float32)
    if g.ndim != 2:
        raise ValueError("gather must have shape (nt, number_of_receivers)")
    x = np.ascontiguousarray(offsets, dtype=np.float64)
    if x.shape != (g.shape[1],):
        raise ValueError("there must be exactly one offset for each trace")
    if not (np.isfinite(c_min) and np.isfinite(c_max) and 0 < c_min <= c_max) or not (dc > 0):
        raise ValueError("velocities must satisfy 0 < c_min <= c_max and dc > 0")
    vel = np.arange(c_min, c_max + dc / 2.0, dc, dtype=np.float64)
    nt, ntr = g.shape
    nf = nt // 2 + 1
    freqs = np.empty(nf)
    amp = np.empty((nf, vel.size))
    dp = C.POINTER(C.c_double)
    check(lib().msfdm_dispersion_image(g.ctypes.data_as(C.POINTER(C.c_float)), nt, ntr, x.ctypes.data_as(dp),
                                       float(fs), vel.ctypes.data_as(dp), vel.size, int(device),
                                       freqs.ctypes.data_as(dp), amp.ctypes.data_as(dp)))
    return freqs, vel, amp


def ncc(a, b) -> float:
    """Zero-lag normalised cross-correlation of two traces after removing their
    means (tools/compare_sem_scholte.py:67-70); 0 for a constant trace."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    a = a - a.mean()
    b = b - b.mean()
    d = np.sqrt(np.dot(a, a) * np.dot(b, b))
    return float(np.dot(a, b) / d) if d > 0 else 0.0
