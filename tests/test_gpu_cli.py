"""GPU: the process boundary bin/main_debug <model.json> <out_root> (src/main.cu:58-339).

Checks the output layout of docs/io.md / src/main.cu (snapshots of vx, vz for
both media at it % snapshot == 0, three time-major record files per shot), that
the output root is cleared first, and that the records equal the in-process
solver's records for every shot (each shot starting from a zero wavefield).
"""
import json
import os
import subprocess

import numpy as np
import pytest

import msfdm
from msfdm import modelgen

pytestmark = pytest.mark.gpu


def test_main_debug_writes_reference_layout(tmp_path):
    m = modelgen.preset("tiny_curvi", nt=250, snapshot=100)
    jpath = modelgen.write_model(m, str(tmp_path / "models" / "tiny"), "tiny", relative_to=str(tmp_path))
    out = tmp_path / "output" / "tiny"
    out.mkdir(parents=True)
    (out / "stale.bin").write_bytes(b"old")
    print("[cli] running main_debug", flush=True)
    r = subprocess.run([msfdm.BIN_PATH, os.path.relpath(jpath, tmp_path), "output/tiny"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert not (out / "stale.bin").exists()
    nx, nzf, nzs = m.nx, m.nz_fluid, m.nz_solid
    gap = int(1e-3 / np.float32(m.dt) + 0.5)
    nrec = (m.nt - 1) // gap + 1
    s = msfdm.Solver(m)
    for shot in range(len(m.src_x)):
        sd = out / str(shot)
        for it in (0, 100, 200):
            for med, rows in (("solid", nzs), ("fluid", nzf)):
                for f in ("vz", "vx"):
                    p = sd / "snapshot" / med / f / f"{f}_{it:06d}.bin"
                    assert p.exists(), p
                    assert p.stat().st_size == rows * nx * 4
        assert not (sd / "snapshot" / "solid" / "vz" / "vz_000050.bin").exists()
        recs = {n: np.fromfile(sd / "record" / f"{n}.bin", dtype="<f4").reshape(-1, nx)
                for n in ("solid_vx", "solid_vz", "fluid_p")}
        for a in recs.values():
            assert a.shape == (nrec, nx)
        s.reset(shot)
        s.run(m.nt)
        vx, vz, p = s.records()
        assert np.array_equal(recs["solid_vx"], vx)
        assert np.array_equal(recs["solid_vz"], vz)
        assert np.array_equal(recs["fluid_p"], p)
        # the last snapshot holds the state after step 200
        s.reset(shot)
        s.run(201)
        snap = np.fromfile(sd / "snapshot" / "solid" / "vz" / "vz_000200.bin", dtype="<f4").reshape(nzs, nx)
        assert np.array_equal(snap, s.get("SVZ"))
