"""GPU parity of the solid-only steppers (grid.calc_type = 0) through the C ABI
against the FP64 oracle: the traction-free top (boundary_type 2, PmlFs::SolidOnly,
include/pml_fs_kernels.cuh:8-307, src/pml_fs_kernels.cu:1-381), PML on four
sides (boundary_type 1, include/cpml_kernels.cuh:8-233) and Dirichlet edges
(boundary_type 0, include/dirichlet_kernels.cuh:7-107).

Tolerances: per-field right-hand sides on random states within 2e-5 of the
field's max |k| (FP32 rounding of ~10-term stencil sums); free-surface vz
records within a relative L2 error of 1e-4 over the whole record.
"""
import dataclasses
import os
import subprocess

import numpy as np
import pytest

import msfdm
from msfdm import modelgen
import pyoracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

RECORD_TOL = 1e-4
SOLID_FIELDS = [f for f in msfdm.FIELDS if f.startswith("S")]


def _general_metrics(m, seed=1):
    """Copy of m with xi_x, xi_z varying in space (general-metric kernel path)."""
    a = dict(m.arrays)
    nz, nx = a["xi_x"].shape
    zz, xx = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    a["xi_x"] = (a["xi_x"] * (1.0 + 0.03 * np.sin(xx / 9.0) * np.cos(zz / 7.0))).astype(np.float32)
    a["xi_z"] = (0.02 * a["xi_x"] * np.sin(zz / 11.0 + xx / 13.0)).astype(np.float32)
    return dataclasses.replace(m, arrays=a)


def _model(kind):
    if kind == "general":
        return _general_metrics(modelgen.preset("solid_tiny"))
    if kind == "cpml_general":
        return _general_metrics(modelgen.preset("solid_cpml_tiny"))
    return modelgen.preset(kind)


@pytest.mark.parametrize("kind", ["solid_tiny", "general", "solid_cpml_tiny", "cpml_general", "solid_diri_tiny"])
@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("tile_rows", [8, 16])
def test_solid_rhs_matches_oracle_on_random_state(kind, direction, tile_rows):
    print(f"[solid rhs] {kind} dir={direction} tile_rows={tile_rows}", flush=True)
    m = _model(kind)
    o = pyoracle.Oracle(m)
    s = msfdm.Solver(m)
    assert s.solid_only
    s.tile_rows = tile_rows
    rng = np.random.default_rng(31 + direction)
    for name in SOLID_FIELDS:
        a32 = rng.standard_normal((m.nz_solid, m.nx)).astype(np.float32)
        o.set(name, a32.astype(np.float64))
        s.set(name, a32)
    amp = 0.731
    o.rhs(direction, amp)
    s.eval_rhs(direction, amp)
    for name in SOLID_FIELDS:
        ko = o.get_k(name)
        kg = s.get_rhs(name).astype(np.float64)
        scale = np.abs(ko).max()
        if m.boundary_type == 0 and name in SOLID_FIELDS[5:]:   # Dirichlet: no PML memory
            assert scale == 0 and not kg.any(), name
            continue
        assert scale > 0, name
        err = np.abs(kg - ko).max() / scale
        assert err < 2e-5, f"{name}: max err {err:.3e} (scale {scale:.3e})"


def test_solid_rhs_surface_traction_is_zero():
    """The HIP kernel's raw surface-row stress rates carry zero traction along
    the surface normal to FP32 rounding (source off)."""
    m = modelgen.preset("solid_tiny")
    s = msfdm.Solver(m)
    rng = np.random.default_rng(5)
    for name in ("SVX", "SVZ", "SSX", "SSZ", "SSXZ"):
        s.set(name, rng.standard_normal((m.nz_solid, m.nx)).astype(np.float32))
    s.eval_rhs(0, 0.0)
    ksx, ksz, ksxz = (s.get_rhs(f)[3].astype(np.float64) for f in ("SSX", "SSZ", "SSXZ"))
    zx = m.arrays["ze_x"][3].astype(np.float64)
    zz = m.arrays["ze_z"][3].astype(np.float64)
    c = slice(3, m.nx - 3)
    scale = np.abs(np.stack([ksx, ksz, ksxz]))[:, c].max() * np.hypot(zx, zz)[c].max()
    assert np.abs(ksx * zx + ksxz * zz)[c].max() < 1e-5 * scale
    assert np.abs(ksxz * zx + ksz * zz)[c].max() < 1e-5 * scale


def _run_both(m, nsteps, isrc=0, graph=True, tile_rows=None):
    o = pyoracle.Oracle(m)
    o.reset(isrc)
    o.run(nsteps)
    s = msfdm.Solver(m)
    if tile_rows:
        s.tile_rows = tile_rows
    s.set_graph(graph)
    s.reset(isrc)
    s.run(nsteps)
    s.sync()
    return o, s


@pytest.mark.parametrize("kind,nsteps,tile_rows,isrc", [("solid_tiny", 300, 8, 0), ("solid_tiny", 300, 16, 1),
                                                        ("general", 200, 8, 0), ("solid_flat301", 800, 16, 0),
                                                        ("solid_cpml_tiny", 300, 8, 1), ("cpml_general", 200, 16, 0),
                                                        ("solid_diri_tiny", 300, 8, 0)])
def test_solid_records_match_oracle(kind, nsteps, tile_rows, isrc):
    print(f"[solid records] {kind} nsteps={nsteps} tile_rows={tile_rows} shot={isrc}", flush=True)
    m = _model(kind)
    o, s = _run_both(m, nsteps, isrc=isrc, tile_rows=tile_rows)
    (ro,) = o.records()
    (rg,) = s.records()
    assert rg.shape == ro.shape
    if m.boundary_type == 0:   # Dirichlet grids record nothing (src/main.cu:199-209)
        assert ro.shape[0] == 0 and s.num_records() == 0
    else:
        assert np.abs(ro).max() > 0
        err = rel_l2(rg, ro)
        print(f"  vz(row {3 if m.boundary_type == 2 else 30}) rel L2 {err:.3e}", flush=True)
        assert err <= RECORD_TOL
    # the final wavefield too
    for name in ("SVX", "SVZ", "SSX", "SSZ", "SSXZ"):
        assert rel_l2(s.get(name), o.get(name)) <= RECORD_TOL, name


def test_solid_invariants_determinism_and_graph():
    """Frozen rows/columns stay zero; two runs are bitwise equal; the graph
    replay equals direct launches; the fluid planes of the record are zero."""
    m = modelgen.preset("solid_tiny")
    runs = []
    for graph in (True, True, False):
        s = msfdm.Solver(m)
        s.set_graph(graph)
        s.reset(0)
        s.run(130)
        s.sync()
        runs.append({f: s.get(f) for f in ("SVX", "SVZ", "SSX", "SSZ", "SSXZ")} | {"rec": s.records()[0]})
    for f, a in runs[0].items():
        assert np.array_equal(a, runs[1][f]), f
        assert np.array_equal(a, runs[2][f]), f
        if f != "rec":
            assert not a[:3].any() and not a[-3:].any() and not a[:, :3].any() and not a[:, -3:].any(), f
            assert np.abs(a).max() > 0, f
    s = msfdm.Solver(m)
    s.reset(0)
    s.run(10)
    n = s.num_records()
    vx, vz, p = (np.zeros((n, m.nx), np.float32) for _ in range(3))
    import ctypes as C
    msfdm._lib.check(msfdm.lib().msfdm_get_records(s.h, 0, n, *[x.ctypes.data_as(C.POINTER(C.c_float))
                                                               for x in (vx, vz, p)]))
    assert not vx.any() and not p.any() and vz.any()


def test_solid_rejects_fluid_solid_only_paths():
    m = modelgen.preset("solid_tiny")
    s = msfdm.Solver(m)
    nrec = 10
    res = [np.zeros((nrec, m.nx), np.float32)] * 3
    with pytest.raises(msfdm.MsfdmError) as e:
        s.adjoint_forward(0, 20)
    assert e.value.code == msfdm._lib.EINVAL and "solid-only" in str(e.value)
    with pytest.raises(msfdm.MsfdmError):
        s.update_media(C11=m.arrays["C11"])
    with pytest.raises(msfdm.MsfdmError):
        msfdm.Batch([m, m])
    del res


def test_solid_cli_layout_and_records(tmp_path):
    """bin/main_debug on a solid-only model: record/solid_vz.bin only (vz of
    row 3, src/main.cu:297-322) and solid snapshots, equal to the solver's."""
    m = modelgen.preset("solid_tiny", nt=120, snapshot=50)
    j = modelgen.write_model(m, str(tmp_path / "model"), "solid_tiny", relative_to=str(tmp_path))
    out = tmp_path / "out"
    r = subprocess.run([msfdm._lib.BIN_PATH, os.path.relpath(j, tmp_path), "out"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for shot in (0, 1):
        rec = out / str(shot) / "record"
        assert sorted(os.listdir(rec)) == ["solid_vz.bin"]
        assert sorted(os.listdir(out / str(shot) / "snapshot")) == ["solid"]
        snaps = sorted(os.listdir(out / str(shot) / "snapshot" / "solid" / "vz"))
        assert snaps == ["vz_000000.bin", "vz_000050.bin", "vz_000100.bin"]
        s = msfdm.Solver(m)
        s.reset(shot)
        s.run(120)
        (vz,) = s.records()
        got = np.fromfile(rec / "solid_vz.bin", dtype="<f4").reshape(-1, m.nx)
        assert np.array_equal(got, vz)


@pytest.mark.parametrize("kind,row", [("solid_cpml_tiny", 30), ("solid_diri_tiny", None)])
def test_solid_cli_record_rows(kind, row, tmp_path):
    """CLI records of the other solid-only kinds (src/main.cu:120-128, 199-209, 303-311):
    vz of row 30 with PML on four sides; an empty solid_vz.bin for Dirichlet edges."""
    m = modelgen.preset(kind, nt=60, snapshot=1000)
    j = modelgen.write_model(m, str(tmp_path / "model"), kind, relative_to=str(tmp_path))
    r = subprocess.run([msfdm._lib.BIN_PATH, os.path.relpath(j, tmp_path), "out"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rec = tmp_path / "out" / "0" / "record"
    assert sorted(os.listdir(rec)) == ["solid_vz.bin"]
    got = np.fromfile(rec / "solid_vz.bin", dtype="<f4")
    if row is None:
        assert got.size == 0
        return
    s = msfdm.Solver(m)
    s.reset(0)
    s.run(60)
    (vz,) = s.records()
    assert np.array_equal(got.reshape(-1, m.nx), vz)
    o = pyoracle.Oracle(m)
    o.reset(0)
    o.run(60)
    assert rel_l2(vz, o.records()[0]) <= RECORD_TOL


@pytest.mark.parametrize("tile_rows", [8, 16])
def test_solid_source_footprint_across_interior_tiles(tile_rows):
    """A tile under the source's 5x5 footprint leaves the interior path, which
    carries no source term: footprints across the column-tile boundary 127 | 128
    and the band boundary at row 19 (bands start at rows 3 + k tile_rows) inside
    the interior of the 301 x 161 grid are applied in full."""
    m = modelgen.preset("solid_flat301", nt=300)
    shots = [(127, 19), (191, 35)]
    m.src_x = [x for x, _ in shots]
    m.src_z = [z for _, z in shots]
    for shot in range(len(shots)):
        print(f"[solid footprint] tile_rows={tile_rows} shot at (ix, iz) = {shots[shot]}", flush=True)
        o = pyoracle.Oracle(m)
        s = msfdm.Solver(m)
        s.tile_rows = tile_rows
        o.reset(shot); o.run(150)
        s.reset(shot); s.run(150)
        for f in ("SVZ", "SSX"):
            e = rel_l2(s.get(f), o.get(f))
            assert e <= 1e-4, (shot, f, e)
        o.run(150)
        s.run(150)
        (ro,) = o.records()
        (rg,) = s.records()
        e = rel_l2(rg, ro)
        assert e <= 1e-4, (shot, e)


def test_solid_topo2001_whole_record_matches_oracle():
    """The solid-only benchmark workload as the benchmark runs it: 2001x1001
    under the sinusoidal free surface, all 6000 steps (1500 surface records),
    against the FP64 oracle (1.5-2 minutes on 16 host cores)."""
    m = modelgen.preset("solid_topo2001")
    print(f"[solid records] solid_topo2001 {m.nx}x{m.nz} {m.nt} steps", flush=True)
    s = msfdm.Solver(m)
    s.reset(0)
    s.run(m.nt)
    o = pyoracle.Oracle(m)
    o.reset(0)
    for done in range(0, m.nt, 500):
        o.run(min(500, m.nt - done))
        print(f"  oracle at step {min(done + 500, m.nt)}", flush=True)
    s.sync()
    (ro,) = o.records()
    (rg,) = s.records()
    assert rg.shape == (1500, 2001)
    err = rel_l2(rg, ro)
    print(f"  vz(row 3) rel L2 {err:.3e}", flush=True)
    assert err <= RECORD_TOL
    for name in ("SVX", "SVZ", "SSX", "SSZ", "SSXZ"):
        e = rel_l2(s.get(name), o.get(name))
        print(f"  {name}: rel L2 {e:.3e}", flush=True)
        assert e <= RECORD_TOL, name


def test_solid_flat301_full_record_matches_oracle():
    """The 301x161 solid-only free-surface workload over its whole record (1500
    steps, 750 surface records).  The final fields are the small residual the PML
    leaves: they are held to 1e-3, the record to 1e-4."""
    m = modelgen.preset("solid_flat301")
    print(f"[solid records] solid_flat301 {m.nt} steps", flush=True)
    o, s = _run_both(m, m.nt)
    (ro,) = o.records()
    (rg,) = s.records()
    assert rg.shape == (750, 301)
    err = rel_l2(rg, ro)
    print(f"  vz(row 3) rel L2 {err:.3e}", flush=True)
    assert err <= RECORD_TOL
    for name in ("SVX", "SVZ", "SSX", "SSZ", "SSXZ"):
        assert rel_l2(s.get(name), o.get(name)) <= 10 * RECORD_TOL, name
