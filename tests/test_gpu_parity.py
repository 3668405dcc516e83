"""GPU parity: the HIP stepper (through the C ABI) against the FP64 oracle.

Tolerances: per-field right-hand sides on random states within 2e-5 of the
field's max |k| (FP32 rounding of ~10-term stencil sums); seismogram records
within a relative L2 error of 1e-4 over the whole record per component
(vx, vz on solid row 0; p on the fluid interface row).
"""
import numpy as np
import pytest

import msfdm
from msfdm import modelgen
import pyoracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

RECORD_TOL = 1e-4


def _random_state(m, seed):
    rng = np.random.default_rng(seed)
    st = {}
    for i, name in enumerate(msfdm.FIELDS):
        rows = m.nz_fluid if msfdm._lib.is_fluid(i) else m.nz_solid
        st[name] = rng.standard_normal((rows, m.nx))
    return st


@pytest.mark.parametrize("preset", ["tiny_curvi", "cut201"])
@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("tile_rows", [8, 16])
def test_rhs_matches_oracle_on_random_state(preset, direction, tile_rows):
    print(f"[rhs] {preset} dir={direction} tile_rows={tile_rows}", flush=True)
    m = modelgen.preset(preset)
    o = pyoracle.Oracle(m)
    s = msfdm.Solver(m)
    s.tile_rows = tile_rows
    st = _random_state(m, 11 + direction)
    for name, a in st.items():
        a32 = a.astype(np.float32)
        o.set(name, a32.astype(np.float64))
        s.set(name, a32)
    amp = 0.731
    o.rhs(direction, amp)
    s.eval_rhs(direction, amp)
    for name in msfdm.FIELDS:
        ko = o.get_k(name)
        kg = s.get_rhs(name).astype(np.float64)
        scale = np.abs(ko).max()
        if scale == 0:
            assert np.abs(kg).max() == 0, name
            continue
        err = np.abs(kg - ko).max() / scale
        assert err < 2e-5, f"{name}: max err {err:.3e} (scale {scale:.3e})"


@pytest.mark.parametrize("water_depth", [20.0, 25.0, 110.0, 115.0, 120.0])
@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("tile_rows", [8, 16])
def test_rhs_fluid_band_heights(water_depth, direction, tile_rows):
    """The fluid bands end at the interface row, so the top band (with the
    surface rows) holds 1..tile_rows rows depending on nz_fluid: 26, 27 and 28
    fluid rows give top bands of 7, 8 and 1 rows at 8-row tiles (7, 8 and 9 at
    16); 8 and 9 fluid rows (the thinnest water layers accepted) give a single
    band that starts above the grid (rows -8.. at 16-row tiles).  The raw RHS
    matches the oracle on every cell for each."""
    m = modelgen.preset("cut201", water_depth=water_depth)
    print(f"[rhs bands] nz_fluid={m.nz_fluid} dir={direction} tile_rows={tile_rows}", flush=True)
    o = pyoracle.Oracle(m)
    s = msfdm.Solver(m)
    s.tile_rows = tile_rows
    st = _random_state(m, 21 + direction)
    for name, a in st.items():
        a32 = a.astype(np.float32)
        o.set(name, a32.astype(np.float64))
        s.set(name, a32)
    o.rhs(direction, 0.731)
    s.eval_rhs(direction, 0.731)
    for name in msfdm.FIELDS:
        ko = o.get_k(name)
        kg = s.get_rhs(name).astype(np.float64)
        scale = np.abs(ko).max()
        if scale == 0:
            assert np.abs(kg).max() == 0, name
            continue
        err = np.abs(kg - ko).max() / scale
        assert err < 2e-5, f"{name}: max err {err:.3e} (scale {scale:.3e})"


def _general_metrics(m, seed=1):
    """Copy of m with xi_x, xi_z varying in space (exercises the general-metric
    path; the oracle reads the same arrays)."""
    import dataclasses
    rng = np.random.default_rng(seed)
    a = dict(m.arrays)
    nz, nx = a["xi_x"].shape
    zz, xx = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    a["xi_x"] = (a["xi_x"] * (1.0 + 0.03 * np.sin(xx / 9.0) * np.cos(zz / 7.0))).astype(np.float32)
    a["xi_z"] = (0.02 * a["xi_x"] * np.sin(zz / 11.0 + xx / 13.0)).astype(np.float32)
    return dataclasses.replace(m, arrays=a)


@pytest.mark.parametrize("direction", [0, 1])
def test_rhs_general_metrics(direction):
    m = _general_metrics(modelgen.preset("tiny_curvi"))
    o = pyoracle.Oracle(m)
    s = msfdm.Solver(m)
    st = _random_state(m, 21 + direction)
    for name, a in st.items():
        a32 = a.astype(np.float32)
        o.set(name, a32.astype(np.float64))
        s.set(name, a32)
    o.rhs(direction, 0.3)
    s.eval_rhs(direction, 0.3)
    for name in msfdm.FIELDS:
        ko = o.get_k(name)
        kg = s.get_rhs(name).astype(np.float64)
        scale = np.abs(ko).max()
        assert np.abs(kg - ko).max() <= 2e-5 * max(scale, 1e-30), name


def test_records_general_metrics():
    m = _general_metrics(modelgen.preset("tiny_curvi"))
    o, s = _run_both(m, 200)
    for a, b in zip(s.records(), o.records()):
        assert rel_l2(a, b) <= RECORD_TOL


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


@pytest.mark.parametrize("preset,nsteps,tile_rows", [("tiny_curvi", 300, 8), ("tiny_curvi", 300, 16),
                                                     ("cut201", 600, 16)])
def test_records_match_oracle(preset, nsteps, tile_rows):
    print(f"[records] {preset} nsteps={nsteps} tile_rows={tile_rows}", flush=True)
    m = modelgen.preset(preset)
    o, s = _run_both(m, nsteps, tile_rows=tile_rows)
    ro = o.records()
    rg = s.records()
    assert s.num_records() == o.records()[0].shape[0]
    for name, a, b in zip(("solid_vx", "solid_vz", "fluid_p"), rg, ro):
        e = rel_l2(a, b)
        assert e <= RECORD_TOL, f"{name}: rel L2 {e:.3e}"
    for name in ("FVX", "FVZ", "FP", "SVX", "SVZ", "SSX", "SSZ", "SSXZ", "SXVX", "SZEX"):
        e = rel_l2(s.get(name), o.get(name))
        assert e <= 10 * RECORD_TOL, f"{name}: rel L2 {e:.3e}"


def test_flat401_full_record_matches_oracle():
    """The 401x201 flat-seafloor workload, all 2000 steps (1000 records)."""
    print("[records] flat401 2000 steps", flush=True)
    m = modelgen.preset("flat401")
    o, s = _run_both(m, m.nt)
    ro, rg = o.records(), s.records()
    assert rg[0].shape == (1000, 401)
    for name, a, b in zip(("solid_vx", "solid_vz", "fluid_p"), rg, ro):
        e = rel_l2(a, b)
        print(f"  {name}: rel L2 {e:.3e}", flush=True)
        assert e <= RECORD_TOL, f"{name}: rel L2 {e:.3e}"


def test_seafloor2001_full_grid_matches_oracle():
    """The 2001x1001 body-fitted seafloor workload at its full size: 160 steps
    (the source sits 4 cells above the seafloor, so the wave crosses the
    interface, the records' rows and the random solid), 40 records, and every
    field of the final state."""
    m = modelgen.preset("seafloor2001", nt=160)
    print(f"[records] seafloor2001 {m.nx}x{m.nz} 160 steps", flush=True)
    o, s = _run_both(m, m.nt)
    ro, rg = o.records(), s.records()
    assert rg[0].shape == (40, 2001)
    for name, a, b in zip(("solid_vx", "solid_vz", "fluid_p"), rg, ro):
        assert np.abs(b).max() > 0, name
        e = rel_l2(a, b)
        print(f"  {name}: rel L2 {e:.3e}", flush=True)
        assert e <= RECORD_TOL, f"{name}: rel L2 {e:.3e}"
    for name in ("FVX", "FVZ", "FP", "SVX", "SVZ", "SSX", "SSZ", "SSXZ"):
        e = rel_l2(s.get(name), o.get(name))
        print(f"  {name}: rel L2 {e:.3e}", flush=True)
        assert e <= 10 * RECORD_TOL, f"{name}: rel L2 {e:.3e}"


def test_second_shot_and_reset():
    m = modelgen.preset("tiny_curvi")
    o, s = _run_both(m, 120, isrc=1)
    for a, b in zip(s.records(), o.records()):
        assert rel_l2(a, b) <= RECORD_TOL
    # a reset brings the solver back to the first shot's result bit for bit
    s.reset(0); s.run(64); s.sync(); first = [r.copy() for r in s.records()]
    s.reset(1); s.run(10); s.reset(0); s.run(64); s.sync()
    for a, b in zip(first, s.records()):
        assert np.array_equal(a, b)


def test_graph_and_direct_launch_agree_bitwise():
    m = modelgen.preset("tiny_curvi")
    s1 = msfdm.Solver(m); s1.set_graph(True); s1.reset(0); s1.run(101)
    s2 = msfdm.Solver(m); s2.set_graph(False); s2.reset(0); s2.run(101)
    for name in msfdm.FIELDS:
        assert np.array_equal(s1.get(name), s2.get(name)), name
    for a, b in zip(s1.records(), s2.records()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("preset", ["tiny_curvi", "cut201", "cpml_fs_tiny"])
def test_chunked_runs_agree_bitwise(preset):
    """Inside one run call the interface rows of a step are finished by the next
    step's stage-1 launch, after its last step by the interface kernel; runs cut
    at other steps (the CLI's snapshot steps, single steps, odd starts, graph on
    and off) must give the same bits."""
    m = modelgen.preset(preset, nt=90)
    print(f"[chunks] {preset} {m.nx}x{m.nz}", flush=True)

    def run(chunks, graph=True):
        s = msfdm.Solver(m); s.set_graph(graph); s.reset(0)
        for c in chunks:
            s.run(c)
        s.sync()
        return [s.get(n) for n in msfdm.FIELDS] + list(s.records())

    ref = run([90])
    for chunks, graph in (([1, 40, 40, 9], True), ([1] * 90, True), ([45, 45], False), ([33, 57], True)):
        for a, b in zip(ref, run(chunks, graph)):
            assert np.array_equal(a, b), (chunks, graph)


def test_repeat_runs_are_bitwise_deterministic():
    m = modelgen.preset("cut201")
    s = msfdm.Solver(m)
    out = []
    for _ in range(2):
        s.reset(0); s.run(200); s.sync()
        out.append([s.get(n) for n in msfdm.FIELDS] + list(s.records()))
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_interface_and_surface_invariants():
    m = modelgen.preset("tiny_curvi")
    s = msfdm.Solver(m)
    s.reset(0)
    s.run(150)
    p = s.get("FP")
    assert np.all(p[3] == 0.0), "pressure-release row must stay exactly zero"
    for name in msfdm.FIELDS:
        a = s.get(name)
        if name.startswith("F"):
            assert np.all(a[:3] == 0.0), name
        else:
            assert np.all(a[-3:] == 0.0), name
        assert np.all(a[:, :3] == 0.0) and np.all(a[:, -3:] == 0.0), name
    # after the filter: v_n and p = -sigma_nn agree across the interface
    nzf = m.nz_fluid
    zx, zz = m.arrays["ze_x"][nzf - 1].astype(np.float64), m.arrays["ze_z"][nzf - 1].astype(np.float64)
    nn = np.hypot(zx, zz); n1, n2 = zx / nn, zz / nn
    c = slice(3, m.nx - 3)
    vnf = n1 * s.get("FVX")[nzf - 1] + n2 * s.get("FVZ")[nzf - 1]
    vns = n1 * s.get("SVX")[0] + n2 * s.get("SVZ")[0]
    assert np.abs(vnf - vns)[c].max() <= 1e-5 * np.abs(vns).max()
    sxx, szz, sxz = s.get("SSX")[0], s.get("SSZ")[0], s.get("SSXZ")[0]
    snn = n1 * n1 * sxx + 2 * n1 * n2 * sxz + n2 * n2 * szz
    assert np.abs(p[nzf - 1] + snn)[c].max() <= 1e-5 * np.abs(snn).max()


def test_state_roundtrip_layout():
    m = modelgen.preset("tiny_curvi")
    s = msfdm.Solver(m)
    rng = np.random.default_rng(3)
    for name in ("FP", "SSXZ", "FXEX", "SZVZ"):
        a = rng.standard_normal(s.shape(msfdm.FID[name])).astype(np.float32)
        s.set(name, a)
        b = s.get(name)
        if name in ("FXEX",):
            npd = m.npd
            cols = list(range(3, 3 + npd)) + list(range(m.nx - 3 - npd, m.nx - 3))
            assert np.array_equal(b[:, cols], a[:, cols])
        elif name == "SZVZ":
            z0 = m.nz_solid - 3 - m.npd
            assert np.array_equal(b[z0:z0 + m.npd], a[z0:z0 + m.npd])
        else:
            assert np.array_equal(a, b)


def test_curved_seafloor_long_run_matches_oracle():
    """6000 steps (2000 records) on the small curved-seafloor, random-solid grid.
    Over long runs a curved interface is where the reference's float interface
    normal and its float filter normalisation show; the solver and the oracle
    both carry them."""
    m = modelgen.preset("tiny_curvi", nt=6000)
    print(f"[records] tiny_curvi {m.nx}x{m.nz} 6000 steps", flush=True)
    o, s = _run_both(m, m.nt)
    ro, rg = o.records(), s.records()
    assert rg[0].shape == (2000, 120)
    for name, a, b in zip(("solid_vx", "solid_vz", "fluid_p"), rg, ro):
        e = rel_l2(a, b)
        print(f"  {name}: rel L2 {e:.3e}", flush=True)
        assert e <= RECORD_TOL, f"{name}: rel L2 {e:.3e}"
    for name in ("FVX", "FVZ", "FP", "SVX", "SVZ", "SSX", "SSZ", "SSXZ"):
        e = rel_l2(s.get(name), o.get(name))
        print(f"  {name}: rel L2 {e:.3e}", flush=True)
        assert e <= 10 * RECORD_TOL, f"{name}: rel L2 {e:.3e}"


def test_seafloor2001_whole_record_matches_oracle():
    """The benchmark workload as the benchmark runs it: 2001x1001, all 6000 steps
    (1500 records), against the FP64 oracle (about 12e9 oracle cell updates,
    1.5-2 minutes on 16 host cores)."""
    m = modelgen.preset("seafloor2001")
    print(f"[records] seafloor2001 {m.nx}x{m.nz} {m.nt} steps", flush=True)
    s = msfdm.Solver(m)
    s.reset(0)
    s.run(m.nt)
    o = pyoracle.Oracle(m)
    o.reset(0)
    for done in range(0, m.nt, 500):
        o.run(min(500, m.nt - done))
        print(f"  oracle at step {min(done + 500, m.nt)}", flush=True)
    s.sync()
    ro, rg = o.records(), s.records()
    assert rg[0].shape == (1500, 2001)
    for name, a, b in zip(("solid_vx", "solid_vz", "fluid_p"), rg, ro):
        e = rel_l2(a, b)
        print(f"  {name}: rel L2 {e:.3e}", flush=True)
        assert e <= RECORD_TOL, f"{name}: rel L2 {e:.3e}"
    for name in ("FVX", "FVZ", "FP", "SVX", "SVZ", "SSX", "SSZ", "SSXZ"):
        e = rel_l2(s.get(name), o.get(name))
        print(f"  {name}: rel L2 {e:.3e}", flush=True)
        assert e <= 10 * RECORD_TOL, f"{name}: rel L2 {e:.3e}"
