"""Host-side checks that need no GPU: the C ABI exports, the JSON/model loader
(semantics of include/json_config.hpp and MemManager::MemManager,
src/elements.cu:64-176) and every error path the reference would exit or
throw on, which the C ABI must turn into error codes."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import msfdm
from msfdm import modelgen
from msfdm._lib import EINVAL, EIO, EPARSE, EDEVICE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_entry_point():
    header = open(os.path.join(ROOT, "include", "msfdm.h")).read()
    declared = set(re.findall(r"\b(msfdm_[a-z_0-9]+)\s*\(", header))
    assert declared, "no declarations parsed"
    L = msfdm.lib()
    for name in sorted(declared):
        assert hasattr(L, name), f"{name} declared in include/msfdm.h but not exported"
    assert declared == set(msfdm._lib.EXPORTS)
    assert msfdm.lib().msfdm_version().startswith(b"msfdm")


@pytest.fixture()
def model_dir(tmp_path, monkeypatch):
    m = modelgen.preset("tiny_curvi")
    jpath = modelgen.write_model(m, str(tmp_path / "models" / "tiny"), "tiny", relative_to=str(tmp_path))
    monkeypatch.chdir(tmp_path)   # file paths in the JSON are CWD-relative, as in the reference
    return m, jpath


def test_pad_file_beside_the_model(model_dir):
    """pad.txt as the reference's generators write it: "key value" lines, the
    outer pad being the 3 frozen columns plus the PML, read the way
    tools/compare_sem_scholte.py (pad_flat) and tools/compare_double_cos.py
    (the token after "shift") read it."""
    m, jpath = model_dir
    text = open(os.path.join(os.path.dirname(jpath), "pad.txt")).read()
    kv = dict(line.split() for line in text.splitlines())
    pad = m.npd + 3
    assert int(kv["pad_flat"]) == pad and int(kv["shift"]) == pad and int(kv["outer"]) == pad
    assert int(text.split("shift")[1].split()[0]) == pad
    assert int(kv["nx_phys"]) == m.nx - 2 * pad
    assert int(kv["phys0"]) == pad and int(kv["phys1"]) == m.nx - pad - 1
    assert int(kv["nz_fluid_full"]) == m.nz_fluid and int(kv["nz_solid_full"]) == m.nz_solid
    assert int(kv["npd"]) == m.npd


def _edit_json(jpath, fn):
    doc = json.load(open(jpath))
    fn(doc)
    json.dump(doc, open(jpath, "w"))


def test_json_model_roundtrip(model_dir):
    m, jpath = model_dir
    hm = msfdm.HostModel.from_json(jpath)
    d = hm.desc()
    assert (d.nx, d.nz_fluid, d.nz_solid) == (m.nx, m.nz_fluid, m.nz_solid)
    assert d.npd == m.npd and d.npow == m.npow and d.nt == m.nt and d.snapshot == m.snapshot
    assert d.dt == np.float32(m.dt) and d.fpeak == np.float32(m.fpeak)
    assert d.beta0 == np.float32(m.beta0) and d.rc == np.float32(m.rc)
    assert [d.src_x[i] for i in range(d.nsrc)] == m.src_x
    assert [d.src_z[i] for i in range(d.nsrc)] == m.src_z
    for name in msfdm.ARRAYS:
        assert np.array_equal(hm.array(name), m.arrays[name]), name


def test_json_number_semantics(model_dir):
    """int keys take cJSON's valueint (saturating, truncating); float keys (float)double."""
    _, jpath = model_dir
    def edit(doc):
        doc["base"]["nt"] = 1234.99
        doc["base"]["snapshot"] = 1e12
        doc["base"]["dt"] = 0.0004000000000000001
        doc["PML"]["NPOW"] = 2.7
    _edit_json(jpath, edit)
    d = msfdm.HostModel.from_json(jpath).desc()
    assert d.nt == 1234
    assert d.snapshot == 2**31 - 1
    assert d.dt == np.float32(0.0004)
    assert d.npow == 2


@pytest.mark.parametrize("mutate,code,needle", [
    (lambda p: open(p, "w").write("{ \"base\": { \"dt\": 1e-4, }"), EPARSE, "parse"),
    (lambda p: open(p, "w").write(""), EPARSE, "empty"),
    (lambda p: _edit_json(p, lambda d: d["base"].pop("dt")), EPARSE, "JSON path not found: base.dt"),
    (lambda p: _edit_json(p, lambda d: d["base"].__setitem__("nt", "2000")), EPARSE, "expected int"),
    (lambda p: _edit_json(p, lambda d: d["PML"].__setitem__("Rc", [1e-3])), EPARSE, "expected float"),
    (lambda p: _edit_json(p, lambda d: d["grid"]["solid"].__setitem__("C11", 5)), EPARSE, "expected string"),
    (lambda p: _edit_json(p, lambda d: d["grid"]["src_pos"].__setitem__("posz", [])), EPARSE, "out of range"),
    (lambda p: _edit_json(p, lambda d: d["grid"]["fluid"].__setitem__("nx", 7)), EINVAL, "nx_fluid doesn't match nx_solid"),
    (lambda p: _edit_json(p, lambda d: d["grid"].__setitem__("calc_type", 2)), EINVAL, "calc_type"),
    (lambda p: _edit_json(p, lambda d: d["grid"].__setitem__("boundary_type", 0)), EINVAL, "boundary_type"),
    (lambda p: _edit_json(p, lambda d: d["grid"]["jacob"].__setitem__("J", "nope.bin")), EIO, "Failed to open file: nope.bin"),
    (lambda p: _edit_json(p, lambda d: d["grid"]["src_pos"].__setitem__("posx", [100000])), EINVAL, "outside"),
])
def test_json_errors_are_codes_not_exits(model_dir, mutate, code, needle):
    _, jpath = model_dir
    mutate(jpath)
    with pytest.raises(msfdm.MsfdmError) as e:
        msfdm.HostModel.from_json(jpath)
    assert e.value.code == code
    assert needle in str(e.value)


def test_short_array_file_is_an_io_error(model_dir):
    m, jpath = model_dir
    doc = json.load(open(jpath))
    path = doc["grid"]["solid"]["rho"]
    data = np.fromfile(path, dtype="<f4")
    data[:-5].tofile(path)
    with pytest.raises(msfdm.MsfdmError) as e:
        msfdm.HostModel.from_json(jpath)
    assert e.value.code == EIO and "Failed to read data from file" in str(e.value)


def test_missing_json_file():
    with pytest.raises(msfdm.MsfdmError) as e:
        msfdm.HostModel.from_json("/nonexistent/model.json")
    assert e.value.code == EPARSE and "Cannot open JSON file" in str(e.value)


def test_model_create_validates_shapes_and_sources():
    m = modelgen.preset("tiny_curvi")
    bad = modelgen.preset("tiny_curvi", src_x=[500], src_z=[10])
    with pytest.raises(msfdm.MsfdmError) as e:
        msfdm.HostModel.from_model(bad)
    assert e.value.code == EINVAL
    hm = msfdm.HostModel.from_model(m)
    assert hm.desc().nsrc == 2


@pytest.mark.parametrize("out", ["", "/", ".", ".."])
def test_cli_refuses_dangerous_output_dirs(out, model_dir):
    _, jpath = model_dir
    assert msfdm.run_cli(jpath, out) == 1
    assert "Invalid output dir" in msfdm.last_error()


def test_cli_failure_before_solver_keeps_output_dir(model_dir, tmp_path):
    """Without a usable device the run fails before the output root is cleared."""
    import ctypes
    _, jpath = model_dir
    out = tmp_path / "out"
    out.mkdir()
    (out / "keep.txt").write_text("x")
    n = ctypes.c_int(0)
    try:
        has_gpu = ctypes.CDLL("libamdhip64.so").hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        has_gpu = False
    if has_gpu:
        pytest.skip("a HIP device is present; the GPU tests cover the CLI run")
    assert msfdm.run_cli(jpath, str(out)) == 1
    assert "no HIP device" in msfdm.last_error()
    assert (out / "keep.txt").exists()


def test_bench_roofline_kernels_are_in_the_committed_profiles():
    """bench.py reads the measured HBM traffic of its roofline kernel from the
    committed rocprofv3 summaries by kernel name; the names must still match."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    for solid, workload in ((False, "seafloor2001"), (True, "solid_topo2001")):
        name, traffic = bench.measured_traffic(workload, bench.stage2_kernel_names(solid, 8))
        assert name is not None, workload
        assert traffic > 0
        # the occupancy copy is named as such and shares its twin's measured traffic
        name5, traffic5 = bench.measured_traffic(workload, bench.stage2_kernel_names(solid, 8, 1))
        assert name5 == name.replace("<", "_w5<", 1) and traffic5 == traffic


def test_main_debug_usage():
    if not os.path.exists(msfdm.BIN_PATH):
        pytest.skip("bin/main_debug not built")
    r = subprocess.run([msfdm.BIN_PATH], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr


def test_solver_create_without_device_is_an_error(model_dir):
    import ctypes
    n = ctypes.c_int(0)
    try:
        has_gpu = ctypes.CDLL("libamdhip64.so").hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        has_gpu = False
    if has_gpu:
        pytest.skip("a HIP device is present")
    m, _ = model_dir
    with pytest.raises(msfdm.MsfdmError) as e:
        msfdm.Solver(m)
    assert e.value.code == EDEVICE


def test_generator_metrics_are_consistent():
    """Body-fitted metrics: J = h z_zeta, ze_z = h / J, xi_x = 1/h; the interface row
    follows the requested seafloor and the bottom PML band is flat."""
    m = modelgen.preset("tiny_curvi")
    h = 5.0
    J, zx, zz = (m.arrays[k].astype(np.float64) for k in ("J", "ze_x", "ze_z"))
    assert np.allclose(zz * J, h, rtol=1e-6)
    assert np.all(m.arrays["xi_z"] == 0) and np.allclose(m.arrays["xi_x"], 1 / h)
    nz = m.nz
    band = slice(nz - 4 - m.npd, nz)
    assert np.allclose(J[band], h * h, rtol=1e-6) and np.allclose(zx[band], 0.0, atol=1e-9)
    # z_xi = -ze_x J follows the derivative of the seafloor shape on the interface row
    zeta0 = m.nz_fluid - 1
    x = h * np.arange(m.nx)
    expect = 12.0 * np.sin(2 * np.pi * x / 200.0) + 4.0 * np.sin(2 * np.pi * x / 70.0)
    dz_dxi = -zx[zeta0] * J[zeta0]
    d_expect = np.gradient(expect, 1.0)
    assert np.abs(dz_dxi - d_expect)[2:-2].max() < 0.05 * np.abs(d_expect).max()


def test_random_field_statistics_and_determinism():
    a = modelgen.gaussian_random_field((200, 300), 2.5, 50.0, 0.05, seed=3)
    b = modelgen.gaussian_random_field((200, 300), 2.5, 50.0, 0.05, seed=3)
    c = modelgen.gaussian_random_field((200, 300), 2.5, 50.0, 0.05, seed=4)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert abs(a.std() - 0.05) < 1e-12 and abs(a.mean()) < 1e-12
    # covariance exp(-r^2/L^2), L = 50 m = 20 cells
    big = modelgen.gaussian_random_field((600, 900), 2.5, 50.0, 0.05, seed=5)
    for lag in (10, 20):
        r = np.corrcoef(big[:, :-lag].ravel(), big[:, lag:].ravel())[0, 1]
        assert abs(r - np.exp(-(lag / 20.0) ** 2)) < 0.06, lag


def test_reference_json_template_loads(tmp_path, monkeypatch):
    """The reference's own configuration template (json_template.json, kept
    unchanged as tests/golden/json_template.json) loads through the C ABI:
    every key lands where include/json_config.hpp puts it, the CWD-relative
    <f4 files of the sizes it names are read in the reference layout."""
    doc = json.load(open(os.path.join(ROOT, "tests", "golden", "json_template.json")))
    g = doc["grid"]
    nx, nzs, nzf = g["solid"]["nx"], g["solid"]["nz"], g["fluid"]["nz"]
    nz = nzf + nzs - 1
    rng = np.random.default_rng(3)
    shapes = {**{k: (nzs, nx) for k in ("C11", "C13", "C33", "C55", "rho")}}
    files = {}
    for sec, rows in (("solid", nzs), ("fluid", nzf), ("jacob", nz)):
        for key, rel in g[sec].items():
            if key == "nx" or key == "nz":
                continue
            a = rng.uniform(1.0, 2.0, (rows, nx)).astype("<f4")
            path = tmp_path / rel
            path.parent.mkdir(parents=True, exist_ok=True)
            a.tofile(path)
            files[(sec, key)] = a
    jpath = tmp_path / "json_template.json"
    json.dump(doc, open(jpath, "w"))
    monkeypatch.chdir(tmp_path)
    hm = msfdm.HostModel.from_json(str(jpath))
    d = hm.desc()
    assert (d.nx, d.nz_fluid, d.nz_solid, d.calc_type, d.boundary_type) == (501, 201, 381, 1, 2)
    assert (d.nt, d.snapshot, d.npd, d.npow) == (30001, 200, 20, 2)
    for got, want in ((d.fpeak, 30.0), (d.dt, 1e-4), (d.thickness, 47.5), (d.vp_max, 3200.0), (d.rc, 1e-3),
                      (d.beta0, 5.0)):
        assert got == np.float32(want)
    assert (d.nsrc, d.src_x[0], d.src_z[0]) == (1, 250, 60)
    names = {("solid", "C11"): "C11", ("solid", "C13"): "C13", ("solid", "C33"): "C33", ("solid", "C55"): "C55",
             ("solid", "rho"): "rho_s", ("fluid", "K"): "K", ("fluid", "rho"): "rho_f", ("jacob", "J"): "J",
             ("jacob", "xi_x"): "xi_x", ("jacob", "xi_z"): "xi_z", ("jacob", "ze_x"): "ze_x", ("jacob", "ze_z"): "ze_z"}
    for k, name in names.items():
        assert np.array_equal(hm.array(name), files[k]), name
    hm.close()


def test_record_count_follows_the_oracle_cadence():
    import pyoracle
    for dt, nt in ((5e-4, 2000), (2.5e-4, 6000), (2e-4, 8000), (4e-4, 37), (1e-4, 30001), (3e-3, 10)):
        m = modelgen.build_model(nx=40, nz=40, h=5.0, water_depth=60.0, npd=5, fpeak=20.0, dt=dt, nt=nt, src_x=[20])
        o = pyoracle.Oracle(m)
        assert msfdm._lib.record_gap(dt) == o.gap, dt
        assert msfdm._lib.num_records(nt, dt) == (nt - 1) // o.gap + 1


def test_residual_shapes_are_checked():
    """A residual that is not [nrec][nx] of the run (a PML-cut gather, a shorter
    record, two planes) is refused before any device call."""
    check = msfdm._lib.check_residuals
    nrec = msfdm._lib.num_records(600, 5e-4)
    assert nrec == 300
    ok = [np.zeros((nrec, 201), np.float64) for _ in range(3)]
    out = check(ok, 600, 5e-4, 201)
    assert all(r.dtype == np.float32 and r.shape == (nrec, 201) for r in out)
    for bad in ([np.zeros((nrec, 161))] * 3, [np.zeros((nrec - 1, 201))] * 3, ok[:2], [np.zeros(nrec * 201)] * 3):
        with pytest.raises(ValueError):
            check(bad, 600, 5e-4, 201)
