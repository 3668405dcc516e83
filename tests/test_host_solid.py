"""Host-side checks of solid-only models (grid.calc_type = 0 with a traction-free
top, PmlFs::SolidOnly) that need no GPU: the JSON layout without a grid.fluid
block (src/elements.cu:87-99, 153-154), the model description, and the error
paths of the C ABI."""
import json
import os

import numpy as np
import pytest

import msfdm
from msfdm import modelgen
from msfdm._lib import EINVAL


@pytest.fixture()
def solid_dir(tmp_path, monkeypatch):
    m = modelgen.preset("solid_tiny")
    jpath = modelgen.write_model(m, str(tmp_path / "models" / "solid"), "solid", relative_to=str(tmp_path))
    monkeypatch.chdir(tmp_path)   # CWD-relative file paths, as in the reference
    return m, jpath


def test_solid_json_has_no_fluid_block_and_round_trips(solid_dir):
    m, jpath = solid_dir
    doc = json.load(open(jpath))
    assert doc["grid"]["calc_type"] == 0 and doc["grid"]["boundary_type"] == 2
    assert "fluid" not in doc["grid"]
    hm = msfdm.HostModel.from_json(jpath)
    d = hm.desc()
    assert (d.calc_type, d.nx, d.nz_fluid, d.nz_solid) == (0, m.nx, 0, m.nz_solid)
    assert [d.src_z[i] for i in range(d.nsrc)] == m.src_z
    for name in ("C11", "C13", "C33", "C55", "rho_s", "J", "xi_x", "xi_z", "ze_x", "ze_z"):
        a = hm.array(name)
        assert a.shape == (m.nz_solid, m.nx), name
        assert np.array_equal(a, m.arrays[name]), name
    assert hm.array("K").shape == (0, m.nx)
    back = modelgen.read_model(jpath)
    assert back.calc_type == 0 and back.nz == m.nz_solid


def test_solid_metrics_follow_the_topography():
    """The generated grid is body-fitted: row 3 follows z = sum A sin(2 pi x / L)
    (J = h z_zeta, ze_x = -z_xi / J, exact for the closed-form map)."""
    m = modelgen.preset("solid_tiny")
    h = 5.0
    x = h * np.arange(m.nx)
    top = 10.0 * np.sin(2 * np.pi * x / 200.0) + 3.0 * np.sin(2 * np.pi * x / 70.0)
    dtop = (10.0 * 2 * np.pi / 200.0 * np.cos(2 * np.pi * x / 200.0)
            + 3.0 * 2 * np.pi / 70.0 * np.cos(2 * np.pi * x / 70.0))
    J = m.arrays["J"][3].astype(np.float64)
    zx = m.arrays["ze_x"][3].astype(np.float64)
    # on row 3 the envelope is 1: z_xi = h * dtop/dx
    assert np.allclose(-zx * J, h * dtop, rtol=1e-5, atol=1e-6)
    assert np.abs(top).max() > 5.0
    # below the bottom-PML edge the grid is flat again
    b = m.nz_solid - 4 - m.npd
    assert np.allclose(m.arrays["ze_x"][b + 1:], 0.0)
    assert np.allclose(m.arrays["J"][b + 1:], h * h, rtol=1e-6)


def _desc_edit(jpath, fn):
    doc = json.load(open(jpath))
    fn(doc)
    json.dump(doc, open(jpath, "w"))


def test_solid_error_paths(solid_dir):
    m, jpath = solid_dir

    def expect(fn, code, text):
        _desc_edit(jpath, fn)
        with pytest.raises(msfdm.MsfdmError) as e:
            msfdm.HostModel.from_json(jpath)
        assert e.value.code == code and text in str(e.value), str(e.value)

    orig = json.load(open(jpath))
    expect(lambda d: d["grid"]["src_pos"]["posz"].__setitem__(0, m.nz_solid), EINVAL, "outside the solid grid")
    json.dump(orig, open(jpath, "w"))
    expect(lambda d: d["grid"].__setitem__("boundary_type", 3), EINVAL, "boundary_type")
    json.dump(orig, open(jpath, "w"))
    expect(lambda d: d["PML"].__setitem__("npd", m.nz_solid - 9), EINVAL, "too small")
    json.dump(orig, open(jpath, "w"))
    assert msfdm.HostModel.from_json(jpath).desc().calc_type == 0


def test_solid_batch_is_rejected_before_any_device_work():
    m = modelgen.preset("solid_tiny")
    with pytest.raises(msfdm.MsfdmError) as e:
        msfdm.Batch([m])
    assert e.value.code == EINVAL and "solid-only" in str(e.value)


def test_solid_boundary_kinds_and_their_limits():
    """Solid-only grids accept boundary_type 2 (free top), 1 (PML on four sides:
    the strips must not meet and the receiver row 30 must be updated) and 0
    (Dirichlet: no PML, npd unused); a fluid-solid grid boundary_type 2 and 1."""
    for bt, nz_ok, nz_bad in ((1, 34, 33), (0, 10, 9), (2, 20, 19)):
        m = modelgen.build_solid_model(nx=60, nz=60, h=5.0, npd=10, fpeak=20.0, dt=4e-4, nt=10, src_x=[30],
                                       src_z=[8], boundary_type=bt)
        for nz, ok in ((nz_ok, True), (nz_bad, False)):
            import dataclasses
            a = {k: (v[:nz] if v.shape[0] == 60 else v) for k, v in m.arrays.items()}
            mm = dataclasses.replace(m, nz_solid=nz, arrays=a)
            if ok:
                assert msfdm.HostModel.from_model(mm).desc().boundary_type == bt
            else:
                with pytest.raises(msfdm.MsfdmError) as e:
                    msfdm.HostModel.from_model(mm)
                assert e.value.code == EINVAL and "too small" in str(e.value)
    import dataclasses
    fs = modelgen.preset("tiny_curvi")
    with pytest.raises(msfdm.MsfdmError) as e:   # Dirichlet edges serve solid-only grids
        msfdm.HostModel.from_model(dataclasses.replace(fs, boundary_type=0))
    assert e.value.code == EINVAL and "supported" in str(e.value)
    assert msfdm.HostModel.from_model(dataclasses.replace(fs, boundary_type=1)).desc().boundary_type == 1
    with pytest.raises(ValueError):
        modelgen.build_solid_model(nx=60, nz=60, h=5.0, npd=10, fpeak=20.0, dt=4e-4, nt=10, src_x=[30],
                                   src_z=[8], boundary_type=1, topography=[(5.0, 100.0)])
