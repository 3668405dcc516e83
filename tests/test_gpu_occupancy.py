"""The stage kernels' occupancy policies (msfdm_set_occupancy) differ only in
the resident waves per SIMD the compiler is held to, so every policy must give
the same bits: every field of the state and every record after a run, on a
fluid-solid grid and on solid-only grids with both top conditions, at both
tile heights.  The autotuner picks a (tile height, policy) pair that the
solver then runs with."""
import numpy as np
import pytest

import msfdm
from msfdm import modelgen

pytestmark = pytest.mark.gpu


def _state(s):
    out = {}
    for f in msfdm.FIELDS:
        try:
            out[f] = s.get(f)
        except msfdm.MsfdmError:   # a memory field the grid does not store
            pass
    return out


@pytest.mark.parametrize("preset", ["tiny_curvi", "solid_tiny", "solid_cpml_tiny"])
@pytest.mark.parametrize("tile_rows", [8, 16])
def test_occupancy_policies_agree_bitwise(preset, tile_rows):
    nsteps = 150
    print(f"[occupancy] {preset} tile_rows={tile_rows} {nsteps} steps", flush=True)
    m = modelgen.preset(preset)
    runs = []
    for policy in (0, 1):
        s = msfdm.Solver(m)
        s.tile_rows = tile_rows
        s.occupancy = policy
        assert s.occupancy == policy
        s.reset(0)
        s.run(nsteps)
        s.sync()
        runs.append((_state(s), [np.array(r) for r in s.records()]))
        s.close()
    (f0, r0), (f1, r1) = runs
    assert f0.keys() == f1.keys() and len(f0) > 0
    for k in f0:
        assert np.array_equal(f0[k], f1[k]), k
    assert any(np.any(v != 0) for v in f0.values())
    for a, b in zip(r0, r1):
        assert np.array_equal(a, b)


def test_occupancy_policy_rules_and_autotune():
    print("[occupancy] policy checks and autotune on the 401 x 201 flat seafloor", flush=True)
    s = msfdm.Solver(modelgen.preset("flat401"))
    with pytest.raises(msfdm.MsfdmError):
        s.occupancy = 2
    rows = s.autotune(32)
    assert rows in (8, 16) and s.tile_rows == rows
    assert s.occupancy in (0, 1)
    s.close()
    c = msfdm.Solver(modelgen.preset("cpml_fs_tiny"))   # fluid CPML grid: default kernels only
    with pytest.raises(msfdm.MsfdmError):
        c.occupancy = 1
    assert c.autotune(32) in (8, 16) and c.occupancy == 0
    c.close()
