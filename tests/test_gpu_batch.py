"""GPU: batched shots (one launch per RK stage for all shots) are bit-identical
to stepping each shot alone, for uneven grid sizes, both tile heights, graph
replay on and off, both occupancy policies; the autotuner returns a valid tile
height and policy."""
import numpy as np
import pytest

import msfdm
from msfdm import modelgen

pytestmark = pytest.mark.gpu


def _shots():
    return modelgen.batch_shots(n=4, seed=5, nx_range=(90, 200), nz_range=(70, 120), nt=150, npd=10,
                                water_depth=100.0)


@pytest.mark.parametrize("tile_rows,graph,occupancy", [(8, True, 0), (16, True, 0), (16, False, 0), (8, True, 1)])
def test_batch_matches_single_shots_bitwise(tile_rows, graph, occupancy):
    models = _shots()
    print(f"[batch] sizes {[(m.nx, m.nz) for m in models]} tile_rows={tile_rows} graph={graph} "
          f"occupancy={occupancy}", flush=True)
    b = msfdm.Batch(models)
    b.tile_rows = tile_rows
    b.occupancy = occupancy   # the single shots below keep policy 0: same bits either way
    assert b.occupancy == occupancy
    b.set_graph(graph)
    b.reset()
    b.run(101)
    b.sync()
    for i, m in enumerate(models):
        s = msfdm.Solver(m)
        s.tile_rows = tile_rows
        s.reset(0)
        s.run(101)
        sb = b.shot(i)
        for a, c in zip(sb.records(), s.records()):
            assert np.array_equal(a, c), i
        for name in ("FP", "SVZ", "SSXZ", "SXEX", "SZVZ"):
            assert np.array_equal(sb.get(name), s.get(name)), (i, name)


def test_batch_reset_and_autotune():
    models = _shots()
    b = msfdm.Batch(models)
    best = b.autotune(32)
    assert best in (8, 16) and b.tile_rows == best and b.occupancy in (0, 1)
    with pytest.raises(msfdm.MsfdmError):
        b.occupancy = 2
    b.run(40); b.sync()
    r1 = [x.copy() for x in b.shot(2).records()]
    b.reset(); b.run(40); b.sync()
    for x, y in zip(r1, b.shot(2).records()):
        assert np.array_equal(x, y)
    s = msfdm.Solver(models[0])
    assert s.autotune(32) in (8, 16)
