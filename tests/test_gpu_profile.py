"""GPU: msfdm_profile_run steps exactly as msfdm_run does (bitwise equal state
and records, the launch sequence of the graph replays) and times every launch
by its own dispatch."""
import numpy as np
import pytest

import msfdm
from msfdm import modelgen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("preset", ["tiny_curvi", "solid_tiny"])
def test_profile_run_equals_run(preset):
    m = modelgen.preset(preset)
    a, b = msfdm.Solver(m), msfdm.Solver(m)
    a.reset(0)
    a.run(70)
    b.reset(0)
    p = b.profile_run(70)
    for x, y in zip(a.records(), b.records()):
        assert np.array_equal(x, y)
    for name in ("SVX", "SVZ", "SSX", "SSZ", "SSXZ") + (() if b.solid_only else ("FVX", "FVZ", "FP")):
        assert np.array_equal(a.get(name), b.get(name)), name
    n = p["launches"]
    # steps 0..69 from an even start: sequences of 32, 32 and 6 steps
    assert n[:4] == [70, 70, 70, 70]
    assert n[4] == (0 if b.solid_only else 3) and n[5] == 3
    assert all(t > 0 for t in p["kernel_ms"][:4])
    assert p["wall_ms"] >= 0.99 * sum(p["kernel_ms"])
    # only the requested classes are timed; the stepping is the same
    c = msfdm.Solver(m)
    c.reset(0)
    q = c.profile_run(70, classes=[1])
    assert q["launches"] == [0, 70, 0, 0, 0, 0] and q["kernel_ms"][1] > 0
    for x, y in zip(a.records(), c.records()):
        assert np.array_equal(x, y)


def test_stream_probe_rates():
    """The streaming probe's rates are positive and below what the part can move
    (1 GB streams do not fit in any cache); bad arguments raise."""
    print("[stream-probe] copy and read3-write1 at 1000 MB", flush=True)
    copy = msfdm.stream_probe(1000, 1, reps=5)
    triad = msfdm.stream_probe(1000, 3, reps=5)
    print(f"  copy {copy:.2f} TB/s, read3+write1 {triad:.2f} TB/s", flush=True)
    for v in (copy, triad):
        assert 0.5 < v < 8.0, v
    with pytest.raises(msfdm.MsfdmError):
        msfdm.stream_probe(1000, 5)
    with pytest.raises(msfdm.MsfdmError):
        msfdm.stream_probe(0, 1)
