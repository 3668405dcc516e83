"""GPU parity on the source and grid edge cases of the solid-only steppers
(add_gauss_source on both normal stresses, include/tools_func.cuh:21-37 via
src/pml_fs_kernels.cu:22-27: rows <= 2 and >= nz-3 and columns <= 2 and >= nx-3
of the 5 x 5 stencil are skipped, and the add runs after the surface row's
ALM stress rates):

* on the free-surface row 3 (the surface stresses take the source: the row is
  no longer traction free),
* one and two rows below it (inside the conservative rows 3..5),
* one row above the last updated row (the lowest stencil row is lost),
* on the first / last updated column and inside the side PML,
* the CPML grid's top strip (row 8) and the Dirichlet grid's first row.

Each shot runs 200 steps; records (and the final vz) within 1e-4 relative L2 of
the FP64 oracle.  Size: the 2001 x 1001 topography grid for 30 steps.
"""
import dataclasses

import numpy as np
import pytest

import msfdm
from msfdm import modelgen
import pyoracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _edge_model():
    m = modelgen.preset("solid_tiny")
    nz, nx = m.nz_solid, m.nx
    shots = [(55, 3), (55, 4), (55, 5), (55, nz - 5), (3, 20), (nx - 4, 20), (8, 20)]
    return dataclasses.replace(m, src_x=[x for x, _ in shots], src_z=[z for _, z in shots]), shots


def _check(m, shot, nsteps):
    o = pyoracle.Oracle(m)
    s = msfdm.Solver(m)
    o.reset(shot); o.run(nsteps)
    s.reset(shot); s.run(nsteps)
    vo, vg = o.get("SVZ"), s.get("SVZ")
    assert np.abs(vo).max() > 0
    e = rel_l2(vg, vo)
    print(f"  final vz rel L2 {e:.3e}", flush=True)
    assert e <= 1e-4
    ro, rg = o.records(), s.records()
    for a, b in zip(rg, ro):
        if b.shape[0] == 0:
            assert a.shape[0] == 0
            continue
        if not np.any(b):
            assert not np.any(a)
            continue
        e = rel_l2(a, b)
        print(f"  record rel L2 {e:.3e}", flush=True)
        assert e <= 1e-4


@pytest.mark.parametrize("shot", range(7))
def test_free_surface_source_positions_match_oracle(shot):
    m, shots = _edge_model()
    print(f"[solid edge] shot {shot} at (ix, iz) = {shots[shot]}, nz = {m.nz_solid}", flush=True)
    if shots[shot][1] == 3:
        # the source enters the surface row's stress rates: not traction free any more
        s = msfdm.Solver(m)
        s.reset(shot)
        s.eval_rhs(0, 1.0)
        assert np.abs(s.get_rhs("SSZ")[3]).max() > 0
    _check(m, shot, 200)


@pytest.mark.parametrize("kind,src", [("solid_cpml_tiny", (48, 8)), ("solid_diri_tiny", (45, 3))])
def test_plain_top_source_positions_match_oracle(kind, src):
    m = dataclasses.replace(modelgen.preset(kind), src_x=[src[0]], src_z=[src[1]])
    print(f"[solid edge] {kind} source at {src}", flush=True)
    _check(m, 0, 200)


def test_topography_grid_matches_oracle():
    m = modelgen.preset("solid_topo2001", nt=30)
    print(f"[solid edge] solid_topo2001 {m.nx}x{m.nz} 30 steps", flush=True)
    _check(m, 0, 30)
