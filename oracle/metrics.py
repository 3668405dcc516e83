"""FP64 restatement of the reference's finite-difference metric block.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU metric
kernel (msfdm_grid_metrics), never by the product path.

Follows tools/run_models/model_ortho_classic_scholte.py:188-195 (the same block
appears in model_Zhao_scholte.py:203-210 and model_FANDS_PANDF_wavy_padmove.py:214-221):
derivatives of the grid coordinates with numpy.gradient(edge_order=2) in unit
computational spacing, then the inverse metric.  The reference writes the
vertical-column form J = h z_ze, xi_x = 1/h, xi_z = 0, ze_x = -z_xi/J,
ze_z = h/J; the general form below reduces to it for x = h xi.
"""
from __future__ import annotations

import numpy as np


def gradient2(f: np.ndarray, axis: int) -> np.ndarray:
    """numpy.gradient(f, axis=axis, edge_order=2) with unit spacing, written out:
    central differences inside, second-order one-sided at both ends (the
    evaluation order is numpy's)."""
    f = np.moveaxis(np.asarray(f, dtype=np.float64), axis, 0)
    out = np.empty_like(f)
    out[1:-1] = (f[2:] - f[:-2]) / 2.0
    out[0] = -1.5 * f[0] + 2.0 * f[1] + -0.5 * f[2]
    out[-1] = 0.5 * f[-3] + -2.0 * f[-2] + 1.5 * f[-1]
    return np.moveaxis(out, 0, axis)


def grid_metrics(x: np.ndarray, z: np.ndarray) -> dict:
    """J, xi_x, xi_z, ze_x, ze_z of the grid (x, z)[nz][nx] (float64)."""
    x_xi, z_xi = gradient2(x, 1), gradient2(z, 1)
    x_ze, z_ze = gradient2(x, 0), gradient2(z, 0)
    J = x_xi * z_ze - x_ze * z_xi
    return {"J": J, "xi_x": z_ze / J, "xi_z": (0.0 - x_ze) / J, "ze_x": (0.0 - z_xi) / J, "ze_z": x_xi / J}
