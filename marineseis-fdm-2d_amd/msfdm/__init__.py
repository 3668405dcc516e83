"""MI355X fluid-solid time stepper (MarineSeis-FDM-2D hot path) - Python side.

``msfdm.Solver`` drives libmsfdm.so (hand-written HIP kernels for gfx950) through
its C ABI; ``msfdm.modelgen`` builds and writes solver input models.
"""
from ._lib import (ARRAYS, BIN_PATH, FID, FIELDS, LIB_PATH, NFIELD, Batch, HostModel, MsfdmError,
                   Solver, grid_metrics, last_error, lib, run_cli, stream_probe)

__all__ = ["ARRAYS", "BIN_PATH", "FID", "FIELDS", "LIB_PATH", "NFIELD", "Batch", "HostModel", "MsfdmError",
           "Solver", "grid_metrics", "last_error", "lib", "run_cli", "stream_probe"]
