"""ctypes wrapper of the FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

# Field numbering (matches oracle.cpp and include/msfdm.h).
FIELDS = ["FVX", "FVZ", "FP", "FXVX", "FXVZ", "FXEX", "FXEZ",
          "SVX", "SVZ", "SSX", "SSZ", "SSXZ", "SXVX", "SXVZ", "SXEX", "SXEZ",
          "SZVX", "SZVZ", "SZEX", "SZEZ", "FZVX", "FZVZ", "FZEX", "FZEZ"]
FID = {n: i for i, n in enumerate(FIELDS)}
NFIELD = len(FIELDS)


def is_fluid(f: int) -> bool:
    return f <= FID["FXEZ"] or f >= FID["FZVX"]


class _Desc(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz_fluid", C.c_int), ("nz_solid", C.c_int), ("npd", C.c_int),
                ("npow", C.c_int), ("fpeak", C.c_float), ("dt", C.c_float),
                ("thickness", C.c_float), ("vp_max", C.c_float), ("rc", C.c_float),
                ("beta0", C.c_float), ("nsrc", C.c_int), ("src_x", C.POINTER(C.c_int)),
                ("src_z", C.POINTER(C.c_int)), ("calc_type", C.c_int), ("boundary_type", C.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        # MSFDM_ORACLE_LIB names another build of oracle.cpp (the mutation test
        # builds copies with planted mistakes and runs the pins against them).
        path = os.environ.get("MSFDM_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        fp = C.POINTER(C.c_float)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(_Desc)] + [fp] * 12
        for name, args, res in [
            ("orc_destroy", [vp], None), ("orc_set_threads", [C.c_int], None),
            ("orc_get_threads", [], C.c_int), ("orc_reset", [vp, C.c_int], None),
            ("orc_set_step", [vp, C.c_int], None), ("orc_get_step", [vp], C.c_int),
            ("orc_gap", [vp], C.c_int), ("orc_run", [vp, C.c_int], None),
            ("orc_num_records", [vp], C.c_int), ("orc_get_records", [vp, dp, dp, dp], None),
            ("orc_get_field", [vp, C.c_int, dp], None), ("orc_set_field", [vp, C.c_int, dp], None),
            ("orc_rhs", [vp, C.c_int, C.c_double], None), ("orc_get_k", [vp, C.c_int, dp], None),
            ("orc_get_ghosts", [vp, C.c_int, dp], None), ("orc_filter", [vp], None),
            ("orc_pml", [vp, C.c_int, dp, dp, dp], None), ("orc_normals", [vp, dp, dp], None),
            ("orc_stage_amp", [vp, C.c_int, C.c_int], C.c_double), ("orc_alm", [vp, dp], None),
            ("orc_op1d", [C.c_int, C.c_int, dp, C.c_int, C.c_int, C.c_int], C.c_double),
            ("orc_set_adj_in", [vp, C.c_int, dp], None), ("orc_get_adj_out", [vp, C.c_int, dp], None),
            ("orc_get_grad", [vp, C.c_int, dp], None), ("orc_zero_grad", [vp], None),
            ("orc_adjoint_op", [vp, C.c_int, C.c_int], None), ("orc_step_nosrc", [vp, C.c_int], None),
            ("orc_gradient", [vp, C.c_int, C.c_int, C.c_int, dp, dp, dp], None),
        ]:
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Oracle:
    """FP64 stepper over a ``msfdm.modelgen.Model``."""

    def __init__(self, model, threads: int | None = None):
        L = lib()
        if threads:
            L.orc_set_threads(int(threads))
        self.m = model
        sx = np.ascontiguousarray(model.src_x, dtype=np.int32)
        sz = np.ascontiguousarray(model.src_z, dtype=np.int32)
        self._keep = (sx, sz)
        d = _Desc(model.nx, model.nz_fluid, model.nz_solid, model.npd, model.npow, model.fpeak,
                  model.dt, model.thickness, model.vp_max, model.rc, model.beta0, len(sx),
                  sx.ctypes.data_as(C.POINTER(C.c_int)), sz.ctypes.data_as(C.POINTER(C.c_int)),
                  int(getattr(model, "calc_type", 1)), int(getattr(model, "boundary_type", 2)))
        a = model.arrays
        arrs = [np.ascontiguousarray(a[k], dtype=np.float32) for k in
                ("K", "rho_f", "C11", "C13", "C33", "C55", "rho_s", "J", "xi_x", "xi_z", "ze_x", "ze_z")]
        self.h = L.orc_create(C.byref(d), *[_fptr(x) for x in arrs])
        self.solid_only = int(getattr(model, "calc_type", 1)) == 0
        self.nx, self.nzs = model.nx, model.nz_solid
        self.nzf = 0 if self.solid_only else model.nz_fluid

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def shape(self, f: int):
        return (self.nzf if is_fluid(f) else self.nzs, self.nx)

    def reset(self, isrc: int = 0):
        lib().orc_reset(self.h, isrc)

    def run(self, nsteps: int):
        lib().orc_run(self.h, int(nsteps))

    @property
    def step(self) -> int:
        return lib().orc_get_step(self.h)

    @step.setter
    def step(self, it: int):
        lib().orc_set_step(self.h, int(it))

    @property
    def gap(self) -> int:
        return lib().orc_gap(self.h)

    def records(self):
        n = lib().orc_num_records(self.h)
        out = [np.zeros((n, self.nx)) for _ in range(3)]
        lib().orc_get_records(self.h, *[_dptr(o) for o in out])
        if self.solid_only:
            return [out[1]]   # vz of the free-surface row 3
        return out   # vx_solid, vz_solid, p_fluid

    def get(self, f) -> np.ndarray:
        f = FID[f] if isinstance(f, str) else f
        out = np.zeros(self.shape(f))
        lib().orc_get_field(self.h, f, _dptr(out))
        return out

    def set(self, f, arr):
        f = FID[f] if isinstance(f, str) else f
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape(f))
        lib().orc_set_field(self.h, f, _dptr(a))

    def rhs(self, direction: int, amp: float = 0.0):
        lib().orc_rhs(self.h, int(direction), float(amp))

    def get_k(self, f) -> np.ndarray:
        f = FID[f] if isinstance(f, str) else f
        out = np.zeros(self.shape(f))
        lib().orc_get_k(self.h, f, _dptr(out))
        return out

    def ghosts(self, which: int) -> np.ndarray:
        out = np.zeros((2, 3, self.nx))
        lib().orc_get_ghosts(self.h, which, _dptr(out))
        return out

    def filter(self):
        lib().orc_filter(self.h)

    def pml(self, axis: int):
        """axis 0 = xi, 1 = solid zeta, 2 = fluid zeta (top strip of the CPML grid)."""
        n = self.nx if axis == 0 else (self.nzs if axis == 1 else self.nzf)
        out = [np.zeros(n) for _ in range(3)]
        lib().orc_pml(self.h, axis, *[_dptr(o) for o in out])
        return out   # alpha, beta, d

    def normals(self):
        a, b = np.zeros(self.nx), np.zeros(self.nx)
        lib().orc_normals(self.h, _dptr(a), _dptr(b))
        return a, b

    def stage_amp(self, it: int, stage: int) -> float:
        return lib().orc_stage_amp(self.h, int(it), int(stage))

    def alm(self) -> np.ndarray:
        """Free-surface ALM coefficients [6][nx] of a solid-only grid."""
        out = np.zeros((6, self.nx))
        lib().orc_alm(self.h, _dptr(out))
        return out


GRADS = ["K", "rho_f", "C11", "C13", "C33", "C55", "rho_s"]


def _adj_methods():
    def set_adj_in(self, f, arr):
        f = FID[f] if isinstance(f, str) else f
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape(f))
        lib().orc_set_adj_in(self.h, f, _dptr(a))

    def get_adj_out(self, f):
        f = FID[f] if isinstance(f, str) else f
        out = np.zeros(self.shape(f))
        lib().orc_get_adj_out(self.h, f, _dptr(out))
        return out

    def adjoint_op(self, op, direction):
        lib().orc_adjoint_op(self.h, int(op), int(direction))

    def step_nosrc(self, it):
        lib().orc_step_nosrc(self.h, int(it))

    def grad(self, name):
        i = GRADS.index(name)
        rows = self.nzf if i < 2 else self.nzs
        out = np.zeros((rows, self.nx))
        lib().orc_get_grad(self.h, i, _dptr(out))
        return out

    def gradient(self, isrc, nsteps, res, every=16):
        """Gradient of sum_r <res_r, rec_r> w.r.t. the model arrays (dict)."""
        rs = [np.ascontiguousarray(r, dtype=np.float64) for r in res]
        lib().orc_gradient(self.h, int(isrc), int(nsteps), int(every), *[_dptr(r) for r in rs])
        return {n: self.grad(n) for n in GRADS}

    for fn in (set_adj_in, get_adj_out, adjoint_op, step_nosrc, grad, gradient):
        setattr(Oracle, fn.__name__, fn)


_adj_methods()


def op1d(kind: int, direction: int, f: np.ndarray, i: int, dist: int = 3, solid_side: bool = False) -> float:
    a = np.ascontiguousarray(f, dtype=np.float64)
    return lib().orc_op1d(kind, direction, _dptr(a), int(i), int(dist), int(solid_side))
