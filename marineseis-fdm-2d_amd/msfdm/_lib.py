"""ctypes binding of libmsfdm.so (the C ABI declared in include/msfdm.h).

The library is loaded from the package directory (built in-tree by
``make -C marineseis-fdm-2d_amd/csrc``).  There is no fallback: a missing or
unloadable library raises, so nothing here can silently run on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libmsfdm.so")
BIN_PATH = os.path.join(PKG_DIR, "bin", "main_debug")

OK, EINVAL, EIO, EPARSE, EDEVICE, ESTATE = range(6)

FIELDS = ["FVX", "FVZ", "FP", "FXVX", "FXVZ", "FXEX", "FXEZ",
          "SVX", "SVZ", "SSX", "SSZ", "SSXZ", "SXVX", "SXVZ", "SXEX", "SXEZ",
          "SZVX", "SZVZ", "SZEX", "SZEZ", "FZVX", "FZVZ", "FZEX", "FZEZ"]
FID = {n: i for i, n in enumerate(FIELDS)}
NFIELD = len(FIELDS)


def is_fluid(f: int) -> bool:
    """Field f lives on the fluid grid ([nz_fluid][nx])."""
    return f <= FID["FXEZ"] or f >= FID["FZVX"]


ARRAYS = ["K", "rho_f", "C11", "C13", "C33", "C55", "rho_s", "J", "xi_x", "xi_z", "ze_x", "ze_z"]

# Exported symbols (checked by the CPU tests against include/msfdm.h).
EXPORTS = [
    "msfdm_last_error", "msfdm_version", "msfdm_model_load_json", "msfdm_model_create",
    "msfdm_model_get_desc", "msfdm_model_get_array", "msfdm_model_destroy",
    "msfdm_solver_create", "msfdm_solver_destroy", "msfdm_reset_shot", "msfdm_run",
    "msfdm_synchronize", "msfdm_step_index", "msfdm_num_records", "msfdm_get_records",
    "msfdm_get_field", "msfdm_set_field", "msfdm_eval_rhs", "msfdm_get_rhs", "msfdm_set_graph",
    "msfdm_set_tile_rows", "msfdm_get_tile_rows", "msfdm_set_occupancy", "msfdm_get_occupancy",
    "msfdm_set_adjoint_launch", "msfdm_get_adjoint_launch",
    "msfdm_time_kernel", "msfdm_time_run", "msfdm_profile_run", "msfdm_bytes_per_step", "msfdm_kernel_bytes",
    "msfdm_run_cli", "msfdm_autotune", "msfdm_batch_create", "msfdm_batch_destroy",
    "msfdm_batch_reset", "msfdm_batch_run", "msfdm_batch_synchronize", "msfdm_batch_size",
    "msfdm_batch_shot", "msfdm_batch_set_tile_rows", "msfdm_batch_get_tile_rows",
    "msfdm_batch_set_occupancy", "msfdm_batch_get_occupancy",
    "msfdm_batch_set_graph", "msfdm_batch_time_run", "msfdm_batch_autotune",
    "msfdm_gradient", "msfdm_adjoint_forward", "msfdm_adjoint_backward",
    "msfdm_adjoint_set_input", "msfdm_adjoint_apply", "msfdm_adjoint_get_output",
    "msfdm_update_media", "msfdm_grid_metrics", "msfdm_gradient_shots", "msfdm_gradient_shots_plan",
    "msfdm_dispersion_image", "msfdm_dispersion_image_noalloc", "msfdm_stream_probe",
]
GRADS = ["K", "rho_f", "C11", "C13", "C33", "C55", "rho_s"]

# Size of the static device scratch of msfdm_dispersion_image_noalloc (MSFDM_DISP_SCRATCH_BYTES).
DISP_SCRATCH_BYTES = 32 << 20


class MsfdmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[msfdm error {code}] {msg}")
        self.code = code


class Desc(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz_fluid", C.c_int), ("nz_solid", C.c_int),
                ("calc_type", C.c_int), ("boundary_type", C.c_int),
                ("fpeak", C.c_float), ("dt", C.c_float), ("nt", C.c_int), ("snapshot", C.c_int),
                ("npd", C.c_int), ("thickness", C.c_float), ("npow", C.c_int),
                ("vp_max", C.c_float), ("rc", C.c_float), ("beta0", C.c_float),
                ("nsrc", C.c_int), ("src_x", C.POINTER(C.c_int)), ("src_z", C.POINTER(C.c_int))]


_LIB = None


def lib():
    """Load libmsfdm.so once and declare the prototypes."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is not built (run make -C marineseis-fdm-2d_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    vp, ip, fp, i = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_float), C.c_int
    protos = {
        "msfdm_last_error": ([], C.c_char_p),
        "msfdm_version": ([], C.c_char_p),
        "msfdm_model_load_json": ([C.c_char_p, C.POINTER(vp)], i),
        "msfdm_model_create": ([C.POINTER(Desc)] + [fp] * 12 + [C.POINTER(vp)], i),
        "msfdm_model_get_desc": ([vp, C.POINTER(Desc)], i),
        "msfdm_model_get_array": ([vp, i, fp], i),
        "msfdm_model_destroy": ([vp], None),
        "msfdm_solver_create": ([vp, i, C.POINTER(vp)], i),
        "msfdm_solver_destroy": ([vp], None),
        "msfdm_reset_shot": ([vp, i], i),
        "msfdm_run": ([vp, i], i),
        "msfdm_synchronize": ([vp], i),
        "msfdm_step_index": ([vp, ip], i),
        "msfdm_num_records": ([vp, ip], i),
        "msfdm_get_records": ([vp, i, i, fp, fp, fp], i),
        "msfdm_get_field": ([vp, i, fp], i),
        "msfdm_set_field": ([vp, i, fp], i),
        "msfdm_eval_rhs": ([vp, i, C.c_float], i),
        "msfdm_get_rhs": ([vp, i, fp], i),
        "msfdm_set_graph": ([vp, i], i),
        "msfdm_set_tile_rows": ([vp, i], i),
        "msfdm_get_tile_rows": ([vp, ip], i),
        "msfdm_set_occupancy": ([vp, i], i),
        "msfdm_get_occupancy": ([vp, ip], i),
        "msfdm_set_adjoint_launch": ([vp, i], i),
        "msfdm_get_adjoint_launch": ([vp, ip], i),
        "msfdm_time_kernel": ([vp, i, i, fp], i),
        "msfdm_time_run": ([vp, i, fp], i),
        "msfdm_profile_run": ([vp, i, i, fp, ip, fp], i),
        "msfdm_bytes_per_step": ([vp, C.POINTER(C.c_double)], i),
        "msfdm_kernel_bytes": ([vp, i, C.POINTER(C.c_double)], i),
        "msfdm_run_cli": ([C.c_char_p, C.c_char_p], i),
        "msfdm_autotune": ([vp, i, ip], i),
        "msfdm_batch_create": ([C.POINTER(vp), ip, i, i, C.POINTER(vp)], i),
        "msfdm_batch_destroy": ([vp], None),
        "msfdm_batch_reset": ([vp], i),
        "msfdm_batch_run": ([vp, i], i),
        "msfdm_batch_synchronize": ([vp], i),
        "msfdm_batch_size": ([vp, ip], i),
        "msfdm_batch_shot": ([vp, i, C.POINTER(vp)], i),
        "msfdm_batch_set_tile_rows": ([vp, i], i),
        "msfdm_batch_get_tile_rows": ([vp, ip], i),
        "msfdm_batch_set_occupancy": ([vp, i], i),
        "msfdm_batch_get_occupancy": ([vp, ip], i),
        "msfdm_batch_set_graph": ([vp, i], i),
        "msfdm_batch_time_run": ([vp, i, fp], i),
        "msfdm_batch_autotune": ([vp, i, ip], i),
        "msfdm_gradient": ([vp, i, i, i, fp, fp, fp] + [C.POINTER(C.c_double)] * 7, i),
        "msfdm_adjoint_forward": ([vp, i, i, i], i),
        "msfdm_adjoint_backward": ([vp, fp, fp, fp] + [C.POINTER(C.c_double)] * 7, i),
        "msfdm_adjoint_set_input": ([vp, i, fp], i),
        "msfdm_adjoint_apply": ([vp, i, i], i),
        "msfdm_adjoint_get_output": ([vp, i, fp], i),
        "msfdm_update_media": ([vp] + [fp] * 7, i),
        "msfdm_grid_metrics": ([C.POINTER(C.c_double)] * 2 + [i, i, i] + [C.POINTER(C.c_double)] * 5, i),
        "msfdm_gradient_shots": ([vp, i, ip, i, i, C.c_double, fp] + [C.POINTER(C.c_double)] * 7 + [ip], i),
        "msfdm_gradient_shots_plan": ([vp, i, i, C.POINTER(C.c_double), ip], i),
        "msfdm_dispersion_image": ([fp, i, i, C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_double), i, i,
                                    C.POINTER(C.c_double), C.POINTER(C.c_double)], i),
        "msfdm_dispersion_image_noalloc": ([vp, i, i, vp, C.c_double, vp, i, vp, vp], i),
        "msfdm_stream_probe": ([i, C.c_double, i, i, C.POINTER(C.c_double)], i),
    }
    for name, (args, res) in protos.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _LIB = L
    return L


def last_error() -> str:
    return lib().msfdm_last_error().decode()


def check(rc: int):
    if rc != OK:
        raise MsfdmError(rc, last_error())


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def record_gap(dt: float) -> int:
    """Record cadence in steps, (int)(1e-3 / dt + 0.5) with dt as float32 (src/main.cu:92)."""
    return max(1, int(1e-3 / float(np.float32(dt)) + 0.5))


def num_records(nsteps: int, dt: float) -> int:
    """Records taken by an nsteps-step run (steps 0, gap, 2 gap, ... < nsteps)."""
    return (int(nsteps) - 1) // record_gap(dt) + 1 if nsteps > 0 else 0


def check_residuals(res, nsteps: int, dt: float, nx: int):
    """The three record residuals (vx, vz, p) as float32 [nrec][nx] arrays, nrec
    the record count of an nsteps-step run; any other shape raises ValueError
    (the C side reads nrec * nx floats of each plane)."""
    want = (num_records(nsteps, dt), int(nx))
    if len(res) != 3:
        raise ValueError(f"three residual planes (vx, vz, p) expected, got {len(res)}")
    rs = [_f32(r) for r in res]
    for name, r in zip(("vx", "vz", "p"), rs):
        if r.shape != want:
            raise ValueError(f"residual {name}: shape {r.shape}, expected {want} "
                             f"(records of a {nsteps}-step run over every grid column)")
    return rs


def stream_probe(mbytes: float, reads: int, reps: int = 20, device: int = 0) -> float:
    """TB/s of a plain streaming kernel: `reads` input streams and one output
    stream of `mbytes` MB each (out = sum of the inputs), timed over `reps`
    launches on the GPU."""
    v = C.c_double()
    check(lib().msfdm_stream_probe(int(device), float(mbytes), int(reads), int(reps), C.byref(v)))
    return v.value


def grid_metrics(x, z, device: int = 0) -> dict:
    """FP64 metric terms of the grid (x, z) ([nz][nx] metres) computed on the GPU:
    numpy.gradient(edge_order=2) derivatives and the inverse metric.  Returns
    float64 arrays J, xi_x, xi_z, ze_x, ze_z."""
    xd = np.ascontiguousarray(x, dtype=np.float64)
    zd = np.ascontiguousarray(z, dtype=np.float64)
    if xd.ndim != 2 or xd.shape != zd.shape:
        raise ValueError("x and z must be 2-D arrays of one shape")
    nz, nx = xd.shape
    out = {k: np.empty((nz, nx)) for k in ("J", "xi_x", "xi_z", "ze_x", "ze_z")}
    dp = C.POINTER(C.c_double)
    check(lib().msfdm_grid_metrics(xd.ctypes.data_as(dp), zd.ctypes.data_as(dp), nz, nx, int(device),
                                   *[out[k].ctypes.data_as(dp) for k in ("J", "xi_x", "xi_z", "ze_x", "ze_z")]))
    return out


class HostModel:
    """A msfdm_model handle (host memory only; no GPU needed)."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def from_json(cls, path: str) -> "HostModel":
        h = C.c_void_p()
        check(lib().msfdm_model_load_json(path.encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def from_model(cls, m) -> "HostModel":
        """From a ``msfdm.modelgen.Model``."""
        sx = np.ascontiguousarray(m.src_x, dtype=np.int32)
        sz = np.ascontiguousarray(m.src_z, dtype=np.int32)
        d = Desc(m.nx, m.nz_fluid, m.nz_solid, m.calc_type, m.boundary_type, m.fpeak, m.dt, m.nt,
                 m.snapshot, m.npd, m.thickness, m.npow, m.vp_max, m.rc, m.beta0, len(sx),
                 sx.ctypes.data_as(C.POINTER(C.c_int)), sz.ctypes.data_as(C.POINTER(C.c_int)))
        arrs = [_f32(m.arrays[k]) for k in ARRAYS]
        h = C.c_void_p()
        check(lib().msfdm_model_create(C.byref(d), *[_fptr(a) for a in arrs], C.byref(h)))
        return cls(h)

    def desc(self) -> Desc:
        d = Desc()
        check(lib().msfdm_model_get_desc(self.h, C.byref(d)))
        return d

    def array(self, name: str) -> np.ndarray:
        d = self.desc()
        which = ARRAYS.index(name)
        solid_only = d.calc_type == 0      # no fluid arrays; the grid is the solid array
        nz = d.nz_solid if solid_only else d.nz_fluid + d.nz_solid - 1
        rows = (0 if solid_only else d.nz_fluid) if which < 2 else (d.nz_solid if which < 7 else nz)
        out = np.zeros((rows, d.nx), dtype=np.float32)
        check(lib().msfdm_model_get_array(self.h, which, _fptr(out)))
        return out

    def close(self):
        if self.h:
            lib().msfdm_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    """Device solver over a HostModel (or a modelgen.Model)."""

    def __init__(self, model, device: int = 0, _handle=None, _owner=None):
        self._hm = model if isinstance(model, HostModel) else HostModel.from_model(model)
        d = self._hm.desc()
        self.solid_only = d.calc_type == 0   # solid grid with a traction-free top
        self.nx, self.nzf, self.nzs = d.nx, (0 if self.solid_only else d.nz_fluid), d.nz_solid
        self.nt, self.dt = d.nt, d.dt
        self._owner = _owner      # a Batch owns the handles of its shots
        # adjoint state: the checkpoints of the last adjoint_forward (token bumped
        # by every call that replaces them or the medium they were taken in)
        self._adj_token = 0
        self._adj_nsteps = 0
        if _handle is not None:
            self.h = _handle
            return
        h = C.c_void_p()
        check(lib().msfdm_solver_create(self._hm.h, int(device), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self._owner is None:
            lib().msfdm_solver_destroy(self.h)
        self.h = None

    def autotune(self, nsteps: int = 64) -> int:
        v = C.c_int()
        check(lib().msfdm_autotune(self.h, int(nsteps), C.byref(v)))
        return v.value

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, f: int):
        return (self.nzf if is_fluid(f) else self.nzs, self.nx)

    def reset(self, isrc: int = 0):
        check(lib().msfdm_reset_shot(self.h, int(isrc)))

    def run(self, nsteps: int):
        check(lib().msfdm_run(self.h, int(nsteps)))

    def sync(self):
        check(lib().msfdm_synchronize(self.h))

    @property
    def step(self) -> int:
        v = C.c_int()
        check(lib().msfdm_step_index(self.h, C.byref(v)))
        return v.value

    def num_records(self) -> int:
        v = C.c_int()
        check(lib().msfdm_num_records(self.h, C.byref(v)))
        return v.value

    def records(self, first: int = 0, count: int | None = None):
        n = self.num_records()
        count = n - first if count is None else count
        out = [np.zeros((count, self.nx), dtype=np.float32) for _ in range(3)]
        check(lib().msfdm_get_records(self.h, int(first), int(count), *[_fptr(o) for o in out]))
        if self.solid_only:
            return [out[1]]   # vz of the free-surface row 3
        return out   # solid vx, solid vz, fluid p

    def get(self, f) -> np.ndarray:
        f = FID[f] if isinstance(f, str) else f
        out = np.zeros(self.shape(f), dtype=np.float32)
        check(lib().msfdm_get_field(self.h, f, _fptr(out)))
        return out

    def set(self, f, arr):
        f = FID[f] if isinstance(f, str) else f
        a = _f32(arr).reshape(self.shape(f))
        check(lib().msfdm_set_field(self.h, f, _fptr(a)))

    def eval_rhs(self, direction: int, amp: float = 0.0):
        check(lib().msfdm_eval_rhs(self.h, int(direction), float(amp)))

    def get_rhs(self, f) -> np.ndarray:
        f = FID[f] if isinstance(f, str) else f
        out = np.zeros(self.shape(f), dtype=np.float32)
        check(lib().msfdm_get_rhs(self.h, f, _fptr(out)))
        return out

    @property
    def tile_rows(self) -> int:
        v = C.c_int()
        check(lib().msfdm_get_tile_rows(self.h, C.byref(v)))
        return v.value

    @tile_rows.setter
    def tile_rows(self, rows: int):
        check(lib().msfdm_set_tile_rows(self.h, int(rows)))

    @property
    def occupancy(self) -> int:
        """Occupancy policy of the stage kernels (0 = tuned per stage, 1 = at
        least 5 waves per SIMD in every stage); msfdm_autotune picks one."""
        v = C.c_int()
        check(lib().msfdm_get_occupancy(self.h, C.byref(v)))
        return v.value

    @occupancy.setter
    def occupancy(self, policy: int):
        check(lib().msfdm_set_occupancy(self.h, int(policy)))

    @property
    def adjoint_launch(self) -> int:
        """Adjoint products per stage: 1 = one launch, 0 = three, -1 = by grid size."""
        v = C.c_int()
        check(lib().msfdm_get_adjoint_launch(self.h, C.byref(v)))
        return v.value

    @adjoint_launch.setter
    def adjoint_launch(self, mode: int):
        check(lib().msfdm_set_adjoint_launch(self.h, int(mode)))

    def set_graph(self, enable: bool):
        check(lib().msfdm_set_graph(self.h, int(bool(enable))))

    def update_media(self, **arrays):
        """Replace medium arrays (K, rho_f, C11, C13, C33, C55, rho_s) in place."""
        keep = {k: _f32(v) for k, v in arrays.items()}
        unknown = set(keep) - set(GRADS)
        if unknown:
            raise KeyError(f"unknown medium arrays {sorted(unknown)}")
        for n, v in keep.items():
            want = ((self.nzf if GRADS.index(n) < 2 else self.nzs), self.nx)
            if v.shape != want:
                raise ValueError(f"{n}: shape {v.shape}, expected {want}")
        ptrs = [(_fptr(keep[n]) if n in keep else None) for n in GRADS]
        self._adj_token += 1
        check(lib().msfdm_update_media(self.h, *ptrs))

    # --- adjoint ------------------------------------------------------------
    def _grad_arrays(self):
        return {n: np.zeros(((self.nzf if k < 2 else self.nzs), self.nx)) for k, n in enumerate(GRADS)}

    def _residuals(self, res, nsteps: int):
        return check_residuals(res, nsteps, self.dt, self.nx)

    @property
    def adjoint_token(self) -> int:
        """Changes whenever the adjoint checkpoints or the medium are replaced."""
        return self._adj_token

    def adjoint_forward(self, isrc: int, nsteps: int, every: int = 0):
        self._adj_token += 1
        check(lib().msfdm_adjoint_forward(self.h, int(isrc), int(nsteps), int(every)))
        self._adj_nsteps = int(nsteps)

    def adjoint_backward(self, res):
        if self._adj_nsteps <= 0:
            raise RuntimeError("adjoint_backward needs a preceding adjoint_forward")
        rs = self._residuals(res, self._adj_nsteps)
        g = self._grad_arrays()
        dp = C.POINTER(C.c_double)
        check(lib().msfdm_adjoint_backward(self.h, *[_fptr(r) for r in rs],
                                           *[g[n].ctypes.data_as(dp) for n in GRADS]))
        return g

    def gradient(self, isrc: int, nsteps: int, res, every: int = 0):
        """Gradient of sum_r <res_r, rec_r> w.r.t. the model arrays (dict of float64)."""
        rs = self._residuals(res, nsteps)
        g = self._grad_arrays()
        self._adj_token += 1
        self._adj_nsteps = int(nsteps)
        dp = C.POINTER(C.c_double)
        check(lib().msfdm_gradient(self.h, int(isrc), int(nsteps), int(every), *[_fptr(r) for r in rs],
                                   *[g[n].ctypes.data_as(dp) for n in GRADS]))
        return g

    def gradient_shots(self, shots, nsteps: int, res, every: int = 0, mem_bytes: float = 0.0):
        """Sum over the shot indices `shots` of the single-shot gradients, all shots
        of a wave stepped together.  res[i] = (res_vx, res_vz, res_p) of shot i, each
        [nrec][nx].  Returns (dict of float64 gradients, shots per wave)."""
        isrc = np.ascontiguousarray(shots, dtype=np.int32)
        if len(res) != isrc.size:
            raise ValueError("one residual triple per shot")
        r = np.ascontiguousarray(np.stack([np.stack(self._residuals(rr, nsteps)) for rr in res]), dtype=np.float32)
        self._adj_token += 1
        g = self._grad_arrays()
        dp = C.POINTER(C.c_double)
        wave = C.c_int()
        check(lib().msfdm_gradient_shots(self.h, int(isrc.size), isrc.ctypes.data_as(C.POINTER(C.c_int)), int(nsteps),
                                         int(every), float(mem_bytes), _fptr(r),
                                         *[g[n].ctypes.data_as(dp) for n in GRADS], C.byref(wave)))
        return g, wave.value

    def gradient_shots_plan(self, nsteps: int, every: int = 0):
        """(device bytes per shot of a multi-shot gradient wave, checkpoint interval)."""
        b, e = C.c_double(), C.c_int()
        check(lib().msfdm_gradient_shots_plan(self.h, int(nsteps), int(every), C.byref(b), C.byref(e)))
        return b.value, e.value

    def adjoint_set_input(self, f, arr):
        f = FID[f] if isinstance(f, str) else f
        a = _f32(arr).reshape(self.shape(f))
        check(lib().msfdm_adjoint_set_input(self.h, f, _fptr(a)))

    def adjoint_apply(self, op: int, direction: int):
        check(lib().msfdm_adjoint_apply(self.h, int(op), int(direction)))

    def adjoint_get_output(self, f) -> np.ndarray:
        f = FID[f] if isinstance(f, str) else f
        out = np.zeros(self.shape(f), dtype=np.float32)
        check(lib().msfdm_adjoint_get_output(self.h, f, _fptr(out)))
        return out

    def time_kernel(self, which: int, reps: int) -> float:
        v = C.c_float()
        check(lib().msfdm_time_kernel(self.h, int(which), int(reps), C.byref(v)))
        return v.value

    def time_run(self, nsteps: int) -> float:
        v = C.c_float()
        check(lib().msfdm_time_run(self.h, int(nsteps), C.byref(v)))
        return v.value

    def profile_run(self, nsteps: int, classes=range(6)) -> dict:
        """Step nsteps as run() does, the launches of the kernel classes in
        `classes` timed by their own dispatch: {"kernel_ms": [6] totals (RK
        stages 1-4, interface finish, other), "launches": [6] counts of timed
        launches, "wall_ms": the whole run}."""
        ms = (C.c_float * 6)()
        n = (C.c_int * 6)()
        wall = C.c_float()
        mask = sum(1 << int(c) for c in classes)
        check(lib().msfdm_profile_run(self.h, int(nsteps), mask, ms, n, C.byref(wall)))
        return {"kernel_ms": list(ms), "launches": list(n), "wall_ms": wall.value}

    def kernel_bytes(self, which: int) -> float:
        v = C.c_double()
        check(lib().msfdm_kernel_bytes(self.h, int(which), C.byref(v)))
        return v.value

    def bytes_per_step(self) -> float:
        v = C.c_double()
        check(lib().msfdm_bytes_per_step(self.h, C.byref(v)))
        return v.value


class Batch:
    """Shots of (possibly) different models stepped together on one device."""

    def __init__(self, models, shot_index=None, device: int = 0):
        self._hms = [m if isinstance(m, HostModel) else HostModel.from_model(m) for m in models]
        n = len(self._hms)
        arr = (C.c_void_p * n)(*[hm.h.value if isinstance(hm.h, C.c_void_p) else hm.h for hm in self._hms])
        idx = (C.c_int * n)(*(shot_index if shot_index is not None else [0] * n))
        h = C.c_void_p()
        check(lib().msfdm_batch_create(arr, idx, n, int(device), C.byref(h)))
        self.h = h
        self.shots = []
        for i in range(n):
            sh = C.c_void_p()
            check(lib().msfdm_batch_shot(self.h, i, C.byref(sh)))
            self.shots.append(Solver(self._hms[i], _handle=sh, _owner=self))

    def close(self):
        if getattr(self, "h", None):
            for s in self.shots:
                s.h = None
            lib().msfdm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        check(lib().msfdm_batch_reset(self.h))

    def run(self, nsteps: int):
        check(lib().msfdm_batch_run(self.h, int(nsteps)))

    def sync(self):
        check(lib().msfdm_batch_synchronize(self.h))

    def shot(self, i: int) -> Solver:
        self.sync()
        return self.shots[i]

    @property
    def tile_rows(self) -> int:
        v = C.c_int()
        check(lib().msfdm_batch_get_tile_rows(self.h, C.byref(v)))
        return v.value

    @tile_rows.setter
    def tile_rows(self, rows: int):
        check(lib().msfdm_batch_set_tile_rows(self.h, int(rows)))

    @property
    def occupancy(self) -> int:
        """Occupancy policy of the batched stage kernels (as Solver.occupancy)."""
        v = C.c_int()
        check(lib().msfdm_batch_get_occupancy(self.h, C.byref(v)))
        return v.value

    @occupancy.setter
    def occupancy(self, policy: int):
        check(lib().msfdm_batch_set_occupancy(self.h, int(policy)))

    def set_graph(self, enable: bool):
        check(lib().msfdm_batch_set_graph(self.h, int(bool(enable))))

    def time_run(self, nsteps: int) -> float:
        v = C.c_float()
        check(lib().msfdm_batch_time_run(self.h, int(nsteps), C.byref(v)))
        return v.value

    def autotune(self, nsteps: int = 64) -> int:
        v = C.c_int()
        check(lib().msfdm_batch_autotune(self.h, int(nsteps), C.byref(v)))
        return v.value


def run_cli(json_path: str, out_root: str) -> int:
    """Run the main_debug flow in-process; returns its exit code."""
    return lib().msfdm_run_cli(json_path.encode(), out_root.encode())
