"""Body-fitted fluid-over-solid model generator (FP64) writing the solver's input layout.

Plays the role of the reference's ``tools/run_models/model_*.py`` scripts: it builds
a body-fitted curvilinear grid whose interface row follows the seafloor, its
metric terms, the water / solid media, the CFS-PML parameters and the source
positions, and writes ``<case>.json`` plus little-endian ``<f4`` row-major
``[nz][nx]`` binaries exactly as ``json_template.json`` / ``docs/io.md`` lay them
out (``tools/run_models/model_FANDS_PANDF_wavy.py:178-225``).

Grid conventions (``include/elements.h:77-87``, ``docs/io.md:53-66``):

* global rows ``0..nz-1``; rows 0-2 are frozen, row 3 is the pressure-release
  sea surface (z = 0), the interface row ``zeta0 = nz_fluid - 1`` is stored both
  as the last fluid row and as solid row 0, so ``nz = nz_fluid + nz_solid - 1``;
* ``x = h * ix`` (vertical grid columns, so ``xi_z = 0``) and
  ``z(ix, zeta) = h * (zeta - 3) + d(x) * w(zeta)`` where ``d`` is the seafloor
  undulation and ``w`` blends it in: 0 on the surface row, 1 on the interface
  row, 0 from the first bottom-PML row down (equal PML thickness everywhere,
  ``tools/run_models/model_ortho_classic_scholte.py:104-133``);
* metrics: ``metrics="analytic"`` (default) takes the closed-form derivatives of
  ``z``: ``J = h z_zeta``, ``xi_x = 1/h``, ``xi_z = 0``, ``ze_x = -z_xi / J``,
  ``ze_z = h / J`` (``model_FANDS_PANDF_wavy.py:130-144``); ``metrics="gpu"``
  differentiates the grid coordinates on the GPU in FP64 with the reference's
  ``numpy.gradient(edge_order=2)`` stencil (``model_ortho_classic_scholte.py:188-195``)
  through ``msfdm.grid_metrics``.

The solid may carry a seeded Gaussian random perturbation ``(1 + delta)`` applied to
vp, vs and rho with vs/vp fixed; ``delta`` comes from a counter-based splitmix64
generator with Box-Muller normals, smoothed by an FFT Gaussian filter and scaled
to the requested standard deviation (no ``numpy.random`` state involved).
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

SOLID, FANDS = 0, 1      # grid.calc_type
DIRI, CPML, PANDF = 0, 1, 2   # grid.boundary_type: Dirichlet, PML on 4 sides, PML sides/bottom + free top

_MASK64 = (1 << 64) - 1


def splitmix64(counter: np.ndarray, seed: int) -> np.ndarray:
    """Counter-based splitmix64: uint64 outputs for uint64 counters."""
    z = (counter.astype(np.uint64) + np.uint64((seed * 0x9E3779B97F4A7C15) & _MASK64)
         + np.uint64(0x9E3779B97F4A7C15))
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def normal_field(shape, seed: int) -> np.ndarray:
    """Standard normal field from splitmix64 uniforms and the Box-Muller transform."""
    n = int(np.prod(shape))
    c = np.arange(2 * n, dtype=np.uint64)
    u = (splitmix64(c, seed) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    u1 = np.maximum(u[:n], 1e-300)
    u2 = u[n:]
    return (np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)).reshape(shape)


def gaussian_random_field(shape, h: float, corr_len: float, std: float, seed: int) -> np.ndarray:
    """Zero-mean field with Gaussian covariance exp(-r^2 / corr_len^2), scaled to ``std``.

    White noise filtered by exp(-r^2 / (2 s^2)) with s = corr_len / 2 has exactly
    that covariance shape.  The filter is applied with zero padding via FFT.
    """
    if std == 0.0:
        return np.zeros(shape)
    s = 0.5 * corr_len / h
    pad = int(math.ceil(4 * s))
    nzp, nxp = shape[0] + 2 * pad, shape[1] + 2 * pad
    w = normal_field((nzp, nxp), seed)
    kz = np.fft.fftfreq(nzp)[:, None]
    kx = np.fft.rfftfreq(nxp)[None, :]
    filt = np.exp(-2.0 * (np.pi * s) ** 2 * (kz ** 2 + kx ** 2))
    g = np.fft.irfft2(np.fft.rfft2(w) * filt, s=(nzp, nxp))[pad:pad + shape[0], pad:pad + shape[1]]
    g -= g.mean()
    return g * (std / g.std())


def _blend(nz: int, zeta0: int, b_pml: float):
    """w(zeta) = u * v and dw/dzeta: smootherstep from the surface row 3 to the
    interface, cos^2 taper from the interface to the bottom PML."""
    zeta = np.arange(nz, dtype=np.float64)
    a = 3.0
    t = np.clip((zeta - a) / (zeta0 - a), 0.0, 1.0)
    u = t * t * t * (t * (6.0 * t - 15.0) + 10.0)
    du = np.where((zeta > a) & (zeta < zeta0), 30.0 * t * t * (1 - t) ** 2 / (zeta0 - a), 0.0)
    v = np.ones(nz)
    dv = np.zeros(nz)
    if b_pml > zeta0:
        seg = zeta > zeta0
        tt = np.clip((zeta[seg] - zeta0) / (b_pml - zeta0), 0.0, 1.0)
        v[seg] = np.cos(0.5 * np.pi * tt) ** 2
        dv[seg] = np.where(tt < 1.0, -0.5 * np.pi * np.sin(np.pi * tt) / (b_pml - zeta0), 0.0)
    return u * v, du * v + u * dv


@dataclass
class Model:
    """A solver input: scalars of the JSON plus the twelve float32 arrays."""
    nx: int
    nz_fluid: int
    nz_solid: int
    h: float
    fpeak: float
    dt: float
    nt: int
    snapshot: int
    npd: int
    thickness: float
    npow: int
    vp_max: float
    rc: float
    beta0: float
    src_x: list
    src_z: list
    arrays: dict = field(default_factory=dict)   # name -> float32 [rows][nx]
    calc_type: int = FANDS
    boundary_type: int = PANDF

    @property
    def nz(self) -> int:
        # src/elements.cu:87-99: the solid-only grid is the solid array
        if self.calc_type == SOLID:
            return self.nz_solid
        return self.nz_fluid + self.nz_solid - 1


def grid_coordinates(nx: int, nz: int, h: float, water_depth: float, *, npd: int, seafloor=()):
    """FP64 grid coordinates (x, z)[nz][nx] in metres of the body-fitted grid:
    x = h xi, z = h (zeta - 3) + d(x) w(zeta) (the grid of ``build_model``)."""
    zeta0 = int(round(water_depth / h)) + 3
    w, _ = _blend(nz, zeta0, float(nz - 4 - npd))
    x = h * np.arange(nx, dtype=np.float64)
    d = np.zeros(nx)
    for amp, lam in seafloor:
        d += amp * np.sin(2 * np.pi * x / lam)
    zeta = np.arange(nz, dtype=np.float64)
    Z = h * (zeta[:, None] - 3.0) + d[None, :] * w[:, None]
    X = np.broadcast_to(x[None, :], (nz, nx)).copy()
    return X, Z


def build_model(nx: int, nz: int, h: float, water_depth: float, *, npd: int, fpeak: float,
                dt: float, nt: int, src_x, src_z=None, src_height: float | None = None,
                seafloor=(), solid=(3000.0, 1700.0, 2200.0), fluid=(1500.0, 1000.0),
                random_std: float = 0.0, corr_len: float = 50.0, seed: int = 0,
                snapshot: int | None = None, ppw0: float = 8.0, rc: float = 1e-3,
                npow: int = 2, vp_max: float | None = None, metrics: str = "analytic",
                device: int = 0, boundary_type: int = PANDF) -> Model:
    """Build a fluid-over-solid model: pressure-release sea surface on row 3
    (``boundary_type`` PANDF) or a top PML in the water (CPML).

    ``seafloor`` is a list of ``(amplitude, wavelength)`` terms of the interface
    undulation ``sum A sin(2 pi x / lambda)`` around ``water_depth``.  Sources are
    given as grid columns ``src_x`` and either fluid-array rows ``src_z`` or a height
    ``src_height`` (m) above the interface row.
    """
    zeta0 = int(round(water_depth / h)) + 3
    nz_fluid = zeta0 + 1
    nz_solid = nz - nz_fluid + 1
    if nz_solid < npd + 8 or nz_fluid < 8:
        raise ValueError("grid too small for the water depth / PML width")
    if boundary_type not in (CPML, PANDF):
        raise ValueError(f"a fluid-solid model takes boundary_type 1 or 2, not {boundary_type}")
    if boundary_type == CPML and nz_fluid < npd + 6:
        raise ValueError("water too shallow for the top PML")
    b_pml = float(nz - 4 - npd)
    w, dw = _blend(nz, zeta0, b_pml)
    x = h * np.arange(nx, dtype=np.float64)
    d = np.zeros(nx)
    dd = np.zeros(nx)
    for amp, lam in seafloor:
        d += amp * np.sin(2 * np.pi * x / lam)
        dd += amp * 2 * np.pi * h / lam * np.cos(2 * np.pi * x / lam)
    dz_dxi = dd[None, :] * w[:, None]
    dz_dze = h + d[None, :] * dw[:, None]
    if np.any(dz_dze <= 0.2 * h):
        raise ValueError("seafloor undulation folds the grid")
    if metrics == "analytic":
        J = h * dz_dze
        xi_x = np.full((nz, nx), 1.0 / h)
        xi_z = np.zeros((nz, nx))
        ze_x = -dz_dxi / J
        ze_z = h / J
    elif metrics == "gpu":
        from ._lib import grid_metrics
        X, Z = grid_coordinates(nx, nz, h, water_depth, npd=npd, seafloor=seafloor)
        gm = grid_metrics(X, Z, device)
        J, xi_x, xi_z, ze_x, ze_z = (gm[k] for k in ("J", "xi_x", "xi_z", "ze_x", "ze_z"))
    else:
        raise ValueError(f"metrics must be 'analytic' or 'gpu', not {metrics!r}")

    vp_s, vs_s, rho_s = solid
    vp_f, rho_f = fluid
    delta = gaussian_random_field((nz_solid, nx), h, corr_len, random_std, seed)
    vp = vp_s * (1 + delta)
    vs = vs_s * (1 + delta)
    rho = rho_s * (1 + delta)
    mu = rho * vs ** 2
    lam_ = rho * (vp ** 2 - 2 * vs ** 2)
    arrays = {
        "C11": lam_ + 2 * mu, "C13": lam_, "C33": lam_ + 2 * mu, "C55": mu, "rho_s": rho,
        "K": np.full((nz_fluid, nx), rho_f * vp_f ** 2), "rho_f": np.full((nz_fluid, nx), rho_f),
        "J": J, "xi_x": xi_x, "xi_z": xi_z, "ze_x": ze_x, "ze_z": ze_z,
    }
    arrays = {k: np.ascontiguousarray(v, dtype="<f4") for k, v in arrays.items()}

    src_x = [int(s) for s in np.atleast_1d(src_x)]
    if src_z is None:
        hh = 10.0 if src_height is None else src_height
        src_z = [zeta0 - int(round(hh / h))] * len(src_x)
    src_z = [int(s) for s in np.atleast_1d(src_z)]
    c_min = min(vs_s, vp_f)
    beta0 = c_min / (0.5 * ppw0 * h * fpeak)      # docs/test.md eq. (41)
    return Model(nx=nx, nz_fluid=nz_fluid, nz_solid=nz_solid, h=h, fpeak=fpeak, dt=dt, nt=nt,
                 snapshot=(nt if snapshot is None else snapshot), npd=npd,
                 thickness=(npd - 1) * h, npow=npow,
                 vp_max=float(vp_max if vp_max is not None else max(vp_s, vp_f)), rc=rc,
                 beta0=round(beta0, 4), src_x=src_x, src_z=src_z, arrays=arrays,
                 boundary_type=boundary_type)


def build_solid_model(nx: int, nz: int, h: float, *, npd: int, fpeak: float, dt: float, nt: int,
                      src_x, src_z, topography=(), solid=(3200.0, 1732.0, 2000.0),
                      random_std: float = 0.0, corr_len: float = 50.0, seed: int = 0,
                      snapshot: int | None = None, ppw0: float = 8.0, rc: float = 1e-3,
                      npow: int = 2, vp_max: float | None = None, boundary_type: int = PANDF) -> Model:
    """Build a solid-only model (calc_type 0): by default with a traction-free
    top (PANDF); ``boundary_type`` CPML puts the PML on all four sides and DIRI
    clamps the grid edges (``model_SOLID_PANDF_ortho.py:29-33``).

    The grid follows ``tools/run_models/model_SOLID_PANDF_wavy_if.py:87-148``:
    ``x = h xi`` and ``z = h zeta + sum A sin(2 pi x / lambda) w(zeta)`` with the
    surface envelope ``w = cos^2(pi t / 2)``, ``t = (zeta - 3) / (b - 3)``, equal
    to 1 on the free-surface row 3 and 0 from the bottom-PML row ``b = nz - 4 - npd``
    down; ``topography`` lists the ``(A, lambda)`` terms.  Metrics are the closed
    form derivatives of that map.  Sources are solid-array ``(src_x, src_z)``.
    """
    if nz < npd + 12 or nx < 2 * npd + 12:
        raise ValueError("grid too small for the PML width")
    if boundary_type not in (DIRI, CPML, PANDF):
        raise ValueError(f"boundary_type must be 0, 1 or 2, not {boundary_type}")
    if topography and boundary_type != PANDF:
        raise ValueError("surface topography needs the free surface (boundary_type 2)")
    zeta = np.arange(nz, dtype=np.float64)
    b_pml = float(nz - 4 - npd)
    w = np.ones(nz)
    dw = np.zeros(nz)
    if topography:
        seg = (zeta >= 3.0) & (zeta <= b_pml)
        t = (zeta[seg] - 3.0) / (b_pml - 3.0)
        w[seg] = np.cos(0.5 * np.pi * t) ** 2
        dw[seg] = -0.5 * np.pi / (b_pml - 3.0) * np.sin(np.pi * t)
        w[zeta > b_pml] = 0.0
    x = h * np.arange(nx, dtype=np.float64)
    d = np.zeros(nx)
    dd = np.zeros(nx)
    for amp, lam in topography:
        d += amp * np.sin(2 * np.pi * x / lam)
        dd += amp * 2 * np.pi * h / lam * np.cos(2 * np.pi * x / lam)
    dz_dxi = dd[None, :] * w[:, None]
    dz_dze = h + d[None, :] * dw[:, None]
    if np.any(dz_dze <= 0.2 * h):
        raise ValueError("topography folds the grid")
    J = h * dz_dze
    vp_s, vs_s, rho_s = solid
    delta = gaussian_random_field((nz, nx), h, corr_len, random_std, seed)
    vp, vs, rho = vp_s * (1 + delta), vs_s * (1 + delta), rho_s * (1 + delta)
    mu = rho * vs ** 2
    lam_ = rho * (vp ** 2 - 2 * vs ** 2)
    arrays = {
        "C11": lam_ + 2 * mu, "C13": lam_, "C33": lam_ + 2 * mu, "C55": mu, "rho_s": rho,
        "K": np.zeros((0, nx)), "rho_f": np.zeros((0, nx)),
        "J": J, "xi_x": np.full((nz, nx), 1.0 / h), "xi_z": np.zeros((nz, nx)),
        "ze_x": -dz_dxi / J, "ze_z": h / J,
    }
    arrays = {k: np.ascontiguousarray(v, dtype="<f4") for k, v in arrays.items()}
    src_x = [int(s) for s in np.atleast_1d(src_x)]
    src_z = [int(s) for s in np.atleast_1d(src_z)]
    if len(src_z) != len(src_x):
        raise ValueError("src_x and src_z differ in length")
    beta0 = vs_s / (0.5 * ppw0 * h * fpeak)      # model_SOLID_PANDF_ortho.py:76
    return Model(nx=nx, nz_fluid=0, nz_solid=nz, h=h, fpeak=fpeak, dt=dt, nt=nt,
                 snapshot=(nt if snapshot is None else snapshot), npd=npd,
                 thickness=(npd - 1) * h, npow=npow,
                 vp_max=float(vp_max if vp_max is not None else vp_s), rc=rc,
                 beta0=round(beta0, 4), src_x=src_x, src_z=src_z, arrays=arrays,
                 calc_type=SOLID, boundary_type=boundary_type)


# Named workloads.  "flat401" and "seafloor2001" are the plain-solver workloads
# (401x201 flat, 2001x1001 body-fitted seafloor); the others are reduced cases
# used for fast parity checks.
def preset(name: str, **over) -> Model:
    if name == "flat401":
        kw = dict(nx=401, nz=201, h=5.0, water_depth=200.0, npd=20, fpeak=15.0, dt=5e-4, nt=2000,
                  src_x=[200], src_height=10.0, solid=(3000.0, 1700.0, 2200.0), vp_max=3000.0)
    elif name == "seafloor2001":
        kw = dict(nx=2001, nz=1001, h=2.5, water_depth=500.0, npd=30, fpeak=20.0, dt=2.5e-4,
                  nt=6000, src_x=[1000], src_height=10.0, seafloor=[(40.0, 800.0)],
                  solid=(3200.0, 1800.0, 2300.0), random_std=0.05, corr_len=50.0, seed=2001,
                  vp_max=3200.0)
    elif name == "cut201":
        kw = dict(nx=201, nz=101, h=5.0, water_depth=100.0, npd=20, fpeak=15.0, dt=5e-4, nt=600,
                  src_x=[100], src_height=10.0, solid=(3000.0, 1700.0, 2200.0), vp_max=3000.0)
    elif name == "tiny_curvi":
        kw = dict(nx=120, nz=90, h=5.0, water_depth=150.0, npd=10, fpeak=25.0, dt=4e-4, nt=300,
                  src_x=[60, 40], src_height=15.0, seafloor=[(12.0, 200.0), (4.0, 70.0)],
                  solid=(3000.0, 1700.0, 2200.0), random_std=0.05, corr_len=25.0, seed=7,
                  vp_max=3000.0)
    elif name == "cpml_fs_tiny":
        kw = dict(nx=120, nz=96, h=5.0, water_depth=150.0, npd=10, fpeak=25.0, dt=4e-4, nt=300,
                  src_x=[60, 40], src_height=15.0, seafloor=[(12.0, 200.0), (4.0, 70.0)],
                  solid=(3000.0, 1700.0, 2200.0), random_std=0.05, corr_len=25.0, seed=8,
                  vp_max=3000.0, boundary_type=CPML)
    elif name == "cpml_fs401":
        kw = dict(nx=401, nz=201, h=5.0, water_depth=200.0, npd=20, fpeak=15.0, dt=5e-4, nt=2000,
                  src_x=[200], src_height=10.0, solid=(3000.0, 1700.0, 2200.0), vp_max=3000.0,
                  boundary_type=CPML)
    elif name == "solid_flat301":
        return build_solid_model(**{**dict(nx=301, nz=161, h=5.0, npd=20, fpeak=15.0, dt=5e-4, nt=1500,
                                           src_x=[150], src_z=[6], solid=(3200.0, 1732.0, 2000.0)), **over})
    elif name == "solid_topo2001":
        return build_solid_model(**{**dict(nx=2001, nz=1001, h=2.5, npd=30, fpeak=20.0, dt=2.5e-4, nt=6000,
                                           src_x=[1000], src_z=[8], topography=[(20.0, 600.0)],
                                           solid=(3200.0, 1800.0, 2300.0), random_std=0.05,
                                           corr_len=50.0, seed=2002), **over})
    elif name == "solid_tiny":
        return build_solid_model(**{**dict(nx=110, nz=80, h=5.0, npd=10, fpeak=25.0, dt=4e-4, nt=300,
                                           src_x=[55, 40], src_z=[5, 20],
                                           topography=[(10.0, 200.0), (3.0, 70.0)],
                                           solid=(3000.0, 1700.0, 2200.0), random_std=0.05,
                                           corr_len=25.0, seed=11), **over})
    elif name == "solid_cpml_tiny":
        return build_solid_model(**{**dict(nx=96, nz=84, h=5.0, npd=10, fpeak=25.0, dt=4e-4, nt=300,
                                           src_x=[48, 30], src_z=[40, 20], boundary_type=CPML,
                                           solid=(3000.0, 1700.0, 2200.0), random_std=0.05,
                                           corr_len=25.0, seed=12), **over})
    elif name == "solid_diri_tiny":
        return build_solid_model(**{**dict(nx=90, nz=70, h=5.0, npd=10, fpeak=25.0, dt=4e-4, nt=300,
                                           src_x=[45], src_z=[30], boundary_type=DIRI,
                                           solid=(3000.0, 1700.0, 2200.0), random_std=0.05,
                                           corr_len=25.0, seed=13), **over})
    elif name == "solid_cpml501":
        # tools/run_models/model_SOLID_PANDF_ortho.py:35-55 (homogeneous, CPML, 2.5 m)
        return build_solid_model(**{**dict(nx=501, nz=501, h=2.5, npd=15, fpeak=30.0, dt=1e-4, nt=25000,
                                           snapshot=200, src_x=[250], src_z=[250], boundary_type=CPML,
                                           solid=(3200.0, 1732.0, 2000.0)), **over})
    elif name == "twoscale4001":
        kw = dict(nx=4001, nz=1501, h=2.5, water_depth=800.0, npd=30, fpeak=25.0, dt=2e-4,
                  nt=8000, src_x=[2000], src_height=10.0,
                  seafloor=[(60.0, 1500.0), (15.0, 230.0)], solid=(3200.0, 1800.0, 2300.0),
                  random_std=0.05, corr_len=50.0, seed=4001, vp_max=3200.0)
    else:
        raise KeyError(name)
    kw.update(over)
    return build_model(**kw)


_SOLID_FILES = [("C11", "C11"), ("C13", "C13"), ("C33", "C33"), ("C55", "C55"), ("rho", "rho_s")]
_FLUID_FILES = [("K", "K"), ("rho", "rho_f")]
_JACOB_FILES = ["J", "xi_x", "xi_z", "ze_x", "ze_z"]


def write_model(m: Model, model_dir: str, case: str, relative_to: str | None = None) -> str:
    """Write ``<model_dir>/{solid,fluid,jacob}/*.bin`` and ``<model_dir>/<case>.json``.

    Paths inside the JSON are relative to ``relative_to`` (the directory the solver
    runs in, like the reference's project root); absolute when None.
    """
    os.makedirs(model_dir, exist_ok=True)
    def path(*p):
        full = os.path.join(model_dir, *p)
        return os.path.relpath(full, relative_to) if relative_to else os.path.abspath(full)
    fluid = m.calc_type != SOLID   # a solid-only model has no grid.fluid block
    for sub in ("solid", "fluid", "jacob") if fluid else ("solid", "jacob"):
        os.makedirs(os.path.join(model_dir, sub), exist_ok=True)
    grid = {"calc_type": m.calc_type, "boundary_type": m.boundary_type,
            "src_pos": {"posx": list(m.src_x), "posz": list(m.src_z)},
            "solid": {"nx": m.nx, "nz": m.nz_solid},
            "jacob": {}}
    if fluid:
        grid["fluid"] = {"nx": m.nx, "nz": m.nz_fluid}
    for key, arr in _SOLID_FILES:
        m.arrays[arr].tofile(os.path.join(model_dir, "solid", key + ".bin"))
        grid["solid"][key] = path("solid", key + ".bin")
    for key, arr in _FLUID_FILES if fluid else ():
        m.arrays[arr].tofile(os.path.join(model_dir, "fluid", key + ".bin"))
        grid["fluid"][key] = path("fluid", key + ".bin")
    for key in _JACOB_FILES:
        m.arrays[key].tofile(os.path.join(model_dir, "jacob", key + ".bin"))
        grid["jacob"][key] = path("jacob", key + ".bin")
    doc = {"base": {"fpeak": m.fpeak, "dt": m.dt, "nt": m.nt, "snapshot": m.snapshot},
           "grid": grid,
           "PML": {"npd": m.npd, "thickness": m.thickness, "NPOW": m.npow, "Vp_max": m.vp_max,
                   "Rc": m.rc, "beta0": m.beta0}}
    jpath = os.path.join(model_dir, case + ".json")
    with open(jpath, "w") as f:
        json.dump(doc, f, indent=2)
        f.write("\n")
    write_pad(m, model_dir)
    return jpath


def write_pad(m: Model, model_dir: str) -> str:
    """Write ``<model_dir>/pad.txt``, the padding description the reference's
    model generators write beside the JSON and its comparison tools read
    (tools/run_models/model_FANDS_PANDF_double_cos_padmove.py:302-314;
    tools/compare_sem_scholte.py:59-64 reads pad_flat, tools/compare_double_cos.py:51
    reads shift): one "key value" line each.  The outer pad is the 3 frozen
    columns plus the npd-column PML on each side, so grid column c holds the
    physical x = (c - pad_flat) h; the fluid array starts with 3 frozen rows and
    the solid array ends with the bottom PML and 3 frozen rows."""
    pad = m.npd + 3
    fluid_full = m.nz_fluid if m.calc_type != SOLID else 0
    lines = [("nx_phys", m.nx - 2 * pad), ("shift", pad), ("pad_flat", pad), ("outer", pad), ("npd", m.npd),
             ("nz_fluid_phys", max(fluid_full - 3, 0)), ("nz_solid_phys", m.nz_solid - pad),
             ("nz_fluid_full", fluid_full), ("nz_solid_full", m.nz_solid), ("src_z_offset", 3),
             ("phys0", pad), ("phys1", m.nx - pad - 1)]
    path = os.path.join(model_dir, "pad.txt")
    with open(path, "w") as f:
        for k, v in lines:
            f.write(f"{k} {v}\n")
    return path


def read_model(json_path: str, cwd: str | None = None) -> Model:
    """Read a JSON + <f4 model (paths relative to ``cwd``, default the process CWD)."""
    with open(json_path) as f:
        doc = json.load(f)
    base = cwd or os.getcwd()
    def load(p, rows, nx):
        a = np.fromfile(os.path.join(base, p), dtype="<f4")
        if a.size != rows * nx:
            raise ValueError(f"{p}: expected {rows * nx} floats, found {a.size}")
        return a.reshape(rows, nx)
    g = doc["grid"]
    calc = int(g["calc_type"])
    nx, nzs = int(g["solid"]["nx"]), int(g["solid"]["nz"])
    nzf = int(g["fluid"]["nz"]) if calc != SOLID else 0
    nz = nzf + nzs - 1 if calc != SOLID else nzs
    arrays = {}
    for key, arr in _SOLID_FILES:
        arrays[arr] = load(g["solid"][key], nzs, nx)
    for key, arr in _FLUID_FILES:
        arrays[arr] = load(g["fluid"][key], nzf, nx) if calc != SOLID else np.zeros((0, nx), dtype="<f4")
    for key in _JACOB_FILES:
        arrays[key] = load(g["jacob"][key], nz, nx)
    b, p = doc["base"], doc["PML"]
    return Model(nx=nx, nz_fluid=nzf, nz_solid=nzs, h=float("nan"), fpeak=b["fpeak"], dt=b["dt"],
                 nt=int(b["nt"]), snapshot=int(b["snapshot"]), npd=int(p["npd"]),
                 thickness=p["thickness"], npow=int(p["NPOW"]), vp_max=p["Vp_max"], rc=p["Rc"],
                 beta0=p["beta0"], src_x=[int(v) for v in g["src_pos"]["posx"]],
                 src_z=[int(v) for v in g["src_pos"]["posz"]], arrays=arrays,
                 calc_type=int(g["calc_type"]), boundary_type=int(g["boundary_type"]))


def batch_shots(n: int = 64, seed: int = 64, nx_range=(800, 2000), nz_range=(400, 1000), nt: int = 4000,
                **over) -> list:
    """Shot windows of the flat 5 m water-over-solid medium (200 m of water,
    vp/vs/rho 3000/1700/2200, 15 Hz, dt 0.5 ms, 20-cell PML) with grid sizes drawn
    uniformly from nx_range x nz_range and the source column drawn uniformly over
    each window (10 m above the seafloor), all from one splitmix64 stream."""
    u = (splitmix64(np.arange(3 * n, dtype=np.uint64), seed) >> np.uint64(11)).astype(np.float64) / (1 << 53)
    shots = []
    npd = over.get("npd", 20)
    for i in range(n):
        nx = nx_range[0] + int(u[3 * i] * (nx_range[1] - nx_range[0] + 1))
        nz = nz_range[0] + int(u[3 * i + 1] * (nz_range[1] - nz_range[0] + 1))
        lo, hi = npd + 8, nx - npd - 9
        sx = lo + int(u[3 * i + 2] * (hi - lo + 1))
        kw = dict(nx=nx, nz=nz, h=5.0, water_depth=200.0, npd=npd, fpeak=15.0, dt=5e-4, nt=nt,
                  src_x=[sx], src_height=10.0, solid=(3000.0, 1700.0, 2200.0), vp_max=3000.0)
        kw.update(over)
        shots.append(build_model(**kw))
    return shots
