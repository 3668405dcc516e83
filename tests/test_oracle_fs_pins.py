"""Pins of the fluid-solid FP64 oracle (PmlFs::FluidSolid: include/pml_fs_kernels.cuh:333-827,
src/pml_fs_kernels.cu:385-525) against mathematics that the oracle does not restate.

* Rotation covariance: on a grid turned by 90 degrees (x = -h zeta, z = h xi) the
  whole run (RHS branches, ghost fluxes, closures, PML projections, interface
  rotation, filter, records) must be the flat run with every vector and tensor
  rotated.  A swapped metric, a sign or a mixed-up component fails it.
* Plane waves on a sheared affine grid and on a curvilinear grid (xi_z, ze_x both
  nonzero): the non-conservative interior operator of all four RHS kernels
  reproduces rho dv/dt = div sigma, dsigma/dt = C : grad v (fluid: -grad p,
  -K div v); in the PML strips with zero memory it reproduces the stretched
  gradient grad + (1/beta - 1) n (n . grad), n the unit normal of the strip,
  and the memory RHS (d/beta) times that projection.
* Bilinear fields on the sheared grid: the conservative rows with ghost fluxes
  (interface and pressure-release surface), the 4-2 / 2-2 closures and the
  interface row (normal components from the side the direction names, tangential
  velocity kept, p = -sigma_nn, sigma_nt = 0) are exact up to the DRP weights'
  zeroth moment.
* The PML memory ODE realises the CFS stretch 1/s(w), s = beta + d/(alpha + i w),
  for every memory field; the side strips absorb what a wide grid would let pass.
* The Scholte wave along a flat seafloor travels at the root of the fluid/solid
  half-space secular equation: the deep-water limit of the layer relation of
  tools/test.py:101-140 (omega_from_c), continued below the water speed.
"""
import dataclasses
import math

import numpy as np
import pytest

import pyoracle
from msfdm import modelgen

F, B = 0, 1
FLUID = ("FVX", "FVZ", "FP")
SOLID = ("SVX", "SVZ", "SSX", "SSZ", "SSXZ")


def _fs(nx=110, nz=84, h=5.0, water=150.0, npd=10, boundary_type=2, random_std=0.05, **kw):
    base = dict(nx=nx, nz=nz, h=h, water_depth=water, npd=npd, fpeak=20.0, dt=4e-4, nt=10,
                src_x=[nx // 2], src_height=10.0, random_std=random_std, corr_len=30.0, seed=5,
                boundary_type=boundary_type)
    base.update(kw)
    return modelgen.build_model(**base)


def _mapped(m, x_xi, x_ze, z_xi, z_ze):
    """The model on the map with derivatives x_xi .. z_ze ([nz][nx], unit
    computational spacing): J = x_xi z_ze - x_ze z_xi and the inverse metrics."""
    J = x_xi * z_ze - x_ze * z_xi
    a = dict(m.arrays)
    for k, v in dict(J=J, xi_x=z_ze / J, xi_z=-x_ze / J, ze_x=-z_xi / J, ze_z=x_xi / J).items():
        a[k] = np.ascontiguousarray(v, dtype=np.float32)
    return dataclasses.replace(m, arrays=a)


def _grid(m, kind, h=5.0):
    """(model, X, Z) on an affine sheared map or a curvilinear one, both with
    xi_z != 0 and ze_x != 0."""
    nz = m.nz_fluid + m.nz_solid - 1
    XI, ZE = np.meshgrid(np.arange(m.nx, dtype=np.float64), np.arange(nz, dtype=np.float64))
    if kind == "affine":
        a, b = 0.3, 0.2
        X, Z = h * (XI + a * ZE), h * (b * XI + ZE)
        d = [np.full_like(XI, v) for v in (h, a * h, b * h, h)]
    else:
        A, Pz, Bm, Px = 3.0, 60.0, 2.5, 80.0
        X = h * (XI + A * np.sin(2 * np.pi * ZE / Pz))
        Z = h * (ZE + Bm * np.sin(2 * np.pi * XI / Px))
        d = [np.full_like(XI, h), h * A * 2 * np.pi / Pz * np.cos(2 * np.pi * ZE / Pz),
             h * Bm * 2 * np.pi / Px * np.cos(2 * np.pi * XI / Px), np.full_like(XI, h)]
    return _mapped(m, *d), X, Z


class _Metrics:
    """Float metric arrays (as the oracle reads them) of the fluid and solid rows."""

    def __init__(self, m):
        nzf = m.nz_fluid
        g = {k: m.arrays[k].astype(np.float64) for k in ("xi_x", "xi_z", "ze_x", "ze_z")}
        self.f = {k: v[:nzf] for k, v in g.items()}
        self.s = {k: v[nzf - 1:] for k, v in g.items()}


def _proj(nx_, nz_, gx, gz):
    """(n . g) n for the unnormalised normal (nx_, nz_)."""
    t = (nx_ * gx + nz_ * gz) / (nx_ * nx_ + nz_ * nz_)
    return t * nx_, t * nz_


class _Stretch:
    """Per-cell stretched gradient grad + sum_strips (1/beta - 1) n (n . grad) of
    the fluid or solid rows, with the oracle's PML profiles (pinned separately by
    the design reflection coefficient)."""

    def __init__(self, o, m, met, solid):
        al, be, dp = o.pml(0)
        rows = m.nz_solid if solid else m.nz_fluid
        cols = np.arange(m.nx)
        side = np.minimum(cols - 3, m.nx - 4 - cols)
        side = (side >= 0) & (side < m.npd)
        self.met = met
        self.sxi = np.broadcast_to(np.where(side, 1.0 / be - 1.0, 0.0)[None, :], (rows, m.nx))
        self.bxi = np.broadcast_to(np.where(side, dp / be, 0.0)[None, :], (rows, m.nx))
        if solid:
            alz, bez, dpz = o.pml(1)
            bot = np.arange(rows) >= rows - 3 - m.npd
            self.sze = np.broadcast_to(np.where(bot, 1.0 / bez - 1.0, 0.0)[:, None], (rows, m.nx))
            self.bze = np.broadcast_to(np.where(bot, dpz / bez, 0.0)[:, None], (rows, m.nx))
        else:
            self.sze = np.zeros((rows, m.nx))
            self.bze = np.zeros((rows, m.nx))

    def grad(self, g):
        gx, gz = g
        M = self.met
        px, pz = _proj(M["xi_x"], M["xi_z"], gx, gz)
        qx, qz = _proj(M["ze_x"], M["ze_z"], gx, gz)
        return gx + self.sxi * px + self.sze * qx, gz + self.sxi * pz + self.sze * qz


def _physics(m, sf, ss, gf, gs):
    """RHS of the first-order system from gradients gf / gs (dict field -> (d/dx,
    d/dz)), each passed through the stretch sf / ss."""
    a = {k: m.arrays[k].astype(np.float64) for k in ("K", "rho_f", "C11", "C13", "C33", "C55", "rho_s")}
    p = sf.grad(gf["FP"]); vxf = sf.grad(gf["FVX"]); vzf = sf.grad(gf["FVZ"])
    vx = ss.grad(gs["SVX"]); vz = ss.grad(gs["SVZ"])
    sx = ss.grad(gs["SSX"]); sz = ss.grad(gs["SSZ"]); sxz = ss.grad(gs["SSXZ"])
    return {
        "FVX": -p[0] / a["rho_f"], "FVZ": -p[1] / a["rho_f"], "FP": -a["K"] * (vxf[0] + vzf[1]),
        "SVX": (sx[0] + sxz[1]) / a["rho_s"], "SVZ": (sxz[0] + sz[1]) / a["rho_s"],
        "SSX": a["C11"] * vx[0] + a["C13"] * vz[1], "SSZ": a["C13"] * vx[0] + a["C33"] * vz[1],
        "SSXZ": a["C55"] * (vx[1] + vz[0]),
    }


def _memory_rhs(m, sf, ss, gf, gs):
    """Memory RHS at zero memory: (d/beta) times the strip projections of the
    gradients (velocity memory: of -grad p / div sigma; stress memory: of grad v
    as (n . grad v_c) / |n|^2)."""
    out = {}
    for st, pre, bxi, bze, g in ((sf, "F", sf.bxi, None, gf), (ss, "S", ss.bxi, ss.bze, gs)):
        M = st.met
        for strip, b, (nx_, nz_) in (("X", bxi, (M["xi_x"], M["xi_z"])), ("Z", bze, (M["ze_x"], M["ze_z"]))):
            if b is None:
                continue
            P = lambda gr: _proj(nx_, nz_, *gr)
            nn = nx_ * nx_ + nz_ * nz_
            dv = lambda gr: (nx_ * gr[0] + nz_ * gr[1]) / nn
            if pre == "F":
                Dx, Dz = P(g["FP"])
                out[f"F{strip}VX"], out[f"F{strip}VZ"] = -b * Dx, -b * Dz
                out[f"F{strip}EX"], out[f"F{strip}EZ"] = b * dv(g["FVX"]), b * dv(g["FVZ"])
            else:
                Pxx, Pxz, Pzz = P(g["SSX"]), P(g["SSXZ"]), P(g["SSZ"])
                out[f"S{strip}VX"] = b * (Pxx[0] + Pxz[1])
                out[f"S{strip}VZ"] = b * (Pxz[0] + Pzz[1])
                out[f"S{strip}EX"], out[f"S{strip}EZ"] = b * dv(g["SVX"]), b * dv(g["SVZ"])
    return out


def _set_state(o, m, vals):
    for k, v in vals.items():
        o.set(k, v)


# ---------------------------------------------------------------------------
# Rotation covariance
# ---------------------------------------------------------------------------

def _rot90(m, h):
    """x = -h zeta, z = h xi: xi_x = 0, xi_z = 1/h, ze_x = -1/h, ze_z = 0."""
    nz = m.nz_fluid + m.nz_solid - 1
    z0 = np.zeros((nz, m.nx))
    return _mapped(m, z0, z0 - h, z0 + h, z0)


# v' = R v, sigma' = R sigma R^T with R = [[0, -1], [1, 0]]; memory fields are
# vectors (velocity memory: projected -grad p / div sigma; stress memory: indexed
# by the velocity component).
_ROT = {"FVX": ("FVZ", -1), "FVZ": ("FVX", 1), "FP": ("FP", 1),
        "SVX": ("SVZ", -1), "SVZ": ("SVX", 1), "SSX": ("SSZ", 1), "SSZ": ("SSX", 1), "SSXZ": ("SSXZ", -1)}
for _p in ("FX", "SX", "SZ", "FZ"):
    _ROT[_p + "VX"] = (_p + "VZ", -1); _ROT[_p + "VZ"] = (_p + "VX", 1)
    _ROT[_p + "EX"] = (_p + "EZ", -1); _ROT[_p + "EZ"] = (_p + "EX", 1)


@pytest.mark.parametrize("boundary_type", [2, 1])
def test_rotated_grid_run_is_the_rotated_flat_run(boundary_type):
    h = 5.0
    flat = _fs(nx=90, nz=80, water=(150.0 if boundary_type == 1 else 110.0), boundary_type=boundary_type,
               src_x=[40], nt=600)
    rot = _rot90(flat, h)
    runs = []
    for m in (flat, rot):
        o = pyoracle.Oracle(m)
        o.reset(0)
        o.run(600)
        runs.append((o, o.records()))
    (of, rf), (orr, rr) = runs
    fields = list(FLUID) + list(SOLID) + ["FXVX", "FXVZ", "FXEX", "FXEZ", "SXVX", "SXVZ", "SXEX", "SXEZ",
                                          "SZVX", "SZVZ", "SZEX", "SZEZ"]
    if boundary_type == 1:
        fields += ["FZVX", "FZVZ", "FZEX", "FZEZ"]
    for f in fields:
        src, sign = _ROT[f]
        a, b = orr.get(f), sign * of.get(src)
        # on the flat grid the xi (zeta) projections have no z (x) component: such
        # pairs are zero on both sides
        scale = max(np.abs(b).max(), np.abs(of.get(f)).max())
        assert scale > 0 or f not in FLUID + SOLID, f
        assert np.abs(a - b).max() <= 1e-12 * scale, f
    # records: solid vx' = -vz, vz' = vx on the interface row; p unchanged
    for a, b in ((rr[0], -rf[1]), (rr[1], rf[0]), (rr[2], rf[2])):
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


# ---------------------------------------------------------------------------
# Plane waves
# ---------------------------------------------------------------------------

_AMP = {"FVX": 0.8, "FVZ": -0.6, "FP": 1.2e6, "SVX": 1.0, "SVZ": -0.7, "SSX": 2.0e6, "SSZ": -1.3e6, "SSXZ": 0.9e6}
_OFF = {"FVX": 0.4, "FVZ": 1.1, "FP": 2.3, "SVX": 0.1, "SVZ": 0.7, "SSX": 1.3, "SSZ": 2.1, "SSXZ": 2.9}


@pytest.mark.parametrize("kind", ["affine", "curvi"])
def test_interior_operator_matches_plane_wave_derivatives(kind):
    m, X, Z = _grid(_fs(), kind)
    nzf, nzs, nx, npd = m.nz_fluid, m.nz_solid, m.nx, m.npd
    o = pyoracle.Oracle(m)
    kx, kz = 2 * np.pi / 120.0, 2 * np.pi / 150.0
    rows = {"F": slice(0, nzf), "S": slice(nzf - 1, None)}
    vals, grads = {}, {}
    for k in FLUID + SOLID:
        r = rows[k[0]]
        ph = kx * X[r] + kz * Z[r] + _OFF[k]
        vals[k] = _AMP[k] * np.sin(ph)
        grads[k] = (_AMP[k] * kx * np.cos(ph), _AMP[k] * kz * np.cos(ph))
    _set_state(o, m, vals)
    met = _Metrics(m)
    sf, ss = _Stretch(o, m, met.f, False), _Stretch(o, m, met.s, True)
    exact = _physics(m, sf, ss, grads, grads)
    mem = _memory_rhs(m, sf, ss, grads, grads)
    ks = {}
    for d in (F, B):
        o.rhs(d, 0.0)
        ks[d] = {k: o.get_k(k) for k in list(exact) + list(mem)}
    # non-conservative DRP rows: fluid 6..nzf-4, solid 3..nzs-4; all updated columns
    region = {"F": (slice(6, nzf - 3), slice(3, nx - 3)), "S": (slice(3, nzs - 3), slice(3, nx - 3))}
    for k, e in exact.items():
        r = region[k[0]]
        scale = np.abs(e[r]).max()
        for d in (F, B):
            assert np.abs(ks[d][k][r] - e[r]).max() < 1e-2 * scale, (k, d)
        avg = 0.5 * (ks[F][k] + ks[B][k])
        assert np.abs(avg[r] - e[r]).max() < 2e-3 * scale, k
    # memory RHS at zero memory, over the strip cells of the same rows
    for k, e in mem.items():
        r = region[k[0]]
        strip = np.abs(e) > 0
        sel = np.zeros_like(strip)
        sel[r] = strip[r]
        scale = np.abs(e[sel]).max()
        assert scale > 0, k
        for d in (F, B):
            assert np.abs(ks[d][k][sel] - e[sel]).max() < 1e-2 * scale, (k, d)
        avg = 0.5 * (ks[F][k] + ks[B][k])
        assert np.abs(avg[sel] - e[sel]).max() < 3e-3 * scale, k
        assert not ks[F][k][~strip].any(), k        # memory is driven in its strip only


# ---------------------------------------------------------------------------
# Bilinear fields: conservative rows, closures, interface row
# ---------------------------------------------------------------------------

def _bilinear(m, z0, x0, seed, surface=False):
    """Fields c1 (xi - x0) + c2 (ze - z0) + c3 (xi - x0)(ze - z0) (+ c0) in
    computational coordinates (global row ze) and their physical gradients.
    surface: the fluid pressure vanishes on row 3 (p = (ze - 3)(a + b (xi - x0)))."""
    rng = np.random.default_rng(seed)
    nzf = m.nz_fluid
    nz = nzf + m.nz_solid - 1
    XI, ZE = np.meshgrid(np.arange(m.nx, dtype=np.float64), np.arange(nz, dtype=np.float64))
    g = {k: m.arrays[k].astype(np.float64) for k in ("xi_x", "xi_z", "ze_x", "ze_z")}
    vals, grads = {}, {}
    for k in FLUID + SOLID:
        c0, c1, c2, c3 = rng.uniform(-1, 1, 4) * np.array([0.2, 1.0, 1.0, 0.1])
        zz = z0
        if surface and k == "FP":        # p = (ze - 3) (c2 + c3 (xi - x0))
            c0, c1, zz = 0.0, 0.0, 3.0
        u, w = XI - x0, ZE - zz
        f = c0 + c1 * u + c2 * w + c3 * u * w
        fxi, fze = c1 + c3 * w, c2 + c3 * u
        gx = g["xi_x"] * fxi + g["ze_x"] * fze
        gz = g["xi_z"] * fxi + g["ze_z"] * fze
        s = slice(0, nzf) if k[0] == "F" else slice(nzf - 1, None)
        amp = _AMP[k]
        vals[k], grads[k] = amp * f[s], (amp * gx[s], amp * gz[s])
    return vals, grads


def _solid_conservative(m, ss, gs):
    """Solid rows 0-2 (conservative form): the rates and xi-strip memory RHS of the
    reference's conservative branch, which projects the divergence vector
    div sigma onto the strip normal (include/pml_fs_kernels.cuh:658-676), where
    the interior rows project each stress gradient."""
    rho = m.arrays["rho_s"].astype(np.float64)
    M = ss.met
    sx = gs["SSX"][0] + gs["SSXZ"][1]
    sz = gs["SSXZ"][0] + gs["SSZ"][1]
    px, pz = _proj(M["xi_x"], M["xi_z"], sx, sz)
    k = {"SVX": (sx + ss.sxi * px) / rho, "SVZ": (sz + ss.sxi * pz) / rho}
    mem = {"SXVX": ss.bxi * px, "SXVZ": ss.bxi * pz}
    return k, mem


@pytest.mark.parametrize("d", [F, B])
def test_boundary_rows_exact_on_bilinear_fields(d):
    """Sheared grid: surface rows 3-5 and interface rows (fluid nzf-3 .. nzf-1,
    solid 0-2) against the continuum values, in a window of interior columns and
    in the left xi strip (stretched gradient and memory RHS at zero memory)."""
    m, X, Z = _grid(_fs(random_std=0.05), "affine")
    nzf, nx, npd = m.nz_fluid, m.nx, m.npd
    met = _Metrics(m)
    n1, n2 = pyoracle.Oracle(m).normals()
    tol = 2e-3
    windows = ((nx // 2, slice(nx // 2 - 12, nx // 2 + 13)), (3 + npd // 2, slice(3, 3 + npd)))
    for x0, cols in windows:
        for surface in (True, False):
            o = pyoracle.Oracle(m)
            vals, grads = _bilinear(m, 3.0 if surface else float(nzf - 1), x0, seed=7 + surface + 2 * x0,
                                    surface=surface)
            _set_state(o, m, vals)
            sf, ss = _Stretch(o, m, met.f, False), _Stretch(o, m, met.s, True)
            ex = _physics(m, sf, ss, grads, grads)
            mem = _memory_rhs(m, sf, ss, grads, grads)
            kc, memc = _solid_conservative(m, ss, grads)
            o.rhs(d, 0.0)
            k = {f: o.get_k(f) for f in list(ex) + list(mem)}
            strip = x0 < 3 + npd

            def check(field, rows, e=None):
                e = (ex if field in ex else mem)[field] if e is None else e
                scale = np.abs(e[rows, cols]).max()
                assert scale > 0, (field, rows)
                err = np.abs(k[field][rows, cols] - e[rows, cols]).max()
                assert err < tol * scale, (field, rows, surface, d, x0, err / scale)

            if surface:
                check("FVX", slice(3, 6)); check("FVZ", slice(3, 6))   # conservative, odd image about row 3
                check("FP", slice(4, 6))                                # 2-2 and 4-2 closures
                assert not k["FP"][3].any()                             # pressure release
                if strip:
                    for f in ("FXVX", "FXVZ"):
                        check(f, slice(3, 6))
                    for f in ("FXEX", "FXEZ"):
                        check(f, slice(4, 6))
                continue
            fl, so = slice(nzf - 3, nzf - 1), slice(1, 3)
            check("FVX", fl); check("FVZ", fl)       # ghosts 2F(0) - F(-n)
            check("FP", fl)                          # 4-2, 2-2 closures
            check("SVX", so, kc["SVX"]); check("SVZ", so, kc["SVZ"])
            for f in ("SSX", "SSZ", "SSXZ"):
                check(f, so)
            if strip:
                for f in ("FXVX", "FXVZ", "FXEX", "FXEZ"):
                    check(f, fl)
                check("SXVX", so, memc["SXVX"]); check("SXVZ", so, memc["SXVZ"])
                for f in ("SXEX", "SXEZ"):
                    check(f, slice(0, 3))
            # interface row: normal from the side the direction names, tangential kept
            i = nzf - 1
            nn = lambda vx, vz: n1 * vx + n2 * vz
            tt = lambda vx, vz: n2 * vx - n1 * vz
            vf = (ex["FVX"][i], ex["FVZ"][i]); vs = (kc["SVX"][0], kc["SVZ"][0])
            sig = (ex["SSX"][0], ex["SSZ"][0], ex["SSXZ"][0])
            snn = n1 * n1 * sig[0] + 2 * n1 * n2 * sig[2] + n2 * n2 * sig[1]
            kvf = (k["FVX"][i], k["FVZ"][i]); kvs = (k["SVX"][0], k["SVZ"][0])
            vn = nn(*vs) if d == F else nn(*vf)
            c = cols
            vscale = max(np.abs(v[c]).max() for v in vf + vs)
            for got, want in ((nn(*kvf), vn), (nn(*kvs), vn), (tt(*kvf), tt(*vf)), (tt(*kvs), tt(*vs))):
                assert np.abs(got[c] - want[c]).max() < tol * vscale, (surface, d, x0)
            p_want = -snn if d == F else ex["FP"][i]
            assert np.abs(k["FP"][i][c] - p_want[c]).max() < tol * np.abs(p_want[c]).max()
            if strip:
                check("FXVX", slice(i, i + 1)); check("FXVZ", slice(i, i + 1))
                check("SXVX", slice(0, 1), memc["SXVX"]); check("SXVZ", slice(0, 1), memc["SXVZ"])


@pytest.mark.parametrize("d", [F, B])
def test_interface_stress_on_flat_seafloor(d):
    """Flat grid: after the interface condition the solid row-0 stress rates have
    sigma_xz = 0, sigma_zz = -p (F: p from the solid's sigma_zz; B: from the
    fluid's own rate) and the solid's own sigma_xx."""
    m = _fs(random_std=0.05)
    nzf, nx = m.nz_fluid, m.nx
    x0 = nx // 2
    c = slice(x0 - 12, x0 + 13)
    o = pyoracle.Oracle(m)
    vals, grads = _bilinear(m, float(nzf - 1), x0, seed=11)
    _set_state(o, m, vals)
    met = _Metrics(m)
    ex = _physics(m, _Stretch(o, m, met.f, False), _Stretch(o, m, met.s, True), grads, grads)
    o.rhs(d, 0.0)
    kf, ks = o.get_k("FP")[nzf - 1], {f: o.get_k(f)[0] for f in ("SSX", "SSZ", "SSXZ")}
    szz = ex["SSZ"][0] if d == F else -ex["FP"][nzf - 1]
    scale = np.abs(ex["SSX"][0][c]).max()
    assert np.abs(ks["SSXZ"][c]).max() < 1e-12 * scale
    assert np.abs(ks["SSZ"][c] - szz[c]).max() < 2e-3 * np.abs(szz[c]).max()
    assert np.abs(kf[c] + ks["SSZ"][c]).max() < 1e-12 * scale
    assert np.abs(ks["SSX"][c] - ex["SSX"][0][c]).max() < 2e-3 * scale


# ---------------------------------------------------------------------------
# CFS stretch of the memory ODE
# ---------------------------------------------------------------------------

def test_pml_memory_realises_the_cfs_stretch():
    """Each memory field m obeys dm/dt = a m + (d/beta) D and feeds its targets
    t with k_t += e m, while the direct path carries c D with c = gamma / beta
    (gamma: the physical coefficient of D in t; the plane-wave test pins c and
    the (d/beta) D drive).  For a harmonic drive D e^{iwt} the response is
    c + e (d/beta) / (iw - a), which must be gamma / s(w) with
    s = beta + d / (alpha + iw) at every frequency."""
    m, X, Z = _grid(_fs(), "affine")
    o = pyoracle.Oracle(m)
    nzf, nzs, nx = m.nz_fluid, m.nz_solid, m.nx
    al, be, dp = o.pml(0)
    alz, bez, dpz = o.pml(1)
    a = {k: m.arrays[k].astype(np.float64) for k in ("K", "rho_f", "C11", "C13", "C33", "C55", "rho_s")}
    met = _Metrics(m)
    # gamma[(memory, target)] on its rows: the coefficient of the memory's D component in the target
    g = {}
    for p_, M, rho in (("F", met.f, a["rho_f"]), ("S", met.s, a["rho_s"])):
        for s_, (nx_, nz_) in (("X", ("xi_x", "xi_z")), ("Z", ("ze_x", "ze_z"))):
            if p_ == "F" and s_ == "Z":
                continue
            ax, az = M[nx_], M[nz_]
            g[(f"{p_}{s_}VX", f"{p_}VX")] = 1.0 / rho
            g[(f"{p_}{s_}VZ", f"{p_}VZ")] = 1.0 / rho
            if p_ == "F":
                g[(f"F{s_}EX", "FP")] = -a["K"] * ax
                g[(f"F{s_}EZ", "FP")] = -a["K"] * az
            else:
                g[(f"S{s_}EX", "SSX")] = a["C11"] * ax
                g[(f"S{s_}EZ", "SSX")] = a["C13"] * az
                g[(f"S{s_}EX", "SSZ")] = a["C13"] * ax
                g[(f"S{s_}EZ", "SSZ")] = a["C33"] * az
                g[(f"S{s_}EX", "SSXZ")] = a["C55"] * az
                g[(f"S{s_}EZ", "SSXZ")] = a["C55"] * ax
    cols = np.arange(nx)
    side = np.minimum(cols - 3, nx - 4 - cols)
    side = (side >= 0) & (side < m.npd)
    omegas = 2 * np.pi * m.fpeak * np.array([0.05, 0.3, 1.0, 3.0, 10.0])
    for mem in sorted({k[0] for k in g}):
        o = pyoracle.Oracle(m)
        rows = nzf if mem[0] == "F" else nzs
        o.set(mem, np.ones((rows, nx)))
        o.rhs(F, 0.0)
        self_k = o.get_k(mem)
        if mem[1] == "X":
            r = slice(6, nzf - 3) if mem[0] == "F" else slice(3, nzs - 3)
            cells = np.zeros((rows, nx), bool)
            cells[r] = side[None, :]
            alpha = np.broadcast_to(al[None, :], (rows, nx)); beta = np.broadcast_to(be[None, :], (rows, nx))
            dd = np.broadcast_to(dp[None, :], (rows, nx))
        else:
            cells = np.zeros((rows, nx), bool)
            cells[nzs - 3 - m.npd:nzs - 3, 3:nx - 3] = True
            alpha = np.broadcast_to(alz[:, None], (rows, nx)); beta = np.broadcast_to(bez[:, None], (rows, nx))
            dd = np.broadcast_to(dpz[:, None], (rows, nx))
        A = self_k[cells]
        for (mm, tgt), gam in g.items():
            if mm != mem:
                continue
            e = o.get_k(tgt)[cells]
            gm, be_, al_, d_ = gam[cells], beta[cells], alpha[cells], dd[cells]
            for w in omegas:
                H = gm / be_ + e * (d_ / be_) / (1j * w - A)
                want = gm / (be_ + d_ / (al_ + 1j * w))
                assert np.abs(H - want).max() <= 1e-9 * np.abs(want).max(), (mem, tgt, w)


# ---------------------------------------------------------------------------
# Side PML absorption on the pressure-release fluid-solid grid
# ---------------------------------------------------------------------------

def test_side_pml_absorbs_what_a_wide_grid_lets_pass():
    """Homogeneous water over solid, source in the water: a 120-column grid with
    15-cell side strips against the same model 300 columns wider (nothing comes
    back from its far edges in the window).  Inside the narrow grid, the
    difference (what the side strips return) stays below 2 % of the peak, and
    at least 10 times below what the same strips return undamped (Rc = 1)."""
    nsteps, pad = 700, 150

    def run(nx, src, rc=1e-3):
        m = _fs(nx=nx, nz=70, water=150.0, npd=15, random_std=0.0, src_x=[src], fpeak=20.0, nt=nsteps, rc=rc)
        o = pyoracle.Oracle(m)
        o.reset(0)
        snaps = []
        for _ in range(nsteps // 50):
            o.run(50)
            snaps.append((o.get("FP"), o.get("SVZ")))
        return m, snaps

    mn, narrow = run(120, 40)
    _, wide = run(120 + 2 * pad, 40 + pad)
    c0, c1 = 3 + 15, 120 - 4 - 15
    peak = max(max(np.abs(p[:, c0:c1]).max() for p, _ in narrow), 1e-30)
    peak_s = max(max(np.abs(s[:, c0:c1]).max() for _, s in narrow), 1e-30)
    err = max(np.abs(pn[:, c0:c1] - pw[:, pad + c0:pad + c1]).max() for (pn, _), (pw, _) in zip(narrow, wide))
    err_s = max(np.abs(sn[:, c0:c1] - sw[:, pad + c0:pad + c1]).max() for (_, sn), (_, sw) in zip(narrow, wide))
    assert err < 0.02 * peak, err / peak
    assert err_s < 0.02 * peak_s, err_s / peak_s
    _, bare = run(120, 40, rc=1.0)
    err_b = max(np.abs(pn[:, c0:c1] - pw[:, pad + c0:pad + c1]).max() for (pn, _), (pw, _) in zip(bare, wide))
    assert err_b > 10 * err, (err_b / peak, err / peak)


# ---------------------------------------------------------------------------
# Scholte wave
# ---------------------------------------------------------------------------

def _scholte_speed(af, rf, a, b, r):
    """Root c < min(vs, vp_f) of the fluid/solid half-space secular equation
    (2 - c^2/b^2)^2 - 4 sqrt(1 - c^2/a^2) sqrt(1 - c^2/b^2)
      + (rf/r) (c^4/b^4) sqrt(1 - c^2/a^2) / sqrt(1 - c^2/af^2) = 0,
    the H -> infinity limit of tools/test.py:101-140 continued to c < af
    (tan(i x) = i tanh(x) -> i)."""
    def f(c):
        q = c * c
        return ((2 - q / b ** 2) ** 2 - 4 * math.sqrt(1 - q / a ** 2) * math.sqrt(1 - q / b ** 2)
                + (rf / r) * (q * q / b ** 4) * math.sqrt(1 - q / a ** 2) / math.sqrt(1 - q / af ** 2))
    lo, hi = 0.3 * b, min(b, af) * (1 - 1e-9)
    assert f(lo) * f(hi) < 0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(lo) * f(mid) <= 0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def test_scholte_wave_speed_along_the_flat_seafloor():
    """Soft sediment (vp 1750, vs 700, rho 1800) under 62.5 m of water, 10 Hz
    source 10 m above the seafloor: the solid vz records of the interface row,
    cross-correlated at 100 m and 400 m offset, move at the Scholte speed within
    1.5 % (the fluid loading slows the wave 8.3 % below the Rayleigh speed)."""
    h = 2.5
    m = modelgen.build_model(nx=300, nz=90, h=h, water_depth=62.5, npd=20, fpeak=10.0, dt=4e-4, nt=2500,
                             src_x=[50], src_height=10.0, solid=(1750.0, 700.0, 1800.0), fluid=(1500.0, 1000.0))
    o = pyoracle.Oracle(m)
    o.reset(0)
    o.run(2500)
    vz = o.records()[1]
    dt_rec = o.gap * m.dt
    x1, x2 = 50 + 40, 50 + 160
    a_, b_ = vz[:, x1], vz[:, x2]
    cc = np.correlate(b_, a_, mode="full")
    k = int(np.argmax(cc))
    y0, y1, y2 = cc[k - 1], cc[k], cc[k + 1]
    frac = 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)
    lag = (k - (len(a_) - 1) + frac) * dt_rec
    c = (x2 - x1) * h / lag
    cs = _scholte_speed(1500.0, 1000.0, 1750.0, 700.0, 1800.0)
    assert cs == pytest.approx(605.0, rel=1e-3)
    assert c == pytest.approx(cs, rel=0.015), (c, cs)
