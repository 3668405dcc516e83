// FP64 host oracle for the MarineSeis-FDM-2D fluid-solid time stepper.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness checker for the HIP
// solver and the CPU baseline that bench.py times beside it.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
// product path (libmsfdm.so, the CLI, the Python binding) never links or calls
// it.
//
// It restates, in double precision and with race-free (sequential) semantics,
// the reference's steppers: the fluid-solid `PmlFs::FluidSolid::update_RK4`; the
// solid-only `PmlFs::SolidOnly::update_RK4` with its traction-free top
// (include/pml_fs_kernels.cuh:8-307, src/pml_fs_kernels.cu:1-381,
// src/elements.cu:907-977); the CPML steppers `Cpml::SolidOnly` and
// `Cpml::FluidSolid` (include/cpml_kernels.cuh, src/cpml_kernels.cu) and the
// solid Dirichlet stepper (include/dirichlet_kernels.cuh, src/dirichlet_kernels.cu).
// Fluid-solid:
//   include/pml_fs_kernels.cuh:331-827   (RHS kernels, interface BC)
//   src/pml_fs_kernels.cu:383-868        (ghost fluxes, filter, source, RK4)
//   include/tools_func.cuh:6-188, src/tools_func.cu:1-66 (operators, fluxes)
//   src/elements.cu:10-40, 345-392, 824-905 (constants, RK reset, Ricker,
//                                            PML profile, interface normal)
//   src/main.cu:91-213, 245-295          (time loop, record cadence/rows)
// Every function below cites the lines it follows.
//
// Deliberate differences from the reference (documented as the race-free
// semantics the oracle defines):
//   * the interface filter reads only pre-filter values (the reference filters
//     in place and races at 128-column block seams, src/pml_fs_kernels.cu:451-524);
//   * every RHS kernel of an RK stage sees the completed previous stage (the
//     reference joins its four streams only pairwise, src/pml_fs_kernels.cu:549-556);
//   * the wavefield is reset between shots (the reference carries state over).
//
// Parity pins (tests/test_oracle_*.py): derivative operators against analytic
// derivatives of polynomials / sinusoids; ghost-flux rows against linear and odd
// extensions; flat-grid conservative == non-conservative forms; PML profile
// against the Rc reflection integral; Ricker closed-form zeros and peak;
// interface invariants after each step; normal-incidence reflection
// coefficients at the seafloor and at the pressure-release surface.  Solid
// only (tests/test_oracle_solid.py): the flat-grid ALM closed form, zero
// traction of the free-surface stress rates on curvilinear grids, the odd-image
// ghost rows, and the Rayleigh-wave speed along the free surface.  CPML and
// Dirichlet (tests/test_oracle_solid_bc.py, tests/test_oracle_pins.py): the
// curvilinear interior operator against analytic plane-wave derivatives, the
// top/bottom profile mirror, CPML absorption against the Dirichlet box, and the
// fluid top PML absorbing the upgoing pulse the pressure-release surface returns.

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

// DRP/opt MacCormack forward weights at offsets -1..3 (include/tools_func.cuh:6-10).
constexpr double DRP_M1 = -0.30874;
constexpr double DRP_0 = -0.63260;
constexpr double DRP_1 = +1.23300;
constexpr double DRP_2 = -0.33340;
constexpr double DRP_3 = +0.04168;

// Interface-row filter taps (src/elements.cu:30-33), the reference's float values.
constexpr double W7[7] = {double(1.1e-13f), double(1.678e-06f), double(0.03419f), double(0.93163f),
                          double(0.03419f), double(1.678e-06f), double(1.1e-13f)};
// The tap sum the filter divides by, accumulated in float over k = klo..khi as
// the reference does (src/pml_fs_kernels.cu:474-477).  Its rounding is part of
// the algorithm: for the seven interior taps the float sum lies 1.8e-8 below
// the exact one, so every step scales the interface rows by 1 + 1.8e-8, about
// 1.1e-4 over 6000 steps.  An exact sum would drift from the reference by that much.
double filter_wsum(int klo, int khi) {
    float s = 0.0f;
    for (int k = klo; k <= khi; ++k) s += (float)W7[k + 3];
    return s;
}

// 5x5 source spreading weights (src/elements.cu:34-40).
constexpr double SRC_W[5][5] = {
    {0.0030, 0.0133, 0.0219, 0.0133, 0.0030},
    {0.0133, 0.0596, 0.0983, 0.0596, 0.0133},
    {0.0219, 0.0983, 0.1621, 0.0983, 0.0219},
    {0.0133, 0.0596, 0.0983, 0.0596, 0.0133},
    {0.0030, 0.0133, 0.0219, 0.0133, 0.0030}};

enum Dir { F = 0, B = 1 };

// Field identifiers (shared numbering with the product's msfdm.h).
enum Field {
    FVX = 0, FVZ, FP,                  // fluid state
    FXVX, FXVZ, FXEX, FXEZ,            // fluid xi-PML memory
    SVX, SVZ, SSX, SSZ, SSXZ,          // solid state
    SXVX, SXVZ, SXEX, SXEZ,            // solid xi-PML memory
    SZVX, SZVZ, SZEX, SZEZ,            // solid zeta-PML memory (bottom)
    FZVX, FZVZ, FZEX, FZEZ,            // fluid zeta-PML memory (top; fluid-solid CPML grid)
    NFIELD
};
inline bool is_fluid(int f) { return f <= FXEZ || f >= FZVX; }

// One-sided 5-point operators along a line; f(m) returns the value at offset m
// (include/tools_func.cuh:51-157).  Unit computational spacing: metrics carry 1/h.
template <class Fn> inline double drp(int dir, Fn f) {
    if (dir == F) return DRP_M1 * f(-1) + DRP_0 * f(0) + DRP_1 * f(1) + DRP_2 * f(2) + DRP_3 * f(3);
    return -(DRP_M1 * f(1) + DRP_0 * f(0) + DRP_1 * f(-1) + DRP_2 * f(-2) + DRP_3 * f(-3));
}
template <class Fn> inline double m42(int dir, Fn f) {
    const double a0 = -7.0 / 6.0, a1 = 8.0 / 6.0, a2 = -1.0 / 6.0;
    if (dir == F) return a0 * f(0) + a1 * f(1) + a2 * f(2);
    return -(a0 * f(0) + a1 * f(-1) + a2 * f(-2));
}
template <class Fn> inline double m22(int dir, Fn f) {
    if (dir == F) return -f(0) + f(1);
    return -(-f(0) + f(-1));
}
// Distance-dependent zeta derivative (include/tools_func.cuh:131-140).
template <class Fn> inline double mac_dze(int dir, int dist, bool solid_side, Fn f) {
    if (dist >= 3) return drp(dir, f);
    if (dist == 2) return m42(dir, f);
    if (dist == 1) return m22(dir, f);
    return solid_side ? m22(F, f) : m22(B, f);
}

struct Oracle {
    int nx = 0, nzf = 0, nzs = 0, nz = 0, npd = 0;
    float dt_f = 0.f, fpeak_f = 0.f;
    int gap = 1;
    double dt = 0.0;
    // model (promoted to double)
    std::vector<double> Kf, rhof, C11, C13, C33, C55, rhos;
    std::vector<double> J, xix, xiz, zex, zez;   // global [nz][nx]
    std::vector<double> al_xi, be_xi, dp_xi, al_zs, be_zs, dp_zs;
    std::vector<double> nifx, nifz;
    std::vector<int> srcx, srcz;
    int isrc = 0;
    int it = 0;
    // state sets (reference MemManager: _d, _b, _tmp, _k)
    std::vector<double> U[NFIELD], Bacc[NFIELD], T[NFIELD], K[NFIELD];
    // Cells of each field the RHS can write, as [begin, end) runs along rows:
    // every cell of a state field, the strip cells of a PML memory field (its k
    // is written nowhere else and stays 0 outside, so the RK combination there
    // leaves U unchanged and is skipped).
    std::vector<std::pair<size_t, size_t>> runs[NFIELD];
    // ghost fluxes [3][nx] x/z components (src/pml_fs_kernels.cu:385-426)
    std::vector<double> Gfx, Gfz, Gsx, Gsz, Gfsx, Gfsz;
    // records
    std::vector<double> rec_vx, rec_vz, rec_p;
    int nrec = 0;
    // adjoint: input / output adjoint fields, model gradients (K, rho_f, C11,
    // C13, C33, C55, rho_s)
    std::vector<double> Ain[NFIELD], Aout[NFIELD];
    std::vector<double> G[7];
    // Solid-only grid (grid.calc_type = 0) with a traction-free top row 3:
    // per-column ALM coefficients al1_xx, al1_zz, al1_xz, al2_xx, al2_zz, al2_xz.
    bool solid_only = false;
    std::vector<double> alm[6];
    // grid.boundary_type: 2 = PML sides/bottom + free surface (PANDF), 1 = PML
    // on all four sides (CPML), 0 = Dirichlet (DIRI, solid-only grids).
    int bdy = 2;
    bool fcpml() const { return !solid_only && bdy == 1; }   // fluid-solid CPML grid
    std::vector<double> al_zf, be_zf, dp_zf;                  // fluid zeta PML (top strip)
    bool zpml(int iz) const {   // zeta-PML rows (bottom; CPML also top, src/elements.cu:866-869)
        if (solid_only && bdy == 0) return false;
        const bool bot = iz >= nzs - 3 - npd;
        return (solid_only && bdy == 1) ? (bot || iz <= npd + 2) : bot;
    }

    size_t fsize(int f) const { return (size_t)(is_fluid(f) ? nzf : nzs) * nx; }
    // metric index of solid row iz (src/tools_func.cu:22, 57)
    // (the whole grid is solid when nzf == 0: no offset)
    size_t gs(int iz, int ix) const { return (size_t)(iz + (nzf > 0 ? nzf - 1 : 0)) * nx + ix; }
    size_t gf(int iz, int ix) const { return (size_t)iz * nx + ix; }
};

// PML profile (src/elements.cu:824-874), evaluated in double.
void pml_profile(int n, int npd, int npow, bool lo, bool hi, double damp0, double beta0,
                 double alpha0, double *al, double *be, double *dp) {
    for (int i = 0; i < n; ++i) {
        int k = -1;
        if (lo && i >= 3 && i <= 2 + npd) k = 3 + npd - i;
        if (hi && i >= n - 3 - npd && i <= n - 4) k = i - (n - 4 - npd);
        if (k <= 0) {
            dp[i] = 0.0; be[i] = 1.0; al[i] = alpha0;
        } else {
            const double r = double(k) / double(npd);
            dp[i] = damp0 * std::pow(r, (double)npow);
            be[i] = 1.0 + (beta0 - 1.0) * r * r;
            al[i] = alpha0 * (1.0 - r);
        }
    }
}

// ---------------------------------------------------------------------------
// Fluxes (src/tools_func.cu:10-65) on double state.
// ---------------------------------------------------------------------------
inline void fxi_f(const Oracle &o, const double *p, int iz, int ix, double &fx, double &fz) {
    const size_t g = o.gf(iz, ix);
    fx = -o.J[g] * o.xix[g] * p[(size_t)iz * o.nx + ix];
    fz = -o.J[g] * o.xiz[g] * p[(size_t)iz * o.nx + ix];
}
inline void fze_f_row(const Oracle &o, const double *p, int iz, int ix, double &fx, double &fz) {
    const size_t g = o.gf(iz, ix);
    fx = -o.J[g] * o.zex[g] * p[(size_t)iz * o.nx + ix];
    fz = -o.J[g] * o.zez[g] * p[(size_t)iz * o.nx + ix];
}
inline void fxi_s(const Oracle &o, const double *const *S, int iz, int ix, double &fx, double &fz) {
    const size_t g = o.gs(iz, ix), i = (size_t)iz * o.nx + ix;
    fx = o.J[g] * (o.xix[g] * S[SSX][i] + o.xiz[g] * S[SSXZ][i]);
    fz = o.J[g] * (o.xix[g] * S[SSXZ][i] + o.xiz[g] * S[SSZ][i]);
}
inline void fze_s_row(const Oracle &o, const double *const *S, int iz, int ix, double &fx, double &fz) {
    const size_t g = o.gs(iz, ix), i = (size_t)iz * o.nx + ix;
    fx = o.J[g] * (o.zex[g] * S[SSX][i] + o.zez[g] * S[SSXZ][i]);
    fz = o.J[g] * (o.zex[g] * S[SSXZ][i] + o.zez[g] * S[SSZ][i]);
}

// Ghost fluxes (src/pml_fs_kernels.cu:385-426): Gf/Gs = 2F(interface) - F(n),
// Gfs = -F(3+n); slot n-1 holds ghost n, columns 3..nx-4.
void build_ghosts(Oracle &o, const double *const *S) {
    const int nx = o.nx;
    for (int ix = 3; ix <= nx - 4; ++ix) {
        double f0x, f0z, s0x, s0z;
        fze_f_row(o, S[FP], o.nzf - 1, ix, f0x, f0z);
        fze_s_row(o, S, 0, ix, s0x, s0z);
        for (int n = 1; n <= 3; ++n) {
            double fx, fz;
            fze_f_row(o, S[FP], o.nzf - 1 - n, ix, fx, fz);
            o.Gfx[(n - 1) * nx + ix] = 2 * f0x - fx;
            o.Gfz[(n - 1) * nx + ix] = 2 * f0z - fz;
            fze_s_row(o, S, n, ix, fx, fz);
            o.Gsx[(n - 1) * nx + ix] = 2 * s0x - fx;
            o.Gsz[(n - 1) * nx + ix] = 2 * s0z - fz;
            fze_f_row(o, S[FP], 3 + n, ix, fx, fz);
            o.Gfsx[(n - 1) * nx + ix] = -fx;
            o.Gfsz[(n - 1) * nx + ix] = -fz;
        }
    }
}

inline bool in_side(const Oracle &o, int ix) {
    if (o.solid_only && o.bdy == 0) return false;   // Dirichlet: no PML (src/elements.cu:825-827)
    const int distx = std::min(ix - 3, o.nx - 4 - ix);
    return distx >= 0 && distx < o.npd;
}

// update_velocity_fluid<DIR> (include/pml_fs_kernels.cuh:333-450).  On the
// fluid-solid CPML grid (include/cpml_kernels.cuh:239-342) only the interface
// rows take the conservative form and rows 3..npd+2 carry the top zeta memory.
void velocity_fluid(Oracle &o, const double *const *S, double *const *Kd, int dir) {
    const int nx = o.nx, nzf = o.nzf;
    const bool cpml = o.fcpml();
#pragma omp parallel for schedule(static)
    for (int iz = 3; iz < nzf; ++iz) {
        for (int ix = 3; ix <= nx - 4; ++ix) {
            const size_t idx = (size_t)iz * nx + ix, g = o.gf(iz, ix);
            const double J = o.J[g], xx = o.xix[g], xz = o.xiz[g], zx = o.zex[g], zz = o.zez[g];
            const int dist_flso = nzf - 1 - iz, dist_fs = iz - 3;
            const double gxixi = xx * xx + xz * xz, gzeze = zx * zx + zz * zz;
            double rvx, rvz, Dx, Dz, Dzx, Dzz;
            if (dist_flso < 3 || (!cpml && dist_fs < 3)) {
                const bool at_if = dist_flso < 3;
                auto Fxi = [&](int m, int c) {
                    double fx, fz; fxi_f(o, S[FP], iz, ix + m, fx, fz); return c ? fz : fx;
                };
                auto Fze = [&](int m, int c) {
                    const int r = iz + m;
                    double fx, fz;
                    if (at_if) {
                        if (r <= nzf - 1) fze_f_row(o, S[FP], r, ix, fx, fz);
                        else { const int n = r - (nzf - 1); fx = o.Gfx[(n - 1) * nx + ix]; fz = o.Gfz[(n - 1) * nx + ix]; }
                    } else {
                        if (r >= 3) fze_f_row(o, S[FP], r, ix, fx, fz);
                        else { const int n = 3 - r; fx = o.Gfsx[(n - 1) * nx + ix]; fz = o.Gfsz[(n - 1) * nx + ix]; }
                    }
                    return c ? fz : fx;
                };
                const double sx = drp(dir, [&](int m) { return Fxi(m, 0); }) + drp(dir, [&](int m) { return Fze(m, 0); });
                const double sz = drp(dir, [&](int m) { return Fxi(m, 1); }) + drp(dir, [&](int m) { return Fze(m, 1); });
                rvx = sx / J;
                rvz = sz / J;
                Dx = xx * (xx * sx + xz * sz) / (gxixi * J);
                Dz = xz * (xx * sx + xz * sz) / (gxixi * J);
                Dzx = zx * (zx * sx + zz * sz) / (gzeze * J);
                Dzz = zz * (zx * sx + zz * sz) / (gzeze * J);
            } else {
                const double *p = S[FP];
                const double dpx = drp(dir, [&](int m) { return p[idx + m]; });
                const double dpz = drp(dir, [&](int m) { return p[idx + (std::ptrdiff_t)m * nx]; });
                rvx = -(xx * dpx + zx * dpz);
                rvz = -(xz * dpx + zz * dpz);
                const double gxize = xx * zx + xz * zz;
                Dx = -xx * (gxixi * dpx + gxize * dpz) / gxixi;
                Dz = -xz * (gxixi * dpx + gxize * dpz) / gxixi;
                Dzx = -zx * (gxize * dpx + gzeze * dpz) / gzeze;
                Dzz = -zz * (gxize * dpx + gzeze * dpz) / gzeze;
            }
            const double rho = o.rhof[idx];
            Kd[FVX][idx] = rvx / rho;
            Kd[FVZ][idx] = rvz / rho;
            if (in_side(o, ix)) {
                const double be = o.be_xi[ix], dpb = o.dp_xi[ix] / be, apdp = o.al_xi[ix] + dpb, invb = 1.0 / be;
                Kd[FXVX][idx] = -apdp * S[FXVX][idx] + dpb * Dx;
                Kd[FXVZ][idx] = -apdp * S[FXVZ][idx] + dpb * Dz;
                Kd[FVX][idx] += ((invb - 1.0) * Dx - invb * S[FXVX][idx]) / rho;
                Kd[FVZ][idx] += ((invb - 1.0) * Dz - invb * S[FXVZ][idx]) / rho;
            }
            if (cpml && iz <= o.npd + 2) {
                const double be = o.be_zf[iz], dpb = o.dp_zf[iz] / be, apdp = o.al_zf[iz] + dpb, invb = 1.0 / be;
                Kd[FZVX][idx] = -apdp * S[FZVX][idx] + dpb * Dzx;
                Kd[FZVZ][idx] = -apdp * S[FZVZ][idx] + dpb * Dzz;
                Kd[FVX][idx] += ((invb - 1.0) * Dzx - invb * S[FZVX][idx]) / rho;
                Kd[FVZ][idx] += ((invb - 1.0) * Dzz - invb * S[FZVZ][idx]) / rho;
            }
        }
    }
}

// update_stress_fluid<DIR> (include/pml_fs_kernels.cuh:452-507).  On the
// fluid-solid CPML grid (include/cpml_kernels.cuh:344-415) there is no
// pressure-release row: the zeta operator closes only at the interface and
// rows 3..npd+2 carry the top zeta memory.
void stress_fluid(Oracle &o, const double *const *S, double *const *Kd, int dir) {
    const int nx = o.nx, nzf = o.nzf;
    const bool cpml = o.fcpml();
#pragma omp parallel for schedule(static)
    for (int iz = 3; iz < nzf; ++iz) {
        for (int ix = 3; ix <= nx - 4; ++ix) {
            const size_t idx = (size_t)iz * nx + ix, g = o.gf(iz, ix);
            if (!cpml && iz == 3) { Kd[FP][idx] = 0.0; continue; }   // pressure-release row
            const double *vx = S[FVX], *vz = S[FVZ];
            const double dvxx = drp(dir, [&](int m) { return vx[idx + m]; });
            const double dvzx = drp(dir, [&](int m) { return vz[idx + m]; });
            const int dzn = cpml ? nzf - 1 - iz : std::min(iz - 3, nzf - 1 - iz);
            const double dvxz = mac_dze(dir, dzn, false, [&](int m) { return vx[idx + (std::ptrdiff_t)m * nx]; });
            const double dvzz = mac_dze(dir, dzn, false, [&](int m) { return vz[idx + (std::ptrdiff_t)m * nx]; });
            const double xx = o.xix[g], xz = o.xiz[g], zx = o.zex[g], zz = o.zez[g];
            const double Km = o.Kf[idx];
            Kd[FP][idx] = -Km * (xx * dvxx + zx * dvxz + xz * dvzx + zz * dvzz);
            if (in_side(o, ix)) {
                const double gxixi = xx * xx + xz * xz, gxize = xx * zx + xz * zz;
                const double Dvx = dvxx + (gxize / gxixi) * dvxz;
                const double Dvz = dvzx + (gxize / gxixi) * dvzz;
                const double be = o.be_xi[ix], dpb = o.dp_xi[ix] / be, apdp = o.al_xi[ix] + dpb, invb = 1.0 / be;
                Kd[FXEX][idx] = -apdp * S[FXEX][idx] + dpb * Dvx;
                Kd[FXEZ][idx] = -apdp * S[FXEZ][idx] + dpb * Dvz;
                Kd[FP][idx] += Km * (-(invb - 1.0) * xx * Dvx - (invb - 1.0) * xz * Dvz
                                     + invb * xx * S[FXEX][idx] + invb * xz * S[FXEZ][idx]);
            }
            if (cpml && iz <= o.npd + 2) {
                const double gxize = xx * zx + xz * zz, gzeze = zx * zx + zz * zz;
                const double Dvx = dvxz + (gxize / gzeze) * dvxx;
                const double Dvz = dvzz + (gxize / gzeze) * dvzx;
                const double be = o.be_zf[iz], dpb = o.dp_zf[iz] / be, apdp = o.al_zf[iz] + dpb, invb = 1.0 / be;
                Kd[FZEX][idx] = -apdp * S[FZEX][idx] + dpb * Dvx;
                Kd[FZEZ][idx] = -apdp * S[FZEZ][idx] + dpb * Dvz;
                Kd[FP][idx] += Km * (-(invb - 1.0) * (zx * Dvx + zz * Dvz) + invb * (zx * S[FZEX][idx] + zz * S[FZEZ][idx]));
            }
        }
    }
}

// update_stress_solid<DIR> (include/pml_fs_kernels.cuh:509-628).
void stress_solid(Oracle &o, const double *const *S, double *const *Kd, int dir) {
    const int nx = o.nx, nzs = o.nzs, npd = o.npd;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < nzs - 3; ++iz) {
        for (int ix = 3; ix <= nx - 4; ++ix) {
            const size_t idx = (size_t)iz * nx + ix, g = o.gs(iz, ix);
            const double *vx = S[SVX], *vz = S[SVZ];
            const double dvxx = drp(dir, [&](int m) { return vx[idx + m]; });
            const double dvxz = mac_dze(dir, iz, true, [&](int m) { return vx[idx + (std::ptrdiff_t)m * nx]; });
            const double dvzx = drp(dir, [&](int m) { return vz[idx + m]; });
            const double dvzz = mac_dze(dir, iz, true, [&](int m) { return vz[idx + (std::ptrdiff_t)m * nx]; });
            const double xx = o.xix[g], xz = o.xiz[g], zx = o.zex[g], zz = o.zez[g];
            const double c11 = o.C11[idx], c13 = o.C13[idx], c33 = o.C33[idx], c55 = o.C55[idx];
            double ksx = c11 * (xx * dvxx + zx * dvxz) + c13 * (xz * dvzx + zz * dvzz);
            double ksz = c13 * (xx * dvxx + zx * dvxz) + c33 * (xz * dvzx + zz * dvzz);
            double ksxz;
            if (o.fcpml()) {
                // include/cpml_kernels.cuh:447-452 reads this term's metrics at the
                // solid-array index (global row iz, not iz + nzf - 1); kept for parity
                const size_t q = idx;
                ksxz = c55 * (o.xiz[q] * dvxx + o.zez[q] * dvxz + o.xix[q] * dvzx + o.zex[q] * dvzz);
            } else {
                ksxz = c55 * (xz * dvxx + zz * dvxz + xx * dvzx + zx * dvzz);
            }
            const bool side = in_side(o, ix), bottom = iz >= nzs - 3 - npd;
            if (side || bottom) {
                const double gxixi = xx * xx + xz * xz, gxize = xx * zx + xz * zz, gzeze = zx * zx + zz * zz;
                if (side) {
                    const double Dvx = dvxx + (gxize / gxixi) * dvxz;
                    const double Dvz = dvzx + (gxize / gxixi) * dvzz;
                    const double be = o.be_xi[ix], dpb = o.dp_xi[ix] / be, apdp = o.al_xi[ix] + dpb, invb = 1.0 / be;
                    const double Ex = S[SXEX][idx], Ez = S[SXEZ][idx];
                    Kd[SXEX][idx] = -apdp * Ex + dpb * Dvx;
                    Kd[SXEZ][idx] = -apdp * Ez + dpb * Dvz;
                    ksx += c11 * (invb - 1.0) * xx * Dvx + c13 * (invb - 1.0) * xz * Dvz - invb * c11 * xx * Ex - invb * c13 * xz * Ez;
                    ksz += c13 * (invb - 1.0) * xx * Dvx + c33 * (invb - 1.0) * xz * Dvz - invb * c13 * xx * Ex - invb * c33 * xz * Ez;
                    ksxz += c55 * (invb - 1.0) * (xz * Dvx + xx * Dvz) - invb * c55 * xz * Ex - invb * c55 * xx * Ez;
                }
                if (bottom) {
                    const double Dvx = dvxz + (gxize / gzeze) * dvxx;
                    const double Dvz = dvzz + (gxize / gzeze) * dvzx;
                    const double be = o.be_zs[iz], dpb = o.dp_zs[iz] / be, apdp = o.al_zs[iz] + dpb, invb = 1.0 / be;
                    const double Ex = S[SZEX][idx], Ez = S[SZEZ][idx];
                    Kd[SZEX][idx] = -apdp * Ex + dpb * Dvx;
                    Kd[SZEZ][idx] = -apdp * Ez + dpb * Dvz;
                    ksx += c11 * (invb - 1.0) * zx * Dvx + c13 * (invb - 1.0) * zz * Dvz - invb * c11 * zx * Ex - invb * c13 * zz * Ez;
                    ksz += c13 * (invb - 1.0) * zx * Dvx + c33 * (invb - 1.0) * zz * Dvz - invb * c13 * zx * Ex - invb * c33 * zz * Ez;
                    ksxz += c55 * (invb - 1.0) * (zz * Dvx + zx * Dvz) - invb * c55 * zz * Ex - invb * c55 * zx * Ez;
                }
            }
            Kd[SSX][idx] = ksx;
            Kd[SSZ][idx] = ksz;
            Kd[SSXZ][idx] = ksxz;
        }
    }
}

// update_velocity_solid<DIR> (include/pml_fs_kernels.cuh:631-752).
void velocity_solid(Oracle &o, const double *const *S, double *const *Kd, int dir) {
    const int nx = o.nx, nzs = o.nzs, npd = o.npd;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < nzs - 3; ++iz) {
        for (int ix = 3; ix <= nx - 4; ++ix) {
            const size_t idx = (size_t)iz * nx + ix, g = o.gs(iz, ix);
            const double J = o.J[g], xx = o.xix[g], xz = o.xiz[g], zx = o.zex[g], zz = o.zez[g];
            const double gxixi = xx * xx + xz * xz, gzeze = zx * zx + zz * zz;
            double rvx, rvz, Dxx, Dxz, Dzx, Dzz;
            if (iz < 3) {
                auto Fxi = [&](int m, int c) {
                    double fx, fz; fxi_s(o, S, iz, ix + m, fx, fz); return c ? fz : fx;
                };
                auto Fze = [&](int m, int c) {
                    const int r = iz + m;
                    double fx, fz;
                    if (r >= 0) fze_s_row(o, S, r, ix, fx, fz);
                    else { const int n = -r; fx = o.Gsx[(n - 1) * nx + ix]; fz = o.Gsz[(n - 1) * nx + ix]; }
                    return c ? fz : fx;
                };
                const double sx = drp(dir, [&](int m) { return Fxi(m, 0); }) + drp(dir, [&](int m) { return Fze(m, 0); });
                const double sz = drp(dir, [&](int m) { return Fxi(m, 1); }) + drp(dir, [&](int m) { return Fze(m, 1); });
                rvx = sx / J;
                rvz = sz / J;
                Dxx = xx * (xx * sx + xz * sz) / (gxixi * J);
                Dxz = xz * (xx * sx + xz * sz) / (gxixi * J);
                Dzx = zx * (zx * sx + zz * sz) / (gzeze * J);
                Dzz = zz * (zx * sx + zz * sz) / (gzeze * J);
            } else {
                const double *sxx = S[SSX], *szz = S[SSZ], *sxz = S[SSXZ];
                auto dx = [&](const double *f) { return drp(dir, [&](int m) { return f[idx + m]; }); };
                auto dz = [&](const double *f) { return drp(dir, [&](int m) { return f[idx + (std::ptrdiff_t)m * nx]; }); };
                const double a = dx(sxx), b = dz(sxx), c = dx(szz), d = dz(szz), e = dx(sxz), f = dz(sxz);
                rvx = xx * a + zx * b + xz * e + zz * f;
                rvz = xx * e + zx * f + xz * c + zz * d;
                const double gxize = xx * zx + xz * zz;
                Dxx = xx * (a + gxize / gxixi * b) + xz * (e + gxize / gxixi * f);
                Dxz = xx * (e + gxize / gxixi * f) + xz * (c + gxize / gxixi * d);
                Dzx = zx * (b + gxize / gzeze * a) + zz * (gxize / gzeze * e + f);
                Dzz = zx * (gxize / gzeze * e + f) + zz * (gxize / gzeze * c + d);
            }
            const double rho = o.rhos[idx];
            double kvx = rvx / rho, kvz = rvz / rho;
            if (in_side(o, ix)) {
                const double be = o.be_xi[ix], dpb = o.dp_xi[ix] / be, apdp = o.al_xi[ix] + dpb, invb = 1.0 / be;
                Kd[SXVX][idx] = -apdp * S[SXVX][idx] + dpb * Dxx;
                Kd[SXVZ][idx] = -apdp * S[SXVZ][idx] + dpb * Dxz;
                kvx += ((invb - 1.0) * Dxx - invb * S[SXVX][idx]) / rho;
                kvz += ((invb - 1.0) * Dxz - invb * S[SXVZ][idx]) / rho;
            }
            if (iz >= nzs - 3 - npd) {
                const double be = o.be_zs[iz], dpb = o.dp_zs[iz] / be, apdp = o.al_zs[iz] + dpb, invb = 1.0 / be;
                Kd[SZVX][idx] = -apdp * S[SZVX][idx] + dpb * Dzx;
                Kd[SZVZ][idx] = -apdp * S[SZVZ][idx] + dpb * Dzz;
                kvx += ((invb - 1.0) * Dzx - invb * S[SZVX][idx]) / rho;
                kvz += ((invb - 1.0) * Dzz - invb * S[SZVZ][idx]) / rho;
            }
            Kd[SVX][idx] = kvx;
            Kd[SVZ][idx] = kvz;
        }
    }
}

// add_gauss_source (include/tools_func.cuh:21-37) on p_k, fluid-array indices.
void add_source(Oracle &o, double *pk, double amp) {
    if (o.srcx.empty()) return;
    const int px = o.srcx[o.isrc], pz = o.srcz[o.isrc];
    for (int j = -2; j <= 2; ++j)
        for (int i = -2; i <= 2; ++i) {
            const int ix = px + i, iz = pz + j;
            if (ix <= 2 || ix >= o.nx - 3 || iz <= 2 || iz >= o.nzf - 3) continue;
            pk[(size_t)iz * o.nx + ix] += amp * SRC_W[j + 2][i + 2];
        }
}

// fl_so_boundary<DIR> (include/pml_fs_kernels.cuh:754-827) on the stage RHS.
// The solid sxz reconstruction keeps the reference's expression, which omits
// the nx*nz*snn term of the inverse rotation (exact on flat interfaces).
void fl_so_boundary(Oracle &o, double *const *Kd, int dir) {
    const int nx = o.nx;
    for (int ix = 3; ix <= nx - 4; ++ix) {
        const size_t f = (size_t)(o.nzf - 1) * nx + ix, s = ix;
        const double n1 = o.nifx[ix], n2 = o.nifz[ix];
        const double vxf = Kd[FVX][f], vzf = Kd[FVZ][f];
        double pf = Kd[FP][f];
        const double vxs = Kd[SVX][s], vzs = Kd[SVZ][s], sxx = Kd[SSX][s], szz = Kd[SSZ][s], sxz = Kd[SSXZ][s];
        double vnf = n1 * vxf + n2 * vzf;
        const double vtf = n2 * vxf - n1 * vzf;
        double vns = n1 * vxs + n2 * vzs;
        const double vts = n2 * vxs - n1 * vzs;
        double snn = n1 * n1 * sxx + 2 * n1 * n2 * sxz + n2 * n2 * szz;
        double snt = n1 * n2 * (sxx - szz) + (n2 * n2 - n1 * n1) * sxz;
        const double stt = n2 * n2 * sxx - 2 * n1 * n2 * sxz + n1 * n1 * szz;
        if (dir == F) { vnf = vns; pf = -snn; snt = 0; }
        else { vns = vnf; snn = -pf; snt = 0; }
        Kd[FVX][f] = n1 * vnf + n2 * vtf;
        Kd[FVZ][f] = n2 * vnf - n1 * vtf;
        Kd[FP][f] = pf;
        Kd[SVX][s] = n1 * vns + n2 * vts;
        Kd[SVZ][s] = n2 * vns - n1 * vts;
        Kd[SSX][s] = n1 * n1 * snn + 2 * n1 * n2 * snt + n2 * n2 * stt;
        Kd[SSZ][s] = n2 * n2 * snn - 2 * n1 * n2 * snt + n1 * n1 * stt;
        Kd[SSXZ][s] = -n1 * n2 * stt + (n2 * n2 - n1 * n1) * snt;
    }
}

// Fields a grid never drives: the fluid zeta memory outside the fluid-solid
// CPML grid (src/elements.cu:866-873: no fluid zeta PML with the
// pressure-release surface).  Their k, stage values and state stay 0, so the
// stepper skips them.
inline bool driven(const Oracle &o, int f) { return o.fsize(f) > 0 && (f < FZVX || o.fcpml()); }
void zero_runs(const Oracle &o, int f, double *a) {
    const auto &r = o.runs[f];
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < r.size(); ++q)
        for (size_t i = r[q].first; i < r[q].second; ++i) a[i] = 0.0;
}

// linear_operate + fl_so_apply (src/pml_fs_kernels.cu:533-660): k = BC(L(S) + src).
void rhs(Oracle &o, double *const *S, double *const *Kd, int dir, double amp) {
    for (int f = 0; f < NFIELD; ++f) zero_runs(o, f, Kd[f]);
    build_ghosts(o, S);
    velocity_fluid(o, S, Kd, dir);
    stress_fluid(o, S, Kd, dir);
    velocity_solid(o, S, Kd, dir);
    stress_solid(o, S, Kd, dir);
    add_source(o, Kd[FP], amp);
    fl_so_boundary(o, Kd, dir);
}

// ---------------------------------------------------------------------------
// Solid-only grid with a traction-free top (grid.calc_type = 0, boundary_type
// = 2): PmlFs::SolidOnly (include/pml_fs_kernels.cuh:8-307,
// src/pml_fs_kernels.cu:1-381).  Rows 0-2 frozen, row 3 the free surface,
// rows 3..nzs-4 and columns 3..nx-4 updated; xi-PML on both sides, zeta-PML
// at the bottom.
// ---------------------------------------------------------------------------

// MemManager::build_free_surface, calc_type == SOLID (src/elements.cu:907-977):
// the traction-free condition at row 3, Dz dv/dzeta + Dx dv/dxi = 0, solved
// for dv/dzeta = -M dv/dxi (M = Dz^-1 Dx), folded into the strain rates:
// exx = al1_xx vx_xi + al2_xx vz_xi, ezz = al1_zz vx_xi + al2_zz vz_xi,
// 2exz = al1_xz vx_xi + al2_xz vz_xi.  Double from the float inputs.
void build_alm(Oracle &o) {
    for (auto &a : o.alm) a.assign(o.nx, 0.0);
    for (int ix = 0; ix < o.nx; ++ix) {
        const size_t i = (size_t)3 * o.nx + ix, g = o.gs(3, ix);
        const double xx = o.xix[g], xz = o.xiz[g], zx = o.zex[g], zz = o.zez[g];
        const double c11 = o.C11[i], c13 = o.C13[i], c33 = o.C33[i], c55 = o.C55[i];
        const double dz11 = c11 * zx * zx + c55 * zz * zz;
        const double dz12 = (c13 + c55) * zx * zz;
        const double dz22 = c55 * zx * zx + c33 * zz * zz;
        const double dx11 = c11 * zx * xx + c55 * zz * xz;
        const double dx12 = c13 * zx * xz + c55 * zz * xx;
        const double dx21 = c13 * zz * xx + c55 * zx * xz;
        const double dx22 = c55 * zx * xx + c33 * zz * xz;
        const double det = dz11 * dz22 - dz12 * dz12;
        double m11 = 0, m12 = 0, m21 = 0, m22 = 0;
        if (std::abs(det) >= 1e-14) {
            const double i11 = dz22 / det, i12 = -dz12 / det, i22 = dz11 / det;
            m11 = i11 * dx11 + i12 * dx21;
            m12 = i11 * dx12 + i12 * dx22;
            m21 = i12 * dx11 + i22 * dx21;
            m22 = i12 * dx12 + i22 * dx22;
        }
        o.alm[0][ix] = xx - zx * m11;
        o.alm[1][ix] = -zz * m21;
        o.alm[2][ix] = xz - zz * m11 - zx * m21;
        o.alm[3][ix] = -zx * m12;
        o.alm[4][ix] = xz - zz * m22;
        o.alm[5][ix] = xx - zz * m12 - zx * m22;
    }
}

// build_ghost_flux (src/pml_fs_kernels.cu:7-19): Gfs ghost n = -F^zeta(3 + n)
// of the solid, slot n-1, columns 3..nx-4.
void build_ghosts_free(Oracle &o, const double *const *S) {
    const int nx = o.nx;
    for (int ix = 3; ix <= nx - 4; ++ix)
        for (int n = 1; n <= 3; ++n) {
            double fx, fz;
            fze_s_row(o, S, 3 + n, ix, fx, fz);
            o.Gfsx[(n - 1) * nx + ix] = -fx;
            o.Gfsz[(n - 1) * nx + ix] = -fz;
        }
}

// update_stress<DIR> (include/pml_fs_kernels.cuh:10-187).  CPML
// (include/cpml_kernels.cuh:10-128) and Dirichlet (include/dirichlet_kernels.cuh:10-47)
// grids take the non-surface branch on every row (DRP in zeta throughout), with
// PML on all four sides (CPML) or none (Dirichlet).
void stress_solid_free(Oracle &o, const double *const *S, double *const *Kd, int dir) {
    const int nx = o.nx, nzs = o.nzs;
    const bool free_top = o.bdy == 2;
#pragma omp parallel for schedule(static)
    for (int iz = 3; iz <= nzs - 4; ++iz) {
        for (int ix = 3; ix <= nx - 4; ++ix) {
            const size_t idx = (size_t)iz * nx + ix, g = o.gs(iz, ix);
            const double *vx = S[SVX], *vz = S[SVZ];
            const int dist = free_top ? iz - 3 : 3;
            const double dvxx = drp(dir, [&](int m) { return vx[idx + m]; });
            const double dvzx = drp(dir, [&](int m) { return vz[idx + m]; });
            const double c11 = o.C11[idx], c13 = o.C13[idx], c33 = o.C33[idx], c55 = o.C55[idx];
            const bool side = in_side(o, ix), bottom = o.zpml(iz);
            if (dist == 0) {
                const double a1xx = o.alm[0][ix], a1zz = o.alm[1][ix], a1xz = o.alm[2][ix];
                const double a2xx = o.alm[3][ix], a2zz = o.alm[4][ix], a2xz = o.alm[5][ix];
                const double cxx_x = c11 * a1xx + c13 * a1zz, cxx_z = c11 * a2xx + c13 * a2zz;
                const double czz_x = c13 * a1xx + c33 * a1zz, czz_z = c13 * a2xx + c33 * a2zz;
                double ksx = cxx_x * dvxx + cxx_z * dvzx;
                double ksz = czz_x * dvxx + czz_z * dvzx;
                double ksxz = c55 * (a1xz * dvxx + a2xz * dvzx);
                if (side) {
                    const double be = o.be_xi[ix], dpb = o.dp_xi[ix] / be, apdp = o.al_xi[ix] + dpb, invb = 1.0 / be;
                    const double Ex = S[SXEX][idx], Ez = S[SXEZ][idx];
                    Kd[SXEX][idx] = -apdp * Ex + dpb * dvxx;
                    Kd[SXEZ][idx] = -apdp * Ez + dpb * dvzx;
                    ksx += (invb - 1.0) * (cxx_x * dvxx + cxx_z * dvzx) - invb * (cxx_x * Ex + cxx_z * Ez);
                    ksz += (invb - 1.0) * (czz_x * dvxx + czz_z * dvzx) - invb * (czz_x * Ex + czz_z * Ez);
                    ksxz += c55 * ((invb - 1.0) * (a1xz * dvxx + a2xz * dvzx) - invb * (a1xz * Ex + a2xz * Ez));
                }
                Kd[SSX][idx] = ksx;
                Kd[SSZ][idx] = ksz;
                Kd[SSXZ][idx] = ksxz;
                continue;
            }
            const double dvxz = mac_dze(dir, dist, true, [&](int m) { return vx[idx + (std::ptrdiff_t)m * nx]; });
            const double dvzz = mac_dze(dir, dist, true, [&](int m) { return vz[idx + (std::ptrdiff_t)m * nx]; });
            const double xx = o.xix[g], xz = o.xiz[g], zx = o.zex[g], zz = o.zez[g];
            double ksx = c11 * (xx * dvxx + zx * dvxz) + c13 * (xz * dvzx + zz * dvzz);
            double ksz = c13 * (xx * dvxx + zx * dvxz) + c33 * (xz * dvzx + zz * dvzz);
            double ksxz = c55 * (xz * dvxx + zz * dvxz + xx * dvzx + zx * dvzz);
            if (side || bottom) {
                const double gxixi = xx * xx + xz * xz, gxize = xx * zx + xz * zz, gzeze = zx * zx + zz * zz;
                if (side) {
                    const double Dvx = dvxx + (gxize / gxixi) * dvxz;
                    const double Dvz = dvzx + (gxize / gxixi) * dvzz;
                    const double be = o.be_xi[ix], dpb = o.dp_xi[ix] / be, apdp = o.al_xi[ix] + dpb, invb = 1.0 / be;
                    const double Ex = S[SXEX][idx], Ez = S[SXEZ][idx];
                    Kd[SXEX][idx] = -apdp * Ex + dpb * Dvx;
                    Kd[SXEZ][idx] = -apdp * Ez + dpb * Dvz;
                    ksx += c11 * (invb - 1.0) * xx * Dvx + c13 * (invb - 1.0) * xz * Dvz - invb * c11 * xx * Ex - invb * c13 * xz * Ez;
                    ksz += c13 * (invb - 1.0) * xx * Dvx + c33 * (invb - 1.0) * xz * Dvz - invb * c13 * xx * Ex - invb * c33 * xz * Ez;
                    ksxz += c55 * (invb - 1.0) * (xz * Dvx + xx * Dvz) - invb * c55 * xz * Ex - invb * c55 * xx * Ez;
                }
                if (bottom) {
                    const double Dvx = dvxz + (gxize / gzeze) * dvxx;
                    const double Dvz = dvzz + (gxize / gzeze) * dvzx;
                    const double be = o.be_zs[iz], dpb = o.dp_zs[iz] / be, apdp = o.al_zs[iz] + dpb, invb = 1.0 / be;
                    const double Ex = S[SZEX][idx], Ez = S[SZEZ][idx];
                    Kd[SZEX][idx] = -apdp * Ex + dpb * Dvx;
                    Kd[SZEZ][idx] = -apdp * Ez + dpb * Dvz;
                    ksx += c11 * (invb - 1.0) * zx * Dvx + c13 * (invb - 1.0) * zz * Dvz - invb * c11 * zx * Ex - invb * c13 * zz * Ez;
                    ksz += c13 * (invb - 1.0) * zx * Dvx + c33 * (invb - 1.0) * zz * Dvz - invb * c13 * zx * Ex - invb * c33 * zz * Ez;
                    ksxz += c55 * (invb - 1.0) * (zz * Dvx + zx * Dvz) - invb * c55 * zz * Ex - invb * c55 * zx * Ez;
                }
            }
            Kd[SSX][idx] = ksx;
            Kd[SSZ][idx] = ksz;
            Kd[SSXZ][idx] = ksxz;
        }
    }
}

// update_velocity<DIR> (include/pml_fs_kernels.cuh:189-305).  Rows 3..5 take
// the conservative flux form with the odd-image ghosts about row 3 and the
// xi-memory source from distance-dependent zeta derivatives of the stresses
// (no zeta-memory source there); deeper rows the non-conservative form.  CPML
// and Dirichlet grids (include/cpml_kernels.cuh:130-217,
// include/dirichlet_kernels.cuh:49-83) take the non-conservative form on every row.
void velocity_solid_free(Oracle &o, const double *const *S, double *const *Kd, int dir) {
    const int nx = o.nx, nzs = o.nzs;
    const bool free_top = o.bdy == 2;
#pragma omp parallel for schedule(static)
    for (int iz = 3; iz <= nzs - 4; ++iz) {
        for (int ix = 3; ix <= nx - 4; ++ix) {
            const size_t idx = (size_t)iz * nx + ix, g = o.gs(iz, ix);
            const double J = o.J[g], xx = o.xix[g], xz = o.xiz[g], zx = o.zex[g], zz = o.zez[g];
            const double gxixi = xx * xx + xz * xz, gzeze = zx * zx + zz * zz, gxize = xx * zx + xz * zz;
            const double *sxx = S[SSX], *szz = S[SSZ], *sxz = S[SSXZ];
            const int dist = free_top ? iz - 3 : 3;
            double rvx, rvz, Dxx, Dxz, Dzx, Dzz;
            if (dist < 3) {
                auto Fxi = [&](int m, int c) {
                    double fx, fz; fxi_s(o, S, iz, ix + m, fx, fz); return c ? fz : fx;
                };
                auto Fze = [&](int m, int c) {
                    const int r = iz + m;
                    double fx, fz;
                    if (r >= 3) fze_s_row(o, S, r, ix, fx, fz);
                    else { const int n = 3 - r; fx = o.Gfsx[(n - 1) * nx + ix]; fz = o.Gfsz[(n - 1) * nx + ix]; }
                    return c ? fz : fx;
                };
                const double sx = drp(dir, [&](int m) { return Fxi(m, 0); }) + drp(dir, [&](int m) { return Fze(m, 0); });
                const double sz = drp(dir, [&](int m) { return Fxi(m, 1); }) + drp(dir, [&](int m) { return Fze(m, 1); });
                rvx = sx / J;
                rvz = sz / J;
                auto dx = [&](const double *f) { return drp(dir, [&](int m) { return f[idx + m]; }); };
                auto dz = [&](const double *f) {
                    return mac_dze(dir, dist, true, [&](int m) { return f[idx + (std::ptrdiff_t)m * nx]; });
                };
                const double dxsx = dx(sxx), dxsz = dx(szz), dxsxz = dx(sxz);
                const double dzsx = dz(sxx), dzsz = dz(szz), dzsxz = dz(sxz);
                Dxx = xx * (dxsx + gxize / gxixi * dzsx) + xz * (dxsxz + gxize / gxixi * dzsxz);
                Dxz = xx * (dxsxz + gxize / gxixi * dzsxz) + xz * (dxsz + gxize / gxixi * dzsz);
                Dzx = 0.0;
                Dzz = 0.0;
            } else {
                auto dx = [&](const double *f) { return drp(dir, [&](int m) { return f[idx + m]; }); };
                auto dz = [&](const double *f) { return drp(dir, [&](int m) { return f[idx + (std::ptrdiff_t)m * nx]; }); };
                const double a = dx(sxx), b = dz(sxx), c = dx(szz), d = dz(szz), e = dx(sxz), f = dz(sxz);
                rvx = xx * a + zx * b + xz * e + zz * f;
                rvz = xx * e + zx * f + xz * c + zz * d;
                Dxx = xx * (a + gxize / gxixi * b) + xz * (e + gxize / gxixi * f);
                Dxz = xx * (e + gxize / gxixi * f) + xz * (c + gxize / gxixi * d);
                Dzx = zx * (b + gxize / gzeze * a) + zz * (gxize / gzeze * e + f);
                Dzz = zx * (gxize / gzeze * e + f) + zz * (gxize / gzeze * c + d);
            }
            const double rho = o.rhos[idx];
            double kvx = rvx / rho, kvz = rvz / rho;
            if (in_side(o, ix)) {
                const double be = o.be_xi[ix], dpb = o.dp_xi[ix] / be, apdp = o.al_xi[ix] + dpb, invb = 1.0 / be;
                Kd[SXVX][idx] = -apdp * S[SXVX][idx] + dpb * Dxx;
                Kd[SXVZ][idx] = -apdp * S[SXVZ][idx] + dpb * Dxz;
                kvx += ((invb - 1.0) * Dxx - invb * S[SXVX][idx]) / rho;
                kvz += ((invb - 1.0) * Dxz - invb * S[SXVZ][idx]) / rho;
            }
            if (o.zpml(iz)) {
                const double be = o.be_zs[iz], dpb = o.dp_zs[iz] / be, apdp = o.al_zs[iz] + dpb, invb = 1.0 / be;
                Kd[SZVX][idx] = -apdp * S[SZVX][idx] + dpb * Dzx;
                Kd[SZVZ][idx] = -apdp * S[SZVZ][idx] + dpb * Dzz;
                kvx += ((invb - 1.0) * Dzx - invb * S[SZVX][idx]) / rho;
                kvz += ((invb - 1.0) * Dzz - invb * S[SZVZ][idx]) / rho;
            }
            Kd[SVX][idx] = kvx;
            Kd[SVZ][idx] = kvz;
        }
    }
}

// add_source (src/pml_fs_kernels.cu:22-27): the 5x5 Gaussian-weighted Ricker
// on both normal stresses, solid-array indices, skipping rows <= 2 and >= nzs-3.
void add_source_solid(Oracle &o, double *const *Kd, double amp) {
    if (o.srcx.empty()) return;
    const int px = o.srcx[o.isrc], pz = o.srcz[o.isrc];
    for (int j = -2; j <= 2; ++j)
        for (int i = -2; i <= 2; ++i) {
            const int ix = px + i, iz = pz + j;
            if (ix <= 2 || ix >= o.nx - 3 || iz <= 2 || iz >= o.nzs - 3) continue;
            const double w = amp * SRC_W[j + 2][i + 2];
            Kd[SSX][(size_t)iz * o.nx + ix] += w;
            Kd[SSZ][(size_t)iz * o.nx + ix] += w;
        }
}

// linear_operate (src/pml_fs_kernels.cu:29-103, src/cpml_kernels.cu,
// src/dirichlet_kernels.cu:12-56): k = L(S) + src.
void rhs_free(Oracle &o, double *const *S, double *const *Kd, int dir, double amp) {
    for (int f = 0; f < NFIELD; ++f) zero_runs(o, f, Kd[f]);
    if (o.bdy == 2) build_ghosts_free(o, S);
    velocity_solid_free(o, S, Kd, dir);
    stress_solid_free(o, S, Kd, dir);
    add_source_solid(o, Kd, amp);
}

// fl_so_filter (src/pml_fs_kernels.cu:428-525), out of place.
void filter(Oracle &o, double *const *Ud) {
    const int nx = o.nx, le = 3, ri = nx - 4;
    const size_t rf = (size_t)(o.nzf - 1) * nx;
    const double *src[8] = {Ud[FVX] + rf, Ud[FVZ] + rf, Ud[FP] + rf, Ud[SVX], Ud[SVZ], Ud[SSX], Ud[SSZ], Ud[SSXZ]};
    std::vector<double> out(8 * (size_t)nx, 0.0);
    for (int ix = le; ix <= ri; ++ix) {
        const int klo = std::max(-3, le - ix), khi = std::min(3, ri - ix);
        const double wsum = filter_wsum(klo, khi);
        double v[8];
        for (int f = 0; f < 8; ++f) {
            double s = 0.0;
            for (int k = klo; k <= khi; ++k) s += W7[k + 3] * src[f][ix + k];
            v[f] = s / wsum;
        }
        const double n1 = o.nifx[ix], n2 = o.nifz[ix];
        const double vnf = n1 * v[0] + n2 * v[1], vtf = n2 * v[0] - n1 * v[1];
        const double vns = n1 * v[3] + n2 * v[4], vts = n2 * v[3] - n1 * v[4];
        const double snn = n1 * n1 * v[5] + 2.0 * n1 * n2 * v[7] + n2 * n2 * v[6];
        const double stt = n2 * n2 * v[5] - 2.0 * n1 * n2 * v[7] + n1 * n1 * v[6];
        const double vn = 0.5 * (vnf + vns), sn = 0.5 * (snn - v[2]);
        out[0 * nx + ix] = n1 * vn + n2 * vtf;
        out[1 * nx + ix] = n2 * vn - n1 * vtf;
        out[2 * nx + ix] = -sn;
        out[3 * nx + ix] = n1 * vn + n2 * vts;
        out[4 * nx + ix] = n2 * vn - n1 * vts;
        out[5 * nx + ix] = n1 * n1 * sn + n2 * n2 * stt;
        out[6 * nx + ix] = n2 * n2 * sn + n1 * n1 * stt;
        out[7 * nx + ix] = n1 * n2 * (sn - stt);
    }
    double *dst[8] = {Ud[FVX] + rf, Ud[FVZ] + rf, Ud[FP] + rf, Ud[SVX], Ud[SVZ], Ud[SSX], Ud[SSZ], Ud[SSXZ]};
    for (int f = 0; f < 8; ++f)
        for (int ix = le; ix <= ri; ++ix) dst[f][ix] = out[(size_t)f * nx + ix];
}

// Ricker at a float stage time (src/elements.cu:385-392): T frozen as float
// 1.2/fpeak, zero once t - T > T; the shape itself is evaluated in double.
double ricker(const Oracle &o, float time) {
    const float T = (float)(1.2 / (double)o.fpeak_f);
    const float t = time - T;
    if (t > T) return 0.0;
    double a = M_PI * (double)o.fpeak_f * (double)t;
    a *= a;
    return (1.0 - 2.0 * a) * std::exp(-a);
}
// Stage times from float it*dt (+dt/2, +dt) (src/pml_fs_kernels.cu:768-844).
double stage_amp(const Oracle &o, int it, int stage) {
    const float time = (float)it * o.dt_f;
    const float hf = (float)(o.dt_f * 0.5);
    float t = time;
    if (stage == 1 || stage == 2) t = time + hf;
    if (stage == 3) t = time + o.dt_f;
    return ricker(o, t);
}

// dst = a + c k, dst += c k, dst = a over the runs of field f.
void axpy(const Oracle &o, int f, std::vector<double> &dst, const std::vector<double> &a, double c,
          const std::vector<double> &k) {
    const auto &r = o.runs[f];
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < r.size(); ++q)
        for (size_t i = r[q].first; i < r[q].second; ++i) dst[i] = a[i] + c * k[i];
}
void self_axpy(const Oracle &o, int f, std::vector<double> &dst, double c, const std::vector<double> &k) {
    const auto &r = o.runs[f];
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < r.size(); ++q)
        for (size_t i = r[q].first; i < r[q].second; ++i) dst[i] += c * k[i];
}
void axpy_all(std::vector<double> &dst, const std::vector<double> &a, double c, const std::vector<double> &k) {
    const size_t n = dst.size();
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) dst[i] = a[i] + c * k[i];
}
void self_axpy_all(std::vector<double> &dst, double c, const std::vector<double> &k) {
    const size_t n = dst.size();
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) dst[i] += c * k[i];
}
void copy_runs(const Oracle &o, int f, std::vector<double> &dst, const std::vector<double> &a) {
    const auto &r = o.runs[f];
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < r.size(); ++q)
        for (size_t i = r[q].first; i < r[q].second; ++i) dst[i] = a[i];
}

// The runs of every field (see Oracle::runs): whole rows of the state fields;
// the xi-strip columns of the xi memory (every row); the zeta-strip rows of the
// zeta memory (every column).  Supersets of the cells the RHS functions write.
void build_runs(Oracle &o) {
    const size_t nx = o.nx;
    for (int f = 0; f < NFIELD; ++f) {
        auto &r = o.runs[f];
        r.clear();
        if (!driven(o, f)) continue;
        const int rows = (int)(o.fsize(f) / nx);
        const bool xi = (f >= FXVX && f <= FXEZ) || (f >= SXVX && f <= SXEZ);
        const bool ze = (f >= SZVX && f <= SZEZ) || f >= FZVX;
        for (int iz = 0; iz < rows; ++iz) {
            const size_t b = (size_t)iz * nx;
            if (!xi && !ze) { r.emplace_back(b, b + nx); continue; }
            if (ze) {
                const bool on = f >= FZVX ? (o.fcpml() && iz >= 3 && iz <= o.npd + 2) : o.zpml(iz);
                if (on) r.emplace_back(b, b + nx);
                continue;
            }
            for (int ix = 0; ix < o.nx;) {
                if (!in_side(o, ix)) { ++ix; continue; }
                int e = ix;
                while (e < o.nx && in_side(o, e)) ++e;
                r.emplace_back(b + ix, b + e);
                ix = e;
            }
        }
    }
}

// One time step it: RK4_reset (src/elements.cu:345-383), update_RK4
// (src/pml_fs_kernels.cu:765-868), record (src/main.cu:245-295).
void step(Oracle &o) {
    double *S[NFIELD], *Kd[NFIELD], *Ud[NFIELD];
    for (int f = 0; f < NFIELD; ++f) {
        copy_runs(o, f, o.Bacc[f], o.U[f]);
        copy_runs(o, f, o.T[f], o.U[f]);
        S[f] = o.T[f].data();
        Kd[f] = o.K[f].data();
        Ud[f] = o.U[f].data();
    }
    const int d1 = (o.it & 1) ? B : F, d2 = (o.it & 1) ? F : B;
    const double dt = o.dt;
    const int dirs[4] = {d1, d2, d1, d2};
    const double ct[3] = {dt / 2, dt / 2, dt};
    const double cb[3] = {dt / 6, dt / 3, dt / 3};
    for (int s = 0; s < 4; ++s) {
        if (o.solid_only) rhs_free(o, S, Kd, dirs[s], stage_amp(o, o.it, s));
        else rhs(o, S, Kd, dirs[s], stage_amp(o, o.it, s));
        for (int f = 0; f < NFIELD; ++f) {
            if (s < 3) {
                axpy(o, f, o.T[f], o.U[f], ct[s], o.K[f]);
                self_axpy(o, f, o.Bacc[f], cb[s], o.K[f]);
            } else {
                axpy(o, f, o.U[f], o.Bacc[f], dt / 6, o.K[f]);
            }
        }
    }
    if (o.solid_only) {
        // SOLID records: vz of row 3 (PANDF) or row 30 (CPML); none for
        // Dirichlet (src/main.cu:199-209, 297-322)
        if (o.bdy != 0 && o.it % o.gap == 0) {
            const size_t r = (size_t)(o.bdy == 1 ? 30 : 3) * o.nx;
            o.rec_vz.insert(o.rec_vz.end(), o.U[SVZ].begin() + r, o.U[SVZ].begin() + r + o.nx);
            o.nrec++;
        }
        o.it++;
        return;
    }
    filter(o, Ud);
    if (o.it % o.gap == 0) {
        const size_t rf = (size_t)(o.nzf - 1) * o.nx;
        o.rec_vx.insert(o.rec_vx.end(), o.U[SVX].begin(), o.U[SVX].begin() + o.nx);
        o.rec_vz.insert(o.rec_vz.end(), o.U[SVZ].begin(), o.U[SVZ].begin() + o.nx);
        o.rec_p.insert(o.rec_p.end(), o.U[FP].begin() + rf, o.U[FP].begin() + rf + o.nx);
        o.nrec++;
    }
    o.it++;
}


// ===========================================================================
// Discrete adjoint (transpose of the scheme above, race-free semantics).
// Scatter form: every forward expression k[i] = sum_j a_ij S[j] becomes
// Sbar[j] += a_ij * qbar[i], branch by branch.  Model gradients are the
// pointwise terms dk/dm . qbar evaluated on the forward stage state.
// ===========================================================================
enum Grad { G_K = 0, G_RHOF, G_C11, G_C13, G_C33, G_C55, G_RHOS };

// Adjoint of the x / z DRP line operators: f at offset m gets w_m * g.
template <class Add> inline void drp_adj(int dir, double g, Add add) {
    if (dir == F) { add(-1, DRP_M1 * g); add(0, DRP_0 * g); add(1, DRP_1 * g); add(2, DRP_2 * g); add(3, DRP_3 * g); }
    else { add(1, -DRP_M1 * g); add(0, -DRP_0 * g); add(-1, -DRP_1 * g); add(-2, -DRP_2 * g); add(-3, -DRP_3 * g); }
}
template <class Add> inline void mac_dze_adj(int dir, int dist, bool solid_side, double g, Add add) {
    if (dist >= 3) { drp_adj(dir, g, add); return; }
    if (dist == 2) {
        const double a0 = -7.0 / 6.0, a1 = 8.0 / 6.0, a2 = -1.0 / 6.0;
        if (dir == F) { add(0, a0 * g); add(1, a1 * g); add(2, a2 * g); }
        else { add(0, -a0 * g); add(-1, -a1 * g); add(-2, -a2 * g); }
        return;
    }
    int d = dir;
    if (dist == 0) d = solid_side ? F : B;
    if (d == F) { add(0, -g); add(1, g); }
    else { add(0, g); add(-1, -g); }
}

// Transpose of update_velocity_fluid (include/pml_fs_kernels.cuh:333-450);
// with S the forward stage state, also dk/d(rho_f) . qbar.
void velocity_fluid_adj(Oracle &o, const double *const *S, const double *const *Q, double *const *Sb, int dir) {
    const int nx = o.nx, nzf = o.nzf;
    for (int iz = 3; iz < nzf; ++iz) {
        for (int ix = 3; ix <= nx - 4; ++ix) {
            const size_t idx = (size_t)iz * nx + ix, g = o.gf(iz, ix);
            const double J = o.J[g], xx = o.xix[g], xz = o.xiz[g], zx = o.zex[g], zz = o.zez[g];
            const int dist_flso = nzf - 1 - iz, dist_fs = iz - 3;
            const double gxixi = xx * xx + xz * xz, rho = o.rhof[idx];
            const double qvx = Q[FVX][idx], qvz = Q[FVZ][idx];
            const bool side = in_side(o, ix);
            double Dxb = 0, Dzb = 0;
            double rvxb = qvx / rho, rvzb = qvz / rho;
            double Dx = 0, Dz = 0, rvx = 0, rvz = 0;   // forward values for the gradient
            double qVx = 0, qVz = 0, invb = 1, dpb = 0, apdp = 0;
            if (side) {
                const double be = o.be_xi[ix];
                dpb = o.dp_xi[ix] / be; apdp = o.al_xi[ix] + dpb; invb = 1.0 / be;
                qVx = Q[FXVX][idx]; qVz = Q[FXVZ][idx];
                Dxb = (invb - 1.0) / rho * qvx + dpb * qVx;
                Dzb = (invb - 1.0) / rho * qvz + dpb * qVz;
                Sb[FXVX][idx] += -invb / rho * qvx - apdp * qVx;
                Sb[FXVZ][idx] += -invb / rho * qvz - apdp * qVz;
            }
            if (dist_flso < 3 || dist_fs < 3) {
                const bool at_if = dist_flso < 3;
                // forward (for the gradient)
                auto fze_fwd = [&](int r, double &fx, double &fz) {
                    if (at_if) {
                        if (r <= nzf - 1) fze_f_row(o, S[FP], r, ix, fx, fz);
                        else { double ax, az, bx, bz; fze_f_row(o, S[FP], nzf - 1, ix, ax, az); fze_f_row(o, S[FP], 2 * (nzf - 1) - r, ix, bx, bz); fx = 2 * ax - bx; fz = 2 * az - bz; }
                    } else {
                        if (r >= 3) fze_f_row(o, S[FP], r, ix, fx, fz);
                        else { double bx, bz; fze_f_row(o, S[FP], 6 - r, ix, bx, bz); fx = -bx; fz = -bz; }
                    }
                };
                const double sx = drp(dir, [&](int m) { double a, b; fxi_f(o, S[FP], iz, ix + m, a, b); return a; }) +
                                  drp(dir, [&](int m) { double a, b; fze_fwd(iz + m, a, b); return a; });
                const double sz = drp(dir, [&](int m) { double a, b; fxi_f(o, S[FP], iz, ix + m, a, b); return b; }) +
                                  drp(dir, [&](int m) { double a, b; fze_fwd(iz + m, a, b); return b; });
                rvx = sx / J; rvz = sz / J;
                const double proj = (xx * sx + xz * sz) / (gxixi * J);
                Dx = xx * proj; Dz = xz * proj;
                // reverse
                const double projb = xx * Dxb + xz * Dzb;
                const double sxb = rvxb / J + xx * projb / (gxixi * J);
                const double szb = rvzb / J + xz * projb / (gxixi * J);
                // xi fluxes at (iz, ix+m): -J p (xi_x, xi_z)
                drp_adj(dir, 1.0, [&](int m, double w) {
                    const size_t gq = o.gf(iz, ix + m);
                    Sb[FP][(size_t)iz * nx + ix + m] += -o.J[gq] * (o.xix[gq] * w * sxb + o.xiz[gq] * w * szb);
                });
                auto fze_add = [&](int r, double bx, double bz) {   // bx, bz: adjoint of the flux at real row r
                    const size_t gq = o.gf(r, ix);
                    Sb[FP][(size_t)r * nx + ix] += -o.J[gq] * (o.zex[gq] * bx + o.zez[gq] * bz);
                };
                drp_adj(dir, 1.0, [&](int m, double w) {
                    const int r = iz + m;
                    const double bx = w * sxb, bz = w * szb;
                    if (at_if) {
                        if (r <= nzf - 1) fze_add(r, bx, bz);
                        else { fze_add(nzf - 1, 2 * bx, 2 * bz); fze_add(2 * (nzf - 1) - r, -bx, -bz); }
                    } else {
                        if (r >= 3) fze_add(r, bx, bz);
                        else fze_add(6 - r, -bx, -bz);
                    }
                });
            } else {
                const double *p = S[FP];
                const double dpx = drp(dir, [&](int m) { return p[idx + m]; });
                const double dpz = drp(dir, [&](int m) { return p[idx + (std::ptrdiff_t)m * nx]; });
                const double gxize = xx * zx + xz * zz;
                rvx = -(xx * dpx + zx * dpz); rvz = -(xz * dpx + zz * dpz);
                const double t = (gxixi * dpx + gxize * dpz) / gxixi;
                Dx = -xx * t; Dz = -xz * t;
                const double tb = -xx * Dxb - xz * Dzb;
                const double dpxb = -xx * rvxb - xz * rvzb + tb;
                const double dpzb = -zx * rvxb - zz * rvzb + tb * gxize / gxixi;
                drp_adj(dir, dpxb, [&](int m, double v) { Sb[FP][idx + m] += v; });
                drp_adj(dir, dpzb, [&](int m, double v) { Sb[FP][idx + (std::ptrdiff_t)m * nx] += v; });
            }
            // gradient wrt rho_f: k_v = X / rho  =>  dk/drho = -k_v / rho
            double kvx = rvx, kvz = rvz;
            if (side) {
                kvx += (invb - 1.0) * Dx - invb * S[FXVX][idx];
                kvz += (invb - 1.0) * Dz - invb * S[FXVZ][idx];
            }
            o.G[G_RHOF][idx] += -(qvx * kvx + qvz * kvz) / (rho * rho);
        }
    }
}

// Transpose of update_stress_fluid (include/pml_fs_kernels.cuh:452-507) + dk/dK.
void stress_fluid_adj(Oracle &o, const double *const *S, const double *const *Q, double *const *Sb, int dir) {
    const int nx = o.nx, nzf = o.nzf;
    for (int iz = 4; iz < nzf; ++iz) {   // row 3: p_k = 0, no dependence
        for (int ix = 3; ix <= nx - 4; ++ix) {
            const size_t idx = (size_t)iz * nx + ix, g = o.gf(iz, ix);
            const double xx = o.xix[g], xz = o.xiz[g], zx = o.zex[g], zz = o.zez[g];
            const double Km = o.Kf[idx], qp = Q[FP][idx];
            const int dzn = std::min(iz - 3, nzf - 1 - iz);
            const double *vx = S[FVX], *vz = S[FVZ];
            const double dvxx = drp(dir, [&](int m) { return vx[idx + m]; });
            const double dvzx = drp(dir, [&](int m) { return vz[idx + m]; });
            const double dvxz = mac_dze(dir, dzn, false, [&](int m) { return vx[idx + (std::ptrdiff_t)m * nx]; });
            const double dvzz = mac_dze(dir, dzn, false, [&](int m) { return vz[idx + (std::ptrdiff_t)m * nx]; });
            double bracket = -(xx * dvxx + zx * dvxz + xz * dvzx + zz * dvzz);
            double a_xx = -Km * xx * qp, a_xz = -Km * zx * qp, a_zx = -Km * xz * qp, a_zz = -Km * zz * qp;
            if (in_side(o, ix)) {
                const double gxixi = xx * xx + xz * xz, gxize = xx * zx + xz * zz, gg = gxize / gxixi;
                const double be = o.be_xi[ix], dpb = o.dp_xi[ix] / be, apdp = o.al_xi[ix] + dpb, invb = 1.0 / be;
                const double qEx = Q[FXEX][idx], qEz = Q[FXEZ][idx];
                const double Dvxb = -Km * (invb - 1.0) * xx * qp + dpb * qEx;
                const double Dvzb = -Km * (invb - 1.0) * xz * qp + dpb * qEz;
                Sb[FXEX][idx] += Km * invb * xx * qp - apdp * qEx;
                Sb[FXEZ][idx] += Km * invb * xz * qp - apdp * qEz;
                a_xx += Dvxb; a_xz += gg * Dvxb; a_zx += Dvzb; a_zz += gg * Dvzb;
                const double Dvx = dvxx + gg * dvxz, Dvz = dvzx + gg * dvzz;
                bracket += -(invb - 1.0) * xx * Dvx - (invb - 1.0) * xz * Dvz + invb * xx * S[FXEX][idx] + invb * xz * S[FXEZ][idx];
            }
            drp_adj(dir, a_xx, [&](int m, double v) { Sb[FVX][idx + m] += v; });
            drp_adj(dir, a_zx, [&](int m, double v) { Sb[FVZ][idx + m] += v; });
            mac_dze_adj(dir, dzn, false, a_xz, [&](int m, double v) { Sb[FVX][idx + (std::ptrdiff_t)m * nx] += v; });
            mac_dze_adj(dir, dzn, false, a_zz, [&](int m, double v) { Sb[FVZ][idx + (std::ptrdiff_t)m * nx] += v; });
            o.G[G_K][idx] += qp * bracket;
        }
    }
}

// Transpose of update_stress_solid (include/pml_fs_kernels.cuh:509-628) + dk/dC.
void stress_solid_adj(Oracle &o, const double *const *S, const double *const *Q, double *const *Sb, int dir) {
    const int nx = o.nx, nzs = o.nzs, npd = o.npd;
    for (int iz = 0; iz < nzs - 3; ++iz) {
        for (int ix = 3; ix <= nx - 4; ++ix) {
            const size_t idx = (size_t)iz * nx + ix, g = o.gs(iz, ix);
            const double *vx = S[SVX], *vz = S[SVZ];
            const double dvxx = drp(dir, [&](int m) { return vx[idx + m]; });
            const double dvxz = mac_dze(dir, iz, true, [&](int m) { return vx[idx + (std::ptrdiff_t)m * nx]; });
            const double dvzx = drp(dir, [&](int m) { return vz[idx + m]; });
            const double dvzz = mac_dze(dir, iz, true, [&](int m) { return vz[idx + (std::ptrdiff_t)m * nx]; });
            const double xx = o.xix[g], xz = o.xiz[g], zx = o.zex[g], zz = o.zez[g];
            const double c11 = o.C11[idx], c13 = o.C13[idx], c33 = o.C33[idx], c55 = o.C55[idx];
            const double qsx = Q[SSX][idx], qsz = Q[SSZ][idx], qsxz = Q[SSXZ][idx];
            const double exx = xx * dvxx + zx * dvxz, ezz = xz * dvzx + zz * dvzz;
            double A11 = exx, A13 = ezz, A55 = xz * dvxx + zz * dvxz + xx * dvzx + zx * dvzz;
            const double exxb = c11 * qsx + c13 * qsz, ezzb = c13 * qsx + c33 * qsz;
            double bxx = xx * exxb + c55 * xz * qsxz, bxz = zx * exxb + c55 * zz * qsxz;
            double bzx = xz * ezzb + c55 * xx * qsxz, bzz = zz * ezzb + c55 * zx * qsxz;
            const bool side = in_side(o, ix), bottom = iz >= nzs - 3 - npd;
            const double gxixi = xx * xx + xz * xz, gxize = xx * zx + xz * zz, gzeze = zx * zx + zz * zz;
            if (side) {
                const double gx = gxize / gxixi;
                const double be = o.be_xi[ix], dpb = o.dp_xi[ix] / be, apdp = o.al_xi[ix] + dpb, invb = 1.0 / be, q1 = invb - 1.0;
                const double Ex = S[SXEX][idx], Ez = S[SXEZ][idx], qEx = Q[SXEX][idx], qEz = Q[SXEZ][idx];
                const double Dvxb = q1 * xx * exxb + c55 * q1 * xz * qsxz + dpb * qEx;
                const double Dvzb = q1 * xz * ezzb + c55 * q1 * xx * qsxz + dpb * qEz;
                Sb[SXEX][idx] += -invb * xx * exxb - invb * c55 * xz * qsxz - apdp * qEx;
                Sb[SXEZ][idx] += -invb * xz * ezzb - invb * c55 * xx * qsxz - apdp * qEz;
                bxx += Dvxb; bxz += gx * Dvxb; bzx += Dvzb; bzz += gx * Dvzb;
                const double Dvx = dvxx + gx * dvxz, Dvz = dvzx + gx * dvzz;
                A11 += q1 * xx * Dvx - invb * xx * Ex;
                A13 += q1 * xz * Dvz - invb * xz * Ez;
                A55 += q1 * (xz * Dvx + xx * Dvz) - invb * xz * Ex - invb * xx * Ez;
            }
            if (bottom) {
                const double gz = gxize / gzeze;
                const double be = o.be_zs[iz], dpb = o.dp_zs[iz] / be, apdp = o.al_zs[iz] + dpb, invb = 1.0 / be, q1 = invb - 1.0;
                const double Ex = S[SZEX][idx], Ez = S[SZEZ][idx], qEx = Q[SZEX][idx], qEz = Q[SZEZ][idx];
                const double Dvxb = q1 * zx * exxb + c55 * q1 * zz * qsxz + dpb * qEx;
                const double Dvzb = q1 * zz * ezzb + c55 * q1 * zx * qsxz + dpb * qEz;
                Sb[SZEX][idx] += -invb * zx * exxb - invb * c55 * zz * qsxz - apdp * qEx;
                Sb[SZEZ][idx] += -invb * zz * ezzb - invb * c55 * zx * qsxz - apdp * qEz;
                bxz += Dvxb; bxx += gz * Dvxb; bzz += Dvzb; bzx += gz * Dvzb;
                const double Dvx = dvxz + gz * dvxx, Dvz = dvzz + gz * dvzx;
                A11 += q1 * zx * Dvx - invb * zx * Ex;
                A13 += q1 * zz * Dvz - invb * zz * Ez;
                A55 += q1 * (zz * Dvx + zx * Dvz) - invb * zz * Ex - invb * zx * Ez;
            }
            drp_adj(dir, bxx, [&](int m, double v) { Sb[SVX][idx + m] += v; });
            drp_adj(dir, bzx, [&](int m, double v) { Sb[SVZ][idx + m] += v; });
            mac_dze_adj(dir, iz, true, bxz, [&](int m, double v) { Sb[SVX][idx + (std::ptrdiff_t)m * nx] += v; });
            mac_dze_adj(dir, iz, true, bzz, [&](int m, double v) { Sb[SVZ][idx + (std::ptrdiff_t)m * nx] += v; });
            o.G[G_C11][idx] += qsx * A11;
            o.G[G_C13][idx] += qsx * A13 + qsz * A11;
            o.G[G_C33][idx] += qsz * A13;
            o.G[G_C55][idx] += qsxz * A55;
        }
    }
}

// Transpose of update_velocity_solid (include/pml_fs_kernels.cuh:631-752) + dk/d(rho_s).
void velocity_solid_adj(Oracle &o, const double *const *S, const double *const *Q, double *const *Sb, int dir) {
    const int nx = o.nx, nzs = o.nzs, npd = o.npd;
    for (int iz = 0; iz < nzs - 3; ++iz) {
        for (int ix = 3; ix <= nx - 4; ++ix) {
            const size_t idx = (size_t)iz * nx + ix, g = o.gs(iz, ix);
            const double J = o.J[g], xx = o.xix[g], xz = o.xiz[g], zx = o.zex[g], zz = o.zez[g];
            const double gxixi = xx * xx + xz * xz, gzeze = zx * zx + zz * zz, gxize = xx * zx + xz * zz;
            const double rho = o.rhos[idx], qvx = Q[SVX][idx], qvz = Q[SVZ][idx];
            const bool side = in_side(o, ix), bottom = iz >= nzs - 3 - npd;
            const double rvxb = qvx / rho, rvzb = qvz / rho;
            double Dxxb = 0, Dxzb = 0, Dzxb = 0, Dzzb = 0;
            double q1x = 0, pxc = 1, q1z = 0, pzc = 1;
            if (side) {
                const double be = o.be_xi[ix], dpb = o.dp_xi[ix] / be, apdp = o.al_xi[ix] + dpb;
                pxc = 1.0 / be; q1x = pxc - 1.0;
                const double qVx = Q[SXVX][idx], qVz = Q[SXVZ][idx];
                Dxxb = q1x / rho * qvx + dpb * qVx;
                Dxzb = q1x / rho * qvz + dpb * qVz;
                Sb[SXVX][idx] += -pxc / rho * qvx - apdp * qVx;
                Sb[SXVZ][idx] += -pxc / rho * qvz - apdp * qVz;
            }
            if (bottom) {
                const double be = o.be_zs[iz], dpb = o.dp_zs[iz] / be, apdp = o.al_zs[iz] + dpb;
                pzc = 1.0 / be; q1z = pzc - 1.0;
                const double qVx = Q[SZVX][idx], qVz = Q[SZVZ][idx];
                Dzxb = q1z / rho * qvx + dpb * qVx;
                Dzzb = q1z / rho * qvz + dpb * qVz;
                Sb[SZVX][idx] += -pzc / rho * qvx - apdp * qVx;
                Sb[SZVZ][idx] += -pzc / rho * qvz - apdp * qVz;
            }
            double rvx, rvz, Dxx, Dxz, Dzx, Dzz;
            if (iz < 3) {
                auto fze_fwd = [&](int r, double &fx, double &fz) {
                    if (r >= 0) fze_s_row(o, S, r, ix, fx, fz);
                    else { double ax, az, bx, bz; fze_s_row(o, S, 0, ix, ax, az); fze_s_row(o, S, -r, ix, bx, bz); fx = 2 * ax - bx; fz = 2 * az - bz; }
                };
                const double sx = drp(dir, [&](int m) { double a, b; fxi_s(o, S, iz, ix + m, a, b); return a; }) +
                                  drp(dir, [&](int m) { double a, b; fze_fwd(iz + m, a, b); return a; });
                const double sz = drp(dir, [&](int m) { double a, b; fxi_s(o, S, iz, ix + m, a, b); return b; }) +
                                  drp(dir, [&](int m) { double a, b; fze_fwd(iz + m, a, b); return b; });
                rvx = sx / J; rvz = sz / J;
                const double pxi = (xx * sx + xz * sz) / (gxixi * J), pze = (zx * sx + zz * sz) / (gzeze * J);
                Dxx = xx * pxi; Dxz = xz * pxi; Dzx = zx * pze; Dzz = zz * pze;
                const double pxib = xx * Dxxb + xz * Dxzb, pzeb = zx * Dzxb + zz * Dzzb;
                const double sxb = rvxb / J + xx * pxib / (gxixi * J) + zx * pzeb / (gzeze * J);
                const double szb = rvzb / J + xz * pxib / (gxixi * J) + zz * pzeb / (gzeze * J);
                drp_adj(dir, 1.0, [&](int m, double w) {
                    const size_t gq = o.gs(iz, ix + m), i = idx + m;
                    const double Fx = w * sxb, Fz = w * szb;   // adjoints of the xi flux components
                    Sb[SSX][i] += o.J[gq] * o.xix[gq] * Fx;
                    Sb[SSXZ][i] += o.J[gq] * (o.xiz[gq] * Fx + o.xix[gq] * Fz);
                    Sb[SSZ][i] += o.J[gq] * o.xiz[gq] * Fz;
                });
                auto fze_add = [&](int r, double Fx, double Fz) {
                    const size_t gq = o.gs(r, ix), i = (size_t)r * nx + ix;
                    Sb[SSX][i] += o.J[gq] * o.zex[gq] * Fx;
                    Sb[SSXZ][i] += o.J[gq] * (o.zez[gq] * Fx + o.zex[gq] * Fz);
                    Sb[SSZ][i] += o.J[gq] * o.zez[gq] * Fz;
                };
                drp_adj(dir, 1.0, [&](int m, double w) {
                    const int r = iz + m;
                    const double Fx = w * sxb, Fz = w * szb;
                    if (r >= 0) fze_add(r, Fx, Fz);
                    else { fze_add(0, 2 * Fx, 2 * Fz); fze_add(-r, -Fx, -Fz); }
                });
            } else {
                const double *sxx = S[SSX], *szz = S[SSZ], *sxz = S[SSXZ];
                auto dx = [&](const double *f) { return drp(dir, [&](int m) { return f[idx + m]; }); };
                auto dz = [&](const double *f) { return drp(dir, [&](int m) { return f[idx + (std::ptrdiff_t)m * nx]; }); };
                const double a = dx(sxx), b = dz(sxx), c = dx(szz), d = dz(szz), e = dx(sxz), f = dz(sxz);
                const double gx = gxize / gxixi, gz = gxize / gzeze;
                rvx = xx * a + zx * b + xz * e + zz * f;
                rvz = xx * e + zx * f + xz * c + zz * d;
                Dxx = xx * (a + gx * b) + xz * (e + gx * f);
                Dxz = xx * (e + gx * f) + xz * (c + gx * d);
                Dzx = zx * (b + gz * a) + zz * (gz * e + f);
                Dzz = zx * (gz * e + f) + zz * (gz * c + d);
                const double ab = xx * rvxb + xx * Dxxb + zx * gz * Dzxb;
                const double bb = zx * rvxb + xx * gx * Dxxb + zx * Dzxb;
                const double cb = xz * rvzb + xz * Dxzb + zz * gz * Dzzb;
                const double db = zz * rvzb + xz * gx * Dxzb + zz * Dzzb;
                const double eb = xz * rvxb + xx * rvzb + xz * Dxxb + xx * Dxzb + zz * gz * Dzxb + zx * gz * Dzzb;
                const double fb = zz * rvxb + zx * rvzb + xz * gx * Dxxb + xx * gx * Dxzb + zz * Dzxb + zx * Dzzb;
                auto ax = [&](int fld, double gv) { drp_adj(dir, gv, [&](int m, double v) { Sb[fld][idx + m] += v; }); };
                auto az = [&](int fld, double gv) { drp_adj(dir, gv, [&](int m, double v) { Sb[fld][idx + (std::ptrdiff_t)m * nx] += v; }); };
                ax(SSX, ab); az(SSX, bb); ax(SSZ, cb); az(SSZ, db); ax(SSXZ, eb); az(SSXZ, fb);
            }
            double kvx = rvx, kvz = rvz;
            if (side) { kvx += q1x * Dxx - pxc * S[SXVX][idx]; kvz += q1x * Dxz - pxc * S[SXVZ][idx]; }
            if (bottom) { kvx += q1z * Dzx - pzc * S[SZVX][idx]; kvz += q1z * Dzz - pzc * S[SZVZ][idx]; }
            o.G[G_RHOS][idx] += -(qvx * kvx + qvz * kvz) / (rho * rho);
        }
    }
}

// The interface BC and the filter are per-column 8x8 linear maps; their
// transposes are built by probing the forward maps with unit vectors.
void bc_matrix(double n1, double n2, int dir, double M[8][8]) {
    for (int j = 0; j < 8; ++j) {
        double k[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        k[j] = 1.0;
        const double vxf = k[0], vzf = k[1], vxs = k[3], vzs = k[4], sxx = k[5], szz = k[6], sxz = k[7];
        double pf = k[2];
        double vnf = n1 * vxf + n2 * vzf;
        const double vtf = n2 * vxf - n1 * vzf;
        double vns = n1 * vxs + n2 * vzs;
        const double vts = n2 * vxs - n1 * vzs;
        double snn = n1 * n1 * sxx + 2 * n1 * n2 * sxz + n2 * n2 * szz;
        double snt = n1 * n2 * (sxx - szz) + (n2 * n2 - n1 * n1) * sxz;
        const double stt = n2 * n2 * sxx - 2 * n1 * n2 * sxz + n1 * n1 * szz;
        if (dir == F) { vnf = vns; pf = -snn; snt = 0; }
        else { vns = vnf; snn = -pf; snt = 0; }
        const double out[8] = {n1 * vnf + n2 * vtf, n2 * vnf - n1 * vtf, pf, n1 * vns + n2 * vts, n2 * vns - n1 * vts,
                               n1 * n1 * snn + 2 * n1 * n2 * snt + n2 * n2 * stt,
                               n2 * n2 * snn - 2 * n1 * n2 * snt + n1 * n1 * stt,
                               -n1 * n2 * stt + (n2 * n2 - n1 * n1) * snt};
        for (int i = 0; i < 8; ++i) M[i][j] = out[i];
    }
}
void filter_matrix(double n1, double n2, double M[8][8]) {
    for (int j = 0; j < 8; ++j) {
        double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        v[j] = 1.0;
        const double vnf = n1 * v[0] + n2 * v[1], vtf = n2 * v[0] - n1 * v[1];
        const double vns = n1 * v[3] + n2 * v[4], vts = n2 * v[3] - n1 * v[4];
        const double snn = n1 * n1 * v[5] + 2.0 * n1 * n2 * v[7] + n2 * n2 * v[6];
        const double stt = n2 * n2 * v[5] - 2.0 * n1 * n2 * v[7] + n1 * n1 * v[6];
        const double vn = 0.5 * (vnf + vns), sn = 0.5 * (snn - v[2]);
        const double out[8] = {n1 * vn + n2 * vtf, n2 * vn - n1 * vtf, -sn, n1 * vn + n2 * vts, n2 * vn - n1 * vts,
                               n1 * n1 * sn + n2 * n2 * stt, n2 * n2 * sn + n1 * n1 * stt, n1 * n2 * (sn - stt)};
        for (int i = 0; i < 8; ++i) M[i][j] = out[i];
    }
}

inline double *iface_ptr(Oracle &o, double *const *A, int f, int ix) {
    static const int fid[8] = {FVX, FVZ, FP, SVX, SVZ, SSX, SSZ, SSXZ};
    const size_t row = f < 3 ? (size_t)(o.nzf - 1) * o.nx : 0;
    return &A[fid[f]][row + ix];
}

// kbar -> qbar = B^T kbar on the interface rows (in place).
void bc_adj(Oracle &o, double *const *A, int dir) {
    for (int ix = 3; ix <= o.nx - 4; ++ix) {
        double M[8][8];
        bc_matrix(o.nifx[ix], o.nifz[ix], dir, M);
        double kb[8], qb[8];
        for (int i = 0; i < 8; ++i) kb[i] = *iface_ptr(o, A, i, ix);
        for (int j = 0; j < 8; ++j) { qb[j] = 0; for (int i = 0; i < 8; ++i) qb[j] += M[i][j] * kb[i]; }
        for (int j = 0; j < 8; ++j) *iface_ptr(o, A, j, ix) = qb[j];
    }
}

// lambda -> Filter^T lambda on the interface rows (out of place).
void filter_adj(Oracle &o, double *const *A) {
    const int nx = o.nx, le = 3, ri = nx - 4;
    std::vector<double> out(8 * (size_t)nx, 0.0);
    for (int ix = le; ix <= ri; ++ix) {
        double M[8][8];
        filter_matrix(o.nifx[ix], o.nifz[ix], M);
        double yb[8], vb[8];
        for (int i = 0; i < 8; ++i) yb[i] = *iface_ptr(o, A, i, ix);
        for (int j = 0; j < 8; ++j) { vb[j] = 0; for (int i = 0; i < 8; ++i) vb[j] += M[i][j] * yb[i]; }
        const int klo = std::max(-3, le - ix), khi = std::min(3, ri - ix);
        const double wsum = filter_wsum(klo, khi);
        for (int f = 0; f < 8; ++f)
            for (int k = klo; k <= khi; ++k) out[(size_t)f * nx + ix + k] += W7[k + 3] / wsum * vb[f];
    }
    for (int f = 0; f < 8; ++f)
        for (int ix = le; ix <= ri; ++ix) *iface_ptr(o, A, f, ix) = out[(size_t)f * nx + ix];
}

// Sbar += L_dir^T(qbar) with gradients on the forward stage state S.
void rhs_adj(Oracle &o, const double *const *S, const double *const *Q, double *const *Sb, int dir) {
    velocity_fluid_adj(o, S, Q, Sb, dir);
    stress_fluid_adj(o, S, Q, Sb, dir);
    velocity_solid_adj(o, S, Q, Sb, dir);
    stress_solid_adj(o, S, Q, Sb, dir);
}

struct Field20 {
    std::vector<double> v[NFIELD];
    std::vector<double> &operator[](int i) { return v[i]; }
    const std::vector<double> &operator[](int i) const { return v[i]; }
};

// Forward step keeping the stage states (U^n, T2, T3, T4) of step it.
void step_keep(Oracle &o, Field20 &T2, Field20 &T3, Field20 &T4) {
    double *S[NFIELD], *Kd[NFIELD];
    for (int f = 0; f < NFIELD; ++f) { Kd[f] = o.K[f].data(); }
    const int d1 = (o.it & 1) ? B : F, d2 = (o.it & 1) ? F : B;
    const int dirs[4] = {d1, d2, d1, d2};
    const double dt = o.dt, ct[3] = {dt / 2, dt / 2, dt};
    Field20 *outs[3] = {&T2, &T3, &T4};
    for (int f = 0; f < NFIELD; ++f) o.Bacc[f] = o.U[f];
    for (int s = 0; s < 4; ++s) {
        for (int f = 0; f < NFIELD; ++f) S[f] = s == 0 ? o.U[f].data() : (*outs[s - 1])[f].data();
        rhs(o, S, Kd, dirs[s], stage_amp(o, o.it, s));
        const double cb = s == 0 || s == 3 ? dt / 6 : dt / 3;
        for (int f = 0; f < NFIELD; ++f) {
            if (s < 3) axpy_all((*outs[s])[f], o.U[f], ct[s], o.K[f]);
            self_axpy_all(o.Bacc[f], cb, o.K[f]);
        }
    }
    for (int f = 0; f < NFIELD; ++f) o.U[f] = o.Bacc[f];
    double *Ud[NFIELD];
    for (int f = 0; f < NFIELD; ++f) Ud[f] = o.U[f].data();
    filter(o, Ud);
    o.it++;
}

// Adjoint of one step (reverse of step_keep), lam in/out.
void step_adj(Oracle &o, int it, const Field20 &Un, const Field20 &T2, const Field20 &T3, const Field20 &T4, Field20 &lam) {
    const int nfz = NFIELD;
    double *L[NFIELD];
    for (int f = 0; f < nfz; ++f) L[f] = lam[f].data();
    filter_adj(o, L);                               // lambda'
    const int d1 = (it & 1) ? B : F, d2 = (it & 1) ? F : B;
    const double dt = o.dt;
    Field20 acc, kb, tb;
    for (int f = 0; f < nfz; ++f) { acc[f] = lam[f]; kb[f].assign(lam[f].size(), 0.0); tb[f].assign(lam[f].size(), 0.0); }
    const Field20 *states[4] = {&Un, &T2, &T3, &T4};
    const int dirs[4] = {d1, d2, d1, d2};
    const double clam[4] = {dt / 6, dt / 3, dt / 3, dt / 6};   // coefficient of lambda' in kbar_s
    const double cprev[4] = {dt / 2, dt / 2, dt, 0.0};         // coefficient of Tbar_{s+1}
    for (int s = 3; s >= 0; --s) {
        for (int f = 0; f < nfz; ++f)
            for (size_t i = 0; i < kb[f].size(); ++i) kb[f][i] = clam[s] * lam[f][i] + cprev[s] * tb[f][i];
        double *K[NFIELD], *TB[NFIELD];
        const double *S[NFIELD], *Q[NFIELD];
        for (int f = 0; f < nfz; ++f) { K[f] = kb[f].data(); }
        bc_adj(o, K, dirs[s]);
        for (int f = 0; f < nfz; ++f) {
            tb[f].assign(lam[f].size(), 0.0);
            TB[f] = tb[f].data();
            S[f] = (*states[s])[f].data();
            Q[f] = kb[f].data();
        }
        rhs_adj(o, S, Q, TB, dirs[s]);
        for (int f = 0; f < nfz; ++f)
            for (size_t i = 0; i < acc[f].size(); ++i) acc[f][i] += tb[f][i];
    }
    for (int f = 0; f < nfz; ++f) lam[f].swap(acc[f]);
}

}  // namespace

// ---------------------------------------------------------------------------
// C interface (ctypes) for tests and the CPU baseline.
// ---------------------------------------------------------------------------
extern "C" {

struct orc_desc {
    int nx, nz_fluid, nz_solid, npd, npow;
    float fpeak, dt, thickness, vp_max, rc, beta0;
    int nsrc;
    const int *src_x;
    const int *src_z;
    int calc_type;       // 1 = fluid over solid (FANDS), 0 = solid only (SOLID)
    int boundary_type;   // 2 = PANDF; solid only also 1 = CPML, 0 = Dirichlet
};

// Model arrays: fluid [nzf][nx] K, rho; solid [nzs][nx] C11, C13, C33, C55,
// rho; metrics [nz][nx] J, xi_x, xi_z, ze_x, ze_z (float32, as read from the
// <f4 model files, src/elements.cu:127-176).
void *orc_create(const orc_desc *d, const float *K, const float *rhof, const float *C11,
                 const float *C13, const float *C33, const float *C55, const float *rhos,
                 const float *J, const float *xix, const float *xiz, const float *zex,
                 const float *zez) {
    Oracle *o = new Oracle();
    o->solid_only = d->calc_type == 0;
    o->bdy = d->boundary_type;
    o->nx = d->nx; o->nzf = o->solid_only ? 0 : d->nz_fluid; o->nzs = d->nz_solid; o->npd = d->npd;
    o->nz = o->solid_only ? o->nzs : o->nzf + o->nzs - 1;   // src/elements.cu:87-99
    o->dt_f = d->dt; o->fpeak_f = d->fpeak; o->dt = (double)d->dt;
    o->gap = (int)(1e-3 / (double)d->dt + 0.5);   // src/main.cu:92
    if (o->gap < 1) o->gap = 1;
    const size_t nf = (size_t)o->nzf * o->nx, ns = (size_t)o->nzs * o->nx, ng = (size_t)o->nz * o->nx;
    auto cp = [](std::vector<double> &v, const float *src, size_t n) { v.assign(src, src + n); };
    cp(o->Kf, K, nf); cp(o->rhof, rhof, nf);
    cp(o->C11, C11, ns); cp(o->C13, C13, ns); cp(o->C33, C33, ns); cp(o->C55, C55, ns); cp(o->rhos, rhos, ns);
    cp(o->J, J, ng); cp(o->xix, xix, ng); cp(o->xiz, xiz, ng); cp(o->zex, zex, ng); cp(o->zez, zez, ng);
    // PML (src/elements.cu:117-124, 824-874): p_beta = 2, p_alpha = 1, alpha0 = 2*pi*fpeak.
    const double damp0 = -(d->npow + 1) * (double)d->vp_max * std::log((double)d->rc) / (2.0 * (double)d->thickness);
    const double alpha0 = 2.0 * M_PI * (double)d->fpeak;
    o->al_xi.resize(o->nx); o->be_xi.resize(o->nx); o->dp_xi.resize(o->nx);
    o->al_zs.resize(o->nzs); o->be_zs.resize(o->nzs); o->dp_zs.resize(o->nzs);
    pml_profile(o->nx, o->npd, d->npow, true, true, damp0, d->beta0, alpha0, o->al_xi.data(), o->be_xi.data(), o->dp_xi.data());
    // fluid zeta: top PML on the fluid-solid CPML grid (src/elements.cu:870-873)
    o->al_zf.assign(o->nzf, 0.0); o->be_zf.assign(o->nzf, 1.0); o->dp_zf.assign(o->nzf, 0.0);
    if (o->fcpml())
        pml_profile(o->nzf, o->npd, d->npow, true, false, damp0, d->beta0, alpha0, o->al_zf.data(), o->be_zf.data(),
                    o->dp_zf.data());
    // zeta: bottom PML; the solid-only CPML grid also has a top PML (src/elements.cu:863-869)
    pml_profile(o->nzs, o->npd, d->npow, o->solid_only && o->bdy == 1, true, damp0, d->beta0, alpha0, o->al_zs.data(),
                o->be_zs.data(), o->dp_zs.data());
    // Interface normal (src/elements.cu:892-905), or the free-surface ALM.
    o->nifx.assign(o->nx, 0.0); o->nifz.assign(o->nx, 0.0);
    if (o->solid_only && o->bdy == 2) build_alm(*o);
    // The reference builds the normal in float and keeps it as float2, so
    // n1^2 + n2^2 differs from 1 by a rounding fixed per column.  Every
    // interface projection applies that factor, so the restatement takes the
    // same float normal.
    for (int ix = 0; ix < o->nx && !o->solid_only; ++ix) {
        const size_t g = (size_t)(o->nzf - 1) * o->nx + ix;
        const float zx = (float)o->zex[g], zz = (float)o->zez[g];
        const float s = std::sqrt(zx * zx + zz * zz);
        o->nifx[ix] = zx / s;
        o->nifz[ix] = zz / s;
    }
    for (int i = 0; i < d->nsrc; ++i) { o->srcx.push_back(d->src_x[i]); o->srcz.push_back(d->src_z[i]); }
    for (int f = 0; f < NFIELD; ++f) {
        o->U[f].assign(o->fsize(f), 0.0);
        o->Bacc[f].assign(o->fsize(f), 0.0);
        o->T[f].assign(o->fsize(f), 0.0);
        o->K[f].assign(o->fsize(f), 0.0);
    }
    for (auto *g : {&o->Gfx, &o->Gfz, &o->Gsx, &o->Gsz, &o->Gfsx, &o->Gfsz}) g->assign(3 * (size_t)o->nx, 0.0);
    for (int f = 0; f < NFIELD; ++f) { o->Ain[f].assign(o->fsize(f), 0.0); o->Aout[f].assign(o->fsize(f), 0.0); }
    for (int g = 0; g < 7; ++g) o->G[g].assign((size_t)(g < 2 ? o->nzf : o->nzs) * o->nx, 0.0);
    build_runs(*o);
    return o;
}

void orc_destroy(void *h) { delete static_cast<Oracle *>(h); }

void orc_set_threads(int n) { omp_set_num_threads(n); }
int orc_get_threads(void) { return omp_get_max_threads(); }

// Zero the wavefield and select the shot (the reference does not reset; see header).
void orc_reset(void *h, int isrc) {
    Oracle *o = static_cast<Oracle *>(h);
    for (int f = 0; f < NFIELD; ++f) std::fill(o->U[f].begin(), o->U[f].end(), 0.0);
    o->isrc = isrc; o->it = 0; o->nrec = 0;
    o->rec_vx.clear(); o->rec_vz.clear(); o->rec_p.clear();
}

void orc_set_step(void *h, int it) { static_cast<Oracle *>(h)->it = it; }
int orc_get_step(void *h) { return static_cast<Oracle *>(h)->it; }
int orc_gap(void *h) { return static_cast<Oracle *>(h)->gap; }

void orc_run(void *h, int nsteps) {
    Oracle *o = static_cast<Oracle *>(h);
    for (int i = 0; i < nsteps; ++i) step(*o);
}

int orc_num_records(void *h) { return static_cast<Oracle *>(h)->nrec; }
void orc_get_records(void *h, double *vx, double *vz, double *p) {
    Oracle *o = static_cast<Oracle *>(h);
    std::memcpy(vx, o->rec_vx.data(), o->rec_vx.size() * sizeof(double));
    std::memcpy(vz, o->rec_vz.data(), o->rec_vz.size() * sizeof(double));
    std::memcpy(p, o->rec_p.data(), o->rec_p.size() * sizeof(double));
}

// Full-layout field access ([nzf][nx] or [nzs][nx]; ADE arrays full size).
void orc_get_field(void *h, int f, double *out) {
    Oracle *o = static_cast<Oracle *>(h);
    std::memcpy(out, o->U[f].data(), o->U[f].size() * sizeof(double));
}
void orc_set_field(void *h, int f, const double *in) {
    Oracle *o = static_cast<Oracle *>(h);
    std::memcpy(o->U[f].data(), in, o->U[f].size() * sizeof(double));
}

// k = BC(L(U) + src) for direction dir (0 = F, 1 = B); read with orc_get_k.
void orc_rhs(void *h, int dir, double amp) {
    Oracle *o = static_cast<Oracle *>(h);
    double *S[NFIELD], *Kd[NFIELD];
    for (int f = 0; f < NFIELD; ++f) { S[f] = o->U[f].data(); Kd[f] = o->K[f].data(); }
    if (o->solid_only) rhs_free(*o, S, Kd, dir, amp);
    else rhs(*o, S, Kd, dir, amp);
}
void orc_get_k(void *h, int f, double *out) {
    Oracle *o = static_cast<Oracle *>(h);
    std::memcpy(out, o->K[f].data(), o->K[f].size() * sizeof(double));
}
// Ghost fluxes of the last rhs call: which 0 = Gf, 1 = Gs, 2 = Gfs; out[2][3][nx].
void orc_get_ghosts(void *h, int which, double *out) {
    Oracle *o = static_cast<Oracle *>(h);
    const std::vector<double> *g[3][2] = {{&o->Gfx, &o->Gfz}, {&o->Gsx, &o->Gsz}, {&o->Gfsx, &o->Gfsz}};
    std::memcpy(out, g[which][0]->data(), 3 * o->nx * sizeof(double));
    std::memcpy(out + 3 * o->nx, g[which][1]->data(), 3 * o->nx * sizeof(double));
}

void orc_filter(void *h) {
    Oracle *o = static_cast<Oracle *>(h);
    double *Ud[NFIELD];
    for (int f = 0; f < NFIELD; ++f) Ud[f] = o->U[f].data();
    filter(*o, Ud);
}

// PML profiles: axis 0 = xi [nx], 1 = solid zeta [nzs], 2 = fluid zeta [nzf];
// out = al, be, dp.
void orc_pml(void *h, int axis, double *al, double *be, double *dp) {
    Oracle *o = static_cast<Oracle *>(h);
    const auto &A = axis == 0 ? o->al_xi : axis == 1 ? o->al_zs : o->al_zf;
    const auto &Bv = axis == 0 ? o->be_xi : axis == 1 ? o->be_zs : o->be_zf;
    const auto &D = axis == 0 ? o->dp_xi : axis == 1 ? o->dp_zs : o->dp_zf;
    std::memcpy(al, A.data(), A.size() * sizeof(double));
    std::memcpy(be, Bv.data(), Bv.size() * sizeof(double));
    std::memcpy(dp, D.data(), D.size() * sizeof(double));
}
void orc_normals(void *h, double *nxo, double *nzo) {
    Oracle *o = static_cast<Oracle *>(h);
    std::memcpy(nxo, o->nifx.data(), o->nx * sizeof(double));
    std::memcpy(nzo, o->nifz.data(), o->nx * sizeof(double));
}
double orc_stage_amp(void *h, int it, int stage) { return stage_amp(*static_cast<Oracle *>(h), it, stage); }
// Free-surface ALM coefficients of a solid-only grid: out[6][nx] (al1_xx,
// al1_zz, al1_xz, al2_xx, al2_zz, al2_xz); zeros for a fluid-solid grid.
void orc_alm(void *h, double *out) {
    Oracle *o = static_cast<Oracle *>(h);
    for (int k = 0; k < 6; ++k)
        for (int ix = 0; ix < o->nx; ++ix) out[(size_t)k * o->nx + ix] = o->alm[k].empty() ? 0.0 : o->alm[k][ix];
}


// --- adjoint probes ------------------------------------------------------
void orc_set_adj_in(void *h, int f, const double *in) {
    Oracle *o = static_cast<Oracle *>(h);
    std::memcpy(o->Ain[f].data(), in, o->Ain[f].size() * sizeof(double));
}
void orc_get_adj_out(void *h, int f, double *out) {
    Oracle *o = static_cast<Oracle *>(h);
    std::memcpy(out, o->Aout[f].data(), o->Aout[f].size() * sizeof(double));
}
void orc_get_grad(void *h, int which, double *out) {
    Oracle *o = static_cast<Oracle *>(h);
    std::memcpy(out, o->G[which].data(), o->G[which].size() * sizeof(double));
}
void orc_zero_grad(void *h) {
    Oracle *o = static_cast<Oracle *>(h);
    for (auto &g : o->G) std::fill(g.begin(), g.end(), 0.0);
}
// op: 0 = L_dir^T (forward stage state = U), 1 = B_dir^T, 2 = Filter^T,
// 3 = the transpose of one whole step it = dir (forward state = U, source off).
void orc_adjoint_op(void *h, int op, int dir) {
    Oracle *o = static_cast<Oracle *>(h);
    for (int f = 0; f < NFIELD; ++f) o->Aout[f] = o->Ain[f];
    double *A[NFIELD];
    for (int f = 0; f < NFIELD; ++f) A[f] = o->Aout[f].data();
    if (op == 0) {
        std::vector<double> Sb[NFIELD];
        double *SB[NFIELD];
        const double *S[NFIELD], *Q[NFIELD];
        for (int f = 0; f < NFIELD; ++f) { Sb[f].assign(o->Ain[f].size(), 0.0); SB[f] = Sb[f].data(); S[f] = o->U[f].data(); Q[f] = o->Ain[f].data(); }
        rhs_adj(*o, S, Q, SB, dir);
        for (int f = 0; f < NFIELD; ++f) o->Aout[f] = Sb[f];
    } else if (op == 1) {
        bc_adj(*o, A, dir);
    } else if (op == 2) {
        filter_adj(*o, A);
    } else {
        Field20 Un, T2, T3, T4;
        for (int f = 0; f < NFIELD; ++f) { Un[f] = o->U[f]; T2[f].assign(o->U[f].size(), 0.0); T3[f] = T2[f]; T4[f] = T2[f]; }
        const int save_it = o->it;
        std::vector<int> sx = o->srcx;
        o->srcx.clear();
        o->it = dir;
        step_keep(*o, T2, T3, T4);
        Field20 lam;
        for (int f = 0; f < NFIELD; ++f) lam[f] = o->Ain[f];
        step_adj(*o, dir, Un, T2, T3, T4, lam);
        for (int f = 0; f < NFIELD; ++f) { o->Aout[f] = lam[f]; o->U[f] = Un[f]; }
        o->srcx = sx;
        o->it = save_it;
    }
}
// Forward in place of the linear step map (source off) for the dot-product test.
void orc_step_nosrc(void *h, int it) {
    Oracle *o = static_cast<Oracle *>(h);
    std::vector<int> sx = o->srcx;
    o->srcx.clear();
    o->it = it;
    step(*o);
    o->srcx = sx;
}

// Gradient of J = sum_r <res_r, rec_r> (res = dJ/drec, records [nrec][nx] for
// solid vx, vz and fluid p) with respect to K, rho_f, C11, C13, C33, C55,
// rho_s for shot isrc and nsteps steps from a zero wavefield.  Checkpoints
// every `every` steps, recomputing each segment.
void orc_gradient(void *h, int isrc, int nsteps, int every, const double *res_vx, const double *res_vz,
                  const double *res_p) {
    Oracle *o = static_cast<Oracle *>(h);
    orc_reset(h, isrc);
    orc_zero_grad(h);
    if (every < 1) every = 1;
    std::vector<Field20 *> ckpt;
    for (int it = 0; it < nsteps; ++it) {
        if (it % every == 0) { auto *c = new Field20; for (int f = 0; f < NFIELD; ++f) (*c)[f] = o->U[f]; ckpt.push_back(c); }
        step(*o);
    }
    Field20 lam;
    for (int f = 0; f < NFIELD; ++f) lam[f].assign(o->U[f].size(), 0.0);
    const int nx = o->nx, gap = o->gap;
    const size_t rf = (size_t)(o->nzf - 1) * nx;
    for (int seg = (int)ckpt.size() - 1; seg >= 0; --seg) {
        const int it0 = seg * every, it1 = std::min(nsteps, it0 + every);
        std::vector<Field20 *> un, t2, t3, t4;
        for (int f = 0; f < NFIELD; ++f) o->U[f] = (*ckpt[seg])[f];
        o->it = it0;
        for (int it = it0; it < it1; ++it) {
            auto *a = new Field20, *b = new Field20, *c = new Field20, *d = new Field20;
            for (int f = 0; f < NFIELD; ++f) { (*a)[f] = o->U[f]; (*b)[f].assign(o->U[f].size(), 0.0); (*c)[f] = (*b)[f]; (*d)[f] = (*b)[f]; }
            step_keep(*o, *b, *c, *d);
            un.push_back(a); t2.push_back(b); t3.push_back(c); t4.push_back(d);
        }
        for (int it = it1 - 1; it >= it0; --it) {
            if (it % gap == 0) {
                const int r = it / gap;
                for (int ix = 0; ix < nx; ++ix) {
                    lam[SVX][ix] += res_vx[(size_t)r * nx + ix];
                    lam[SVZ][ix] += res_vz[(size_t)r * nx + ix];
                    lam[FP][rf + ix] += res_p[(size_t)r * nx + ix];
                }
            }
            const int k = it - it0;
            step_adj(*o, it, *un[k], *t2[k], *t3[k], *t4[k], lam);
            delete un[k]; delete t2[k]; delete t3[k]; delete t4[k];
        }
        delete ckpt[seg];
    }
    for (int f = 0; f < NFIELD; ++f) o->Aout[f] = lam[f];   // adjoint of the initial state
}

// 1-D operator probes for pinning: kind 0 = DRP, 1 = M42, 2 = M22,
// 3 = distance-dependent zeta operator (dist, solid_side); f sampled at i.
double orc_op1d(int kind, int dir, const double *f, int i, int dist, int solid_side) {
    auto fn = [&](int m) { return f[i + m]; };
    switch (kind) {
        case 0: return drp(dir, fn);
        case 1: return m42(dir, fn);
        case 2: return m22(dir, fn);
        default: return mac_dze(dir, dist, solid_side != 0, fn);
    }
}

}  // extern "C"
