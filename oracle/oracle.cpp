// FP64 host oracle for the MarineSeis-FDM-2D fluid-solid time stepper.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness checker for the HIP
// solver and the CPU baseline that bench.py times beside it.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
// product path (libmsfdm.so, the CLI, the Python binding) never links or calls
// it.
//
// It restates, in double precision and with race-free (sequential) semantics,
// the reference's PANDF + FANDS stepper `PmlFs::FluidSolid::update_RK4`:
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
// coefficients at the seafloor and at the pressure-release surface.

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

// Interface-row filter taps (src/elements.cu:30-33).
constexpr double W7[7] = {1.1e-13, 1.678e-06, 0.03419, 0.93163, 0.03419, 1.678e-06, 1.1e-13};

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
    NFIELD
};
inline bool is_fluid(int f) { return f <= FXEZ; }

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
    // ghost fluxes [3][nx] x/z components (src/pml_fs_kernels.cu:385-426)
    std::vector<double> Gfx, Gfz, Gsx, Gsz, Gfsx, Gfsz;
    // records
    std::vector<double> rec_vx, rec_vz, rec_p;
    int nrec = 0;

    size_t fsize(int f) const { return (size_t)(is_fluid(f) ? nzf : nzs) * nx; }
    // metric index of solid row iz (src/tools_func.cu:22, 57)
    size_t gs(int iz, int ix) const { return (size_t)(iz + nzf - 1) * nx + ix; }
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
    const int distx = std::min(ix - 3, o.nx - 4 - ix);
    return distx >= 0 && distx < o.npd;
}

// update_velocity_fluid<DIR> (include/pml_fs_kernels.cuh:333-450).
void velocity_fluid(Oracle &o, const double *const *S, double *const *Kd, int dir) {
    const int nx = o.nx, nzf = o.nzf;
#pragma omp parallel for schedule(static)
    for (int iz = 3; iz < nzf; ++iz) {
        for (int ix = 3; ix <= nx - 4; ++ix) {
            const size_t idx = (size_t)iz * nx + ix, g = o.gf(iz, ix);
            const double J = o.J[g], xx = o.xix[g], xz = o.xiz[g], zx = o.zex[g], zz = o.zez[g];
            const int dist_flso = nzf - 1 - iz, dist_fs = iz - 3;
            const double gxixi = xx * xx + xz * xz;
            double rvx, rvz, Dx, Dz;
            if (dist_flso < 3 || dist_fs < 3) {
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
            } else {
                const double *p = S[FP];
                const double dpx = drp(dir, [&](int m) { return p[idx + m]; });
                const double dpz = drp(dir, [&](int m) { return p[idx + (std::ptrdiff_t)m * nx]; });
                rvx = -(xx * dpx + zx * dpz);
                rvz = -(xz * dpx + zz * dpz);
                const double gxize = xx * zx + xz * zz;
                Dx = -xx * (gxixi * dpx + gxize * dpz) / gxixi;
                Dz = -xz * (gxixi * dpx + gxize * dpz) / gxixi;
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
        }
    }
}

// update_stress_fluid<DIR> (include/pml_fs_kernels.cuh:452-507).
void stress_fluid(Oracle &o, const double *const *S, double *const *Kd, int dir) {
    const int nx = o.nx, nzf = o.nzf;
#pragma omp parallel for schedule(static)
    for (int iz = 3; iz < nzf; ++iz) {
        for (int ix = 3; ix <= nx - 4; ++ix) {
            const size_t idx = (size_t)iz * nx + ix, g = o.gf(iz, ix);
            if (iz == 3) { Kd[FP][idx] = 0.0; continue; }   // pressure-release row
            const double *vx = S[FVX], *vz = S[FVZ];
            const double dvxx = drp(dir, [&](int m) { return vx[idx + m]; });
            const double dvzx = drp(dir, [&](int m) { return vz[idx + m]; });
            const int dzn = std::min(iz - 3, nzf - 1 - iz);
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
            double ksxz = c55 * (xz * dvxx + zz * dvxz + xx * dvzx + zx * dvzz);
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

// linear_operate + fl_so_apply (src/pml_fs_kernels.cu:533-660): k = BC(L(S) + src).
void rhs(Oracle &o, double *const *S, double *const *Kd, int dir, double amp) {
    for (int f = 0; f < NFIELD; ++f) std::fill(Kd[f], Kd[f] + o.fsize(f), 0.0);
    build_ghosts(o, S);
    velocity_fluid(o, S, Kd, dir);
    stress_fluid(o, S, Kd, dir);
    velocity_solid(o, S, Kd, dir);
    stress_solid(o, S, Kd, dir);
    add_source(o, Kd[FP], amp);
    fl_so_boundary(o, Kd, dir);
}

// fl_so_filter (src/pml_fs_kernels.cu:428-525), out of place.
void filter(Oracle &o, double *const *Ud) {
    const int nx = o.nx, le = 3, ri = nx - 4;
    const size_t rf = (size_t)(o.nzf - 1) * nx;
    const double *src[8] = {Ud[FVX] + rf, Ud[FVZ] + rf, Ud[FP] + rf, Ud[SVX], Ud[SVZ], Ud[SSX], Ud[SSZ], Ud[SSXZ]};
    std::vector<double> out(8 * (size_t)nx, 0.0);
    for (int ix = le; ix <= ri; ++ix) {
        const int klo = std::max(-3, le - ix), khi = std::min(3, ri - ix);
        double wsum = 0.0;
        for (int k = klo; k <= khi; ++k) wsum += W7[k + 3];
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

void axpy(std::vector<double> &dst, const std::vector<double> &a, double c, const std::vector<double> &k) {
    const size_t n = dst.size();
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) dst[i] = a[i] + c * k[i];
}
void self_axpy(std::vector<double> &dst, double c, const std::vector<double> &k) {
    const size_t n = dst.size();
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) dst[i] += c * k[i];
}

// One time step it: RK4_reset (src/elements.cu:345-383), update_RK4
// (src/pml_fs_kernels.cu:765-868), record (src/main.cu:245-295).
void step(Oracle &o) {
    double *S[NFIELD], *Kd[NFIELD], *Ud[NFIELD];
    for (int f = 0; f < NFIELD; ++f) {
        o.Bacc[f] = o.U[f];
        o.T[f] = o.U[f];
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
        rhs(o, S, Kd, dirs[s], stage_amp(o, o.it, s));
        for (int f = 0; f < NFIELD; ++f) {
            if (s < 3) {
                axpy(o.T[f], o.U[f], ct[s], o.K[f]);
                self_axpy(o.Bacc[f], cb[s], o.K[f]);
            } else {
                axpy(o.U[f], o.Bacc[f], dt / 6, o.K[f]);
            }
        }
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
};

// Model arrays: fluid [nzf][nx] K, rho; solid [nzs][nx] C11, C13, C33, C55,
// rho; metrics [nz][nx] J, xi_x, xi_z, ze_x, ze_z (float32, as read from the
// <f4 model files, src/elements.cu:127-176).
void *orc_create(const orc_desc *d, const float *K, const float *rhof, const float *C11,
                 const float *C13, const float *C33, const float *C55, const float *rhos,
                 const float *J, const float *xix, const float *xiz, const float *zex,
                 const float *zez) {
    Oracle *o = new Oracle();
    o->nx = d->nx; o->nzf = d->nz_fluid; o->nzs = d->nz_solid; o->npd = d->npd;
    o->nz = o->nzf + o->nzs - 1;
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
    pml_profile(o->nzs, o->npd, d->npow, false, true, damp0, d->beta0, alpha0, o->al_zs.data(), o->be_zs.data(), o->dp_zs.data());
    // Interface normal (src/elements.cu:892-905).
    o->nifx.resize(o->nx); o->nifz.resize(o->nx);
    for (int ix = 0; ix < o->nx; ++ix) {
        const size_t g = (size_t)(o->nzf - 1) * o->nx + ix;
        const double s = std::sqrt(o->zex[g] * o->zex[g] + o->zez[g] * o->zez[g]);
        o->nifx[ix] = o->zex[g] / s;
        o->nifz[ix] = o->zez[g] / s;
    }
    for (int i = 0; i < d->nsrc; ++i) { o->srcx.push_back(d->src_x[i]); o->srcz.push_back(d->src_z[i]); }
    for (int f = 0; f < NFIELD; ++f) {
        o->U[f].assign(o->fsize(f), 0.0);
        o->Bacc[f].assign(o->fsize(f), 0.0);
        o->T[f].assign(o->fsize(f), 0.0);
        o->K[f].assign(o->fsize(f), 0.0);
    }
    for (auto *g : {&o->Gfx, &o->Gfz, &o->Gsx, &o->Gsz, &o->Gfsx, &o->Gfsz}) g->assign(3 * (size_t)o->nx, 0.0);
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
    rhs(*o, S, Kd, dir, amp);
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

// PML profiles: axis 0 = xi [nx], 1 = solid zeta [nzs]; out = al, be, dp.
void orc_pml(void *h, int axis, double *al, double *be, double *dp) {
    Oracle *o = static_cast<Oracle *>(h);
    const auto &A = axis ? o->al_zs : o->al_xi, &Bv = axis ? o->be_zs : o->be_xi, &D = axis ? o->dp_zs : o->dp_xi;
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
