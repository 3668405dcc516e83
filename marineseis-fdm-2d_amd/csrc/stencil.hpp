// Device code of the fluid-solid RK4 stepper for gfx950.
//
// One RK stage = one launch:
//   stage_kernel<DIR,STAGE,TZ,MET> runs, per 64 x TZ tile,
//   fluid_body                 fluid RHS (update_velocity_fluid + update_stress_fluid,
//                              include/pml_fs_kernels.cuh:333-507) + xi-PML memory
//                              + source (src/pml_fs_kernels.cu:527-531, 618-621)
//                              + the RK4 combination (src/pml_fs_kernels.cu:663-763)
//   solid_body                 solid RHS (update_stress_solid + update_velocity_solid,
//                              include/pml_fs_kernels.cuh:509-752) + xi/zeta-PML memory
//                              + the RK4 combination
//                              + the previous stage's interface rows, rebuilt from
//                              its raw interface k with fl_so_boundary
//                              (include/pml_fs_kernels.cuh:754-827)
//   iface_final<DIR>           after the last step of a launch sequence: the interface
//                              rows of U' (fl_so_boundary of all four stages),
//                              fl_so_filter (src/pml_fs_kernels.cu:428-525, out of
//                              place) and the receiver-row record.  Inside a sequence
//                              the next step's stage-1 launch (FIN) does this work in
//                              the tiles that hold an interface row, so a step costs
//                              four launches.
//   free_stage_kernel<DIR,STAGE,TZ,MET,TOP>
//                              solid-only grids, one launch per stage: TOP_FREE, the
//                              traction-free top (PmlFs::SolidOnly,
//                              include/pml_fs_kernels.cuh:8-307: ALM surface row,
//                              odd-image ghost fluxes); TOP_PLAIN, the CPML and
//                              Dirichlet grids (include/cpml_kernels.cuh:8-233,
//                              include/dirichlet_kernels.cuh:7-107: DRP on every row,
//                              PML on four sides or none).  The source goes on both
//                              normal stresses; stage 4 records vz of the receiver row.
//
// The stage kernels read the stage state once through an LDS tile with a 3-cell
// halo on every side, build the interface / free-surface ghost fluxes
// (src/pml_fs_kernels.cu:385-426) from that tile, and never materialise k: the
// classical-RK4 accumulator and the next stage state are written directly.
// Interface-row k values (fluid row nzf-1, solid row 0) go to a small side
// buffer because the boundary condition couples both media.
//
// STAGE 1..4 are the RK stages; STAGE 0 writes the raw k (used to check the
// right-hand side against the oracle).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "msfdm.h"

namespace msfdm {
namespace dev {

// DRP/opt MacCormack weights, offsets -1..3 (include/tools_func.cuh:6-10).
constexpr float DRP_M1 = -0.30874f;
constexpr float DRP_0 = -0.63260f;
constexpr float DRP_1 = +1.23300f;
constexpr float DRP_2 = -0.33340f;
constexpr float DRP_3 = +0.04168f;
// 4-2 and 2-2 closures (include/tools_func.cuh:12-17).
constexpr float M42_0 = -7.0f / 6.0f;
constexpr float M42_1 = +8.0f / 6.0f;
constexpr float M42_2 = -1.0f / 6.0f;

constexpr int TX = 64;          // tile width = one wavefront
constexpr int HALO = 3;
constexpr int LXW = TX + 2 * HALO;  // 70 loaded columns
constexpr int LXP = 72;         // LDS row pitch (floats)
constexpr int XOFF = 4;         // LDS column of the tile's first column (16-byte aligned interior)
constexpr int NTHR = 256;       // 64 x 4 threads per stage block

struct Geo {
    int nx, nzf, nzs, P, npd, W2;
    int ztop_pml;                 // 1: the solid zeta memory also has a top strip (solid-only CPML grid)
    int fcpml;                    // 1: fluid-solid CPML grid (top PML in the water, fluid zeta memory)
    long long off[MSFDM_NFIELD];  // float offset of each field inside a state buffer
};

struct Coef {
    const float *rhof_inv, *Kf;                          // [nzf][P]
    const float *rhos_inv, *C11, *C13, *C33, *C55;       // [nzs][P]
    const float *J, *xix, *xiz, *zex, *zez;              // [nz][P] global rows
    const float *pxa, *pxb, *pxc;                        // xi PML: al+d/be, d/be, 1/be  [nx]
    const float *pza, *pzb, *pzc;                        // solid zeta PML               [nzs]
    const float *pzaf, *pzbf, *pzcf;                     // fluid zeta PML (CPML top)    [nzf]
    const float2 *nif;                                   // interface normal            [nx]
    const float *alm;                                    // free-surface ALM, solid-only grid [6][P]
    // Vertical grid columns (x = h xi everywhere): xi_x is the constant xix0 and
    // xi_z = 0, so neither array is read.
    int vert;
    float xix0;
};

// MET: 0 = general metrics (arrays read), 1 = vertical columns (compile-time
// constants), 2 = decided at run time from c.vert.
template <int MET>
__device__ __forceinline__ void xi_metrics(const Coef &c, long long gi, float &xx, float &xz) {
    if (MET == 1 || (MET == 2 && c.vert)) { xx = c.xix0; xz = 0.0f; }
    else { xx = c.xix[gi]; xz = c.xiz[gi]; }
}

struct Ctl {
    const float *amp;     // Ricker amplitude per (step, stage): [4 * cap]
    const int *base;      // first step index of the current launch batch
    float *rec;           // records [3][rec_cap][nx]: solid vx, solid vz, fluid p
    int rec_cap;
    int gap;
    int srcx, srcz;       // fluid-array source position of the current shot
    int rec_row;          // solid-only grid: recorded row of vz (-1: no records)
};

// 5-point one-sided derivative from values at ascending offsets
// (F: -1..3, B: -3..1), include/tools_func.cuh:51-71.
template <int DIR>
__device__ __forceinline__ float drp5(float v0, float v1, float v2, float v3, float v4) {
    if constexpr (DIR == 0) return DRP_M1 * v0 + DRP_0 * v1 + DRP_1 * v2 + DRP_2 * v3 + DRP_3 * v4;
    else return -(DRP_M1 * v4 + DRP_0 * v3 + DRP_1 * v2 + DRP_2 * v1 + DRP_3 * v0);
}

// Offsets read by the DRP stencil of direction DIR.
template <int DIR> __device__ __forceinline__ constexpr int lo_off() { return DIR == 0 ? -1 : -3; }

// x-derivative of an LDS row at column c (tile coordinates).
template <int DIR>
__device__ __forceinline__ float dx_lds(const float *row, int c) {
    constexpr int o = lo_off<DIR>();
    return drp5<DIR>(row[c + o], row[c + o + 1], row[c + o + 2], row[c + o + 3], row[c + o + 4]);
}
// z-derivative at tile row r, column c.
template <int DIR>
__device__ __forceinline__ float dz_lds(const float (*t)[LXP], int r, int c) {
    constexpr int o = lo_off<DIR>();
    return drp5<DIR>(t[r + o][c], t[r + o + 1][c], t[r + o + 2][c], t[r + o + 3][c], t[r + o + 4][c]);
}
// Distance-dependent zeta derivative Mac_dv_dze (include/tools_func.cuh:131-140).
template <int DIR>
__device__ __forceinline__ float mac_dz(const float (*t)[LXP], int r, int c, int dist, bool solid_side) {
    if (dist >= 3) return dz_lds<DIR>(t, r, c);
    if (dist == 2) {
        if constexpr (DIR == 0) return M42_0 * t[r][c] + M42_1 * t[r + 1][c] + M42_2 * t[r + 2][c];
        else return -(M42_0 * t[r][c] + M42_1 * t[r - 1][c] + M42_2 * t[r - 2][c]);
    }
    if (dist == 1) {
        if constexpr (DIR == 0) return -t[r][c] + t[r + 1][c];
        else return -(-t[r][c] + t[r - 1][c]);
    }
    return solid_side ? (-t[r][c] + t[r + 1][c]) : -(-t[r][c] + t[r - 1][c]);
}

// Classical RK4 (src/pml_fs_kernels.cu:663-763) on stored stage states:
//   T2 = U + dt/2 k1,  T3 = U + dt/2 k2,  T4 = U + dt k3,
//   U' = U + dt/6 (k1 + 2 k2 + 2 k3 + k4)
//      = U + ((T2 - U) + 2 (T3 - U) + (T4 - U)) / 3 + dt/6 k4.
// Writing the stage states instead of an accumulator moves 13 instead of 16
// floats per field and step; the rounding of U' is no larger than that of the
// accumulator form.  s is the stage state at the cell (U in stage 1, T4 in
// stage 4); STAGE 0 writes the raw k.
// U^n (U) and the stage output (O) are distinct buffers in stages 0-3 and carry
// __restrict__ there.  Stage 4 writes U' over U^n (O == U): each cell of U is
// read by the thread that writes it, before the write, and by no other thread
// of the launch; the pointers are plain, and the stage-4 cell update (rk4_cell)
// loads every input of a cell before its first store.
template <int STAGE> using UPtr = std::conditional_t<STAGE == 4, const float *, const float *__restrict__>;
template <int STAGE> using OPtr = std::conditional_t<STAGE == 4, float *, float *__restrict__>;

template <int STAGE>
__device__ __forceinline__ void rk(UPtr<STAGE> U, const float *__restrict__ T2, const float *__restrict__ T3,
                                   OPtr<STAGE> O, long long i, float s, float k, float ct) {
    if constexpr (STAGE == 0) {
        O[i] = k;
    } else if constexpr (STAGE == 1) {
        O[i] = s + ct * k;
    } else if constexpr (STAGE == 4) {
        const float u = U[i];
        O[i] = u + ((T2[i] - u) + 2.0f * (T3[i] - u) + (s - u)) * (1.0f / 3.0f) + ct * k;
    } else {
        O[i] = U[i] + ct * k;
    }
}

// rk with U[i] already in a register (loaded ahead of the row's arithmetic in
// stages 2 and 3; stage 4 loads it here).
template <int STAGE>
__device__ __forceinline__ constexpr bool prefetch_u() { return STAGE == 2 || STAGE == 3; }
template <int STAGE>
__device__ __forceinline__ float rk_u(UPtr<STAGE> U, const float *__restrict__ T2, const float *__restrict__ T3,
                                      OPtr<STAGE> O, long long i, float s, float k, float ct, float u) {
    float v;
    if constexpr (STAGE == 0) {
        v = k;
        O[i] = v;
        return v;
    } else if constexpr (STAGE == 1) {
        v = s + ct * k;
    } else if constexpr (STAGE == 4) {
        if constexpr (!prefetch_u<STAGE>()) u = U[i];
        v = u + ((T2[i] - u) + 2.0f * (T3[i] - u) + (s - u)) * (1.0f / 3.0f) + ct * k;
    } else {
        v = u + ct * k;
    }
    O[i] = v;
    return v;
}

// Stage 4 of NF fields at one cell (the classical combination of rk, same
// expression): all U^n, T2 and T3 values are loaded before the first store.
template <int NF>
__device__ __forceinline__ void rk4_cell(const float *U, const float *__restrict__ T2, const float *__restrict__ T3,
                                         float *O, const long long (&idx)[NF], const float (&s)[NF],
                                         const float (&k)[NF], float ct, float (&v)[NF]) {
    float u[NF], a[NF], b[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        u[f] = U[idx[f]];
        a[f] = T2[idx[f]];
        b[f] = T3[idx[f]];
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        v[f] = u[f] + ((a[f] - u[f]) + 2.0f * (b[f] - u[f]) + (s[f] - u[f])) * (1.0f / 3.0f) + ct * k[f];
        O[idx[f]] = v[f];
    }
}

// Pointwise inputs of one row of a thread, loaded one row ahead of their use so
// that their latency overlaps the previous row's arithmetic (and, for the first
// row, the tile load).  u: U^n of the row's state fields (stages 2-4).
template <int NF, int NC>
struct RowPre {
    float xx, xz, zx, zz;
    float cf[NC];
    float u[NF];
};

template <int STAGE, int MET, int NF, int NC>
__device__ __forceinline__ void prefetch_row(RowPre<NF, NC> &p, const Coef &c, const float *const (&cp)[NC],
                                             UPtr<STAGE> U, const long long *off, int f0,
                                             long long ai, long long gi) {
    xi_metrics<MET>(c, gi, p.xx, p.xz);
    p.zx = c.zex[gi];
    p.zz = c.zez[gi];
#pragma unroll
    for (int q = 0; q < NC; ++q) p.cf[q] = cp[q][ai];
#pragma unroll
    for (int f = 0; f < NF; ++f) p.u[f] = prefetch_u<STAGE>() ? U[off[f0 + f] + ai] : 0.0f;
}

// Compact PML-strip column of ix (left strip first), or -1.
__device__ __forceinline__ int side_col(int ix, int nx, int npd) {
    if (ix - 3 < npd) return ix - 3;
    if (nx - 4 - ix < npd) return npd + (ix - (nx - 3 - npd));
    return -1;
}

__constant__ float SRC_W[5][5] = {   // src/elements.cu:34-40
    {0.0030f, 0.0133f, 0.0219f, 0.0133f, 0.0030f},
    {0.0133f, 0.0596f, 0.0983f, 0.0596f, 0.0133f},
    {0.0219f, 0.0983f, 0.1621f, 0.0983f, 0.0219f},
    {0.0133f, 0.0596f, 0.0983f, 0.0596f, 0.0133f},
    {0.0030f, 0.0133f, 0.0219f, 0.0133f, 0.0030f}};
// GAUSS_W7_PML, src/elements.cu:30-33 (compile-time constants of iface_filter)
constexpr float FILT_W[7] = {1.1e-13f, 1.678e-06f, 0.03419f, 0.93163f, 0.03419f, 1.678e-06f, 1.1e-13f};

// Load NF fields of the stage state (rows z0-3 .. z0+TZ+2, columns x0-3 .. x0+66)
// into LDS (tile column x0 at LDS column XOFF); cells outside [0,nrows) x [0,P)
// read as 0.  The 64 interior columns move as 16-byte vectors (x0, the pitch P
// and every field offset are multiples of 64 floats), the 2 x 3 halo columns as
// single floats.  INB: every cell of the tile is known to lie inside the array
// (an interior tile), so the bounds tests are compiled out.
template <int NF, int TZ, bool INB = false>
__device__ __forceinline__ void load_tile(float (*sh)[TZ + 2 * HALO][LXP], const float *__restrict__ S,
                                          const long long *off, int f0, int z0, int x0, int nrows, int P) {
    constexpr int LZ = TZ + 2 * HALO;
    constexpr int NV = LZ * (TX / 4);              // float4 interior elements per field
    constexpr int IV = (NV + NTHR - 1) / NTHR;     // ... per thread
    constexpr int NH = LZ * 2 * HALO;              // halo floats per field
    constexpr int IH = (NH + NTHR - 1) / NTHR;
    // every global load of the tile is issued before the first LDS store
    float4 v[NF][IV];
    float hv[NF][IH];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const float *src = S + off[f0 + f];
#pragma unroll
        for (int i = 0; i < IV; ++i) {
            const int e = threadIdx.x + i * NTHR;
            const int r = e / (TX / 4), q = e - r * (TX / 4);
            const int gz = z0 - HALO + r;
            v[f][i] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (e < NV && (INB || (gz >= 0 && gz < nrows)))
                v[f][i] = *reinterpret_cast<const float4 *>(src + (long long)gz * P + x0 + 4 * q);
        }
#pragma unroll
        for (int i = 0; i < IH; ++i) {
            const int e = threadIdx.x + i * NTHR;
            const int r = e / (2 * HALO), h = e - r * (2 * HALO);
            const int gz = z0 - HALO + r;
            const int gx = x0 + (h < HALO ? h - HALO : TX + (h - HALO));
            hv[f][i] = 0.0f;
            if (e < NH && (INB || (gz >= 0 && gz < nrows && gx >= 0 && gx < P))) hv[f][i] = src[(long long)gz * P + gx];
        }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
#pragma unroll
        for (int i = 0; i < IV; ++i) {
            const int e = threadIdx.x + i * NTHR;
            const int r = e / (TX / 4), q = e - r * (TX / 4);
            if (e < NV) *reinterpret_cast<float4 *>(&sh[f][r][XOFF + 4 * q]) = v[f][i];
        }
#pragma unroll
        for (int i = 0; i < IH; ++i) {
            const int e = threadIdx.x + i * NTHR;
            const int r = e / (2 * HALO), h = e - r * (2 * HALO);
            const int dx = h < HALO ? h - HALO : TX + (h - HALO);
            if (e < NH) sh[f][r][XOFF + dx] = hv[f][i];
        }
    }
}

// fl_so_boundary<DIR> on the eight interface k values (include/pml_fs_kernels.cuh:754-827).
// The sxz reconstruction keeps the reference's expression (no nx*nz*snn term).
// The interface arithmetic (interface_bc, iface_pre, iface_filter) is compiled
// without multiply-add contraction: the kernels that inline it then round it
// alike, so iface_final and a FIN stage-1 tile give the same bits.
template <int DIR>
__device__ __forceinline__ void interface_bc(float k[8], float2 n) {
#pragma clang fp contract(off)
    const float n1 = n.x, n2 = n.y;
    const float vxf = k[0], vzf = k[1];
    float pf = k[2];
    const float vxs = k[3], vzs = k[4], sxx = k[5], szz = k[6], sxz = k[7];
    float vnf = n1 * vxf + n2 * vzf;
    const float vtf = n2 * vxf - n1 * vzf;
    float vns = n1 * vxs + n2 * vzs;
    const float vts = n2 * vxs - n1 * vzs;
    float snn = n1 * n1 * sxx + 2.0f * n1 * n2 * sxz + n2 * n2 * szz;
    float snt = n1 * n2 * (sxx - szz) + (n2 * n2 - n1 * n1) * sxz;
    const float stt = n2 * n2 * sxx - 2.0f * n1 * n2 * sxz + n1 * n1 * szz;
    if constexpr (DIR == 0) { vnf = vns; pf = -snn; snt = 0.0f; }
    else { vns = vnf; snn = -pf; snt = 0.0f; }
    k[0] = n1 * vnf + n2 * vtf;
    k[1] = n2 * vnf - n1 * vtf;
    k[2] = pf;
    k[3] = n1 * vns + n2 * vts;
    k[4] = n2 * vns - n1 * vts;
    k[5] = n1 * n1 * snn + 2.0f * n1 * n2 * snt + n2 * n2 * stt;
    k[6] = n2 * n2 * snn - 2.0f * n1 * n2 * snt + n1 * n1 * stt;
    k[7] = -n1 * n2 * stt + (n2 * n2 - n1 * n1) * snt;
}

__device__ __forceinline__ long long iface_index(const Geo &g, int f, int ix) {
    // f 0..2: fluid vx, vz, p at row nzf-1; 3..7: solid vx, vz, sx, sz, sxz at row 0
    return f < 3 ? g.off[MSFDM_FVX + f] + (long long)(g.nzf - 1) * g.P + ix : g.off[MSFDM_SVX + f - 3] + ix;
}

// Interface side buffer kif ([KIF_ROWS][P]), sets of eight rows.  Row f of a
// set: 0..2 = fluid vx, vz, p on fluid row nzf-1; 3..7 = solid vx, vz, sxx, szz,
// sxz on solid row 0.  Sets: the raw interface-row k of RK stages 2, 3, 4 at
// rows 8, 16, 24; the stage-1 k and U^n (saved by stage 1) twice, by step parity
// par = the direction of the step's stages 1 and 3: k1 at rows 0 (par 0) / 40
// (par 1), U^n at KIF_U (par 0) / KIF_U + 16 (par 1).  A stage-1 launch that
// also finishes the previous step's interface rows reads that step's sets while
// it writes its own (raw-k mode, stage 0: rows 0..7).
constexpr int KIF_U = 32, KIF_ROWS = 56;
__device__ __forceinline__ constexpr int kif_set(int stage, int par) {
    return stage == 0 ? 0 : stage == 1 ? (par ? 40 : 0) : 8 * (stage - 1);
}
__device__ __forceinline__ constexpr int kif_u(int par) { return par ? KIF_U + 16 : KIF_U; }
// Step parity of RK stage STAGE (1..4) run in direction DIR (stages go d1 d2 d1 d2).
template <int DIR, int STAGE>
__device__ __forceinline__ constexpr int step_par() { return (STAGE & 1) ? DIR : 1 - DIR; }

// Stage-4 interface rows + fl_so_filter (out of place) + receiver records.
// The interface rows of T2, T3, T4 and the stage-4 k are rebuilt from the four
// raw k sets and U^n (the interface condition of each stage's direction: the
// stages run d1 d2 d1 d2 with d2 = DIR), with the same expressions the stage
// kernels use: T2 = U + dt/2 B(k1), T3 = U + dt/2 B(k2), T4 = U + dt B(k3),
// U' = U + ((T2 - U) + 2 (T3 - U) + (T4 - U)) / 3 + dt/6 B(k4).
// iface_pre gives U' before the filter at one inner column, iface_filter the
// filtered, re-projected U' from the 7 columns around it, iface_record the
// receiver record of step it.  Two callers run them: iface_final (after the last
// step of a launch sequence) and a stage-1 tile of the next step (FIN), with the
// same arithmetic, so both give the same bits.
template <int DIR>
__device__ __forceinline__ void iface_pre(const Geo &g, const Coef &c, const float *__restrict__ kif, float c12,
                                          float c3, float cb, int ix, float v[8]) {
#pragma clang fp contract(off)
    constexpr int par = 1 - DIR;   // the step's stage 1 ran in direction 1 - DIR
    const float2 n = c.nif[ix];
    const long long P = g.P;
    float k1[8], k2[8], k3[8], k4[8];
#pragma unroll
    for (int f = 0; f < 8; ++f) {
        k1[f] = kif[(long long)(kif_set(1, par) + f) * P + ix];
        k2[f] = kif[(long long)(kif_set(2, par) + f) * P + ix];
        k3[f] = kif[(long long)(kif_set(3, par) + f) * P + ix];
        k4[f] = kif[(long long)(kif_set(4, par) + f) * P + ix];
    }
    interface_bc<1 - DIR>(k1, n);
    interface_bc<DIR>(k2, n);
    interface_bc<1 - DIR>(k3, n);
    interface_bc<DIR>(k4, n);
#pragma unroll
    for (int f = 0; f < 8; ++f) {
        const float u = kif[(long long)(kif_u(par) + f) * P + ix];   // U^n saved in stage 1
        const float t2 = u + c12 * k1[f], t3 = u + c12 * k2[f], t4 = u + c3 * k3[f];
        v[f] = u + ((t2 - u) + 2.0f * (t3 - u) + (t4 - u)) * (1.0f / 3.0f) + cb * k4[f];
    }
}

// pre[f * pitch + t + kk], kk = -3..3: U' before the filter around inner column
// ix (0 at columns that are not inner).
__device__ __forceinline__ void iface_filter(const float *pre, int pitch, int t, int ix, int nx, float2 n,
                                             float o[8]) {
#pragma clang fp contract(off)
    // taps outside the grid (klo..khi clips them) weigh 0; the 7 taps are
    // unrolled with constant weights, so no tap waits on a memory load
    const int klo = max(-3, 3 - ix), khi = min(3, nx - 4 - ix);
    float w[7];
    float wsum = 0.0f;
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        w[q] = (q - 3 >= klo && q - 3 <= khi) ? FILT_W[q] : 0.0f;
        wsum += w[q];
    }
    float v[8];
#pragma unroll
    for (int f = 0; f < 8; ++f) {
        float s = 0.0f;
#pragma unroll
        for (int q = 0; q < 7; ++q) s += w[q] * pre[f * pitch + t + q - 3];
        v[f] = s / wsum;
    }
    const float n1 = n.x, n2 = n.y;
    const float vnf = n1 * v[0] + n2 * v[1], vtf = n2 * v[0] - n1 * v[1];
    const float vns = n1 * v[3] + n2 * v[4], vts = n2 * v[3] - n1 * v[4];
    const float snn = n1 * n1 * v[5] + 2.0f * n1 * n2 * v[7] + n2 * n2 * v[6];
    const float stt = n2 * n2 * v[5] - 2.0f * n1 * n2 * v[7] + n1 * n1 * v[6];
    const float vn = 0.5f * (vnf + vns), sn = 0.5f * (snn - v[2]);
    o[0] = n1 * vn + n2 * vtf;
    o[1] = n2 * vn - n1 * vtf;
    o[2] = -sn;
    o[3] = n1 * vn + n2 * vts;
    o[4] = n2 * vn - n1 * vts;
    o[5] = n1 * n1 * sn + n2 * n2 * stt;
    o[6] = n2 * n2 * sn + n1 * n1 * stt;
    o[7] = n1 * n2 * (sn - stt);
}

// Receiver record of step it at column ix: solid vx, vz and fluid p of the
// interface rows after the step (src/main.cu:297-322).
__device__ __forceinline__ void iface_record(const Ctl &ctl, int it, int nx, int ix, const float o[8]) {
    if (it % ctl.gap != 0) return;
    const int slot = it / ctl.gap;
    if (slot >= ctl.rec_cap) return;
    const long long plane = (long long)ctl.rec_cap * nx;
    float *rr = ctl.rec + (long long)slot * nx + ix;
    rr[0] = o[3];
    rr[plane] = o[4];
    rr[2 * plane] = o[2];
}

// Stage 1 of a step that follows another step of the same launch sequence
// (FIN): no iface_final ran after the previous step, so the interface rows of
// U^n still lack its interface condition and filter.  Every tile that holds an
// interface row computes them in LDS for its 70 loaded columns from the previous
// step's side-buffer sets (iface_pre on the 76 columns x0-6 .. x0+69, then
// iface_filter); the tile that owns the row (wb) stores its 64 columns of fields
// f0 .. f0+NF-1 to U^n, and the solid one writes the previous step's record.
// As in rebuild_interface_row, other blocks may load the row before or after
// that store and replace it in LDS with the same bits they compute themselves.
constexpr int FW = TX + 4 * HALO;   // pre-filter columns of a tile
template <int DIR, int NF, int TZ>
__device__ __forceinline__ void finish_interface_row(float (*sh)[TZ + 2 * HALO][LXP], float (*pre)[FW], const Geo &g,
                                                     const Coef &c, const Ctl &ctl, const float *__restrict__ S,
                                                     const float *__restrict__ kif, int lrow, int f0, int x0, bool wb,
                                                     float dt, int jstep) {
    const int nx = g.nx;
    const float c12 = dt * 0.5f, c3 = dt, cb = __fdiv_rn(dt, 6.0f);   // the host's dt / 2, dt, dt / 6
    for (int t = threadIdx.x; t < FW; t += NTHR) {
        const int ix = x0 - 2 * HALO + t;
        float v[8];
        if (ix >= 3 && ix <= nx - 4) {
            iface_pre<DIR>(g, c, kif, c12, c3, cb, ix, v);
        } else {
#pragma unroll
            for (int f = 0; f < 8; ++f) v[f] = 0.0f;
        }
#pragma unroll
        for (int f = 0; f < 8; ++f) pre[f][t] = v[f];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < LXW; t += NTHR) {
        const int dx = t - HALO, ix = x0 + dx;
        if (ix < 3 || ix > nx - 4) continue;
        float o[8];
        iface_filter(&pre[0][0], FW, t + HALO, ix, nx, c.nif[ix], o);
#pragma unroll
        for (int f = 0; f < NF; ++f) sh[f][lrow][XOFF + dx] = o[f0 + f];
        if (wb && dx >= 0 && dx < TX) {
#pragma unroll
            for (int f = 0; f < NF; ++f) const_cast<float *>(S)[iface_index(g, f0 + f, ix)] = o[f0 + f];
            if (f0 == 3) iface_record(ctl, ctl.base[0] + jstep - 1, nx, ix, o);
        }
    }
}


// The interface rows of stage state S (RK stage STAGE >= 2) are not written by
// the previous stage's launch, whose k there still lacks the interface
// condition; every tile that holds an interface row rebuilds it in its LDS tile
// from that stage's raw k (fl_so_boundary of the previous stage's direction
// 1 - DIR) and U^n:  S_if = U^n_if + cprev * B(k_prev).  The tile whose own 64
// columns hold the row (wb) also stores the rebuilt values to S, so the stage
// buffers are complete for later readers (the RK4 combination of stage 4 uses
// the same values through iface_final, the adjoint reads the stored states).
// Other blocks of the same launch may load that row before or after the store;
// they overwrite it in LDS either way.  This is race-free only because every
// block rebuilds bit-identical values from the same side-buffer inputs with the
// same arithmetic (interface_bc is compiled without contraction): a block that
// loads the stored row and one that loads the stale row end with the same LDS
// contents.  The bitwise chunked-run tests guard this.
template <int DIR, int STAGE, int NF, int TZ>
__device__ __forceinline__ void rebuild_interface_row(float (*sh)[TZ + 2 * HALO][LXP], const Geo &g, const Coef &c,
                                                      const float *__restrict__ S, const float *__restrict__ kif,
                                                      int lrow, int f0, int x0, float cprev, bool wb) {
    const long long P = g.P;
    constexpr int par = step_par<DIR, STAGE>();
    const float *kp = kif + (long long)kif_set(STAGE - 1, par) * P;
    for (int t = threadIdx.x; t < LXW; t += NTHR) {
        const int dx = t - HALO, ix = x0 + dx;
        if (ix < 3 || ix > g.nx - 4) continue;
        float k[8];
#pragma unroll
        for (int f = 0; f < 8; ++f) k[f] = kp[f * P + ix];
        interface_bc<1 - DIR>(k, c.nif[ix]);
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const float v = kif[(kif_u(par) + f0 + f) * P + ix] + cprev * k[f0 + f];
            sh[f][lrow][XOFF + dx] = v;
            if (wb && dx >= 0 && dx < TX) const_cast<float *>(S)[iface_index(g, f0 + f, ix)] = v;
        }
    }
}

// First row of fluid band `band` of TZ-row tiles: the bands end at the interface
// row nzf-1, so only the top band (holding the surface rows) is partial and a
// single band holds the interface row; rows above 3 are not updated.
__device__ __forceinline__ int fluid_z0(const Geo &g, int band, int TZ) {
    const int nbf = (g.nzf - 3 + TZ - 1) / TZ;
    return g.nzf - (nbf - band) * TZ;
}

// ---------------------------------------------------------------------------
// Fluid stage: rows 3..nzf-1, columns 3..nx-4.  FT = 0: pressure-release sea
// surface on row 3 (PmlFs::FluidSolid); FT = 1: a top PML in the water on rows
// 3..npd+2 instead (Cpml::FluidSolid, include/cpml_kernels.cuh:239-415): no
// surface rows, the zeta operator closes only at the interface, and the top
// strip carries the fluid zeta memory.
// ---------------------------------------------------------------------------
// FIN (stage 1 only): also finish the previous step's interface rows
// (finish_interface_row, LDS scratch fin, cprev = dt).  INT: an interior tile
// (fluid_interior: no PML strip, surface, interface or boundary row in it), whose
// boundary branches are compiled out.
template <int DIR, int STAGE, int TZ, int MET, int FT = 0, int FIN = 0, int INT = 0>
__device__ __forceinline__ void
fluid_body(float (*sh)[TZ + 2 * HALO][LXP], const Geo &g, const Coef &c, const Ctl &ctl,
           const float *__restrict__ S, UPtr<STAGE> U, const float *__restrict__ T2,
           const float *__restrict__ T3, OPtr<STAGE> O, float *__restrict__ kif, int jstep,
           float ct, float amp_direct, float cprev, int bx, int bz, float (*fin)[FW] = nullptr) {
    // sh[0..2] = vx, vz, p of the stage state
    // a wave holds one row of the tile at a time: ty (and every row test) is
    // uniform per wave, which readfirstlane tells the compiler, so the row
    // branches are scalar branches rather than exec-masked ones
    const int tx = threadIdx.x & (TX - 1), ty = __builtin_amdgcn_readfirstlane(threadIdx.x / TX);
    const int x0 = bx * TX, z0 = fluid_z0(g, bz, TZ);
    const int P = g.P, nx = g.nx, nzf = g.nzf;
    const int ix = x0 + tx;
    const bool active = ix >= 3 && ix <= nx - 4;
    const int scol = INT ? -1 : side_col(ix, nx, g.npd);
    float amp = 0.0f, pxa = 0.f, pxb = 0.f, pxc = 0.f;
    const float *const cp[2] = {c.rhof_inv, c.Kf};
    RowPre<3, 2> nxt;
    if (INT || active) {   // an interior tile has every column active
        amp = (STAGE == 0) ? amp_direct : ctl.amp[4 * (ctl.base[0] + jstep) + (STAGE > 0 ? STAGE - 1 : 0)];
        if (scol >= 0) { pxa = c.pxa[ix]; pxb = c.pxb[ix]; pxc = c.pxc[ix]; }
        const long long a0 = (long long)min(max(z0 + ty, 3), nzf - 1) * P + ix;
        prefetch_row<STAGE, MET>(nxt, c, cp, U, g.off, MSFDM_FVX, a0, a0);
    }
    load_tile<3, TZ, INT>(sh, S, g.off, MSFDM_FVX, z0, x0, nzf, P);
    __syncthreads();
    if constexpr (STAGE >= 2 && !INT) {
        const int lrow = (nzf - 1) - (z0 - HALO);   // tile row of the interface (the tile starts inside the fluid)
        if (lrow < TZ + 2 * HALO) {
            rebuild_interface_row<DIR, STAGE, 3, TZ>(sh, g, c, S, kif, lrow, 0, x0, cprev, lrow - HALO < TZ);
            __syncthreads();
        }
    } else if constexpr (STAGE == 1 && FIN && !INT) {
        const int lrow = (nzf - 1) - (z0 - HALO);
        if (lrow < TZ + 2 * HALO) {
            finish_interface_row<DIR, 3, TZ>(sh, fin, g, c, ctl, S, kif, lrow, 0, x0, lrow - HALO < TZ, cprev, jstep);
            __syncthreads();
        }
    }
    if (!INT && !active) return;
    const int cc = tx + XOFF;
    const float (*shp)[LXP] = sh[2];
    const float (*shvx)[LXP] = sh[0];
    const float (*shvz)[LXP] = sh[1];

#pragma unroll
    for (int j = 0; j < TZ / 4; ++j) {
        const int lz = ty + 4 * j;
        const int iz = z0 + lz;
        if (!INT && iz > nzf - 1) break;   // (an interior tile ends 4 rows above the interface)
        const RowPre<3, 2> cur = nxt;
        if (j + 1 < TZ / 4) {
            const long long an = (long long)min(max(iz + 4, 3), nzf - 1) * P + ix;
            prefetch_row<STAGE, MET>(nxt, c, cp, U, g.off, MSFDM_FVX, an, an);
        }
        if (!INT && iz < 3) continue;   // above the first updated row (top band)
        const int r = lz + HALO;
        const long long gi = (long long)iz * P + ix;
        const float xx = cur.xx, xz = cur.xz, zx = cur.zx, zz = cur.zz;
        const float gxixi = xx * xx + xz * xz;
        const int dist_flso = INT ? 3 : nzf - 1 - iz, dist_fs = INT ? 3 : iz - 3;
        const bool top = !INT && FT == 1 && iz <= g.npd + 2;   // fluid zeta-PML strip (CPML grid)

        // ---- velocity: rho dv/dt = -grad p (include/pml_fs_kernels.cuh:333-450)
        float rvx, rvz, Dx, Dz, Dzx = 0.f, Dzz = 0.f;
        if (dist_flso < 3 || (FT == 0 && dist_fs < 3)) {
            const bool at_if = dist_flso < 3;
            // zeta flux -J p (ze_x, ze_z) at real fluid row rr of this column
            auto fze_row = [&](int rr, float &fx, float &fz) {
                const long long q = (long long)rr * P + ix;
                const float mJp = -c.J[q] * shp[r + (rr - iz)][cc];
                fx = mJp * c.zex[q];
                fz = mJp * c.zez[q];
            };
            float ax[5], az[5], bx[5], bz[5];
            constexpr int o = lo_off<DIR>();
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                const int m = o + q;
                const long long gq = gi + m;
                const float mJp = -c.J[gq] * shp[r][cc + m];
                float xxq, xzq;
                xi_metrics<MET>(c, gq, xxq, xzq);
                ax[q] = mJp * xxq;
                az[q] = mJp * xzq;
                const int rr = iz + m;
                if (at_if) {
                    if (rr <= nzf - 1) fze_row(rr, bx[q], bz[q]);
                    else {   // Gf ghost n = rr - (nzf-1): 2 F(interface) - F(interface - n)
                        float fx0, fz0, fxn, fzn;
                        fze_row(nzf - 1, fx0, fz0);
                        fze_row(2 * (nzf - 1) - rr, fxn, fzn);
                        bx[q] = 2.0f * fx0 - fxn;
                        bz[q] = 2.0f * fz0 - fzn;
                    }
                } else {
                    if (rr >= 3) fze_row(rr, bx[q], bz[q]);
                    else {   // Gfs ghost n = 3 - rr: -F(3 + n)
                        float fxn, fzn;
                        fze_row(6 - rr, fxn, fzn);
                        bx[q] = -fxn;
                        bz[q] = -fzn;
                    }
                }
            }
            const float sx = drp5<DIR>(ax[0], ax[1], ax[2], ax[3], ax[4]) + drp5<DIR>(bx[0], bx[1], bx[2], bx[3], bx[4]);
            const float sz = drp5<DIR>(az[0], az[1], az[2], az[3], az[4]) + drp5<DIR>(bz[0], bz[1], bz[2], bz[3], bz[4]);
            const float Jc = c.J[gi];
            rvx = sx / Jc;
            rvz = sz / Jc;
            const float proj = (xx * sx + xz * sz) / (gxixi * Jc);
            Dx = xx * proj;
            Dz = xz * proj;
            if (top) {   // include/cpml_kernels.cuh:286-291
                const float pze = (zx * sx + zz * sz) / ((zx * zx + zz * zz) * Jc);
                Dzx = zx * pze;
                Dzz = zz * pze;
            }
        } else {
            const float dpx = dx_lds<DIR>(shp[r], cc);
            const float dpz = dz_lds<DIR>(shp, r, cc);
            rvx = -(xx * dpx + zx * dpz);
            rvz = -(xz * dpx + zz * dpz);
            const float gxize = xx * zx + xz * zz;
            Dx = Dz = 0.0f;
            if (scol >= 0) {   // xi-PML source: only the side-strip cells pay for the division
                const float t = (gxixi * dpx + gxize * dpz) / gxixi;
                Dx = -xx * t;
                Dz = -xz * t;
            }
            if (top) {   // include/cpml_kernels.cuh:304-305
                const float gzeze = zx * zx + zz * zz;
                const float tz = (gxize * dpx + gzeze * dpz) / gzeze;
                Dzx = -zx * tz;
                Dzz = -zz * tz;
            }
        }
        const float rinv = cur.cf[0];
        float kvx = rvx * rinv, kvz = rvz * rinv;

        // ---- pressure: dp/dt = -K div v (include/pml_fs_kernels.cuh:452-507)
        float kp = 0.0f, kEx = 0.0f, kEz = 0.0f;
        const float Km = cur.cf[1];
        float dvxx = 0.f, dvzx = 0.f, dvxz = 0.f, dvzz = 0.f;
        if (FT == 1 || dist_fs != 0) {
            dvxx = dx_lds<DIR>(shvx[r], cc);
            dvzx = dx_lds<DIR>(shvz[r], cc);
            const int dzn = FT == 1 ? dist_flso : min(dist_fs, dist_flso);
            dvxz = mac_dz<DIR>(shvx, r, cc, dzn, false);
            dvzz = mac_dz<DIR>(shvz, r, cc, dzn, false);
            kp = -Km * (xx * dvxx + zx * dvxz + xz * dvzx + zz * dvzz);
        }
        // ---- xi-PML memory (side strips)
        if (scol >= 0) {
            const long long ai = (long long)iz * g.W2 + scol;
            const float Vx = S[g.off[MSFDM_FXVX] + ai], Vz = S[g.off[MSFDM_FXVZ] + ai];
            const float kVx = -pxa * Vx + pxb * Dx;
            const float kVz = -pxa * Vz + pxb * Dz;
            kvx += ((pxc - 1.0f) * Dx - pxc * Vx) * rinv;
            kvz += ((pxc - 1.0f) * Dz - pxc * Vz) * rinv;
            const float Ex = S[g.off[MSFDM_FXEX] + ai], Ez = S[g.off[MSFDM_FXEZ] + ai];
            if (FT == 1 || dist_fs != 0) {
                const float gxize = xx * zx + xz * zz;
                const float Dvx = dvxx + (gxize / gxixi) * dvxz;
                const float Dvz = dvzx + (gxize / gxixi) * dvzz;
                kEx = -pxa * Ex + pxb * Dvx;
                kEz = -pxa * Ez + pxb * Dvz;
                kp += Km * (-(pxc - 1.0f) * xx * Dvx - (pxc - 1.0f) * xz * Dvz + pxc * xx * Ex + pxc * xz * Ez);
            }
            rk<STAGE>(U, T2, T3, O, g.off[MSFDM_FXVX] + ai, Vx, kVx, ct);
            rk<STAGE>(U, T2, T3, O, g.off[MSFDM_FXVZ] + ai, Vz, kVz, ct);
            rk<STAGE>(U, T2, T3, O, g.off[MSFDM_FXEX] + ai, Ex, kEx, ct);
            rk<STAGE>(U, T2, T3, O, g.off[MSFDM_FXEZ] + ai, Ez, kEz, ct);
        }
        // ---- zeta-PML memory, top strip of the CPML grid (include/cpml_kernels.cuh:331-342, 399-413)
        if (top) {
            const long long ai = (long long)(iz - 3) * P + ix;
            const float pza = c.pzaf[iz], pzb = c.pzbf[iz], pzc = c.pzcf[iz];
            const float Vx = S[g.off[MSFDM_FZVX] + ai], Vz = S[g.off[MSFDM_FZVZ] + ai];
            const float Ex = S[g.off[MSFDM_FZEX] + ai], Ez = S[g.off[MSFDM_FZEZ] + ai];
            const float q1 = pzc - 1.0f;
            const float kVx = -pza * Vx + pzb * Dzx;
            const float kVz = -pza * Vz + pzb * Dzz;
            kvx += (q1 * Dzx - pzc * Vx) * rinv;
            kvz += (q1 * Dzz - pzc * Vz) * rinv;
            const float gxize = xx * zx + xz * zz, gzeze = zx * zx + zz * zz;
            const float Dvx = dvxz + (gxize / gzeze) * dvxx;
            const float Dvz = dvzz + (gxize / gzeze) * dvzx;
            const float kEx = -pza * Ex + pzb * Dvx;
            const float kEz = -pza * Ez + pzb * Dvz;
            kp += Km * (-q1 * (zx * Dvx + zz * Dvz) + pzc * (zx * Ex + zz * Ez));
            rk<STAGE>(U, T2, T3, O, g.off[MSFDM_FZVX] + ai, Vx, kVx, ct);
            rk<STAGE>(U, T2, T3, O, g.off[MSFDM_FZVZ] + ai, Vz, kVz, ct);
            rk<STAGE>(U, T2, T3, O, g.off[MSFDM_FZEX] + ai, Ex, kEx, ct);
            rk<STAGE>(U, T2, T3, O, g.off[MSFDM_FZEZ] + ai, Ez, kEz, ct);
        }
        // ---- source: 5x5 Gaussian-weighted Ricker on p_k (include/tools_func.cuh:21-37)
        // (never in an interior tile, whose rows then form one basic block)
        if constexpr (!INT) {
            const int sxo = ix - ctl.srcx, szo = iz - ctl.srcz;
            if (sxo >= -2 && sxo <= 2 && szo >= -2 && szo <= 2 && iz <= nzf - 4)
                kp += amp * SRC_W[szo + 2][sxo + 2];
        }
        if (dist_flso == 0) {
            constexpr int par = STAGE == 0 ? 0 : step_par<DIR, STAGE>();
            float *kk = kif + (long long)kif_set(STAGE, par) * P;
            kk[0 * (long long)P + ix] = kvx;
            kk[1 * (long long)P + ix] = kvz;
            kk[2 * (long long)P + ix] = kp;
            if constexpr (STAGE == 1) {   // S = U^n here
                kif[(long long)(kif_u(par) + 0) * P + ix] = shvx[r][cc];
                kif[(long long)(kif_u(par) + 1) * P + ix] = shvz[r][cc];
                kif[(long long)(kif_u(par) + 2) * P + ix] = shp[r][cc];
            }
            if constexpr (STAGE != 0) continue;
        }
        if constexpr (STAGE == 4) {
            const long long idx[3] = {g.off[MSFDM_FVX] + gi, g.off[MSFDM_FVZ] + gi, g.off[MSFDM_FP] + gi};
            const float sv[3] = {shvx[r][cc], shvz[r][cc], shp[r][cc]}, kv[3] = {kvx, kvz, kp};
            float v[3];
            rk4_cell<3>(U, T2, T3, O, idx, sv, kv, ct, v);
        } else {
            rk_u<STAGE>(U, T2, T3, O, g.off[MSFDM_FVX] + gi, shvx[r][cc], kvx, ct, cur.u[0]);
            rk_u<STAGE>(U, T2, T3, O, g.off[MSFDM_FVZ] + gi, shvz[r][cc], kvz, ct, cur.u[1]);
            rk_u<STAGE>(U, T2, T3, O, g.off[MSFDM_FP] + gi, shp[r][cc], kp, ct, cur.u[2]);
        }
    }
}

// ---------------------------------------------------------------------------
// Solid stage.  TOP_IFACE: the solid below the fluid, rows 0..nzs-4 (row 0 the
// interface), metrics at global row iz+nzf-1 (src/tools_func.cu:22, 57).  The
// solid-only grids (rows 3..nzs-4, metrics at row iz, the source on both normal
// stresses, the receiver row recorded by stage 4): TOP_FREE with the
// traction-free surface on row 3 (PmlFs::SolidOnly, include/pml_fs_kernels.cuh:8-307),
// TOP_PLAIN with DRP on every row and zeta PML at the top and bottom (CPML,
// include/cpml_kernels.cuh:8-233) or no PML (Dirichlet, npd = 0,
// include/dirichlet_kernels.cuh:7-107).  Columns 3..nx-4 in all.
// ---------------------------------------------------------------------------
// TOP_IFACE_CPML: TOP_IFACE on the fluid-solid CPML grid, whose sigma_xz rate
// reads its metrics at the solid-array row (include/cpml_kernels.cuh:447-452).
constexpr int TOP_IFACE = 0, TOP_FREE = 1, TOP_PLAIN = 2, TOP_IFACE_CPML = 3;
template <int DIR, int STAGE, int TZ, int MET, int TOP = TOP_IFACE, int FIN = 0, int INT = 0>
__device__ __forceinline__ void
solid_body(float (*sh)[TZ + 2 * HALO][LXP], const Geo &g, const Coef &c, const Ctl &ctl, const float *__restrict__ S,
           UPtr<STAGE> U, const float *__restrict__ T2, const float *__restrict__ T3,
           OPtr<STAGE> O, float *__restrict__ kif, int jstep, float ct, float amp_direct, float cprev, int bx,
           int bz, float (*fin)[FW] = nullptr) {
    // sh[0..4] = vx, vz, sx, sz, sxz of the stage state
    // a wave holds one row of the tile at a time: ty (and every row test) is
    // uniform per wave, which readfirstlane tells the compiler, so the row
    // branches are scalar branches rather than exec-masked ones
    const int tx = threadIdx.x & (TX - 1), ty = __builtin_amdgcn_readfirstlane(threadIdx.x / TX);
    constexpr bool SOLO = TOP == TOP_FREE || TOP == TOP_PLAIN;   // a solid-only grid
    const int x0 = bx * TX, z0 = (SOLO ? 3 : 0) + bz * TZ;
    const int P = g.P, nx = g.nx, nzs = g.nzs, npd = g.npd;
    const int ix = x0 + tx;
    const bool active = ix >= 3 && ix <= nx - 4;
    const int scol = INT ? -1 : side_col(ix, nx, npd);
    const long long moff = SOLO ? 0 : (long long)(g.nzf - 1) * P;   // solid row -> global metric row
    constexpr int ztop = SOLO ? 3 : 0;                               // first updated row
    float pxa = 0.f, pxb = 0.f, pxc = 0.f;
    float amp = 0.0f;
    const float *const cp[5] = {c.C11, c.C13, c.C33, c.C55, c.rhos_inv};
    RowPre<5, 5> nxt;
    if (INT || active) {   // an interior tile has every column active
        if (scol >= 0) { pxa = c.pxa[ix]; pxb = c.pxb[ix]; pxc = c.pxc[ix]; }
        if constexpr (SOLO)
            amp = (STAGE == 0) ? amp_direct : ctl.amp[4 * (ctl.base[0] + jstep) + (STAGE > 0 ? STAGE - 1 : 0)];
        const long long a0 = (long long)min(z0 + ty, nzs - 4) * P + ix;
        prefetch_row<STAGE, MET>(nxt, c, cp, U, g.off, MSFDM_SVX, a0, moff + a0);
    }
    load_tile<5, TZ, INT>(sh, S, g.off, MSFDM_SVX, z0, x0, nzs, P);
    __syncthreads();
    if constexpr (STAGE >= 2 && !SOLO && !INT) {
        if (z0 == 0) {   // the band holding solid row 0 (tile row HALO)
            rebuild_interface_row<DIR, STAGE, 5, TZ>(sh, g, c, S, kif, HALO, 3, x0, cprev, true);
            __syncthreads();
        }
    } else if constexpr (STAGE == 1 && FIN && !SOLO && !INT) {
        if (z0 == 0) {
            finish_interface_row<DIR, 5, TZ>(sh, fin, g, c, ctl, S, kif, HALO, 3, x0, true, cprev, jstep);
            __syncthreads();
        }
    }
    if (!INT && !active) return;
    const int cc = tx + XOFF;
    const float (*svx)[LXP] = sh[0];
    const float (*svz)[LXP] = sh[1];
    const float (*ssx)[LXP] = sh[2];
    const float (*ssz)[LXP] = sh[3];
    const float (*ssxz)[LXP] = sh[4];
    const int zb0 = nzs - 3 - npd;   // first bottom-PML row

#pragma unroll
    for (int j = 0; j < TZ / 4; ++j) {
        const int lz = ty + 4 * j;
        const int iz = z0 + lz;
        if (!INT && iz > nzs - 4) break;   // (an interior tile ends above the bottom strip)
        // Boundary tiles load a row's inputs at the row rather than one row
        // ahead: the boundary path sets the kernel's register budget, and the
        // prefetched row would stay live through all of its branches.
        RowPre<5, 5> cur = nxt;
        if constexpr (!INT) {
            const long long ac = (long long)min(iz, nzs - 4) * P + ix;
            if (j > 0) prefetch_row<STAGE, MET>(cur, c, cp, U, g.off, MSFDM_SVX, ac, moff + ac);
        } else if (j + 1 < TZ / 4) {
            const long long an = (long long)min(iz + 4, nzs - 4) * P + ix;
            prefetch_row<STAGE, MET>(nxt, c, cp, U, g.off, MSFDM_SVX, an, moff + an);
        }
        const int r = lz + HALO;
        const long long si = (long long)iz * P + ix;
        const long long gi = moff + si;
        const float xx = cur.xx, xz = cur.xz, zx = cur.zx, zz = cur.zz;
        const float gxixi = xx * xx + xz * xz, gzeze = zx * zx + zz * zz, gxize = xx * zx + xz * zz;
        const bool side = scol >= 0, bottom = !INT && iz >= zb0;
        const bool top = !INT && TOP == TOP_PLAIN && g.ztop_pml && iz <= npd + 2;
        // rows below the interface / free surface (the plain top never closes the stencils)
        const int dist = (INT || TOP == TOP_PLAIN) ? 3 : iz - ztop;
        const bool surf = !INT && TOP == TOP_FREE && dist == 0;

        // ---- stress (include/pml_fs_kernels.cuh:509-628; solid-only: 10-187)
        const float dvxx = dx_lds<DIR>(svx[r], cc);
        const float dvzx = dx_lds<DIR>(svz[r], cc);
        float dvxz = 0.0f, dvzz = 0.0f;
        const float c11 = cur.cf[0], c13 = cur.cf[1], c33 = cur.cf[2], c55 = cur.cf[3];
        float ksx, ksz, ksxz;
        float a1xx = 0.f, a1zz = 0.f, a1xz = 0.f, a2xx = 0.f, a2zz = 0.f, a2xz = 0.f;
        if (surf) {
            // traction-free row: zeta derivatives eliminated through the ALM
            // coefficients (src/elements.cu:907-977)
            a1xx = c.alm[0 * (long long)P + ix]; a1zz = c.alm[1 * (long long)P + ix]; a1xz = c.alm[2 * (long long)P + ix];
            a2xx = c.alm[3 * (long long)P + ix]; a2zz = c.alm[4 * (long long)P + ix]; a2xz = c.alm[5 * (long long)P + ix];
            ksx = (c11 * a1xx + c13 * a1zz) * dvxx + (c11 * a2xx + c13 * a2zz) * dvzx;
            ksz = (c13 * a1xx + c33 * a1zz) * dvxx + (c13 * a2xx + c33 * a2zz) * dvzx;
            ksxz = c55 * (a1xz * dvxx + a2xz * dvzx);
        } else {
            dvxz = mac_dz<DIR>(svx, r, cc, dist, true);
            dvzz = mac_dz<DIR>(svz, r, cc, dist, true);
            const float exx = xx * dvxx + zx * dvxz, ezz = xz * dvzx + zz * dvzz;
            ksx = c11 * exx + c13 * ezz;
            ksz = c13 * exx + c33 * ezz;
            if constexpr (TOP == TOP_IFACE_CPML) {   // metrics of global row iz (not iz + nzf - 1)
                float xxq, xzq;
                xi_metrics<MET>(c, si, xxq, xzq);
                ksxz = c55 * (xzq * dvxx + c.zez[si] * dvxz + xxq * dvzx + c.zex[si] * dvzz);
            } else {
                ksxz = c55 * (xz * dvxx + zz * dvxz + xx * dvzx + zx * dvzz);
            }
        }

        // ---- velocity (include/pml_fs_kernels.cuh:631-752; solid-only: 189-305)
        float rvx, rvz, Dxx, Dxz, Dzx, Dzz;
        if (dist < 3) {
            // zeta traction flux at real solid row rr of this column
            auto fze_row = [&](int rr, float &fx, float &fz) {
                const long long q = moff + (long long)rr * P + ix;
                const int tr = r + (rr - iz);
                const float Jq = c.J[q], zxq = c.zex[q], zzq = c.zez[q];
                fx = Jq * (zxq * ssx[tr][cc] + zzq * ssxz[tr][cc]);
                fz = Jq * (zxq * ssxz[tr][cc] + zzq * ssz[tr][cc]);
            };
            float ax[5], az[5], bx[5], bz[5];
            constexpr int o = lo_off<DIR>();
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                const int m = o + q;
                const long long gq = gi + m;
                const float Jq = c.J[gq];
                float xxq, xzq;
                xi_metrics<MET>(c, gq, xxq, xzq);
                ax[q] = Jq * (xxq * ssx[r][cc + m] + xzq * ssxz[r][cc + m]);
                az[q] = Jq * (xxq * ssxz[r][cc + m] + xzq * ssz[r][cc + m]);
                const int rr = iz + m;
                if constexpr (TOP == TOP_FREE) {
                    if (rr >= 3) fze_row(rr, bx[q], bz[q]);
                    else {   // Gfs ghost n = 3 - rr: -F(3 + n) (src/pml_fs_kernels.cu:7-19)
                        float fxn, fzn;
                        fze_row(6 - rr, fxn, fzn);
                        bx[q] = -fxn;
                        bz[q] = -fzn;
                    }
                } else {
                    if (rr >= 0) fze_row(rr, bx[q], bz[q]);
                    else {   // Gs ghost n = -rr: 2 F(0) - F(n)
                        float fx0, fz0, fxn, fzn;
                        fze_row(0, fx0, fz0);
                        fze_row(-rr, fxn, fzn);
                        bx[q] = 2.0f * fx0 - fxn;
                        bz[q] = 2.0f * fz0 - fzn;
                    }
                }
            }
            const float sx = drp5<DIR>(ax[0], ax[1], ax[2], ax[3], ax[4]) + drp5<DIR>(bx[0], bx[1], bx[2], bx[3], bx[4]);
            const float sz = drp5<DIR>(az[0], az[1], az[2], az[3], az[4]) + drp5<DIR>(bz[0], bz[1], bz[2], bz[3], bz[4]);
            const float Jc = c.J[gi];
            rvx = sx / Jc;
            rvz = sz / Jc;
            if constexpr (TOP == TOP_FREE) {
                // xi-memory source from distance-dependent zeta derivatives, no
                // zeta-memory source (include/pml_fs_kernels.cuh:212-233)
                const float dxsx = dx_lds<DIR>(ssx[r], cc), dxsz = dx_lds<DIR>(ssz[r], cc);
                const float dxsxz = dx_lds<DIR>(ssxz[r], cc);
                const float dzsx = mac_dz<DIR>(ssx, r, cc, dist, true), dzsz = mac_dz<DIR>(ssz, r, cc, dist, true);
                const float dzsxz = mac_dz<DIR>(ssxz, r, cc, dist, true);
                Dxx = xx * (dxsx + gxize / gxixi * dzsx) + xz * (dxsxz + gxize / gxixi * dzsxz);
                Dxz = xx * (dxsxz + gxize / gxixi * dzsxz) + xz * (dxsz + gxize / gxixi * dzsz);
                Dzx = 0.0f;
                Dzz = 0.0f;
            } else {
                const float pxi = (xx * sx + xz * sz) / (gxixi * Jc);
                const float pze = (zx * sx + zz * sz) / (gzeze * Jc);
                Dxx = xx * pxi; Dxz = xz * pxi;
                Dzx = zx * pze; Dzz = zz * pze;
            }
        } else {
            const float a = dx_lds<DIR>(ssx[r], cc), b = dz_lds<DIR>(ssx, r, cc);
            const float cz = dx_lds<DIR>(ssz[r], cc), d = dz_lds<DIR>(ssz, r, cc);
            const float e = dx_lds<DIR>(ssxz[r], cc), f = dz_lds<DIR>(ssxz, r, cc);
            rvx = xx * a + zx * b + xz * e + zz * f;
            rvz = xx * e + zx * f + xz * cz + zz * d;
            // PML sources: only the strip cells pay for the metric divisions
            Dxx = Dxz = Dzx = Dzz = 0.0f;
            if (side) {
                const float gx = gxize / gxixi;
                Dxx = xx * (a + gx * b) + xz * (e + gx * f);
                Dxz = xx * (e + gx * f) + xz * (cz + gx * d);
            }
            if (bottom || top) {
                const float gz = gxize / gzeze;
                Dzx = zx * (b + gz * a) + zz * (gz * e + f);
                Dzz = zx * (gz * e + f) + zz * (gz * cz + d);
            }
        }
        const float rinv = cur.cf[4];
        float kvx = rvx * rinv, kvz = rvz * rinv;

        // ---- xi-PML memory (side strips)
        if (side) {
            const long long ai = (long long)iz * g.W2 + scol;
            const float Vx = S[g.off[MSFDM_SXVX] + ai], Vz = S[g.off[MSFDM_SXVZ] + ai];
            const float Ex = S[g.off[MSFDM_SXEX] + ai], Ez = S[g.off[MSFDM_SXEZ] + ai];
            const float q1 = pxc - 1.0f;
            float kEx, kEz;
            if (surf) {   // include/pml_fs_kernels.cuh:50-79: memory of the xi derivatives
                kEx = -pxa * Ex + pxb * dvxx;
                kEz = -pxa * Ez + pxb * dvzx;
                const float cxx_x = c11 * a1xx + c13 * a1zz, cxx_z = c11 * a2xx + c13 * a2zz;
                const float czz_x = c13 * a1xx + c33 * a1zz, czz_z = c13 * a2xx + c33 * a2zz;
                ksx += q1 * (cxx_x * dvxx + cxx_z * dvzx) - pxc * (cxx_x * Ex + cxx_z * Ez);
                ksz += q1 * (czz_x * dvxx + czz_z * dvzx) - pxc * (czz_x * Ex + czz_z * Ez);
                ksxz += c55 * (q1 * (a1xz * dvxx + a2xz * dvzx) - pxc * (a1xz * Ex + a2xz * Ez));
            } else {
                const float Dvx = dvxx + (gxize / gxixi) * dvxz;
                const float Dvz = dvzx + (gxize / gxixi) * dvzz;
                kEx = -pxa * Ex + pxb * Dvx;
                kEz = -pxa * Ez + pxb * Dvz;
                ksx += c11 * q1 * xx * Dvx + c13 * q1 * xz * Dvz - pxc * c11 * xx * Ex - pxc * c13 * xz * Ez;
                ksz += c13 * q1 * xx * Dvx + c33 * q1 * xz * Dvz - pxc * c13 * xx * Ex - pxc * c33 * xz * Ez;
                ksxz += c55 * q1 * (xz * Dvx + xx * Dvz) - pxc * c55 * xz * Ex - pxc * c55 * xx * Ez;
            }
            const float kVx = -pxa * Vx + pxb * Dxx;
            const float kVz = -pxa * Vz + pxb * Dxz;
            kvx += (q1 * Dxx - pxc * Vx) * rinv;
            kvz += (q1 * Dxz - pxc * Vz) * rinv;
            rk<STAGE>(U, T2, T3, O, g.off[MSFDM_SXVX] + ai, Vx, kVx, ct);
            rk<STAGE>(U, T2, T3, O, g.off[MSFDM_SXVZ] + ai, Vz, kVz, ct);
            rk<STAGE>(U, T2, T3, O, g.off[MSFDM_SXEX] + ai, Ex, kEx, ct);
            rk<STAGE>(U, T2, T3, O, g.off[MSFDM_SXEZ] + ai, Ez, kEz, ct);
        }
        // ---- zeta-PML memory (bottom strip; CPML solid-only grid: top strip
        // first, rows 3..npd+2, then the bottom strip)
        if (bottom || top) {
            const long long ai = (long long)(top ? iz - 3 : (g.ztop_pml ? npd : 0) + iz - zb0) * P + ix;
            const float pza = c.pza[iz], pzb = c.pzb[iz], pzc = c.pzc[iz];
            const float Vx = S[g.off[MSFDM_SZVX] + ai], Vz = S[g.off[MSFDM_SZVZ] + ai];
            const float Ex = S[g.off[MSFDM_SZEX] + ai], Ez = S[g.off[MSFDM_SZEZ] + ai];
            const float Dvx = dvxz + (gxize / gzeze) * dvxx;
            const float Dvz = dvzz + (gxize / gzeze) * dvzx;
            const float kEx = -pza * Ex + pzb * Dvx;
            const float kEz = -pza * Ez + pzb * Dvz;
            const float q1 = pzc - 1.0f;
            ksx += c11 * q1 * zx * Dvx + c13 * q1 * zz * Dvz - pzc * c11 * zx * Ex - pzc * c13 * zz * Ez;
            ksz += c13 * q1 * zx * Dvx + c33 * q1 * zz * Dvz - pzc * c13 * zx * Ex - pzc * c33 * zz * Ez;
            ksxz += c55 * q1 * (zz * Dvx + zx * Dvz) - pzc * c55 * zz * Ex - pzc * c55 * zx * Ez;
            const float kVx = -pza * Vx + pzb * Dzx;
            const float kVz = -pza * Vz + pzb * Dzz;
            kvx += (q1 * Dzx - pzc * Vx) * rinv;
            kvz += (q1 * Dzz - pzc * Vz) * rinv;
            rk<STAGE>(U, T2, T3, O, g.off[MSFDM_SZVX] + ai, Vx, kVx, ct);
            rk<STAGE>(U, T2, T3, O, g.off[MSFDM_SZVZ] + ai, Vz, kVz, ct);
            rk<STAGE>(U, T2, T3, O, g.off[MSFDM_SZEX] + ai, Ex, kEx, ct);
            rk<STAGE>(U, T2, T3, O, g.off[MSFDM_SZEZ] + ai, Ez, kEz, ct);
        }
        // ---- source: 5x5 Gaussian-weighted Ricker on both normal stresses
        // (solid only, src/pml_fs_kernels.cu:22-27, include/tools_func.cuh:21-37)
        if constexpr (SOLO && !INT) {
            const int sxo = ix - ctl.srcx, szo = iz - ctl.srcz;
            if (sxo >= -2 && sxo <= 2 && szo >= -2 && szo <= 2) {
                const float w = amp * SRC_W[szo + 2][sxo + 2];
                ksx += w;
                ksz += w;
            }
        }
        if (!SOLO && !INT && iz == 0) {
            constexpr int par = STAGE == 0 ? 0 : step_par<DIR, STAGE>();
            float *kk = kif + (long long)kif_set(STAGE, par) * P;
            kk[3 * (long long)P + ix] = kvx;
            kk[4 * (long long)P + ix] = kvz;
            kk[5 * (long long)P + ix] = ksx;
            kk[6 * (long long)P + ix] = ksz;
            kk[7 * (long long)P + ix] = ksxz;
            if constexpr (STAGE == 1) {   // S = U^n here
                kif[(long long)(kif_u(par) + 3) * P + ix] = svx[r][cc];
                kif[(long long)(kif_u(par) + 4) * P + ix] = svz[r][cc];
                kif[(long long)(kif_u(par) + 5) * P + ix] = ssx[r][cc];
                kif[(long long)(kif_u(par) + 6) * P + ix] = ssz[r][cc];
                kif[(long long)(kif_u(par) + 7) * P + ix] = ssxz[r][cc];
            }
            if constexpr (STAGE != 0) continue;
        }
        float vz_new;
        if constexpr (STAGE == 4) {
            const long long idx[5] = {g.off[MSFDM_SVX] + si, g.off[MSFDM_SVZ] + si, g.off[MSFDM_SSX] + si,
                                      g.off[MSFDM_SSZ] + si, g.off[MSFDM_SSXZ] + si};
            const float sv[5] = {svx[r][cc], svz[r][cc], ssx[r][cc], ssz[r][cc], ssxz[r][cc]};
            const float kv[5] = {kvx, kvz, ksx, ksz, ksxz};
            float v[5];
            rk4_cell<5>(U, T2, T3, O, idx, sv, kv, ct, v);
            vz_new = v[1];
        } else {
            rk_u<STAGE>(U, T2, T3, O, g.off[MSFDM_SVX] + si, svx[r][cc], kvx, ct, cur.u[0]);
            vz_new = rk_u<STAGE>(U, T2, T3, O, g.off[MSFDM_SVZ] + si, svz[r][cc], kvz, ct, cur.u[1]);
            rk_u<STAGE>(U, T2, T3, O, g.off[MSFDM_SSX] + si, ssx[r][cc], ksx, ct, cur.u[2]);
            rk_u<STAGE>(U, T2, T3, O, g.off[MSFDM_SSZ] + si, ssz[r][cc], ksz, ct, cur.u[3]);
            rk_u<STAGE>(U, T2, T3, O, g.off[MSFDM_SSXZ] + si, ssxz[r][cc], ksxz, ct, cur.u[4]);
        }
        if constexpr (SOLO && STAGE == 4) {
            // receiver row (3 with the free surface, 30 with CPML), vz after the
            // step (src/main.cu:297-322), plane 1 of the record buffer
            const int it = ctl.base[0] + jstep;
            if (iz == ctl.rec_row && it % ctl.gap == 0) {
                const int slot = it / ctl.gap;
                if (slot < ctl.rec_cap) ctl.rec[(long long)ctl.rec_cap * nx + (long long)slot * nx + ix] = vz_new;
            }
        }
        (void)vz_new;
    }
}

// One RK stage over both media in a single launch.  The grid is 1-D with a
// multiple of 8 blocks; consecutive blocks go round-robin to the 8 XCDs.  Each
// XCD takes, in this order, one eighth of the interface-band tiles (the last
// nfi fluid bands, whose loaded rows hold the interface row, and solid band 0:
// they also rebuild or finish interface rows, so they are the slowest tiles),
// column by column so that an XCD holds the whole interface stack of its
// columns; then one eighth of the other fluid tiles and one eighth of the other
// solid tiles, each as a contiguous run of row bands, so every XCD finds the
// halo rows of its neighbours in its own L2 and gets the same mix of cheap,
// dear and slow tiles.
// Tile (bx, band) of block b, band -1 past the end of a run; bands 0..nbf-1
// are fluid, nbf..nbf+nbs-1 solid.
__device__ __forceinline__ int2 xcd_tile(int b, int nbx, int nbf, int nbs, int nfi) {
    const int x = b & 7, l = b >> 3;
    const int ni = nbx * (nfi + 1), nf = nbx * (nbf - nfi), ns = nbx * (nbs - 1);
    const int ci = (ni + 7) >> 3, cf = (nf + 7) >> 3, cs = (ns + 7) >> 3;
    if (l < ci) {
        const int t = x * ci + l;
        if (t >= ni) return make_int2(0, -1);
        const int bx = t / (nfi + 1);
        return make_int2(bx, nbf - nfi + (t - bx * (nfi + 1)));
    }
    if (l < ci + cf) {
        const int t = x * cf + (l - ci);
        if (t >= nf) return make_int2(0, -1);
        const int band = t / nbx;
        return make_int2(t - band * nbx, band);
    }
    const int t = x * cs + (l - ci - cf);
    if (t >= ns) return make_int2(0, -1);
    const int band = t / nbx;
    return make_int2(t - band * nbx, nbf + 1 + band);
}

// Interior tiles: every column outside the side strips, every row at least 3
// rows from a surface, the interface or a closed top, outside the zeta strips,
// and (fluid) the loaded rows clear of the interface row, so no rebuild either.
// A tile under the source's 5 x 5 footprint is not interior either.
__device__ __forceinline__ bool cols_interior(const Geo &g, int x0) {
    return x0 >= 3 + g.npd && x0 + TX - 1 <= g.nx - 4 - g.npd;
}
__device__ __forceinline__ bool clear_of_source(const Ctl &ctl, int x0, int z0, int TZ) {
    return ctl.srcx + 2 < x0 || ctl.srcx - 2 > x0 + TX - 1 || ctl.srcz + 2 < z0 || ctl.srcz - 2 > z0 + TZ - 1;
}
template <int FT>
__device__ __forceinline__ bool fluid_interior(const Geo &g, int x0, int z0, int TZ) {
    return cols_interior(g, x0) && (FT == 1 ? z0 > g.npd + 2 : z0 >= 6) && z0 + TZ + 2 < g.nzf - 1;
}
template <int TOP>
__device__ __forceinline__ bool solid_interior(const Geo &g, int x0, int z0, int TZ) {
    constexpr int ztop = (TOP == TOP_FREE || TOP == TOP_PLAIN) ? 3 : 0;
    return cols_interior(g, x0) && (TOP == TOP_PLAIN || z0 - ztop >= 3) && !(g.ztop_pml && z0 <= g.npd + 2) &&
           z0 + TZ - 1 < g.nzs - 3 - g.npd;
}

// Resident waves per SIMD of the fluid-solid stage kernels: at most 4 for stage
// 4, the stage that streams the most bytes (its registers would allow 5, which
// measured slower); at least 5 for stage 1, which moves the fewest bytes per
// cell and so gains from more waves in flight (22.7 against 23.3 us on the
// 2001 x 1001 seafloor grid, with a 40-byte spill in its boundary path); the
// compiler's choice (4) for stages 2 and 3, which lose 2.8 us at 5.
template <int STAGE> __host__ __device__ constexpr int stage_max_waves() { return STAGE == 4 ? 4 : 8; }
template <int STAGE> __host__ __device__ constexpr int stage_min_waves() { return STAGE == 1 ? 5 : 1; }

// The other occupancy policy the autotuner may pick for a grid
// (msfdm_set_occupancy 1): at least 5 resident waves per SIMD in every stage,
// compiled as stage_kernel_w5 / free_stage_kernel_w5 from the same tile code.
constexpr int MORE_WAVES_MIN = 5;

// One block's tile of a fluid-solid RK stage.  FIN = 1 (stage 1): the launch
// also finishes the previous step's interface rows and record
// (finish_interface_row); cprev is then the time step.
template <int DIR, int STAGE, int TZ, int MET, int CPML, int FIN>
__device__ __forceinline__ void stage_block(float (*sh)[TZ + 2 * HALO][LXP], float (*fin)[FW], const Geo &g,
                                            const Coef &c, const Ctl &ctl, const float *__restrict__ S,
                                            UPtr<STAGE> U, const float *__restrict__ T2,
                                            const float *__restrict__ T3, OPtr<STAGE> O, float *__restrict__ kif,
                                            int jstep, float ct, float amp_direct, float cprev, int nbx, int nbz_f,
                                            int nbz_s, int nfi) {
    static_assert(FIN == 0 || STAGE == 1, "only stage 1 finishes the previous step");
    const int2 tb = xcd_tile(blockIdx.x, nbx, nbz_f, nbz_s, nfi);
    if (tb.y < 0) return;
    const int bx = tb.x, band = tb.y;
    constexpr int TOP = CPML ? TOP_IFACE_CPML : TOP_IFACE;
    if (band < nbz_f) {
        const int z0 = fluid_z0(g, band, TZ);
        if (fluid_interior<CPML>(g, bx * TX, z0, TZ) && clear_of_source(ctl, bx * TX, z0, TZ))
            fluid_body<DIR, STAGE, TZ, MET, CPML, FIN, 1>(sh, g, c, ctl, S, U, T2, T3, O, kif, jstep, ct, amp_direct,
                                                          cprev, bx, band, fin);
        else
            fluid_body<DIR, STAGE, TZ, MET, CPML, FIN, 0>(sh, g, c, ctl, S, U, T2, T3, O, kif, jstep, ct, amp_direct,
                                                          cprev, bx, band, fin);
    } else {
        const int sb = band - nbz_f;
        if (solid_interior<TOP>(g, bx * TX, sb * TZ, TZ))
            solid_body<DIR, STAGE, TZ, MET, TOP, FIN, 1>(sh, g, c, ctl, S, U, T2, T3, O, kif, jstep, ct, amp_direct,
                                                         cprev, bx, sb, fin);
        else
            solid_body<DIR, STAGE, TZ, MET, TOP, FIN, 0>(sh, g, c, ctl, S, U, T2, T3, O, kif, jstep, ct, amp_direct,
                                                         cprev, bx, sb, fin);
    }
}

template <int DIR, int STAGE, int TZ, int MET, int CPML = 0, int FIN = 0>
__global__ void __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(stage_min_waves<STAGE>(), stage_max_waves<STAGE>())))
stage_kernel(Geo g, Coef c, Ctl ctl, const float *__restrict__ S, UPtr<STAGE> U,
             const float *__restrict__ T2, const float *__restrict__ T3, OPtr<STAGE> O,
             float *__restrict__ kif, int jstep, float ct, float amp_direct, float cprev, int nbx, int nbz_f,
             int nbz_s, int nfi) {
    __shared__ float sh[5][TZ + 2 * HALO][LXP];
    __shared__ float fin[FIN ? 8 : 1][FW];
    stage_block<DIR, STAGE, TZ, MET, CPML, FIN>(sh, fin, g, c, ctl, S, U, T2, T3, O, kif, jstep, ct, amp_direct,
                                                cprev, nbx, nbz_f, nbz_s, nfi);
}

template <int DIR, int STAGE, int TZ, int MET, int CPML = 0, int FIN = 0>
__global__ void __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(MORE_WAVES_MIN, 8)))
stage_kernel_w5(Geo g, Coef c, Ctl ctl, const float *__restrict__ S, UPtr<STAGE> U,
                const float *__restrict__ T2, const float *__restrict__ T3, OPtr<STAGE> O,
                float *__restrict__ kif, int jstep, float ct, float amp_direct, float cprev, int nbx, int nbz_f,
                int nbz_s, int nfi) {
    __shared__ float sh[5][TZ + 2 * HALO][LXP];
    __shared__ float fin[FIN ? 8 : 1][FW];
    stage_block<DIR, STAGE, TZ, MET, CPML, FIN>(sh, fin, g, c, ctl, S, U, T2, T3, O, kif, jstep, ct, amp_direct,
                                                cprev, nbx, nbz_f, nbz_s, nfi);
}

// One RK stage of a solid-only grid (PmlFs::SolidOnly, src/pml_fs_kernels.cu:29-381;
// Cpml::SolidOnly; Dirichlet, src/dirichlet_kernels.cu:12-163): 64 x TZ tiles
// of rows 3..nzs-4, the same XCD-aware block order.  Stage 4 also records vz
// of the receiver row.
template <int DIR, int STAGE, int TZ, int MET, int TOP>
__device__ __forceinline__ void free_stage_block(float (*sh)[TZ + 2 * HALO][LXP], const Geo &g, const Coef &c,
                                                 const Ctl &ctl, const float *__restrict__ S, UPtr<STAGE> U,
                                                 const float *__restrict__ T2, const float *__restrict__ T3,
                                                 OPtr<STAGE> O, int jstep, float ct, float amp_direct, int nbx,
                                                 int nb) {
    // band 0 (the surface rows, the dearest tiles) spread over the XCDs and
    // dispatched first, then contiguous runs of the other bands (xcd_tile with
    // no fluid bands)
    const int2 tb = xcd_tile(blockIdx.x, nbx, 0, nb, 0);
    if (tb.y < 0) return;
    const int bx = tb.x, band = tb.y;
    if (solid_interior<TOP>(g, bx * TX, 3 + band * TZ, TZ) && clear_of_source(ctl, bx * TX, 3 + band * TZ, TZ))
        solid_body<DIR, STAGE, TZ, MET, TOP, 0, 1>(sh, g, c, ctl, S, U, T2, T3, O, nullptr, jstep, ct, amp_direct, 0.0f,
                                                   bx, band);
    else
        solid_body<DIR, STAGE, TZ, MET, TOP, 0, 0>(sh, g, c, ctl, S, U, T2, T3, O, nullptr, jstep, ct, amp_direct, 0.0f,
                                                   bx, band);
}

template <int DIR, int STAGE, int TZ, int MET, int TOP>
__global__ void __launch_bounds__(NTHR)
free_stage_kernel(Geo g, Coef c, Ctl ctl, const float *__restrict__ S, UPtr<STAGE> U,
                  const float *__restrict__ T2, const float *__restrict__ T3, OPtr<STAGE> O, int jstep,
                  float ct, float amp_direct, int nbx, int nb) {
    __shared__ float sh[5][TZ + 2 * HALO][LXP];
    free_stage_block<DIR, STAGE, TZ, MET, TOP>(sh, g, c, ctl, S, U, T2, T3, O, jstep, ct, amp_direct, nbx, nb);
}

template <int DIR, int STAGE, int TZ, int MET, int TOP>
__global__ void __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(MORE_WAVES_MIN, 8)))
free_stage_kernel_w5(Geo g, Coef c, Ctl ctl, const float *__restrict__ S, UPtr<STAGE> U,
                     const float *__restrict__ T2, const float *__restrict__ T3, OPtr<STAGE> O, int jstep,
                     float ct, float amp_direct, int nbx, int nb) {
    __shared__ float sh[5][TZ + 2 * HALO][LXP];
    free_stage_block<DIR, STAGE, TZ, MET, TOP>(sh, g, c, ctl, S, U, T2, T3, O, jstep, ct, amp_direct, nbx, nb);
}

// Raw-k mode (STAGE 0, msfdm_eval_rhs): the interface condition applied to the
// raw k of the interface rows at column ix.
template <int DIR>
__device__ __forceinline__ void iface_rhs_body(const Geo &g, const Coef &c, float *__restrict__ O,
                                               const float *__restrict__ kif, int ix) {
    if (ix < 3 || ix > g.nx - 4) return;
    float k[8];
#pragma unroll
    for (int f = 0; f < 8; ++f) k[f] = kif[(long long)f * g.P + ix];
    interface_bc<DIR>(k, c.nif[ix]);
#pragma unroll
    for (int f = 0; f < 8; ++f) O[iface_index(g, f, ix)] = k[f];
}

template <int DIR>
__global__ void __launch_bounds__(128)
iface_rhs(Geo g, Coef c, float *__restrict__ O, const float *__restrict__ kif) {
    iface_rhs_body<DIR>(g, c, O, kif, blockIdx.x * blockDim.x + threadIdx.x);
}

// iface_final: the interface rows of U' and the record after the last step of a
// launch sequence (within a sequence the next step's stage-1 tiles do it, FIN).
// A block of FB threads loads FB consecutive interface columns starting at
// x0 - 3 (one per thread, all loads issued at once) and filters the FO = FB - 6
// columns x0 .. x0 + FO - 1 that have their full 7-tap neighbourhood in LDS.
constexpr int FB = 128;
constexpr int FO = FB - 2 * HALO;
template <int DIR>
__device__ __forceinline__ void iface_final_body(float (*pre)[FB], const Geo &g, const Coef &c,
                                                 const Ctl &ctl, float *__restrict__ Uo,
                                                 const float *__restrict__ kif, float c12, float c3, float cb,
                                                 int jstep, int x0) {
    const int nx = g.nx;
    const int t = threadIdx.x;
    const int ix = x0 - HALO + t;
    const bool inner = ix >= 3 && ix <= nx - 4;
    {
        float v[8];
        if (inner) {
            iface_pre<DIR>(g, c, kif, c12, c3, cb, ix, v);
        } else {
#pragma unroll
            for (int f = 0; f < 8; ++f) v[f] = 0.0f;
        }
#pragma unroll
        for (int f = 0; f < 8; ++f) pre[f][t] = v[f];
    }
    __syncthreads();
    if (!inner || t < HALO || t >= FB - HALO) return;
    float o[8];
    iface_filter(&pre[0][0], FB, t, ix, nx, c.nif[ix], o);
#pragma unroll
    for (int f = 0; f < 8; ++f) Uo[iface_index(g, f, ix)] = o[f];
    iface_record(ctl, ctl.base[0] + jstep, nx, ix, o);
}

template <int DIR>
__global__ void __launch_bounds__(FB)
iface_final(Geo g, Coef c, Ctl ctl, float *__restrict__ Uo, const float *__restrict__ kif, float c12, float c3,
            float cb, int jstep) {
    __shared__ float pre[8][FB];
    iface_final_body<DIR>(pre, g, c, ctl, Uo, kif, c12, c3, cb, jstep, blockIdx.x * FO);
}

// ---------------------------------------------------------------------------
// Batched launches: many shots (each with its own grid, model and buffers)
// stepped together, one launch per RK stage for all of them.
// ---------------------------------------------------------------------------
struct Shot {
    Geo g;
    Coef c;
    Ctl ctl;
    float *buf[5];    // U, T2, T3, T4, then where stage 4 writes U^{n+1}: U itself, or (the
                      // gradient's segment recomputation) the next step's U^n slot
    float *kif;
};

// tiles[t] = (shot, band, bx, fluid bands of that shot); FIN as in stage_kernel.
// MET as in stage_kernel: 1 when every shot of the launch has vertical grid
// columns (xi_x a constant, xi_z = 0), 2 to decide per shot at run time.
template <int DIR, int STAGE, int TZ, int MET, int FIN>
__device__ __forceinline__ void batch_stage_block(float (*sh)[TZ + 2 * HALO][LXP], float (*fin)[FW],
                                                  const Shot *__restrict__ shots, const int4 *__restrict__ tiles,
                                                  int ntiles, int rS, int rO, int jstep, float ct, float cprev) {
    static_assert(FIN == 0 || STAGE == 1, "only stage 1 finishes the previous step");
    const int per_xcd = gridDim.x >> 3;
    const int t = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (t >= ntiles) return;
    const int4 tl = tiles[t];
    const Shot &s = shots[tl.x];
    const float *S = s.buf[rS];
    float *O = s.buf[rO];
    if (tl.y < tl.w) {
        const int z0 = fluid_z0(s.g, tl.y, TZ);
        if (fluid_interior<0>(s.g, tl.z * TX, z0, TZ) && clear_of_source(s.ctl, tl.z * TX, z0, TZ))
            fluid_body<DIR, STAGE, TZ, MET, 0, FIN, 1>(sh, s.g, s.c, s.ctl, S, s.buf[0], s.buf[1], s.buf[2], O, s.kif,
                                                       jstep, ct, 0.0f, cprev, tl.z, tl.y, fin);
        else
            fluid_body<DIR, STAGE, TZ, MET, 0, FIN, 0>(sh, s.g, s.c, s.ctl, S, s.buf[0], s.buf[1], s.buf[2], O, s.kif,
                                                       jstep, ct, 0.0f, cprev, tl.z, tl.y, fin);
    } else {
        const int sb = tl.y - tl.w;
        if (solid_interior<TOP_IFACE>(s.g, tl.z * TX, sb * TZ, TZ))
            solid_body<DIR, STAGE, TZ, MET, TOP_IFACE, FIN, 1>(sh, s.g, s.c, s.ctl, S, s.buf[0], s.buf[1], s.buf[2], O,
                                                               s.kif, jstep, ct, 0.0f, cprev, tl.z, sb, fin);
        else
            solid_body<DIR, STAGE, TZ, MET, TOP_IFACE, FIN, 0>(sh, s.g, s.c, s.ctl, S, s.buf[0], s.buf[1], s.buf[2], O,
                                                               s.kif, jstep, ct, 0.0f, cprev, tl.z, sb, fin);
    }
}


template <int DIR, int STAGE, int TZ, int MET, int FIN = 0>
__global__ void __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(stage_min_waves<STAGE>(), stage_max_waves<STAGE>())))
batch_stage_kernel(const Shot *__restrict__ shots, const int4 *__restrict__ tiles, int ntiles, int rS, int rO,
                   int jstep, float ct, float cprev) {
    __shared__ float sh[5][TZ + 2 * HALO][LXP];
    __shared__ float fin[FIN ? 8 : 1][FW];
    batch_stage_block<DIR, STAGE, TZ, MET, FIN>(sh, fin, shots, tiles, ntiles, rS, rO, jstep, ct, cprev);
}

// The batch under occupancy policy 1 (at least 5 waves per SIMD), vertical
// columns only.
template <int DIR, int STAGE, int TZ, int FIN = 0>
__global__ void __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(MORE_WAVES_MIN, 8)))
batch_stage_kernel_w5(const Shot *__restrict__ shots, const int4 *__restrict__ tiles, int ntiles, int rS, int rO,
                      int jstep, float ct, float cprev) {
    __shared__ float sh[5][TZ + 2 * HALO][LXP];
    __shared__ float fin[FIN ? 8 : 1][FW];
    batch_stage_block<DIR, STAGE, TZ, 1, FIN>(sh, fin, shots, tiles, ntiles, rS, rO, jstep, ct, cprev);
}

// grid (column blocks of the widest shot, shots)
template <int DIR>
__global__ void __launch_bounds__(FB)
batch_iface_final(const Shot *__restrict__ shots, float c12, float c3, float cb, int jstep) {
    __shared__ float pre[8][FB];
    const Shot &s = shots[blockIdx.y];
    if ((int)blockIdx.x * FO > s.g.nx) return;   // whole block beyond this shot's grid
    iface_final_body<DIR>(pre, s.g, s.c, s.ctl, s.buf[4], s.kif, c12, c3, cb, jstep, blockIdx.x * FO);
}

__global__ void advance_base(int *base, int n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) base[0] += n;
}

__global__ void set_base(int *base, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) base[0] = v;
}

}  // namespace dev
}  // namespace msfdm
