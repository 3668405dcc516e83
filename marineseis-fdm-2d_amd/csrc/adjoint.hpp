// Device code of the discrete adjoint (reverse-time) stepper.
//
// The forward stage is k = B_dir(L_dir(S) + src) (interface BC B, RHS L of
// include/pml_fs_kernels.cuh:333-827).  One adjoint stage takes
// kbar = clam * LP + cprev * TBprev (the RK4 stage weights), applies B^T on the
// interface rows (iface_qbar, inside the products kernel) and forms
// TB = L^T(qbar) in two launches:
//   adj_prod_kernel    per source cell (64 x TZ tiles, the forward state S of
//                      the tile and its 3-cell halo in LDS): every forward
//                      expression k[i] = c(i) D(S)[i] (D a one-sided stencil
//                      along x or z) has the transpose
//                      Sbar[j] += w * c(i) qbar[i] over the stencil's cells j;
//                      the kernel evaluates the pointwise products c(i) qbar[i]
//                      of all of a cell's stencils (the "adjoint products") and
//                      stores them, together with the pointwise parts of the
//                      transpose (PML memory) and the model-gradient terms
//                      dk/dm . qbar, which accumulate in FP64.
//   adj_gather_kernel  per output cell (64 x TZ tiles, the products of the
//                      tile and its 3-cell halo in LDS): Sbar[j] = the sum of
//                      w * product over every source cell whose stencil holds
//                      j, with the weights of that source row's stencil -- the
//                      DRP taps, the 4-2 / 2-2 closures, and the ghost-flux
//                      rows of the interface and free surface with their
//                      reflections.
// Every output cell has one writer and sums its terms in a fixed order, so
// repeated runs give the same bits; nothing needs zeroing between stages.
//   adj_filter        Filter^T on the interface rows (LDS; its input rows are
//                     staged by iface_rows_copy, so it can write them in place).
// The per-column 8x8 interface BC and filter maps are transposed by probing
// the forward maps with unit vectors.
#pragma once

#include "stencil.hpp"

namespace msfdm {
namespace dev {

struct Grad {
    double *K, *rho_f;                 // [nzf][P]
    double *C11, *C13, *C33, *C55, *rho_s;   // [nzs][P]
    long long gstride;                 // doubles between the gradients of consecutive shots
};

// Gradients of shot `shot` of a multi-shot launch (blockIdx.y).
__device__ __forceinline__ Grad grad_of_shot(Grad g, int shot) {
    const long long o = (long long)shot * g.gstride;
    g.K += o; g.rho_f += o; g.C11 += o; g.C13 += o; g.C33 += o; g.C55 += o; g.rho_s += o;
    return g;
}

// weights of the one-sided operators at offsets (ascending)
template <int DIR> __device__ __forceinline__ float drp_w(int q) {   // q = 0..4 for offsets lo_off..lo_off+4
    if constexpr (DIR == 0) {
        return q == 0 ? DRP_M1 : q == 1 ? DRP_0 : q == 2 ? DRP_1 : q == 3 ? DRP_2 : DRP_3;
    } else {   // offsets -3..1: -(A3, A2, A1, A0, AM1)
        return q == 0 ? -DRP_3 : q == 1 ? -DRP_2 : q == 2 ? -DRP_1 : q == 3 ? -DRP_0 : -DRP_M1;
    }
}

// Weight of row offset m in the DRP stencil of direction DIR (0 outside it).
template <int DIR> __device__ __forceinline__ float drp_at(int m) {
    const int q = m - lo_off<DIR>();
    return (q >= 0 && q <= 4) ? drp_w<DIR>(q) : 0.0f;
}

// Weight of row offset m in the distance-dependent zeta operator mac_dz
// (Mac_dv_dze, include/tools_func.cuh:131-140) of a source row at distance
// dist from the boundary.
template <int DIR>
__device__ __forceinline__ float mac_at(int dist, bool solid_side, int m) {
    if (dist >= 3) return drp_at<DIR>(m);
    if (dist == 2) {
        if constexpr (DIR == 0) return m == 0 ? M42_0 : m == 1 ? M42_1 : m == 2 ? M42_2 : 0.0f;
        else return m == 0 ? -M42_0 : m == -1 ? -M42_1 : m == -2 ? -M42_2 : 0.0f;
    }
    bool fwd = DIR == 0;
    if (dist == 0) fwd = solid_side;
    if (fwd) return m == 0 ? -1.0f : m == 1 ? 1.0f : 0.0f;
    return m == 0 ? 1.0f : m == -1 ? -1.0f : 0.0f;
}

struct AdjIn {
    const float *S;      // forward stage state
    const float *LP;     // lambda'
    const float *TBp;    // previous stage's Tbar (LP with cprev = 0 when there is none)
    float *TB;           // output Tbar (null: the last stage of a step, whose Tbar only goes to ACC)
    float *ACC;          // null, or the step's accumulator: ACC = ACCin + TB wherever the stage writes TB
    const float *ACCin;  // the accumulator's input: ACC itself, or LP for the first stage of a step
                         // (which so sets ACC = LP + TB on every entry a stage writes)
    float *prod;         // adjoint products: [ADJ_NQF][nzf][P], then [ADJ_NQS][nzs][P]
    float clam, cprev;
    long long bstride;   // floats between the state buffers of consecutive shots
    long long pstride;   // floats between the product buffers of consecutive shots
};

// Adjoint products per cell.  Fluid: 0, 1 = the velocity equations' pressure
// stencils (x, z: dpx_bar, dpz_bar; on the ghost-flux rows near the surface
// and the interface the flux sums sx_bar, sz_bar instead); 2..5 = the pressure
// equation's velocity stencils (vx along x, vz along x, vx along z, vz along
// z).  Solid: 0..3 = the stress equations' velocity stencils (vx x, vz x,
// vx z, vz z); 4..9 = the velocity equations' stress stencils (sxx x, szz x,
// sxz x, sxx z, szz z, sxz z; on the ghost-flux rows 0..2 the flux sums
// sx_bar, sz_bar in 4, 5 and zeros in 6..9).
constexpr int ADJ_NQF = 6, ADJ_NQS = 10;

__device__ __forceinline__ long long prod_fluid(const Geo &g, int k) { return (long long)k * g.nzf * g.P; }
__device__ __forceinline__ long long prod_solid(const Geo &g, int k) {
    return (long long)ADJ_NQF * g.nzf * g.P + (long long)k * g.nzs * g.P;
}

// Inputs of shot `shot` of a multi-shot launch (blockIdx.y); every shot's
// buffers sit at a fixed stride from shot 0's.
__device__ __forceinline__ AdjIn adj_of_shot(AdjIn a, int shot) {
    const long long o = (long long)shot * a.bstride;
    a.S += o; a.LP += o; a.TBp += o;
    if (a.TB) a.TB += o;
    if (a.ACC) { a.ACC += o; a.ACCin += o; }
    a.prod += (long long)shot * a.pstride;
    return a;
}

// Load from memory that no thread writes during the kernel (the adjoint
// inputs lambda', the previous stage's Tbar, the interface qbar, the products
// in the gather kernel, the model coefficients and metrics): through the
// constant address space, so the compiler may issue it ahead of earlier
// stores to the outputs instead of waiting for them.
template <class T>
__device__ __forceinline__ T ldc(const T *p) {
    return *(const __attribute__((address_space(4))) T *)p;
}

// Tbar[i] = v (unless the stage keeps no Tbar), and ACC[i] = ACCin[i] + v
// when the stage accumulates into ACC (each entry has one writer per stage).
__device__ __forceinline__ void tb_out(const AdjIn &a, long long i, float v) {
    if (a.TB) a.TB[i] = v;
    if (a.ACC) a.ACC[i] = a.ACCin[i] + v;
}

// kbar = clam LP + cprev TBp (the host passes TBp = LP and cprev = 0 for the
// first stage of a step, which has no previous Tbar).
__device__ __forceinline__ float kbar(const AdjIn &a, long long i) {
    float v = a.clam * ldc(&a.LP[i]);
    v += a.cprev * ldc(&a.TBp[i]);
    return v;
}

// 8x8 interface BC matrix of one column (columns = unit inputs).
template <int DIR>
__device__ __forceinline__ void bc_matrix(float2 n, float M[8][8]) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        float k[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) k[i] = i == j ? 1.0f : 0.0f;
        interface_bc<DIR>(k, n);
#pragma unroll
        for (int i = 0; i < 8; ++i) M[i][j] = k[i];
    }
}

// qbar = B^T kbar of interface column ix (3 <= ix <= nx - 4): q[0..2] for the
// fluid row nzf - 1, q[3..7] for solid row 0, from the kbar of all 8
// interface-row fields.  The products tiles that hold an interface row
// evaluate it for their own columns.
template <int DIR>
__device__ __forceinline__ void iface_qbar(const Geo &g, const Coef &c, const AdjIn &a, int ix, float q[8]) {
    float kb[8];
#pragma unroll
    for (int f = 0; f < 8; ++f) kb[f] = kbar(a, iface_index(g, f, ix));
    float M[8][8];
    bc_matrix<DIR>(c.nif[ix], M);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        float s = 0.0f;
#pragma unroll
        for (int i = 0; i < 8; ++i) s += M[i][j] * kb[i];
        q[j] = s;
    }
}

// Interior source tiles (the products of a cell come from the plain DRP
// stencils and no PML term): every column outside the side strips, every row
// at least 3 rows from the surface, the interface and the bottom strip.  Fluid
// tiles start at row 3 + band * TZ, solid tiles at row band * TZ.  The host
// sorts the tiles of a grid into an interior and a boundary list with these.
__host__ __device__ inline bool adj_cols_interior(int nx, int npd, int x0) {
    return x0 >= 3 + npd && x0 + TX - 1 <= nx - 4 - npd;
}
__host__ __device__ inline bool adj_fluid_interior(int nx, int npd, int nzf, int x0, int z0, int TZ) {
    return adj_cols_interior(nx, npd, x0) && z0 >= 6 && z0 + TZ - 1 <= nzf - 4;
}
__host__ __device__ inline bool adj_solid_interior(int nx, int npd, int nzs, int x0, int z0, int TZ) {
    return adj_cols_interior(nx, npd, x0) && z0 >= 3 && z0 + TZ - 1 < nzs - 3 - npd;
}

// ---------------------------------------------------------------------------
// Products, fluid source tile: rows 3..nzf-1 (z0 = 3 + bz * TZ).  The
// arithmetic is the transpose of fluid_body, expression by expression.
// ---------------------------------------------------------------------------
template <int DIR, int TZ, int MET, int INT>
__device__ __forceinline__ void fluid_prod_body(const float (*sh)[TZ + 2 * HALO][LXP], const Geo &g, const Coef &c,
                                                const AdjIn &a, const Grad &gr, int bx, int bz) {
    // a wave holds consecutive rows of the tile: ty (and every row test) is
    // uniform per wave, which readfirstlane tells the compiler
    const int tx = threadIdx.x & (TX - 1), ty = __builtin_amdgcn_readfirstlane(threadIdx.x / TX);
    constexpr int RW = TZ / 4;   // consecutive tile rows per wave
    const int x0 = bx * TX, z0 = 3 + bz * TZ;
    const int P = g.P, nx = g.nx, nzf = g.nzf;
    const float *S = a.S;
    const float (*shvx)[LXP] = sh[0];
    const float (*shvz)[LXP] = sh[1];
    const float (*shp)[LXP] = sh[2];
    const int sc = tx + XOFF;
    const int ix = x0 + tx;
    const bool act = INT || (ix >= 3 && ix <= nx - 4);
    const int scol = (INT || !act) ? -1 : side_col(ix, nx, g.npd);
    float *Q = a.prod;
    const long long pP = (long long)nzf * P;
    // the FP64 gradient sums of the wave's cells, loaded before any store of
    // the tile (no load waits behind an earlier row's stores)
    double grho[RW], gK[RW];
#pragma unroll
    for (int j = 0; j < RW; ++j) {
        const int iz = z0 + ty * RW + j;
        const long long gi = (long long)iz * P + ix;
        const bool in = act && (INT || iz <= nzf - 1);
        grho[j] = in ? gr.rho_f[gi] : 0.0;
        gK[j] = in && iz != 3 ? gr.K[gi] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < RW; ++j) {
        const int iz = z0 + ty * RW + j;
        if (!INT && iz > nzf - 1) break;   // uniform over the wave
        const long long gi = (long long)iz * P + ix;
        const int r = iz - z0 + HALO;   // tile row of the cell in sh
        float q[ADJ_NQF] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (act) {
            const int dist_flso = nzf - 1 - iz, dist_fs = iz - 3;
            const double g_rho = grho[j], g_K = gK[j];
            const float J = ldc(&c.J[gi]), zx = ldc(&c.zex[gi]), zz = ldc(&c.zez[gi]);
            float xx, xz;
            xi_metrics<MET>(c, gi, xx, xz);
            const float gxixi = xx * xx + xz * xz;
            const bool at_row = !INT && dist_flso == 0;
            float qi[8];
            if (at_row) iface_qbar<DIR>(g, c, a, ix, qi);
            const float qvx = at_row ? qi[0] : kbar(a, g.off[MSFDM_FVX] + gi);
            const float qvz = at_row ? qi[1] : kbar(a, g.off[MSFDM_FVZ] + gi);
            const float qp = at_row ? qi[2] : kbar(a, g.off[MSFDM_FP] + gi);
            const float rinv = ldc(&c.rhof_inv[gi]);
            // ---- velocity equations
            float Dxb = 0.f, Dzb = 0.f;
            const float rvxb = qvx * rinv, rvzb = qvz * rinv;
            float pxa = 0.f, pxb = 0.f, pxc = 1.f;
            long long ai = 0;
            float Vx = 0.f, Vz = 0.f, Ex = 0.f, Ez = 0.f;
            if (scol >= 0) {
                pxa = ldc(&c.pxa[ix]); pxb = ldc(&c.pxb[ix]); pxc = ldc(&c.pxc[ix]);
                ai = (long long)iz * g.W2 + scol;
                const float qVx = kbar(a, g.off[MSFDM_FXVX] + ai), qVz = kbar(a, g.off[MSFDM_FXVZ] + ai);
                Dxb = (pxc - 1.0f) * rinv * qvx + pxb * qVx;
                Dzb = (pxc - 1.0f) * rinv * qvz + pxb * qVz;
                tb_out(a, g.off[MSFDM_FXVX] + ai, -pxc * rinv * qvx - pxa * qVx);
                tb_out(a, g.off[MSFDM_FXVZ] + ai, -pxc * rinv * qvz - pxa * qVz);
                Vx = S[g.off[MSFDM_FXVX] + ai]; Vz = S[g.off[MSFDM_FXVZ] + ai];
                Ex = S[g.off[MSFDM_FXEX] + ai]; Ez = S[g.off[MSFDM_FXEZ] + ai];
            }
            float rvx, rvz, Dx = 0.f, Dz = 0.f;   // forward values for the gradient
            if (!INT && (dist_flso < 3 || dist_fs < 3)) {
                // ghost-flux rows: rho v_k = (D_xi(-J p xi) + D_ze(-J p ze)) / J
                const bool at_if = dist_flso < 3;
                auto fze_row = [&](int rr, float &fx, float &fz) {
                    const long long qq = (long long)rr * P + ix;
                    const float mJp = -ldc(&c.J[qq]) * shp[r + (rr - iz)][sc];
                    fx = mJp * ldc(&c.zex[qq]);
                    fz = mJp * ldc(&c.zez[qq]);
                };
                float sx = 0.f, sz = 0.f;
                constexpr int o = lo_off<DIR>();
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    const int m = o + k;
                    const float w = drp_w<DIR>(k);
                    const long long gq = gi + m;
                    const float mJp = -ldc(&c.J[gq]) * shp[r][sc + m];
                    float xxq, xzq;
                    xi_metrics<MET>(c, gq, xxq, xzq);
                    sx += w * mJp * xxq;
                    sz += w * mJp * xzq;
                    const int rr = iz + m;
                    float fbx, fbz;
                    if (at_if) {
                        if (rr <= nzf - 1) fze_row(rr, fbx, fbz);
                        else { float f0x, f0z, fnx, fnz; fze_row(nzf - 1, f0x, f0z); fze_row(2 * (nzf - 1) - rr, fnx, fnz); fbx = 2.f * f0x - fnx; fbz = 2.f * f0z - fnz; }
                    } else {
                        if (rr >= 3) fze_row(rr, fbx, fbz);
                        else { float fnx, fnz; fze_row(6 - rr, fnx, fnz); fbx = -fnx; fbz = -fnz; }
                    }
                    sx += w * fbx;
                    sz += w * fbz;
                }
                rvx = sx / J; rvz = sz / J;
                const float proj = (xx * sx + xz * sz) / (gxixi * J);
                Dx = xx * proj; Dz = xz * proj;
                const float projb = xx * Dxb + xz * Dzb;
                q[0] = rvxb / J + xx * projb / (gxixi * J);   // sx_bar
                q[1] = rvzb / J + xz * projb / (gxixi * J);   // sz_bar
            } else {
                const float dpx = dx_lds<DIR>(shp[r], sc), dpz = dz_lds<DIR>(shp, r, sc);
                const float gxize = xx * zx + xz * zz;
                rvx = -(xx * dpx + zx * dpz); rvz = -(xz * dpx + zz * dpz);
                float tb = 0.0f, tgx = 0.0f;
                if (scol >= 0) {
                    const float t = (gxixi * dpx + gxize * dpz) / gxixi;
                    Dx = -xx * t; Dz = -xz * t;
                    tb = -xx * Dxb - xz * Dzb;
                    tgx = tb * gxize / gxixi;
                }
                q[0] = -xx * rvxb - xz * rvzb + tb;   // dpx_bar
                q[1] = -zx * rvxb - zz * rvzb + tgx;  // dpz_bar
            }
            float kvx = rvx, kvz = rvz;
            if (scol >= 0) { kvx += (pxc - 1.0f) * Dx - pxc * Vx; kvz += (pxc - 1.0f) * Dz - pxc * Vz; }
            gr.rho_f[gi] = g_rho + -(double)(qvx * kvx + qvz * kvz) * (double)rinv * (double)rinv;
            // ---- pressure equation (row 3: p_k = 0, no dependence)
            if (dist_fs != 0) {
                const float Km = ldc(&c.Kf[gi]);
                const int dzn = INT ? 3 : min(dist_fs, dist_flso);
                const float dvxx = dx_lds<DIR>(shvx[r], sc), dvzx = dx_lds<DIR>(shvz[r], sc);
                const float dvxz = mac_dz<DIR>(shvx, r, sc, dzn, false), dvzz = mac_dz<DIR>(shvz, r, sc, dzn, false);
                float bracket = -(xx * dvxx + zx * dvxz + xz * dvzx + zz * dvzz);
                float bxx = -Km * xx * qp, bxz = -Km * zx * qp, bzx = -Km * xz * qp, bzz = -Km * zz * qp;
                if (scol >= 0) {
                    const float gg = (xx * zx + xz * zz) / gxixi;
                    const float qEx = kbar(a, g.off[MSFDM_FXEX] + ai), qEz = kbar(a, g.off[MSFDM_FXEZ] + ai);
                    const float Dvxb = -Km * (pxc - 1.0f) * xx * qp + pxb * qEx;
                    const float Dvzb = -Km * (pxc - 1.0f) * xz * qp + pxb * qEz;
                    tb_out(a, g.off[MSFDM_FXEX] + ai, Km * pxc * xx * qp - pxa * qEx);
                    tb_out(a, g.off[MSFDM_FXEZ] + ai, Km * pxc * xz * qp - pxa * qEz);
                    bxx += Dvxb; bxz += gg * Dvxb; bzx += Dvzb; bzz += gg * Dvzb;
                    const float Dvx = dvxx + gg * dvxz, Dvz = dvzx + gg * dvzz;
                    bracket += -(pxc - 1.0f) * xx * Dvx - (pxc - 1.0f) * xz * Dvz + pxc * xx * Ex + pxc * xz * Ez;
                }
                q[2] = bxx; q[3] = bzx; q[4] = bxz; q[5] = bzz;
                gr.K[gi] = g_K + (double)qp * (double)bracket;
            }
        }
#pragma unroll
        for (int k = 0; k < ADJ_NQF; ++k) Q[k * pP + gi] = q[k];
    }
}

// ---------------------------------------------------------------------------
// Products, solid source tile: rows 0..nzs-4 (z0 = bz * TZ), metrics at the
// global row iz + nzf - 1.  The transpose of solid_body, expression by
// expression.
// ---------------------------------------------------------------------------
template <int DIR, int TZ, int MET, int INT>
__device__ __forceinline__ void solid_prod_body(const float (*sh)[TZ + 2 * HALO][LXP], const Geo &g, const Coef &c,
                                                const AdjIn &a, const Grad &gr, int bx, int bz) {
    const int tx = threadIdx.x & (TX - 1), ty = __builtin_amdgcn_readfirstlane(threadIdx.x / TX);
    constexpr int RW = TZ / 4;
    const int x0 = bx * TX, z0 = bz * TZ;
    const int P = g.P, nx = g.nx, nzs = g.nzs, npd = g.npd;
    const float *S = a.S;
    const float (*svx)[LXP] = sh[0];
    const float (*svz)[LXP] = sh[1];
    const float (*ssx)[LXP] = sh[2];
    const float (*ssz)[LXP] = sh[3];
    const float (*ssxz)[LXP] = sh[4];
    const int sc = tx + XOFF;
    const int ix = x0 + tx;
    const bool act = INT || (ix >= 3 && ix <= nx - 4);
    const int scol = (INT || !act) ? -1 : side_col(ix, nx, npd);
    const long long moff = (long long)(g.nzf - 1) * P;
    const int zb0 = nzs - 3 - npd;
    float *Q = a.prod + prod_solid(g, 0);
    const long long pP = (long long)nzs * P;
    constexpr int o = lo_off<DIR>();
    // the FP64 gradient sums of the wave's cells, loaded before any store of
    // the tile (no load waits behind an earlier row's stores)
    double gc[RW][5];
#pragma unroll
    for (int j = 0; j < RW; ++j) {
        const int iz = z0 + ty * RW + j;
        const long long si = (long long)iz * P + ix;
        const bool in = act && (INT || iz <= nzs - 4);
        gc[j][0] = in ? gr.C11[si] : 0.0;
        gc[j][1] = in ? gr.C13[si] : 0.0;
        gc[j][2] = in ? gr.C33[si] : 0.0;
        gc[j][3] = in ? gr.C55[si] : 0.0;
        gc[j][4] = in ? gr.rho_s[si] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < RW; ++j) {
        const int iz = z0 + ty * RW + j;
        if (!INT && iz > nzs - 4) break;   // uniform over the wave
        const long long si = (long long)iz * P + ix, gi = moff + si;
        const int r = iz - z0 + HALO;
        float q[ADJ_NQS] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (act) {
            const double g11 = gc[j][0], g13 = gc[j][1], g33 = gc[j][2], g55 = gc[j][3], grs = gc[j][4];
            const float J = ldc(&c.J[gi]), zx = ldc(&c.zex[gi]), zz = ldc(&c.zez[gi]);
            float xx, xz;
            xi_metrics<MET>(c, gi, xx, xz);
            const float gxixi = xx * xx + xz * xz, gzeze = zx * zx + zz * zz, gxize = xx * zx + xz * zz;
            const bool side = scol >= 0, bottom = !INT && iz >= zb0;
            const bool at_row = !INT && iz == 0;
            float qi[8];
            if (at_row) iface_qbar<DIR>(g, c, a, ix, qi);
            const float qvx = at_row ? qi[3] : kbar(a, g.off[MSFDM_SVX] + si);
            const float qvz = at_row ? qi[4] : kbar(a, g.off[MSFDM_SVZ] + si);
            const float qsx = at_row ? qi[5] : kbar(a, g.off[MSFDM_SSX] + si);
            const float qsz = at_row ? qi[6] : kbar(a, g.off[MSFDM_SSZ] + si);
            const float qsxz = at_row ? qi[7] : kbar(a, g.off[MSFDM_SSXZ] + si);
            const float c11 = ldc(&c.C11[si]), c13 = ldc(&c.C13[si]), c33 = ldc(&c.C33[si]), c55 = ldc(&c.C55[si]);
            const float rinv = ldc(&c.rhos_inv[si]);
            const int dist = INT ? 3 : iz;

            // ---- stress equations
            {
                const float dvxx = dx_lds<DIR>(svx[r], sc), dvzx = dx_lds<DIR>(svz[r], sc);
                const float dvxz = mac_dz<DIR>(svx, r, sc, dist, true), dvzz = mac_dz<DIR>(svz, r, sc, dist, true);
                const float exx = xx * dvxx + zx * dvxz, ezz = xz * dvzx + zz * dvzz;
                float A11 = exx, A13 = ezz, A55 = xz * dvxx + zz * dvxz + xx * dvzx + zx * dvzz;
                const float exxb = c11 * qsx + c13 * qsz, ezzb = c13 * qsx + c33 * qsz;
                float bxx = xx * exxb + c55 * xz * qsxz, bxz = zx * exxb + c55 * zz * qsxz;
                float bzx = xz * ezzb + c55 * xx * qsxz, bzz = zz * ezzb + c55 * zx * qsxz;
                if (side) {
                    const float gx = gxize / gxixi;
                    const float pxa = ldc(&c.pxa[ix]), pxb = ldc(&c.pxb[ix]), pxc = ldc(&c.pxc[ix]), q1 = pxc - 1.0f;
                    const long long ai = (long long)iz * g.W2 + scol;
                    const float Ex = S[g.off[MSFDM_SXEX] + ai], Ez = S[g.off[MSFDM_SXEZ] + ai];
                    const float qEx = kbar(a, g.off[MSFDM_SXEX] + ai), qEz = kbar(a, g.off[MSFDM_SXEZ] + ai);
                    const float Dvxb = q1 * xx * exxb + c55 * q1 * xz * qsxz + pxb * qEx;
                    const float Dvzb = q1 * xz * ezzb + c55 * q1 * xx * qsxz + pxb * qEz;
                    tb_out(a, g.off[MSFDM_SXEX] + ai, -pxc * xx * exxb - pxc * c55 * xz * qsxz - pxa * qEx);
                    tb_out(a, g.off[MSFDM_SXEZ] + ai, -pxc * xz * ezzb - pxc * c55 * xx * qsxz - pxa * qEz);
                    bxx += Dvxb; bxz += gx * Dvxb; bzx += Dvzb; bzz += gx * Dvzb;
                    const float Dvx = dvxx + gx * dvxz, Dvz = dvzx + gx * dvzz;
                    A11 += q1 * xx * Dvx - pxc * xx * Ex;
                    A13 += q1 * xz * Dvz - pxc * xz * Ez;
                    A55 += q1 * (xz * Dvx + xx * Dvz) - pxc * xz * Ex - pxc * xx * Ez;
                }
                if (bottom) {
                    const float gz = gxize / gzeze;
                    const float pza = ldc(&c.pza[iz]), pzb = ldc(&c.pzb[iz]), pzc = ldc(&c.pzc[iz]), q1 = pzc - 1.0f;
                    const long long ai = (long long)(iz - zb0) * P + ix;
                    const float Ex = S[g.off[MSFDM_SZEX] + ai], Ez = S[g.off[MSFDM_SZEZ] + ai];
                    const float qEx = kbar(a, g.off[MSFDM_SZEX] + ai), qEz = kbar(a, g.off[MSFDM_SZEZ] + ai);
                    const float Dvxb = q1 * zx * exxb + c55 * q1 * zz * qsxz + pzb * qEx;
                    const float Dvzb = q1 * zz * ezzb + c55 * q1 * zx * qsxz + pzb * qEz;
                    tb_out(a, g.off[MSFDM_SZEX] + ai, -pzc * zx * exxb - pzc * c55 * zz * qsxz - pza * qEx);
                    tb_out(a, g.off[MSFDM_SZEZ] + ai, -pzc * zz * ezzb - pzc * c55 * zx * qsxz - pza * qEz);
                    bxz += Dvxb; bxx += gz * Dvxb; bzz += Dvzb; bzx += gz * Dvzb;
                    const float Dvx = dvxz + gz * dvxx, Dvz = dvzz + gz * dvzx;
                    A11 += q1 * zx * Dvx - pzc * zx * Ex;
                    A13 += q1 * zz * Dvz - pzc * zz * Ez;
                    A55 += q1 * (zz * Dvx + zx * Dvz) - pzc * zz * Ex - pzc * zx * Ez;
                }
                q[0] = bxx; q[1] = bzx; q[2] = bxz; q[3] = bzz;
                gr.C11[si] = g11 + (double)qsx * (double)A11;
                gr.C13[si] = g13 + ((double)qsx * (double)A13 + (double)qsz * (double)A11);
                gr.C33[si] = g33 + (double)qsz * (double)A13;
                gr.C55[si] = g55 + (double)qsxz * (double)A55;
            }
            // ---- velocity equations
            {
                const float rvxb = qvx * rinv, rvzb = qvz * rinv;
                float Dxxb = 0.f, Dxzb = 0.f, Dzxb = 0.f, Dzzb = 0.f;
                float q1x = 0.f, pxc = 1.f, q1z = 0.f, pzc = 1.f;
                float Vxs = 0.f, Vzs = 0.f, Vxb = 0.f, Vzb = 0.f;
                if (side) {
                    const float pxa = ldc(&c.pxa[ix]), pxb = ldc(&c.pxb[ix]);
                    pxc = ldc(&c.pxc[ix]); q1x = pxc - 1.0f;
                    const long long ai = (long long)iz * g.W2 + scol;
                    const float qVx = kbar(a, g.off[MSFDM_SXVX] + ai), qVz = kbar(a, g.off[MSFDM_SXVZ] + ai);
                    Dxxb = q1x * rinv * qvx + pxb * qVx;
                    Dxzb = q1x * rinv * qvz + pxb * qVz;
                    tb_out(a, g.off[MSFDM_SXVX] + ai, -pxc * rinv * qvx - pxa * qVx);
                    tb_out(a, g.off[MSFDM_SXVZ] + ai, -pxc * rinv * qvz - pxa * qVz);
                    Vxs = S[g.off[MSFDM_SXVX] + ai]; Vzs = S[g.off[MSFDM_SXVZ] + ai];
                }
                if (bottom) {
                    const float pza = ldc(&c.pza[iz]), pzb = ldc(&c.pzb[iz]);
                    pzc = ldc(&c.pzc[iz]); q1z = pzc - 1.0f;
                    const long long ai = (long long)(iz - zb0) * P + ix;
                    const float qVx = kbar(a, g.off[MSFDM_SZVX] + ai), qVz = kbar(a, g.off[MSFDM_SZVZ] + ai);
                    Dzxb = q1z * rinv * qvx + pzb * qVx;
                    Dzzb = q1z * rinv * qvz + pzb * qVz;
                    tb_out(a, g.off[MSFDM_SZVX] + ai, -pzc * rinv * qvx - pza * qVx);
                    tb_out(a, g.off[MSFDM_SZVZ] + ai, -pzc * rinv * qvz - pza * qVz);
                    Vxb = S[g.off[MSFDM_SZVX] + ai]; Vzb = S[g.off[MSFDM_SZVZ] + ai];
                }
                float rvx, rvz, Dxx = 0.f, Dxz = 0.f, Dzx = 0.f, Dzz = 0.f;
                if (!INT && iz < 3) {
                    // ghost-flux rows 0..2: rho v_k = (D_xi(J sigma xi) + D_ze(J sigma ze)) / J
                    auto fze_row = [&](int rr, float &fx, float &fz) {
                        const long long qq = moff + (long long)rr * P + ix;
                        const int tr = r + (rr - iz);
                        const float Jq = ldc(&c.J[qq]), zxq = ldc(&c.zex[qq]), zzq = ldc(&c.zez[qq]);
                        fx = Jq * (zxq * ssx[tr][sc] + zzq * ssxz[tr][sc]);
                        fz = Jq * (zxq * ssxz[tr][sc] + zzq * ssz[tr][sc]);
                    };
                    float sx = 0.f, sz = 0.f;
#pragma unroll
                    for (int k = 0; k < 5; ++k) {
                        const int m = o + k;
                        const float w = drp_w<DIR>(k);
                        const long long gq = gi + m;
                        const float Jq = ldc(&c.J[gq]);
                        float xxq, xzq;
                        xi_metrics<MET>(c, gq, xxq, xzq);
                        sx += w * Jq * (xxq * ssx[r][sc + m] + xzq * ssxz[r][sc + m]);
                        sz += w * Jq * (xxq * ssxz[r][sc + m] + xzq * ssz[r][sc + m]);
                        const int rr = iz + m;
                        float fbx, fbz;
                        if (rr >= 0) fze_row(rr, fbx, fbz);
                        else { float f0x, f0z, fnx, fnz; fze_row(0, f0x, f0z); fze_row(-rr, fnx, fnz); fbx = 2.f * f0x - fnx; fbz = 2.f * f0z - fnz; }
                        sx += w * fbx;
                        sz += w * fbz;
                    }
                    rvx = sx / J; rvz = sz / J;
                    const float pxi = (xx * sx + xz * sz) / (gxixi * J), pze = (zx * sx + zz * sz) / (gzeze * J);
                    Dxx = xx * pxi; Dxz = xz * pxi; Dzx = zx * pze; Dzz = zz * pze;
                    const float pxib = xx * Dxxb + xz * Dxzb, pzeb = zx * Dzxb + zz * Dzzb;
                    q[4] = rvxb / J + xx * pxib / (gxixi * J) + zx * pzeb / (gzeze * J);   // sx_bar
                    q[5] = rvzb / J + xz * pxib / (gxixi * J) + zz * pzeb / (gzeze * J);   // sz_bar
                } else {
                    const float A = dx_lds<DIR>(ssx[r], sc), Bv = dz_lds<DIR>(ssx, r, sc);
                    const float Cv = dx_lds<DIR>(ssz[r], sc), Dv = dz_lds<DIR>(ssz, r, sc);
                    const float E = dx_lds<DIR>(ssxz[r], sc), Fv = dz_lds<DIR>(ssxz, r, sc);
                    rvx = xx * A + zx * Bv + xz * E + zz * Fv;
                    rvz = xx * E + zx * Fv + xz * Cv + zz * Dv;
                    if (INT) {
                        q[4] = xx * rvxb;                   // sxx along x
                        q[5] = xz * rvzb;                   // szz along x
                        q[6] = xz * rvxb + xx * rvzb;       // sxz along x
                        q[7] = zx * rvxb;                   // sxx along z
                        q[8] = zz * rvzb;                   // szz along z
                        q[9] = zz * rvxb + zx * rvzb;       // sxz along z
                    } else {
                        const float gx = gxize / gxixi, gz = gxize / gzeze;
                        Dxx = xx * (A + gx * Bv) + xz * (E + gx * Fv);
                        Dxz = xx * (E + gx * Fv) + xz * (Cv + gx * Dv);
                        Dzx = zx * (Bv + gz * A) + zz * (gz * E + Fv);
                        Dzz = zx * (gz * E + Fv) + zz * (gz * Cv + Dv);
                        q[4] = xx * rvxb + xx * Dxxb + zx * gz * Dzxb;
                        q[5] = xz * rvzb + xz * Dxzb + zz * gz * Dzzb;
                        q[6] = xz * rvxb + xx * rvzb + xz * Dxxb + xx * Dxzb + zz * gz * Dzxb + zx * gz * Dzzb;
                        q[7] = zx * rvxb + xx * gx * Dxxb + zx * Dzxb;
                        q[8] = zz * rvzb + xz * gx * Dxzb + zz * Dzzb;
                        q[9] = zz * rvxb + zx * rvzb + xz * gx * Dxxb + xx * gx * Dxzb + zz * Dzxb + zx * Dzzb;
                    }
                }
                float kvx = rvx, kvz = rvz;
                if (side) { kvx += q1x * Dxx - pxc * Vxs; kvz += q1x * Dxz - pxc * Vzs; }
                if (bottom) { kvx += q1z * Dzx - pzc * Vxb; kvz += q1z * Dzz - pzc * Vzb; }
                gr.rho_s[si] = grs + -(double)(qvx * kvx + qvz * kvz) * (double)rinv * (double)rinv;
            }
        }
#pragma unroll
        for (int k = 0; k < ADJ_NQS; ++k) Q[k * pP + si] = q[k];
    }
}

// Products over the source tiles of a list: tiles[t] = (column tile, band),
// bands 0..nbz_f-1 fluid, then solid.  Three lists and three kernels, so the
// registers and code of one kind of tile do not weigh on another:
// ADJ_INTERIOR, the interior tiles of both media (adj_fluid_interior /
// adj_solid_interior; bodies without a boundary branch); ADJ_FLUID and
// ADJ_SOLID, the other tiles of each medium.  ADJ_ALL runs every tile in one
// launch from a list in eight equal parts, one per XCD, each with its share of
// the solid and fluid boundary tiles first and then of the interior tiles; the
// band of an entry carries its kind (band + kind * ADJ_KIND_SHIFT) and entries
// with a negative column pad the parts.  On a grid of a few hundred tiles each
// split launch is shorter than one wave of blocks and lasts as long as its
// slowest tile, so one launch that overlaps them beats three.
constexpr int ADJ_INTERIOR = 0, ADJ_FLUID = 1, ADJ_SOLID = 2, ADJ_ALL = 3;
constexpr int ADJ_KIND_SHIFT = 1 << 24;
template <int DIR, int TZ, int MET, int KIND>
__global__ void __launch_bounds__(NTHR)
adj_prod_kernel(Geo g, Coef c, AdjIn a0, Grad gr0, const int2 *__restrict__ tiles, int ntiles, int nbz_f) {
    __shared__ float sh[5][TZ + 2 * HALO][LXP];   // forward stage state of the tile + halo
    const AdjIn a = adj_of_shot(a0, blockIdx.y);
    const Grad gr = grad_of_shot(gr0, blockIdx.y);
    const int per_xcd = gridDim.x >> 3;
    const int t = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (t >= ntiles) return;
    const int2 tl = tiles[t];
    if (KIND == ADJ_ALL && tl.x < 0) return;   // padding of an XCD's part
    const int bx = tl.x;
    const int band = KIND == ADJ_ALL ? tl.y % ADJ_KIND_SHIFT : tl.y;
    const int kind = KIND == ADJ_ALL ? tl.y / ADJ_KIND_SHIFT : KIND;
    if (kind == ADJ_INTERIOR) {
        if (band < nbz_f) {
            load_tile<3, TZ, true>(sh, a.S, g.off, MSFDM_FVX, 3 + band * TZ, bx * TX, g.nzf, g.P);
            __syncthreads();
            fluid_prod_body<DIR, TZ, MET, 1>(sh, g, c, a, gr, bx, band);
        } else {
            const int sb = band - nbz_f;
            load_tile<5, TZ, true>(sh, a.S, g.off, MSFDM_SVX, sb * TZ, bx * TX, g.nzs, g.P);
            __syncthreads();
            solid_prod_body<DIR, TZ, MET, 1>(sh, g, c, a, gr, bx, sb);
        }
    } else if (KIND == ADJ_FLUID || (KIND == ADJ_ALL && kind == ADJ_FLUID)) {
        load_tile<3, TZ, false>(sh, a.S, g.off, MSFDM_FVX, 3 + band * TZ, bx * TX, g.nzf, g.P);
        __syncthreads();
        fluid_prod_body<DIR, TZ, MET, 0>(sh, g, c, a, gr, bx, band);
    } else if (KIND == ADJ_SOLID || KIND == ADJ_ALL) {
        const int sb = band - nbz_f;
        load_tile<5, TZ, false>(sh, a.S, g.off, MSFDM_SVX, sb * TZ, bx * TX, g.nzs, g.P);
        __syncthreads();
        solid_prod_body<DIR, TZ, MET, 0>(sh, g, c, a, gr, bx, sb);
    }
}

// ---------------------------------------------------------------------------
// Gather: Sbar at the output cells of a 64 x TZ tile from the products of the
// tile and its 3-cell halo (LDS).  A source row's stencil along x is always
// the DRP (the ghost-flux rows weigh it with the output cell's J xi); along z
// it depends on the source row: DRP, a closure, or a ghost-flux row whose
// taps beyond the boundary reflect into rows inside.
// ---------------------------------------------------------------------------
// Weight of ghost-flux source row s at output row t: the taps rr = s + m of
// its DRP stencil, reflected as the forward ghosts are (interface: 2 F(nzf-1)
// - F(2 (nzf-1) - rr); fluid surface: -F(6 - rr); solid below the interface:
// 2 F(0) - F(-rr)).
constexpr int GH_IF = 0, GH_SURF = 1, GH_SOLID = 2;
template <int DIR, int GH>
__device__ __forceinline__ float ghost_at(int s, int t, int nzf) {
    float w = 0.0f;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int rr = s + lo_off<DIR>() + k;
        const float wk = drp_w<DIR>(k);
        if constexpr (GH == GH_IF) {
            if (rr <= nzf - 1) w += rr == t ? wk : 0.0f;
            else w += (t == nzf - 1 ? 2.0f * wk : 0.0f) - (t == 2 * (nzf - 1) - rr ? wk : 0.0f);
        } else if constexpr (GH == GH_SURF) {
            if (rr >= 3) w += rr == t ? wk : 0.0f;
            else w -= t == 6 - rr ? wk : 0.0f;
        } else {
            if (rr >= 0) w += rr == t ? wk : 0.0f;
            else w += (t == 0 ? 2.0f * wk : 0.0f) - (t == -rr ? wk : 0.0f);
        }
    }
    return w;
}

// sum over k of w_k * v[c - o - k]: the x-direction transpose of a DRP row.
template <int DIR>
__device__ __forceinline__ float xsum(const float *row, int c) {
    constexpr int o = lo_off<DIR>();
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 5; ++k) s += drp_w<DIR>(k) * row[c - o - k];
    return s;
}
// sum over k of w_k * v[r - o - k][c]: the z-direction transpose of DRP rows.
template <int DIR>
__device__ __forceinline__ float zsum(const float (*t)[LXP], int r, int c) {
    constexpr int o = lo_off<DIR>();
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 5; ++k) s += drp_w<DIR>(k) * t[r - o - k][c];
    return s;
}

// Fluid output rows 0..nzf-1 of one tile (z0 = bz * TZ).
template <int DIR, int TZ, int MET>
__device__ __forceinline__ void fluid_gather_body(const float (*sh)[TZ + 2 * HALO][LXP], const Geo &g, const Coef &c,
                                                  const AdjIn &a, int bx, int bz) {
    const int tx = threadIdx.x & (TX - 1), ty = __builtin_amdgcn_readfirstlane(threadIdx.x / TX);
    constexpr int RW = TZ / 4;
    const int x0 = bx * TX, z0 = bz * TZ;
    const int P = g.P, nzf = g.nzf;
    const int sc = tx + XOFF, ix = x0 + tx;
    auto flux_row = [&](int s) { return s >= 3 && s <= nzf - 1 && (nzf - 1 - s < 3 || s - 3 < 3); };
    // the accumulator entries of the wave's cells, loaded before any store
    float acc[RW][3];
#pragma unroll
    for (int j = 0; j < RW; ++j) {
        const int t = z0 + ty * RW + j;
        const long long gi = (long long)t * P + ix;
#pragma unroll
        for (int f = 0; f < 3; ++f) acc[j][f] = (a.ACC && t <= nzf - 1) ? a.ACCin[g.off[MSFDM_FVX + f] + gi] : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < RW; ++j) {
        const int t = z0 + ty * RW + j;
        if (t > nzf - 1) break;   // uniform over the wave
        const int r = t - z0 + HALO;
        const long long gi = (long long)t * P + ix;
        float bp, bvx, bvz;
        if (t >= 9 && t <= nzf - 7) {
            // every source row within 3 rows is a plain DRP row
            bp = xsum<DIR>(sh[0][r], sc) + zsum<DIR>(sh[1], r, sc);
            bvx = xsum<DIR>(sh[2][r], sc) + zsum<DIR>(sh[4], r, sc);
            bvz = xsum<DIR>(sh[3][r], sc) + zsum<DIR>(sh[5], r, sc);
        } else {
            // x: the source row is this row
            if (flux_row(t)) {
                float xx, xz;
                xi_metrics<MET>(c, gi, xx, xz);
                bp = -ldc(&c.J[gi]) * (xx * xsum<DIR>(sh[0][r], sc) + xz * xsum<DIR>(sh[1][r], sc));
            } else {
                bp = xsum<DIR>(sh[0][r], sc);
            }
            bvx = xsum<DIR>(sh[2][r], sc);
            bvz = xsum<DIR>(sh[3][r], sc);
            // z: every source row s within 3 rows, with its own stencil
            float gx = 0.0f, gz = 0.0f;
            for (int s = max(t - 3, 3); s <= min(t + 3, nzf - 1); ++s) {
                const int rs = s - z0 + HALO;
                if (flux_row(s)) {
                    const float w = nzf - 1 - s < 3 ? ghost_at<DIR, GH_IF>(s, t, nzf) : ghost_at<DIR, GH_SURF>(s, t, nzf);
                    gx += w * sh[0][rs][sc];
                    gz += w * sh[1][rs][sc];
                } else {
                    bp += drp_at<DIR>(t - s) * sh[1][rs][sc];
                }
                if (s > 3) {   // pressure equation (none on the surface row 3)
                    const float w = mac_at<DIR>(min(s - 3, nzf - 1 - s), false, t - s);
                    bvx += w * sh[4][rs][sc];
                    bvz += w * sh[5][rs][sc];
                }
            }
            bp += -ldc(&c.J[gi]) * (ldc(&c.zex[gi]) * gx + ldc(&c.zez[gi]) * gz);
        }
        if (a.TB) {
            a.TB[g.off[MSFDM_FVX] + gi] = bvx;
            a.TB[g.off[MSFDM_FVZ] + gi] = bvz;
            a.TB[g.off[MSFDM_FP] + gi] = bp;
        }
        if (a.ACC) {
            a.ACC[g.off[MSFDM_FVX] + gi] = acc[j][0] + bvx;
            a.ACC[g.off[MSFDM_FVZ] + gi] = acc[j][1] + bvz;
            a.ACC[g.off[MSFDM_FP] + gi] = acc[j][2] + bp;
        }
    }
}

// Solid output rows 0..nzs-1 of one tile (z0 = bz * TZ).
template <int DIR, int TZ, int MET>
__device__ __forceinline__ void solid_gather_body(const float (*sh)[TZ + 2 * HALO][LXP], const Geo &g, const Coef &c,
                                                  const AdjIn &a, int bx, int bz) {
    const int tx = threadIdx.x & (TX - 1), ty = __builtin_amdgcn_readfirstlane(threadIdx.x / TX);
    constexpr int RW = TZ / 4;
    const int x0 = bx * TX, z0 = bz * TZ;
    const int P = g.P, nzs = g.nzs;
    const int sc = tx + XOFF, ix = x0 + tx;
    const long long moff = (long long)(g.nzf - 1) * P;
    // the accumulator entries of the wave's cells, loaded before any store
    float acc[RW][5];
#pragma unroll
    for (int j = 0; j < RW; ++j) {
        const int t = z0 + ty * RW + j;
        const long long si = (long long)t * P + ix;
#pragma unroll
        for (int f = 0; f < 5; ++f) acc[j][f] = (a.ACC && t <= nzs - 1) ? a.ACCin[g.off[MSFDM_SVX + f] + si] : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < RW; ++j) {
        const int t = z0 + ty * RW + j;
        if (t > nzs - 1) break;   // uniform over the wave
        const int r = t - z0 + HALO;
        const long long si = (long long)t * P + ix;
        float bvx, bvz, bsx, bsz, bsxz;
        if (t >= 6) {
            // every source row within 3 rows is a plain DRP row
            bvx = xsum<DIR>(sh[0][r], sc) + zsum<DIR>(sh[2], r, sc);
            bvz = xsum<DIR>(sh[1][r], sc) + zsum<DIR>(sh[3], r, sc);
            bsx = xsum<DIR>(sh[4][r], sc) + zsum<DIR>(sh[7], r, sc);
            bsz = xsum<DIR>(sh[5][r], sc) + zsum<DIR>(sh[8], r, sc);
            bsxz = xsum<DIR>(sh[6][r], sc) + zsum<DIR>(sh[9], r, sc);
        } else {
            bvx = xsum<DIR>(sh[0][r], sc);
            bvz = xsum<DIR>(sh[1][r], sc);
            const long long gi = moff + si;
            const float Jt = ldc(&c.J[gi]), zxt = ldc(&c.zex[gi]), zzt = ldc(&c.zez[gi]);
            if (t <= 2) {   // ghost-flux row: x taps weighed with this cell's J xi
                float xx, xz;
                xi_metrics<MET>(c, gi, xx, xz);
                const float fx = xsum<DIR>(sh[4][r], sc), fz = xsum<DIR>(sh[5][r], sc);
                bsx = Jt * xx * fx;
                bsz = Jt * xz * fz;
                bsxz = Jt * (xz * fx + xx * fz);
            } else {
                bsx = xsum<DIR>(sh[4][r], sc);
                bsz = xsum<DIR>(sh[5][r], sc);
                bsxz = xsum<DIR>(sh[6][r], sc);
            }
            float fx = 0.0f, fz = 0.0f;
            for (int s = max(t - 3, 0); s <= min(t + 3, nzs - 4); ++s) {
                const int rs = s - z0 + HALO;
                const float wm = mac_at<DIR>(s, true, t - s);   // stress equations: mac_dz at distance s
                bvx += wm * sh[2][rs][sc];
                bvz += wm * sh[3][rs][sc];
                if (s <= 2) {   // ghost-flux velocity rows
                    const float w = ghost_at<DIR, GH_SOLID>(s, t, 0);
                    fx += w * sh[4][rs][sc];
                    fz += w * sh[5][rs][sc];
                } else {
                    const float w = drp_at<DIR>(t - s);
                    bsx += w * sh[7][rs][sc];
                    bsz += w * sh[8][rs][sc];
                    bsxz += w * sh[9][rs][sc];
                }
            }
            bsx += Jt * zxt * fx;
            bsz += Jt * zzt * fz;
            bsxz += Jt * (zzt * fx + zxt * fz);
        }
        const float out[5] = {bvx, bvz, bsx, bsz, bsxz};
#pragma unroll
        for (int f = 0; f < 5; ++f)
            if (a.TB) a.TB[g.off[MSFDM_SVX + f] + si] = out[f];
        if (a.ACC) {
#pragma unroll
            for (int f = 0; f < 5; ++f) a.ACC[g.off[MSFDM_SVX + f] + si] = acc[j][f] + out[f];
        }
    }
}

// Output tiles: P / 64 columns x (fluid bands of rows 0..nzf-1, then solid
// bands of rows 0..nzs-1).
template <int DIR, int TZ, int MET>
__global__ void __launch_bounds__(NTHR)
adj_gather_kernel(Geo g, Coef c, AdjIn a0, int nbx, int nbz_f, int nbz_s) {
    __shared__ float sh[ADJ_NQS][TZ + 2 * HALO][LXP];
    const AdjIn a = adj_of_shot(a0, blockIdx.y);
    const int per_xcd = gridDim.x >> 3;
    const int t = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (t >= nbx * (nbz_f + nbz_s)) return;
    const int band = t / nbx, bx = t - band * nbx;
    long long poff[ADJ_NQS];
    if (band < nbz_f) {
#pragma unroll
        for (int k = 0; k < ADJ_NQF; ++k) poff[k] = prod_fluid(g, k);
        load_tile<ADJ_NQF, TZ>(sh, a.prod, poff, 0, band * TZ, bx * TX, g.nzf, g.P);
        __syncthreads();
        fluid_gather_body<DIR, TZ, MET>(sh, g, c, a, bx, band);
    } else {
#pragma unroll
        for (int k = 0; k < ADJ_NQS; ++k) poff[k] = prod_solid(g, k);
        load_tile<ADJ_NQS, TZ>(sh, a.prod, poff, 0, (band - nbz_f) * TZ, bx * TX, g.nzs, g.P);
        __syncthreads();
        solid_gather_body<DIR, TZ, MET>(sh, g, c, a, bx, band - nbz_f);
    }
}

// Compensated (Kahan) FP64 accumulation of segment partial sums: total += part
// with the running compensation comp, then part = 0.  Over a gradient of N
// steps the partial sums of one checkpoint segment (about sqrt(N) steps, four
// stages each) are plain FP64; the totals over the segments carry the
// compensation, so their rounding does not grow with the number of segments.
__global__ void kahan_add_kernel(double *__restrict__ total, double *__restrict__ comp, double *__restrict__ part,
                                 long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double y = part[i] - comp[i];
        const double t = total[i] + y;
        comp[i] = (t - total[i]) - y;
        total[i] = t;
        part[i] = 0.0;
    }
}

// Filter map of one column (the post-filter values from the 8 filtered inputs).
__device__ __forceinline__ void filter_map(float2 n, const float v[8], float o[8]) {
    const float n1 = n.x, n2 = n.y;
    const float vnf = n1 * v[0] + n2 * v[1], vtf = n2 * v[0] - n1 * v[1];
    const float vns = n1 * v[3] + n2 * v[4], vts = n2 * v[3] - n1 * v[4];
    const float snn = n1 * n1 * v[5] + 2.0f * n1 * n2 * v[7] + n2 * n2 * v[6];
    const float stt = n2 * n2 * v[5] - 2.0f * n1 * n2 * v[7] + n1 * n1 * v[6];
    const float vn = 0.5f * (vnf + vns), sn = 0.5f * (snn - v[2]);
    o[0] = n1 * vn + n2 * vtf; o[1] = n2 * vn - n1 * vtf; o[2] = -sn;
    o[3] = n1 * vn + n2 * vts; o[4] = n2 * vn - n1 * vts;
    o[5] = n1 * n1 * sn + n2 * n2 * stt; o[6] = n2 * n2 * sn + n1 * n1 * stt; o[7] = n1 * n2 * (sn - stt);
}

// LP interface rows (columns 3..nx-4) = Filter^T(LAM interface rows).
// The eight interface-row fields of src copied to the same entries of dst
// (so that adj_filter can read them from dst while it writes src's).
__global__ void __launch_bounds__(256) iface_rows_copy(Geo g, const float *__restrict__ src, float *__restrict__ dst,
                                                       long long bstride) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    if (ix >= g.nx) return;
    src += (long long)blockIdx.y * bstride;
    dst += (long long)blockIdx.y * bstride;
#pragma unroll
    for (int f = 0; f < 8; ++f) dst[iface_index(g, f, ix)] = src[iface_index(g, f, ix)];
}

__global__ void __launch_bounds__(FB) adj_filter(Geo g, Coef c, const float *__restrict__ LAM, float *__restrict__ LP,
                                                  long long bstride) {
    __shared__ float vb[8][FB + 2 * HALO];
    __shared__ float ws[FB + 2 * HALO];
    LAM += (long long)blockIdx.y * bstride;
    LP += (long long)blockIdx.y * bstride;
    const int nx = g.nx, x0 = blockIdx.x * FB;
    for (int t = threadIdx.x; t < FB + 2 * HALO; t += FB) {
        const int ix = x0 - HALO + t;
        float out[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        float wsum = 1.0f;
        if (ix >= 3 && ix <= nx - 4) {
            float yb[8];
#pragma unroll
            for (int f = 0; f < 8; ++f) yb[f] = LAM[iface_index(g, f, ix)];
            const float2 n = c.nif[ix];
#pragma unroll
            for (int j = 0; j < 8; ++j) {   // column j of the map, dotted with yb
                float v[8], o[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) v[i] = i == j ? 1.0f : 0.0f;
                filter_map(n, v, o);
                float s = 0.0f;
#pragma unroll
                for (int i = 0; i < 8; ++i) s += o[i] * yb[i];
                out[j] = s;
            }
            const int klo = max(-3, 3 - ix), khi = min(3, nx - 4 - ix);
            wsum = 0.0f;
            for (int kk = klo; kk <= khi; ++kk) wsum += FILT_W[kk + 3];
        }
#pragma unroll
        for (int f = 0; f < 8; ++f) vb[f][t] = out[f];
        ws[t] = wsum;
    }
    __syncthreads();
    const int jx = x0 + threadIdx.x;
    if (jx < 3 || jx > nx - 4) return;
    const int t = threadIdx.x + HALO;
#pragma unroll
    for (int f = 0; f < 8; ++f) {
        float s = 0.0f;
        for (int kk = -3; kk <= 3; ++kk) {
            const int ix = jx - kk;   // contributing column, tap offset kk = jx - ix
            if (ix < 3 || ix > nx - 4) continue;
            s += FILT_W[kk + 3] / ws[t - kk] * vb[f][t - kk];
        }
        LP[iface_index(g, f, jx)] = s;
    }
}

// total += sum over nshot per-shot gradient arrays of n doubles each (stride
// gstride), added in shot order.
__global__ void grad_sum_kernel(double *__restrict__ total, const double *__restrict__ g, long long n, long long gstride,
                                int nshot) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        double t = total[i];
        for (int k = 0; k < nshot; ++k) t += g[(long long)k * gstride + i];
        total[i] = t;
    }
}

// Record residual injection into lambda at record slot r (shot blockIdx.y:
// lambda at bstride floats and residuals [3][nrec][nx] at rstride floats per shot).
__global__ void adj_inject(Geo g, float *__restrict__ LAM, const float *__restrict__ res, int nrec, int r,
                           long long bstride, long long rstride) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    if (ix >= g.nx) return;
    LAM += (long long)blockIdx.y * bstride;
    res += (long long)blockIdx.y * rstride;
    const long long plane = (long long)nrec * g.nx, o = (long long)r * g.nx + ix;
    LAM[g.off[MSFDM_SVX] + ix] += res[o];
    LAM[g.off[MSFDM_SVZ] + ix] += res[plane + o];
    LAM[g.off[MSFDM_FP] + (long long)(g.nzf - 1) * g.P + ix] += res[2 * plane + o];
}

}  // namespace dev
}  // namespace msfdm
