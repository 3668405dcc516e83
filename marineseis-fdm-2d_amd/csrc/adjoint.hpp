// Device code of the discrete adjoint (reverse-time) stepper.
//
// The forward stage is k = B_dir(L_dir(S) + src) (interface BC B, RHS L of
// include/pml_fs_kernels.cuh:333-827).  One adjoint stage takes
// kbar = clam * LP + cprev * TBprev (the RK4 stage weights), applies B^T on the
// interface rows (adj_iface_q) and scatters L^T(qbar) into TB:
//   adj_stage_kernel  per 64 x TZ tile, every forward expression
//                     k[i] = sum_j a_ij S[j] is run in reverse,
//                     Sbar[j] += a_ij qbar[i]; stencil contributions are
//                     gathered in an LDS accumulator covering the tile and its
//                     3-cell halo (ghost-flux reflections included) and added
//                     to the output (one LDS copy per wave, x-direction taps
//                     gathered across lanes: no LDS atomics).  The tiles run in
//                     four launches, one per colour (tile column parity x band
//                     parity): the accumulators of one colour never overlap, so
//                     the flush is a plain read-modify-write and the colours
//                     add in a fixed order -- repeated runs give the same bits.
//                     PML memory terms and the model-gradient terms dk/dm . qbar
//                     (evaluated on the forward stage state S) are pointwise.
//                     Gradients accumulate in FP64.
//   adj_filter        Filter^T on the interface rows (out of place, LDS).
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

// Adjoint of mac_dz: add(offset, weight*g) for the distance-dependent zeta operator.
template <int DIR, class Add>
__device__ __forceinline__ void mac_dz_adj(int dist, bool solid_side, float g, Add add) {
    if (dist >= 3) {
        constexpr int o = lo_off<DIR>();
#pragma unroll
        for (int q = 0; q < 5; ++q) add(o + q, drp_w<DIR>(q) * g);
        return;
    }
    if (dist == 2) {
        if constexpr (DIR == 0) { add(0, M42_0 * g); add(1, M42_1 * g); add(2, M42_2 * g); }
        else { add(0, -M42_0 * g); add(-1, -M42_1 * g); add(-2, -M42_2 * g); }
        return;
    }
    bool fwd = DIR == 0;
    if (dist == 0) fwd = solid_side;
    if (fwd) { add(0, -g); add(1, g); }
    else { add(0, g); add(-1, -g); }
}

// Forward one-sided derivative reading global memory (stride 1 or P).
template <int DIR>
__device__ __forceinline__ float drp_g(const float *f, long long i, long long stride) {
    constexpr int o = lo_off<DIR>();
    return drp5<DIR>(f[i + o * stride], f[i + (o + 1) * stride], f[i + (o + 2) * stride], f[i + (o + 3) * stride],
                     f[i + (o + 4) * stride]);
}
template <int DIR>
__device__ __forceinline__ float mac_g(const float *f, long long i, long long P, int dist, bool solid_side) {
    if (dist >= 3) return drp_g<DIR>(f, i, P);
    if (dist == 2) {
        if constexpr (DIR == 0) return M42_0 * f[i] + M42_1 * f[i + P] + M42_2 * f[i + 2 * P];
        else return -(M42_0 * f[i] + M42_1 * f[i - P] + M42_2 * f[i - 2 * P]);
    }
    bool fwd = DIR == 0;
    if (dist == 0) fwd = solid_side;
    return fwd ? (-f[i] + f[i + P]) : -(-f[i] + f[i - P]);
}

// LDS accumulator of one wave: 5 fields x (its RW rows + 3-row halos) x the
// tile's 70 columns (64 + 3-column halos).
template <int TZ>
struct WaveAcc {
    float v[5][TZ / 4 + 2 * HALO][LXP];
};

struct AdjIn {
    const float *S;      // forward stage state
    const float *LP;     // lambda'
    const float *TBp;    // previous stage's Tbar (may be null when cprev == 0)
    const float *qif;    // [8][P] interface qbar = B^T kbar
    float *TB;           // output Tbar (zeroed)
    float clam, cprev;
    long long bstride;   // floats between the state buffers of consecutive shots
    long long qstride;   // floats between the qif buffers of consecutive shots
};

// Inputs of shot `shot` of a multi-shot launch (blockIdx.y); every shot's
// buffers sit at a fixed stride from shot 0's.
__device__ __forceinline__ AdjIn adj_of_shot(AdjIn a, int shot) {
    const long long o = (long long)shot * a.bstride;
    a.S += o; a.LP += o; a.TB += o;
    if (a.TBp) a.TBp += o;
    a.qif += (long long)shot * a.qstride;
    return a;
}

__device__ __forceinline__ float kbar(const AdjIn &a, long long i) {
    float v = a.clam * a.LP[i];
    if (a.TBp) v += a.cprev * a.TBp[i];
    return v;
}

// x-direction scatter of one field on one row: tap q of lane t lands on tile
// column t + o + q (o = lo_off<DIR>).  It runs as a gather over the wave: lane u
// sums the taps landing on column u and lanes 0..3 the taps landing on the
// four halo columns (F: -1, 64, 65, 66; B: -3, -2, -1, 64), so every
// accumulator cell has a single writing lane and the read-modify-writes need
// no ordering.  v[] is zero on lanes that contribute nothing; all 64 lanes
// must reach the call (the wave's row is uniform).
template <int DIR>
__device__ __forceinline__ void xgather(float (*row)[LXP], int lr, int tx, const float v[5]) {
    constexpr int o = lo_off<DIR>();
    const int hc = DIR == 0 ? (tx == 0 ? -1 : TX - 1 + tx) : (tx < 3 ? tx - 3 : TX);
    float s = 0.0f, sh = 0.0f;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const int src = tx - o - q, hsrc = hc - o - q;
        const float x = __shfl(v[q], src & (TX - 1));
        const float y = __shfl(v[q], hsrc & (TX - 1));
        s += (src >= 0 && src < TX) ? x : 0.0f;
        sh += (hsrc >= 0 && hsrc < TX) ? y : 0.0f;
    }
    row[lr][tx + HALO] += s;
    if (tx < 4) row[lr][hc + HALO] += sh;
}

// ---------------------------------------------------------------------------
// Fluid tile: rows 3..nzf-1.  acc[0..2] = vx, vz, p adjoint accumulators.
// ---------------------------------------------------------------------------
template <int DIR, int TZ>
__device__ __forceinline__ void fluid_adj_body(WaveAcc<TZ> *acc, const Geo &g, const Coef &c,
                                               const AdjIn &a, const Grad &gr, int bx, int bz) {
    const int tx = threadIdx.x & (TX - 1), ty = threadIdx.x / TX;
    constexpr int RW = TZ / 4;   // consecutive tile rows per wave
    const int x0 = bx * TX, z0 = 3 + bz * TZ;
    const int P = g.P, nx = g.nx, nzf = g.nzf;
    const float *S = a.S;
    const float *Sp = S + g.off[MSFDM_FP], *Svx = S + g.off[MSFDM_FVX], *Svz = S + g.off[MSFDM_FVZ];
    const int ix = x0 + tx;
    const bool active_col = ix >= 3 && ix <= nx - 4;
    const int scol = active_col ? side_col(ix, nx, g.npd) : -1;
    const int cc = tx + HALO;
    // this wave's accumulator; its rows start at tile row ty * RW - HALO
    float (*wa)[RW + 2 * HALO][LXP] = acc[ty].v;
    // Every accumulator cell has one writing lane: single-cell updates go to the
    // lane's own column (z-direction taps, ghost reflections), the x-direction
    // taps go through xgather.  Lanes outside columns 3..nx-4 run the row with
    // in-bounds loads but contribute nothing.
    const bool act = active_col;
    auto add = [&](int fld, int iz_to, float v) {
        if (act) wa[fld][iz_to - z0 - ty * RW + HALO][tx + HALO] += v;
    };
    auto addx = [&](int fld, int iz_to, const float (&v)[5]) {
        float w[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) w[q] = act ? v[q] : 0.0f;
        xgather<DIR>(wa[fld], iz_to - z0 - ty * RW + HALO, tx, w);
    };
    for (int j = 0; j < RW; ++j) {
        const int iz = z0 + ty * RW + j;
        if (iz > nzf - 1) continue;   // uniform over the wave
        const long long gi = (long long)iz * P + ix;
        const float J = c.J[gi], zx = c.zex[gi], zz = c.zez[gi];
        float xx, xz;
        xi_metrics<2>(c, gi, xx, xz);
        const float gxixi = xx * xx + xz * xz;
        const int dist_flso = nzf - 1 - iz, dist_fs = iz - 3;
        const bool at_row = dist_flso == 0;
        const float qvx = at_row ? a.qif[0 * (long long)P + ix] : kbar(a, g.off[MSFDM_FVX] + gi);
        const float qvz = at_row ? a.qif[1 * (long long)P + ix] : kbar(a, g.off[MSFDM_FVZ] + gi);
        const float qp = at_row ? a.qif[2 * (long long)P + ix] : kbar(a, g.off[MSFDM_FP] + gi);
        const float rinv = c.rhof_inv[gi];
        // ---- velocity equations
        float Dxb = 0.f, Dzb = 0.f;
        const float rvxb = qvx * rinv, rvzb = qvz * rinv;
        float pxa = 0.f, pxb = 0.f, pxc = 1.f;
        long long ai = 0;
        float Vx = 0.f, Vz = 0.f, Ex = 0.f, Ez = 0.f;
        if (scol >= 0) {
            pxa = c.pxa[ix]; pxb = c.pxb[ix]; pxc = c.pxc[ix];
            ai = (long long)iz * g.W2 + scol;
            const float qVx = kbar(a, g.off[MSFDM_FXVX] + ai), qVz = kbar(a, g.off[MSFDM_FXVZ] + ai);
            Dxb = (pxc - 1.0f) * rinv * qvx + pxb * qVx;
            Dzb = (pxc - 1.0f) * rinv * qvz + pxb * qVz;
            a.TB[g.off[MSFDM_FXVX] + ai] += -pxc * rinv * qvx - pxa * qVx;
            a.TB[g.off[MSFDM_FXVZ] + ai] += -pxc * rinv * qvz - pxa * qVz;
            Vx = S[g.off[MSFDM_FXVX] + ai]; Vz = S[g.off[MSFDM_FXVZ] + ai];
            Ex = S[g.off[MSFDM_FXEX] + ai]; Ez = S[g.off[MSFDM_FXEZ] + ai];
        }
        float rvx, rvz, Dx, Dz;   // forward values for the gradient
        if (dist_flso < 3 || dist_fs < 3) {
            const bool at_if = dist_flso < 3;
            auto fze_row = [&](int rr, float &fx, float &fz) {
                const long long q = (long long)rr * P + ix;
                const float mJp = -c.J[q] * Sp[q];
                fx = mJp * c.zex[q];
                fz = mJp * c.zez[q];
            };
            float sx = 0.f, sz = 0.f;
            constexpr int o = lo_off<DIR>();
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                const int m = o + q;
                const float w = drp_w<DIR>(q);
                const long long gq = gi + m;
                const float mJp = -c.J[gq] * Sp[gq];
                float xxq, xzq;
                xi_metrics<2>(c, gq, xxq, xzq);
                sx += w * mJp * xxq;
                sz += w * mJp * xzq;
                const int rr = iz + m;
                float bx, bz;
                if (at_if) {
                    if (rr <= nzf - 1) fze_row(rr, bx, bz);
                    else { float f0x, f0z, fnx, fnz; fze_row(nzf - 1, f0x, f0z); fze_row(2 * (nzf - 1) - rr, fnx, fnz); bx = 2.f * f0x - fnx; bz = 2.f * f0z - fnz; }
                } else {
                    if (rr >= 3) fze_row(rr, bx, bz);
                    else { float fnx, fnz; fze_row(6 - rr, fnx, fnz); bx = -fnx; bz = -fnz; }
                }
                sx += w * bx;
                sz += w * bz;
            }
            rvx = sx / J; rvz = sz / J;
            const float proj = (xx * sx + xz * sz) / (gxixi * J);
            Dx = xx * proj; Dz = xz * proj;
            const float projb = xx * Dxb + xz * Dzb;
            const float sxb = rvxb / J + xx * projb / (gxixi * J);
            const float szb = rvzb / J + xz * projb / (gxixi * J);
            auto fze_add = [&](int rr, float fbx, float fbz) {
                const long long q = (long long)rr * P + ix;
                add(2, rr, -c.J[q] * (c.zex[q] * fbx + c.zez[q] * fbz));
            };
            float xv[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                const int m = o + q;
                const float w = drp_w<DIR>(q);
                const long long gq = gi + m;
                float xxq, xzq;
                xi_metrics<2>(c, gq, xxq, xzq);
                xv[q] = -c.J[gq] * (xxq * w * sxb + xzq * w * szb);
                const int rr = iz + m;
                const float fbx = w * sxb, fbz = w * szb;
                if (at_if) {
                    if (rr <= nzf - 1) fze_add(rr, fbx, fbz);
                    else { fze_add(nzf - 1, 2.f * fbx, 2.f * fbz); fze_add(2 * (nzf - 1) - rr, -fbx, -fbz); }
                } else {
                    if (rr >= 3) fze_add(rr, fbx, fbz);
                    else fze_add(6 - rr, -fbx, -fbz);
                }
            }
            addx(2, iz, xv);
        } else {
            const float dpx = drp_g<DIR>(Sp, gi, 1), dpz = drp_g<DIR>(Sp, gi, P);
            const float gxize = xx * zx + xz * zz;
            rvx = -(xx * dpx + zx * dpz); rvz = -(xz * dpx + zz * dpz);
            const float t = (gxixi * dpx + gxize * dpz) / gxixi;
            Dx = -xx * t; Dz = -xz * t;
            const float tb = -xx * Dxb - xz * Dzb;
            const float dpxb = -xx * rvxb - xz * rvzb + tb;
            const float dpzb = -zx * rvxb - zz * rvzb + tb * gxize / gxixi;
            constexpr int o = lo_off<DIR>();
            float xv[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                xv[q] = drp_w<DIR>(q) * dpxb;
                add(2, iz + o + q, drp_w<DIR>(q) * dpzb);
            }
            addx(2, iz, xv);
        }
        float kvx = rvx, kvz = rvz;
        if (scol >= 0) { kvx += (pxc - 1.0f) * Dx - pxc * Vx; kvz += (pxc - 1.0f) * Dz - pxc * Vz; }
        if (act) gr.rho_f[gi] += -(double)(qvx * kvx + qvz * kvz) * (double)rinv * (double)rinv;
        // ---- pressure equation (row 3: p_k = 0, no dependence)
        if (dist_fs != 0) {
            const float Km = c.Kf[gi];
            const int dzn = min(dist_fs, dist_flso);
            const float dvxx = drp_g<DIR>(Svx, gi, 1), dvzx = drp_g<DIR>(Svz, gi, 1);
            const float dvxz = mac_g<DIR>(Svx, gi, P, dzn, false), dvzz = mac_g<DIR>(Svz, gi, P, dzn, false);
            float bracket = -(xx * dvxx + zx * dvxz + xz * dvzx + zz * dvzz);
            float bxx = -Km * xx * qp, bxz = -Km * zx * qp, bzx = -Km * xz * qp, bzz = -Km * zz * qp;
            if (scol >= 0) {
                const float gg = (xx * zx + xz * zz) / gxixi;
                const float qEx = kbar(a, g.off[MSFDM_FXEX] + ai), qEz = kbar(a, g.off[MSFDM_FXEZ] + ai);
                const float Dvxb = -Km * (pxc - 1.0f) * xx * qp + pxb * qEx;
                const float Dvzb = -Km * (pxc - 1.0f) * xz * qp + pxb * qEz;
                a.TB[g.off[MSFDM_FXEX] + ai] += Km * pxc * xx * qp - pxa * qEx;
                a.TB[g.off[MSFDM_FXEZ] + ai] += Km * pxc * xz * qp - pxa * qEz;
                bxx += Dvxb; bxz += gg * Dvxb; bzx += Dvzb; bzz += gg * Dvzb;
                const float Dvx = dvxx + gg * dvxz, Dvz = dvzx + gg * dvzz;
                bracket += -(pxc - 1.0f) * xx * Dvx - (pxc - 1.0f) * xz * Dvz + pxc * xx * Ex + pxc * xz * Ez;
            }
            float xa[5], xb[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                xa[q] = drp_w<DIR>(q) * bxx;
                xb[q] = drp_w<DIR>(q) * bzx;
            }
            addx(0, iz, xa);
            addx(1, iz, xb);
            mac_dz_adj<DIR>(dzn, false, bxz, [&](int m, float v) { add(0, iz + m, v); });
            mac_dz_adj<DIR>(dzn, false, bzz, [&](int m, float v) { add(1, iz + m, v); });
            if (act) gr.K[gi] += (double)qp * (double)bracket;
        }
    }
}

// ---------------------------------------------------------------------------
// Solid tile: rows 0..nzs-4.  acc[0..4] = vx, vz, sx, sz, sxz accumulators.
// ---------------------------------------------------------------------------
template <int DIR, int TZ>
__device__ __forceinline__ void solid_adj_body(WaveAcc<TZ> *acc, const Geo &g, const Coef &c,
                                               const AdjIn &a, const Grad &gr, int bx, int bz) {
    const int tx = threadIdx.x & (TX - 1), ty = threadIdx.x / TX;
    constexpr int RW = TZ / 4;   // consecutive tile rows per wave
    const int x0 = bx * TX, z0 = bz * TZ;
    const int P = g.P, nx = g.nx, nzs = g.nzs, npd = g.npd;
    const float *S = a.S;
    const float *Svx = S + g.off[MSFDM_SVX], *Svz = S + g.off[MSFDM_SVZ];
    const float *Ssx = S + g.off[MSFDM_SSX], *Ssz = S + g.off[MSFDM_SSZ], *Ssxz = S + g.off[MSFDM_SSXZ];
    const int ix = x0 + tx;
    const bool active_col = ix >= 3 && ix <= nx - 4;
    const int scol = active_col ? side_col(ix, nx, npd) : -1;
    const long long moff = (long long)(g.nzf - 1) * P;
    const int zb0 = nzs - 3 - npd;
    // this wave's accumulator; its rows start at tile row ty * RW - HALO
    float (*wa)[RW + 2 * HALO][LXP] = acc[ty].v;
    // Every accumulator cell has one writing lane: single-cell updates go to the
    // lane's own column (z-direction taps, ghost reflections), the x-direction
    // taps go through xgather.  Lanes outside columns 3..nx-4 run the row with
    // in-bounds loads but contribute nothing.
    const bool act = active_col;
    auto add = [&](int fld, int iz_to, float v) {
        if (act) wa[fld][iz_to - z0 - ty * RW + HALO][tx + HALO] += v;
    };
    auto addx = [&](int fld, int iz_to, const float (&v)[5]) {
        float w[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) w[q] = act ? v[q] : 0.0f;
        xgather<DIR>(wa[fld], iz_to - z0 - ty * RW + HALO, tx, w);
    };
    for (int j = 0; j < RW; ++j) {
        const int iz = z0 + ty * RW + j;
        if (iz > nzs - 4) continue;   // uniform over the wave
        const long long si = (long long)iz * P + ix, gi = moff + si;
        const float J = c.J[gi], zx = c.zex[gi], zz = c.zez[gi];
        float xx, xz;
        xi_metrics<2>(c, gi, xx, xz);
        const float gxixi = xx * xx + xz * xz, gzeze = zx * zx + zz * zz, gxize = xx * zx + xz * zz;
        const bool side = scol >= 0, bottom = act && iz >= zb0;   // PML memory only in updated columns
        const bool at_row = iz == 0;
        const float qvx = at_row ? a.qif[3 * (long long)P + ix] : kbar(a, g.off[MSFDM_SVX] + si);
        const float qvz = at_row ? a.qif[4 * (long long)P + ix] : kbar(a, g.off[MSFDM_SVZ] + si);
        const float qsx = at_row ? a.qif[5 * (long long)P + ix] : kbar(a, g.off[MSFDM_SSX] + si);
        const float qsz = at_row ? a.qif[6 * (long long)P + ix] : kbar(a, g.off[MSFDM_SSZ] + si);
        const float qsxz = at_row ? a.qif[7 * (long long)P + ix] : kbar(a, g.off[MSFDM_SSXZ] + si);
        const float c11 = c.C11[si], c13 = c.C13[si], c33 = c.C33[si], c55 = c.C55[si], rinv = c.rhos_inv[si];

        // ---- stress equations
        {
            const float dvxx = drp_g<DIR>(Svx, si, 1), dvzx = drp_g<DIR>(Svz, si, 1);
            const float dvxz = mac_g<DIR>(Svx, si, P, iz, true), dvzz = mac_g<DIR>(Svz, si, P, iz, true);
            const float exx = xx * dvxx + zx * dvxz, ezz = xz * dvzx + zz * dvzz;
            float A11 = exx, A13 = ezz, A55 = xz * dvxx + zz * dvxz + xx * dvzx + zx * dvzz;
            const float exxb = c11 * qsx + c13 * qsz, ezzb = c13 * qsx + c33 * qsz;
            float bxx = xx * exxb + c55 * xz * qsxz, bxz = zx * exxb + c55 * zz * qsxz;
            float bzx = xz * ezzb + c55 * xx * qsxz, bzz = zz * ezzb + c55 * zx * qsxz;
            if (side) {
                const float gx = gxize / gxixi;
                const float pxa = c.pxa[ix], pxb = c.pxb[ix], pxc = c.pxc[ix], q1 = pxc - 1.0f;
                const long long ai = (long long)iz * g.W2 + scol;
                const float Ex = S[g.off[MSFDM_SXEX] + ai], Ez = S[g.off[MSFDM_SXEZ] + ai];
                const float qEx = kbar(a, g.off[MSFDM_SXEX] + ai), qEz = kbar(a, g.off[MSFDM_SXEZ] + ai);
                const float Dvxb = q1 * xx * exxb + c55 * q1 * xz * qsxz + pxb * qEx;
                const float Dvzb = q1 * xz * ezzb + c55 * q1 * xx * qsxz + pxb * qEz;
                a.TB[g.off[MSFDM_SXEX] + ai] += -pxc * xx * exxb - pxc * c55 * xz * qsxz - pxa * qEx;
                a.TB[g.off[MSFDM_SXEZ] + ai] += -pxc * xz * ezzb - pxc * c55 * xx * qsxz - pxa * qEz;
                bxx += Dvxb; bxz += gx * Dvxb; bzx += Dvzb; bzz += gx * Dvzb;
                const float Dvx = dvxx + gx * dvxz, Dvz = dvzx + gx * dvzz;
                A11 += q1 * xx * Dvx - pxc * xx * Ex;
                A13 += q1 * xz * Dvz - pxc * xz * Ez;
                A55 += q1 * (xz * Dvx + xx * Dvz) - pxc * xz * Ex - pxc * xx * Ez;
            }
            if (bottom) {
                const float gz = gxize / gzeze;
                const float pza = c.pza[iz], pzb = c.pzb[iz], pzc = c.pzc[iz], q1 = pzc - 1.0f;
                const long long ai = (long long)(iz - zb0) * P + ix;
                const float Ex = S[g.off[MSFDM_SZEX] + ai], Ez = S[g.off[MSFDM_SZEZ] + ai];
                const float qEx = kbar(a, g.off[MSFDM_SZEX] + ai), qEz = kbar(a, g.off[MSFDM_SZEZ] + ai);
                const float Dvxb = q1 * zx * exxb + c55 * q1 * zz * qsxz + pzb * qEx;
                const float Dvzb = q1 * zz * ezzb + c55 * q1 * zx * qsxz + pzb * qEz;
                a.TB[g.off[MSFDM_SZEX] + ai] += -pzc * zx * exxb - pzc * c55 * zz * qsxz - pza * qEx;
                a.TB[g.off[MSFDM_SZEZ] + ai] += -pzc * zz * ezzb - pzc * c55 * zx * qsxz - pza * qEz;
                bxz += Dvxb; bxx += gz * Dvxb; bzz += Dvzb; bzx += gz * Dvzb;
                const float Dvx = dvxz + gz * dvxx, Dvz = dvzz + gz * dvzx;
                A11 += q1 * zx * Dvx - pzc * zx * Ex;
                A13 += q1 * zz * Dvz - pzc * zz * Ez;
                A55 += q1 * (zz * Dvx + zx * Dvz) - pzc * zz * Ex - pzc * zx * Ez;
            }
            float xa[5], xb[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                xa[q] = drp_w<DIR>(q) * bxx;
                xb[q] = drp_w<DIR>(q) * bzx;
            }
            addx(0, iz, xa);
            addx(1, iz, xb);
            mac_dz_adj<DIR>(iz, true, bxz, [&](int m, float v) { add(0, iz + m, v); });
            mac_dz_adj<DIR>(iz, true, bzz, [&](int m, float v) { add(1, iz + m, v); });
            if (act) {
                gr.C11[si] += (double)qsx * (double)A11;
                gr.C13[si] += (double)qsx * (double)A13 + (double)qsz * (double)A11;
                gr.C33[si] += (double)qsz * (double)A13;
                gr.C55[si] += (double)qsxz * (double)A55;
            }
        }
        // ---- velocity equations
        {
            const float rvxb = qvx * rinv, rvzb = qvz * rinv;
            float Dxxb = 0.f, Dxzb = 0.f, Dzxb = 0.f, Dzzb = 0.f;
            float q1x = 0.f, pxc = 1.f, q1z = 0.f, pzc = 1.f;
            float Vxs = 0.f, Vzs = 0.f, Vxb = 0.f, Vzb = 0.f;
            if (side) {
                const float pxa = c.pxa[ix], pxb = c.pxb[ix];
                pxc = c.pxc[ix]; q1x = pxc - 1.0f;
                const long long ai = (long long)iz * g.W2 + scol;
                const float qVx = kbar(a, g.off[MSFDM_SXVX] + ai), qVz = kbar(a, g.off[MSFDM_SXVZ] + ai);
                Dxxb = q1x * rinv * qvx + pxb * qVx;
                Dxzb = q1x * rinv * qvz + pxb * qVz;
                a.TB[g.off[MSFDM_SXVX] + ai] += -pxc * rinv * qvx - pxa * qVx;
                a.TB[g.off[MSFDM_SXVZ] + ai] += -pxc * rinv * qvz - pxa * qVz;
                Vxs = S[g.off[MSFDM_SXVX] + ai]; Vzs = S[g.off[MSFDM_SXVZ] + ai];
            }
            if (bottom) {
                const float pza = c.pza[iz], pzb = c.pzb[iz];
                pzc = c.pzc[iz]; q1z = pzc - 1.0f;
                const long long ai = (long long)(iz - zb0) * P + ix;
                const float qVx = kbar(a, g.off[MSFDM_SZVX] + ai), qVz = kbar(a, g.off[MSFDM_SZVZ] + ai);
                Dzxb = q1z * rinv * qvx + pzb * qVx;
                Dzzb = q1z * rinv * qvz + pzb * qVz;
                a.TB[g.off[MSFDM_SZVX] + ai] += -pzc * rinv * qvx - pza * qVx;
                a.TB[g.off[MSFDM_SZVZ] + ai] += -pzc * rinv * qvz - pza * qVz;
                Vxb = S[g.off[MSFDM_SZVX] + ai]; Vzb = S[g.off[MSFDM_SZVZ] + ai];
            }
            float rvx, rvz, Dxx, Dxz, Dzx, Dzz;
            constexpr int o = lo_off<DIR>();
            if (iz < 3) {
                auto fze_row = [&](int rr, float &fx, float &fz) {
                    const long long q = moff + (long long)rr * P + ix, sq = (long long)rr * P + ix;
                    const float Jq = c.J[q], zxq = c.zex[q], zzq = c.zez[q];
                    fx = Jq * (zxq * Ssx[sq] + zzq * Ssxz[sq]);
                    fz = Jq * (zxq * Ssxz[sq] + zzq * Ssz[sq]);
                };
                float sx = 0.f, sz = 0.f;
#pragma unroll
                for (int q = 0; q < 5; ++q) {
                    const int m = o + q;
                    const float w = drp_w<DIR>(q);
                    const long long gq = gi + m, sq = si + m;
                    const float Jq = c.J[gq];
                    float xxq, xzq;
                    xi_metrics<2>(c, gq, xxq, xzq);
                    sx += w * Jq * (xxq * Ssx[sq] + xzq * Ssxz[sq]);
                    sz += w * Jq * (xxq * Ssxz[sq] + xzq * Ssz[sq]);
                    const int rr = iz + m;
                    float bxv, bzv;
                    if (rr >= 0) fze_row(rr, bxv, bzv);
                    else { float f0x, f0z, fnx, fnz; fze_row(0, f0x, f0z); fze_row(-rr, fnx, fnz); bxv = 2.f * f0x - fnx; bzv = 2.f * f0z - fnz; }
                    sx += w * bxv;
                    sz += w * bzv;
                }
                rvx = sx / J; rvz = sz / J;
                const float pxi = (xx * sx + xz * sz) / (gxixi * J), pze = (zx * sx + zz * sz) / (gzeze * J);
                Dxx = xx * pxi; Dxz = xz * pxi; Dzx = zx * pze; Dzz = zz * pze;
                const float pxib = xx * Dxxb + xz * Dxzb, pzeb = zx * Dzxb + zz * Dzzb;
                const float sxb = rvxb / J + xx * pxib / (gxixi * J) + zx * pzeb / (gzeze * J);
                const float szb = rvzb / J + xz * pxib / (gxixi * J) + zz * pzeb / (gzeze * J);
                auto fze_add = [&](int rr, float Fx, float Fz) {
                    const long long q = moff + (long long)rr * P + ix;
                    const float Jq = c.J[q], zxq = c.zex[q], zzq = c.zez[q];
                    add(2, rr, Jq * zxq * Fx);
                    add(4, rr, Jq * (zzq * Fx + zxq * Fz));
                    add(3, rr, Jq * zzq * Fz);
                };
                float x2[5], x3[5], x4[5];
#pragma unroll
                for (int q = 0; q < 5; ++q) {
                    const int m = o + q;
                    const float w = drp_w<DIR>(q);
                    const long long gq = gi + m;
                    const float Jq = c.J[gq];
                    float xxq, xzq;
                    xi_metrics<2>(c, gq, xxq, xzq);
                    const float Fx = w * sxb, Fz = w * szb;
                    x2[q] = Jq * xxq * Fx;
                    x4[q] = Jq * (xzq * Fx + xxq * Fz);
                    x3[q] = Jq * xzq * Fz;
                    const int rr = iz + m;
                    if (rr >= 0) fze_add(rr, Fx, Fz);
                    else { fze_add(0, 2.f * Fx, 2.f * Fz); fze_add(-rr, -Fx, -Fz); }
                }
                addx(2, iz, x2);
                addx(4, iz, x4);
                addx(3, iz, x3);
            } else {
                const float A = drp_g<DIR>(Ssx, si, 1), Bv = drp_g<DIR>(Ssx, si, P);
                const float Cv = drp_g<DIR>(Ssz, si, 1), Dv = drp_g<DIR>(Ssz, si, P);
                const float E = drp_g<DIR>(Ssxz, si, 1), Fv = drp_g<DIR>(Ssxz, si, P);
                const float gx = gxize / gxixi, gz = gxize / gzeze;
                rvx = xx * A + zx * Bv + xz * E + zz * Fv;
                rvz = xx * E + zx * Fv + xz * Cv + zz * Dv;
                Dxx = xx * (A + gx * Bv) + xz * (E + gx * Fv);
                Dxz = xx * (E + gx * Fv) + xz * (Cv + gx * Dv);
                Dzx = zx * (Bv + gz * A) + zz * (gz * E + Fv);
                Dzz = zx * (gz * E + Fv) + zz * (gz * Cv + Dv);
                const float ab = xx * rvxb + xx * Dxxb + zx * gz * Dzxb;
                const float bb = zx * rvxb + xx * gx * Dxxb + zx * Dzxb;
                const float cb = xz * rvzb + xz * Dxzb + zz * gz * Dzzb;
                const float db = zz * rvzb + xz * gx * Dxzb + zz * Dzzb;
                const float eb = xz * rvxb + xx * rvzb + xz * Dxxb + xx * Dxzb + zz * gz * Dzxb + zx * gz * Dzzb;
                const float fb = zz * rvxb + zx * rvzb + xz * gx * Dxxb + xx * gx * Dxzb + zz * Dzxb + zx * Dzzb;
                float x2[5], x3[5], x4[5];
#pragma unroll
                for (int q = 0; q < 5; ++q) {
                    const float w = drp_w<DIR>(q);
                    x2[q] = w * ab;
                    add(2, iz + o + q, w * bb);
                    x3[q] = w * cb;
                    add(3, iz + o + q, w * db);
                    x4[q] = w * eb;
                    add(4, iz + o + q, w * fb);
                }
                addx(2, iz, x2);
                addx(3, iz, x3);
                addx(4, iz, x4);
            }
            float kvx = rvx, kvz = rvz;
            if (side) { kvx += q1x * Dxx - pxc * Vxs; kvz += q1x * Dxz - pxc * Vzs; }
            if (bottom) { kvx += q1z * Dzx - pzc * Vxb; kvz += q1z * Dzz - pzc * Vzb; }
            if (act) gr.rho_s[si] += -(double)(qvx * kvx + qvz * kvz) * (double)rinv * (double)rinv;
        }
    }
}

// Zero the per-wave LDS accumulators, run the medium body, add the tile and
// its halo to the output.  Wave w owns tile rows w*RW .. w*RW+RW-1 and
// accumulates into its own copy covering rows w*RW-3 .. w*RW+RW+2; no other
// wave touches the copy and each cell of it has one writing lane, so every LDS
// update is a plain read-modify-write.  The flush adds the overlapping copies
// in wave order; no other block of the launch (one tile colour) covers the
// same output cells.
template <int NF, int TZ>
__device__ __forceinline__ void acc_zero(WaveAcc<TZ> *acc) {
    constexpr int RW = TZ / 4, N = (RW + 2 * HALO) * LXW;
    for (int w = 0; w < 4; ++w)
        for (int f = 0; f < NF; ++f)
            for (int e = threadIdx.x; e < N; e += NTHR) {
                const int r = e / LXW, cc = e - r * LXW;
                acc[w].v[f][r][cc] = 0.0f;
            }
}
template <int NF, int TZ>
__device__ __forceinline__ void acc_flush(WaveAcc<TZ> *acc, float *TB, const long long *off, int f0, int z0, int x0,
                                          int nrows, int P) {
    constexpr int RW = TZ / 4, N = (TZ + 2 * HALO) * LXW;
    for (int f = 0; f < NF; ++f) {
        float *dst = TB + off[f0 + f];
        for (int e = threadIdx.x; e < N; e += NTHR) {
            const int r = e / LXW, cc = e - r * LXW;   // tile row r - HALO
            float v = 0.0f;
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const int lr = r - w * RW;
                if (lr >= 0 && lr < RW + 2 * HALO) v += acc[w].v[f][lr][cc];
            }
            const int gz = z0 - HALO + r, gx = x0 - HALO + cc;
            if (v != 0.0f && gz >= 0 && gz < nrows && gx >= 0 && gx < P) dst[(long long)gz * P + gx] += v;
        }
    }
}

// Tiles of colour `color` (bit 0: tile column parity, bit 1: band parity within
// its medium).  Two tiles of one colour are at least one tile apart in x (64
// columns) and in z (TZ >= 8 rows), more than the 2 x 3-cell halos of their
// accumulators.
__host__ __device__ __forceinline__ int adj_color_tiles(int nbx, int nbz_f, int nbz_s, int color) {
    const int px = color & 1, pz = color >> 1;
    return ((nbx - px + 1) / 2) * ((nbz_f - pz + 1) / 2 + (nbz_s - pz + 1) / 2);
}

template <int DIR, int TZ>
__global__ void __launch_bounds__(NTHR)
adj_stage_kernel(Geo g, Coef c, AdjIn a0, Grad gr0, int nbx, int nbz_f, int nbz_s, int color) {
    static_assert(TZ >= 2 * HALO, "tiles of one colour must not share accumulator cells");
    __shared__ WaveAcc<TZ> acc[4];
    const AdjIn a = adj_of_shot(a0, blockIdx.y);
    const Grad gr = grad_of_shot(gr0, blockIdx.y);
    const int per_xcd = gridDim.x >> 3;
    const int t = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (t >= adj_color_tiles(nbx, nbz_f, nbz_s, color)) return;
    const int px = color & 1, pz = color >> 1;
    const int nbxc = (nbx - px + 1) / 2, nfc = (nbz_f - pz + 1) / 2;
    const int bc = t / nbxc, bx = 2 * (t - bc * nbxc) + px;
    if (bc < nfc) {
        const int band = 2 * bc + pz;
        acc_zero<3, TZ>(acc);
        __syncthreads();
        fluid_adj_body<DIR, TZ>(acc, g, c, a, gr, bx, band);
        __syncthreads();
        acc_flush<3, TZ>(acc, a.TB, g.off, MSFDM_FVX, 3 + band * TZ, bx * TX, g.nzf, g.P);
    } else {
        const int band = 2 * (bc - nfc) + pz;
        acc_zero<5, TZ>(acc);
        __syncthreads();
        solid_adj_body<DIR, TZ>(acc, g, c, a, gr, bx, band);
        __syncthreads();
        acc_flush<5, TZ>(acc, a.TB, g.off, MSFDM_SVX, band * TZ, bx * TX, g.nzs, g.P);
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

// qif = B^T kbar on the interface rows, kbar = clam * LP + cprev * TBp.
template <int DIR>
__global__ void __launch_bounds__(128) adj_iface_q(Geo g, Coef c, AdjIn a0, float *qif) {
    const AdjIn a = adj_of_shot(a0, blockIdx.y);
    qif += (long long)blockIdx.y * a0.qstride;
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    if (ix < 3 || ix > g.nx - 4) return;
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
        qif[(long long)j * g.P + ix] = s;
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

// ACC += TB over n floats.
__global__ void axpy_kernel(float *__restrict__ acc, const float *__restrict__ tb, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        acc[i] += tb[i];
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
