// Device solver: HBM layout, PML/normal/source tables, RK4 stage launches,
// hipGraph replay of the time loop and the msfdm_* device entry points.
//
// Replaces MemManager's device half (src/elements.cu:557-905, 979-987),
// PmlFs::FluidSolid::update_RK4 (src/pml_fs_kernels.cu:765-868) and the
// record sampling of src/main.cu:245-295.  One stream, explicit ordering (no
// cross-stream races), records kept on the device.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "model.hpp"
#include "stencil.hpp"
#include "adjoint.hpp"

using namespace msfdm;
namespace D = msfdm::dev;

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return set_error(MSFDM_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

namespace {

constexpr int GRAPH_STEPS = 32;   // steps captured per graph replay (even: F/B parity is baked in)

enum Role { R_U = 0, R_T2 = 1, R_T3 = 2, R_T4 = 3 };
constexpr int SHOT_UOUT = 4;   // D::Shot::buf entry that stage 4 writes

inline long long align64(long long n) { return (n + 63) / 64 * 64; }

}  // namespace

struct msfdm_solver {
    int device = 0;
    hipStream_t stream = nullptr;
    msfdm_desc d{};
    std::vector<int> srcx, srcz;
    D::Geo geo{};
    D::Coef coef{};
    D::Ctl ctl{};
    long long buf_floats = 0;
    float *buf[4] = {nullptr, nullptr, nullptr, nullptr};
    float *kbuf = nullptr;       // raw-k output of msfdm_eval_rhs
    void *stage = nullptr;       // packed rows of a host transfer (host_rows)
    size_t stage_bytes = 0;
    float *kif = nullptr;        // [KIF_ROWS][P]: raw interface-row k of the 4 stages + U^n interface rows
    float *coefmem = nullptr;
    float *amp = nullptr;
    int *base = nullptr;
    float *rec = nullptr;
    int amp_cap = 0;             // steps covered by the amplitude table
    int rec_cap = 0;             // record slots
    int gap = 1;
    int host_it = 0;
    int isrc = 0;
    int nrec_done = 0;
    int tz = 16;                 // tile height (rows per block)
    int occ = 0;                 // occupancy policy of the stage kernels (msfdm_set_occupancy)
    float *u_out = nullptr;      // non-null: stage 4 writes U^{n+1} here instead of over U^n
                                 // (the adjoint's segment recomputation; never in a graph)
    int adj_launch = -1;         // adjoint products: -1 by grid size, 0 three launches, 1 one launch
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double bytes_step = 0.0;
    double bytes_kernel[3] = {0.0, 0.0, 0.0};
    // adjoint state: lambda, lambda', accumulator, two Tbar buffers, interface qbar
    float *adj[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    float *aprod = nullptr;          // adjoint products of one stage (adj_prod_kernel -> adj_gather_kernel)
    int2 *atiles = nullptr;          // adjoint source tiles: interior, fluid boundary, solid boundary
    int atiles_n[3] = {0, 0, 0};
    int atiles_xcd = 0;              // entries per XCD of the one-launch list (after the three lists)
    double *gmem = nullptr;          // 7 gradient arrays, [rows][P] doubles, three times:
                                     // segment partial sums, Kahan totals, compensations
    long long gfl = 0;               // doubles of one set of 7 arrays
    D::Grad grad{};
    std::vector<float *> ckpt;       // checkpoints of the full state every adj_every steps (the first
                                     // adj_nseg in use; kept allocated across gradients for reuse)
    std::vector<float *> seg;        // per-step U^n, T2, T3, T4 of one segment
    int adj_every = 0, adj_nsteps = 0, adj_isrc = 0, adj_nseg = 0;
    bool solid_only = false;         // grid.calc_type = 0: a solid grid
    int top = D::TOP_IFACE;          // top kind of the solid stage: interface, free surface or plain
    // In-sequence kernel timing (msfdm_profile_run): while `on`, every launch of
    // the stepper carries a start / stop event pair stamped by its own dispatch.
    struct KProf {
        bool on = false;
        int mask = 0;                // kernel classes timed (bit c: class c)
        std::vector<hipEvent_t> ev;  // 2 per timed launch, kept for reuse
        std::vector<int> cls;        // kernel class of each timed launch of the current run
    } prof;
};

namespace {

// ---------------------------------------------------------------------------
// Tables
// ---------------------------------------------------------------------------

// Ricker at a float stage time (src/elements.cu:385-392): T frozen as float
// 1.2/fpeak, zero once t - T > T; the wavelet is evaluated in double.
double ricker(float fpeak, float time) {
    const float T = (float)(1.2 / (double)fpeak);
    const float t = time - T;
    if (t > T) return 0.0;
    double a = M_PI * (double)fpeak * (double)t;
    a *= a;
    return (1.0 - 2.0 * a) * std::exp(-a);
}
// Stage times: float it*dt, +dt/2, +dt/2, +dt (src/pml_fs_kernels.cu:768-844).
double stage_amp(float fpeak, float dt, int it, int stage) {
    const float time = (float)it * dt;
    const float hf = (float)(dt * 0.5);
    float t = time;
    if (stage == 1 || stage == 2) t = time + hf;
    if (stage == 3) t = time + dt;
    return ricker(fpeak, t);
}

// PML profile (src/elements.cu:824-874) in double: returns al + d/be, d/be, 1/be.
void pml_profile(int n, int npd, int npow, bool lo, bool hi, double damp0, double beta0, double alpha0,
                 std::vector<float> &pa, std::vector<float> &pb, std::vector<float> &pc) {
    pa.assign(n, 0.f); pb.assign(n, 0.f); pc.assign(n, 1.f);
    for (int i = 0; i < n; ++i) {
        int k = -1;
        if (lo && i >= 3 && i <= 2 + npd) k = 3 + npd - i;
        if (hi && i >= n - 3 - npd && i <= n - 4) k = i - (n - 4 - npd);
        double dp = 0.0, be = 1.0, al = alpha0;
        if (k > 0) {
            const double r = double(k) / double(npd);
            dp = damp0 * std::pow(r, (double)npow);
            be = 1.0 + (beta0 - 1.0) * r * r;
            al = alpha0 * (1.0 - r);
        }
        pa[i] = (float)(al + dp / be);
        pb[i] = (float)(dp / be);
        pc[i] = (float)(1.0 / be);
    }
}

// Kernel classes of msfdm_profile_run: RK stages 1..4 (0..3), the interface
// finish (4), the step counter (5).
constexpr int KCLASS_IFACE = 4, KCLASS_OTHER = 5, NKCLASS = 6;

// Launch `kernel` on the solver's stream; during a profiled run with a start /
// stop event pair recorded by the dispatch itself (hipExtLaunchKernelGGL), so
// the events time the kernel and not the gaps between launches.
template <class F, class... A>
void klaunch(msfdm_solver *s, int cls, F kernel, dim3 grid, dim3 block, A... args) {
    if (!s->prof.on || !((s->prof.mask >> cls) & 1)) {
        hipLaunchKernelGGL(kernel, grid, block, 0, s->stream, args...);
        return;
    }
    const size_t i = s->prof.cls.size();
    while (s->prof.ev.size() < 2 * i + 2) {
        // timing-only events: without the system-scope fence the stamped
        // dispatch ends with the device-scope release of every other launch
        // (no L2 write-back / invalidate inside the timed kernel or the next)
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) {   // out of events: launch untimed
            hipLaunchKernelGGL(kernel, grid, block, 0, s->stream, args...);
            return;
        }
        s->prof.ev.push_back(e);
    }
    hipExtLaunchKernelGGL(kernel, grid, block, 0, s->stream, s->prof.ev[2 * i], s->prof.ev[2 * i + 1], 0, args...);
    s->prof.cls.push_back(cls);
}

int grow_tables(msfdm_solver *s, int steps_needed) {
    // amplitude table: 4 floats per step
    if (steps_needed > s->amp_cap) {
        int cap = std::max(steps_needed, s->amp_cap * 2);
        cap = std::max(cap, 64);
        std::vector<float> h(4 * (size_t)cap);
        for (int it = 0; it < cap; ++it)
            for (int st = 0; st < 4; ++st) h[4 * (size_t)it + st] = (float)stage_amp(s->d.fpeak, s->d.dt, it, st);
        float *na = nullptr;
        HIPCHK(hipMalloc(&na, h.size() * sizeof(float)));
        HIPCHK(hipMemcpy(na, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
        if (s->amp) { HIPCHK(hipStreamSynchronize(s->stream)); HIPCHK(hipFree(s->amp)); }
        s->amp = na;
        s->amp_cap = cap;
        s->ctl.amp = na;
        if (s->gexec) { (void)hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
        if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
    }
    const int slots = steps_needed > 0 ? (steps_needed - 1) / s->gap + 1 : 0;
    if (slots > s->rec_cap) {
        const int cap = std::max({slots, s->rec_cap * 2, 16});
        const size_t nx = (size_t)s->d.nx;
        float *nr = nullptr;
        HIPCHK(hipMalloc(&nr, 3 * (size_t)cap * nx * sizeof(float)));
        HIPCHK(hipMemsetAsync(nr, 0, 3 * (size_t)cap * nx * sizeof(float), s->stream));
        if (s->rec) {
            for (int p = 0; p < 3; ++p)
                HIPCHK(hipMemcpyAsync(nr + (size_t)p * cap * nx, s->rec + (size_t)p * s->rec_cap * nx,
                                      (size_t)s->rec_cap * nx * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
            HIPCHK(hipStreamSynchronize(s->stream));
            HIPCHK(hipFree(s->rec));
        }
        s->rec = nr;
        s->rec_cap = cap;
        s->ctl.rec = nr;
        s->ctl.rec_cap = cap;
        if (s->gexec) { (void)hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
        if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
    }
    return MSFDM_OK;
}

// ---------------------------------------------------------------------------
// Launches
// ---------------------------------------------------------------------------
template <int DIR, int STAGE, int TZ, int FIN = 0>
void launch_media(msfdm_solver *s, const float *S, const float *U, const float *T2, const float *T3, float *O, int j,
                  float ct, float amp, float cprev) {
    const int nbx = (s->d.nx - 4) / D::TX + 1;
    if (s->solid_only) {   // rows 3..nzs-4 of the solid-only grid
        const int nb = (s->d.nz_solid - 6 + TZ - 1) / TZ;
        const int grid = 8 * ((nbx + 7) / 8 + (nbx * (nb - 1) + 7) / 8);   // D::xcd_tile with no fluid bands
#define MSFDM_FREE_LAUNCH(KERNEL, MET, TOP)                                                                    \
    klaunch(s, STAGE - 1, D::KERNEL<DIR, STAGE, TZ, MET, TOP>, dim3(grid), dim3(D::NTHR), s->geo, s->coef, s->ctl, S, \
            U, T2, T3, O, j, ct, amp, nbx, nb)
        const bool w5 = s->occ == 1 && s->coef.vert;   // the occupancy variants exist for vertical columns
        if (s->top == D::TOP_FREE) {
            if (w5) MSFDM_FREE_LAUNCH(free_stage_kernel_w5, 1, D::TOP_FREE);
            else if (s->coef.vert) MSFDM_FREE_LAUNCH(free_stage_kernel, 1, D::TOP_FREE);
            else MSFDM_FREE_LAUNCH(free_stage_kernel, 0, D::TOP_FREE);
        } else {
            if (w5) MSFDM_FREE_LAUNCH(free_stage_kernel_w5, 1, D::TOP_PLAIN);
            else if (s->coef.vert) MSFDM_FREE_LAUNCH(free_stage_kernel, 1, D::TOP_PLAIN);
            else MSFDM_FREE_LAUNCH(free_stage_kernel, 0, D::TOP_PLAIN);
        }
#undef MSFDM_FREE_LAUNCH
        return;
    }
    const int nzf = s->d.nz_fluid;
    const int nbf = (nzf - 3 + TZ - 1) / TZ, nbs = (s->d.nz_solid - 3 + TZ - 1) / TZ;
    int nfi = 0;   // fluid bands whose loaded rows (z0 - 3 .. z0 + TZ + 2, z0 = D::fluid_z0) hold row nzf - 1
    for (int b = 0; b < nbf; ++b) nfi += (nzf - 1) - (nzf - (nbf - b) * TZ) < TZ + 2 * D::HALO;
    // each XCD: one eighth of the interface-band, the other fluid and the other solid tiles (D::xcd_tile)
    const int grid = 8 * ((nbx * (nfi + 1) + 7) / 8 + (nbx * (nbf - nfi) + 7) / 8 + (nbx * (nbs - 1) + 7) / 8);
    const int cls = STAGE > 0 ? STAGE - 1 : KCLASS_OTHER;
    if (s->geo.fcpml)   // top PML in the water; metric form decided at run time
        klaunch(s, cls, D::stage_kernel<DIR, STAGE, TZ, 2, 1, FIN>, dim3(grid), dim3(D::NTHR), s->geo, s->coef, s->ctl,
                S, U, T2, T3, O, s->kif, j, ct, amp, cprev, nbx, nbf, nbs, nfi);
    else if (s->coef.vert && s->occ == 1 && STAGE > 0)   // the other occupancy policy
        klaunch(s, cls, D::stage_kernel_w5<DIR, STAGE, TZ, 1, 0, FIN>, dim3(grid), dim3(D::NTHR), s->geo, s->coef,
                s->ctl, S, U, T2, T3, O, s->kif, j, ct, amp, cprev, nbx, nbf, nbs, nfi);
    else if (s->coef.vert)   // vertical grid columns: xi_x, xi_z are constants of the kernel
        klaunch(s, cls, D::stage_kernel<DIR, STAGE, TZ, 1, 0, FIN>, dim3(grid), dim3(D::NTHR), s->geo, s->coef, s->ctl,
                S, U, T2, T3, O, s->kif, j, ct, amp, cprev, nbx, nbf, nbs, nfi);
    else
        klaunch(s, cls, D::stage_kernel<DIR, STAGE, TZ, 0, 0, FIN>, dim3(grid), dim3(D::NTHR), s->geo, s->coef, s->ctl,
                S, U, T2, T3, O, s->kif, j, ct, amp, cprev, nbx, nbf, nbs, nfi);
}
// One RK stage; cprev is the previous stage's weight (its interface rows are
// rebuilt by this stage's tiles).  Stage 4 adds the interface / filter / record
// kernel unless the next step's stage 1 does that work (fin_next), the raw-k
// mode (STAGE 0) the interface condition on the raw k.  FIN: a stage 1 that
// finishes the previous step (cprev = dt).
template <int DIR, int STAGE, int FIN = 0>
void launch_stage(msfdm_solver *s, const float *S, float *O, int j, float ct, float amp, float cprev,
                  bool fin_next = false) {
    const float *U = s->buf[R_U], *T2 = s->buf[R_T2], *T3 = s->buf[R_T3];
    if (s->tz == 8) launch_media<DIR, STAGE, 8, FIN>(s, S, U, T2, T3, O, j, ct, amp, cprev);
    else launch_media<DIR, STAGE, 16, FIN>(s, S, U, T2, T3, O, j, ct, amp, cprev);
    if (s->solid_only) return;   // no interface: the stage kernel writes every row and the record
    if constexpr (STAGE == 4) {
        if (fin_next) return;
        const float dt = s->d.dt;
        klaunch(s, KCLASS_IFACE, D::iface_final<DIR>, dim3((s->d.nx + D::FO - 1) / D::FO), dim3(D::FB), s->geo, s->coef,
                s->ctl, O, (const float *)s->kif, dt / 2, dt, ct, j);
    } else if constexpr (STAGE == 0) {
        D::iface_rhs<DIR><<<(s->d.nx + 127) / 128, 128, 0, s->stream>>>(s->geo, s->coef, O, s->kif);
    }
}

// One time step (update_RK4 FBFB / BFBF, src/pml_fs_kernels.cu:771-844) at
// offset j of the current batch; it is the absolute step index.  Within one
// launch sequence the interface rows of a step are finished by the next step's
// stage-1 launch (first: no step before it in the sequence; last: none after).
template <int D1>
void enqueue_step_dir(msfdm_solver *s, int j, bool first, bool last) {
    constexpr int D2 = 1 - D1;
    const float dt = s->d.dt;
    float *U = s->buf[R_U], *T2 = s->buf[R_T2], *T3 = s->buf[R_T3], *T4 = s->buf[R_T4];
    if (first || s->solid_only) launch_stage<D1, 1>(s, U, T2, j, dt / 2, 0.f, 0.f);
    else launch_stage<D1, 1, 1>(s, U, T2, j, dt / 2, 0.f, dt);
    launch_stage<D2, 2>(s, T2, T3, j, dt / 2, 0.f, dt / 2);
    launch_stage<D1, 3>(s, T3, T4, j, dt, 0.f, dt / 2);
    launch_stage<D2, 4>(s, T4, s->u_out ? s->u_out : U, j, dt / 6, 0.f, dt, !last);
}
// n steps from absolute step it0; U, the records and every interface row are
// complete when the sequence ends (nothing reads them in between).
void enqueue_steps(msfdm_solver *s, int it0, int n) {
    for (int j = 0; j < n; ++j) {
        if ((it0 + j) & 1) enqueue_step_dir<1>(s, j, j == 0, j == n - 1);
        else enqueue_step_dir<0>(s, j, j == 0, j == n - 1);
    }
    klaunch(s, KCLASS_OTHER, D::advance_base, dim3(1), dim3(64), s->base, n);
}

int build_graph(msfdm_solver *s) {
    HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    enqueue_steps(s, 0, GRAPH_STEPS);
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamEndCapture(s->stream, &g));
    s->graph = g;
    HIPCHK(hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0));
    return MSFDM_OK;
}

void drop_graph(msfdm_solver *s) {
    if (s->gexec) { (void)hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
    if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
}

int field_rows(const msfdm_solver *s, int f) {
    if (f <= MSFDM_FXEZ || f >= MSFDM_FZVX) return s->d.nz_fluid;
    return s->d.nz_solid;
}

// `rows` rows of row_bytes between packed host memory and a device array of
// pitch dpitch.  A pitched copy straight into or out of pageable host memory
// is split by the runtime into one transfer per row (some 6500 for the
// gradients of a 4001 x 1501 grid, 0.4-0.9 s); packing the rows on the device
// first leaves a single transfer.
int host_rows(msfdm_solver *s, void *host, void *dev, size_t dpitch, size_t row_bytes, int rows, bool to_host) {
    if (rows <= 0) return MSFDM_OK;
    const size_t n = row_bytes * (size_t)rows;
    HIPCHK(hipStreamSynchronize(s->stream));
    if (s->stage_bytes < n) {
        if (s->stage) HIPCHK(hipFree(s->stage));
        s->stage = nullptr;
        s->stage_bytes = 0;
        HIPCHK(hipMalloc(&s->stage, n));
        s->stage_bytes = n;
    }
    if (to_host) {
        HIPCHK(hipMemcpy2DAsync(s->stage, row_bytes, dev, dpitch, row_bytes, rows, hipMemcpyDeviceToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        HIPCHK(hipMemcpy(host, s->stage, n, hipMemcpyDeviceToHost));
    } else {
        HIPCHK(hipMemcpy(s->stage, host, n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy2DAsync(dev, dpitch, s->stage, row_bytes, row_bytes, rows, hipMemcpyDeviceToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    return MSFDM_OK;
}

// Copy between the reference layout [rows][nx] and the device layout of field f.
int copy_field(msfdm_solver *s, float *dbuf, int f, float *host_out, const float *host_in) {
    const int nx = s->d.nx, P = s->geo.P, npd = s->geo.npd, W2 = s->geo.W2;
    const int rows = field_rows(s, f);
    float *dp = dbuf + s->geo.off[f];
    const bool side = (f >= MSFDM_FXVX && f <= MSFDM_FXEZ) || (f >= MSFDM_SXVX && f <= MSFDM_SXEZ);
    if (f >= MSFDM_FZVX) {
        // fluid zeta memory: top strip, fluid rows 3 .. npd+2 (stored on the CPML grid only)
        HIPCHK(hipStreamSynchronize(s->stream));
        if (host_out) std::fill(host_out, host_out + (size_t)rows * nx, 0.0f);
        if (!s->geo.fcpml || npd == 0) return MSFDM_OK;
        return host_rows(s, host_out ? host_out + 3 * (size_t)nx : (void *)(host_in + 3 * (size_t)nx), dp,
                         P * sizeof(float), nx * sizeof(float), npd, host_out != nullptr);
    }
    const bool bottom = f >= MSFDM_SZVX;
    HIPCHK(hipStreamSynchronize(s->stream));
    if (!side && !bottom) {
        return host_rows(s, host_out ? host_out : (void *)host_in, dp, P * sizeof(float), nx * sizeof(float), rows,
                         host_out != nullptr);
    }
    if (side) {
        std::vector<float> c((size_t)rows * W2);
        if (host_out) {
            HIPCHK(hipMemcpy(c.data(), dp, c.size() * sizeof(float), hipMemcpyDeviceToHost));
            std::fill(host_out, host_out + (size_t)rows * nx, 0.0f);
            for (int r = 0; r < rows; ++r)
                for (int k = 0; k < W2; ++k) {
                    const int ix = k < npd ? 3 + k : nx - 3 - npd + (k - npd);
                    host_out[(size_t)r * nx + ix] = c[(size_t)r * W2 + k];
                }
        } else {
            for (int r = 0; r < rows; ++r)
                for (int k = 0; k < W2; ++k) {
                    const int ix = k < npd ? 3 + k : nx - 3 - npd + (k - npd);
                    c[(size_t)r * W2 + k] = host_in[(size_t)r * nx + ix];
                }
            HIPCHK(hipMemcpy(dp, c.data(), c.size() * sizeof(float), hipMemcpyHostToDevice));
        }
        return MSFDM_OK;
    }
    // bottom strip: solid rows nzs-3-npd .. nzs-4 (after the top strip, rows 3 .. npd+2,
    // on the solid-only CPML grid)
    const int zb0 = s->d.nz_solid - 3 - npd;
    if (host_out) std::fill(host_out, host_out + (size_t)rows * nx, 0.0f);
    if (npd == 0) return MSFDM_OK;
    float *bp = dp;
    if (s->geo.ztop_pml) {
        if (int rc = host_rows(s, host_out ? host_out + 3 * (size_t)nx : (void *)(host_in + 3 * (size_t)nx), dp,
                               P * sizeof(float), nx * sizeof(float), npd, host_out != nullptr))
            return rc;
        bp = dp + (size_t)npd * P;
    }
    return host_rows(s, host_out ? host_out + (size_t)zb0 * nx : (void *)(host_in + (size_t)zb0 * nx), bp,
                     P * sizeof(float), nx * sizeof(float), npd, host_out != nullptr);
}

void free_adjoint(msfdm_solver *s) {
    for (auto &p : s->adj) if (p) { (void)hipFree(p); p = nullptr; }
    if (s->aprod) { (void)hipFree(s->aprod); s->aprod = nullptr; }
    if (s->atiles) { (void)hipFree(s->atiles); s->atiles = nullptr; }
    if (s->gmem) { (void)hipFree(s->gmem); s->gmem = nullptr; }
    for (float *p : s->ckpt) (void)hipFree(p);
    for (float *p : s->seg) (void)hipFree(p);
    s->ckpt.clear();
    s->seg.clear();
}

void free_all(msfdm_solver *s) {
    drop_graph(s);
    for (hipEvent_t e : s->prof.ev) (void)hipEventDestroy(e);
    s->prof.ev.clear();
    free_adjoint(s);
    for (auto &b : s->buf) if (b) { (void)hipFree(b); b = nullptr; }
    for (void *p : {(void *)s->kbuf, s->stage, (void *)s->kif, (void *)s->coefmem, (void *)s->amp, (void *)s->base,
                    (void *)s->rec})
        if (p) (void)hipFree(p);
    if (s->stream) (void)hipStreamDestroy(s->stream);
}

// The batched, adjoint and multi-shot paths serve the fluid-solid grid with
// the pressure-release surface only.
int require_fluid_solid(const msfdm_solver *s) {
    if (s->solid_only)
        return set_error(MSFDM_EINVAL, "solid-only grids (grid.calc_type = 0) support the forward stepper only");
    if (s->geo.fcpml)
        return set_error(MSFDM_EINVAL, "the fluid-solid CPML grid (grid.boundary_type = 1) supports the forward stepper only");
    return MSFDM_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int msfdm_solver_create(const msfdm_model *m, int device, msfdm_solver **out) {
    clear_error();
    if (!m || !out) return set_error(MSFDM_EINVAL, "null argument");
    *out = nullptr;
    if (int rc = validate_desc(m->d)) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(MSFDM_EDEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return set_error(MSFDM_EINVAL, "device index out of range");
    HIPCHK(hipSetDevice(device));

    msfdm_solver *s = new msfdm_solver();
    s->device = device;
    s->d = m->d;
    s->srcx = m->src_x;
    s->srcz = m->src_z;
    s->d.src_x = s->srcx.data();
    s->d.src_z = s->srcz.data();
    s->gap = record_gap(s->d.dt);
    s->solid_only = s->d.calc_type == 0;
    if (s->solid_only) s->d.nz_fluid = 0;
    // solid-only boundary kinds: 2 free top + PML (PANDF), 1 PML on four sides
    // (CPML), 0 Dirichlet: no PML (src/elements.cu:825-827)
    s->top = !s->solid_only ? D::TOP_IFACE : (s->d.boundary_type == 2 ? D::TOP_FREE : D::TOP_PLAIN);
    const bool fcpml = !s->solid_only && s->d.boundary_type == 1;   // top PML in the water
    const bool cpml = s->solid_only && s->d.boundary_type == 1, diri = s->solid_only && s->d.boundary_type == 0;
    auto fail = [&](int rc) { free_all(s); delete s; return rc; };
#define CK(expr) do { hipError_t e_ = (expr); if (e_ != hipSuccess) return fail(set_error(MSFDM_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_))); } while (0)
    CK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));

    const int nx = s->d.nx, nzf = s->d.nz_fluid, nzs = s->d.nz_solid, npd = diri ? 0 : s->d.npd;
    const int nz = s->solid_only ? nzs : nzf + nzs - 1;   // src/elements.cu:87-99
    const long long P = align64(nx);
    D::Geo &g = s->geo;
    g.nx = nx; g.nzf = nzf; g.nzs = nzs; g.P = (int)P; g.npd = npd; g.W2 = 2 * npd;
    g.ztop_pml = cpml ? 1 : 0;
    g.fcpml = fcpml ? 1 : 0;
    long long off = 0;
    for (int f = 0; f < MSFDM_NFIELD; ++f) {
        g.off[f] = off;
        long long n;
        if (f <= MSFDM_FP) n = nzf * P;
        else if (f <= MSFDM_FXEZ) n = (long long)nzf * g.W2;
        else if (f <= MSFDM_SSXZ) n = nzs * P;
        else if (f <= MSFDM_SXEZ) n = (long long)nzs * g.W2;
        else if (f <= MSFDM_SZEZ) n = (long long)(cpml ? 2 : 1) * npd * P;   // solid zeta: bottom (CPML: + top)
        else n = fcpml ? (long long)npd * P : 0;                               // fluid zeta: top strip
        off += align64(n);
    }
    s->buf_floats = off;
    for (int r = 0; r < 4; ++r) {
        CK(hipMalloc(&s->buf[r], off * sizeof(float)));
        CK(hipMemset(s->buf[r], 0, off * sizeof(float)));
    }
    CK(hipMalloc(&s->kif, D::KIF_ROWS * P * sizeof(float)));
    CK(hipMemset(s->kif, 0, D::KIF_ROWS * P * sizeof(float)));
    CK(hipMalloc(&s->base, 64 * sizeof(int)));
    CK(hipMemset(s->base, 0, 64 * sizeof(int)));

    // Coefficients, padded to pitch P (host prep in double where derived).
    const long long nF = align64(nzf * P), nS = align64(nzs * P), nG = align64(nz * P);
    const long long nX = align64(P), nZ = align64(nzs), nZf = align64(nzf);
    const long long total = 2 * nF + 5 * nS + 5 * nG + 3 * nX + 3 * nZ + 3 * nZf + 2 * nX + 6 * nX;
    CK(hipMalloc(&s->coefmem, total * sizeof(float)));
    CK(hipMemset(s->coefmem, 0, total * sizeof(float)));
    std::vector<float> h;
    float *cur = s->coefmem;
    auto upload = [&](const std::vector<float> &v, long long n) -> const float * {
        float *p = cur;
        cur += n;
        const long long cnt = std::min<long long>((long long)v.size(), n);
        if (cnt > 0 && hipMemcpy(p, v.data(), cnt * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            return nullptr;
        return p;
    };
    auto pad = [&](const std::vector<float> &src, int rows, bool invert) {
        std::vector<float> o((size_t)rows * P, 0.0f);
        for (int r = 0; r < rows; ++r)
            for (int x = 0; x < nx; ++x) {
                const float v = src[(size_t)r * nx + x];
                o[(size_t)r * P + x] = invert ? (float)(1.0 / (double)v) : v;
            }
        return o;
    };
    D::Coef &c = s->coef;
    const float *ptrs[32];
    int np = 0;
    ptrs[np++] = c.rhof_inv = upload(pad(m->a[MSFDM_A_RHO_F], nzf, true), nF);
    ptrs[np++] = c.Kf = upload(pad(m->a[MSFDM_A_K], nzf, false), nF);
    ptrs[np++] = c.rhos_inv = upload(pad(m->a[MSFDM_A_RHO_S], nzs, true), nS);
    ptrs[np++] = c.C11 = upload(pad(m->a[MSFDM_A_C11], nzs, false), nS);
    ptrs[np++] = c.C13 = upload(pad(m->a[MSFDM_A_C13], nzs, false), nS);
    ptrs[np++] = c.C33 = upload(pad(m->a[MSFDM_A_C33], nzs, false), nS);
    ptrs[np++] = c.C55 = upload(pad(m->a[MSFDM_A_C55], nzs, false), nS);
    ptrs[np++] = c.J = upload(pad(m->a[MSFDM_A_J], nz, false), nG);
    ptrs[np++] = c.xix = upload(pad(m->a[MSFDM_A_XI_X], nz, false), nG);
    ptrs[np++] = c.xiz = upload(pad(m->a[MSFDM_A_XI_Z], nz, false), nG);
    ptrs[np++] = c.zex = upload(pad(m->a[MSFDM_A_ZE_X], nz, false), nG);
    ptrs[np++] = c.zez = upload(pad(m->a[MSFDM_A_ZE_Z], nz, false), nG);
    // PML (src/elements.cu:117-124, 824-874): alpha0 = 2 pi fpeak, p_beta = 2, p_alpha = 1.
    const double damp0 = -(s->d.npow + 1) * (double)s->d.vp_max * std::log((double)s->d.rc) / (2.0 * (double)s->d.thickness);
    const double alpha0 = 2.0 * M_PI * (double)s->d.fpeak;
    std::vector<float> pa, pb, pc;
    pml_profile(nx, npd, s->d.npow, true, true, damp0, s->d.beta0, alpha0, pa, pb, pc);
    ptrs[np++] = c.pxa = upload(pa, nX);
    ptrs[np++] = c.pxb = upload(pb, nX);
    ptrs[np++] = c.pxc = upload(pc, nX);
    pml_profile(nzs, npd, s->d.npow, cpml, true, damp0, s->d.beta0, alpha0, pa, pb, pc);
    ptrs[np++] = c.pza = upload(pa, nZ);
    ptrs[np++] = c.pzb = upload(pb, nZ);
    ptrs[np++] = c.pzc = upload(pc, nZ);
    // fluid zeta PML: the top strip of the CPML grid (src/elements.cu:870-873)
    pml_profile(nzf, fcpml ? npd : 0, s->d.npow, true, false, damp0, s->d.beta0, alpha0, pa, pb, pc);
    ptrs[np++] = c.pzaf = upload(pa, nZf);
    ptrs[np++] = c.pzbf = upload(pb, nZf);
    ptrs[np++] = c.pzcf = upload(pc, nZf);
    // Interface normal (src/elements.cu:892-905; zero on a solid-only grid),
    // computed in float as the reference does, so that it holds the same
    // float2 bit for bit: every interface projection applies its n1^2 + n2^2.
    std::vector<float> nif(2 * (size_t)nx, 0.0f);
    for (int ix = 0; ix < nx && !s->solid_only; ++ix) {
        const size_t gi = (size_t)(nzf - 1) * nx + ix;
        const float zx = m->a[MSFDM_A_ZE_X][gi], zz = m->a[MSFDM_A_ZE_Z][gi];
        const float n = std::sqrt(zx * zx + zz * zz);
        nif[2 * ix] = zx / n;
        nif[2 * ix + 1] = zz / n;
    }
    // vertical grid columns: xi_z == 0 and xi_x constant everywhere
    {
        const std::vector<float> &X = m->a[MSFDM_A_XI_X], &Z = m->a[MSFDM_A_XI_Z];
        bool vert = !X.empty();
        for (size_t i = 0; i < X.size() && vert; ++i) vert = (Z[i] == 0.0f) && (X[i] == X[0]);
        c.vert = vert ? 1 : 0;
        c.xix0 = X.empty() ? 0.0f : X[0];
    }
    const float *nifp = upload(nif, 2 * nX);
    c.nif = reinterpret_cast<const float2 *>(nifp);
    ptrs[np++] = nifp;
    // Free-surface ALM coefficients of a solid-only grid (src/elements.cu:907-977):
    // double arithmetic on the float model values of row 3, rounded to float.
    std::vector<float> alm(6 * (size_t)P, 0.0f);
    for (int ix = 0; ix < nx && s->top == D::TOP_FREE; ++ix) {
        const size_t i = (size_t)3 * nx + ix;
        const double xx = m->a[MSFDM_A_XI_X][i], xz = m->a[MSFDM_A_XI_Z][i];
        const double zx = m->a[MSFDM_A_ZE_X][i], zz = m->a[MSFDM_A_ZE_Z][i];
        const double c11 = m->a[MSFDM_A_C11][i], c13 = m->a[MSFDM_A_C13][i];
        const double c33 = m->a[MSFDM_A_C33][i], c55 = m->a[MSFDM_A_C55][i];
        const double dz11 = c11 * zx * zx + c55 * zz * zz, dz12 = (c13 + c55) * zx * zz;
        const double dz22 = c55 * zx * zx + c33 * zz * zz;
        const double dx11 = c11 * zx * xx + c55 * zz * xz, dx12 = c13 * zx * xz + c55 * zz * xx;
        const double dx21 = c13 * zz * xx + c55 * zx * xz, dx22 = c55 * zx * xx + c33 * zz * xz;
        const double det = dz11 * dz22 - dz12 * dz12;
        double m11 = 0, m12 = 0, m21 = 0, m22 = 0;
        if (std::abs(det) >= 1e-14) {
            const double i11 = dz22 / det, i12 = -dz12 / det, i22 = dz11 / det;
            m11 = i11 * dx11 + i12 * dx21;
            m12 = i11 * dx12 + i12 * dx22;
            m21 = i12 * dx11 + i22 * dx21;
            m22 = i12 * dx12 + i22 * dx22;
        }
        const double al[6] = {xx - zx * m11, -zz * m21, xz - zz * m11 - zx * m21,
                              -zx * m12, xz - zz * m22, xx - zz * m12 - zx * m22};
        for (int k = 0; k < 6; ++k) alm[(size_t)k * P + ix] = (float)al[k];
    }
    ptrs[np++] = c.alm = upload(alm, 6 * nX);
    for (int i = 0; i < np; ++i)
        if (!ptrs[i]) return fail(set_error(MSFDM_EDEVICE, "coefficient upload failed"));

    s->ctl.base = s->base;
    s->ctl.gap = s->gap;
    s->ctl.srcx = s->srcx.empty() ? -100 : s->srcx[0];
    s->ctl.srcz = s->srcz.empty() ? -100 : s->srcz[0];
    // solid-only receiver row (src/main.cu:303-311): 3 (PANDF), 30 (CPML), none (Dirichlet)
    s->ctl.rec_row = !s->solid_only ? -1 : (s->d.boundary_type == 2 ? 3 : (s->d.boundary_type == 1 ? 30 : -1));
    if (int rc = grow_tables(s, std::max(s->d.nt, 1))) return fail(rc);

    // Tile heights: small grids need more blocks to fill 256 CUs.
    const long long cells = (long long)nx * nz;
    s->tz = 8;

    // Algorithmic bytes per step: 13 state floats per field (stage 1: read U,
    // write T2; stages 2-3: read stage state + U, write next; stage 4: read
    // T4, U, T2, T3, write U) + 4 coefficient floats (fluid 3 fields /
    // 6 coefficients, solid 5 / 9) per updated cell, plus 13 floats x 4
    // memory fields per active PML direction.
    // Updated rows: fluid 3..nzf-1, solid 0..nzs-4 (solid only: 3..nzs-4).
    const int srows = s->solid_only ? nzs - 6 : nzs - 3;
    const double cf = s->solid_only ? 0.0 : (double)(nzf - 3) * (nx - 6), cs = (double)srows * (nx - 6);
    const double pmlf = s->solid_only ? 0.0 : (double)(nzf - 3) * 2 * npd + (fcpml ? (double)npd * (nx - 6) : 0.0);
    const double pmls = (double)srows * 2 * npd + (double)(cpml ? 2 : 1) * npd * (nx - 6);
    // Vertical-column grids read neither xi_x nor xi_z: 2 coefficients fewer.
    const double ncf = c.vert ? 4 : 6, ncs = c.vert ? 7 : 9;
    s->bytes_step = 4.0 * (cf * (13 * 3 + 4 * ncf) + cs * (13 * 5 + 4 * ncs) + (pmlf + pmls) * 13 * 4);
    // stage-2 launch: 3 state floats per field + coefficients (+3 per PML memory field)
    s->bytes_kernel[0] = 4.0 * (cf * (3 * 3 + ncf) + cs * (3 * 5 + ncs) + (pmlf + pmls) * 3 * 4);
    // stage-4 launch: 5 state floats per field (read T4, U^n, T2, T3, write U) + coefficients
    s->bytes_kernel[1] = 4.0 * (cf * (5 * 3 + ncf) + cs * (5 * 5 + ncs) + (pmlf + pmls) * 5 * 4);
    // interface / filter / record launch: 4 k sets + U^n, normal, output rows
    s->bytes_kernel[2] = s->solid_only ? 0.0 : 4.0 * (double)(nx - 6) * (8 * 5 + 2 + 8);
    CK(hipStreamSynchronize(s->stream));
#undef CK
    *out = s;
    return MSFDM_OK;
}

extern "C" void msfdm_solver_destroy(msfdm_solver *s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
}

extern "C" int msfdm_reset_shot(msfdm_solver *s, int isrc) {
    clear_error();
    if (!s) return set_error(MSFDM_EINVAL, "null solver");
    if (isrc < 0 || (isrc >= (int)s->srcx.size() && !(s->srcx.empty() && isrc == 0)))
        return set_error(MSFDM_EINVAL, "shot index out of range");
    HIPCHK(hipSetDevice(s->device));
    for (int r = 0; r < 4; ++r) HIPCHK(hipMemsetAsync(s->buf[r], 0, s->buf_floats * sizeof(float), s->stream));
    HIPCHK(hipMemsetAsync(s->base, 0, sizeof(int), s->stream));
    HIPCHK(hipMemsetAsync(s->rec, 0, 3 * (size_t)s->rec_cap * s->d.nx * sizeof(float), s->stream));
    const int nsx = s->srcx.empty() ? -100 : s->srcx[isrc], nsz = s->srcz.empty() ? -100 : s->srcz[isrc];
    if (nsx != s->ctl.srcx || nsz != s->ctl.srcz) drop_graph(s);
    s->ctl.srcx = nsx;
    s->ctl.srcz = nsz;
    s->isrc = isrc;
    s->host_it = 0;
    s->nrec_done = 0;
    HIPCHK(hipStreamSynchronize(s->stream));
    return MSFDM_OK;
}

extern "C" int msfdm_run(msfdm_solver *s, int nsteps) {
    clear_error();
    if (!s) return set_error(MSFDM_EINVAL, "null solver");
    if (nsteps < 0) return set_error(MSFDM_EINVAL, "negative step count");
    HIPCHK(hipSetDevice(s->device));
    if (int rc = grow_tables(s, s->host_it + nsteps)) return rc;
    int left = nsteps;
    while (left > 0) {
        if (s->use_graph && left >= GRAPH_STEPS && (s->host_it & 1) == 0) {
            if (!s->gexec)
                if (int rc = build_graph(s)) return rc;
            HIPCHK(hipGraphLaunch(s->gexec, s->stream));
            s->host_it += GRAPH_STEPS;
            left -= GRAPH_STEPS;
        } else {
            const int n = (s->use_graph && (s->host_it & 1)) ? 1 : std::min(left, GRAPH_STEPS);
            enqueue_steps(s, s->host_it, n);
            HIPCHK(hipGetLastError());
            s->host_it += n;
            left -= n;
        }
    }
    s->nrec_done = s->host_it > 0 ? (s->host_it - 1) / s->gap + 1 : 0;
    if (s->solid_only && s->ctl.rec_row < 0) s->nrec_done = 0;   // Dirichlet grids record nothing
    return MSFDM_OK;
}

extern "C" int msfdm_synchronize(msfdm_solver *s) {
    if (!s) return set_error(MSFDM_EINVAL, "null solver");
    HIPCHK(hipStreamSynchronize(s->stream));
    return MSFDM_OK;
}

extern "C" int msfdm_step_index(const msfdm_solver *s, int *it) {
    if (!s || !it) return set_error(MSFDM_EINVAL, "null argument");
    *it = s->host_it;
    return MSFDM_OK;
}

extern "C" int msfdm_num_records(const msfdm_solver *s, int *n) {
    if (!s || !n) return set_error(MSFDM_EINVAL, "null argument");
    *n = s->nrec_done;
    return MSFDM_OK;
}

extern "C" int msfdm_get_records(msfdm_solver *s, int first, int count, float *vx, float *vz, float *p) {
    clear_error();
    if (!s) return set_error(MSFDM_EINVAL, "null solver");
    if (first < 0 || count < 0 || first + count > s->nrec_done) return set_error(MSFDM_EINVAL, "record range out of bounds");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    const size_t nx = (size_t)s->d.nx, bytes = (size_t)count * nx * sizeof(float);
    float *outs[3] = {vx, vz, p};
    for (int k = 0; k < 3; ++k)
        if (outs[k] && count > 0)
            HIPCHK(hipMemcpy(outs[k], s->rec + ((size_t)k * s->rec_cap + first) * nx, bytes, hipMemcpyDeviceToHost));
    return MSFDM_OK;
}

extern "C" int msfdm_get_field(msfdm_solver *s, int field, float *out) {
    clear_error();
    if (!s || !out) return set_error(MSFDM_EINVAL, "null argument");
    if (field < 0 || field >= MSFDM_NFIELD) return set_error(MSFDM_EINVAL, "bad field id");
    HIPCHK(hipSetDevice(s->device));
    return copy_field(s, s->buf[R_U], field, out, nullptr);
}

extern "C" int msfdm_set_field(msfdm_solver *s, int field, const float *in) {
    clear_error();
    if (!s || !in) return set_error(MSFDM_EINVAL, "null argument");
    if (field < 0 || field >= MSFDM_NFIELD) return set_error(MSFDM_EINVAL, "bad field id");
    HIPCHK(hipSetDevice(s->device));
    // The stage buffers keep the same values in cells the stepper never
    // updates (frozen bands), as the reference's RK4_reset copies do.
    for (int r : {R_U, R_T2, R_T3, R_T4})
        if (int rc = copy_field(s, s->buf[r], field, nullptr, in)) return rc;
    return MSFDM_OK;
}

extern "C" int msfdm_eval_rhs(msfdm_solver *s, int dir, float amp) {
    clear_error();
    if (!s) return set_error(MSFDM_EINVAL, "null solver");
    if (dir != 0 && dir != 1) return set_error(MSFDM_EINVAL, "dir must be 0 (F) or 1 (B)");
    HIPCHK(hipSetDevice(s->device));
    if (!s->kbuf) {
        HIPCHK(hipMalloc(&s->kbuf, s->buf_floats * sizeof(float)));
    }
    HIPCHK(hipMemsetAsync(s->kbuf, 0, s->buf_floats * sizeof(float), s->stream));
    float *U = s->buf[R_U];
    if (dir == 0) launch_stage<0, 0>(s, U, s->kbuf, 0, 0.f, amp, 0.f);
    else launch_stage<1, 0>(s, U, s->kbuf, 0, 0.f, amp, 0.f);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return MSFDM_OK;
}

extern "C" int msfdm_get_rhs(msfdm_solver *s, int field, float *out) {
    clear_error();
    if (!s || !out) return set_error(MSFDM_EINVAL, "null argument");
    if (!s->kbuf) return set_error(MSFDM_ESTATE, "msfdm_eval_rhs has not been called");
    if (field < 0 || field >= MSFDM_NFIELD) return set_error(MSFDM_EINVAL, "bad field id");
    HIPCHK(hipSetDevice(s->device));
    return copy_field(s, s->kbuf, field, out, nullptr);
}

extern "C" int msfdm_set_tile_rows(msfdm_solver *s, int rows) {
    if (!s) return set_error(MSFDM_EINVAL, "null solver");
    if (rows != 8 && rows != 16) return set_error(MSFDM_EINVAL, "tile rows must be 8 or 16");
    if (rows != s->tz) {
        HIPCHK(hipStreamSynchronize(s->stream));
        drop_graph(s);
        s->tz = rows;
    }
    return MSFDM_OK;
}

extern "C" int msfdm_get_tile_rows(const msfdm_solver *s, int *rows) {
    if (!s || !rows) return set_error(MSFDM_EINVAL, "null argument");
    *rows = s->tz;
    return MSFDM_OK;
}

// Occupancy policies with a kernel of their own; a grid without vertical
// columns, or the fluid-solid CPML grid, has the default kernels only.
static bool occupancy_variants(const msfdm_solver *s) { return s->coef.vert && !s->geo.fcpml; }

extern "C" int msfdm_set_occupancy(msfdm_solver *s, int policy) {
    clear_error();
    if (!s) return set_error(MSFDM_EINVAL, "null solver");
    if (policy != 0 && policy != 1) return set_error(MSFDM_EINVAL, "occupancy policy must be 0 or 1");
    if (policy == 1 && !occupancy_variants(s))
        return set_error(MSFDM_EINVAL, "occupancy policy 1 needs vertical grid columns and no fluid CPML");
    if (policy != s->occ) {
        HIPCHK(hipStreamSynchronize(s->stream));
        drop_graph(s);
        s->occ = policy;
    }
    return MSFDM_OK;
}

extern "C" int msfdm_get_occupancy(const msfdm_solver *s, int *policy) {
    if (!s || !policy) return set_error(MSFDM_EINVAL, "null argument");
    *policy = s->occ;
    return MSFDM_OK;
}

extern "C" int msfdm_set_adjoint_launch(msfdm_solver *s, int mode) {
    clear_error();
    if (!s) return set_error(MSFDM_EINVAL, "null solver");
    if (mode < -1 || mode > 1) return set_error(MSFDM_EINVAL, "adjoint launch mode must be -1, 0 or 1");
    s->adj_launch = mode;
    return MSFDM_OK;
}

extern "C" int msfdm_get_adjoint_launch(const msfdm_solver *s, int *mode) {
    if (!s || !mode) return set_error(MSFDM_EINVAL, "null argument");
    *mode = s->adj_launch;
    return MSFDM_OK;
}

extern "C" int msfdm_set_graph(msfdm_solver *s, int enable) {
    if (!s) return set_error(MSFDM_EINVAL, "null solver");
    s->use_graph = enable != 0;
    if (!s->use_graph) drop_graph(s);
    return MSFDM_OK;
}

extern "C" int msfdm_time_kernel(msfdm_solver *s, int which, int reps, float *ms) {
    clear_error();
    if (!s || !ms || reps <= 0) return set_error(MSFDM_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(s->device));
    if (int rc = grow_tables(s, s->host_it + 1)) return rc;   // the stage reads amplitude row host_it
    float *U = s->buf[R_U], *T2 = s->buf[R_T2], *T3 = s->buf[R_T3], *T4 = s->buf[R_T4];
    const float dt = s->d.dt;
    if (!s->kbuf) {   // timing output (the state buffers are left as they are)
        HIPCHK(hipMalloc(&s->kbuf, s->buf_floats * sizeof(float)));
        HIPCHK(hipMemset(s->kbuf, 0, s->buf_floats * sizeof(float)));
    }
    float *X = s->kbuf;
    if (s->solid_only && which == 2) return set_error(MSFDM_EINVAL, "a solid-only grid has no interface kernel");
    // no record slot is written while timing (the solid-only stage-4 kernel records row 3)
    struct CtlRestore {
        msfdm_solver *s;
        D::Ctl c;
        ~CtlRestore() { s->ctl = c; }
    } restore{s, s->ctl};
    s->ctl.gap = 0x7fffffff;
    auto one = [&]() {
        switch (which) {
            case 0:   // RK stage 2: reads T2 (+ U^n), writes the next stage state
                if (s->tz == 8) launch_media<0, 2, 8>(s, T2, U, T2, T3, X, 0, dt / 2, 0.f, dt / 2);
                else launch_media<0, 2, 16>(s, T2, U, T2, T3, X, 0, dt / 2, 0.f, dt / 2);
                break;
            case 1:   // RK stage 4: reads T4, U^n, T2, T3, writes U'
                if (s->tz == 8) launch_media<1, 4, 8>(s, T4, U, T2, T3, X, 0, dt / 6, 0.f, dt);
                else launch_media<1, 4, 16>(s, T4, U, T2, T3, X, 0, dt / 6, 0.f, dt);
                break;
            default: {
                D::Ctl quiet = s->ctl;   // no record slot is written while timing
                quiet.gap = 0x7fffffff;
                D::iface_final<1><<<(s->d.nx + D::FO - 1) / D::FO, D::FB, 0, s->stream>>>(s->geo, s->coef, quiet, X,
                                                                                           s->kif, dt / 2, dt, dt / 6, 0);
                break;
            }
        }
    };
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    one();   // warm-up
    HIPCHK(hipEventRecord(e0, s->stream));
    for (int i = 0; i < reps; ++i) one();
    HIPCHK(hipEventRecord(e1, s->stream));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / reps;
    return MSFDM_OK;
}

extern "C" int msfdm_time_run(msfdm_solver *s, int nsteps, float *ms) {
    clear_error();
    if (!s || !ms) return set_error(MSFDM_EINVAL, "null argument");
    HIPCHK(hipSetDevice(s->device));
    if (int rc = grow_tables(s, s->host_it + nsteps)) return rc;
    if (s->use_graph && !s->gexec && (s->host_it & 1) == 0 && nsteps >= GRAPH_STEPS)
        if (int rc = build_graph(s)) return rc;   // instantiate outside the timed region
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, s->stream));
    if (int rc = msfdm_run(s, nsteps)) return rc;
    HIPCHK(hipEventRecord(e1, s->stream));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t;
    return MSFDM_OK;
}

// Streaming probe: o[i] = in_0[i] + ... + in_{R-1}[i] over float4 elements,
// four per thread, every load issued before the first add; the inputs are R
// consecutive streams of n4 float4 in one allocation.
template <int R>
__global__ __launch_bounds__(256) void stream_probe_kernel(const float4 *__restrict__ a, long long n4,
                                                           float4 *__restrict__ o) {
    const long long base = (long long)blockIdx.x * 1024 + threadIdx.x;
    float4 v[R][4];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const long long i = base + k * 256;
            v[r][k] = i < n4 ? a[r * n4 + i] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const long long i = base + k * 256;
        if (i >= n4) continue;
        float4 t = v[0][k];
#pragma unroll
        for (int r = 1; r < R; ++r) {
            t.x += v[r][k].x; t.y += v[r][k].y; t.z += v[r][k].z; t.w += v[r][k].w;
        }
        o[i] = t;
    }
}

namespace {
struct ProbeRes {
    float4 *in = nullptr, *out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~ProbeRes() {
        if (st) (void)hipStreamSynchronize(st);
        if (in) (void)hipFree(in);
        if (out) (void)hipFree(out);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
    }
};

template <int R>
void stream_probe_launch(const ProbeRes &p, long long n4) {
    const unsigned blocks = (unsigned)((n4 + 1023) / 1024);
    stream_probe_kernel<R><<<blocks, 256, 0, p.st>>>(p.in, n4, p.out);
}
}  // namespace

extern "C" int msfdm_stream_probe(int device, double mbytes, int reads, int reps, double *tbps) {
    clear_error();
    if (!tbps || reads < 1 || reads > 4 || reps < 1 || !(mbytes > 0.0) || mbytes > 65536.0)
        return set_error(MSFDM_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(device));
    ProbeRes p;
    const long long n4 = std::max(1LL, (long long)(mbytes * 1e6 / 16.0));
    HIPCHK(hipMalloc(&p.in, (size_t)reads * n4 * sizeof(float4)));
    HIPCHK(hipMalloc(&p.out, (size_t)n4 * sizeof(float4)));
    HIPCHK(hipStreamCreateWithFlags(&p.st, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&p.e0));
    HIPCHK(hipEventCreate(&p.e1));
    // non-zero data (1.0f): kernels on zero-filled buffers run measurably faster
    HIPCHK(hipMemsetD32Async((hipDeviceptr_t)p.in, 0x3f800000, (size_t)reads * n4 * 4, p.st));
    auto launch = [&]() {
        switch (reads) {
            case 1: stream_probe_launch<1>(p, n4); break;
            case 2: stream_probe_launch<2>(p, n4); break;
            case 3: stream_probe_launch<3>(p, n4); break;
            default: stream_probe_launch<4>(p, n4); break;
        }
    };
    for (int i = 0; i < 3; ++i) launch();
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(p.e0, p.st));
    for (int i = 0; i < reps; ++i) launch();
    HIPCHK(hipEventRecord(p.e1, p.st));
    HIPCHK(hipEventSynchronize(p.e1));
    HIPCHK(hipGetLastError());
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, p.e0, p.e1));
    *tbps = (double)(reads + 1) * n4 * sizeof(float4) * reps / (ms * 1e-3) / 1e12;
    return MSFDM_OK;
}

// The launch sequence of msfdm_run (GRAPH_STEPS-step sequences where msfdm_run
// replays its graph, single steps at odd step indices) with direct launches;
// the launches of the classes in class_mask are stamped by their own dispatch.
// kernel_ms[c] / launches[c]: total milliseconds and count of the timed
// launches of kernel class c (RK stages 1..4, interface finish, other);
// wall_ms: the whole run between two stream events.
extern "C" int msfdm_profile_run(msfdm_solver *s, int nsteps, int class_mask, float *kernel_ms, int *launches,
                                 float *wall_ms) {
    clear_error();
    if (!s || nsteps < 0) return set_error(MSFDM_EINVAL, "bad argument");
    s->prof.mask = class_mask;
    HIPCHK(hipSetDevice(s->device));
    if (int rc = grow_tables(s, s->host_it + nsteps)) return rc;
    hipEvent_t w0, w1;
    HIPCHK(hipEventCreate(&w0));
    HIPCHK(hipEventCreate(&w1));
    s->prof.cls.clear();
    s->prof.on = true;
    (void)hipEventRecord(w0, s->stream);
    int left = nsteps;
    while (left > 0) {
        int n = std::min(left, GRAPH_STEPS);
        if (s->use_graph) n = (s->host_it & 1) ? 1 : (left >= GRAPH_STEPS ? GRAPH_STEPS : std::min(left, GRAPH_STEPS));
        enqueue_steps(s, s->host_it, n);
        s->host_it += n;
        left -= n;
    }
    (void)hipEventRecord(w1, s->stream);
    s->prof.on = false;
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventSynchronize(w1));
    float wall = 0.f;
    HIPCHK(hipEventElapsedTime(&wall, w0, w1));
    (void)hipEventDestroy(w0);
    (void)hipEventDestroy(w1);
    double tot[NKCLASS] = {0, 0, 0, 0, 0, 0};
    int cnt[NKCLASS] = {0, 0, 0, 0, 0, 0};
    for (size_t i = 0; i < s->prof.cls.size(); ++i) {
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, s->prof.ev[2 * i], s->prof.ev[2 * i + 1]));
        tot[s->prof.cls[i]] += t;
        cnt[s->prof.cls[i]] += 1;
    }
    for (int c = 0; c < NKCLASS; ++c) {
        if (kernel_ms) kernel_ms[c] = (float)tot[c];
        if (launches) launches[c] = cnt[c];
    }
    if (wall_ms) *wall_ms = wall;
    s->nrec_done = s->host_it > 0 ? (s->host_it - 1) / s->gap + 1 : 0;
    if (s->solid_only && s->ctl.rec_row < 0) s->nrec_done = 0;
    return MSFDM_OK;
}

extern "C" int msfdm_kernel_bytes(const msfdm_solver *s, int which, double *bytes) {
    if (!s || !bytes) return set_error(MSFDM_EINVAL, "null argument");
    if (which < 0 || which > 2) return set_error(MSFDM_EINVAL, "bad kernel id");
    *bytes = s->bytes_kernel[which];
    return MSFDM_OK;
}

extern "C" int msfdm_bytes_per_step(const msfdm_solver *s, double *bytes) {
    if (!s || !bytes) return set_error(MSFDM_EINVAL, "null argument");
    *bytes = s->bytes_step;
    return MSFDM_OK;
}

// ---------------------------------------------------------------------------
// Batched shots: independent shots (own grid, model, source and buffers) that
// share dt are stepped together, one launch per RK stage for all of them, the
// time loop replayed from a hipGraph.  Each shot is a full msfdm_solver whose
// device state the batch drives; records and fields are read per shot.
// ---------------------------------------------------------------------------
struct msfdm_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<msfdm_solver *> shots;
    std::vector<D::Shot> host_shots;
    D::Shot *d_shots = nullptr;
    int4 *d_tiles[2] = {nullptr, nullptr};   // tile tables for 8- and 16-row tiles
    int ntiles[2] = {0, 0};
    int tz = 16;
    int occ = 0;      // occupancy policy of the batched stage kernels (msfdm_batch_set_occupancy)
    int maxnx = 0;
    int *base = nullptr;
    int host_it = 0;
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    float dt = 0.f;
};

namespace {

void batch_drop_graph(msfdm_batch *b) {
    if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
}

// Refresh the device shot table (table pointers change when a solver grows its
// amplitude / record tables).
int batch_upload(msfdm_batch *b) {
    std::vector<D::Shot> h(b->shots.size());
    for (size_t i = 0; i < b->shots.size(); ++i) {
        msfdm_solver *s = b->shots[i];
        D::Shot &t = h[i];
        t.g = s->geo;
        t.c = s->coef;
        t.ctl = s->ctl;
        t.ctl.base = b->base;
        for (int r = 0; r < 4; ++r) t.buf[r] = s->buf[r];
        t.buf[SHOT_UOUT] = s->buf[R_U];   // in place
        t.kif = s->kif;
    }
    if (h.size() == b->host_shots.size() && std::memcmp(h.data(), b->host_shots.data(), h.size() * sizeof(D::Shot)) == 0)
        return MSFDM_OK;
    HIPCHK(hipStreamSynchronize(b->stream));
    b->host_shots = h;
    HIPCHK(hipMemcpy(b->d_shots, b->host_shots.data(), h.size() * sizeof(D::Shot), hipMemcpyHostToDevice));
    batch_drop_graph(b);
    return MSFDM_OK;
}

// What a batched launch sequence needs: the device shot table, the tile tables
// (8- and 16-row tiles), the tile height in use and the widest grid.
struct BatchView {
    const D::Shot *shots;
    const int4 *tiles[2];
    int ntiles[2];
    int tz, ns, maxnx;
    int *base;
    float dt;
    hipStream_t stream;
    bool vert;   // every shot has vertical grid columns: the MET = 1 kernels
    int occ;     // occupancy policy (1: the _w5 kernels, vertical columns only)
};
bool all_vertical(const std::vector<msfdm_solver *> &shots) {
    for (const msfdm_solver *s : shots)
        if (!s->coef.vert) return false;
    return true;
}
BatchView view_of(const msfdm_batch *b) {
    return BatchView{b->d_shots, {b->d_tiles[0], b->d_tiles[1]}, {b->ntiles[0], b->ntiles[1]}, b->tz,
                     (int)b->shots.size(), b->maxnx, b->base, b->dt, b->stream, all_vertical(b->shots), b->occ};
}

template <int DIR, int STAGE, int TZ, int FIN>
void batch_launch_media(const BatchView &b, int rS, int rO, int j, float ct, float cprev) {
    const int k = TZ == 8 ? 0 : 1;
    const int grid = 8 * ((b.ntiles[k] + 7) / 8);
    if (b.vert && b.occ == 1)
        D::batch_stage_kernel_w5<DIR, STAGE, TZ, FIN><<<grid, D::NTHR, 0, b.stream>>>(b.shots, b.tiles[k], b.ntiles[k],
                                                                                       rS, rO, j, ct, cprev);
    else if (b.vert)
        D::batch_stage_kernel<DIR, STAGE, TZ, 1, FIN><<<grid, D::NTHR, 0, b.stream>>>(b.shots, b.tiles[k], b.ntiles[k],
                                                                                       rS, rO, j, ct, cprev);
    else
        D::batch_stage_kernel<DIR, STAGE, TZ, 2, FIN><<<grid, D::NTHR, 0, b.stream>>>(b.shots, b.tiles[k], b.ntiles[k],
                                                                                       rS, rO, j, ct, cprev);
}
// As launch_stage: FIN, a stage 1 that finishes the previous step (cprev = dt);
// fin_next, stage 4 leaves its interface rows to the next step's stage 1.
template <int DIR, int STAGE, int FIN = 0>
void batch_launch_stage(const BatchView &b, int rS, int rO, int j, float ct, float cprev, bool fin_next = false) {
    if (b.tz == 8) batch_launch_media<DIR, STAGE, 8, FIN>(b, rS, rO, j, ct, cprev);
    else batch_launch_media<DIR, STAGE, 16, FIN>(b, rS, rO, j, ct, cprev);
    if constexpr (STAGE == 4) {
        if (fin_next) return;
        D::batch_iface_final<DIR><<<dim3((b.maxnx + D::FO - 1) / D::FO, b.ns), D::FB, 0, b.stream>>>(b.shots, b.dt / 2,
                                                                                                   b.dt, ct, j);
    }
}
template <int D1>
void batch_step_dir(const BatchView &b, int j, bool first = true, bool last = true) {
    constexpr int D2 = 1 - D1;
    const float dt = b.dt;
    if (first) batch_launch_stage<D1, 1>(b, R_U, R_T2, j, dt / 2, 0.f);
    else batch_launch_stage<D1, 1, 1>(b, R_U, R_T2, j, dt / 2, dt);
    batch_launch_stage<D2, 2>(b, R_T2, R_T3, j, dt / 2, dt / 2);
    batch_launch_stage<D1, 3>(b, R_T3, R_T4, j, dt, dt / 2);
    batch_launch_stage<D2, 4>(b, R_T4, SHOT_UOUT, j, dt / 6, dt, !last);
}
void batch_enqueue(const BatchView &b, int it0, int n) {
    for (int j = 0; j < n; ++j) {
        if ((it0 + j) & 1) batch_step_dir<1>(b, j, j == 0, j == n - 1);
        else batch_step_dir<0>(b, j, j == 0, j == n - 1);
    }
    D::advance_base<<<1, 64, 0, b.stream>>>(b.base, n);
}

void batch_free(msfdm_batch *b) {
    batch_drop_graph(b);
    for (msfdm_solver *s : b->shots) msfdm_solver_destroy(s);
    for (void *p : {(void *)b->d_shots, (void *)b->d_tiles[0], (void *)b->d_tiles[1], (void *)b->base})
        if (p) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

}  // namespace

extern "C" int msfdm_batch_create(const msfdm_model *const *models, const int *shot_index, int nshots, int device,
                                  msfdm_batch **out) {
    clear_error();
    if (!models || !out || nshots <= 0) return set_error(MSFDM_EINVAL, "bad argument");
    *out = nullptr;
    for (int i = 0; i < nshots; ++i) {
        if (models[i] && models[i]->d.calc_type == 0)
            return set_error(MSFDM_EINVAL, "solid-only grids (grid.calc_type = 0) support the forward stepper only");
        if (models[i] && models[i]->d.boundary_type != 2)
            return set_error(MSFDM_EINVAL, "the fluid-solid CPML grid (grid.boundary_type = 1) supports the forward stepper only");
    }
    msfdm_batch *b = new msfdm_batch();
    b->device = device;
    auto fail = [&](int rc) { std::string msg = msfdm_last_error(); batch_free(b); delete b; return set_error(rc, msg); };
    for (int i = 0; i < nshots; ++i) {
        if (!models[i]) return fail(set_error(MSFDM_EINVAL, "null model " + std::to_string(i)));
        if (i > 0 && models[i]->d.dt != models[0]->d.dt)
            return fail(set_error(MSFDM_EINVAL, "all shots of a batch must share dt"));
        msfdm_solver *s = nullptr;
        if (int rc = msfdm_solver_create(models[i], device, &s)) return fail(rc);
        b->shots.push_back(s);
        const int isrc = shot_index ? shot_index[i] : 0;
        if (int rc = msfdm_reset_shot(s, isrc)) return fail(rc);
        s->use_graph = false;
        b->maxnx = std::max(b->maxnx, s->d.nx);
    }
    b->dt = models[0]->d.dt;
#define BCK(expr) do { hipError_t e_ = (expr); if (e_ != hipSuccess) return fail(set_error(MSFDM_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_))); } while (0)
    BCK(hipSetDevice(device));
    BCK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    BCK(hipMalloc(&b->base, 64 * sizeof(int)));
    BCK(hipMemset(b->base, 0, 64 * sizeof(int)));
    BCK(hipMalloc(&b->d_shots, nshots * sizeof(D::Shot)));
    for (int k = 0; k < 2; ++k) {
        const int tz = k == 0 ? 8 : 16;
        std::vector<int4> tiles;
        for (int i = 0; i < nshots; ++i) {
            const msfdm_solver *s = b->shots[i];
            const int nbx = (s->d.nx - 4) / D::TX + 1;
            const int nbf = (s->d.nz_fluid - 3 + tz - 1) / tz, nbs = (s->d.nz_solid - 3 + tz - 1) / tz;
            for (int band = 0; band < nbf + nbs; ++band)
                for (int bx = 0; bx < nbx; ++bx) tiles.push_back(make_int4(i, band, bx, nbf));
        }
        b->ntiles[k] = (int)tiles.size();
        BCK(hipMalloc(&b->d_tiles[k], tiles.size() * sizeof(int4)));
        BCK(hipMemcpy(b->d_tiles[k], tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
#undef BCK
    long long cells = 0;
    for (auto *s : b->shots) cells += (long long)s->d.nx * (s->d.nz_fluid + s->d.nz_solid - 1);
    b->tz = 8;
    if (int rc = batch_upload(b)) return fail(rc);
    *out = b;
    return MSFDM_OK;
}

extern "C" void msfdm_batch_destroy(msfdm_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_free(b);
    delete b;
}

extern "C" int msfdm_batch_reset(msfdm_batch *b) {
    clear_error();
    if (!b) return set_error(MSFDM_EINVAL, "null batch");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (msfdm_solver *s : b->shots)
        if (int rc = msfdm_reset_shot(s, s->isrc)) return rc;
    HIPCHK(hipMemset(b->base, 0, sizeof(int)));
    b->host_it = 0;
    return batch_upload(b);
}

extern "C" int msfdm_batch_run(msfdm_batch *b, int nsteps) {
    clear_error();
    if (!b) return set_error(MSFDM_EINVAL, "null batch");
    if (nsteps < 0) return set_error(MSFDM_EINVAL, "negative step count");
    HIPCHK(hipSetDevice(b->device));
    bool grow = false;
    for (msfdm_solver *s : b->shots)
        grow |= (b->host_it + nsteps > s->amp_cap) || ((b->host_it + nsteps - 1) / s->gap + 1 > s->rec_cap);
    if (grow) {
        HIPCHK(hipStreamSynchronize(b->stream));
        for (msfdm_solver *s : b->shots)
            if (int rc = grow_tables(s, b->host_it + nsteps)) return rc;
    }
    if (int rc = batch_upload(b)) return rc;
    int left = nsteps;
    while (left > 0) {
        if (b->use_graph && left >= GRAPH_STEPS && (b->host_it & 1) == 0) {
            if (!b->gexec) {
                HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                batch_enqueue(view_of(b), 0, GRAPH_STEPS);
                hipGraph_t g = nullptr;
                HIPCHK(hipStreamEndCapture(b->stream, &g));
                b->graph = g;
                HIPCHK(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
            }
            HIPCHK(hipGraphLaunch(b->gexec, b->stream));
            b->host_it += GRAPH_STEPS;
            left -= GRAPH_STEPS;
        } else {
            const int n = (b->use_graph && (b->host_it & 1)) ? 1 : std::min(left, GRAPH_STEPS);
            batch_enqueue(view_of(b), b->host_it, n);
            HIPCHK(hipGetLastError());
            b->host_it += n;
            left -= n;
        }
    }
    for (msfdm_solver *s : b->shots) {
        s->host_it = b->host_it;
        s->nrec_done = b->host_it > 0 ? (b->host_it - 1) / s->gap + 1 : 0;
    }
    return MSFDM_OK;
}

extern "C" int msfdm_batch_synchronize(msfdm_batch *b) {
    if (!b) return set_error(MSFDM_EINVAL, "null batch");
    HIPCHK(hipStreamSynchronize(b->stream));
    return MSFDM_OK;
}

extern "C" int msfdm_batch_size(const msfdm_batch *b, int *nshots) {
    if (!b || !nshots) return set_error(MSFDM_EINVAL, "null argument");
    *nshots = (int)b->shots.size();
    return MSFDM_OK;
}

extern "C" int msfdm_batch_shot(msfdm_batch *b, int shot, msfdm_solver **out) {
    if (!b || !out) return set_error(MSFDM_EINVAL, "null argument");
    if (shot < 0 || shot >= (int)b->shots.size()) return set_error(MSFDM_EINVAL, "shot out of range");
    HIPCHK(hipStreamSynchronize(b->stream));
    *out = b->shots[shot];
    return MSFDM_OK;
}

extern "C" int msfdm_batch_set_tile_rows(msfdm_batch *b, int rows) {
    if (!b) return set_error(MSFDM_EINVAL, "null batch");
    if (rows != 8 && rows != 16) return set_error(MSFDM_EINVAL, "tile rows must be 8 or 16");
    if (rows != b->tz) {
        HIPCHK(hipStreamSynchronize(b->stream));
        batch_drop_graph(b);
        b->tz = rows;
    }
    return MSFDM_OK;
}

extern "C" int msfdm_batch_get_tile_rows(const msfdm_batch *b, int *rows) {
    if (!b || !rows) return set_error(MSFDM_EINVAL, "null argument");
    *rows = b->tz;
    return MSFDM_OK;
}

extern "C" int msfdm_batch_set_occupancy(msfdm_batch *b, int policy) {
    clear_error();
    if (!b) return set_error(MSFDM_EINVAL, "null batch");
    if (policy != 0 && policy != 1) return set_error(MSFDM_EINVAL, "occupancy policy must be 0 or 1");
    if (policy == 1 && !all_vertical(b->shots))
        return set_error(MSFDM_EINVAL, "occupancy policy 1 needs vertical grid columns in every shot");
    if (policy != b->occ) {
        HIPCHK(hipStreamSynchronize(b->stream));
        batch_drop_graph(b);
        b->occ = policy;
    }
    return MSFDM_OK;
}

extern "C" int msfdm_batch_get_occupancy(const msfdm_batch *b, int *policy) {
    if (!b || !policy) return set_error(MSFDM_EINVAL, "null argument");
    *policy = b->occ;
    return MSFDM_OK;
}

extern "C" int msfdm_batch_set_graph(msfdm_batch *b, int enable) {
    if (!b) return set_error(MSFDM_EINVAL, "null batch");
    HIPCHK(hipStreamSynchronize(b->stream));
    b->use_graph = enable != 0;
    if (!b->use_graph) batch_drop_graph(b);
    return MSFDM_OK;
}

extern "C" int msfdm_batch_time_run(msfdm_batch *b, int nsteps, float *ms) {
    clear_error();
    if (!b || !ms) return set_error(MSFDM_EINVAL, "null argument");
    HIPCHK(hipSetDevice(b->device));
    if (int rc = msfdm_batch_run(b, 0)) return rc;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, b->stream));
    if (int rc = msfdm_batch_run(b, nsteps)) return rc;
    HIPCHK(hipEventRecord(e1, b->stream));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t;
    return MSFDM_OK;
}

// Tile-shape autotuner: every candidate (tile height, and for the single-shot
// solver the occupancy policy) is timed AUTOTUNE_ROUNDS times on `nsteps`
// steps, the candidates interleaved round by round (each timing after a
// graph-building warm-up from a fresh shot, so every candidate sees the same
// early, partly quiet wavefield); the fastest single timing of a candidate is
// its score, so a clock or power transient during one timing cannot decide the
// choice.  The state is reset afterwards.
constexpr int AUTOTUNE_ROUNDS = 3;
constexpr int AUTOTUNE_ROWS[2] = {8, 16};

extern "C" int msfdm_batch_autotune(msfdm_batch *b, int nsteps, int *best_rows) {
    clear_error();
    if (!b || nsteps <= 0) return set_error(MSFDM_EINVAL, "bad argument");
    const int nocc = all_vertical(b->shots) ? 2 : 1;
    float best[2][2] = {{1e30f, 1e30f}, {1e30f, 1e30f}};   // [tile height][policy]
    for (int round = 0; round < AUTOTUNE_ROUNDS; ++round)
        for (int k = 0; k < 2; ++k)
            for (int o = 0; o < nocc; ++o) {
                if (int rc = msfdm_batch_set_tile_rows(b, AUTOTUNE_ROWS[k])) return rc;
                if (int rc = msfdm_batch_set_occupancy(b, o)) return rc;
                if (int rc = msfdm_batch_reset(b)) return rc;
                if (int rc = msfdm_batch_run(b, 2 * GRAPH_STEPS)) return rc;
                float ms = 0.f;
                if (int rc = msfdm_batch_time_run(b, nsteps, &ms)) return rc;
                best[k][o] = std::min(best[k][o], ms);
            }
    int bk = 0, bo = 0;
    for (int k = 0; k < 2; ++k)
        for (int o = 0; o < nocc; ++o)
            if (best[k][o] < best[bk][bo]) { bk = k; bo = o; }
    if (int rc = msfdm_batch_set_tile_rows(b, AUTOTUNE_ROWS[bk])) return rc;
    if (int rc = msfdm_batch_set_occupancy(b, bo)) return rc;
    if (int rc = msfdm_batch_reset(b)) return rc;
    if (best_rows) *best_rows = AUTOTUNE_ROWS[bk];
    return MSFDM_OK;
}

extern "C" int msfdm_autotune(msfdm_solver *s, int nsteps, int *best_rows) {
    clear_error();
    if (!s || nsteps <= 0) return set_error(MSFDM_EINVAL, "bad argument");
    const int nocc = occupancy_variants(s) ? 2 : 1;
    float best[2][2] = {{1e30f, 1e30f}, {1e30f, 1e30f}};   // [tile height][policy]
    for (int round = 0; round < AUTOTUNE_ROUNDS; ++round)
        for (int k = 0; k < 2; ++k)
            for (int o = 0; o < nocc; ++o) {
                if (int rc = msfdm_set_tile_rows(s, AUTOTUNE_ROWS[k])) return rc;
                if (int rc = msfdm_set_occupancy(s, o)) return rc;
                if (int rc = msfdm_reset_shot(s, s->isrc)) return rc;
                if (int rc = msfdm_run(s, 2 * GRAPH_STEPS)) return rc;
                float ms = 0.f;
                if (int rc = msfdm_time_run(s, nsteps, &ms)) return rc;
                best[k][o] = std::min(best[k][o], ms);
            }
    int bk = 0, bo = 0;
    for (int k = 0; k < 2; ++k)
        for (int o = 0; o < nocc; ++o)
            if (best[k][o] < best[bk][bo]) { bk = k; bo = o; }
    if (int rc = msfdm_set_tile_rows(s, AUTOTUNE_ROWS[bk])) return rc;
    if (int rc = msfdm_set_occupancy(s, bo)) return rc;
    if (int rc = msfdm_reset_shot(s, s->isrc)) return rc;
    if (best_rows) *best_rows = AUTOTUNE_ROWS[bk];
    return MSFDM_OK;
}

// ---------------------------------------------------------------------------
// Discrete adjoint: gradient of J = sum_r <res_r, rec_r> with respect to the
// model arrays, by reverse-time stepping through the transposed scheme with
// checkpointed forward recomputation.  No reference counterpart (the
// reference has no adjoint); the forward it differentiates is msfdm_run.
// ---------------------------------------------------------------------------
namespace {

enum AdjRole { A_LAM = 0, A_LP = 1, A_ACC = 2, A_TBA = 3, A_TBB = 4 };

// Tile height of the adjoint stage kernels (4-row product tiles measured
// slower: 186 against 168 us on 32 shots of 401 x 201).
constexpr int ADJ_TZ = 8;
// Up to this many source tiles (over all shots of a launch) the adjoint
// products run as one launch (ADJ_ALL), above it as three.  Measured: one
// launch is 1.32x faster for a gradient of one 401 x 201 shot (182 tiles),
// 1.10x for 32 of them, 1.02-1.06x for 1-4 shots of 4001 x 1501 (11844 tiles
// each), 0.96x for 6 and 0.94x for 32 of those.
constexpr long long ADJ_ONE_LAUNCH_TILES = 65536;

// Floats of one shot's adjoint product buffer (fluid products, then solid).
long long adj_prod_floats(const msfdm_solver *s) {
    const long long P = s->geo.P;
    return align64((long long)D::ADJ_NQF * s->d.nz_fluid * P + (long long)D::ADJ_NQS * s->d.nz_solid * P);
}

// The adjoint source tiles of the grid (column tile, band; fluid bands first)
// in three lists, one per products kernel: the interior tiles of both media
// (D::ADJ_INTERIOR), then the other fluid tiles (D::ADJ_FLUID) and the other
// solid tiles (D::ADJ_SOLID).  After them, the list of the one-launch form
// (D::ADJ_ALL): eight equal parts, one per XCD, each holding one eighth of
// the solid boundary tiles, then of the fluid boundary tiles, then of the
// interior tiles, every eighth a contiguous run of its list (neighbouring
// tiles share halo rows in the XCD's L2); the band carries the tile's kind
// (band + kind * D::ADJ_KIND_SHIFT), entries (-1, -1) pad the parts.
int ensure_adj_tiles(msfdm_solver *s) {
    if (s->atiles) return MSFDM_OK;
    const int nx = s->d.nx, npd = s->geo.npd, nzf = s->d.nz_fluid, nzs = s->d.nz_solid;
    const int nbx = (nx - 4) / D::TX + 1;
    const int nbf = (nzf - 3 + ADJ_TZ - 1) / ADJ_TZ, nbs = (nzs - 3 + ADJ_TZ - 1) / ADJ_TZ;
    std::vector<int2> list[3];
    for (int band = 0; band < nbf + nbs; ++band)
        for (int bx = 0; bx < nbx; ++bx) {
            const bool fluid = band < nbf;
            const bool inner = fluid ? D::adj_fluid_interior(nx, npd, nzf, bx * D::TX, 3 + band * ADJ_TZ, ADJ_TZ)
                                     : D::adj_solid_interior(nx, npd, nzs, bx * D::TX, (band - nbf) * ADJ_TZ, ADJ_TZ);
            list[inner ? D::ADJ_INTERIOR : fluid ? D::ADJ_FLUID : D::ADJ_SOLID].push_back(make_int2(bx, band));
        }
    std::vector<int2> all;
    for (int k = 0; k < 3; ++k) {
        s->atiles_n[k] = (int)list[k].size();
        all.insert(all.end(), list[k].begin(), list[k].end());
    }
    std::vector<int2> part[8];
    for (int k : {D::ADJ_SOLID, D::ADJ_FLUID, D::ADJ_INTERIOR}) {
        const size_t n = list[k].size();
        for (int x = 0; x < 8; ++x)
            for (size_t i = n * x / 8; i < n * (x + 1) / 8; ++i)
                part[x].push_back(make_int2(list[k][i].x, list[k][i].y + k * D::ADJ_KIND_SHIFT));
    }
    size_t per = 0;
    for (auto &q : part) per = std::max(per, q.size());
    s->atiles_xcd = (int)per;
    for (auto &q : part) {
        q.resize(per, make_int2(-1, -1));
        all.insert(all.end(), q.begin(), q.end());
    }
    HIPCHK(hipMalloc(&s->atiles, all.size() * sizeof(int2)));
    HIPCHK(hipMemcpy(s->atiles, all.data(), all.size() * sizeof(int2), hipMemcpyHostToDevice));
    return MSFDM_OK;
}

int ensure_adjoint(msfdm_solver *s) {
    if (int rc = require_fluid_solid(s)) return rc;
    const size_t bytes = s->buf_floats * sizeof(float);
    for (auto &p : s->adj)
        if (!p) { HIPCHK(hipMalloc(&p, bytes)); HIPCHK(hipMemset(p, 0, bytes)); }
    const long long P = s->geo.P;
    if (!s->aprod) {   // cells no source covers stay zero
        const size_t n = (size_t)adj_prod_floats(s);
        HIPCHK(hipMalloc(&s->aprod, n * sizeof(float)));
        HIPCHK(hipMemset(s->aprod, 0, n * sizeof(float)));
    }
    if (int rc = ensure_adj_tiles(s)) return rc;
    if (!s->gmem) {
        const long long nF = (long long)s->d.nz_fluid * P, nS = (long long)s->d.nz_solid * P;
        const long long total = 2 * nF + 5 * nS;
        s->gfl = total;
        HIPCHK(hipMalloc(&s->gmem, 3 * total * sizeof(double)));
        HIPCHK(hipMemset(s->gmem, 0, 3 * total * sizeof(double)));
        double *p = s->gmem;
        s->grad.K = p; p += nF;
        s->grad.rho_f = p; p += nF;
        s->grad.C11 = p; p += nS;
        s->grad.C13 = p; p += nS;
        s->grad.C33 = p; p += nS;
        s->grad.C55 = p; p += nS;
        s->grad.rho_s = p;
        s->grad.gstride = 0;
    }
    return MSFDM_OK;
}

// One adjoint stage's launches: the products of the interior source tiles,
// of the other fluid and the other solid source tiles, then the gather into
// every output tile (Tbar of the state fields; the products kernels write the
// PML memory entries).
template <int DIR, int MET>
void launch_adj_met(msfdm_solver *s, hipStream_t st, const D::AdjIn &a, const D::Grad &gr, int nshot) {
    constexpr int TZ = ADJ_TZ;
    const D::Geo &g = s->geo;
    const D::Coef &c = s->coef;
    const int nzf = s->d.nz_fluid, nzs = s->d.nz_solid;
    const int nbf = (nzf - 3 + TZ - 1) / TZ;
    const int2 *tl = s->atiles;
    const int *n = s->atiles_n;
    auto grid = [&](int nt) { return dim3(8 * ((nt + 7) / 8), nshot); };
    const int nall = n[0] + n[1] + n[2];
    const bool one_launch = s->adj_launch >= 0 ? s->adj_launch == 1 : (long long)nall * nshot <= ADJ_ONE_LAUNCH_TILES;
    if (one_launch) {
        D::adj_prod_kernel<DIR, TZ, MET, D::ADJ_ALL><<<dim3(8 * s->atiles_xcd, nshot), D::NTHR, 0, st>>>(
            g, c, a, gr, tl + nall, 8 * s->atiles_xcd, nbf);
    } else {
        if (n[0])
            D::adj_prod_kernel<DIR, TZ, MET, D::ADJ_INTERIOR><<<grid(n[0]), D::NTHR, 0, st>>>(g, c, a, gr, tl, n[0], nbf);
        if (n[1])
            D::adj_prod_kernel<DIR, TZ, MET, D::ADJ_FLUID><<<grid(n[1]), D::NTHR, 0, st>>>(g, c, a, gr, tl + n[0], n[1],
                                                                                          nbf);
        if (n[2])
            D::adj_prod_kernel<DIR, TZ, MET, D::ADJ_SOLID><<<grid(n[2]), D::NTHR, 0, st>>>(g, c, a, gr, tl + n[0] + n[1],
                                                                                          n[2], nbf);
    }
    const int gbx = g.P / D::TX;            // output tiles cover every column of the pitch
    const int gbf = (nzf + TZ - 1) / TZ, gbs = (nzs + TZ - 1) / TZ;
    D::adj_gather_kernel<DIR, TZ, MET><<<grid(gbx * (gbf + gbs)), D::NTHR, 0, st>>>(g, c, a, gbx, gbf, gbs);
}
template <int DIR>
void launch_adj(msfdm_solver *s, hipStream_t st, const D::AdjIn &a, const D::Grad &gr, int nshot) {
    if (s->coef.vert) launch_adj_met<DIR, 1>(s, st, a, gr, nshot);   // vertical grid columns: xi constants
    else launch_adj_met<DIR, 0>(s, st, a, gr, nshot);
}
// total += segment partial sums (Kahan), partial sums = 0; n doubles from p.
void kahan_fold(hipStream_t st, double *part, long long n) {
    D::kahan_add_kernel<<<(int)std::min<long long>((n + 255) / 256, 8192), 256, 0, st>>>(part + n, part + 2 * n, part, n);
}

// One adjoint stage: TB = L_dir^T B_dir^T (clam LP + cprev TBp) at forward state S.
// Every entry of TB a stage can reach is written by it (the others stay zero).
// ACC (may be null): the step's accumulator, ACC += TB; the first stage of a
// step (first) sets ACC = LP + TB instead.  LP: lambda' (null: the probe
// input buffer).
template <int DIR>
int adj_stage(msfdm_solver *s, const float *S, const float *TBp, float *TB, float clam, float cprev,
              float *ACC = nullptr, const float *LP = nullptr, bool first = false) {
    D::AdjIn a;
    a.S = S; a.LP = LP ? LP : s->adj[A_LP]; a.TB = TB; a.ACC = ACC; a.prod = s->aprod;
    a.ACCin = first ? a.LP : ACC;
    a.TBp = TBp ? TBp : a.LP;   // no previous Tbar: LP with weight 0
    a.clam = clam; a.cprev = TBp ? cprev : 0.0f;
    a.bstride = 0; a.pstride = 0;
    launch_adj<DIR>(s, s->stream, a, s->grad, 1);
    HIPCHK(hipGetLastError());
    return MSFDM_OK;
}

// Reverse of one RK4 step `it` (forward stage states Un, T2, T3, T4):
// lambda <- lambda' + sum of the four stage transposes, lambda' = Filter^T lambda.
// Filter^T changes only the interface rows, so it runs in place on lambda
// (those rows staged in TA, which the first stage overwrites), and lambda'
// is lambda; the first stage writes ACC = lambda' + Tbar over every entry a
// stage writes, so ACC needs no copy of lambda' (the entries no stage writes
// are zero in every adjoint buffer).
int adj_step(msfdm_solver *s, int it, const float *Un, const float *T2, const float *T3, const float *T4) {
    float *LAM = s->adj[A_LAM], *ACC = s->adj[A_ACC], *TA = s->adj[A_TBA], *TBb = s->adj[A_TBB];
    D::iface_rows_copy<<<(s->d.nx + 255) / 256, 256, 0, s->stream>>>(s->geo, LAM, TA, 0);
    D::adj_filter<<<(s->d.nx + D::FB - 1) / D::FB, D::FB, 0, s->stream>>>(s->geo, s->coef, TA, LAM, 0);
    const float *LP = LAM;
    const float dt = s->d.dt;
    const bool odd = it & 1;                 // forward stage directions: d1 d2 d1 d2
    int rc;
    // stage 4 (d2, state T4), stage 3 (d1, T3), stage 2 (d2, T2), stage 1 (d1, U^n);
    // each adds its Tbar to ACC
    if ((rc = odd ? adj_stage<0>(s, T4, nullptr, TA, dt / 6, 0.f, ACC, LP, true)
                  : adj_stage<1>(s, T4, nullptr, TA, dt / 6, 0.f, ACC, LP, true))) return rc;
    if ((rc = odd ? adj_stage<1>(s, T3, TA, TBb, dt / 3, dt, ACC, LP) : adj_stage<0>(s, T3, TA, TBb, dt / 3, dt, ACC, LP))) return rc;
    if ((rc = odd ? adj_stage<0>(s, T2, TBb, TA, dt / 3, dt / 2, ACC, LP) : adj_stage<1>(s, T2, TBb, TA, dt / 3, dt / 2, ACC, LP))) return rc;
    // (the last stage's Tbar only goes to ACC)
    if ((rc = odd ? adj_stage<1>(s, Un, TA, nullptr, dt / 6, dt / 2, ACC, LP)
                  : adj_stage<0>(s, Un, TA, nullptr, dt / 6, dt / 2, ACC, LP))) return rc;
    std::swap(s->adj[A_LAM], s->adj[A_ACC]);
    HIPCHK(hipGetLastError());
    return MSFDM_OK;
}

int copy_grad_out(msfdm_solver *s, const double *dsrc, int rows, double *out) {
    if (!out) return MSFDM_OK;
    return host_rows(s, out, (void *)dsrc, s->geo.P * sizeof(double), s->d.nx * sizeof(double), rows, true);
}

}  // namespace

extern "C" int msfdm_adjoint_forward(msfdm_solver *s, int isrc, int nsteps, int every) {
    clear_error();
    if (!s || nsteps <= 0) return set_error(MSFDM_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(s->device));
    if (every <= 0) every = std::max(2, (int)std::ceil(std::sqrt((double)nsteps)));
    if (every & 1) ++every;                  // keep segment starts even (F/B parity of the graph)
    if (int rc = ensure_adjoint(s)) return rc;
    s->adj_nseg = 0;   // no valid forward until the checkpoints are written
    if (int rc = msfdm_reset_shot(s, isrc)) return rc;
    const size_t bytes = s->buf_floats * sizeof(float);
    // checkpoint buffers stay allocated between gradients (an FWI loop calls
    // this with the same nsteps every iteration); more are added when needed
    const int nseg = (nsteps + every - 1) / every;
    while ((int)s->ckpt.size() < nseg) {
        float *c = nullptr;
        if (hipMalloc(&c, bytes) != hipSuccess)
            return set_error(MSFDM_EDEVICE, "out of device memory for adjoint checkpoints");
        s->ckpt.push_back(c);
    }
    for (int seg = 0; seg < nseg; ++seg) {
        const int it0 = seg * every;
        HIPCHK(hipMemcpyAsync(s->ckpt[seg], s->buf[R_U], bytes, hipMemcpyDeviceToDevice, s->stream));
        if (int rc = msfdm_run(s, std::min(every, nsteps - it0))) return rc;
    }
    s->adj_nseg = nseg;
    s->adj_every = every;
    s->adj_nsteps = nsteps;
    s->adj_isrc = isrc;
    HIPCHK(hipStreamSynchronize(s->stream));
    return MSFDM_OK;
}

extern "C" int msfdm_adjoint_backward(msfdm_solver *s, const float *res_vx, const float *res_vz, const float *res_p,
                                      double *gK, double *grho_f, double *gC11, double *gC13, double *gC33,
                                      double *gC55, double *grho_s) {
    clear_error();
    if (!s || !res_vx || !res_vz || !res_p) return set_error(MSFDM_EINVAL, "null argument");
    if (s->adj_nseg == 0) return set_error(MSFDM_ESTATE, "msfdm_adjoint_forward has not been called");
    HIPCHK(hipSetDevice(s->device));
    const int nsteps = s->adj_nsteps, every = s->adj_every, gap = s->gap;
    const int nrec = (nsteps - 1) / gap + 1;
    const size_t bytes = s->buf_floats * sizeof(float);
    const size_t nx = (size_t)s->d.nx;
    // segment storage: U^n, T2, T3, T4 of each step of one segment
    if ((int)s->seg.size() < 4 * every) {
        for (int i = (int)s->seg.size(); i < 4 * every; ++i) {
            float *p = nullptr;
            if (hipMalloc(&p, bytes) != hipSuccess)
                return set_error(MSFDM_EDEVICE, "out of device memory for adjoint segment storage");
            HIPCHK(hipMemset(p, 0, bytes));
            s->seg.push_back(p);
        }
    }
    float *res = nullptr;
    HIPCHK(hipMalloc(&res, 3 * (size_t)nrec * nx * sizeof(float)));
    HIPCHK(hipMemcpy(res, res_vx, (size_t)nrec * nx * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(res + (size_t)nrec * nx, res_vz, (size_t)nrec * nx * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(res + 2 * (size_t)nrec * nx, res_p, (size_t)nrec * nx * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemsetAsync(s->adj[A_LAM], 0, bytes, s->stream));
    HIPCHK(hipMemsetAsync(s->gmem, 0, 3 * s->gfl * sizeof(double), s->stream));
    int rc = MSFDM_OK;
    for (int seg = s->adj_nseg - 1; seg >= 0 && rc == MSFDM_OK; --seg) {
        const int it0 = seg * every, it1 = std::min(nsteps, it0 + every);
        // recompute the segment, keeping every step's stage states: the
        // checkpoint goes straight into the segment's first U^n slot, and each
        // step's stage 4 writes U^{n+1} into the next slot (no copy per step)
        if (hipMemcpyAsync(s->seg[0], s->ckpt[seg], bytes, hipMemcpyDeviceToDevice, s->stream) != hipSuccess) {
            rc = set_error(MSFDM_EDEVICE, "checkpoint restore failed");
            break;
        }
        D::set_base<<<1, 64, 0, s->stream>>>(s->base, it0);
        float *keep[4] = {s->buf[R_U], s->buf[R_T2], s->buf[R_T3], s->buf[R_T4]};
        for (int it = it0; it < it1; ++it) {
            const int j = it - it0;
            s->buf[R_U] = s->seg[4 * j];
            s->buf[R_T2] = s->seg[4 * j + 1];
            s->buf[R_T3] = s->seg[4 * j + 2];
            s->buf[R_T4] = s->seg[4 * j + 3];
            s->u_out = j + 1 < every ? s->seg[4 * (j + 1)] : keep[0];
            // one complete step at offset j from the counter set to it0 above
            if (it & 1) enqueue_step_dir<1>(s, j, true, true);
            else enqueue_step_dir<0>(s, j, true, true);
        }
        s->u_out = nullptr;
        s->buf[R_U] = keep[0]; s->buf[R_T2] = keep[1]; s->buf[R_T3] = keep[2]; s->buf[R_T4] = keep[3];
        // reverse sweep over the segment
        for (int it = it1 - 1; it >= it0 && rc == MSFDM_OK; --it) {
            if (it % gap == 0)
                D::adj_inject<<<((int)nx + 255) / 256, 256, 0, s->stream>>>(s->geo, s->adj[A_LAM], res, nrec, it / gap, 0, 0);
            const int j = it - it0;
            rc = adj_step(s, it, s->seg[4 * j], s->seg[4 * j + 1], s->seg[4 * j + 2], s->seg[4 * j + 3]);
        }
        kahan_fold(s->stream, s->gmem, s->gfl);   // this segment's sums into the compensated totals
    }
    (void)hipStreamSynchronize(s->stream);
    (void)hipFree(res);
    if (rc) return rc;
    // leave the step counter where the forward run ended
    s->host_it = nsteps;
    s->nrec_done = nrec;
    D::set_base<<<1, 64, 0, s->stream>>>(s->base, nsteps);
    HIPCHK(hipStreamSynchronize(s->stream));
    const int nzf = s->d.nz_fluid, nzs = s->d.nz_solid;
    const long long T = s->gfl;   // the totals follow the partial sums
    if ((rc = copy_grad_out(s, s->grad.K + T, nzf, gK)) || (rc = copy_grad_out(s, s->grad.rho_f + T, nzf, grho_f)) ||
        (rc = copy_grad_out(s, s->grad.C11 + T, nzs, gC11)) || (rc = copy_grad_out(s, s->grad.C13 + T, nzs, gC13)) ||
        (rc = copy_grad_out(s, s->grad.C33 + T, nzs, gC33)) || (rc = copy_grad_out(s, s->grad.C55 + T, nzs, gC55)) ||
        (rc = copy_grad_out(s, s->grad.rho_s + T, nzs, grho_s)))
        return rc;
    return MSFDM_OK;
}

extern "C" int msfdm_gradient(msfdm_solver *s, int isrc, int nsteps, int every, const float *res_vx,
                              const float *res_vz, const float *res_p, double *gK, double *grho_f, double *gC11,
                              double *gC13, double *gC33, double *gC55, double *grho_s) {
    if (int rc = msfdm_adjoint_forward(s, isrc, nsteps, every)) return rc;
    return msfdm_adjoint_backward(s, res_vx, res_vz, res_p, gK, grho_f, gC11, gC13, gC33, gC55, grho_s);
}

// Adjoint probes: the input field set is kept in lambda'; op 0 applies
// L_dir^T B_dir^T (forward state = U), op 1 applies Filter^T.
extern "C" int msfdm_adjoint_set_input(msfdm_solver *s, int field, const float *in) {
    clear_error();
    if (!s || !in || field < 0 || field >= MSFDM_NFIELD) return set_error(MSFDM_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(s->device));
    if (int rc = ensure_adjoint(s)) return rc;
    return copy_field(s, s->adj[A_LP], field, nullptr, in);
}

extern "C" int msfdm_adjoint_apply(msfdm_solver *s, int op, int dir) {
    clear_error();
    if (!s || (op != 0 && op != 1) || (dir != 0 && dir != 1)) return set_error(MSFDM_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(s->device));
    if (int rc = ensure_adjoint(s)) return rc;
    float *out = s->adj[A_TBA];
    if (op == 0) {
        const int rc = dir == 0 ? adj_stage<0>(s, s->buf[R_U], nullptr, out, 1.0f, 0.0f)
                                : adj_stage<1>(s, s->buf[R_U], nullptr, out, 1.0f, 0.0f);
        if (rc) return rc;
    } else {
        HIPCHK(hipMemcpyAsync(out, s->adj[A_LP], s->buf_floats * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
        D::adj_filter<<<(s->d.nx + D::FB - 1) / D::FB, D::FB, 0, s->stream>>>(s->geo, s->coef, s->adj[A_LP], out, 0);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return MSFDM_OK;
}

extern "C" int msfdm_adjoint_get_output(msfdm_solver *s, int field, float *out) {
    clear_error();
    if (!s || !out || field < 0 || field >= MSFDM_NFIELD) return set_error(MSFDM_EINVAL, "bad argument");
    if (!s->adj[A_TBA]) return set_error(MSFDM_ESTATE, "no adjoint output yet");
    HIPCHK(hipSetDevice(s->device));
    return copy_field(s, s->adj[A_TBA], field, out, nullptr);
}

extern "C" int msfdm_update_media(msfdm_solver *s, const float *K, const float *rho_f, const float *C11,
                                  const float *C13, const float *C33, const float *C55, const float *rho_s) {
    clear_error();
    if (!s) return set_error(MSFDM_EINVAL, "null solver");
    if (int rc = require_fluid_solid(s)) return rc;
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    const int nx = s->d.nx, P = s->geo.P, nzf = s->d.nz_fluid, nzs = s->d.nz_solid;
    struct Item { const float *src; const float *dst; int rows; bool invert; } items[] = {
        {K, s->coef.Kf, nzf, false}, {rho_f, s->coef.rhof_inv, nzf, true}, {C11, s->coef.C11, nzs, false},
        {C13, s->coef.C13, nzs, false}, {C33, s->coef.C33, nzs, false}, {C55, s->coef.C55, nzs, false},
        {rho_s, s->coef.rhos_inv, nzs, true}};
    std::vector<float> h;
    for (const Item &it : items) {
        if (!it.src) continue;
        h.assign((size_t)it.rows * P, 0.0f);
        for (int r = 0; r < it.rows; ++r)
            for (int x = 0; x < nx; ++x) {
                const float v = it.src[(size_t)r * nx + x];
                if (it.invert && !(v > 0.0f)) return set_error(MSFDM_EINVAL, "densities must be positive");
                h[(size_t)r * P + x] = it.invert ? (float)(1.0 / (double)v) : v;
            }
        HIPCHK(hipMemcpy(const_cast<float *>(it.dst), h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    return MSFDM_OK;
}

// ---------------------------------------------------------------------------
// Multi-shot gradient: the shots of one model are stepped together, forward
// (checkpointed) and adjoint, one launch per RK stage for all shots of a
// wave.  Every per-shot buffer of a wave is a slab with a fixed stride per
// shot; the forward recomputation runs the batched stage kernels from per-step
// shot tables pointing into the segment storage, the adjoint kernels take the
// shot from blockIdx.y.  Each shot accumulates its own FP64 gradient (the
// single-shot arithmetic), and the per-shot gradients are added in shot order,
// so the result is deterministic.  The wave size follows from a memory budget.
// No reference counterpart.
// ---------------------------------------------------------------------------
namespace {

struct ShotsPlan {
    long long bf = 0;     // floats per state buffer
    long long gfl = 0;    // doubles per gradient set
    long long P = 0;
    long long rfl = 0;    // residual floats per shot
    long long pfl = 0;    // adjoint product floats per shot
    int every = 0, nseg = 0, nrec = 0;
    long long bufs = 0;   // state buffers per shot: U, 5 adjoint, checkpoints, 4 per segment step
    size_t per_shot = 0;  // device bytes per shot of a wave
    size_t shared = 0;    // device bytes independent of the wave size
};

ShotsPlan plan_shots(const msfdm_solver *s, int nsteps, int every) {
    ShotsPlan p;
    p.P = s->geo.P;
    p.bf = s->buf_floats;
    p.gfl = 2 * (long long)s->d.nz_fluid * p.P + 5 * (long long)s->d.nz_solid * p.P;
    p.nrec = (nsteps - 1) / s->gap + 1;
    p.rfl = 3 * (long long)p.nrec * s->d.nx;
    if (every <= 0) every = std::max(2, (int)std::lround(std::sqrt(nsteps / 4.0)));
    if (every & 1) ++every;
    p.every = every;
    p.nseg = (nsteps + every - 1) / every;
    p.bufs = 6 + p.nseg + 4LL * every;
    p.pfl = adj_prod_floats(s);
    p.per_shot = sizeof(float) * (size_t)(p.bufs * p.bf + D::KIF_ROWS * p.P + p.rfl + p.pfl) +
                 3 * sizeof(double) * (size_t)p.gfl + (size_t)(every + 1) * sizeof(D::Shot);
    p.shared = sizeof(double) * (size_t)p.gfl + ((size_t)64 << 20);
    return p;
}

enum { W_LAM = 0, W_LP = 1, W_ACC = 2, W_TA = 3, W_TB = 4 };

struct ShotsWave {
    msfdm_solver *s = nullptr;
    ShotsPlan p;
    int kc = 0;                        // shots per wave (slab capacity)
    float *fmem = nullptr;             // all float slabs
    double *gmem = nullptr;            // kc per-shot gradient sets three times (segment partial sums,
                                       // Kahan totals, compensations), then the sum over shots
    float *U = nullptr, *ckpt = nullptr, *seg = nullptr, *kif = nullptr, *res = nullptr;
    float *prod = nullptr;             // adjoint products, p.pfl floats per shot
    float *adj[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    double *grad = nullptr, *total = nullptr;
    D::Shot *tables = nullptr;         // p.every + 1 tables of kc shots (recomputation slots, forward)
    int4 *tiles = nullptr;             // tiles of kc shots (tile height s->tz)
    int nt_shot = 0, nbx = 0, nbf = 0;
    // slab idx of a region holding one state buffer per shot
    float *slab(float *base, long long idx) const { return base + idx * kc * p.bf; }
    ~ShotsWave() {
        if (s) (void)hipStreamSynchronize(s->stream);   // nothing queued may still use the slabs
        for (void *q : {(void *)fmem, (void *)gmem, (void *)tables, (void *)tiles})
            if (q) (void)hipFree(q);
    }
};

int wave_alloc(ShotsWave &w) {
    msfdm_solver *s = w.s;
    const ShotsPlan &p = w.p;
    const long long kc = w.kc;
    const long long nfl = kc * (p.bufs * p.bf + D::KIF_ROWS * p.P + p.rfl + p.pfl);
    if (hipMalloc(&w.fmem, nfl * sizeof(float)) != hipSuccess)
        return set_error(MSFDM_EDEVICE, "out of device memory for the multi-shot gradient slabs");
    if (hipMalloc(&w.gmem, (3 * kc + 1) * p.gfl * sizeof(double)) != hipSuccess)
        return set_error(MSFDM_EDEVICE, "out of device memory for the multi-shot gradients");
    HIPCHK(hipMemsetAsync(w.fmem, 0, nfl * sizeof(float), s->stream));
    HIPCHK(hipMemsetAsync(w.gmem, 0, (3 * kc + 1) * p.gfl * sizeof(double), s->stream));
    float *f = w.fmem;
    w.U = f; f += kc * p.bf;
    for (auto &a : w.adj) { a = f; f += kc * p.bf; }
    w.ckpt = f; f += (long long)p.nseg * kc * p.bf;
    w.seg = f; f += 4LL * p.every * kc * p.bf;
    w.kif = f; f += kc * D::KIF_ROWS * p.P;
    w.prod = f; f += kc * p.pfl;
    w.res = f;
    w.grad = w.gmem;
    w.total = w.gmem + 3 * kc * p.gfl;
    HIPCHK(hipMalloc(&w.tables, (size_t)(p.every + 1) * kc * sizeof(D::Shot)));
    const int tz = s->tz;
    w.nbx = (s->d.nx - 4) / D::TX + 1;
    w.nbf = (s->d.nz_fluid - 3 + tz - 1) / tz;
    const int nbs = (s->d.nz_solid - 3 + tz - 1) / tz;
    w.nt_shot = w.nbx * (w.nbf + nbs);
    std::vector<int4> t;
    for (int i = 0; i < w.kc; ++i)
        for (int band = 0; band < w.nbf + nbs; ++band)
            for (int bx = 0; bx < w.nbx; ++bx) t.push_back(make_int4(i, band, bx, w.nbf));
    HIPCHK(hipMalloc(&w.tiles, t.size() * sizeof(int4)));
    HIPCHK(hipMemcpy(w.tiles, t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice));
    return MSFDM_OK;
}

BatchView wave_view(const ShotsWave &w, int kk, int slot) {
    BatchView v{};
    v.shots = w.tables + (size_t)slot * w.kc;
    const int k = w.s->tz == 8 ? 0 : 1;
    v.tiles[k] = w.tiles;
    v.ntiles[k] = kk * w.nt_shot;
    v.tz = w.s->tz;
    v.ns = kk;
    v.maxnx = w.s->d.nx;
    v.base = w.s->base;
    v.dt = w.s->d.dt;
    v.stream = w.s->stream;
    v.vert = w.s->coef.vert != 0;   // every shot of a wave shares the solver's grid
    v.occ = w.s->occ;               // and its occupancy policy
    return v;
}

// Step `it` of the wave with the tables of `slot`; jstep: the step's offset
// from the device step counter (set once per pass, not advanced per step).
void wave_step(const ShotsWave &w, int kk, int slot, int it, int jstep) {
    const BatchView v = wave_view(w, kk, slot);
    if (it & 1) batch_step_dir<1>(v, jstep);
    else batch_step_dir<0>(v, jstep);
}

D::Grad wave_grad(const ShotsWave &w) {
    const long long nF = (long long)w.s->d.nz_fluid * w.p.P, nS = (long long)w.s->d.nz_solid * w.p.P;
    D::Grad g;
    g.K = w.grad; g.rho_f = w.grad + nF;
    g.C11 = w.grad + 2 * nF; g.C13 = g.C11 + nS; g.C33 = g.C13 + nS; g.C55 = g.C33 + nS; g.rho_s = g.C55 + nS;
    g.gstride = w.p.gfl;
    return g;
}

// One adjoint stage of the wave (adj_stage's arithmetic; LP: lambda', first:
// the step's first stage, which sets ACC = LP + Tbar).
template <int DIR>
void wave_adj_stage(ShotsWave &w, int kk, const float *S, const float *TBp, float *TB, float clam, float cprev,
                    const float *LP, bool first = false) {
    msfdm_solver *s = w.s;
    D::AdjIn a;
    a.S = S; a.LP = LP; a.TB = TB; a.ACC = w.adj[W_ACC]; a.prod = w.prod;
    a.ACCin = first ? LP : a.ACC;
    a.TBp = TBp ? TBp : a.LP;   // no previous Tbar: LP with weight 0
    a.clam = clam; a.cprev = TBp ? cprev : 0.0f;
    a.bstride = w.p.bf;
    a.pstride = w.p.pfl;
    launch_adj<DIR>(s, s->stream, a, wave_grad(w), kk);
}

// Reverse of step `it` for the kk shots of the wave (segment slot j holds the
// forward stage states of that step).
// As adj_step: Filter^T in place on lambda (interface rows staged in TA),
// lambda' = lambda, and the first stage writes ACC = lambda' + Tbar.
void wave_adj_step(ShotsWave &w, int kk, int it, int j) {
    msfdm_solver *s = w.s;
    float *LAM = w.adj[W_LAM], *TA = w.adj[W_TA], *TBb = w.adj[W_TB];
    D::iface_rows_copy<<<dim3((s->d.nx + 255) / 256, kk), 256, 0, s->stream>>>(s->geo, LAM, TA, w.p.bf);
    D::adj_filter<<<dim3((s->d.nx + D::FB - 1) / D::FB, kk), D::FB, 0, s->stream>>>(s->geo, s->coef, TA, LAM, w.p.bf);
    const float *LP = LAM;
    const float dt = s->d.dt;
    const bool odd = it & 1;
    const float *Un = w.slab(w.seg, 4LL * j), *T2 = w.slab(w.seg, 4LL * j + 1), *T3 = w.slab(w.seg, 4LL * j + 2),
                *T4 = w.slab(w.seg, 4LL * j + 3);
    // each stage adds its Tbar to ACC (wave_adj_stage)
    if (odd) wave_adj_stage<0>(w, kk, T4, nullptr, TA, dt / 6, 0.f, LP, true);
    else wave_adj_stage<1>(w, kk, T4, nullptr, TA, dt / 6, 0.f, LP, true);
    if (odd) wave_adj_stage<1>(w, kk, T3, TA, TBb, dt / 3, dt, LP); else wave_adj_stage<0>(w, kk, T3, TA, TBb, dt / 3, dt, LP);
    if (odd) wave_adj_stage<0>(w, kk, T2, TBb, TA, dt / 3, dt / 2, LP); else wave_adj_stage<1>(w, kk, T2, TBb, TA, dt / 3, dt / 2, LP);
    // (the last stage's Tbar only goes to ACC)
    if (odd) wave_adj_stage<1>(w, kk, Un, TA, nullptr, dt / 6, dt / 2, LP);
    else wave_adj_stage<0>(w, kk, Un, TA, nullptr, dt / 6, dt / 2, LP);
    std::swap(w.adj[W_LAM], w.adj[W_ACC]);
}

// One wave: shots isrc[0..kk) with residuals res_host ([kk][3][nrec][nx]).
int wave_run(ShotsWave &w, int kk, const int *isrc, int nsteps, const float *res_host) {
    msfdm_solver *s = w.s;
    const ShotsPlan &p = w.p;
    const int every = p.every, gap = s->gap;
    const size_t sbytes = (size_t)kk * p.bf * sizeof(float);
    // table j < every: step j of a segment's recomputation, U^n in slot j of the
    // segment storage and U^{n+1} written into slot j + 1 (no copy per step);
    // table `every`: the checkpointed forward, in place on U
    std::vector<D::Shot> h((size_t)(every + 1) * w.kc);
    for (int j = 0; j <= every; ++j)
        for (int i = 0; i < kk; ++i) {
            D::Shot &t = h[(size_t)j * w.kc + i];
            t.g = s->geo;
            t.c = s->coef;
            t.ctl = s->ctl;
            t.ctl.base = s->base;
            t.ctl.rec = nullptr;        // records are not kept in the gradient run
            t.ctl.rec_cap = 0;
            t.ctl.gap = 0x7fffffff;
            t.ctl.srcx = s->srcx.empty() ? -100 : s->srcx[isrc[i]];
            t.ctl.srcz = s->srcz.empty() ? -100 : s->srcz[isrc[i]];
            const long long o = (long long)i * p.bf;
            const int js = j < every ? j : 0;   // the forward's stage states go to slot 0
            t.buf[0] = (j < every ? w.slab(w.seg, 4LL * j) : w.U) + o;
            for (int r = 1; r < 4; ++r) t.buf[r] = w.slab(w.seg, 4LL * js + r) + o;
            t.buf[SHOT_UOUT] = (j + 1 < every ? w.slab(w.seg, 4LL * (j + 1)) : w.U) + o;
            t.kif = w.kif + (long long)i * D::KIF_ROWS * p.P;
        }
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(w.tables, h.data(), h.size() * sizeof(D::Shot), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(w.res, res_host, (size_t)kk * p.rfl * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemsetAsync(w.U, 0, sbytes, s->stream));
    HIPCHK(hipMemsetAsync(w.adj[W_LAM], 0, sbytes, s->stream));
    HIPCHK(hipMemsetAsync(w.kif, 0, (size_t)kk * D::KIF_ROWS * p.P * sizeof(float), s->stream));
    HIPCHK(hipMemsetAsync(w.grad, 0, 3 * (size_t)w.kc * p.gfl * sizeof(double), s->stream));
    // checkpointed forward
    D::set_base<<<1, 64, 0, s->stream>>>(s->base, 0);
    for (int it = 0; it < nsteps; ++it) {
        if (it % every == 0)
            HIPCHK(hipMemcpyAsync(w.slab(w.ckpt, it / every), w.U, sbytes, hipMemcpyDeviceToDevice, s->stream));
        wave_step(w, kk, every, it, it);
    }
    HIPCHK(hipGetLastError());
    // reverse sweep, segment by segment
    for (int sg = p.nseg - 1; sg >= 0; --sg) {
        const int it0 = sg * every, it1 = std::min(nsteps, it0 + every);
        HIPCHK(hipMemcpyAsync(w.slab(w.seg, 0), w.slab(w.ckpt, sg), sbytes, hipMemcpyDeviceToDevice, s->stream));
        D::set_base<<<1, 64, 0, s->stream>>>(s->base, it0);
        for (int it = it0; it < it1; ++it) wave_step(w, kk, it - it0, it, it - it0);
        for (int it = it1 - 1; it >= it0; --it) {
            if (it % gap == 0)
                D::adj_inject<<<dim3((s->d.nx + 255) / 256, kk), 256, 0, s->stream>>>(s->geo, w.adj[W_LAM], w.res, p.nrec,
                                                                                       it / gap, p.bf, p.rfl);
            wave_adj_step(w, kk, it, it - it0);
        }
        // every shot's segment sums into its compensated totals
        D::kahan_add_kernel<<<(int)std::min<long long>((kk * p.gfl + 255) / 256, 8192), 256, 0, s->stream>>>(
            w.grad + w.kc * p.gfl, w.grad + 2 * w.kc * p.gfl, w.grad, kk * p.gfl);
        HIPCHK(hipGetLastError());
    }
    D::grad_sum_kernel<<<(int)std::min<long long>((p.gfl + 255) / 256, 8192), 256, 0, s->stream>>>(
        w.total, w.grad + w.kc * p.gfl, p.gfl, p.gfl, kk);
    HIPCHK(hipGetLastError());
    return MSFDM_OK;
}

}  // namespace

extern "C" int msfdm_gradient_shots(msfdm_solver *s, int nshot, const int *isrc, int nsteps, int every,
                                    double mem_bytes, const float *res, double *gK, double *grho_f, double *gC11,
                                    double *gC13, double *gC33, double *gC55, double *grho_s, int *wave_shots) {
    clear_error();
    if (!s || !isrc || !res || nshot <= 0 || nsteps <= 0) return set_error(MSFDM_EINVAL, "bad argument");
    if (int rc = require_fluid_solid(s)) return rc;
    const int nsrc = s->srcx.empty() ? 1 : (int)s->srcx.size();
    for (int i = 0; i < nshot; ++i)
        if (isrc[i] < 0 || isrc[i] >= nsrc) return set_error(MSFDM_EINVAL, "shot index out of range");
    HIPCHK(hipSetDevice(s->device));
    if (int rc = grow_tables(s, nsteps)) return rc;
    if (int rc = ensure_adj_tiles(s)) return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    // the single-shot gradient's checkpoint and segment buffers (kept between
    // msfdm_gradient calls, tens of GB on a large grid) go back to the pool the
    // waves are sized from; a pending msfdm_adjoint_forward has to be repeated
    for (float *p : s->ckpt) (void)hipFree(p);
    for (float *p : s->seg) (void)hipFree(p);
    s->ckpt.clear();
    s->seg.clear();
    s->adj_nseg = 0;
    ShotsWave w;
    w.s = s;
    w.p = plan_shots(s, nsteps, every);
    size_t budget = 0;
    if (mem_bytes > 0) {
        budget = (size_t)mem_bytes;
    } else {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        budget = (size_t)(0.9 * (double)fr);
    }
    if (budget <= w.p.shared + w.p.per_shot)
        return set_error(MSFDM_EDEVICE, "device memory budget too small for one shot (" +
                                            std::to_string((w.p.shared + w.p.per_shot) >> 20) + " MiB needed)");
    w.kc = (int)std::min<size_t>((size_t)nshot, (budget - w.p.shared) / w.p.per_shot);
    if (wave_shots) *wave_shots = w.kc;
    if (int rc = wave_alloc(w)) return rc;
    for (int w0 = 0; w0 < nshot; w0 += w.kc) {
        const int kk = std::min(w.kc, nshot - w0);
        if (int rc = wave_run(w, kk, isrc + w0, nsteps, res + (size_t)w0 * w.p.rfl)) return rc;
    }
    D::set_base<<<1, 64, 0, s->stream>>>(s->base, s->host_it);   // the solver's own run is left where it was
    HIPCHK(hipStreamSynchronize(s->stream));
    const long long nF = (long long)s->d.nz_fluid * w.p.P, nS = (long long)s->d.nz_solid * w.p.P;
    const int nzf = s->d.nz_fluid, nzs = s->d.nz_solid;
    const double *t = w.total;
    int rc;
    if ((rc = copy_grad_out(s, t, nzf, gK)) || (rc = copy_grad_out(s, t + nF, nzf, grho_f)) ||
        (rc = copy_grad_out(s, t + 2 * nF, nzs, gC11)) || (rc = copy_grad_out(s, t + 2 * nF + nS, nzs, gC13)) ||
        (rc = copy_grad_out(s, t + 2 * nF + 2 * nS, nzs, gC33)) ||
        (rc = copy_grad_out(s, t + 2 * nF + 3 * nS, nzs, gC55)) ||
        (rc = copy_grad_out(s, t + 2 * nF + 4 * nS, nzs, grho_s)))
        return rc;
    return MSFDM_OK;
}

extern "C" int msfdm_gradient_shots_plan(const msfdm_solver *s, int nsteps, int every, double *bytes_per_shot,
                                         int *every_used) {
    clear_error();
    if (!s || nsteps <= 0) return set_error(MSFDM_EINVAL, "bad argument");
    if (int rc = require_fluid_solid(s)) return rc;
    const ShotsPlan p = plan_shots(s, nsteps, every);
    if (bytes_per_shot) *bytes_per_shot = (double)p.per_shot;
    if (every_used) *every_used = p.every;
    return MSFDM_OK;
}
