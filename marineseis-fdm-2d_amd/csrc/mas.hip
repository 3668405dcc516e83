// Phase-shift dispersion image of a receiver gather on the GPU, in FP64: the
// consumer-side analysis of the seafloor records (tools/MAS.py:50-98, also
// tools/test.py:50-99).  For a time-major gather g[nt][ntr]:
//   S(f, x) = sum_t g[t][x] exp(-2 pi i f t / nt)          (numpy.fft.rfft over t)
//   P(f, x) = conj(S) / |S|   (0 where |S| = 0)
//   A(f, c) = |sum_x P(f, x) exp(-i w_f x / c)| / ntr,      w_f = 2 pi f fs / nt
// Two kernels: the DFT from a twiddle table indexed by (f t) mod nt (one thread
// per trace, one block row per frequency), and the stack (one thread per test
// velocity, the frequency's phase row read by every thread of the block).
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "model.hpp"

// Scratch of msfdm_dispersion_image_noalloc: the twiddle table (nt complex
// doubles, padded to 256 bytes) followed by the phase spectrum (nf x ntr complex
// doubles).  Each GPU the library runs on holds one copy until the process
// ends.  In host code `scratch` names a host-side stand-in: its device address
// on each GPU comes from hipGetSymbolAddress and is kept in scratch_dev.
static __device__ __align__(256) unsigned char scratch[MSFDM_DISP_SCRATCH_BYTES];

namespace {

using msfdm::set_error;
constexpr size_t DISP_SCRATCH_BYTES = MSFDM_DISP_SCRATCH_BYTES;
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
unsigned char *scratch_dev[MAX_DEVICES] = {};

// Device address of `scratch` on the current device (looked up once per device,
// with that device current).
int scratch_of_current_device(unsigned char **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(MSFDM_EDEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES) return set_error(MSFDM_EDEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(MSFDM_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        scratch_dev[dev] = static_cast<unsigned char *>(p);
    }
    *out = scratch_dev[dev];
    return MSFDM_OK;
}

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

__global__ void twiddle_kernel(double2 *w, int n) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    double s, c;
    sincospi(2.0 * (double)m / (double)n, &s, &c);
    w[m] = make_double2(c, -s);   // exp(-2 pi i m / n)
}

__global__ void __launch_bounds__(256) phase_spectrum_kernel(const float *__restrict__ g, int nt, int ntr, int nf,
                                                             const double2 *__restrict__ w, double2 *__restrict__ P) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (x >= ntr || f >= nf) return;
    double re = 0.0, im = 0.0;
    int idx = 0;                       // (f * t) mod nt, f < nt
    for (int t = 0; t < nt; ++t) {
        const double v = g[(long long)t * ntr + x];
        const double2 tw = w[idx];
        re = fma(v, tw.x, re);
        im = fma(v, tw.y, im);
        idx += f;
        if (idx >= nt) idx -= nt;
    }
    const double mag = hypot(re, im);
    P[(long long)f * ntr + x] = mag > 0.0 ? make_double2(re / mag, -im / mag) : make_double2(0.0, 0.0);
}

__global__ void __launch_bounds__(256) stack_kernel(const double2 *__restrict__ P, const double *__restrict__ xo,
                                                    const double *__restrict__ vel, int nf, int ntr, int nc,
                                                    double fstep, double *__restrict__ amp) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (c >= nc || f >= nf) return;
    const double omega = 2.0 * M_PI * ((double)f * fstep);   // numpy: 2 pi * rfftfreq
    const double v = vel[c];
    const double2 *row = P + (long long)f * ntr;
    double re = 0.0, im = 0.0;
    for (int x = 0; x < ntr; ++x) {
        double s, co;
        sincos(-(omega * xo[x]) / v, &s, &co);
        const double2 p = row[x];
        re += p.x * co - p.y * s;
        im += p.x * s + p.y * co;
    }
    amp[(long long)f * nc + c] = hypot(re, im) / (double)ntr;
}

}  // namespace

using namespace msfdm;

extern "C" int msfdm_dispersion_image(const float *gather, int nt, int ntr, const double *offsets, double fs,
                                      const double *velocities, int nc, int device, double *freqs, double *amp) {
    clear_error();
    if (!gather || !offsets || !velocities || !amp) return set_error(MSFDM_EINVAL, "null argument");
    if (nt < 2 || ntr < 1 || nc < 1) return set_error(MSFDM_EINVAL, "need nt >= 2, ntr >= 1 and nc >= 1");
    if (!(fs > 0.0) || !std::isfinite(fs)) return set_error(MSFDM_EINVAL, "fs must be a positive finite value");
    for (int i = 0; i < ntr; ++i)
        if (!std::isfinite(offsets[i])) return set_error(MSFDM_EINVAL, "all offsets must be finite");
    for (int i = 0; i < nc; ++i)
        if (!(velocities[i] > 0.0) || !std::isfinite(velocities[i]))
            return set_error(MSFDM_EINVAL, "test velocities must be positive and finite");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(MSFDM_EDEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return set_error(MSFDM_EINVAL, "device index out of range");
    const int nf = nt / 2 + 1;
    const double fstep = 1.0 / ((double)nt * (1.0 / fs));   // numpy.fft.rfftfreq spacing
    float *dg = nullptr;
    double2 *dw = nullptr, *dP = nullptr;
    double *dx = nullptr, *dv = nullptr, *da = nullptr;
    hipError_t e = hipSetDevice(device);
    auto fail = [&](hipError_t err, const char *what) {
        (void)hipFree(dg); (void)hipFree(dw); (void)hipFree(dP);
        (void)hipFree(dx); (void)hipFree(dv); (void)hipFree(da);
        return set_error(MSFDM_EDEVICE, std::string(what) + ": " + hipGetErrorString(err));
    };
    if (e != hipSuccess) return fail(e, "hipSetDevice");
    const size_t ng = (size_t)nt * ntr, np = (size_t)nf * ntr, na = (size_t)nf * nc;
    if ((e = hipMalloc(&dg, ng * sizeof(float))) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMalloc(&dw, (size_t)nt * sizeof(double2))) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMalloc(&dP, np * sizeof(double2))) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMalloc(&dx, (size_t)ntr * sizeof(double))) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMalloc(&dv, (size_t)nc * sizeof(double))) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMalloc(&da, na * sizeof(double))) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMemcpy(dg, gather, ng * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "hipMemcpy");
    if ((e = hipMemcpy(dx, offsets, (size_t)ntr * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "hipMemcpy");
    if ((e = hipMemcpy(dv, velocities, (size_t)nc * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "hipMemcpy");
    twiddle_kernel<<<(nt + 255) / 256, 256>>>(dw, nt);
    phase_spectrum_kernel<<<dim3((ntr + 255) / 256, nf), 256>>>(dg, nt, ntr, nf, dw, dP);
    stack_kernel<<<dim3((nc + 255) / 256, nf), 256>>>(dP, dx, dv, nf, ntr, nc, fstep, da);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e, "kernel launch");
    if ((e = hipMemcpy(amp, da, na * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess) return fail(e, "hipMemcpy");
    if (freqs)
        for (int f = 0; f < nf; ++f) freqs[f] = (double)f * fstep;
    (void)hipFree(dg); (void)hipFree(dw); (void)hipFree(dP);
    (void)hipFree(dx); (void)hipFree(dv); (void)hipFree(da);
    return MSFDM_OK;
}

// The same three kernels over caller-owned device buffers, with the twiddle
// table and the phase spectrum in the library's static scratch: no device
// memory is allocated or freed.  Every intermediate cell is written before it is
// read, so nothing in the scratch needs clearing between calls.
extern "C" int msfdm_dispersion_image_noalloc(const float *d_gather, int nt, int ntr, const double *d_offsets,
                                              double fs, const double *d_velocities, int nc, double *d_amp,
                                              void *stream) {
    clear_error();
    if (!d_gather || !d_offsets || !d_velocities || !d_amp) return set_error(MSFDM_EINVAL, "null argument");
    if (nt < 2 || ntr < 1 || nc < 1) return set_error(MSFDM_EINVAL, "need nt >= 2, ntr >= 1 and nc >= 1");
    if (!(fs > 0.0) || !std::isfinite(fs)) return set_error(MSFDM_EINVAL, "fs must be a positive finite value");
    const int nf = nt / 2 + 1;
    const size_t tw_bytes = align256((size_t)nt * sizeof(double2));
    const size_t need = tw_bytes + (size_t)nf * (size_t)ntr * sizeof(double2);
    if (need > DISP_SCRATCH_BYTES)
        return set_error(MSFDM_EINVAL, "gather too large for the allocation-free scratch: " + std::to_string(need) +
                                           " bytes needed, " + std::to_string(DISP_SCRATCH_BYTES) + " available");
    unsigned char *base = nullptr;
    if (int rc = scratch_of_current_device(&base)) return rc;
    double2 *dw = reinterpret_cast<double2 *>(base);
    double2 *dP = reinterpret_cast<double2 *>(base + tw_bytes);
    const double fstep = 1.0 / ((double)nt * (1.0 / fs));   // numpy.fft.rfftfreq spacing
    hipStream_t st = static_cast<hipStream_t>(stream);
    twiddle_kernel<<<(nt + 255) / 256, 256, 0, st>>>(dw, nt);
    phase_spectrum_kernel<<<dim3((ntr + 255) / 256, nf), 256, 0, st>>>(d_gather, nt, ntr, nf, dw, dP);
    stack_kernel<<<dim3((nc + 255) / 256, nf), 256, 0, st>>>(dP, d_offsets, d_velocities, nf, ntr, nc, fstep, d_amp);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(MSFDM_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    return MSFDM_OK;
}
