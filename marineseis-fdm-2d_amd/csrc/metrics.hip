// FP64 curvilinear metric terms of a body-fitted grid, on the GPU.
//
// Follows the reference's model generators (tools/run_models/model_ortho_classic_scholte.py:188-195,
// tools/run_models/model_Zhao_scholte.py:203-210): grid-coordinate derivatives in unit
// computational spacing with numpy.gradient's edge_order=2 stencil
//   interior        f'(i) = (f[i+1] - f[i-1]) / 2
//   first point     f'(0) = -1.5 f[0] + 2 f[1] - 0.5 f[2]
//   last point      f'(n-1) = 0.5 f[n-3] - 2 f[n-2] + 1.5 f[n-1]
// then the inverse metric of the map (xi, zeta) -> (x, z):
//   J = x_xi z_ze - x_ze z_xi,  xi_x = z_ze / J,  xi_z = -x_ze / J,
//   ze_x = -z_xi / J,           ze_z = x_xi / J.
// For vertical grid columns (x = h xi) this is the reference's J = h z_ze,
// xi_x = 1/h, xi_z = 0, ze_x = -z_xi / J, ze_z = h / J.  Every operation is
// a separately rounded FP64 multiply / add / divide in numpy's evaluation
// order, so the results are those of the same formulas evaluated in numpy
// float64.  Contraction into FMAs is switched off for this file (pragma and
// -ffp-contract=off): a fused multiply-add rounds once where numpy rounds twice,
// which moves the cancelling one-sided end stencil by up to an ulp of z.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "model.hpp"
#include "msfdm.h"

namespace {

__device__ __forceinline__ double grad_at(const double *__restrict__ f, long long stride, int i, int n) {
    // numpy.gradient(f, edge_order=2) along one axis with unit spacing, n >= 3
#pragma clang fp contract(off)
    if (i == 0) {
        const double a = -1.5 * f[0], b = 2.0 * f[stride], c = -0.5 * f[2 * stride];
        return (a + b) + c;
    }
    if (i == n - 1) {
        const double a = 0.5 * f[-2 * stride], b = -2.0 * f[-stride], c = 1.5 * f[0];
        return (a + b) + c;
    }
    return (f[stride] - f[-stride]) / 2.0;
}

__global__ void __launch_bounds__(256)
metric_kernel(const double *__restrict__ X, const double *__restrict__ Z, int nz, int nx, double *__restrict__ J,
              double *__restrict__ xix, double *__restrict__ xiz, double *__restrict__ zex, double *__restrict__ zez) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= nx || k >= nz) return;
    const long long q = (long long)k * nx + i;
    const double x_xi = grad_at(X + q, 1, i, nx), z_xi = grad_at(Z + q, 1, i, nx);
    const double x_ze = grad_at(X + q, nx, k, nz), z_ze = grad_at(Z + q, nx, k, nz);
    const double p1 = x_xi * z_ze, p2 = x_ze * z_xi;
    const double j = p1 - p2;
    J[q] = j;
    xix[q] = z_ze / j;
    xiz[q] = (0.0 - x_ze) / j;
    zex[q] = (0.0 - z_xi) / j;
    zez[q] = x_xi / j;
}

#define MCHK(expr)                                                                                      \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) { rc = set_error(MSFDM_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); goto done; } \
    } while (0)

}  // namespace

using msfdm::clear_error;
using msfdm::set_error;

extern "C" int msfdm_grid_metrics(const double *x, const double *z, int nz, int nx, int device, double *J,
                                  double *xi_x, double *xi_z, double *ze_x, double *ze_z) {
    clear_error();
    if (!x || !z || !J || !xi_x || !xi_z || !ze_x || !ze_z) return set_error(MSFDM_EINVAL, "null argument");
    if (nz < 3 || nx < 3) return set_error(MSFDM_EINVAL, "the grid needs at least 3 x 3 points");
    if ((long long)nz * nx > (1LL << 31) || nz > 65535) return set_error(MSFDM_EINVAL, "grid too large");
    int rc = MSFDM_OK;
    const size_t n = (size_t)nz * nx, bytes = n * sizeof(double);
    double *d = nullptr;
    double *out[5] = {J, xi_x, xi_z, ze_x, ze_z};
    int ndev = 0;
    MCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return set_error(MSFDM_EINVAL, "no such device");
    MCHK(hipSetDevice(device));
    MCHK(hipMalloc(&d, 7 * bytes));
    MCHK(hipMemcpy(d, x, bytes, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(d + n, z, bytes, hipMemcpyHostToDevice));
    metric_kernel<<<dim3((nx + 255) / 256, nz), 256>>>(d, d + n, nz, nx, d + 2 * n, d + 3 * n, d + 4 * n, d + 5 * n,
                                                       d + 6 * n);
    MCHK(hipGetLastError());
    for (int a = 0; a < 5; ++a) MCHK(hipMemcpy(out[a], d + (2 + a) * n, bytes, hipMemcpyDeviceToHost));
done:
    if (d) (void)hipFree(d);
    return rc;
}
