// The FFT_X Poisson solve for callers that cannot allocate device memory (code under HIP
// graph capture, tight loops, harnesses that forbid hipMalloc): the caller brings the charge
// densities and the potential as device arrays, and every work array (twiddles, Thomas
// coefficients, wall rows, the half spectrum) lives in one static device buffer of a fixed
// size.  The same kernels as edgpu_solve_poisson_fftx, so the results are the same bits.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>

#include "kernels.h"

namespace edgpu {

namespace {

// largest grid: N = 2048 periodic cells, gmj = 1024 rows between the walls
constexpr int kNaMaxN = 2048;
constexpr int kNaMaxGmj = 1024;
constexpr size_t kNaModes = kNaMaxN / 2 + 1;
constexpr size_t kNaRows = kNaMaxGmj + 1;
constexpr size_t kNaSpecBytes = sizeof(double2) * kNaModes * kNaRows;  // half spectrum
constexpr size_t kNaAeqBytes = sizeof(double) * kNaModes * kNaRows;    // Thomas coefficients
constexpr size_t kNaTwBytes = sizeof(double2) * (kNaMaxN / 2);          // twiddles
constexpr size_t kNaWallBytes = sizeof(double) * 2 * (kNaMaxN + 2);     // wall rows below, above
constexpr size_t kNaBytes = kNaSpecBytes + kNaAeqBytes + kNaTwBytes + kNaWallBytes;

// Every call on a device shares this buffer; a call leaves its contents behind, and every
// array in it is rewritten before it is read, so nothing has to start at zero.
static __device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kNaMaxDevices = 64;
void *g_scratch[kNaMaxDevices] = {};  // device address of scratch per device ordinal

// scratch's device address on the current device, looked up once per device
hipError_t scratch_ptr(void **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kNaMaxDevices) return hipErrorInvalidDevice;
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

__global__ void k_fill_walls(double *below, double *above, int nx, double phi_below, double phi_above) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx) return;
    below[i] = phi_below;
    above[i] = phi_above;
}

}  // namespace

}  // namespace edgpu

using namespace edgpu;

extern "C" int edgpu_noalloc_scratch_bytes(int64_t *bytes) {
    if (bytes) *bytes = (int64_t)kNaBytes;
    return EDGPU_OK;
}

extern "C" int edgpu_solve_poisson_fftx_noalloc(int32_t N, int32_t gmj, int32_t N_of_particles_cell,
                                                double phi_below, double phi_above, const double *rho_e,
                                                const double *rho_i, double *phi, void *stream) {
    if (N < 4 || (N & (N - 1)) != 0 || N > kNaMaxN || gmj < 2 || gmj > kNaMaxGmj)
        return edgpu_set_error(EDGPU_ERR_ARG, "edgpu_solve_poisson_fftx_noalloc: needs N a power of two in [4, 2048] "
                                              "and 2 <= gmj <= 1024 (the size of its static work buffer)");
    if (N_of_particles_cell < 1 || !rho_e || !rho_i || !phi)
        return edgpu_set_error(EDGPU_ERR_ARG, "edgpu_solve_poisson_fftx_noalloc: null grid or N_of_particles_cell < 1");
    void *base = nullptr;
    hipError_t e = scratch_ptr(&base);
    if (e != hipSuccess) {
        char buf[160];
        std::snprintf(buf, sizeof(buf), "edgpu_solve_poisson_fftx_noalloc: hipGetSymbolAddress: %s",
                      hipGetErrorString(e));
        return edgpu_set_error(EDGPU_ERR_HIP, buf);
    }
    unsigned char *b = (unsigned char *)base;
    double2 *spec = (double2 *)b;
    double *a_eq = (double *)(b + kNaSpecBytes);
    double2 *tw = (double2 *)(b + kNaSpecBytes + kNaAeqBytes);
    double *wall_below = (double *)(b + kNaSpecBytes + kNaAeqBytes + kNaTwBytes);
    double *wall_above = wall_below + (kNaMaxN + 2);
    hipStream_t s = (hipStream_t)stream;
    Geo g{};
    g.gmi = N + 1;
    g.gmj = gmj;
    g.nx = N + 2;
    g.ny = gmj + 1;
    g.periodic_x = 1;
    k_fill_walls<<<(g.nx + kBlock - 1) / kBlock, kBlock, 0, s>>>(wall_below, wall_above, g.nx, phi_below, phi_above);
    launch_fftx_tables(tw, a_eq, N, gmj, s);
    launch_poisson_forward(rho_e, rho_i, 1.0 / (double)N_of_particles_cell, wall_below, wall_above, tw, spec, g, N, s);
    launch_poisson_thomas(spec, a_eq, N, gmj, s);
    launch_poisson_inverse(spec, tw, phi, g, N, s);
    e = hipGetLastError();
    if (e != hipSuccess) {
        char buf[160];
        std::snprintf(buf, sizeof(buf), "edgpu_solve_poisson_fftx_noalloc: launch: %s", hipGetErrorString(e));
        return edgpu_set_error(EDGPU_ERR_HIP, buf);
    }
    return EDGPU_OK;
}
