// Velocity moments on the grid (COLLECT_ELECTRON_MOMENTS, pic2d_ElectronMoments.f90:3-424;
// COLLECT_ION_MOMENTS, pic2d_IonMoments.f90:3-347): 13 FP32 accumulators per node with FP32
// bilinear weights.  Electrons use the half-step velocity 0.5*(v + v_pushed) obtained by
// re-running the Boris update on the current field (pic2d_ElectronMoments.f90:148-229).
#include "kernels.h"

namespace edgpu {

__global__ __launch_bounds__(kBlock) void k_moments_fp64(ParticlesF64 p, int64_t n, Geo g, const double *EX,
                                                         const double *EY, Kmat K, double EZ, int electron,
                                                         float *acc) {
    const int64_t nn = (int64_t)g.nx * g.ny;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const double x = p.x[k], y = p.y[k];
        int i = (int)x, j = (int)y;
        if (x == g.xmax) i = g.gmi - 1;
        if (y == g.ymax) j = g.gmj - 1;
        if (i < 0 || i > g.gmi - 1 || j < 0 || j > g.gmj - 1) continue;
        const int64_t q0 = (int64_t)i + (int64_t)j * g.nx;
        const float ax_ip1 = (float)(x - (double)i), ax_i = 1.0f - ax_ip1;
        const float ay_jp1 = (float)(y - (double)j), ay_j = 1.0f - ay_jp1;
        float w[4];
        w[0] = ax_i * ay_j;
        w[1] = ax_ip1 * ay_j;
        w[2] = ax_i * ay_jp1;
        w[3] = 1.0f - w[0] - w[1] - w[2];
        double vx = p.vx[k], vy = p.vy[k], vz = p.vz[k];
        if (electron) {
            const double dax_ip1 = x - (double)i, dax_i = 1.0 - dax_ip1;
            const double day_jp1 = y - (double)j, day_j = 1.0 - day_jp1;
            const double E_X = EX[q0] * dax_i * day_j + EX[q0 + 1] * dax_ip1 * day_j + EX[q0 + g.nx] * dax_i * day_jp1 +
                               EX[q0 + g.nx + 1] * dax_ip1 * day_jp1;
            const double E_Y = EY[q0] * dax_i * day_j + EY[q0 + 1] * dax_ip1 * day_j + EY[q0 + g.nx] * dax_i * day_jp1 +
                               EY[q0 + g.nx + 1] * dax_ip1 * day_jp1;
            const double vxm = vx - 0.5 * E_X, vym = vy - 0.5 * E_Y, vzm = vz - 0.5 * EZ;
            const double ux = K.k[0] * vxm + K.k[1] * vym + K.k[2] * vzm - 0.5 * E_X;
            const double uy = K.k[3] * vxm + K.k[4] * vym + K.k[5] * vzm - 0.5 * E_Y;
            const double uz = K.k[6] * vxm + K.k[7] * vym + K.k[8] * vzm - 0.5 * EZ;
            vx = 0.5 * (vx + ux);
            vy = 0.5 * (vy + uy);
            vz = 0.5 * (vz + uz);
        }
        const float rvx = (float)vx, rvy = (float)vy, rvz = (float)vz;
        const float v2 = rvx * rvx + rvy * rvy + rvz * rvz;
        const float m[13] = {1.0f,      rvx,       rvy,       rvz,      rvx * rvx, rvy * rvy, rvz * rvz,
                             rvx * rvy, rvx * rvz, rvy * rvz, v2 * rvx, v2 * rvy,  v2 * rvz};
        const int64_t q[4] = {q0, q0 + 1, q0 + g.nx, q0 + g.nx + 1};
#pragma unroll
        for (int a = 0; a < 13; ++a)
#pragma unroll
            for (int c = 0; c < 4; ++c) atomicAdd(&acc[a * nn + q[c]], a == 0 ? w[c] : m[a] * w[c]);
    }
}

void launch_moments_fp64(const ParticlesF64 &p, int64_t n, const Geo &g, const double *EX, const double *EY,
                         const double K[9], double EZ, bool electron, float *acc, hipStream_t s) {
    if (n <= 0) return;
    Kmat km;
    for (int q = 0; q < 9; ++q) km.k[q] = K[q];
    int64_t b = (n + kBlock - 1) / kBlock;
    k_moments_fp64<<<(int)(b > 65536 ? 65536 : b), kBlock, 0, s>>>(p, n, g, EX, EY, km, EZ, electron ? 1 : 0, acc);
}

}  // namespace edgpu
