// Preconditioned conjugate gradients for the potential on geometries with inner objects
// (the solve SOLVE_POTENTIAL_WITH_PETSC hands to KSPSolve, pic2d_PETSc_Solver.F90:31-40).
//
// Operator: 5-point Laplacian with eps=1 (pic2d_PETSc_Solver.F90:140-330), Dirichlet rows for
// wall nodes and nodes of metal inner boxes.  The Dirichlet values are moved to the
// right-hand side, which leaves -A restricted to the free nodes symmetric positive
// definite.  Vectors are full grids (x fastest); only free nodes are updated.  Scalars of
// iteration k live in red[4k..4k+3] on the device so no host round trip is needed per
// iteration; the host reads the residual every few iterations.
#include "kernels.h"

namespace edgpu {

// mask: 1 free node, 0 Dirichlet or periodic image (columns gmi-1, gmi when x-periodic)
__global__ void k_krylov_mask(uint8_t *mask, Geo g) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)g.nx * g.ny) return;
    const int i = (int)(q % g.nx), j = (int)(q / g.nx);
    uint8_t f = 1;
    if (j == 0 || j == g.gmj) f = 0;
    if (!g.periodic_x && (i == 0 || i == g.gmi)) f = 0;
    if (g.periodic_x && i >= g.gmi - 1) f = 0;
    for (int m = 0; m < g.n_inner; ++m)
        if (i >= g.box[m][0] && i <= g.box[m][1] && j >= g.box[m][2] && j <= g.box[m][3]) f = 0;
    mask[q] = f;
}

void launch_krylov_setup_mask(uint8_t *mask, const Geo &g, hipStream_t s) {
    int64_t n = (int64_t)g.nx * g.ny;
    k_krylov_mask<<<(int)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(mask, g);
}

// b on free nodes = -(rho_i - rho_e)/N_ppc (pic2d_ElectricFieldCalc_PETSc.F90:49,116);
// x on Dirichlet nodes = their potential
__global__ void k_krylov_rhs(const double *rho_e, const double *rho_i, double factor_rho, const double *dval,
                             const uint8_t *mask, double *b, double *x, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    if (mask[q]) {
        b[q] = factor_rho * (rho_i[q] - rho_e[q]);
    } else {
        b[q] = 0.0;
        if (dval[q] == dval[q]) x[q] = dval[q];
    }
}

void launch_krylov_rhs(const double *rho_e, const double *rho_i, double factor_rho, const double *dirichlet_val,
                       const uint8_t *mask, double *b, double *x, const Geo &g, hipStream_t s) {
    int64_t n = (int64_t)g.nx * g.ny;
    k_krylov_rhs<<<(int)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(rho_e, rho_i, factor_rho, dirichlet_val, mask, b,
                                                                     x, n);
}

__device__ inline double stencil(const double *v, int64_t q, int i, const Geo &g) {
    int64_t ql = q - 1, qr = q + 1;
    if (g.periodic_x) {
        if (i == 0) ql = q + (g.gmi - 2);
        if (i == g.gmi - 2) qr = q - (g.gmi - 2);
    }
    return v[ql] + v[qr] + v[q - g.nx] + v[q + g.nx] - 4.0 * v[q];
}

__device__ inline void block_sum_atomic(double v, double *dst) {
    __shared__ double sh[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
        unsafeAtomicAdd(dst, s);
    }
}

// r = A x - b on free nodes (residual of -A x = -b), z = r/4, red[0] += r.z, red[3] += r.r
__global__ __launch_bounds__(kBlock) void k_pcg_init(const double *x, const double *b, const uint8_t *mask, double *r,
                                                     double *z, double *p, double *red, Geo g) {
    const int64_t n = (int64_t)g.nx * g.ny;
    double rz = 0.0, rr = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        double rv = 0.0, zv = 0.0;
        if (mask[q]) {
            const int i = (int)(q % g.nx);
            rv = stencil(x, q, i, g) - b[q];
            zv = 0.25 * rv;
        }
        r[q] = rv;
        z[q] = zv;
        p[q] = zv;
        rz += rv * zv;
        rr += rv * rv;
    }
    block_sum_atomic(rz, &red[0]);
    block_sum_atomic(rr, &red[3]);
}

// q = -A p on free nodes, red[k*4+1] += p.q
__global__ __launch_bounds__(kBlock) void k_pcg_matvec(const double *p, double *qv, const uint8_t *mask, double *red,
                                                       Geo g) {
    const int64_t n = (int64_t)g.nx * g.ny;
    double pq = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        if (mask[q]) {
            const int i = (int)(q % g.nx);
            v = -stencil(p, q, i, g);
            pq += p[q] * v;
        }
        qv[q] = v;
    }
    block_sum_atomic(pq, &red[1]);
}

// alpha = rz/pq; x += alpha p; r -= alpha q; z = r/4; next rz, rr
__global__ __launch_bounds__(kBlock) void k_pcg_update(double *x, double *r, double *z, const double *p,
                                                       const double *qv, const uint8_t *mask, const double *red_k,
                                                       double *red_next, Geo g) {
    const int64_t n = (int64_t)g.nx * g.ny;
    const double alpha = red_k[0] / red_k[1];
    double rz = 0.0, rr = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        if (!mask[q]) continue;
        x[q] += alpha * p[q];
        const double rv = r[q] - alpha * qv[q];
        const double zv = 0.25 * rv;
        r[q] = rv;
        z[q] = zv;
        rz += rv * zv;
        rr += rv * rv;
    }
    block_sum_atomic(rz, &red_next[0]);
    block_sum_atomic(rr, &red_next[3]);
}

// p = z + beta p with beta = rz_next / rz
__global__ __launch_bounds__(kBlock) void k_pcg_direction(double *p, const double *z, const uint8_t *mask,
                                                          const double *red_k, const double *red_next, int64_t n) {
    const double beta = red_next[0] / red_k[0];
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        if (mask[q]) p[q] = z[q] + beta * p[q];
}

// copy the sources of the periodic images into the image columns
__global__ void k_sync_images(double *x, Geo g) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= g.ny) return;
    x[(int64_t)j * g.nx + g.gmi - 1] = x[(int64_t)j * g.nx + 0];
    x[(int64_t)j * g.nx + g.gmi] = x[(int64_t)j * g.nx + 1];
}

__global__ void k_zero_free(double *v, const uint8_t *mask, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n && mask[q]) v[q] = 0.0;
}

// Returns the iteration count (>=0) on convergence, -1 when max_iter is reached.
int krylov_pcg(KrylovWork &w, double *x, const Geo &g, double rtol, int max_iter, double *relres, hipStream_t s) {
    const int64_t n = (int64_t)g.nx * g.ny;
    int grid = (int)((n + kBlock - 1) / kBlock);
    if (grid > 4096) grid = 4096;
    const int eg = (int)((n + kBlock - 1) / kBlock);
    if (max_iter > w.nred - 3) max_iter = w.nred - 3;
    double *red = w.red;
    double *red_b = w.red + 4 * (size_t)(w.nred - 1);
    hipMemsetAsync(w.red, 0, sizeof(double) * 4 * (size_t)w.nred, s);
    // ||b'||: residual of the vector that holds only the Dirichlet values
    hipMemcpyAsync(w.z, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s);
    k_zero_free<<<eg, kBlock, 0, s>>>(w.z, w.mask, n);
    k_pcg_init<<<grid, kBlock, 0, s>>>(w.z, w.b, w.mask, w.q, w.r, w.p, red_b, g);
    k_pcg_init<<<grid, kBlock, 0, s>>>(x, w.b, w.mask, w.r, w.z, w.p, red, g);
    double h[4], hb[4];
    hipMemcpyAsync(hb, red_b, sizeof(double) * 4, hipMemcpyDeviceToHost, s);
    hipMemcpyAsync(h, red, sizeof(double) * 4, hipMemcpyDeviceToHost, s);
    hipStreamSynchronize(s);
    double bnorm = sqrt(hb[3]);
    if (!(bnorm > 0.0)) bnorm = 1.0;
    double res = sqrt(h[3]) / bnorm;
    int it = 0;
    while (res > rtol && it < max_iter) {
        double *rk = red + 4 * (size_t)it, *rn = red + 4 * (size_t)(it + 1);
        k_pcg_matvec<<<grid, kBlock, 0, s>>>(w.p, w.q, w.mask, rk, g);
        k_pcg_update<<<grid, kBlock, 0, s>>>(x, w.r, w.z, w.p, w.q, w.mask, rk, rn, g);
        k_pcg_direction<<<grid, kBlock, 0, s>>>(w.p, w.z, w.mask, rk, rn, n);
        ++it;
        if ((it & 7) == 0 || it == max_iter) {
            hipMemcpyAsync(h, rn, sizeof(double) * 4, hipMemcpyDeviceToHost, s);
            hipStreamSynchronize(s);
            res = sqrt(h[3]) / bnorm;
        }
    }
    // the last checked residual may lag by up to 7 iterations: report the exact one
    hipMemcpyAsync(h, red + 4 * (size_t)it, sizeof(double) * 4, hipMemcpyDeviceToHost, s);
    if (g.periodic_x) k_sync_images<<<(g.ny + kBlock - 1) / kBlock, kBlock, 0, s>>>(x, g);
    hipStreamSynchronize(s);
    res = sqrt(h[3]) / bnorm;
    if (relres) *relres = res;
    return res <= rtol ? it : -1;
}

}  // namespace edgpu
