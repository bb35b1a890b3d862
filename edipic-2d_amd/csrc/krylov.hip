// Preconditioned conjugate gradients for the potential on geometries with inner objects
// (the solve SOLVE_POTENTIAL_WITH_PETSC hands to KSPSolve, pic2d_PETSc_Solver.F90:31-40).
//
// Operator: 5-point Laplacian with eps=1 (pic2d_PETSc_Solver.F90:140-330), Dirichlet rows for
// wall nodes and nodes of metal inner boxes.  The Dirichlet values are moved to the
// right-hand side, which leaves -A restricted to the free nodes symmetric positive
// definite.  Vectors are full grids (x fastest); only free nodes are updated.  Scalars of
// iteration k live in red[4k..4k+3] on the device so no host round trip is needed per
// iteration; the host reads the residual every few iterations.
#include <algorithm>

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

// Reductions are two-stage and deterministic: each workgroup stores its partial sums,
// part[k * gridDim.x + blockIdx.x] for quantity k, and k_finish adds them in a fixed order.
// (Same-address FP64 atomics from thousands of workgroups serialise at the memory side and
// cost more than the streaming they follow.)
template <int NQ>
__device__ inline void block_sums(const double (&v)[NQ], double *part) {
    __shared__ double sh[NQ][kBlock / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        double x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) sh[k][wv] = x;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[threadIdx.x][w];
        part[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
    }
}

// Control words of the iteration loop, in device memory (KrylovCtl in kernels.h): stop is
// set by the reduction whose residual reaches the tolerance, after which every kernel of the
// later iterations returns at once; iters counts the iterations whose update ran.  They are
// written by one-workgroup kernels and read by the kernels after them on the stream.  (A
// reduction finished inside its own kernel by the last workgroup to arrive costs more than
// the separate launch on this chip: each workgroup's device-scope fence writes back its
// XCD's L2, and thousands of tickets queue on one address.)
//
// The kernels of an iteration return at once after convergence (uniform for the workgroup).
__device__ inline bool stopped(const KrylovCtl *ctl) { return ctl && *(const volatile int *)&ctl->stop; }

// after the reduction that produced r.r of an iteration (in the one workgroup that wrote it):
// count the iteration and stop unless ||r|| / ||b'|| > rtol (a NaN stops too)
__device__ inline void check_converged(KrylovCtl *ctl, const double *rr, double bnorm, double rtol) {
    if (threadIdx.x == 0) {
        ctl->iters += 1;
        if (!(sqrt(*(const volatile double *)rr) / bnorm > rtol)) ctl->stop = 1;
    }
}

// *dst[k] = sum over the nb partials of quantity k; one workgroup per quantity.  With a
// control block: nothing after convergence, and the workgroup of quantity conv (0 or 1; -1:
// none) runs the residual test on its sum.
__global__ __launch_bounds__(1024) void k_finish(const double *part, int nb, double *d0, double *d1,
                                                 KrylovCtl *ctl = nullptr, int conv = -1, double bnorm = 1.0,
                                                 double rtol = 0.0) {
    if (stopped(ctl)) return;
    __shared__ double sh[16];
    const double *src = part + (int64_t)blockIdx.x * nb;
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s += src[b];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sh[w];
        double *d = blockIdx.x == 0 ? d0 : d1;
        *d = t;
        if ((int)blockIdx.x == conv) check_converged(ctl, d, bnorm, rtol);
    }
}

// r = A x - b on free nodes (residual of -A x = -b), z = r/4; partials of r.z and r.r
__global__ __launch_bounds__(kBlock) void k_pcg_init(const double *x, const double *b, const uint8_t *mask, double *r,
                                                     double *z, double *p, double *part, Geo g) {
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
    block_sums<2>({rz, rr}, part);
}

// q = -A p on free nodes; partials of p.q
__global__ __launch_bounds__(kBlock) void k_pcg_matvec(const double *p, double *qv, const uint8_t *mask, double *part,
                                                       Geo g, const KrylovCtl *ctl) {
    if (stopped(ctl)) return;
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
    block_sums<1>({pq}, part);
}


// p = z + beta p on the free nodes with beta = rz_next / rz
__global__ __launch_bounds__(kBlock) void k_pcg_direction(double *p, const double *z, const uint8_t *mask,
                                                          const double *red_k, const double *red_next, int64_t n,
                                                          const KrylovCtl *ctl) {
    if (stopped(ctl)) return;
    const double beta = red_next[0] / red_k[0];
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        if (mask[q]) p[q] = z[q] + beta * p[q];
}

// alpha = rz/pq; x += alpha p; r -= alpha q; z = r/4; partials of the next rz and rr
__global__ __launch_bounds__(kBlock) void k_pcg_update(double *x, double *r, double *z, const double *p,
                                                       const double *qv, const uint8_t *mask, const double *red_k,
                                                       double *part, Geo g, const KrylovCtl *ctl) {
    if (stopped(ctl)) return;
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
    block_sums<2>({rz, rr}, part);
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

// ---------------------------------------------------------------------------------
// Geometric multigrid V-cycle used as the preconditioner M ~ (-L)^{-1}: vertex-centred
// coarsening by 2 in both directions, rediscretised 5-point operators (a quarter of the
// finer one per level in fine-grid units), full-weighting restriction, bilinear
// prolongation (the restriction is a quarter of its transpose), two damped-Jacobi sweeps
// before and after the coarse correction and a fixed number of sweeps from zero on the
// coarsest level.  Every piece is a symmetric linear map, so M is symmetric positive
// definite and CG stays CG.  A coarse node is free when the fine node at the same place
// is free; corrections are zero on Dirichlet nodes of every level.
constexpr double kMgOmega = 0.8;
constexpr int kMgSweeps = 2;
constexpr int kMgFusedThreads = 1024;
constexpr int64_t kMgFusedLdsBytes = 64 * 1024;  // levels fused into k_mg_fused fit in this

// A u at (i, j) for u given as a function of (i, j, q)
template <class U>
__device__ inline double mg_op_of(const U &u, int i, int j, const MgLevel &L, bool periodic) {
    const int64_t q = (int64_t)i + (int64_t)j * L.stride;
    int il = i - 1, ir = i + 1;
    if (periodic) {
        if (il < 0) il += L.ncol;
        if (ir >= L.ncol) ir -= L.ncol;
    }
    const int64_t row = (int64_t)j * L.stride;
    return L.scale * (4.0 * u(i, j, q) - u(il, j, row + il) - u(ir, j, row + ir) - u(i, j - 1, q - L.stride) -
                      u(i, j + 1, q + L.stride));
}

__device__ inline double mg_op(const double *u, int i, int j, const MgLevel &L, bool periodic) {
    return mg_op_of([u](int, int, int64_t q) { return u[q]; }, i, j, L, periodic);
}

// Each operation is written once per node t (i = t % ncol, j = t / ncol) and used both by
// the one-launch-per-operation kernels on the fine levels and by k_mg_fused on the coarse
// levels held in LDS, so the two give bit-identical corrections.

// out = in + omega D^{-1} (f - A in) on free nodes (in == nullptr: in = 0), 0 elsewhere
__device__ inline void mg_jacobi_at(const double *in, double *out, const double *f, const MgLevel &L, bool periodic,
                                    int64_t t) {
    const int i = (int)(t % L.ncol), j = (int)(t / L.ncol);
    const int64_t q = (int64_t)i + (int64_t)j * L.stride;
    if (!L.mask[q]) {
        out[q] = 0.0;
        return;
    }
    const double inv_d = kMgOmega / (4.0 * L.scale);
    out[q] = in ? in[q] + inv_d * (f[q] - mg_op(in, i, j, L, periodic)) : inv_d * f[q];
}

__device__ inline void mg_residual_at(const double *u, const double *f, double *r, const MgLevel &L, bool periodic,
                                      int64_t t) {
    const int i = (int)(t % L.ncol), j = (int)(t / L.ncol);
    const int64_t q = (int64_t)i + (int64_t)j * L.stride;
    r[q] = L.mask[q] ? f[q] - mg_op(u, i, j, L, periodic) : 0.0;
}

// coarse f = full weighting of the fine residual (weights 1/4, 1/8, 1/16); t over the coarse level
__device__ inline void mg_restrict_at(const double *rf, const MgLevel &F, double *fc, const MgLevel &Cl,
                                      bool periodic, int64_t t) {
    const int I = (int)(t % Cl.ncol), J = (int)(t / Cl.ncol);
    const int64_t qc = (int64_t)I + (int64_t)J * Cl.stride;
    if (!Cl.mask[qc]) {
        fc[qc] = 0.0;
        return;
    }
    double s = 0.0;
    for (int b = -1; b <= 1; ++b) {
        const double wb = b == 0 ? 0.5 : 0.25;
        const int64_t row = (int64_t)(2 * J + b) * F.stride;
        for (int a = -1; a <= 1; ++a) {
            int i = 2 * I + a;
            if (periodic) {
                if (i < 0) i += F.ncol;
                if (i >= F.ncol) i -= F.ncol;
            }
            s += (a == 0 ? 0.5 : 0.25) * wb * rf[row + i];
        }
    }
    fc[qc] = s;
}

// fine u += bilinear interpolation of the coarse correction (free fine nodes); t over the fine level
__device__ inline void mg_prolong_at(const double *uc, const MgLevel &Cl, double *uf, const MgLevel &F, bool periodic,
                                     int64_t t) {
    const int i = (int)(t % F.ncol), j = (int)(t / F.ncol);
    const int64_t q = (int64_t)i + (int64_t)j * F.stride;
    if (!F.mask[q]) return;
    const int I0 = i >> 1, J0 = j >> 1;
    int I1 = I0 + (i & 1);
    const int J1 = J0 + (j & 1);
    if (periodic && I1 >= Cl.ncol) I1 -= Cl.ncol;
    const double v = 0.25 * (uc[I0 + (int64_t)J0 * Cl.stride] + uc[I1 + (int64_t)J0 * Cl.stride] +
                             uc[I0 + (int64_t)J1 * Cl.stride] + uc[I1 + (int64_t)J1 * Cl.stride]);
    uf[q] += v;
}

__global__ void k_mg_jacobi(const double *in, double *out, const double *f, MgLevel L, bool periodic,
                            const KrylovCtl *ctl) {
    if (stopped(ctl)) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (int64_t)L.ncol * L.ny) mg_jacobi_at(in, out, f, L, periodic, t);
}

// The two pre-smoothing sweeps from zero and the residual in one pass, over tiles of
// kPsW x kPsH nodes.  The first sweep is u1 = omega D^{-1} f on free nodes, so it is
// computed in LDS on the tile with a two-node halo; the second, u2 = u1 + omega D^{-1}
// (f - A u1), on the tile with a one-node halo; then u = u2 and r = f - A u2 on the tile.
// The halo columns wrap on x-periodic levels, nodes outside the level count as Dirichlet
// (a free node never reaches them).  Each value is the same expression, in the same order,
// as in mg_jacobi_at / mg_op_of / mg_residual_at, so the result has the same bits.
static_assert(kMgSweeps == 2, "k_mg_presmooth and k_mg_postsmooth fuse exactly two sweeps");
constexpr int kPsW = 64, kPsH = 4;

__global__ __launch_bounds__(kPsW * kPsH) void k_mg_presmooth(const double *__restrict__ f, double *__restrict__ u,
                                                              double *__restrict__ r, MgLevel L, bool periodic,
                                                              const KrylovCtl *ctl) {
    if (stopped(ctl)) return;
    constexpr int W2 = kPsW + 4, H2 = kPsH + 4, W1 = kPsW + 2, H1 = kPsH + 2;
    __shared__ double u1s[H2][W2], fs[H2][W2], u2s[H1][W1];
    __shared__ uint8_t ms[H2][W2];
    const int ntx = (L.ncol + kPsW - 1) / kPsW;
    const int x0 = (blockIdx.x % ntx) * kPsW, y0 = (blockIdx.x / ntx) * kPsH;
    const double inv_d = kMgOmega / (4.0 * L.scale);
    for (int k = threadIdx.x; k < H2 * W2; k += blockDim.x) {
        const int c = k % W2, rr = k / W2;
        int gx = x0 - 2 + c;
        const int gy = y0 - 2 + rr;
        if (periodic) gx = ((gx % L.ncol) + L.ncol) % L.ncol;
        uint8_t m = 0;
        double fv = 0.0;
        if (gx >= 0 && gx < L.ncol && gy >= 0 && gy < L.ny) {
            const int64_t q = (int64_t)gx + (int64_t)gy * L.stride;
            m = L.mask[q];
            fv = f[q];
        }
        ms[rr][c] = m;
        fs[rr][c] = fv;
        u1s[rr][c] = m ? inv_d * fv : 0.0;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < H1 * W1; k += blockDim.x) {
        const int c = k % W1 + 1, rr = k / W1 + 1;  // coordinates in the halo-2 arrays
        double v = 0.0;
        if (ms[rr][c]) {
            const double op = L.scale * (4.0 * u1s[rr][c] - u1s[rr][c - 1] - u1s[rr][c + 1] - u1s[rr - 1][c] -
                                         u1s[rr + 1][c]);
            v = u1s[rr][c] + inv_d * (fs[rr][c] - op);
        }
        u2s[rr - 1][c - 1] = v;
    }
    __syncthreads();
    const int c = threadIdx.x % kPsW, rr = threadIdx.x / kPsW;
    const int gx = x0 + c, gy = y0 + rr;
    if (gx >= L.ncol || gy >= L.ny) return;  // no barrier follows
    const int64_t q = (int64_t)gx + (int64_t)gy * L.stride;
    const int a = c + 1, b = rr + 1;  // coordinates in u2s
    u[q] = u2s[b][a];
    double rv = 0.0;
    if (ms[b + 1][a + 1]) {
        const double op = L.scale * (4.0 * u2s[b][a] - u2s[b][a - 1] - u2s[b][a + 1] - u2s[b - 1][a] - u2s[b + 1][a]);
        rv = fs[b + 1][a + 1] - op;
    }
    r[q] = rv;
}

__global__ void k_mg_residual(const double *u, const double *f, double *r, MgLevel L, bool periodic,
                              const KrylovCtl *ctl) {
    if (stopped(ctl)) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (int64_t)L.ncol * L.ny) mg_residual_at(u, f, r, L, periodic, t);
}

__global__ void k_mg_restrict(const double *rf, MgLevel F, double *fc, MgLevel Cl, bool periodic,
                              const KrylovCtl *ctl) {
    if (stopped(ctl)) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (int64_t)Cl.ncol * Cl.ny) mg_restrict_at(rf, F, fc, Cl, periodic, t);
}

__global__ void k_mg_prolong(const double *uc, MgLevel Cl, double *uf, MgLevel F, bool periodic,
                             const KrylovCtl *ctl) {
    if (stopped(ctl)) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (int64_t)F.ncol * F.ny) mg_prolong_at(uc, Cl, uf, F, periodic, t);
}

// Prolongation and the two post-smoothing sweeps in one pass over the same tiles:
// c = u + bilinear(coarse correction) on free nodes (k_mg_prolong) on the tile with a
// two-node halo, the first sweep on a one-node halo, the second on the tile, written to
// out (which must not alias u).  Same expressions in the same order as the launched
// kernels, so the same bits.
__global__ __launch_bounds__(kPsW * kPsH) void k_mg_postsmooth(const double *__restrict__ u,
                                                               const double *__restrict__ uc, MgLevel Cl,
                                                               const double *__restrict__ f, double *__restrict__ out,
                                                               MgLevel L, bool periodic, const KrylovCtl *ctl,
                                                               double *dot_part) {
    if (stopped(ctl)) return;
    constexpr int W2 = kPsW + 4, H2 = kPsH + 4, W1 = kPsW + 2, H1 = kPsH + 2;
    __shared__ double cs[H2][W2], fs[H2][W2], s1[H1][W1];
    __shared__ uint8_t ms[H2][W2];
    const int ntx = (L.ncol + kPsW - 1) / kPsW;
    const int x0 = (blockIdx.x % ntx) * kPsW, y0 = (blockIdx.x / ntx) * kPsH;
    const double inv_d = kMgOmega / (4.0 * L.scale);
    for (int k = threadIdx.x; k < H2 * W2; k += blockDim.x) {
        const int c = k % W2, rr = k / W2;
        int gx = x0 - 2 + c;
        const int gy = y0 - 2 + rr;
        if (periodic) gx = ((gx % L.ncol) + L.ncol) % L.ncol;
        uint8_t m = 0;
        double fv = 0.0, cv = 0.0;
        if (gx >= 0 && gx < L.ncol && gy >= 0 && gy < L.ny) {
            const int64_t q = (int64_t)gx + (int64_t)gy * L.stride;
            m = L.mask[q];
            fv = f[q];
            cv = u[q];
            if (m) {
                const int I0 = gx >> 1, J0 = gy >> 1;
                int I1 = I0 + (gx & 1);
                const int J1 = J0 + (gy & 1);
                if (periodic && I1 >= Cl.ncol) I1 -= Cl.ncol;
                const double v = 0.25 * (uc[I0 + (int64_t)J0 * Cl.stride] + uc[I1 + (int64_t)J0 * Cl.stride] +
                                         uc[I0 + (int64_t)J1 * Cl.stride] + uc[I1 + (int64_t)J1 * Cl.stride]);
                cv += v;
            }
        }
        ms[rr][c] = m;
        fs[rr][c] = fv;
        cs[rr][c] = cv;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < H1 * W1; k += blockDim.x) {
        const int c = k % W1 + 1, rr = k / W1 + 1;
        double v = 0.0;
        if (ms[rr][c]) {
            const double op =
                L.scale * (4.0 * cs[rr][c] - cs[rr][c - 1] - cs[rr][c + 1] - cs[rr - 1][c] - cs[rr + 1][c]);
            v = cs[rr][c] + inv_d * (fs[rr][c] - op);
        }
        s1[rr - 1][c - 1] = v;
    }
    __syncthreads();
    const int c = threadIdx.x % kPsW, rr = threadIdx.x / kPsW;
    const int gx = x0 + c, gy = y0 + rr;
    double rz = 0.0;
    if (gx < L.ncol && gy < L.ny) {
        const int a = c + 1, b = rr + 1;
        double v = 0.0;
        if (ms[b + 1][a + 1]) {
            const double op = L.scale * (4.0 * s1[b][a] - s1[b][a - 1] - s1[b][a + 1] - s1[b - 1][a] - s1[b + 1][a]);
            v = s1[b][a] + inv_d * (fs[b + 1][a + 1] - op);
            rz = fs[b + 1][a + 1] * v;
        }
        out[(int64_t)gx + (int64_t)gy * L.stride] = v;
    }
    // the PCG's r.z (f = r on level 0), one partial per tile: k_dot_tiles' order
    if (dot_part) block_sums<1>({rz}, dot_part);
}

// r.z over the free nodes of level 0 with the partials of k_mg_postsmooth (same tiles, same
// node per thread, same products), for the V-cycle that does not end in that kernel
__global__ __launch_bounds__(kPsW * kPsH) void k_dot_tiles(const double *__restrict__ r, const double *__restrict__ z,
                                                          MgLevel L, const KrylovCtl *ctl, double *dot_part) {
    if (stopped(ctl)) return;
    const int ntx = (L.ncol + kPsW - 1) / kPsW;
    const int x0 = (blockIdx.x % ntx) * kPsW, y0 = (blockIdx.x / ntx) * kPsH;
    const int gx = x0 + (int)threadIdx.x % kPsW, gy = y0 + (int)threadIdx.x / kPsW;
    double rz = 0.0;
    if (gx < L.ncol && gy < L.ny) {
        const int64_t q = (int64_t)gx + (int64_t)gy * L.stride;
        if (L.mask[q]) rz = r[q] * z[q];
    }
    block_sums<1>({rz}, dot_part);
}

static int presmooth_grid(const MgLevel &L) {
    return ((L.ncol + kPsW - 1) / kPsW) * ((L.ny + kPsH - 1) / kPsH);
}

// The bottom of the V-cycle in one workgroup: levels first..last (each with stride = ncol)
// live in LDS (u, u2, f, r as doubles, then the mask, each piece padded to 16 bytes), and
// every sweep, residual, restriction and prolongation of the recursive cycle is one pass
// over a level followed by a barrier, in the order the launched version runs them.  The
// right-hand side of level `first` comes from global memory, its correction goes back there.
// Saves ~7 launches per level and the coarsest level's 4 max(ncol, ny) sweep launches.
struct MgFusedArgs {
    MgLevel lev[kMgMaxLevels];  // geometry and global masks; array pointers are replaced by LDS
    int first, last;
    bool per;
    const double *f_in;
    double *u_out;
    const KrylovCtl *ctl;
};

__device__ inline int64_t mg_fused_piece(int64_t n) { return (n * 8 + 15) & ~(int64_t)15; }

__device__ inline MgLevel mg_fused_level(const MgFusedArgs &a, unsigned char *smem, int l) {
    int64_t off = 0;
    for (int k = a.first; k < l; ++k) {
        const int64_t nk = (int64_t)a.lev[k].ncol * a.lev[k].ny;
        off += 4 * mg_fused_piece(nk) + ((nk + 15) & ~(int64_t)15);
    }
    MgLevel x = a.lev[l];
    const int64_t n = (int64_t)x.ncol * x.ny, P = mg_fused_piece(n);
    x.u = reinterpret_cast<double *>(smem + off);
    x.u2 = reinterpret_cast<double *>(smem + off + P);
    x.f = reinterpret_cast<double *>(smem + off + 2 * P);
    x.r = reinterpret_cast<double *>(smem + off + 3 * P);
    x.mask = smem + off + 4 * P;
    return x;
}

__global__ __launch_bounds__(kMgFusedThreads) void k_mg_fused(MgFusedArgs a) {
    if (stopped(a.ctl)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, nt = blockDim.x;
    const bool per = a.per;
    for (int l = a.first; l <= a.last; ++l) {
        const MgLevel x = mg_fused_level(a, smem, l);
        const int64_t n = (int64_t)x.ncol * x.ny;
        for (int64_t t = tid; t < n; t += nt) x.mask[t] = a.lev[l].mask[t];
        if (l == a.first)
            for (int64_t t = tid; t < n; t += nt) x.f[t] = a.f_in[t];
    }
    __syncthreads();
    unsigned sel = 0;  // bit l set: the current correction of level l is in u2
    auto cur = [&](const MgLevel &x, int l) { return ((sel >> l) & 1u) ? x.u2 : x.u; };
    auto oth = [&](const MgLevel &x, int l) { return ((sel >> l) & 1u) ? x.u : x.u2; };
    auto sweeps = [&](const MgLevel &x, int l, int k0, int count) {
        const int64_t n = (int64_t)x.ncol * x.ny;
        for (int k = k0; k < count; ++k) {
            const double *in = k == 0 ? nullptr : cur(x, l);
            double *out = k == 0 ? x.u : oth(x, l);
            for (int64_t t = tid; t < n; t += nt) mg_jacobi_at(in, out, x.f, x, per, t);
            __syncthreads();
            if (k == 0)
                sel &= ~(1u << l);
            else
                sel ^= 1u << l;
        }
    };
    for (int l = a.first; l < a.last; ++l) {
        const MgLevel x = mg_fused_level(a, smem, l), c = mg_fused_level(a, smem, l + 1);
        sweeps(x, l, 0, kMgSweeps);
        const int64_t n = (int64_t)x.ncol * x.ny, nc = (int64_t)c.ncol * c.ny;
        for (int64_t t = tid; t < n; t += nt) mg_residual_at(cur(x, l), x.f, x.r, x, per, t);
        __syncthreads();
        for (int64_t t = tid; t < nc; t += nt) mg_restrict_at(x.r, x, c.f, c, per, t);
        __syncthreads();
    }
    {
        const MgLevel x = mg_fused_level(a, smem, a.last);
        sweeps(x, a.last, 0, 4 * max(x.ncol, x.ny));
    }
    for (int l = a.last - 1; l >= a.first; --l) {
        const MgLevel x = mg_fused_level(a, smem, l), c = mg_fused_level(a, smem, l + 1);
        const int64_t n = (int64_t)x.ncol * x.ny;
        for (int64_t t = tid; t < n; t += nt) mg_prolong_at(cur(c, l + 1), c, cur(x, l), x, per, t);
        __syncthreads();
        // post-smoothing continues from the current correction: sweeps 1..kMgSweeps
        sweeps(x, l, 1, kMgSweeps + 1);
    }
    const MgLevel x = mg_fused_level(a, smem, a.first);
    const double *u = cur(x, a.first);
    for (int64_t t = tid; t < (int64_t)x.ncol * x.ny; t += nt) a.u_out[t] = u[t];
}

// coarse mask: free where the fine node at (2I, 2J) is free
__global__ void k_mg_coarse_mask(const uint8_t *mf, MgLevel F, uint8_t *mc, MgLevel Cl) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)Cl.ncol * Cl.ny) return;
    const int I = (int)(t % Cl.ncol), J = (int)(t / Cl.ncol);
    mc[(int64_t)I + (int64_t)J * Cl.stride] = mf[(int64_t)(2 * I) + (int64_t)(2 * J) * F.stride];
}


// alpha = rz/pq; x += alpha p; r -= alpha q; partials of r.r
__global__ __launch_bounds__(kBlock) void k_pcg_update_nz(double *x, double *r, const double *p, const double *qv,
                                                          const uint8_t *mask, const double *red_k, double *part,
                                                          int64_t n, const KrylovCtl *ctl) {
    if (stopped(ctl)) return;
    const double alpha = red_k[0] / red_k[1];
    double rr = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        if (!mask[q]) continue;
        x[q] += alpha * p[q];
        const double rv = r[q] - alpha * qv[q];
        r[q] = rv;
        rr += rv * rv;
    }
    block_sums<1>({rr}, part);
}

__global__ void k_copy_free(const double *src, double *dst, const uint8_t *mask, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = mask[q] ? src[q] : 0.0;
}

static int mg_grid(const MgLevel &L) { return (int)(((int64_t)L.ncol * L.ny + kBlock - 1) / kBlock); }

int krylov_setup_multigrid(KrylovWork &w, const Geo &g, hipStream_t s) {
    const bool per = g.periodic_x != 0;
    MgLevel L0;
    L0.ncol = per ? g.gmi - 1 : g.nx;
    L0.ny = g.ny;
    L0.stride = g.nx;
    L0.scale = 1.0;
    L0.mask = w.mask;
    const int64_t n0 = (int64_t)g.nx * g.ny;
    if (hipMalloc(&L0.u, sizeof(double) * n0) != hipSuccess || hipMalloc(&L0.u2, sizeof(double) * n0) != hipSuccess ||
        hipMalloc(&L0.r, sizeof(double) * n0) != hipSuccess)
        return -1;
    if (hipMemsetAsync(L0.u, 0, sizeof(double) * n0, s) != hipSuccess ||
        hipMemsetAsync(L0.u2, 0, sizeof(double) * n0, s) != hipSuccess ||
        hipMemsetAsync(L0.r, 0, sizeof(double) * n0, s) != hipSuccess)
        return -1;
    L0.f = nullptr;
    w.lev[0] = L0;
    int nl = 1;
    while (nl < kMgMaxLevels) {
        const MgLevel &F = w.lev[nl - 1];
        const int cx = per ? F.ncol : F.ncol - 1, cy = F.ny - 1;
        if ((cx & 1) || (cy & 1) || cx / 2 < 4 || cy / 2 < 4) break;
        MgLevel Cl;
        Cl.ncol = per ? F.ncol / 2 : (F.ncol - 1) / 2 + 1;
        Cl.ny = (F.ny - 1) / 2 + 1;
        Cl.stride = Cl.ncol;
        Cl.scale = F.scale * 0.25;
        const int64_t nc = (int64_t)Cl.ncol * Cl.ny;
        if (hipMalloc(&Cl.u, sizeof(double) * nc) != hipSuccess || hipMalloc(&Cl.u2, sizeof(double) * nc) != hipSuccess ||
            hipMalloc(&Cl.f, sizeof(double) * nc) != hipSuccess || hipMalloc(&Cl.r, sizeof(double) * nc) != hipSuccess ||
            hipMalloc(&Cl.mask, nc) != hipSuccess)
            return -1;
        if (hipMemsetAsync(Cl.u, 0, sizeof(double) * nc, s) != hipSuccess ||
            hipMemsetAsync(Cl.u2, 0, sizeof(double) * nc, s) != hipSuccess)
            return -1;
        k_mg_coarse_mask<<<mg_grid(Cl), kBlock, 0, s>>>(F.mask, F, Cl.mask, Cl);
        w.lev[nl++] = Cl;
    }
    w.nlev = nl;
    // the coarsest levels whose LDS copies fit together go to k_mg_fused (level 0 never:
    // its stride is the grid's row length and it is the PCG's own vector)
    w.mg_fused_first = nl;
    w.mg_fused_lds = 0;
    int64_t bytes = 0;
    for (int l = nl - 1; l >= 1; --l) {
        const int64_t n = (int64_t)w.lev[l].ncol * w.lev[l].ny;
        const int64_t b = 4 * ((n * 8 + 15) & ~(int64_t)15) + ((n + 15) & ~(int64_t)15);
        if (bytes + b > kMgFusedLdsBytes) break;
        bytes += b;
        w.mg_fused_first = l;
    }
    if (w.mg_fused_first < nl) w.mg_fused_lds = (int)bytes;
    return nl;
}

// z = M f on level l; the result ends in *result (one of the level's two buffers, or `out`
// on level 0 when given).  dot_part (level 0 with out): the tile partials of f.out, by the
// post-smoothing kernel or by k_dot_tiles after the launched sweeps (level0_tiles of them).
static void mg_vcycle(KrylovWork &w, int l, const double *f, double *out, double **result, bool per, hipStream_t s,
                      double *dot_part = nullptr) {
    MgLevel &L = w.lev[l];
    L.f = const_cast<double *>(f);
    const int grid = mg_grid(L);
    double *cur = L.u, *oth = L.u2;
    if (w.precond == kPrecondMultigrid && l == w.mg_fused_first && l >= 1) {
        MgFusedArgs a;
        for (int k = 0; k < w.nlev; ++k) a.lev[k] = w.lev[k];
        a.first = l;
        a.last = w.nlev - 1;
        a.per = per;
        a.f_in = f;
        a.u_out = L.u;
        a.ctl = w.ctl;
        k_mg_fused<<<1, kMgFusedThreads, w.mg_fused_lds, s>>>(a);
        *result = L.u;
        return;
    }
    if (l == w.nlev - 1) {
        const int sweeps = 4 * std::max(L.ncol, L.ny);
        k_mg_jacobi<<<grid, kBlock, 0, s>>>(nullptr, cur, f, L, per, w.ctl);
        for (int k = 1; k < sweeps; ++k) {
            k_mg_jacobi<<<grid, kBlock, 0, s>>>(cur, oth, f, L, per, w.ctl);
            std::swap(cur, oth);
        }
        *result = cur;
        return;
    }
    if (w.precond == kPrecondMultigrid) {
        k_mg_presmooth<<<presmooth_grid(L), kPsW * kPsH, 0, s>>>(f, cur, L.r, L, per, w.ctl);
    } else {
        k_mg_jacobi<<<grid, kBlock, 0, s>>>(nullptr, cur, f, L, per, w.ctl);
        for (int k = 1; k < kMgSweeps; ++k) {
            k_mg_jacobi<<<grid, kBlock, 0, s>>>(cur, oth, f, L, per, w.ctl);
            std::swap(cur, oth);
        }
        k_mg_residual<<<grid, kBlock, 0, s>>>(cur, f, L.r, L, per, w.ctl);
    }
    MgLevel &Cl = w.lev[l + 1];
    k_mg_restrict<<<mg_grid(Cl), kBlock, 0, s>>>(L.r, L, Cl.f, Cl, per, w.ctl);
    double *uc = nullptr;
    mg_vcycle(w, l + 1, Cl.f, nullptr, &uc, per, s);
    if (w.precond == kPrecondMultigrid) {
        double *dst = (l == 0 && out) ? out : oth;
        k_mg_postsmooth<<<presmooth_grid(L), kPsW * kPsH, 0, s>>>(cur, uc, Cl, f, dst, L, per, w.ctl,
                                                                  dst == out ? dot_part : nullptr);
        *result = dst;
        return;
    }
    k_mg_prolong<<<grid, kBlock, 0, s>>>(uc, Cl, cur, L, per, w.ctl);
    for (int k = 0; k < kMgSweeps; ++k) {
        double *dst = (l == 0 && out && k == kMgSweeps - 1) ? out : oth;
        k_mg_jacobi<<<grid, kBlock, 0, s>>>(cur, dst, f, L, per, w.ctl);
        if (dst == out) {
            cur = out;
        } else {
            std::swap(cur, oth);
        }
    }
    if (dot_part && cur == out) k_dot_tiles<<<presmooth_grid(L), kPsW * kPsH, 0, s>>>(f, out, L, w.ctl, dot_part);
    *result = cur;
}

// partials the r.z of a multigrid iteration leaves (one per level-0 tile)
static int level0_tiles(const KrylovWork &w) { return presmooth_grid(w.lev[0]); }

// Returns the iteration count (>=0) on convergence, -1 when max_iter is reached.
//
// The loop stops after the first iteration whose residual satisfies ||r|| <= rtol ||b'||,
// as a loop that read the residual after every iteration would; the test runs on the device
// (check_converged), so the host queues several iterations between its reads of the control
// block: first batch_hint of them (the caller's count of its previous solve of the same kind,
// updated here), then one at a time (multigrid) or 16 at a time (Jacobi).  Iterations queued
// after convergence return from each kernel at once.
int krylov_pcg(KrylovWork &w, double *x, const Geo &g, double rtol, int max_iter, double *relres, int &batch_hint,
               hipStream_t s) {
    // the first failing HIP call; the residuals read after it would be meaningless
    hipError_t err = hipSuccess;
    auto ck = [&err](hipError_t e) {
        if (e != hipSuccess && err == hipSuccess) err = e;
    };
    const int64_t n = (int64_t)g.nx * g.ny;
    int grid = (int)((n + kBlock - 1) / kBlock);
    if (grid > kKrylovRedGrid) grid = kKrylovRedGrid;
    double *part = w.part;
    // red slot k of the set-up <- the partials of quantity 0 (and 1)
    auto finish = [&](double *d0, double *d1) { k_finish<<<d1 ? 2 : 1, 1024, 0, s>>>(part, grid, d0, d1); };
    const int eg = (int)((n + kBlock - 1) / kBlock);
    if (max_iter > w.nred - 3) max_iter = w.nred - 3;
    double *red = w.red;
    double *red_b = w.red + 4 * (size_t)(w.nred - 1);
    ck(hipMemsetAsync(w.red, 0, sizeof(double) * 4 * (size_t)w.nred, s));
    ck(hipMemsetAsync(w.ctl, 0, sizeof(KrylovCtl), s));
    // ||b'||: residual of the vector that holds only the Dirichlet values
    ck(hipMemcpyAsync(w.z, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    k_zero_free<<<eg, kBlock, 0, s>>>(w.z, w.mask, n);
    k_pcg_init<<<grid, kBlock, 0, s>>>(w.z, w.b, w.mask, w.q, w.r, w.p, part, g);
    finish(&red_b[0], &red_b[3]);
    k_pcg_init<<<grid, kBlock, 0, s>>>(x, w.b, w.mask, w.r, w.z, w.p, part, g);
    finish(&red[0], &red[3]);
    double *h = w.hred, *hb = w.hred + 4;  // page-locked
    KrylovCtl *hc = reinterpret_cast<KrylovCtl *>(w.hred + 8);
    ck(hipMemcpyAsync(hb, red_b, sizeof(double) * 4, hipMemcpyDeviceToHost, s));
    ck(hipMemcpyAsync(h, red, sizeof(double) * 4, hipMemcpyDeviceToHost, s));
    ck(hipStreamSynchronize(s));
    if (err != hipSuccess) return kKrylovHipError;
    double bnorm = sqrt(hb[3]);
    if (!(bnorm > 0.0)) bnorm = 1.0;
    double res = sqrt(h[3]) / bnorm;
    const bool mg = (w.precond == kPrecondMultigrid || w.precond == kPrecondMultigridLaunched) && w.nlev > 1;
    const bool per = g.periodic_x != 0;
    int it = 0;
    if (res > rtol) {
        if (mg) {
            // z = M r, red[0] = r.z, p = z
            double *zres = nullptr;
            mg_vcycle(w, 0, w.r, w.z, &zres, per, s, part);
            k_finish<<<1, 1024, 0, s>>>(part, level0_tiles(w), &red[0], nullptr);
            k_copy_free<<<eg, kBlock, 0, s>>>(w.z, w.p, w.mask, n);
        }
        int issued = 0;
        int batch = std::max(1, batch_hint);
        for (;;) {
            const int nb = std::min(batch, max_iter - issued);
            for (int b = 0; b < nb; ++b, ++issued) {
                double *rk = red + 4 * (size_t)issued, *rn = red + 4 * (size_t)(issued + 1);
                k_pcg_matvec<<<grid, kBlock, 0, s>>>(w.p, w.q, w.mask, part, g, w.ctl);
                k_finish<<<1, 1024, 0, s>>>(part, grid, &rk[1], nullptr, w.ctl);
                if (mg) {
                    k_pcg_update_nz<<<grid, kBlock, 0, s>>>(x, w.r, w.p, w.q, w.mask, rk, part, n, w.ctl);
                    k_finish<<<1, 1024, 0, s>>>(part, grid, &rn[3], nullptr, w.ctl, 0, bnorm, rtol);
                    double *zres = nullptr;
                    mg_vcycle(w, 0, w.r, w.z, &zres, per, s, part);
                    k_finish<<<1, 1024, 0, s>>>(part, level0_tiles(w), &rn[0], nullptr, w.ctl);
                } else {
                    k_pcg_update<<<grid, kBlock, 0, s>>>(x, w.r, w.z, w.p, w.q, w.mask, rk, part, g, w.ctl);
                    k_finish<<<2, 1024, 0, s>>>(part, grid, &rn[0], &rn[3], w.ctl, 1, bnorm, rtol);
                }
                k_pcg_direction<<<grid, kBlock, 0, s>>>(w.p, w.z, w.mask, rk, rn, n, w.ctl);
            }
            ck(hipMemcpyAsync(hc, w.ctl, sizeof(KrylovCtl), hipMemcpyDeviceToHost, s));
            ck(hipStreamSynchronize(s));
            if (err != hipSuccess) return kKrylovHipError;
            it = hc->iters;
            if (hc->stop || issued >= max_iter) break;
            batch = mg ? 1 : 16;
        }
        // a solve that ran out of iterations says nothing about the next one
        batch_hint = hc->stop ? it : 0;
    }
    // the residual of the last iteration that ran
    ck(hipMemcpyAsync(h, red + 4 * (size_t)it, sizeof(double) * 4, hipMemcpyDeviceToHost, s));
    if (g.periodic_x) k_sync_images<<<(g.ny + kBlock - 1) / kBlock, kBlock, 0, s>>>(x, g);
    ck(hipStreamSynchronize(s));
    if (err != hipSuccess) return kKrylovHipError;
    res = sqrt(h[3]) / bnorm;
    if (relres) *relres = res;
    return res <= rtol ? it : -1;
}

}  // namespace edgpu
