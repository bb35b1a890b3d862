// Grid kernels: density fold, FFT_X + tridiagonal Poisson solve, electric field.
#include "kernels.h"

namespace edgpu {

__global__ void k_fill(double *a, int64_t n, double v) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) a[k] = v;
}

// plain streaming copy (16 B per lane per iteration): the reference point for HBM bandwidth
// and for the scale of the FETCH_SIZE / WRITE_SIZE counters
// one 16-byte element per thread: the fastest plain copy measured on MI355X (a grid-stride
// loop over 8192 workgroups reaches about three quarters of it)
__global__ __launch_bounds__(kBlock) void k_copy_probe(const double2 *__restrict__ a, double2 *__restrict__ b, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) b[k] = a[k];
}

void launch_copy_probe(const void *a, void *b, int64_t nbytes, hipStream_t s) {
    const int64_t n = nbytes / 16;
    const int64_t blocks = (n + kBlock - 1) / kBlock;  // < 2^31 for any size hipMalloc can return here
    k_copy_probe<<<(unsigned)blocks, kBlock, 0, s>>>((const double2 *)a, (double2 *)b, n);
}

void launch_fill(double *a, int64_t n, double v, hipStream_t s) {
    if (n <= 0) return;
    int64_t b = (n + kBlock - 1) / kBlock;
    k_fill<<<(int)(b > 65536 ? 65536 : b), kBlock, 0, s>>>(a, n, v);
}

// ---------------------------------------------------------------------------------
// periodic self-fold (pic2d_ElectronDynamics.f90:1660-1668), one thread per row
__global__ void k_fold_periodic(double *rho, int nx, int ny) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ny) return;
    double *r = rho + (int64_t)j * nx;
    const int xmax = nx - 1;
    r[1] = r[1] + r[xmax];
    r[xmax - 1] = r[xmax - 1] + r[0];
    r[0] = r[xmax - 1];
    r[xmax] = r[1];
}

// wall rows x2 after the fold (pic2d_ElectronDynamics.f90:1804-1815)
__global__ void k_double_rows(double *rho, int nx, int ny) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx) return;
    rho[i] = 2.0 * rho[i];
    rho[(int64_t)(ny - 1) * nx + i] = 2.0 * rho[(int64_t)(ny - 1) * nx + i];
}

// walls on all sides, corners SURROUNDED_BY_WALL (pic2d_ElectronDynamics.f90:1817-1878)
__global__ void k_walls_box(double *rho, int nx, int ny) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nx * ny) return;
    int i = (int)(q % nx), j = (int)(q / nx);
    bool ex = (i == 0 || i == nx - 1), ey = (j == 0 || j == ny - 1);
    if (ex && ey) rho[q] = 4.0 * rho[q];
    else if (ex || ey) rho[q] = 2.0 * rho[q];
}

void launch_fold_walls(double *rho, const Geo &g, hipStream_t s) {
    if (g.periodic_x) {
        k_fold_periodic<<<(g.ny + kBlock - 1) / kBlock, kBlock, 0, s>>>(rho, g.nx, g.ny);
        k_double_rows<<<(g.nx + kBlock - 1) / kBlock, kBlock, 0, s>>>(rho, g.nx, g.ny);
    } else {
        int64_t n = (int64_t)g.nx * g.ny;
        k_walls_box<<<(int)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(rho, g.nx, g.ny);
    }
}

// ---------------------------------------------------------------------------------
// Radix-2 complex FFT of length N held in LDS (in place, bit-reversed load).
// tw[k] = exp(-2 pi i k/N), k < N/2; sign = -1 forward, +1 inverse.
__device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ void lds_fft(double2 *buf, const double2 *tw, int N, int logN, bool inverse) {
    for (int st = 1; st <= logN; ++st) {
        const int m = 1 << st, half = m >> 1, tstride = N >> st;
        for (int b = threadIdx.x; b < (N >> 1); b += blockDim.x) {
            const int grp = b >> (st - 1), pos = b & (half - 1);
            const int i0 = grp * m + pos, i1 = i0 + half;
            double2 w = tw[pos * tstride];
            if (inverse) w.y = -w.y;
            const double2 t = cmul(w, buf[i1]);
            const double2 u = buf[i0];
            buf[i0] = make_double2(u.x + t.x, u.y + t.y);
            buf[i1] = make_double2(u.x - t.x, u.y - t.y);
        }
        __syncthreads();
    }
}

__device__ inline int bitrev(int v, int logN) { return (int)(__brev((unsigned)v) >> (32 - logN)); }

// forward transform of row j of rho = factor*(rho_i - rho_e) with wall rows replaced by the
// wall potential (pic2d_ElectricFieldCalc_FFT_X.f90:72-145, 160-324): xband(0)=rho(N).  On
// the nodes of a dielectric wall (no potential, and a surface-charge row given) the row keeps
// half its doubled wall density plus the surface charge, 0.5 rho + factor sigma
// (pic2d_ElectricFieldCalc_FFT_X.f90:104-108, 127-131).
__global__ __launch_bounds__(kBlock) void k_poisson_forward(const double *rho_e, const double *rho_i, double factor,
                                                            const double *wall_below, const double *wall_above,
                                                            const double *sigma_below, const double *sigma_above,
                                                            const double2 *tw, double2 *spec, int nx, int gmj, int N,
                                                            int logN) {
    extern __shared__ double2 buf[];
    const int j = blockIdx.x;
    const int64_t row = (int64_t)j * nx;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const int src = (i == 0) ? N : i;
        double v = factor * (rho_i[row + src] - rho_e[row + src]);
        if (j == 0) {
            if (wall_below[src] == wall_below[src]) v = wall_below[src];
            else if (sigma_below) v = 0.5 * v + factor * sigma_below[src];
        }
        if (j == gmj) {
            if (wall_above[src] == wall_above[src]) v = wall_above[src];
            else if (sigma_above) v = 0.5 * v + factor * sigma_above[src];
        }
        buf[bitrev(i, logN)] = make_double2(v, 0.0);
    }
    __syncthreads();
    lds_fft(buf, tw, N, logN, false);
    const int nm = N / 2 + 1;
    for (int n = threadIdx.x; n < nm; n += blockDim.x) {
        double2 c = buf[n];
        if (n == 0 || n == N / 2) c.y = 0.0;
        spec[(int64_t)j * nm + n] = c;
    }
}

// Thomas sweep per Fourier mode (pic2d_ElectricFieldCalc_FFT_X.f90:680-759): rhs_j =
// (-rho_j - rhs_{j-1}) a_j, phi_j = rhs_j - a_j phi_{j+1}.  A metal wall has a_0 = 0 (a_gmj =
// 0), so its row keeps the wall value; a dielectric wall below starts from rhs_0 = -rho_0 a_0
// (zero for mode 0 when both walls are dielectric), one above ends with rhs_gmj = (-rho_gmj -
// rhs_{gmj-1}) a_gmj, and row 0 becomes phi_0 = rhs_0 - a_0 phi_1.
//
// One mode has only gmj sequential rows, and one thread per mode leaves most of the GPU idle
// waiting on memory.  The interior rows 1..gmj-1 are therefore split into kThChunks chunks
// with a thread per (mode, chunk).  Both recurrences are affine in the carried value
// (forward r -> -a (y + r), backward p -> r - a p), so each thread first composes its chunk
// into one map v_out = M v_in + C, the maps of the chunks before it (in sweep order) give its
// entry value, and it then runs the reference's recurrence over its own rows from there.
// Only the chunk entry values are rounded differently from the sequential sweep; every |a| < 1,
// so the composed maps are contractions.
constexpr int kThModes = 16, kThChunks = 64, kThBatch = 8;

// rhs_0 of mode n from row 0 (the wall row's transform)
__device__ inline double2 thomas_rhs0(double2 y0, double a0, bool two_dielectric, int n) {
    if (two_dielectric && n == 0) return make_double2(0.0, 0.0);
    if (a0 == 0.0) return y0;
    return make_double2(-y0.x * a0, -y0.y * a0);
}

__global__ __launch_bounds__(kThModes *kThChunks) void k_poisson_thomas(double2 *spec, const double *__restrict__ a_eq,
                                                                        int nm, int gmj, int two_dielectric) {
    __shared__ double sM[kThChunks][kThModes];
    __shared__ double2 sC[kThChunks][kThModes];
    const int lm = threadIdx.x % kThModes, c = threadIdx.x / kThModes;
    const int n = blockIdx.x * kThModes + lm;
    const bool on = n < nm;
    const int per = (gmj - 1 + kThChunks - 1) / kThChunks;
    const int j0 = 1 + c * per, j1 = on ? min(gmj, j0 + per) : j0;  // rows [j0, j1), maybe empty
    auto at = [nm, n](int j) { return (int64_t)j * nm + n; };

    // forward sweep: compose the chunk, r_end = M r_in + C
    double M = 1.0;
    double2 Cc = make_double2(0.0, 0.0);
    for (int jb = j0; jb < j1; jb += kThBatch) {
        double a[kThBatch];
        double2 y[kThBatch];
#pragma unroll
        for (int q = 0; q < kThBatch; ++q) {
            const int j = min(jb + q, j1 - 1);
            a[q] = a_eq[at(j)];
            y[q] = spec[at(j)];
        }
#pragma unroll
        for (int q = 0; q < kThBatch; ++q) {
            if (jb + q < j1) {
                M = -a[q] * M;
                Cc = make_double2(-a[q] * (y[q].x + Cc.x), -a[q] * (y[q].y + Cc.y));
            }
        }
    }
    sM[c][lm] = M;
    sC[c][lm] = Cc;
    __syncthreads();
    const double2 rhs0 = on ? thomas_rhs0(spec[at(0)], a_eq[at(0)], two_dielectric != 0, n) : make_double2(0.0, 0.0);
    double2 r = rhs0;
    for (int k = 0; k < c; ++k) r = make_double2(sM[k][lm] * r.x + sC[k][lm].x, sM[k][lm] * r.y + sC[k][lm].y);
    for (int jb = j0; jb < j1; jb += kThBatch) {
        double a[kThBatch];
        double2 y[kThBatch];
#pragma unroll
        for (int q = 0; q < kThBatch; ++q) {
            const int j = min(jb + q, j1 - 1);
            a[q] = a_eq[at(j)];
            y[q] = spec[at(j)];
        }
#pragma unroll
        for (int q = 0; q < kThBatch; ++q) {
            if (jb + q < j1) {
                r = make_double2((-y[q].x - r.x) * a[q], (-y[q].y - r.y) * a[q]);
                spec[at(jb + q)] = r;
            }
        }
    }
    // the wall above: a dielectric one continues the sweep from rhs_{gmj-1} (the thread whose
    // chunk ends at row gmj-1 holds it)
    if (on && j1 == gmj && j0 < j1) {
        const double ag = a_eq[at(gmj)];
        if (ag != 0.0) {
            const double2 yg = spec[at(gmj)];
            spec[at(gmj)] = make_double2((-yg.x - r.x) * ag, (-yg.y - r.y) * ag);
        }
    }
    __syncthreads();  // the LDS maps are reused below; rhs_gmj is read below

    // back substitution, rows j1-1 down to j0: p_out = M p_in + C
    M = 1.0;
    Cc = make_double2(0.0, 0.0);
    for (int jt = j1; jt > j0; jt -= kThBatch) {
        double a[kThBatch];
        double2 y[kThBatch];
#pragma unroll
        for (int q = 0; q < kThBatch; ++q) {
            const int j = max(jt - 1 - q, j0);
            a[q] = a_eq[at(j)];
            y[q] = spec[at(j)];  // this thread's own forward results
        }
#pragma unroll
        for (int q = 0; q < kThBatch; ++q) {
            if (jt - 1 - q >= j0) {
                M = -a[q] * M;
                Cc = make_double2(y[q].x - a[q] * Cc.x, y[q].y - a[q] * Cc.y);
            }
        }
    }
    sM[c][lm] = M;
    sC[c][lm] = Cc;
    __syncthreads();
    double2 p = on ? spec[at(gmj)] : make_double2(0.0, 0.0);  // rhs_gmj
    for (int k = kThChunks - 1; k > c; --k)
        p = make_double2(sM[k][lm] * p.x + sC[k][lm].x, sM[k][lm] * p.y + sC[k][lm].y);
    for (int jt = j1; jt > j0; jt -= kThBatch) {
        double a[kThBatch];
        double2 y[kThBatch];
#pragma unroll
        for (int q = 0; q < kThBatch; ++q) {
            const int j = max(jt - 1 - q, j0);
            a[q] = a_eq[at(j)];
            y[q] = spec[at(j)];
        }
#pragma unroll
        for (int q = 0; q < kThBatch; ++q) {
            if (jt - 1 - q >= j0) {
                p = make_double2(y[q].x - a[q] * p.x, y[q].y - a[q] * p.y);
                spec[at(jt - 1 - q)] = p;
            }
        }
    }
    // row 0: phi_0 = rhs_0 - a_0 phi_1 (the wall value itself when a_0 = 0), by the thread that
    // holds phi_1
    if (on && j0 == 1 && j0 < j1) {
        const double a0 = a_eq[at(0)];
        spec[at(0)] = make_double2(rhs0.x - a0 * p.x, rhs0.y - a0 * p.y);
    }
}

// inverse transform of row j; phi(N)=phi(0), phi(N+1)=phi(1)
// (pic2d_ElectricFieldCalc_FFT_X.f90:1113-1130)
__global__ __launch_bounds__(kBlock) void k_poisson_inverse(const double2 *spec, const double2 *tw, double *phi, int nx,
                                                            int N, int logN) {
    extern __shared__ double2 buf[];
    const int j = blockIdx.x;
    const int nm = N / 2 + 1;
    for (int n = threadIdx.x; n < N; n += blockDim.x) {
        double2 c;
        if (n < nm) {
            c = spec[(int64_t)j * nm + n];
            if (n == 0 || n == N / 2) c.y = 0.0;
        } else {
            c = spec[(int64_t)j * nm + (N - n)];
            c.y = -c.y;
        }
        buf[bitrev(n, logN)] = c;
    }
    __syncthreads();
    lds_fft(buf, tw, N, logN, true);
    const double inv = 1.0 / (double)N;
    double *out = phi + (int64_t)j * nx;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const double v = buf[i].x * inv;
        out[i] = v;
        if (i == 0) out[N] = v;
        if (i == 1) out[N + 1] = v;
    }
}

// Twiddles tw[k] = exp(-2 pi i k / N), k < N/2, and the Thomas coefficients of PREPARE_SYS_Y
// (pic2d_Prepare_FFTX_SYSY.f90:826-897), a[j * (N/2+1) + n]: per mode n,
// a_j = 1 / (-2 - 4 pi^2 n^2 / N^2 - a_{j-1}) for 0 < j < gmj; a metal wall has a_0 = 0
// (a_gmj = 0); a dielectric wall below a_0 = 1 / prim_lower (0 for mode 0 when both walls are
// dielectric), one above a_gmj = 1 / (prim_upper - a_{gmj-1}), with
// prim = -1 - 2 pi (pi n + eps) n / N^2 for the wall's permittivity eps.
__global__ void k_fftx_tables(double2 *tw, double *a_eq, int N, int gmj, int diel_below, int diel_above,
                              double eps_below, double eps_above) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nm = N / 2 + 1;
    if (t < N / 2) {
        const double ang = 2.0 * kPi * (double)t / (double)N;
        tw[t] = make_double2(cos(ang), -sin(ang));
    }
    if (t < nm) {
        const int n = t;
        const double main_diag = -2.0 - 4.0 * kPi * kPi * (double)n * (double)n / ((double)N * (double)N);
        const double prim_upper =
            -1.0 - 2.0 * kPi * (kPi * (double)n + eps_above) * (double)n / ((double)N * (double)N);
        const double prim_lower =
            -1.0 - 2.0 * kPi * (kPi * (double)n + eps_below) * (double)n / ((double)N * (double)N);
        double a = 0.0;
        if (diel_below && !(diel_above && n == 0)) a = 1.0 / prim_lower;
        a_eq[n] = a;
        for (int j = 1; j <= gmj - 1; ++j) {
            a = 1.0 / (main_diag - a);
            a_eq[(int64_t)j * nm + n] = a;
        }
        a_eq[(int64_t)gmj * nm + n] = diel_above ? 1.0 / (prim_upper - a) : 0.0;
    }
}

void launch_fftx_tables(double2 *tw, double *a_eq, int N, int gmj, hipStream_t s, int diel_below, int diel_above,
                        double eps_below, double eps_above) {
    const int nm = N / 2 + 1;
    k_fftx_tables<<<(nm + kBlock - 1) / kBlock, kBlock, 0, s>>>(tw, a_eq, N, gmj, diel_below, diel_above, eps_below,
                                                                 eps_above);
}

static int ilog2(int N) {
    int l = 0;
    while ((1 << l) < N) ++l;
    return l;
}

void launch_poisson_forward(const double *rho_e, const double *rho_i, double factor_rho, const double *wall_below,
                            const double *wall_above, const double2 *twiddle, double2 *spec, const Geo &g, int N,
                            hipStream_t s, const double *sigma_below, const double *sigma_above) {
    k_poisson_forward<<<g.gmj + 1, kBlock, sizeof(double2) * N, s>>>(rho_e, rho_i, factor_rho, wall_below, wall_above,
                                                                     sigma_below, sigma_above, twiddle, spec, g.nx,
                                                                     g.gmj, N, ilog2(N));
}

void launch_poisson_thomas(double2 *spec, const double *a_eq, int N, int gmj, hipStream_t s, int two_dielectric) {
    const int nm = N / 2 + 1;
    k_poisson_thomas<<<(nm + kThModes - 1) / kThModes, kThModes * kThChunks, 0, s>>>(spec, a_eq, nm, gmj,
                                                                                     two_dielectric);
}

void launch_poisson_inverse(const double2 *spec, const double2 *twiddle, double *phi, const Geo &g, int N,
                            hipStream_t s) {
    k_poisson_inverse<<<g.gmj + 1, kBlock, sizeof(double2) * N, s>>>(spec, twiddle, phi, g.nx, N, ilog2(N));
}

// ---------------------------------------------------------------------------------
// EX, EY by central differences, one-sided at walls (pic2d_ElectricFieldCalc_FFT_X.f90:1350-1379,
// pic2d_ElectricFieldCalc_PETSc.F90:381-460); x-periodic edge columns take the values of their
// periodic images; acc_E += E (:1536-1545).
__global__ void k_efield(const double *phi, double *EX, double *EY, double *accEX, double *accEY, double f1, double f2,
                         int nx, int ny, int periodic) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nx * ny) return;
    const int jn = (int)(q / nx), in = (int)(q % nx);
    const int imax = nx - 1, jmax = ny - 1;
    int i = in;
    if (periodic) {
        if (in == 0) i = imax - 1;
        if (in == imax) i = 1;
    }
    const int j = jn;
#define P(a, b) phi[(int64_t)(b) * nx + (a)]
    double ex, ey;
    if (i == 0) ex = f1 * (P(i, j) - P(i + 1, j));
    else if (i == imax) ex = f1 * (P(i - 1, j) - P(i, j));
    else ex = f2 * (P(i - 1, j) - P(i + 1, j));
    if (j == 0) ey = f1 * (P(i, j) - P(i, j + 1));
    else if (j == jmax) ey = f1 * (P(i, j - 1) - P(i, j));
    else ey = f2 * (P(i, j - 1) - P(i, j + 1));
#undef P
    EX[q] = ex;
    EY[q] = ey;
    if (accEX) {
        accEX[q] = accEX[q] + ex;
        accEY[q] = accEY[q] + ey;
    }
}

void launch_efield(const double *phi, double *EX, double *EY, double *accEX, double *accEY, double f1, double f2,
                   const Geo &g, hipStream_t s) {
    int64_t n = (int64_t)g.nx * g.ny;
    k_efield<<<(int)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(phi, EX, EY, accEX, accEY, f1, f2, g.nx, g.ny,
                                                                 g.periodic_x);
}

}  // namespace edgpu
