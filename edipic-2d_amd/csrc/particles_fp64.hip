// FP64 particle kernels: fused push + wall routing + deposit (+ e-n MCC / RCX),
// the stand-alone deposit, collision and compaction kernels.
//
// Compiled with -ffp-contract=off: every FP64 expression rounds in the order it
// is written, which is the order of the reference Fortran, so positions,
// velocities and cell indices agree bit for bit with a CPU evaluation of the
// same expressions.
#include "collide.h"
#include "sortwin.h"

#include <algorithm>

namespace edgpu {

// Weights of the bilinear count deposit with the charge-number cases of
// GATHER_ION_CHARGE_DENSITY (pic2d_IonDynamics.f90:1653-1686); Qs=1 is
// GATHER_ELECTRON_CHARGE_DENSITY (pic2d_ElectronDynamics.f90:1631-1636).
__device__ inline void qs_weights(double vij, double vip1j, double vijp1, int Qs, double w[4]) {
    if (Qs == 1) {
        w[0] = vij;
        w[1] = vip1j;
        w[2] = vijp1;
        w[3] = 1.0 - vij - vip1j - vijp1;
    } else if (Qs == 2) {
        w[0] = vij + vij;
        w[1] = vip1j + vip1j;
        w[2] = vijp1 + vijp1;
        w[3] = 2.0 - vij - vip1j - vijp1 - vij - vip1j - vijp1;
    } else if (Qs == 3) {
        w[0] = vij + vij + vij;
        w[1] = vip1j + vip1j + vip1j;
        w[2] = vijp1 + vijp1 + vijp1;
        w[3] = 3.0 - vij - vip1j - vijp1 - vij - vip1j - vijp1 - vij - vip1j - vijp1;
    } else {  // Qs == -1
        w[0] = -vij;
        w[1] = -vip1j;
        w[2] = -vijp1;
        w[3] = -1.0 + vij + vip1j + vijp1;
    }
}

__device__ inline void deposit_fp64(double *rho, int nx, int64_t p, double vij, double vip1j, double vijp1, int Qs) {
    double w[4];
    qs_weights(vij, vip1j, vijp1, Qs, w);
    unsafeAtomicAdd(&rho[p], w[0]);
    unsafeAtomicAdd(&rho[p + 1], w[1]);
    unsafeAtomicAdd(&rho[p + nx], w[2]);
    unsafeAtomicAdd(&rho[p + nx + 1], w[3]);
}

// ---------------------------------------------------------------------------------
// Push of one particle: gather (from the LDS window when the cell is inside it), Boris
// push, move, inner-object and wall routing, collision candidacy / RCX, deposit (into
// the LDS window when possible).  COLL: 0 none, 1 e-n candidacy (electrons), 2 RCX (ions).
struct StateF64 {
    double x, y, vx, vy, vz;
};

// VZM: kVzReadWrite, kVzRead or kVzSkip (kernels.h)
template <int VZM>
__device__ inline StateF64 load_state(const ParticlesF64 &p, int64_t k) {
    return StateF64{p.x[k], p.y[k], p.vx[k], p.vy[k], VZM == kVzSkip ? 0.0 : p.vz[k]};
}

template <int VZM>
__device__ inline void store_state(const ParticlesF64 &p, int64_t k, const StateF64 &o) {
    p.x[k] = o.x;
    p.y[k] = o.y;
    p.vx[k] = o.vx;
    p.vy[k] = o.vy;
    if (VZM == kVzReadWrite) p.vz[k] = o.vz;
}

// Field windows of a tile in LDS, read through LDS pointers (tile.h): the gather then waits
// on the LDS counter only, so neither the stores of the previous particle nor the loads of
// the next one are waited for inside the loop.
struct FieldWindow {
    lds<const double> *ex, *ey;
};

// Returns the state to store in slot k: every path stores the five fields, so the number of
// vector-memory operations issued after the next particle's loads is the same on all paths
// and the wait for those loads at the end of the iteration leaves the stores in flight.
// Error paths store the unchanged state; a removed particle stores the NaN removal mark in x.
// COPIES: the number of copies of the caller's LDS deposit window (window_clear/window_copy).
template <bool ELECTRON, int COLL, int VZM, bool MOM, int COPIES = kCopiesF64>
__device__ inline StateF64 advance_one_fp64(const AdvanceArgs &a, int64_t k, const StateF64 &st, double vz2b,
                                            const Window *win,
                                            FieldWindow fw, double *sR, BlockEvents *ev,
                                            const double *sBound, float *sM) {
    const Geo &g = a.g;
    double x = st.x, y = st.y;
    if (x < 0.0 || x > g.xmax || y < 0.0 || y > g.ymax) {
        set_error(a.cnt, EDGPU_ERR_OUT_OF_BOUNDS, k);
        return st;
    }
    int i, j;
    if (!cell_of(g, x, y, i, j)) {
        set_error(a.cnt, EDGPU_ERR_INDEX, k);
        return st;
    }
    const double ax_ip1 = x - (double)i, ax_i = 1.0 - ax_ip1;
    const double ay_jp1 = y - (double)j, ay_j = 1.0 - ay_jp1;
    double ex00, ex10, ex01, ex11, ey00, ey10, ey01, ey11;
    const int s = win ? window_slot(*win, i, j) : -1;
    if (s >= 0) {
        ex00 = fw.ex[s];
        ex10 = fw.ex[s + 1];
        ex01 = fw.ex[s + kWinW];
        ex11 = fw.ex[s + kWinW + 1];
        ey00 = fw.ey[s];
        ey10 = fw.ey[s + 1];
        ey01 = fw.ey[s + kWinW];
        ey11 = fw.ey[s + kWinW + 1];
    } else {
        const int64_t p = (int64_t)i + (int64_t)j * g.nx;
        ex00 = a.EX[p];
        ex10 = a.EX[p + 1];
        ex01 = a.EX[p + g.nx];
        ex11 = a.EX[p + g.nx + 1];
        ey00 = a.EY[p];
        ey10 = a.EY[p + 1];
        ey01 = a.EY[p + g.nx];
        ey11 = a.EY[p + g.nx + 1];
        // wait here, on the rare path, so that the registers these loads fill carry no
        // pending load past the join: the common LDS path then waits on the LDS counter
        // alone and the next particle's loads stay in flight
        wait_vector_memory();
    }
    // pic2d_ElectronDynamics.f90:79-86
    const double E_X = ex00 * ax_i * ay_j + ex10 * ax_ip1 * ay_j + ex01 * ax_i * ay_jp1 + ex11 * ax_ip1 * ay_jp1;
    const double E_Y = ey00 * ax_i * ay_j + ey10 * ax_ip1 * ay_j + ey01 * ax_i * ay_jp1 + ey11 * ax_ip1 * ay_jp1;
    double vx = st.vx, vy = st.vy, vz = st.vz;
    const double *K = a.K;
    if (ELECTRON) {
        // pic2d_ElectronDynamics.f90:116-149
        const double vxm = vx - 0.5 * E_X, vym = vy - 0.5 * E_Y, vzm = vz - 0.5 * a.EZ;
        const double vxp = K[0] * vxm + K[1] * vym + K[2] * vzm;
        const double vyp = K[3] * vxm + K[4] * vym + K[5] * vzm;
        const double vzp = K[6] * vxm + K[7] * vym + K[8] * vzm;
        vx = vxp - 0.5 * E_X;
        vy = vyp - 0.5 * E_Y;
        vz = vzp - 0.5 * a.EZ;
        x = x + vx;
        y = y + vy;
    } else if (a.sense_B) {
        // pic2d_IonDynamics.f90:128-155
        const double q = a.QM2s;
        const double vxm = vx + q * E_X, vym = vy + q * E_Y, vzm = vz + q * a.EZ;
        const double vxp = K[0] * vxm + K[1] * vym + K[2] * vzm;
        const double vyp = K[3] * vxm + K[4] * vym + K[5] * vzm;
        const double vzp = K[6] * vxm + K[7] * vym + K[8] * vzm;
        vx = vxp + q * E_X;
        vy = vyp + q * E_Y;
        vz = vzp + q * a.EZ;
        x = x + vx * a.Nsub;
        y = y + vy * a.Nsub;
    } else {
        // pic2d_IonDynamics.f90:144-155
        const double q2 = a.QM2s + a.QM2s;
        vx = vx + q2 * E_X;
        vy = vy + q2 * E_Y;
        vz = vz + q2 * a.EZ;
        x = x + vx * a.Nsub;
        y = y + vy * a.Nsub;
    }
    if (MOM) {
        // at the position before the step, with the half-step velocity (the fused moments'
        // order: pic2d_ElectronDynamicsWithVDFMoments.f90:330-460)
        const MomentTerms t = moment_terms((float)(ax_i * ay_j), (float)(ax_ip1 * ay_j), (float)(ax_i * ay_jp1),
                                           0.5 * (vx + st.vx), 0.5 * (vy + st.vy), 0.5 * (vz + st.vz));
        moments_add(sM, s, (int64_t)i + (int64_t)j * g.nx, g, a.mom, t);
    }
    // inner metal objects: open-box test (pic2d_ElectronDynamics.f90:161-174); the object whose
    // face the step's trajectory crossed first is found after the push (k_inner_hits_fp64)
    int obj = inside_inner(g, x, y) ? kInnerPending : -1, side = -1;
    double rx = x, ry = y;
    if (obj < 0 && !(x >= 0.0 && x <= g.xmax && y >= 0.0 && y <= g.ymax)) {
        side = exit_side(g, x, y);
        rx = x;
        if (side >= 0) {
            obj = wall_object(g, side, x, y);
            if (obj < 0) {
                set_error(a.cnt, EDGPU_ERR_NO_WALL, k);
                return st;
            }
        }
    }
    if (obj >= 0) {
        event_removed(ev, a, k, obj, side, rx, ry);
        return StateF64{__longlong_as_double(0x7ff8000000000000ll), y, vx, vy, vz};  // removal mark (NaN)
    }
    if (COLL == 1) en_candidate(a, ev, k, vx, vy, vz, sBound);
    if (COLL == 2) {
        // kVzSkip: vz2b bounds vz^2 of every ion, so an ion this test rejects is rejected by
        // rcx_collide's own test as well; the others read vz and take the exact decision
        if (VZM != kVzSkip || !rcx_rejects_any_vz(a.rcx, a.rng, k, vx * vx + vy * vy, vz2b)) {
            if (VZM == kVzSkip) {
                vz = a.p.vz[k];
                wait_vector_memory();  // on this rare path, as in the field gather
            }
            if (rcx_collide(a.rcx, a.maxw, a.rng, k, vx, vy, vz)) {
                atomicAdd(&ev->coll[0], 1u);
                if (VZM != kVzReadWrite) a.p.vz[k] = vz;  // the exchange draws a new vz
                raise_vz_bound(a.vz2max, vz);
            }
        }
    }
    const StateF64 out{x, y, vx, vy, vz};
    if (a.rho) {
        int ii, jj;
        if (!cell_of(g, x, y, ii, jj)) {
            set_error(a.cnt, EDGPU_ERR_INDEX, k);
            return out;
        }
        const double bx_ip1 = x - (double)ii, bx_i = 1.0 - bx_ip1;
        const double by_jp1 = y - (double)jj, by_j = 1.0 - by_jp1;
        const int d = win ? window_slot(*win, ii, jj) : -1;
        if (d >= 0) {
            double w[4];
            qs_weights(bx_i * by_j, bx_ip1 * by_j, bx_i * by_jp1, a.Qs, w);
            window_deposit<double, COPIES>(sR, d, w[0], w[1], w[2], w[3]);
        } else {
            deposit_fp64(a.rho, g.nx, (int64_t)ii + (int64_t)jj * g.nx, bx_i * by_j, bx_ip1 * by_j, bx_i * by_jp1,
                         a.Qs);
        }
    }
    return out;
}

// Particles k, k+stride, ... < end of one thread.  The state of the next particle is loaded
// before the current one is pushed, so its loads are in flight during the push.  Vector
// memory operations complete in order on gfx950 and the compiler waits for a load by
// counting the operations issued after it on the path with the fewest: the prefetch is
// therefore unconditional (the last iteration reloads its own slot) and every path stores
// the same fields, so the wait at the end of an iteration is "all but those stores".
template <bool ELECTRON, int COLL, int VZM, bool MOM>
__device__ inline void run_range_fp64(const AdvanceArgs &a, int64_t k, int64_t end, int64_t stride, const Window *win,
                                      FieldWindow fw, double *sR, BlockEvents *ev,
                                      const double *sBound, float *sM) {
    if (k >= end) return;
    // the bound only grows during the launch (exchanges raise it), so a copy read before the
    // loop still bounds every ion not yet pushed
    const double vz2b = (COLL == 2 && VZM == kVzSkip) ? vz_bound(a.vz2max) : 0.0;
    StateF64 cur = load_state<VZM>(a.p, k);
    wait_vector_memory();
    for (;;) {
        const int64_t kn = k + stride;
        const bool more = kn < end;
        const StateF64 nxt = load_state<VZM>(a.p, more ? kn : k);
        store_state<VZM>(a.p, k,
                         advance_one_fp64<ELECTRON, COLL, VZM, MOM>(a, k, cur, vz2b, win, fw, sR, ev, sBound, sM));
        if (!more) break;
        cur = nxt;
        k = kn;
    }
}

// Workgroups [0, ntiles) own one tile each; the others stride over the untiled particles.
// MOM: also accumulate the moments (a.mom) of a snapshot step.
template <bool ELECTRON, int COLL, int VZM, bool MOM>
__global__ __launch_bounds__(kAdvBlock) void k_advance_fp64(AdvanceArgs a) {
    __shared__ double sR[kWinN * kCopiesF64];
    __shared__ double sEX[kWinN], sEY[kWinN];
    __shared__ BlockEvents ev;
    __shared__ double sBound[COLL == 1 ? kEnBins : 1];
    __shared__ float sM[MOM ? kMoments * kWinN : 1];
    events_init(&ev);
    if (COLL == 1) stage_bound(a, sBound);
    const int b = blockIdx.x;
    if (b < a.tiles.ntiles) {
        const int t = tile_of_block(b, a.tiles.ntiles);
        const Window w = tile_window(t, a.tiles.ntx, a.g);
        window_clear<double, kCopiesF64>(sR);
        if (MOM) moments_clear(sM);
        window_load_field(w, a.g, a.EX, a.EY, sEX, sEY);
        __syncthreads();
        const int64_t s0 = a.tiles.start[t];
        const int64_t s1 = min((int64_t)a.tiles.start[t + 1], a.n);
        const FieldWindow fw{to_lds<const double>(sEX), to_lds<const double>(sEY)};
        run_range_fp64<ELECTRON, COLL, VZM, MOM>(a, s0 + threadIdx.x, s1, blockDim.x, &w, fw,
                                            window_copy<double, kCopiesF64>(sR), &ev, sBound, sM);
        __syncthreads();
        if (a.rho) window_flush<double, kCopiesF64>(w, a.g, sR, a.rho);
        if (MOM) moments_flush(w, a.g, sM, a.mom);
    } else {
        __syncthreads();
        const int64_t begin = min(a.tiles.covered, a.n);
        const int64_t stride = (int64_t)(gridDim.x - a.tiles.ntiles) * blockDim.x;
        run_range_fp64<ELECTRON, COLL, VZM, MOM>(a, begin + (int64_t)(b - a.tiles.ntiles) * blockDim.x + threadIdx.x,
                                            a.n, stride, nullptr, FieldWindow{nullptr, nullptr}, nullptr, &ev, sBound,
                                            nullptr);
    }
    events_flush(&ev, a);
}

template <bool ELECTRON, int COLL, int VZM>
static void launch_advance_kernel(const AdvanceArgs &a, int grid, hipStream_t s) {
    if (a.mom) k_advance_fp64<ELECTRON, COLL, VZM, true><<<grid, kAdvBlock, 0, s>>>(a);
    else k_advance_fp64<ELECTRON, COLL, VZM, false><<<grid, kAdvBlock, 0, s>>>(a);
}

// ---------------------------------------------------------------------------------
// The sorting push of the electrons: the push of a sort step writes the particles to the
// spare arrays in the tile order of their positions before the push (k_advance_sort_mixed
// describes the scheme).  Two particles per thread and chunk: the FP64 states of the current
// and the next chunk stay in registers.
constexpr int kSortPushPer = 2;
constexpr int kSortPushChunk = kSortPushPer * kAdvBlock;
// Its LDS deposit window keeps 4 copies instead of the plain push's 16: at 46 KB of LDS a CU
// held 3 of its workgroups, at 26 KB it holds the 4 its registers allow, and the sort step's
// push ran 10.4 ms instead of 12.3 ms at 2048x1024 and 200 particles per cell.
constexpr int kCopiesSortF64 = 4;

__device__ inline void store_particle(const ParticlesF64 &p, int64_t d, const StateF64 &o, int32_t tag) {
    p.x[d] = o.x;
    p.y[d] = o.y;
    p.vx[d] = o.vx;
    p.vy[d] = o.vy;
    p.vz[d] = o.vz;
    p.tag[d] = tag;
}

// the untiled particles k, k+stride, ... < end of one thread: single slots from the cursors
template <int COLL, int VZM>
__device__ inline void sort_push_untiled_fp64(const AdvanceArgs &a, const ParticlesF64 &out, int32_t *cursor,
                                              int64_t k, int64_t end, int64_t stride, BlockEvents *ev,
                                              const double *sBound) {
    const int ntx = a.tiles.ntx;
    for (; k < end; k += stride) {
        const StateF64 st = load_state<kVzReadWrite>(a.p, k);
        const int32_t tag = a.p.tag[k];
        int tx, j;
        sort_tile_f64(st.x, st.y, a.g, tx, j);
        const int64_t d = atomicAdd(&cursor[tx + j * ntx], 1);
        if (d >= a.n) {
            set_error(a.cnt, EDGPU_ERR_INDEX, k);
            continue;
        }
        store_particle(out, d,
                       advance_one_fp64<true, COLL, VZM, false>(a, d, st, 0.0, nullptr, FieldWindow{nullptr, nullptr},
                                                                nullptr, ev, sBound, nullptr),
                       tag);
    }
}

// the states of one thread's particles of a chunk: the positions first, so that the ranking
// waits for them alone
struct ChunkF64 {
    StateF64 st[kSortPushPer];
    int32_t tag[kSortPushPer];
};

__device__ inline void load_chunk_fp64(const ParticlesF64 &p, int64_t c0, int64_t c1, ChunkF64 &c) {
#pragma unroll
    for (int u = 0; u < kSortPushPer; ++u) {
        const int64_t k = min(c0 + threadIdx.x + (int64_t)u * kAdvBlock, c1 - 1);
        c.st[u].x = p.x[k];
        c.st[u].y = p.y[k];
    }
#pragma unroll
    for (int u = 0; u < kSortPushPer; ++u) {
        const int64_t k = min(c0 + threadIdx.x + (int64_t)u * kAdvBlock, c1 - 1);
        c.st[u].vx = p.vx[k];
        c.st[u].vy = p.vy[k];
        c.st[u].vz = p.vz[k];
        c.tag[u] = p.tag[k];
    }
}

template <int COLL, int VZM>
__global__ __launch_bounds__(kAdvBlock) void k_advance_sort_fp64(AdvanceArgs a, ParticlesF64 out, int32_t *cursor,
                                                                 const int32_t *wcount) {
    __shared__ double sR[kWinN * kCopiesSortF64];
    __shared__ double sEX[kWinN], sEY[kWinN];
    __shared__ BlockEvents ev;
    __shared__ double sBound[COLL == 1 ? kEnBins : 1];
    __shared__ PushStage<kSortPushChunk> ps;
    __shared__ double stage[kSortPushChunk];
    events_init(&ev);
    if (COLL == 1) stage_bound(a, sBound);
    const Geo &g = a.g;
    const int ntx = a.tiles.ntx;
    const int b = blockIdx.x;
    if (b >= a.tiles.ntiles) {
        __syncthreads();
        const int64_t begin = min(a.tiles.covered, a.n);
        const int64_t stride = (int64_t)(gridDim.x - a.tiles.ntiles) * blockDim.x;
        sort_push_untiled_fp64<COLL, VZM>(a, out, cursor, begin + (int64_t)(b - a.tiles.ntiles) * blockDim.x + threadIdx.x,
                                          a.n, stride, &ev, sBound);
        events_flush(&ev, a);
        return;
    }
    const int t = tile_of_block(b, a.tiles.ntiles);
    const Window w = tile_window(t, ntx, g);
    const SortWindow sw = sort_window(t, ntx, g);
    window_clear<double, kCopiesSortF64>(sR);
    window_load_field(w, g, a.EX, a.EY, sEX, sEY);
    double *sRc = window_copy<double, kCopiesSortF64>(sR);
    const FieldWindow fw{to_lds<const double>(sEX), to_lds<const double>(sEY)};
    const int64_t s0 = a.tiles.start[t];
    const int64_t s1 = min((int64_t)a.tiles.start[t + 1], a.n);
    ChunkF64 cur;
    if (s0 < s1) load_chunk_fp64(a.p, s0, min(s0 + (int64_t)kSortPushChunk, s1), cur);
    for (int q = threadIdx.x; q < kSortWinN; q += blockDim.x) {
        const int c = wcount[(int64_t)t * kSortWinN + q];
        ps.runs[q] = c ? atomicAdd(&cursor[(sw.x0 + q % kSortWinW) + (sw.y0 + q / kSortWinW) * ntx], c) : 0;
        ps.cnt[q] = 0;
    }
    __syncthreads();
    for (int64_t c0 = s0; c0 < s1; c0 += kSortPushChunk) {
        const int64_t c1 = min(c0 + (int64_t)kSortPushChunk, s1);
        // slot: rank in the destination tile, then the stage slot, or -1 outside the window;
        // dst: destination index, or -1 beyond the chunk; cq: window bin, or -1
        int slot[kSortPushPer], cq[kSortPushPer];
        int32_t dst[kSortPushPer];
#pragma unroll
        for (int u = 0; u < kSortPushPer; ++u) {
            slot[u] = -1;
            cq[u] = -1;
            dst[u] = -1;
            if (c0 + threadIdx.x + (int64_t)u * kAdvBlock < c1) {
                int tx, j;
                sort_tile_f64(cur.st[u].x, cur.st[u].y, g, tx, j);
                const int d = sort_bin(sw, tx, j);
                if (d >= 0) {
                    cq[u] = d;
                    slot[u] = atomicAdd(&ps.cnt[d], 1);
                } else {
                    dst[u] = atomicAdd(&cursor[tx + j * ntx], 1);
                }
            }
        }
        lds_barrier();
        scan_window_counts(ps);
        lds_barrier();
#pragma unroll
        for (int u = 0; u < kSortPushPer; ++u) {
            if (cq[u] >= 0) {
                const int s = ps.loc[cq[u]] + slot[u];
                dst[u] = ps.runs[cq[u]] + slot[u];
                ps.dest[s] = dst[u];
                slot[u] = s;
            }
        }
        const int m = ps.m_in;
        lds_barrier();
        // the runs advance past this chunk; the counts start again from zero
        for (int q = threadIdx.x; q < kSortWinN; q += blockDim.x) {
            ps.runs[q] += ps.cnt[q];
            ps.cnt[q] = 0;
        }
        // the next chunk's loads go out before this chunk is pushed and placed
        ChunkF64 nxt;
        const int64_t n0 = c0 + kSortPushChunk;
        if (n0 < s1) load_chunk_fp64(a.p, n0, min(n0 + (int64_t)kSortPushChunk, s1), nxt);
#pragma unroll
        for (int u = 0; u < kSortPushPer; ++u) {
            if (dst[u] >= a.n) {
                set_error(a.cnt, EDGPU_ERR_INDEX, c0 + threadIdx.x + (int64_t)u * kAdvBlock);
            } else if (dst[u] >= 0) {
                cur.st[u] = advance_one_fp64<true, COLL, VZM, false, kCopiesSortF64>(a, dst[u], cur.st[u], 0.0, &w,
                                                                     fw, sRc, &ev, sBound, nullptr);
            }
        }
        double v[kSortPushPer];
#define EDGPU_PLACE(FIELD, OUT)                                         \
    for (int u = 0; u < kSortPushPer; ++u) v[u] = cur.st[u].FIELD;      \
    place_array<kSortPushPer, double>(v, OUT, slot, dst, stage, ps.dest, m, a.n);
        EDGPU_PLACE(x, out.x)
        EDGPU_PLACE(y, out.y)
        EDGPU_PLACE(vx, out.vx)
        EDGPU_PLACE(vy, out.vy)
        EDGPU_PLACE(vz, out.vz)
#undef EDGPU_PLACE
        place_array<kSortPushPer, int32_t>(cur.tag, out.tag, slot, dst, (int32_t *)stage, ps.dest, m, a.n);
        cur = nxt;
    }
    __syncthreads();
    if (a.rho) window_flush<double, kCopiesSortF64>(w, g, sR, a.rho);
    events_flush(&ev, a);
}

// Process selection of the candidate slots listed by the electron push (before compaction):
// one thread per candidate, no kinematics, so the kernel stays small and keeps many waves
// in flight over the random reads; slots that picked a process join the selection list.
constexpr int kSelBuf = 256;

__global__ __launch_bounds__(kBlock) void k_select_fp64(AdvanceArgs a) {
    __shared__ BlockList<kSelBuf> sl;
    block_list_init(&sl);
    const int64_t nc = (int64_t)min(a.cnt->n_cand, (unsigned long long)a.cand_cap);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nc; t += stride) {
        const int64_t k = a.cand[t];
        double u0[4];
        draw_en(a.rng, k, u0);
        double en;
        const int coll = en_select(*a.en, u0, a.p.vx[k], a.p.vy[k], a.p.vz[k], k, a.cnt, en);
        if (coll >= 1 && coll <= a.en->n_proc) block_list_push(&sl, k, &a.cnt->n_sel, a.sel, a.cand_cap, a.cnt);
    }
    block_list_flush(&sl, &a.cnt->n_sel, a.sel, a.cand_cap, a.cnt);
}

// kinematics of the selected slots
__global__ __launch_bounds__(kBlock) void k_candidates_fp64(AdvanceArgs a) {
    __shared__ unsigned int sc[EDGPU_MAX_PROC];
    coll_count_init(sc);
    const int64_t ns = (int64_t)min(a.cnt->n_sel, (unsigned long long)a.cand_cap);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ns; t += stride) {
        const int64_t k = a.sel[t];
        double u0[4];
        draw_en(a.rng, k, u0);
        double vx = a.p.vx[k], vy = a.p.vy[k], vz = a.p.vz[k];
        if (en_collide_candidate(a.en, u0, a.maxw, a.rng, k, a.p.x[k], a.p.y[k], vx, vy, vz, a.add_e, a.add_i,
                                 a.cnt, sc)) {
            a.p.vx[k] = vx;
            a.p.vy[k] = vy;
            a.p.vz[k] = vz;
        }
    }
    coll_count_flush(sc, a.cnt);
}

// stand-alone deposit of particles [begin, n)
__global__ __launch_bounds__(kBlock) void k_deposit_fp64(ParticlesF64 p, int64_t begin, int64_t n, Geo g, double *rho,
                                                          int Qs, Counters *cnt) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        double x = p.x[k], y = p.y[k];
        int i, j;
        if (!cell_of(g, x, y, i, j)) {
            set_error(cnt, EDGPU_ERR_INDEX, k);
            continue;
        }
        const double ax_ip1 = x - (double)i, ax_i = 1.0 - ax_ip1;
        const double ay_jp1 = y - (double)j, ay_j = 1.0 - ay_jp1;
        deposit_fp64(rho, g.nx, (int64_t)i + (int64_t)j * g.nx, ax_i * ay_j, ax_ip1 * ay_j, ax_i * ay_jp1, Qs);
    }
}

// stand-alone e-n MCC over electrons [begin, n)
__global__ __launch_bounds__(kBlock) void k_mcc_en_fp64(ParticlesF64 p, int64_t begin, int64_t n, const EnTable *t, double pmax,
                                                         const double *maxw, Rng rng, AddList ae, IonAddLists ai,
                                                         Counters *cnt) {
    __shared__ unsigned int sc[EDGPU_MAX_PROC];
    coll_count_init(sc);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        double vx = p.vx[k], vy = p.vy[k], vz = p.vz[k];
        if (en_collide(t, pmax, maxw, rng, k, p.x[k], p.y[k], vx, vy, vz, ae, ai, cnt, sc)) {
            p.vx[k] = vx;
            p.vy[k] = vy;
            p.vz[k] = vz;
        }
    }
    coll_count_flush(sc, cnt);
}

__global__ __launch_bounds__(kBlock) void k_rcx_fp64(ParticlesF64 p, int64_t begin, int64_t n, RcxPar r,
                                                      const double *maxw, Rng rng, Counters *cnt,
                                                      unsigned long long *vz2max) {
    __shared__ unsigned int sc[EDGPU_MAX_PROC];
    coll_count_init(sc);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        double vx = p.vx[k], vy = p.vy[k], vz = p.vz[k];
        if (rcx_collide(r, maxw, rng, k, vx, vy, vz)) {
            p.vx[k] = vx;
            p.vy[k] = vy;
            p.vz[k] = vz;
            atomicAdd(&sc[0], 1u);
            raise_vz_bound(vz2max, vz);
        }
    }
    coll_count_flush(sc, cnt);
}

// ---------------------------------------------------------------------------------
// compaction after removals: removed slots below new_n are holes, surviving slots at or
// above new_n are movers; each mover is copied into one hole.
__global__ void k_find_holes(const int64_t *removed, int64_t nrem, int64_t new_n, int64_t *holes, Counters *cnt) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrem) return;
    int64_t idx = removed[t];
    if (idx < new_n) holes[atomicAdd(&cnt->n_holes, 1ull)] = idx;
}

__global__ void k_find_movers(const double *x, int64_t new_n, int64_t old_n, int64_t *movers, Counters *cnt) {
    int64_t t = new_n + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= old_n) return;
    double v = x[t];
    if (v == v) movers[atomicAdd(&cnt->n_movers, 1ull)] = t;
}

__global__ void k_move_fp64(ParticlesF64 p, const int64_t *holes, const int64_t *movers, const Counters *cnt) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t count = (int64_t)min(cnt->n_holes, cnt->n_movers);
    if (t >= count) return;
    int64_t d = holes[t], s = movers[t];
    p.x[d] = p.x[s];
    p.y[d] = p.y[s];
    p.vx[d] = p.vx[s];
    p.vy[d] = p.vy[s];
    p.vz[d] = p.vz[s];
    p.tag[d] = p.tag[s];
}

// The initial loading's removal of the particles placed inside inner objects
// (pic2d_CurProblemValues.f90:3245-3258): mark them as a push marks a removal (x = NaN) and
// count them; the marked slots below the new count are then filled from the unmarked slots
// above it, as after a push.
__global__ void k_mark_inside_fp64(ParticlesF64 p, int64_t n, Geo g, Counters *cnt) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    unsigned long long c = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        if (inside_inner(g, p.x[k], p.y[k])) {
            p.x[k] = __longlong_as_double(0x7ff8000000000000ll);
            ++c;
        }
    }
    if (c) atomicAdd(&cnt->n_removed, c);
}

__global__ void k_marked_holes_fp64(const double *x, int64_t new_n, int64_t *holes, Counters *cnt) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < new_n; k += stride) {
        const double v = x[k];
        if (v != v) holes[atomicAdd(&cnt->n_holes, 1ull)] = k;
    }
}

// The particles a push left inside an inner box (records with object kInnerPending, at their
// position after the step): the step's trajectory from (x - vx back, y - vy back) to (x, y),
// back = 1 for electrons and N_subcycles for ions, against the boxes' faces
// (TRY_ELECTRON_COLL_WITH_INNER_OBJECT, pic2d_ElectronWallCollisions.f90:470-557;
// TRY_ION_COLL_WITH_INNER_OBJECT, pic2d_IonWallCollisions.f90:391-480).  The record gets
// the object and face crossed first and the crossing, the object its hit.
__global__ void k_inner_hits_fp64(AdvanceArgs a, double back) {
    const int64_t n = min((int64_t)a.cnt->n_removed, a.removed_cap);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        if ((a.removed_obj[q] & 0xff) != kInnerPending) continue;
        const int64_t k = a.removed_idx[q];
        const double x = a.removed_x[q], y = a.removed_y[q];
        const double vx = a.p.vx[k], vy = a.p.vy[k];
        int obj = -1, side = -1;
        double rx = x, ry = y;
        if (!inner_crossing(a.g, x - vx * back, y - vy * back, x, y, obj, side, rx, ry)) {
            set_error(a.cnt, EDGPU_ERR_NO_WALL, k);
            continue;
        }
        a.removed_obj[q] = obj | ((side + 1) << 8);
        a.removed_x[q] = rx;
        a.removed_y[q] = ry;
        atomicAdd(&a.cnt->hit[obj], 1ull);
    }
}

// append the add list to the particle array
__global__ void k_append_fp64(ParticlesF64 p, int64_t n0, AddList a, int64_t count) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    p.x[n0 + t] = a.x[t];
    p.y[n0 + t] = a.y[t];
    p.vx[n0 + t] = a.vx[t];
    p.vy[n0 + t] = a.vy[t];
    p.vz[n0 + t] = a.vz[t];
    p.tag[n0 + t] = a.tag[t];
}

// DISTRIBUTE_PARTICLES with the counter-based RNG (draw order per particle: x, y, vx, vy, vz)
__global__ __launch_bounds__(kBlock) void k_distribute_fp64(ParticlesF64 p, int64_t n0, int64_t n, double x0, double x1,
                                                             double y0, double y1, double xmax, double ymax, double fc,
                                                             const double *maxw, Rng rng, const double *ex,
                                                             const double *ey) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        double u[8];
        draw4(rng, k, 0, u);
        draw4(rng, k, 1, u + 4);
        double x, y, vz;
        if (ex) {
            x = ex[k];
            y = ey[k];
            vz = 0.0;
        } else {
            x = fmax(0.0, fmin(xmax, x0 + (x1 - x0) * u[0]));
            y = fmax(0.0, fmin(ymax, y0 + (y1 - y0) * u[1]));
            vz = maxwell_sample(maxw, u[4]) * fc;
        }
        p.x[n0 + k] = x;
        p.y[n0 + k] = y;
        p.vx[n0 + k] = maxwell_sample(maxw, u[2]) * fc;
        p.vy[n0 + k] = maxwell_sample(maxw, u[3]) * fc;
        p.vz[n0 + k] = vz;
        p.tag[n0 + k] = 0;
    }
}

// ---------------------------------------------------------------------------------
// host launchers
static int grid_for(int64_t n, int cap = 65536) {
    int64_t b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

int advance_grid(const AdvanceArgs &a) {
    const int64_t untiled = a.n - std::min<int64_t>(a.tiles.covered, a.n);
    int tail = (int)std::min<int64_t>((untiled + kAdvBlock - 1) / kAdvBlock, 65536);
    if (a.tiles.ntiles == 0 && tail < 1) tail = 1;
    return a.tiles.ntiles + tail;
}

void launch_advance_sort_fp64(const AdvanceArgs &a, const ParticlesF64 &out, const SortWork &w, int coll,
                              int vz_mode, hipStream_t s) {
    if (a.n <= 0) return;
    const int grid = advance_grid(a);
    if (vz_mode == kVzReadWrite) {
        if (coll == 1) k_advance_sort_fp64<1, kVzReadWrite><<<grid, kAdvBlock, 0, s>>>(a, out, w.count, w.wcount);
        else k_advance_sort_fp64<0, kVzReadWrite><<<grid, kAdvBlock, 0, s>>>(a, out, w.count, w.wcount);
    } else {
        if (coll == 1) k_advance_sort_fp64<1, kVzRead><<<grid, kAdvBlock, 0, s>>>(a, out, w.count, w.wcount);
        else k_advance_sort_fp64<0, kVzRead><<<grid, kAdvBlock, 0, s>>>(a, out, w.count, w.wcount);
    }
}

void launch_advance_fp64(const AdvanceArgs &a, bool electron, int coll, int vz_mode, hipStream_t s) {
    if (a.n <= 0) return;
    const int grid = advance_grid(a);
    if (electron) {
        // electrons read vz for the collision test (kVzSkip is an ion mode)
        if (vz_mode == kVzReadWrite) {
            if (coll == 1) launch_advance_kernel<true, 1, kVzReadWrite>(a, grid, s);
            else launch_advance_kernel<true, 0, kVzReadWrite>(a, grid, s);
        } else {
            if (coll == 1) launch_advance_kernel<true, 1, kVzRead>(a, grid, s);
            else launch_advance_kernel<true, 0, kVzRead>(a, grid, s);
        }
    } else if (vz_mode == kVzReadWrite) {
        if (coll == 2) launch_advance_kernel<false, 2, kVzReadWrite>(a, grid, s);
        else launch_advance_kernel<false, 0, kVzReadWrite>(a, grid, s);
    } else if (vz_mode == kVzRead) {
        if (coll == 2) launch_advance_kernel<false, 2, kVzRead>(a, grid, s);
        else launch_advance_kernel<false, 0, kVzRead>(a, grid, s);
    } else {
        if (coll == 2) launch_advance_kernel<false, 2, kVzSkip>(a, grid, s);
        else launch_advance_kernel<false, 0, kVzSkip>(a, grid, s);
    }
}

void launch_vz_bound_fp64(const ParticlesF64 &p, int64_t begin, int64_t end, unsigned long long *bound,
                          hipStream_t s) {
    if (end <= begin) return;
    k_vz_bound<double><<<grid_for(end - begin, 8192), kBlock, 0, s>>>(p.vz, begin, end, bound);
}

// The list lengths are only known on the device: the selection grid covers the candidate
// capacity (one thread per entry, threads past the length exit at once); the selected slots
// are few and the kinematics kernel strides over them.
void launch_candidates_fp64(const AdvanceArgs &a, hipStream_t s) {
    const int64_t blocks = std::min<int64_t>(std::max<int64_t>((a.cand_cap + kBlock - 1) / kBlock, 1), 65536);
    k_select_fp64<<<(int)blocks, kBlock, 0, s>>>(a);
    k_candidates_fp64<<<(int)std::min<int64_t>(blocks, 4096), kBlock, 0, s>>>(a);
}

void launch_deposit_fp64(const ParticlesF64 &p, int64_t begin, int64_t n, const Geo &g, double *rho, int Qs,
                         Counters *cnt, hipStream_t s) {
    if (n <= begin) return;
    k_deposit_fp64<<<grid_for(n - begin, 1 << 20), kBlock, 0, s>>>(p, begin, n, g, rho, Qs, cnt);
}

void launch_mcc_en_fp64(const ParticlesF64 &p, int64_t begin, int64_t n, const EnTable *t, double pmax, const double *maxw,
                        const Rng &rng, const AddList &ae, const IonAddLists &ai, Counters *cnt, hipStream_t s) {
    if (n <= begin) return;
    k_mcc_en_fp64<<<grid_for(n - begin, 1 << 20), kBlock, 0, s>>>(p, begin, n, t, pmax, maxw, rng, ae, ai, cnt);
}

void launch_rcx_fp64(const ParticlesF64 &p, int64_t begin, int64_t n, const RcxPar &r, const double *maxw,
                     const Rng &rng, Counters *cnt, unsigned long long *vz2max, hipStream_t s) {
    if (n <= begin) return;
    k_rcx_fp64<<<grid_for(n - begin, 1 << 20), kBlock, 0, s>>>(p, begin, n, r, maxw, rng, cnt, vz2max);
}

void launch_compact_fp64(const ParticlesF64 &p, const int64_t *removed, int64_t nrem, int64_t old_n, int64_t *holes,
                         int64_t *movers, Counters *cnt, hipStream_t s) {
    if (nrem <= 0) return;
    int64_t new_n = old_n - nrem;
    int g = (int)((nrem + kBlock - 1) / kBlock);
    k_find_holes<<<g, kBlock, 0, s>>>(removed, nrem, new_n, holes, cnt);
    k_find_movers<<<g, kBlock, 0, s>>>(p.x, new_n, old_n, movers, cnt);
    // the number of holes equals the number of movers; the move kernel reads the pair
    // count the two kernels above accumulated
    k_move_fp64<<<g, kBlock, 0, s>>>(p, holes, movers, cnt);
}

void launch_mark_inside_fp64(const ParticlesF64 &p, int64_t n, const Geo &g, Counters *cnt, hipStream_t s) {
    if (n <= 0 || g.n_inner == 0) return;
    k_mark_inside_fp64<<<grid_for(n, 8192), kBlock, 0, s>>>(p, n, g, cnt);
}

void launch_fill_marked_fp64(const ParticlesF64 &p, int64_t new_n, int64_t old_n, int64_t *holes, int64_t *movers,
                             Counters *cnt, hipStream_t s) {
    const int64_t m = old_n - new_n;
    if (m <= 0) return;
    const int g = (int)((m + kBlock - 1) / kBlock);
    k_marked_holes_fp64<<<grid_for(new_n, 8192), kBlock, 0, s>>>(p.x, new_n, holes, cnt);
    k_find_movers<<<g, kBlock, 0, s>>>(p.x, new_n, old_n, movers, cnt);
    k_move_fp64<<<g, kBlock, 0, s>>>(p, holes, movers, cnt);
}

void launch_inner_hits_fp64(const AdvanceArgs &a, bool electron, hipStream_t s) {
    if (a.g.n_inner == 0 || a.n <= 0) return;
    k_inner_hits_fp64<<<64, kBlock, 0, s>>>(a, electron ? 1.0 : a.Nsub);
}

void launch_append_fp64(const ParticlesF64 &p, int64_t n0, const AddList &a, int64_t count, hipStream_t s) {
    if (count <= 0) return;
    k_append_fp64<<<(int)((count + kBlock - 1) / kBlock), kBlock, 0, s>>>(p, n0, a, count);
}

void launch_distribute_fp64(const ParticlesF64 &p, int64_t n0, int64_t n, double x0, double x1, double y0, double y1,
                            double xmax, double ymax, double fc, const double *maxw, const Rng &rng, const double *ex,
                            const double *ey, hipStream_t s) {
    if (n <= 0) return;
    k_distribute_fp64<<<grid_for(n, 1 << 20), kBlock, 0, s>>>(p, n0, n, x0, x1, y0, y1, xmax, ymax, fc, maxw, rng, ex,
                                                              ey);
}

}  // namespace edgpu
