// Mixed-precision particle kernels: a particle is a packed int32 cell (j<<16 | i) plus FP32
// in-cell offsets dx, dy in [0,1] and FP32 velocities (24 bytes).  The gather uses FP32
// copies of the field, the Boris push runs in FP32, deposits are accumulated in FP64,
// collisions evaluate their physics in FP64 on the FP32 state.  Cell crossings, walls and
// periodic wraps are decided with integer cell arithmetic, so positions never lose
// resolution far from the origin.
#include "collide.h"
#include "sortwin.h"

namespace edgpu {

__device__ inline int m_i(int32_t c) { return c & 0xffff; }
__device__ inline int m_j(int32_t c) { return c >> 16; }
__device__ inline int32_t m_pack(int i, int j) { return (j << 16) | i; }

__device__ inline void deposit_mixed(double *rho, int nx, int64_t p, float fx, float fy, int Qs) {
    const float vij = (1.0f - fx) * (1.0f - fy), vip1j = fx * (1.0f - fy), vijp1 = (1.0f - fx) * fy;
    const float w3 = 1.0f - vij - vip1j - vijp1;
    const double q = (double)Qs;
    unsafeAtomicAdd(&rho[p], q * (double)vij);
    unsafeAtomicAdd(&rho[p + 1], q * (double)vip1j);
    unsafeAtomicAdd(&rho[p + nx], q * (double)vijp1);
    unsafeAtomicAdd(&rho[p + nx + 1], q * (double)w3);
}

// wall object for a position given as cell + offset (clamped along the wall)
__device__ inline int wall_object_mixed(const Geo &g, int side, int i, float fx, int j, float fy) {
    int ci;
    float cf;
    if (side == EDGPU_SIDE_LEFT || side == EDGPU_SIDE_RIGHT) {
        ci = j;
        cf = fy;
        if (ci < 0) { ci = 0; cf = 0.0f; }
        if (ci > g.gmj || (ci == g.gmj && cf > 0.0f)) { ci = g.gmj; cf = 0.0f; }
    } else {
        ci = i;
        cf = fx;
        if (ci < 0) { ci = 0; cf = 0.0f; }
        if (ci > g.gmi || (ci == g.gmi && cf > 0.0f)) { ci = g.gmi; cf = 0.0f; }
    }
    for (int s = 0; s < g.n_seg; ++s) {
        if (g.seg_side[s] != side) continue;
        if (ci >= g.seg_start[s] && (ci < g.seg_end[s] || (ci == g.seg_end[s] && cf == 0.0f))) return g.seg_obj[s];
    }
    return -1;
}

struct StateMixed {
    int32_t cell;
    float dx, dy, vx, vy, vz;
};

// VZM: kVzReadWrite, kVzRead or kVzSkip (kernels.h)
template <int VZM>
__device__ inline StateMixed load_state(const ParticlesMixed &m, int64_t k) {
    return StateMixed{m.cell[k], m.dx[k], m.dy[k], m.vx[k], m.vy[k], VZM == kVzSkip ? 0.0f : m.vz[k]};
}

template <int VZM>
__device__ inline void store_state(const ParticlesMixed &m, int64_t k, const StateMixed &o) {
    m.cell[k] = o.cell;
    m.dx[k] = o.dx;
    m.dy[k] = o.dy;
    m.vx[k] = o.vx;
    m.vy[k] = o.vy;
    if (VZM == kVzReadWrite) m.vz[k] = o.vz;
}

struct FieldWindowF {
    lds<const float> *ex, *ey;
};

// Returns the state to store in slot k; every path stores all six fields (see
// advance_one_fp64).  A removed particle stores the removal mark cell = -1.  COPIES: the
// number of copies of the caller's LDS deposit window.
template <bool ELECTRON, int COLL, int VZM, bool MOM, int COPIES = kCopiesMixed>
__device__ inline StateMixed advance_one_mixed(const AdvanceArgs &a, int64_t k, const StateMixed &st, double vz2b,
                                               const Window *win,
                                               FieldWindowF fw, double *sR, BlockEvents *ev,
                                               const double *sBound, float *sM) {
    const Geo &g = a.g;
    const int32_t c = st.cell;
    int i = m_i(c), j = m_j(c);
    float fx = st.dx, fy = st.dy;
    if (i < 0 || i > g.gmi - 1 || j < 0 || j > g.gmj - 1) {
        set_error(a.cnt, EDGPU_ERR_INDEX, k);
        return st;
    }
    const float ax_ip1 = fx, ax_i = 1.0f - fx, ay_jp1 = fy, ay_j = 1.0f - fy;
    float ex00, ex10, ex01, ex11, ey00, ey10, ey01, ey11;
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
        ex00 = a.EXf[p];
        ex10 = a.EXf[p + 1];
        ex01 = a.EXf[p + g.nx];
        ex11 = a.EXf[p + g.nx + 1];
        ey00 = a.EYf[p];
        ey10 = a.EYf[p + 1];
        ey01 = a.EYf[p + g.nx];
        ey11 = a.EYf[p + g.nx + 1];
        // wait here, on the rare path, so that the registers these loads fill carry no
        // pending load past the join: the common LDS path then waits on the LDS counter
        // alone and the next particle's loads stay in flight
        wait_vector_memory();
    }
    // The file is compiled without FMA contraction and the fused multiply-adds are written
    // out, so every kernel that inlines this push rounds it the same way (the sorting push
    // must leave the same bits as the plain one).
    const float w00 = ax_i * ay_j, w10 = ax_ip1 * ay_j, w01 = ax_i * ay_jp1, w11 = ax_ip1 * ay_jp1;
    const float E_X = fmaf(ex00, w00, fmaf(ex10, w10, fmaf(ex01, w01, ex11 * w11)));
    const float E_Y = fmaf(ey00, w00, fmaf(ey10, w10, fmaf(ey01, w01, ey11 * w11)));
    float vx = st.vx, vy = st.vy, vz = st.vz;
    const float EZ = (float)a.EZ, qm = (float)a.QM2s, nsub = (float)a.Nsub;
    float K[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) K[q] = (float)a.K[q];
    if (ELECTRON) {
        const float vxm = fmaf(-0.5f, E_X, vx), vym = fmaf(-0.5f, E_Y, vy), vzm = fmaf(-0.5f, EZ, vz);
        vx = fmaf(K[0], vxm, fmaf(K[1], vym, fmaf(K[2], vzm, -0.5f * E_X)));
        vy = fmaf(K[3], vxm, fmaf(K[4], vym, fmaf(K[5], vzm, -0.5f * E_Y)));
        vz = fmaf(K[6], vxm, fmaf(K[7], vym, fmaf(K[8], vzm, -0.5f * EZ)));
        fx += vx;
        fy += vy;
    } else if (a.sense_B) {
        const float vxm = fmaf(qm, E_X, vx), vym = fmaf(qm, E_Y, vy), vzm = fmaf(qm, EZ, vz);
        vx = fmaf(K[0], vxm, fmaf(K[1], vym, fmaf(K[2], vzm, qm * E_X)));
        vy = fmaf(K[3], vxm, fmaf(K[4], vym, fmaf(K[5], vzm, qm * E_Y)));
        vz = fmaf(K[6], vxm, fmaf(K[7], vym, fmaf(K[8], vzm, qm * EZ)));
        fx = fmaf(vx, nsub, fx);
        fy = fmaf(vy, nsub, fy);
    } else {
        const float q2 = qm + qm;
        vx = fmaf(q2, E_X, vx);
        vy = fmaf(q2, E_Y, vy);
        vz = fmaf(q2, EZ, vz);
        fx = fmaf(vx, nsub, fx);
        fy = fmaf(vy, nsub, fy);
    }
    if (MOM) {
        // at the position before the step, with the half-step velocity (the fused moments'
        // order: pic2d_ElectronDynamicsWithVDFMoments.f90:330-460)
        const MomentTerms t =
            moment_terms(ax_i * ay_j, ax_ip1 * ay_j, ax_i * ay_jp1, 0.5 * ((double)vx + (double)st.vx),
                         0.5 * ((double)vy + (double)st.vy), 0.5 * ((double)vz + (double)st.vz));
        moments_add(sM, s, (int64_t)i + (int64_t)j * g.nx, g, a.mom, t);
    }
    // renormalise the offsets into [0,1)
    const float flx = floorf(fx), fly = floorf(fy);
    i += (int)flx;
    j += (int)fly;
    fx -= flx;
    fy -= fly;
    // inner metal objects: open box ileft < x < iright, jbottom < y < jtop; the object whose
    // face the step's trajectory crossed first is found after the push (k_inner_hits_mixed)
    int obj = -1, side = -1;
    double rx = 0.0, ry = 0.0;
    for (int m = 0; m < g.n_inner; ++m) {
        const bool gx = (i > g.box[m][0]) || (i == g.box[m][0] && fx > 0.0f);
        const bool lx = i < g.box[m][1];
        const bool gy = (j > g.box[m][2]) || (j == g.box[m][2] && fy > 0.0f);
        const bool ly = j < g.box[m][3];
        if (gx && lx && gy && ly) {
            obj = kInnerPending;
            break;
        }
    }
    const bool x_lo = i < 0, x_hi = (i > g.gmi) || (i == g.gmi && fx > 0.0f);
    const bool y_lo = j < 0, y_hi = (j > g.gmj) || (j == g.gmj && fy > 0.0f);
    if (obj < 0 && (x_lo || x_hi || y_lo || y_hi)) {
        const int L = g.gmi - 1;
        if (x_lo || x_hi) {
            if (g.periodic_x) {
                i += x_lo ? L : -L;
                if (y_lo) side = EDGPU_SIDE_BELOW;
                else if (y_hi) side = EDGPU_SIDE_ABOVE;
            } else {
                // distance outside the left/right edge and below/above the y edges
                const float dxo = x_lo ? -((float)i + fx) : ((float)(i - g.gmi) + fx);
                const bool right = x_hi;
                const float yrel_lo = (float)j + fy;            // y - ymin
                const float yrel_hi = (float)(j - g.gmj) + fy;  // y - ymax
                if (yrel_lo >= 1.0f && yrel_hi <= -1.0f) side = right ? EDGPU_SIDE_RIGHT : EDGPU_SIDE_LEFT;
                else if (yrel_lo < 1.0f)
                    side = (dxo < -yrel_lo) ? EDGPU_SIDE_BELOW : (right ? EDGPU_SIDE_RIGHT : EDGPU_SIDE_LEFT);
                else side = (dxo < yrel_hi) ? EDGPU_SIDE_ABOVE : (right ? EDGPU_SIDE_RIGHT : EDGPU_SIDE_LEFT);
            }
        } else {
            side = y_hi ? EDGPU_SIDE_ABOVE : EDGPU_SIDE_BELOW;
        }
        if (side >= 0) {
            obj = wall_object_mixed(g, side, i, fx, j, fy);
            if (obj < 0) {
                set_error(a.cnt, EDGPU_ERR_NO_WALL, k);
                return st;
            }
        }
    }
    if (obj >= 0) {
        if (side < 0 || obj < g.n_boundary_obj) {
            rx = (double)i + (double)fx;
            ry = (double)j + (double)fy;
        }
        event_removed(ev, a, k, obj, side, rx, ry);
        return StateMixed{-1, fx, fy, vx, vy, vz};  // removal mark
    }
    // exact upper edges belong to the last cell (pic2d_ElectronDynamics.f90:65-66)
    if (i == g.gmi) {
        i = g.gmi - 1;
        fx = 1.0f;
    }
    if (j == g.gmj) {
        j = g.gmj - 1;
        fy = 1.0f;
    }
    if (COLL == 1) en_candidate(a, ev, k, (double)vx, (double)vy, (double)vz, sBound);
    if (COLL == 2) {
        double dvx = vx, dvy = vy;
        // kVzSkip: vz2b bounds vz^2 of every ion (see run_range_fp64)
        if (VZM != kVzSkip || !rcx_rejects_any_vz(a.rcx, a.rng, k, dvx * dvx + dvy * dvy, vz2b)) {
            if (VZM == kVzSkip) {
                vz = a.m.vz[k];
                wait_vector_memory();  // on this rare path, as in the field gather
            }
            double dvz = vz;
            if (rcx_collide(a.rcx, a.maxw, a.rng, k, dvx, dvy, dvz)) {
                vx = (float)dvx;
                vy = (float)dvy;
                vz = (float)dvz;
                atomicAdd(&ev->coll[0], 1u);
                if (VZM != kVzReadWrite) a.m.vz[k] = vz;  // the exchange draws a new vz
                raise_vz_bound(a.vz2max, (double)vz);
            }
        }
    }
    if (a.rho) {
        const int d = win ? window_slot(*win, i, j) : -1;
        if (d >= 0) {
            const float vij = (1.0f - fx) * (1.0f - fy), vip1j = fx * (1.0f - fy), vijp1 = (1.0f - fx) * fy;
            const float w3 = 1.0f - vij - vip1j - vijp1;
            const float q = (float)a.Qs;
            window_deposit<double, COPIES>(sR, d, q * vij, q * vip1j, q * vijp1, q * w3);
        } else {
            deposit_mixed(a.rho, g.nx, (int64_t)i + (int64_t)j * g.nx, fx, fy, a.Qs);
        }
    }
    return StateMixed{m_pack(i, j), fx, fy, vx, vy, vz};
}

// particles k, k+stride, ... < end of one thread, the next state loaded before the current
// push (see run_range_fp64)
template <bool ELECTRON, int COLL, int VZM, bool MOM>
__device__ inline void run_range_mixed(const AdvanceArgs &a, int64_t k, int64_t end, int64_t stride, const Window *win,
                                       FieldWindowF fw, double *sR, BlockEvents *ev,
                                        const double *sBound, float *sM) {
    if (k >= end) return;
    const double vz2b = (COLL == 2 && VZM == kVzSkip) ? vz_bound(a.vz2max) : 0.0;
    StateMixed cur = load_state<VZM>(a.m, k);
    wait_vector_memory();
    for (;;) {
        const int64_t kn = k + stride;
        const bool more = kn < end;
        const StateMixed nxt = load_state<VZM>(a.m, more ? kn : k);
        store_state<VZM>(a.m, k,
                         advance_one_mixed<ELECTRON, COLL, VZM, MOM>(a, k, cur, vz2b, win, fw, sR, ev, sBound, sM));
        if (!more) break;
        cur = nxt;
        k = kn;
    }
}

// Workgroups [0, ntiles) own one tile each (FP32 weights accumulated in an FP64 LDS window:
// LDS FP64 atomic adds are several times cheaper than FP32 ones on gfx950); the others
// stride over the untiled particles with global memory.
template <bool ELECTRON, int COLL, int VZM, bool MOM>
__global__ __launch_bounds__(kAdvBlock) void k_advance_mixed(AdvanceArgs a) {
    __shared__ double sR[kWinN * kCopiesMixed];
    __shared__ float sEX[kWinN], sEY[kWinN];
    __shared__ BlockEvents ev;
    __shared__ double sBound[COLL == 1 ? kEnBins : 1];
    __shared__ float sM[MOM ? kMoments * kWinN : 1];
    events_init(&ev);
    if (COLL == 1) stage_bound(a, sBound);
    const int b = blockIdx.x;
    if (b < a.tiles.ntiles) {
        const int t = tile_of_block(b, a.tiles.ntiles);
        const Window w = tile_window(t, a.tiles.ntx, a.g);
        window_clear<double, kCopiesMixed>(sR);
        if (MOM) moments_clear(sM);
        window_load_field(w, a.g, a.EXf, a.EYf, sEX, sEY);
        __syncthreads();
        const int64_t s0 = a.tiles.start[t];
        const int64_t s1 = min((int64_t)a.tiles.start[t + 1], a.n);
        const FieldWindowF fw{to_lds<const float>(sEX), to_lds<const float>(sEY)};
        run_range_mixed<ELECTRON, COLL, VZM, MOM>(a, s0 + threadIdx.x, s1, blockDim.x, &w, fw,
                                             window_copy<double, kCopiesMixed>(sR), &ev, sBound, sM);
        __syncthreads();
        if (a.rho) window_flush<double, kCopiesMixed>(w, a.g, sR, a.rho);
        if (MOM) moments_flush(w, a.g, sM, a.mom);
    } else {
        __syncthreads();
        const int64_t begin = min(a.tiles.covered, a.n);
        const int64_t stride = (int64_t)(gridDim.x - a.tiles.ntiles) * blockDim.x;
        run_range_mixed<ELECTRON, COLL, VZM, MOM>(a, begin + (int64_t)(b - a.tiles.ntiles) * blockDim.x + threadIdx.x,
                                             a.n, stride, nullptr, FieldWindowF{nullptr, nullptr}, nullptr, &ev,
                                             sBound, nullptr);
    }
    events_flush(&ev, a);
}

template <bool ELECTRON, int COLL, int VZM>
static void launch_advance_kernel(const AdvanceArgs &a, int grid, hipStream_t s) {
    if (a.mom) k_advance_mixed<ELECTRON, COLL, VZM, true><<<grid, kAdvBlock, 0, s>>>(a);
    else k_advance_mixed<ELECTRON, COLL, VZM, false><<<grid, kAdvBlock, 0, s>>>(a);
}

// ---------------------------------------------------------------------------------
// The sorting push (electrons): on a step whose sort is due the push itself writes the
// particles to the spare arrays in tile order, so the sort needs no pass over the particles
// of its own.  The tiles are those of the positions before the push, counted beforehand
// (launch_tile_counts_mixed: per-tile cursors from the scanned histogram, and every source
// tile's counts over its destination window): the order is one step old when the next push
// reads it, which the window margins absorb.
//
// A tile workgroup reserves, once, one run per destination tile of its window from the
// cursors and keeps the runs' next free slots in LDS.  It takes its particles in chunks of
// kSortPushChunk, the next chunk's loads in flight while the current one is ranked, pushed
// and placed: every particle takes a rank in its destination tile (LDS atomic), a scan of the
// window gives the stage offsets, the particles are pushed with their destination slot as
// index (events and collision draws are keyed by the slot the particle holds from now on), and
// each array goes out through LDS as contiguous runs (sortwin.h).  The state moves whole (vz
// and the tag too).  Particles outside the window and the untiled ones take single slots from
// their tile's cursor.
constexpr int kSortPushPer = 2;
constexpr int kSortPushChunk = kSortPushPer * kAdvBlock;

__device__ inline void store_particle(const ParticlesMixed &m, int64_t d, const StateMixed &o, int32_t tag) {
    m.cell[d] = o.cell;
    m.dx[d] = o.dx;
    m.dy[d] = o.dy;
    m.vx[d] = o.vx;
    m.vy[d] = o.vy;
    m.vz[d] = o.vz;
    m.tag[d] = tag;
}

// the untiled particles k, k+stride, ... < end of one thread: single slots from the cursors
template <int COLL, int VZM>
__device__ inline void sort_push_untiled_mixed(const AdvanceArgs &a, const ParticlesMixed &out, int32_t *cursor,
                                               int64_t k, int64_t end, int64_t stride, BlockEvents *ev,
                                               const double *sBound) {
    const int ntx = a.tiles.ntx;
    for (; k < end; k += stride) {
        const StateMixed st = load_state<kVzReadWrite>(a.m, k);
        const int32_t tag = a.m.tag[k];
        int tx, j;
        sort_tile_mixed(st.cell, a.g, tx, j);
        const int64_t d = atomicAdd(&cursor[tx + j * ntx], 1);
        if (d >= a.n) {
            set_error(a.cnt, EDGPU_ERR_INDEX, k);
            continue;
        }
        store_particle(out, d,
                       advance_one_mixed<true, COLL, VZM, false>(a, d, st, 0.0, nullptr, FieldWindowF{nullptr, nullptr},
                                                                 nullptr, ev, sBound, nullptr),
                       tag);
    }
}

// the states of one thread's particles of a chunk: the cells first, so that the ranking waits
// for them alone
struct ChunkMixed {
    StateMixed st[kSortPushPer];
    int32_t tag[kSortPushPer];
};

__device__ inline void load_chunk_mixed(const ParticlesMixed &m, int64_t c0, int64_t c1, ChunkMixed &c) {
#pragma unroll
    for (int u = 0; u < kSortPushPer; ++u) c.st[u].cell = m.cell[min(c0 + threadIdx.x + (int64_t)u * kAdvBlock, c1 - 1)];
#pragma unroll
    for (int u = 0; u < kSortPushPer; ++u) {
        const int64_t k = min(c0 + threadIdx.x + (int64_t)u * kAdvBlock, c1 - 1);
        c.st[u].dx = m.dx[k];
        c.st[u].dy = m.dy[k];
        c.st[u].vx = m.vx[k];
        c.st[u].vy = m.vy[k];
        c.st[u].vz = m.vz[k];
        c.tag[u] = m.tag[k];
    }
}

template <int COLL, int VZM>
__global__ __launch_bounds__(kAdvBlock) void k_advance_sort_mixed(AdvanceArgs a, ParticlesMixed out,
                                                                  int32_t *cursor, const int32_t *wcount) {
    __shared__ double sR[kWinN * kCopiesMixed];
    __shared__ float sEX[kWinN], sEY[kWinN];
    __shared__ BlockEvents ev;
    __shared__ double sBound[COLL == 1 ? kEnBins : 1];
    __shared__ PushStage<kSortPushChunk> ps;
    __shared__ uint32_t stage[kSortPushChunk];
    events_init(&ev);
    if (COLL == 1) stage_bound(a, sBound);
    const Geo &g = a.g;
    const int ntx = a.tiles.ntx;
    const int b = blockIdx.x;
    if (b >= a.tiles.ntiles) {
        __syncthreads();
        const int64_t begin = min(a.tiles.covered, a.n);
        const int64_t stride = (int64_t)(gridDim.x - a.tiles.ntiles) * blockDim.x;
        sort_push_untiled_mixed<COLL, VZM>(a, out, cursor, begin + (int64_t)(b - a.tiles.ntiles) * blockDim.x + threadIdx.x,
                                           a.n, stride, &ev, sBound);
        events_flush(&ev, a);
        return;
    }
    const int t = tile_of_block(b, a.tiles.ntiles);
    const Window w = tile_window(t, ntx, g);
    const SortWindow sw = sort_window(t, ntx, g);
    window_clear<double, kCopiesMixed>(sR);
    window_load_field(w, g, a.EXf, a.EYf, sEX, sEY);
    double *sRc = window_copy<double, kCopiesMixed>(sR);
    const FieldWindowF fw{to_lds<const float>(sEX), to_lds<const float>(sEY)};
    const int64_t s0 = a.tiles.start[t];
    const int64_t s1 = min((int64_t)a.tiles.start[t + 1], a.n);
    ChunkMixed cur;
    if (s0 < s1) load_chunk_mixed(a.m, s0, min(s0 + (int64_t)kSortPushChunk, s1), cur);
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
                sort_tile_mixed(cur.st[u].cell, g, tx, j);
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
        ChunkMixed nxt;
        const int64_t n0 = c0 + kSortPushChunk;
        if (n0 < s1) load_chunk_mixed(a.m, n0, min(n0 + (int64_t)kSortPushChunk, s1), nxt);
#pragma unroll
        for (int u = 0; u < kSortPushPer; ++u) {
            if (dst[u] >= a.n) {
                set_error(a.cnt, EDGPU_ERR_INDEX, c0 + threadIdx.x + (int64_t)u * kAdvBlock);
            } else if (dst[u] >= 0) {
                cur.st[u] = advance_one_mixed<true, COLL, VZM, false>(a, dst[u], cur.st[u], 0.0, &w,
                                                                      fw, sRc, &ev, sBound, nullptr);
            }
        }
        uint32_t v[kSortPushPer];
#define EDGPU_PLACE(FIELD, OUT)                                                                           \
    for (int u = 0; u < kSortPushPer; ++u) v[u] = __builtin_bit_cast(uint32_t, cur.st[u].FIELD);          \
    place_array<kSortPushPer, uint32_t>(v, (uint32_t *)OUT, slot, dst, stage, ps.dest, m, a.n);
        EDGPU_PLACE(cell, out.cell)
        EDGPU_PLACE(dx, out.dx)
        EDGPU_PLACE(dy, out.dy)
        EDGPU_PLACE(vx, out.vx)
        EDGPU_PLACE(vy, out.vy)
        EDGPU_PLACE(vz, out.vz)
#undef EDGPU_PLACE
        place_array<kSortPushPer, uint32_t>((const uint32_t *)cur.tag, (uint32_t *)out.tag, slot, dst, stage, ps.dest,
                                            m, a.n);
        cur = nxt;
    }
    __syncthreads();
    if (a.rho) window_flush<double, kCopiesMixed>(w, g, sR, a.rho);
    events_flush(&ev, a);
}

// process selection of the listed candidates, then the kinematics of the selected slots
// (the FP64 kernels explain the split)
constexpr int kSelBuf = 256;

__global__ __launch_bounds__(kBlock) void k_select_mixed(AdvanceArgs a) {
    __shared__ BlockList<kSelBuf> sl;
    block_list_init(&sl);
    const int64_t nc = (int64_t)min(a.cnt->n_cand, (unsigned long long)a.cand_cap);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nc; t += stride) {
        const int64_t k = a.cand[t];
        double u0[4];
        draw_en(a.rng, k, u0);
        double en;
        const int coll = en_select(*a.en, u0, (double)a.m.vx[k], (double)a.m.vy[k], (double)a.m.vz[k], k, a.cnt, en);
        if (coll >= 1 && coll <= a.en->n_proc) block_list_push(&sl, k, &a.cnt->n_sel, a.sel, a.cand_cap, a.cnt);
    }
    block_list_flush(&sl, &a.cnt->n_sel, a.sel, a.cand_cap, a.cnt);
}

__global__ __launch_bounds__(kBlock) void k_candidates_mixed(AdvanceArgs a) {
    __shared__ unsigned int sc[EDGPU_MAX_PROC];
    coll_count_init(sc);
    const int64_t ns = (int64_t)min(a.cnt->n_sel, (unsigned long long)a.cand_cap);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ns; t += stride) {
        const int64_t k = a.sel[t];
        double u0[4];
        draw_en(a.rng, k, u0);
        const int32_t c = a.m.cell[k];
        double vx = a.m.vx[k], vy = a.m.vy[k], vz = a.m.vz[k];
        if (en_collide_candidate(a.en, u0, a.maxw, a.rng, k, (double)m_i(c) + (double)a.m.dx[k],
                                 (double)m_j(c) + (double)a.m.dy[k], vx, vy, vz, a.add_e, a.add_i, a.cnt, sc)) {
            a.m.vx[k] = (float)vx;
            a.m.vy[k] = (float)vy;
            a.m.vz[k] = (float)vz;
        }
    }
    coll_count_flush(sc, a.cnt);
}

__global__ __launch_bounds__(kBlock) void k_deposit_mixed(ParticlesMixed p, int64_t begin, int64_t n, Geo g,
                                                           double *rho, int Qs) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const int32_t c = p.cell[k];
        deposit_mixed(rho, g.nx, (int64_t)m_i(c) + (int64_t)m_j(c) * g.nx, p.dx[k], p.dy[k], Qs);
    }
}

__global__ __launch_bounds__(kBlock) void k_mcc_en_mixed(ParticlesMixed p, int64_t begin, int64_t n, const EnTable *t, double pmax,
                                                          const double *maxw, Rng rng, AddList ae, IonAddLists ai,
                                                          Counters *cnt) {
    __shared__ unsigned int sc[EDGPU_MAX_PROC];
    coll_count_init(sc);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        double vx = p.vx[k], vy = p.vy[k], vz = p.vz[k];
        const int32_t c = p.cell[k];
        if (en_collide(t, pmax, maxw, rng, k, (double)m_i(c) + (double)p.dx[k], (double)m_j(c) + (double)p.dy[k], vx, vy, vz,
                       ae, ai, cnt, sc)) {
            p.vx[k] = (float)vx;
            p.vy[k] = (float)vy;
            p.vz[k] = (float)vz;
        }
    }
    coll_count_flush(sc, cnt);
}

__global__ __launch_bounds__(kBlock) void k_rcx_mixed(ParticlesMixed p, int64_t begin, int64_t n, RcxPar r,
                                                       const double *maxw, Rng rng, Counters *cnt,
                                                       unsigned long long *vz2max) {
    __shared__ unsigned int sc[EDGPU_MAX_PROC];
    coll_count_init(sc);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        double vx = p.vx[k], vy = p.vy[k], vz = p.vz[k];
        if (rcx_collide(r, maxw, rng, k, vx, vy, vz)) {
            p.vx[k] = (float)vx;
            p.vy[k] = (float)vy;
            p.vz[k] = (float)vz;
            atomicAdd(&sc[0], 1u);
            raise_vz_bound(vz2max, (double)(float)vz);
        }
    }
    coll_count_flush(sc, cnt);
}

__global__ void k_find_movers_mixed(const int32_t *cell, int64_t new_n, int64_t old_n, int64_t *movers, Counters *cnt) {
    int64_t t = new_n + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= old_n) return;
    if (cell[t] != -1) movers[atomicAdd(&cnt->n_movers, 1ull)] = t;
}

__global__ void k_find_holes_mixed(const int64_t *removed, int64_t nrem, int64_t new_n, int64_t *holes, Counters *cnt) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrem) return;
    int64_t idx = removed[t];
    if (idx < new_n) holes[atomicAdd(&cnt->n_holes, 1ull)] = idx;
}

// The initial loading's removal of the particles placed inside inner objects
// (pic2d_CurProblemValues.f90:3245-3258), tested on the FP64 positions a loaded chunk was
// converted from: the converted particle is marked as a push marks a removal (cell = -1).
__global__ void k_mark_inside_mixed(const double *x, const double *y, int32_t *cell, int64_t n, Geo g,
                                    Counters *cnt) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    unsigned long long c = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        if (inside_inner(g, x[k], y[k])) {
            cell[k] = -1;
            ++c;
        }
    }
    if (c) atomicAdd(&cnt->n_removed, c);
}

__global__ void k_marked_holes_mixed(const int32_t *cell, int64_t new_n, int64_t *holes, Counters *cnt) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < new_n; k += stride)
        if (cell[k] == -1) holes[atomicAdd(&cnt->n_holes, 1ull)] = k;
}

// The particles a push left inside an inner box, as k_inner_hits_fp64 (the positions were
// recorded from cell and offset, the velocities are the pushed FP32 ones)
__global__ void k_inner_hits_mixed(AdvanceArgs a, double back) {
    const int64_t n = min((int64_t)a.cnt->n_removed, a.removed_cap);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        if ((a.removed_obj[q] & 0xff) != kInnerPending) continue;
        const int64_t k = a.removed_idx[q];
        const double x = a.removed_x[q], y = a.removed_y[q];
        const double vx = (double)a.m.vx[k], vy = (double)a.m.vy[k];
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

__global__ void k_move_mixed(ParticlesMixed p, const int64_t *holes, const int64_t *movers, const Counters *cnt) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t count = (int64_t)min(cnt->n_holes, cnt->n_movers);
    if (t >= count) return;
    int64_t d = holes[t], s = movers[t];
    p.cell[d] = p.cell[s];
    p.dx[d] = p.dx[s];
    p.dy[d] = p.dy[s];
    p.vx[d] = p.vx[s];
    p.vy[d] = p.vy[s];
    p.vz[d] = p.vz[s];
    p.tag[d] = p.tag[s];
}

// FP64 position -> cell + offset, with x == xmax in the last cell
__device__ inline void to_cell(double x, double xmax, int imax, int &i, float &f) {
    i = (int)x;
    if (x == xmax) i = imax - 1;
    f = (float)(x - (double)i);
}

__global__ void k_append_mixed(ParticlesMixed p, int64_t n0, Geo g, AddList a, int64_t count) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    int i, j;
    float fx, fy;
    to_cell(a.x[t], g.xmax, g.gmi, i, fx);
    to_cell(a.y[t], g.ymax, g.gmj, j, fy);
    p.cell[n0 + t] = m_pack(i, j);
    p.dx[n0 + t] = fx;
    p.dy[n0 + t] = fy;
    p.vx[n0 + t] = (float)a.vx[t];
    p.vy[n0 + t] = (float)a.vy[t];
    p.vz[n0 + t] = (float)a.vz[t];
    p.tag[n0 + t] = a.tag[t];
}

__global__ void k_to_mixed(ParticlesF64 s, ParticlesMixed d, int64_t n, Geo g) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        int i, j;
        float fx, fy;
        to_cell(s.x[k], g.xmax, g.gmi, i, fx);
        to_cell(s.y[k], g.ymax, g.gmj, j, fy);
        d.cell[k] = m_pack(i, j);
        d.dx[k] = fx;
        d.dy[k] = fy;
        d.vx[k] = (float)s.vx[k];
        d.vy[k] = (float)s.vy[k];
        d.vz[k] = (float)s.vz[k];
        d.tag[k] = s.tag[k];
    }
}

__global__ void k_to_fp64(ParticlesMixed s, ParticlesF64 d, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const int32_t c = s.cell[k];
        d.x[k] = (double)m_i(c) + (double)s.dx[k];
        d.y[k] = (double)m_j(c) + (double)s.dy[k];
        d.vx[k] = s.vx[k];
        d.vy[k] = s.vy[k];
        d.vz[k] = s.vz[k];
        d.tag[k] = s.tag[k];
    }
}

__global__ void k_field_to_f32(const double *a, const double *b, float *fa, float *fb, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        fa[k] = (float)a[k];
        fb[k] = (float)b[k];
    }
}

static int grid_m(int64_t n) {
    int64_t b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    if (b > (1 << 20)) b = 1 << 20;
    return (int)b;
}

void launch_advance_mixed(const AdvanceArgs &a, bool electron, int coll, int vz_mode, hipStream_t s) {
    if (a.n <= 0) return;
    const int grid = advance_grid(a);
    if (electron) {
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

void launch_advance_sort_mixed(const AdvanceArgs &a, const ParticlesMixed &out, const SortWork &w, int coll,
                               int vz_mode, hipStream_t s) {
    if (a.n <= 0) return;
    const int grid = advance_grid(a);
    if (vz_mode == kVzReadWrite) {
        if (coll == 1) k_advance_sort_mixed<1, kVzReadWrite><<<grid, kAdvBlock, 0, s>>>(a, out, w.count, w.wcount);
        else k_advance_sort_mixed<0, kVzReadWrite><<<grid, kAdvBlock, 0, s>>>(a, out, w.count, w.wcount);
    } else {
        if (coll == 1) k_advance_sort_mixed<1, kVzRead><<<grid, kAdvBlock, 0, s>>>(a, out, w.count, w.wcount);
        else k_advance_sort_mixed<0, kVzRead><<<grid, kAdvBlock, 0, s>>>(a, out, w.count, w.wcount);
    }
}

void launch_vz_bound_mixed(const ParticlesMixed &p, int64_t begin, int64_t end, unsigned long long *bound,
                           hipStream_t s) {
    if (end <= begin) return;
    k_vz_bound<float><<<grid_m(end - begin), kBlock, 0, s>>>(p.vz, begin, end, bound);
}

void launch_candidates_mixed(const AdvanceArgs &a, hipStream_t s) {
    const int64_t blocks = std::min<int64_t>(std::max<int64_t>((a.cand_cap + kBlock - 1) / kBlock, 1), 65536);
    k_select_mixed<<<(int)blocks, kBlock, 0, s>>>(a);
    k_candidates_mixed<<<(int)std::min<int64_t>(blocks, 4096), kBlock, 0, s>>>(a);
}

void launch_deposit_mixed(const ParticlesMixed &p, int64_t begin, int64_t n, const Geo &g, double *rho, int Qs,
                          hipStream_t s) {
    if (n <= begin) return;
    k_deposit_mixed<<<grid_m(n - begin), kBlock, 0, s>>>(p, begin, n, g, rho, Qs);
}

void launch_mcc_en_mixed(const ParticlesMixed &p, int64_t begin, int64_t n, const Geo &g, const EnTable *t, double pmax,
                         const double *maxw, const Rng &rng, const AddList &ae, const IonAddLists &ai, Counters *cnt,
                         hipStream_t s) {
    (void)g;
    if (n <= begin) return;
    k_mcc_en_mixed<<<grid_m(n - begin), kBlock, 0, s>>>(p, begin, n, t, pmax, maxw, rng, ae, ai, cnt);
}

void launch_rcx_mixed(const ParticlesMixed &p, int64_t begin, int64_t n, const RcxPar &r, const double *maxw,
                      const Rng &rng, Counters *cnt, unsigned long long *vz2max, hipStream_t s) {
    if (n <= begin) return;
    k_rcx_mixed<<<grid_m(n - begin), kBlock, 0, s>>>(p, begin, n, r, maxw, rng, cnt, vz2max);
}

void launch_compact_mixed(const ParticlesMixed &p, const int64_t *removed, int64_t nrem, int64_t old_n,
                          int64_t *holes, int64_t *movers, Counters *cnt, hipStream_t s) {
    if (nrem <= 0) return;
    int64_t new_n = old_n - nrem;
    int g = (int)((nrem + kBlock - 1) / kBlock);
    k_find_holes_mixed<<<g, kBlock, 0, s>>>(removed, nrem, new_n, holes, cnt);
    k_find_movers_mixed<<<g, kBlock, 0, s>>>(p.cell, new_n, old_n, movers, cnt);
    k_move_mixed<<<g, kBlock, 0, s>>>(p, holes, movers, cnt);
}

void launch_mark_inside_mixed(const ParticlesF64 &src, const ParticlesMixed &dst, int64_t n, const Geo &g,
                              Counters *cnt, hipStream_t s) {
    if (n <= 0 || g.n_inner == 0) return;
    k_mark_inside_mixed<<<grid_m(n), kBlock, 0, s>>>(src.x, src.y, dst.cell, n, g, cnt);
}

void launch_fill_marked_mixed(const ParticlesMixed &p, int64_t new_n, int64_t old_n, int64_t *holes, int64_t *movers,
                              Counters *cnt, hipStream_t s) {
    const int64_t m = old_n - new_n;
    if (m <= 0) return;
    const int g = (int)((m + kBlock - 1) / kBlock);
    k_marked_holes_mixed<<<grid_m(new_n), kBlock, 0, s>>>(p.cell, new_n, holes, cnt);
    k_find_movers_mixed<<<g, kBlock, 0, s>>>(p.cell, new_n, old_n, movers, cnt);
    k_move_mixed<<<g, kBlock, 0, s>>>(p, holes, movers, cnt);
}

void launch_inner_hits_mixed(const AdvanceArgs &a, bool electron, hipStream_t s) {
    if (a.g.n_inner == 0 || a.n <= 0) return;
    k_inner_hits_mixed<<<64, kBlock, 0, s>>>(a, electron ? 1.0 : a.Nsub);
}

void launch_append_mixed(const ParticlesMixed &p, int64_t n0, const Geo &g, const AddList &a, int64_t count,
                         hipStream_t s) {
    if (count <= 0) return;
    k_append_mixed<<<(int)((count + kBlock - 1) / kBlock), kBlock, 0, s>>>(p, n0, g, a, count);
}

void launch_to_mixed(const ParticlesF64 &src, const ParticlesMixed &dst, int64_t n, const Geo &g, hipStream_t s) {
    if (n <= 0) return;
    k_to_mixed<<<grid_m(n), kBlock, 0, s>>>(src, dst, n, g);
}

void launch_to_fp64(const ParticlesMixed &src, const ParticlesF64 &dst, int64_t n, const Geo &g, hipStream_t s) {
    (void)g;
    if (n <= 0) return;
    k_to_fp64<<<grid_m(n), kBlock, 0, s>>>(src, dst, n);
}

void launch_field_to_f32(const double *a, const double *b, float *fa, float *fb, int64_t n, hipStream_t s) {
    k_field_to_f32<<<grid_m(n), kBlock, 0, s>>>(a, b, fa, fb, n);
}

}  // namespace edgpu
