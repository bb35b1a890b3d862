// Tile geometry shared by the sorted push and the counting sort.
//
// After a counting sort the particles of a species are ordered by tile: a tile is a run of
// kTileW cells of one cell row, and its particles (in no particular order among themselves)
// occupy the contiguous index range [start[t], start[t+1]).  A workgroup that owns a tile
// stages the field nodes of a window around the tile in LDS and accumulates its deposit
// in an LDS copy of the same window; particles that drifted further than the window margin
// (kMarginX cells beyond the tile ends, kMarginY rows above and below) since the sort fall
// back to global memory.  The y margin is the larger one: the tile is one row high, so
// drift in y leaves it after a few steps while drift in x mostly stays inside 16 cells.  Particles appended after the sort
// (and any index range the tiles do not cover) are handled by "tail" workgroups that
// use global memory only.
#pragma once

#include "collide.h"

namespace edgpu {

constexpr int kTileW = 16;
// threads of a push workgroup (one tile)
constexpr int kAdvBlock = 256;
constexpr int kMarginX = 2;
constexpr int kMarginY = 4;
constexpr int kWinW = kTileW + 2 * kMarginX + 1;  // window nodes in x
constexpr int kWinH = 2 * kMarginY + 2;           // window nodes in y
constexpr int kWinN = kWinW * kWinH;
constexpr int kWinCW = kTileW + 2 * kMarginX;     // window cells in x
constexpr int kWinCH = 2 * kMarginY + 1;          // window cells in y
constexpr int kWinCN = kWinCW * kWinCH;

struct Window {
    int x0, y0, x1, y1;  // node window [x0,x1] x [y0,y1] (inclusive), clamped to the grid
};

// Pointers into LDS.  A push reads each particle's field nodes from the LDS window or, for
// a particle outside it, from global memory; through generic pointers the compiler merges
// the two branches into flat loads, whose wait (vmcnt and lgkmcnt both to zero) also drains
// the next particle's state loads that were issued ahead of the push.  Loads and atomics
// through these pointers compile to ds_* instructions and wait on the LDS counter alone.
template <typename T>
using lds = __attribute__((address_space(3))) T;

template <typename T>
__device__ inline lds<T> *to_lds(T *p) {
    return (lds<T> *)p;
}

__device__ inline void lds_add(lds<float> *p, float v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ inline void lds_add(lds<double> *p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Tile owned by workgroup b < ntiles.  Workgroups are dispatched round-robin over the 8
// XCDs (b % 8); each XCD gets a contiguous run of tiles so that neighbouring tiles, whose
// windows overlap in the grid and whose particles land in neighbouring slots of a sorted
// array, share one L2.  A bijection of [0, ntiles) for any ntiles.
constexpr int kXcds = 8;

__device__ inline int tile_of_block(int b, int ntiles) {
    const int q = ntiles / kXcds, rem = ntiles % kXcds;
    const int x = b % kXcds, r = b / kXcds;
    return x * q + min(x, rem) + r;
}

__device__ inline Window tile_window(int t, int ntx, const Geo &g) {
    const int j = t / ntx, i0 = (t % ntx) * kTileW;
    Window w;
    w.x0 = max(0, i0 - kMarginX);
    w.x1 = min(g.gmi, i0 + kTileW + kMarginX);
    w.y0 = max(0, j - kMarginY);
    w.y1 = min(g.gmj, j + 1 + kMarginY);
    return w;
}

// LDS slot of the lower-left node (i,j) of a cell when all four nodes are in the window
__device__ inline int window_slot(const Window &w, int i, int j) {
    if (i < w.x0 || i + 1 > w.x1 || j < w.y0 || j + 1 > w.y1) return -1;
    return (i - w.x0) + (j - w.y0) * kWinW;
}

// cell window (for the sort histogram): cells [x0, x1-1] x [y0, y1-1]
__device__ inline int window_cell(const Window &w, int i, int j) {
    if (i < w.x0 || i > w.x1 - 1 || j < w.y0 || j > w.y1 - 1) return -1;
    return (i - w.x0) + (j - w.y0) * kWinCW;
}


// ---------------------------------------------------------------------------------
// Per-workgroup accumulation of the rare events of a push (wall hits, removals, collision
// candidates) in LDS: one global reservation per workgroup instead of one global atomic per
// event on a handful of addresses.  Entries that do not fit the LDS buffers go straight to
// global memory.
constexpr int kCandBuf = 512;
constexpr int kRemBuf = 128;

struct BlockEvents {
    unsigned int hit[EDGPU_MAX_OBJ];
    unsigned int coll[EDGPU_MAX_PROC];
    int n_cand, n_rem;
    unsigned long long base_cand, base_rem;
    int64_t cand[kCandBuf];
    int64_t rem_idx[kRemBuf];
    int32_t rem_obj[kRemBuf];
    double rem_x[kRemBuf], rem_y[kRemBuf];
};

__device__ inline void events_init(BlockEvents *ev) {
    for (int q = threadIdx.x; q < EDGPU_MAX_OBJ; q += blockDim.x) ev->hit[q] = 0;
    for (int q = threadIdx.x; q < EDGPU_MAX_PROC; q += blockDim.x) ev->coll[q] = 0;
    if (threadIdx.x == 0) {
        ev->n_cand = 0;
        ev->n_rem = 0;
    }
}

__device__ inline void set_error_t(Counters *c, int code, int64_t k) {
    if (atomicCAS(&c->err, 0, code) == 0) c->err_index = (long long)k;
}

// Object of a particle the push found inside an inner box: the box and face its trajectory
// crossed first, the crossing and the hit count are filled in after the push
// (k_inner_hits_*), which keeps that rare and register-hungry path out of the push.
constexpr int kInnerPending = 0xff;

// Particle k leaves through object obj, crossing wall `side` (-1: an inner object) at the
// position (x, y).  The record keeps object and side packed as obj | (side + 1) << 8.
__device__ inline void event_removed(BlockEvents *ev, const AdvanceArgs &a, int64_t k, int obj, int side, double x,
                                     double y) {
    if (obj != kInnerPending) atomicAdd(&ev->hit[obj], 1u);
    const int code = obj | ((side + 1) << 8);
    const int r = atomicAdd(&ev->n_rem, 1);
    if (r < kRemBuf) {
        ev->rem_idx[r] = k;
        ev->rem_obj[r] = code;
        ev->rem_x[r] = x;
        ev->rem_y[r] = y;
        return;
    }
    const unsigned long long g = atomicAdd(&a.cnt->n_removed, 1ull);
    if ((int64_t)g < a.removed_cap) {
        a.removed_idx[g] = k;
        a.removed_obj[g] = code;
        a.removed_x[g] = x;
        a.removed_y[g] = y;
    } else {
        set_error_t(a.cnt, EDGPU_ERR_CAPACITY, k);
    }
}

__device__ inline void event_candidate(BlockEvents *ev, Counters *cnt, int64_t *gcand, int64_t cap, int64_t k) {
    const int r = atomicAdd(&ev->n_cand, 1);
    if (r < kCandBuf) {
        ev->cand[r] = k;
        return;
    }
    const unsigned long long g = atomicAdd(&cnt->n_cand, 1ull);
    if ((int64_t)g < cap) gcand[g] = k;
    else set_error_t(cnt, EDGPU_ERR_CAPACITY, k);
}

// after the workgroup's particles: reserve global ranges and copy the buffered events
__device__ inline void events_flush(BlockEvents *ev, const AdvanceArgs &a) {
    Counters *cnt = a.cnt;
    int64_t *gcand = a.cand;
    const int64_t ccap = a.cand_cap, rcap = a.removed_cap;
    __syncthreads();
    const int nc = min(ev->n_cand, kCandBuf), nr = min(ev->n_rem, kRemBuf);
    if (threadIdx.x == 0) {
        ev->base_cand = nc ? atomicAdd(&cnt->n_cand, (unsigned long long)nc) : 0ull;
        ev->base_rem = nr ? atomicAdd(&cnt->n_removed, (unsigned long long)nr) : 0ull;
    }
    for (int q = threadIdx.x; q < EDGPU_MAX_OBJ; q += blockDim.x)
        if (ev->hit[q]) atomicAdd(&cnt->hit[q], (unsigned long long)ev->hit[q]);
    for (int q = threadIdx.x; q < EDGPU_MAX_PROC; q += blockDim.x)
        if (ev->coll[q]) atomicAdd(&cnt->coll[q], (unsigned long long)ev->coll[q]);
    __syncthreads();
    for (int q = threadIdx.x; q < nc; q += blockDim.x) {
        const int64_t at = (int64_t)ev->base_cand + q;
        if (at < ccap) gcand[at] = ev->cand[q];
        else set_error_t(cnt, EDGPU_ERR_CAPACITY, ev->cand[q]);
    }
    for (int q = threadIdx.x; q < nr; q += blockDim.x) {
        const int64_t at = (int64_t)ev->base_rem + q;
        if (at < rcap) {
            a.removed_idx[at] = ev->rem_idx[q];
            a.removed_obj[at] = ev->rem_obj[q];
            a.removed_x[at] = ev->rem_x[q];
            a.removed_y[at] = ev->rem_y[q];
        } else {
            set_error_t(cnt, EDGPU_ERR_CAPACITY, ev->rem_idx[q]);
        }
    }
}

// null-collision candidacy inside a push, on the pushed velocity.  A candidate whose
// process-selection draw exceeds the bound of the total selection probability over its
// energy bin (sBound, staged in LDS) is certainly a null collision and is dropped here;
// the others join the candidate list, where the selection kernel repeats the exact test
// with the same draws.  The exact selection (dependent table loads) stays out of the push:
// nearly every wave would stall on it.
__device__ inline void en_candidate(const AdvanceArgs &a, BlockEvents *ev, int64_t k, double vx, double vy,
                                    double vz, const double *sBound) {
    double u0[2];
    draw2(a.rng, k, u0);
    if (!(u0[0] < a.en_pmax)) return;
    const double en = (vx * vx + vy * vy + vz * vz) * a.en_V2 * kMeKg * 0.5 / kECl;
    const int b = en_bin(en);
    if (b >= 0 && u0[1] > sBound[b]) return;
    event_candidate(ev, a.cnt, a.cand, a.cand_cap, k);
}

// stage the bound table of the e-n collisions in LDS
__device__ inline void stage_bound(const AdvanceArgs &a, double *sBound) {
    for (int q = threadIdx.x; q < kEnBins; q += blockDim.x) sBound[q] = a.en->bound[q];
}

// ---------------------------------------------------------------------------------
// Grouping of the active lanes of a wave by an integer key: rank of the lane among the
// lanes with the same key, the group size and the lowest lane of the group (the leader).
// The loop runs once per distinct key, which for sorted particles is one or two.
struct WaveGroup {
    int rank, size, leader;
};

template <typename F>
__device__ inline WaveGroup wave_group(int key, bool active, F &&on_leader) {
    const int lane = __lane_id();
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    unsigned long long rem = __ballot(active);
    WaveGroup wg{0, 0, -1};
    while (rem) {
        const int leader = __ffsll((unsigned long long)rem) - 1;
        const int lk = __shfl(key, leader);
        const unsigned long long m = __ballot(active && key == lk);
        const int size = __popcll(m);
        int ret = 0;
        if (lane == leader) ret = on_leader(lk, size);
        ret = __shfl(ret, leader);
        if (active && key == lk) {
            wg.rank = ret + __popcll(m & lt);
            wg.size = size;
            wg.leader = leader;
        }
        rem &= ~m;
    }
    return wg;
}

// stage the field nodes of the window in LDS (nodes outside the grid read as 0)
template <typename T>
__device__ inline void window_load_field(const Window &w, const Geo &g, const T *EX, const T *EY, T *sEX, T *sEY) {
    for (int q = threadIdx.x; q < kWinN; q += blockDim.x) {
        const int i = w.x0 + q % kWinW, j = w.y0 + q / kWinW;
        T ex = T(0), ey = T(0);
        if (i <= w.x1 && j <= w.y1) {
            const int64_t p = (int64_t)i + (int64_t)j * g.nx;
            ex = EX[p];
            ey = EY[p];
        }
        sEX[q] = ex;
        sEY[q] = ey;
    }
}

// The LDS deposit window is replicated: node q of copy c sits at q*C + c and lane l of a
// workgroup adds into copy l % C.  Lanes of a wave then spread over C addresses (and
// banks) whatever cells they deposit into, so the atomics of a wave whose particles share
// a cell (the common case after a sort) are not serialised on one address.  More copies
// mean fewer conflicts but more LDS per workgroup: the FP64 push uses 16, the lighter
// mixed-precision push 8 (more workgroups per CU).
constexpr int kCopiesF64 = 16;
constexpr int kCopiesMixed = 8;

template <typename T, int C>
__device__ inline void window_clear(T *sR) {
    for (int q = threadIdx.x; q < kWinN * C; q += blockDim.x) sR[q] = T(0);
}

// base of the copy used by this lane
template <typename T, int C>
__device__ inline T *window_copy(T *sR) {
    return sR + (threadIdx.x % C);
}

// add the four node weights of the cell whose lower-left node has slot d
template <typename T, int C>
__device__ inline void window_deposit(T *mine, int d, T w0, T w1, T w2, T w3) {
    lds<T> *m = to_lds(mine);
    lds_add(m + d * C, w0);
    lds_add(m + (d + 1) * C, w1);
    lds_add(m + (d + kWinW) * C, w2);
    lds_add(m + (d + kWinW + 1) * C, w3);
}

// sum the copies of every node (in FP64) and add the window to the grid; the copy index is
// rotated by the node so the threads of a wave read distinct banks
template <typename T, int C>
__device__ inline void window_flush(const Window &w, const Geo &g, const T *sR, double *rho) {
    for (int q = threadIdx.x; q < kWinN; q += blockDim.x) {
        double v = 0.0;
#pragma unroll 8
        for (int c = 0; c < C; ++c) v += (double)sR[q * C + ((c + q) % C)];
        if (v != 0.0) {
            const int i = w.x0 + q % kWinW, j = w.y0 + q / kWinW;
            unsafeAtomicAdd(&rho[(int64_t)i + (int64_t)j * g.nx], v);
        }
    }
}

// ---------------------------------------------------------------------------------
// Moments accumulated by the push itself on snapshot steps (ADVANCE_ELECTRONS_AND_CALCULATE_
// MOMENTS_2D, pic2d_ElectronDynamicsWithVDFMoments.f90:156-1010; ADVANCE_IONS_AND_CALCULATE_
// MOMENTS_2D, pic2d_IonDynamicsWithVDFMoments.f90:180-1026): every particle, at its position
// before the step, adds with the half-step velocity 0.5 (v + v') rounded to FP32 the 13 sums
// N, VX, VY, VZ, VX^2, VY^2, VZ^2, VXVY, VXVZ, VYVZ, v^2 VX, v^2 VY, v^2 VZ times its FP32
// bilinear weights (the fourth is 1 minus the other three).  Accumulators: kMoments node
// arrays.

struct MomentTerms {
    float w[4];
    float m[kMoments];
};

__device__ inline MomentTerms moment_terms(float w0, float w1, float w2, double vx, double vy, double vz) {
    MomentTerms t;
    t.w[0] = w0;
    t.w[1] = w1;
    t.w[2] = w2;
    t.w[3] = 1.0f - w0 - w1 - w2;
    const float rvx = (float)vx, rvy = (float)vy, rvz = (float)vz;
    const float rvx2 = rvx * rvx, rvy2 = rvy * rvy, rvz2 = rvz * rvz;
    const float v2 = rvx2 + rvy2 + rvz2;
    t.m[0] = 1.0f;
    t.m[1] = rvx;
    t.m[2] = rvy;
    t.m[3] = rvz;
    t.m[4] = rvx2;
    t.m[5] = rvy2;
    t.m[6] = rvz2;
    t.m[7] = rvx * rvy;
    t.m[8] = rvx * rvz;
    t.m[9] = rvy * rvz;
    t.m[10] = v2 * rvx;
    t.m[11] = v2 * rvy;
    t.m[12] = v2 * rvz;
    return t;
}

__device__ inline float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Adds the terms of the wave's particles.  The cell whose lower-left node has window slot d
// (-1: its nodes are not all in the window; q0 its lower-left node in the grid): when all 64
// lanes push a particle of one window cell (the common case after a sort) each of the 52 sums
// is reduced over the wave and added by one lane; otherwise every lane adds into the window
// with LDS atomics, or into the grid when its cell is outside the window.
__device__ inline void moments_add(float *sM, int d, int64_t q0, const Geo &g, float *mom, const MomentTerms &t) {
    const int64_t nn = (int64_t)g.nx * g.ny;
    const unsigned long long all = __ballot(1);
    const int d0 = __shfl(d, 0);
    if (all == ~0ull && d0 >= 0 && __ballot(d == d0) == ~0ull) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int slot = d0 + (c & 1) + (c >> 1) * kWinW;
#pragma unroll
            for (int a = 0; a < kMoments; ++a) {
                const float v = wave_sum(a == 0 ? t.w[c] : t.m[a] * t.w[c]);
                if (__lane_id() == 0) lds_add(to_lds(sM) + a * kWinN + slot, v);
            }
        }
        return;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int off = (c & 1) + (c >> 1) * kWinW;
        const int64_t node = q0 + (c & 1) + (int64_t)(c >> 1) * g.nx;
#pragma unroll
        for (int a = 0; a < kMoments; ++a) {
            const float v = a == 0 ? t.w[c] : t.m[a] * t.w[c];
            if (d >= 0) lds_add(to_lds(sM) + a * kWinN + d + off, v);
            else atomicAdd(&mom[a * nn + node], v);
        }
    }
}

__device__ inline void moments_clear(float *sM) {
    for (int q = threadIdx.x; q < kMoments * kWinN; q += blockDim.x) sM[q] = 0.0f;
}

// adds the window's sums to the grid accumulators
__device__ inline void moments_flush(const Window &w, const Geo &g, const float *sM, float *mom) {
    const int64_t nn = (int64_t)g.nx * g.ny;
    for (int q = threadIdx.x; q < kMoments * kWinN; q += blockDim.x) {
        const float v = sM[q];
        if (v != 0.0f) {
            const int a = q / kWinN, r = q % kWinN;
            const int i = w.x0 + r % kWinW, j = w.y0 + r / kWinW;
            atomicAdd(&mom[a * nn + (int64_t)i + (int64_t)j * g.nx], v);
        }
    }
}

}  // namespace edgpu
