// Destination windows of the counting sort by tile, shared by the sort kernels (sort.hip) and
// the sorting pushes, which write the particles of a sort step in tile order themselves
// (particles_fp64.hip, particles_mixed.hip).
//
// A source tile's particles, sorted a few steps of thermal drift ago, land in its own tile
// or a neighbour: the window of a source tile is its column and the two next to it, and
// kSortMarginY rows above and below.  A workgroup places the particles of a chunk through
// LDS: it ranks them inside their destination tiles, gives each destination a run of slots and
// writes each particle array out as contiguous runs; particles outside the window take single
// slots from their tile's global cursor.
#pragma once

#include "tile.h"

namespace edgpu {

constexpr int kSortMarginY = 8;
constexpr int kSortWinW = 3;
constexpr int kSortWinH = 2 * kSortMarginY + 1;
constexpr int kSortWinN = kSortWinW * kSortWinH;

struct SortWindow {
    int x0, y0, x1, y1;  // tile columns [x0, x1) x rows [y0, y1), clamped to the grid
};

__device__ inline SortWindow sort_window(int t, int ntx, const Geo &g) {
    const int j = t / ntx, tx = t % ntx;
    return SortWindow{max(0, tx - 1), max(0, j - kSortMarginY), min(ntx, tx + 2), min(g.gmj, j + 1 + kSortMarginY)};
}

__device__ inline int sort_bin(const SortWindow &w, int tx, int j) {
    if (tx < w.x0 || tx >= w.x1 || j < w.y0 || j >= w.y1) return -1;
    return (tx - w.x0) + (j - w.y0) * kSortWinW;
}

// The tile of a particle, from its FP64 position or its packed mixed-precision cell, clamped
// into the grid (x == xmax belongs to the last cell).  The histogram and the placement of one
// sort must give every particle the same tile: both call these.
__device__ inline void sort_tile_f64(double x, double y, const Geo &g, int &tx, int &j) {
    int i = (int)x;
    j = (int)y;
    if (x == g.xmax) i = g.gmi - 1;
    if (y == g.ymax) j = g.gmj - 1;
    i = max(0, min(i, g.gmi - 1));
    j = max(0, min(j, g.gmj - 1));
    tx = i / kTileW;
}

__device__ inline void sort_tile_mixed(int32_t c, const Geo &g, int &tx, int &j) {
    const int i = max(0, min(c & 0xffff, g.gmi - 1));
    j = max(0, min(c >> 16, g.gmj - 1));
    tx = i / kTileW;
}

// Workgroup barrier that orders LDS only: global stores and loads still in flight are not
// waited for (nothing in the workgroup reads the placement's global stores), so loads issued
// before it stay in flight across it.
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// LDS of one chunk's placement
template <int CHUNK>
struct SortStage {
    int32_t cnt[kSortWinN], loc[kSortWinN], base[kSortWinN];
    int32_t dest[CHUNK];  // destination index of each stage slot
    int32_t m_in;         // particles of the chunk inside the window
};

// Exclusive scan of the window's counts into stage offsets (wave 0, total into m_in) and one
// global reservation per non-empty destination tile.  Called between two barriers, after
// every particle of the chunk has taken its rank in cnt.
template <int CHUNK>
__device__ inline void scan_and_reserve(const SortWindow &w, int ntx, SortStage<CHUNK> &ss, int32_t *cursor) {
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        int carry = 0;
        for (int q0 = 0; q0 < kSortWinN; q0 += 64) {
            const int q = q0 + lane;
            const int v = q < kSortWinN ? ss.cnt[q] : 0;
            int x = v;
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(x, o);
                if (lane >= o) x += y;
            }
            if (q < kSortWinN) ss.loc[q] = carry + x - v;
            carry += __shfl(x, 63);
        }
        if (lane == 0) ss.m_in = carry;
    }
    for (int q = threadIdx.x; q < kSortWinN; q += blockDim.x) {
        const int v = ss.cnt[q];
        ss.base[q] = v ? atomicAdd(&cursor[(w.x0 + q % kSortWinW) + (w.y0 + q / kSortWinW) * ntx], v) : 0;
    }
}

// The placement of one chunk of a sorting push, in LDS: the window counts of the chunk, their
// exclusive scan (stage offsets), the destination of every stage slot, and the next free slot
// of the workgroup's run in each destination tile of the window (reserved once per tile).
template <int CHUNK>
struct PushStage {
    int32_t runs[kSortWinN], cnt[kSortWinN], loc[kSortWinN];
    int32_t dest[CHUNK];
    int32_t m_in;
};

// Exclusive scan of cnt into loc by wave 0 (total into m_in); called between two barriers.
template <int CHUNK>
__device__ inline void scan_window_counts(PushStage<CHUNK> &ps) {
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        int carry = 0;
        for (int q0 = 0; q0 < kSortWinN; q0 += 64) {
            const int q = q0 + lane;
            const int v = q < kSortWinN ? ps.cnt[q] : 0;
            int x = v;
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(x, o);
                if (lane >= o) x += y;
            }
            if (q < kSortWinN) ps.loc[q] = carry + x - v;
            carry += __shfl(x, 63);
        }
        if (lane == 0) ps.m_in = carry;
    }
}

// One particle array of a chunk out through LDS: the values (P per thread) go to their stage
// slots, or straight to their destination when outside the window (slot < 0 <= dst), then the
// workgroup writes the stage out in slot order, which is destination order inside each run.
// Destinations outside [0, n) are not written (a count that disagreed with the placement is
// reported by the placement).  The barriers order LDS only: loads of the next chunk stay in
// flight across them.
template <int P, typename T>
__device__ inline void place_array(const T *v, T *out, const int *slot, const int32_t *dst, T *stage,
                                   const int32_t *dest, int m_in, int64_t n) {
#pragma unroll
    for (int u = 0; u < P; ++u) {
        if (slot[u] >= 0) stage[slot[u]] = v[u];
        else if (dst[u] >= 0 && dst[u] < n) out[dst[u]] = v[u];
    }
    lds_barrier();
    for (int q = threadIdx.x; q < m_in; q += blockDim.x) {
        const int32_t d = dest[q];
        if (d >= 0 && d < n) out[d] = stage[q];
    }
    lds_barrier();
}

}  // namespace edgpu
