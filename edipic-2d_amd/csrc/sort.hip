// Per-tile counting sort of a species (key = tile: kTileW cells of one cell row, numbered
// tx + j*ntx): histogram, exclusive scan, scatter into the spare arrays; the scan's offsets
// are the tile ranges of the new order.  The push works tile by tile (its LDS windows cover a
// tile and its margins), so the order of the particles inside a tile does not matter to it.
//
// The histogram runs over the tiles of the previous sort: a workgroup counts its particles
// in LDS copies of a window of destination tiles around its own and adds the window to the
// global counts once.  The scatter stages each tile through LDS: the workgroup ranks its
// particles inside their destination tiles, reserves one global range per destination tile
// with a single atomic, and then, one particle array at a time, places the values in LDS in
// destination order and writes them out as contiguous runs (one run per destination tile:
// most of a tile's particles stay in it or its neighbours, so the runs are hundreds of
// particles long).  Particles outside the window and untiled particles take a slot from the
// global cursor of their tile directly.  The order inside a tile depends on the order the
// atomics complete in.
#include <hipcub/hipcub.hpp>

#include "sortwin.h"

namespace edgpu {

// particles per thread of one staged chunk of the scatter
constexpr int kStagePer = 16;
constexpr int kStageMax = kStagePer * kBlock;

template <typename T>
__device__ inline void load_chunk(const T *src, int64_t c0, int64_t c1, T *v) {
#pragma unroll
    for (int u = 0; u < kStagePer; ++u) v[u] = src[min(c0 + threadIdx.x + (int64_t)u * kBlock, c1 - 1)];
}

// values already in registers: through LDS in destination order and out as contiguous runs
template <typename T>
__device__ inline void place_chunk(const T *v, T *out, const int *slot, const int *gslot, void *stage_raw,
                                   const int32_t *dest, int m_in) {
    T *st = static_cast<T *>(stage_raw);
#pragma unroll
    for (int u = 0; u < kStagePer; ++u) {
        if (slot[u] >= 0) st[slot[u]] = v[u];
        else if (gslot[u] >= 0) out[gslot[u]] = v[u];
    }
    lds_barrier();
    for (int q = threadIdx.x; q < m_in; q += kBlock) out[dest[q]] = st[q];
    lds_barrier();
}

// Arrays of one element type, one after another, software-pipelined: the loads of array a + 1
// are issued before array a goes through LDS, so every thread keeps two arrays' loads in
// flight across the barriers of the placement.
template <typename T, int NA>
__device__ inline void stage_arrays(const T *const *src, T *const *dst, const int *slot, const int *gslot, int64_t c0,
                                    int64_t c1, void *stage_raw, const int32_t *dest, int m_in) {
    T v[2][kStagePer];
    load_chunk(src[0], c0, c1, v[0]);
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        if (a + 1 < NA) load_chunk(src[a + 1], c0, c1, v[(a + 1) & 1]);
        place_chunk(v[a & 1], dst[a], slot, gslot, stage_raw, dest, m_in);
    }
}

struct KeyF64 {
    ParticlesF64 p;
    __device__ void tile(int64_t k, const Geo &g, int &tx, int &j) const { sort_tile_f64(p.x[k], p.y[k], g, tx, j); }
    __device__ void move(int64_t from, int64_t to, const ParticlesF64 &d) const {
        d.x[to] = p.x[from];
        d.y[to] = p.y[from];
        d.vx[to] = p.vx[from];
        d.vy[to] = p.vy[from];
        d.vz[to] = p.vz[from];
        d.tag[to] = p.tag[from];
    }
    // every particle array through the staged placement: the five FP64 arrays pipelined, then the tag
    __device__ void stage(const ParticlesF64 &d, const int *slot, const int *gslot, int64_t c0, int64_t c1,
                          void *st, const int32_t *dest, int m) const {
        const double *src[5] = {p.x, p.y, p.vx, p.vy, p.vz};
        double *const dst[5] = {d.x, d.y, d.vx, d.vy, d.vz};
        stage_arrays<double, 5>(src, dst, slot, gslot, c0, c1, st, dest, m);
        const int32_t *ts[1] = {p.tag};
        int32_t *const td[1] = {d.tag};
        stage_arrays<int32_t, 1>(ts, td, slot, gslot, c0, c1, st, dest, m);
    }
};

struct KeyMixed {
    ParticlesMixed p;
    __device__ void tile(int64_t k, const Geo &g, int &tx, int &j) const { sort_tile_mixed(p.cell[k], g, tx, j); }
    __device__ void move(int64_t from, int64_t to, const ParticlesMixed &d) const {
        d.cell[to] = p.cell[from];
        d.dx[to] = p.dx[from];
        d.dy[to] = p.dy[from];
        d.vx[to] = p.vx[from];
        d.vy[to] = p.vy[from];
        d.vz[to] = p.vz[from];
        d.tag[to] = p.tag[from];
    }
    // every particle array through the staged placement, all seven as 32-bit words in one pipeline
    __device__ void stage(const ParticlesMixed &d, const int *slot, const int *gslot, int64_t c0, int64_t c1,
                          void *st, const int32_t *dest, int m) const {
        const uint32_t *src[7] = {(const uint32_t *)p.cell, (const uint32_t *)p.dx, (const uint32_t *)p.dy,
                                  (const uint32_t *)p.vx,   (const uint32_t *)p.vy, (const uint32_t *)p.vz,
                                  (const uint32_t *)p.tag};
        uint32_t *const dst[7] = {(uint32_t *)d.cell, (uint32_t *)d.dx, (uint32_t *)d.dy, (uint32_t *)d.vx,
                                  (uint32_t *)d.vy,   (uint32_t *)d.vz, (uint32_t *)d.tag};
        stage_arrays<uint32_t, 7>(src, dst, slot, gslot, c0, c1, st, dest, m);
    }
};

// particle range of workgroup b: a tile of the previous order, or a tail stride.  `begin`
// is the first index handled by thread 0; every thread of the workgroup runs the same
// number of iterations (index begin + threadIdx.x + t*stride, active while < end).
struct Range {
    int64_t begin, end, stride;
};

__device__ inline Range range_of(int b, const Tiles &t, int64_t n) {
    Range r;
    if (b < t.ntiles) {
        const int tb = tile_of_block(b, t.ntiles);
        r.begin = t.start[tb];
        r.end = min((int64_t)t.start[tb + 1], n);
        r.stride = blockDim.x;
    } else {
        r.begin = min(t.covered, n) + (int64_t)(b - t.ntiles) * blockDim.x;
        r.end = n;
        r.stride = (int64_t)(gridDim.x - t.ntiles) * blockDim.x;
    }
    return r;
}

// The histogram's LDS window is replicated kSortCopies times, copy-major with an odd pitch:
// counter (copy c, bin d) sits at c*kSortPitch + d and thread t uses copy t % 32.  The lanes of
// a wave then hit 32 distinct banks whatever tiles they count, so the particles of one tile
// (nearly all of a wave's) do not serialise on one LDS address.
constexpr int kSortCopies = 32;
constexpr int kSortPitch = kSortWinN;
static_assert(kSortPitch % 2 == 1, "odd pitch spreads the copies over the banks");

// packed tile key: (j << 16) | tx; the tile's index is tx + j * ntx
__device__ inline int32_t pack_key(int tx, int j) { return (j << 16) | tx; }

// keys (may be null): the packed tile of every particle, for the scatter; wcount (may be null):
// the window counts of every source tile, kSortWinN per tile, for the sorting push
template <typename KEY>
__global__ __launch_bounds__(kBlock) void k_histogram(KEY key, int64_t n, Geo g, Tiles tiles, int32_t *keys,
                                                      int32_t *count, int32_t *wcount) {
    __shared__ int32_t sc[kSortCopies * kSortPitch];
    const int b = blockIdx.x;
    const bool tiled = b < tiles.ntiles;
    const int copy = threadIdx.x % kSortCopies;
    const int ntx = tiles.ntx;
    SortWindow w{};
    const int tb = tiled ? tile_of_block(b, tiles.ntiles) : 0;
    if (tiled) {
        w = sort_window(tb, ntx, g);
        for (int q = threadIdx.x; q < kSortCopies * kSortPitch; q += blockDim.x) sc[q] = 0;
        __syncthreads();
    }
    const Range r = range_of(b, tiles, n);
    for (int64_t k = r.begin + threadIdx.x; k < r.end; k += r.stride) {
        int tx, j;
        key.tile(k, g, tx, j);
        if (keys) keys[k] = pack_key(tx, j);
        const int d = tiled ? sort_bin(w, tx, j) : -1;
        if (d >= 0) atomicAdd(&sc[copy * kSortPitch + d], 1);
        else atomicAdd(&count[tx + j * ntx], 1);
    }
    if (tiled) {
        __syncthreads();
        for (int q = threadIdx.x; q < kSortWinN; q += blockDim.x) {
            int v = 0;
            for (int c = 0; c < kSortCopies; ++c) v += sc[c * kSortPitch + q];
            if (v) atomicAdd(&count[(w.x0 + q % kSortWinW) + (w.y0 + q / kSortWinW) * ntx], v);
            if (wcount) wcount[(int64_t)tb * kSortWinN + q] = v;
        }
    }
}

// Staged placement.  A tiled workgroup handles its tile in chunks of up to kStageMax
// particles (kStagePer per thread).  Per chunk: every particle inside the window takes the
// next rank of its destination tile (LDS atomic); an exclusive scan of the window counts gives
// each destination's run in stage order and one global atomic per destination reserves the
// run; then every particle array goes through LDS in stage order and out to the runs.  A
// particle outside the window takes a slot from the global cursor of its tile and is written
// straight from registers.  Untiled workgroups use the global cursors only.
// The loads of a thread's kStagePer particles are unconditional (the index is clamped into
// the chunk), so they are issued back to back and their latencies overlap.

template <typename KEY, typename P>
__global__ __launch_bounds__(kBlock) void k_scatter(KEY key, P dst, int64_t n, Geo g, Tiles tiles,
                                                    const int32_t *keys, int32_t *cursor) {
    __shared__ SortStage<kStageMax> ss;
    __shared__ double stage[kStageMax];
    const int b = blockIdx.x;
    const int ntx = tiles.ntx;
    const Range r = range_of(b, tiles, n);
    if (b >= tiles.ntiles) {
        for (int64_t k = r.begin + threadIdx.x; k < r.end; k += r.stride) {
            const int32_t c = keys[k];
            key.move(k, atomicAdd(&cursor[(c & 0xffff) + (c >> 16) * ntx], 1), dst);
        }
        return;
    }
    const SortWindow w = sort_window(tile_of_block(b, tiles.ntiles), ntx, g);
    for (int64_t c0 = r.begin; c0 < r.end; c0 += kStageMax) {
        const int64_t c1 = min(c0 + (int64_t)kStageMax, r.end);
        for (int q = threadIdx.x; q < kSortWinN; q += kBlock) ss.cnt[q] = 0;
        int32_t kc[kStagePer];
#pragma unroll
        for (int u = 0; u < kStagePer; ++u) kc[u] = keys[min(c0 + threadIdx.x + (int64_t)u * kBlock, c1 - 1)];
        // all keys in registers before the ranking: the global atomics below are then the only
        // vector memory operations in flight and nothing waits on them until they are used
        wait_vector_memory();
        __syncthreads();
        // slot: rank in the destination tile (then stage slot), or -1; gslot: destination of
        // a particle outside the window, or -1; cq: window bin, or -1
        int slot[kStagePer], gslot[kStagePer], cq[kStagePer];
#pragma unroll
        for (int u = 0; u < kStagePer; ++u) {
            slot[u] = -1;
            gslot[u] = -1;
            cq[u] = -1;
            if (c0 + threadIdx.x + (int64_t)u * kBlock < c1) {
                const int tx = kc[u] & 0xffff, j = kc[u] >> 16;
                const int d = sort_bin(w, tx, j);
                if (d >= 0) {
                    cq[u] = d;
                    slot[u] = atomicAdd(&ss.cnt[d], 1);
                } else {
                    gslot[u] = atomicAdd(&cursor[tx + j * ntx], 1);
                }
            }
        }
        __syncthreads();
        scan_and_reserve(w, ntx, ss, cursor);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kStagePer; ++u) {
            if (cq[u] >= 0) {
                const int s = ss.loc[cq[u]] + slot[u];
                ss.dest[s] = ss.base[cq[u]] + slot[u];
                slot[u] = s;
            }
        }
        const int m = ss.m_in;
        key.stage(dst, slot, gslot, c0, c1, stage, ss.dest, m);
    }
}

size_t sort_scan_bytes(int32_t ncell) {
    size_t bytes = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (int32_t *)nullptr, (int32_t *)nullptr, ncell + 1) !=
        hipSuccess)
        return 0;
    return bytes;
}

int tiles_in_x(const Geo &g) { return (g.gmi + kTileW - 1) / kTileW; }

int64_t sort_window_counts(const Geo &g) { return (int64_t)tiles_in_x(g) * g.gmj * kSortWinN; }

// returns the first failing HIP call's status (kernel launch errors surface in hipGetLastError)
// The keys are tiles: ntiles = ntx * gmj counters (SortWork holds gmi * gmj + 1).  The
// exclusive scan of the counts over ntiles + 1 entries (the last one 0) is the new tile
// ranges itself: start[t] = offset[t], start[ntiles] = n.
template <typename KEY, typename P>
static hipError_t sort_impl(KEY key, const P &dst, int64_t n, const Geo &g, SortWork &w, const Tiles &old_tiles,
                            int32_t *new_start, hipStream_t s) {
    const int32_t ntiles = tiles_in_x(g) * g.gmj;
    Tiles t = old_tiles;
    t.ntx = tiles_in_x(g);
    const int64_t untiled = n - std::min<int64_t>(t.covered, n);
    int tail = (int)std::min<int64_t>((untiled + kBlock - 1) / kBlock, 65536);
    if (t.ntiles == 0 && tail < 1) tail = 1;
    const int grid = t.ntiles + tail;
    hipError_t e = hipMemsetAsync(w.count, 0, sizeof(int32_t) * (size_t)(ntiles + 1), s);
    if (e != hipSuccess) return e;
    k_histogram<<<grid, kBlock, 0, s>>>(key, n, g, t, w.key, w.count, (int32_t *)nullptr);
    size_t bytes = w.scan_tmp_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(w.scan_tmp, bytes, w.count, w.offset, ntiles + 1, s);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(w.count, w.offset, sizeof(int32_t) * (size_t)(ntiles + 1), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    k_scatter<<<grid, kBlock, 0, s>>>(key, dst, n, g, t, w.key, w.count);
    return hipMemcpyAsync(new_start, w.offset, sizeof(int32_t) * (size_t)(ntiles + 1), hipMemcpyDeviceToDevice, s);
}

// The counts of a sorting push: histogram by tile (no key array; the window counts of every
// source tile into SortWork::wcount), scan, the scanned offsets copied to the cursors
// (SortWork::count) and to the new tile ranges.  The push then places every particle of
// [0, n): a tile's workgroup reserves one run per destination tile of its window from the
// cursors, the other particles take single slots.
template <typename KEY>
static hipError_t tile_counts_impl(KEY key, int64_t n, const Geo &g, SortWork &w, const Tiles &old_tiles,
                                   int32_t *new_start, hipStream_t s) {
    const int32_t ntiles = tiles_in_x(g) * g.gmj;
    Tiles t = old_tiles;
    t.ntx = tiles_in_x(g);
    const int64_t untiled = n - std::min<int64_t>(t.covered, n);
    int tail = (int)std::min<int64_t>((untiled + kBlock - 1) / kBlock, 65536);
    if (t.ntiles == 0 && tail < 1) tail = 1;
    hipError_t e = hipMemsetAsync(w.count, 0, sizeof(int32_t) * (size_t)(ntiles + 1), s);
    if (e != hipSuccess) return e;
    k_histogram<<<t.ntiles + tail, kBlock, 0, s>>>(key, n, g, t, (int32_t *)nullptr, w.count, w.wcount);
    size_t bytes = w.scan_tmp_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(w.scan_tmp, bytes, w.count, w.offset, ntiles + 1, s);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(w.count, w.offset, sizeof(int32_t) * (size_t)(ntiles + 1), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(new_start, w.offset, sizeof(int32_t) * (size_t)(ntiles + 1), hipMemcpyDeviceToDevice, s);
}

hipError_t launch_tile_counts_fp64(const ParticlesF64 &p, int64_t n, const Geo &g, SortWork &w,
                                   const Tiles &old_tiles, int32_t *new_start, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    return tile_counts_impl(KeyF64{p}, n, g, w, old_tiles, new_start, s);
}

hipError_t launch_tile_counts_mixed(const ParticlesMixed &p, int64_t n, const Geo &g, SortWork &w,
                                    const Tiles &old_tiles, int32_t *new_start, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    return tile_counts_impl(KeyMixed{p}, n, g, w, old_tiles, new_start, s);
}

hipError_t launch_sort_fp64(const ParticlesF64 &p, const ParticlesF64 &tmp, int64_t n, const Geo &g, SortWork &w,
                            const Tiles &old_tiles, int32_t *new_start, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    return sort_impl(KeyF64{p}, tmp, n, g, w, old_tiles, new_start, s);
}

hipError_t launch_sort_mixed(const ParticlesMixed &p, const ParticlesMixed &tmp, int64_t n, const Geo &g,
                             SortWork &w, const Tiles &old_tiles, int32_t *new_start, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    return sort_impl(KeyMixed{p}, tmp, n, g, w, old_tiles, new_start, s);
}

}  // namespace edgpu
