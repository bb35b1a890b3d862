// Per-cell counting sort of a species (key = cell i + j*gmi): histogram, exclusive scan,
// scatter into the spare arrays, and the tile ranges of the new order.
//
// The histogram and the scatter run over the tiles of the previous sort: a workgroup
// counts its particles in LDS copies of its cell window and adds the window to the
// global counts once; in the scatter it reserves one range per window cell with a single
// global atomic and places its particles inside the range with LDS atomics.  Particles
// outside the window (and untiled particles) use global atomics directly.  The order of
// the particles inside a cell depends on the order the atomics complete in.
#include <hipcub/hipcub.hpp>

#include "tile.h"

namespace edgpu {

struct KeyF64 {
    ParticlesF64 p;
    __device__ void cell(int64_t k, const Geo &g, int &i, int &j) const {
        const double x = p.x[k], y = p.y[k];
        i = (int)x;
        j = (int)y;
        if (x == g.xmax) i = g.gmi - 1;
        if (y == g.ymax) j = g.gmj - 1;
        i = max(0, min(i, g.gmi - 1));
        j = max(0, min(j, g.gmj - 1));
    }
    __device__ void move(int64_t from, int64_t to, const ParticlesF64 &d) const {
        d.x[to] = p.x[from];
        d.y[to] = p.y[from];
        d.vx[to] = p.vx[from];
        d.vy[to] = p.vy[from];
        d.vz[to] = p.vz[from];
        d.tag[to] = p.tag[from];
    }
};

struct KeyMixed {
    ParticlesMixed p;
    __device__ void cell(int64_t k, const Geo &g, int &i, int &j) const {
        const int32_t c = p.cell[k];
        i = c & 0xffff;
        j = c >> 16;
    }
    __device__ void move(int64_t from, int64_t to, const ParticlesMixed &d) const {
        d.cell[to] = p.cell[from];
        d.dx[to] = p.dx[from];
        d.dy[to] = p.dy[from];
        d.vx[to] = p.vx[from];
        d.vy[to] = p.vy[from];
        d.vz[to] = p.vz[from];
        d.tag[to] = p.tag[from];
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

// The LDS cell window is replicated kSortCopies times, copy-major with an odd row pitch:
// counter (copy c, window cell d) sits at c*kSortPitch + d and thread t uses copy t % 32.
// The lanes of a wave then hit 32 distinct banks whatever cells they count, so particles of
// one cell (the common case after a sort) do not serialise on one LDS address.
constexpr int kSortCopies = 32;
constexpr int kSortPitch = kWinCN + 1;
static_assert(kSortPitch % 2 == 1, "odd pitch spreads the copies over the banks");

// packed cell key: (j << 16) | i
__device__ inline int32_t pack_key(int i, int j) { return (j << 16) | i; }

template <typename KEY>
__global__ __launch_bounds__(kBlock) void k_histogram(KEY key, int64_t n, Geo g, Tiles tiles, int32_t *keys,
                                                      int32_t *count) {
    __shared__ int32_t sc[kSortCopies * kSortPitch];
    const int b = blockIdx.x;
    const bool tiled = b < tiles.ntiles;
    const int copy = threadIdx.x % kSortCopies;
    Window w{};
    if (tiled) {
        w = tile_window(tile_of_block(b, tiles.ntiles), tiles.ntx, g);
        for (int q = threadIdx.x; q < kSortCopies * kSortPitch; q += blockDim.x) sc[q] = 0;
        __syncthreads();
    }
    const Range r = range_of(b, tiles, n);
    for (int64_t k = r.begin + threadIdx.x; k < r.end; k += r.stride) {
        int i, j;
        key.cell(k, g, i, j);
        keys[k] = pack_key(i, j);
        const int d = tiled ? window_cell(w, i, j) : -1;
        if (d >= 0) atomicAdd(&sc[copy * kSortPitch + d], 1);
        else atomicAdd(&count[i + j * g.gmi], 1);
    }
    if (tiled) {
        __syncthreads();
        for (int q = threadIdx.x; q < kWinCN; q += blockDim.x) {
            int v = 0;
            for (int c = 0; c < kSortCopies; ++c) v += sc[c * kSortPitch + q];
            if (v) atomicAdd(&count[(w.x0 + q % kWinCW) + (w.y0 + q / kWinCW) * g.gmi], v);
        }
    }
}

// Placement: a tiled workgroup counts its particles per (copy, window cell) again, reserves
// one global range per window cell with a single atomic and splits it between the copies;
// every particle then takes the next slot of its (copy, cell) range.  Particles outside the
// window (and untiled particles) take a slot from the global cursor of their cell.
template <typename KEY, typename P>
__global__ __launch_bounds__(kBlock) void k_scatter(KEY key, P dst, int64_t n, Geo g, Tiles tiles,
                                                    const int32_t *keys, int32_t *cursor) {
    __shared__ int32_t sc[kSortCopies * kSortPitch];
    const int b = blockIdx.x;
    const bool tiled = b < tiles.ntiles;
    const int copy = threadIdx.x % kSortCopies;
    Window w{};
    const Range r = range_of(b, tiles, n);
    if (tiled) {
        w = tile_window(tile_of_block(b, tiles.ntiles), tiles.ntx, g);
        for (int q = threadIdx.x; q < kSortCopies * kSortPitch; q += blockDim.x) sc[q] = 0;
        __syncthreads();
        for (int64_t k = r.begin + threadIdx.x; k < r.end; k += r.stride) {
            const int32_t c = keys[k];
            const int d = window_cell(w, c & 0xffff, c >> 16);
            if (d >= 0) atomicAdd(&sc[copy * kSortPitch + d], 1);
        }
        __syncthreads();
        for (int q = threadIdx.x; q < kWinCN; q += blockDim.x) {
            int tot = 0;
            for (int c = 0; c < kSortCopies; ++c) tot += sc[c * kSortPitch + q];
            int at = tot ? atomicAdd(&cursor[(w.x0 + q % kWinCW) + (w.y0 + q / kWinCW) * g.gmi], tot) : 0;
            for (int c = 0; c < kSortCopies; ++c) {
                const int v = sc[c * kSortPitch + q];
                sc[c * kSortPitch + q] = at;
                at += v;
            }
        }
        __syncthreads();
    }
    for (int64_t k = r.begin + threadIdx.x; k < r.end; k += r.stride) {
        const int32_t c = keys[k];
        const int i = c & 0xffff, j = c >> 16;
        const int d = tiled ? window_cell(w, i, j) : -1;
        const int at = d >= 0 ? atomicAdd(&sc[copy * kSortPitch + d], 1) : atomicAdd(&cursor[i + j * g.gmi], 1);
        key.move(k, at, dst);
    }
}

// first particle of every tile of the new order: start[t] = offset of its first cell
__global__ void k_tile_starts(const int32_t *offset, int32_t *start, int ntx, int ntiles, int gmi, int32_t n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntiles) return;
    if (t == ntiles) {
        start[t] = n;
        return;
    }
    const int j = t / ntx, i0 = (t % ntx) * kTileW;
    start[t] = offset[i0 + j * gmi];
}

size_t sort_scan_bytes(int32_t ncell) {
    size_t bytes = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (int32_t *)nullptr, (int32_t *)nullptr, ncell + 1);
    return bytes;
}

int tiles_in_x(const Geo &g) { return (g.gmi + kTileW - 1) / kTileW; }

template <typename KEY, typename P>
static void sort_impl(KEY key, const P &dst, int64_t n, const Geo &g, SortWork &w, const Tiles &old_tiles,
                      int32_t *new_start, hipStream_t s) {
    const int32_t ncell = g.gmi * g.gmj;
    const int64_t untiled = n - std::min<int64_t>(old_tiles.covered, n);
    int tail = (int)std::min<int64_t>((untiled + kBlock - 1) / kBlock, 65536);
    if (old_tiles.ntiles == 0 && tail < 1) tail = 1;
    const int grid = old_tiles.ntiles + tail;
    hipMemsetAsync(w.count, 0, sizeof(int32_t) * (size_t)(ncell + 1), s);
    k_histogram<<<grid, kBlock, 0, s>>>(key, n, g, old_tiles, w.key, w.count);
    size_t bytes = w.scan_tmp_bytes;
    hipcub::DeviceScan::ExclusiveSum(w.scan_tmp, bytes, w.count, w.offset, ncell + 1, s);
    hipMemcpyAsync(w.count, w.offset, sizeof(int32_t) * (size_t)(ncell + 1), hipMemcpyDeviceToDevice, s);
    k_scatter<<<grid, kBlock, 0, s>>>(key, dst, n, g, old_tiles, w.key, w.count);
    const int ntx = tiles_in_x(g), ntiles = ntx * g.gmj;
    k_tile_starts<<<(ntiles + kBlock) / kBlock, kBlock, 0, s>>>(w.offset, new_start, ntx, ntiles, g.gmi, (int32_t)n);
}

void launch_sort_fp64(const ParticlesF64 &p, const ParticlesF64 &tmp, int64_t n, const Geo &g, SortWork &w,
                      const Tiles &old_tiles, int32_t *new_start, hipStream_t s) {
    if (n <= 0) return;
    sort_impl(KeyF64{p}, tmp, n, g, w, old_tiles, new_start, s);
}

void launch_sort_mixed(const ParticlesMixed &p, const ParticlesMixed &tmp, int64_t n, const Geo &g, SortWork &w,
                       const Tiles &old_tiles, int32_t *new_start, hipStream_t s) {
    if (n <= 0) return;
    sort_impl(KeyMixed{p}, tmp, n, g, w, old_tiles, new_start, s);
}

}  // namespace edgpu
