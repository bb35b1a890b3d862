// Per-cell counting sort of a species (cell index i + j*gmi as key): count, exclusive scan,
// scatter into the spare arrays.  Keeps the particles of one cell contiguous so that the
// gather reads and the deposit writes of consecutive threads touch neighbouring nodes.
#include <hipcub/hipcub.hpp>

#include "kernels.h"

namespace edgpu {

__global__ void k_key_fp64(ParticlesF64 p, int64_t n, Geo g, int32_t *key, int32_t *count) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const double x = p.x[k], y = p.y[k];
        int i = (int)x, j = (int)y;
        if (x == g.xmax) i = g.gmi - 1;
        if (y == g.ymax) j = g.gmj - 1;
        i = max(0, min(i, g.gmi - 1));
        j = max(0, min(j, g.gmj - 1));
        const int32_t c = i + j * g.gmi;
        key[k] = c;
        atomicAdd(&count[c], 1);
    }
}

__global__ void k_key_mixed(ParticlesMixed p, int64_t n, Geo g, int32_t *key, int32_t *count) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const int32_t pc = p.cell[k];
        const int32_t c = (pc & 0xffff) + (pc >> 16) * g.gmi;
        key[k] = c;
        atomicAdd(&count[c], 1);
    }
}

__global__ void k_scatter_fp64(ParticlesF64 s, ParticlesF64 d, int64_t n, const int32_t *key, int32_t *offset) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const int32_t at = atomicAdd(&offset[key[k]], 1);
        d.x[at] = s.x[k];
        d.y[at] = s.y[k];
        d.vx[at] = s.vx[k];
        d.vy[at] = s.vy[k];
        d.vz[at] = s.vz[k];
        d.tag[at] = s.tag[k];
    }
}

__global__ void k_scatter_mixed(ParticlesMixed s, ParticlesMixed d, int64_t n, const int32_t *key, int32_t *offset) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const int32_t at = atomicAdd(&offset[key[k]], 1);
        d.cell[at] = s.cell[k];
        d.dx[at] = s.dx[k];
        d.dy[at] = s.dy[k];
        d.vx[at] = s.vx[k];
        d.vy[at] = s.vy[k];
        d.vz[at] = s.vz[k];
        d.tag[at] = s.tag[k];
    }
}

size_t sort_scan_bytes(int32_t ncell) {
    size_t bytes = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (int32_t *)nullptr, (int32_t *)nullptr, ncell + 1);
    return bytes;
}

static int grid_s(int64_t n) {
    int64_t b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    if (b > 65536) b = 65536;
    return (int)b;
}

static void scan_counts(SortWork &w, int32_t ncell, hipStream_t s) {
    size_t bytes = w.scan_tmp_bytes;
    hipcub::DeviceScan::ExclusiveSum(w.scan_tmp, bytes, w.count, w.offset, ncell + 1, s);
}

void launch_sort_fp64(const ParticlesF64 &p, const ParticlesF64 &tmp, int64_t n, const Geo &g, SortWork w,
                      hipStream_t s) {
    if (n <= 0) return;
    const int32_t ncell = g.gmi * g.gmj;
    hipMemsetAsync(w.count, 0, sizeof(int32_t) * (size_t)(ncell + 1), s);
    k_key_fp64<<<grid_s(n), kBlock, 0, s>>>(p, n, g, w.key, w.count);
    scan_counts(w, ncell, s);
    k_scatter_fp64<<<grid_s(n), kBlock, 0, s>>>(p, tmp, n, w.key, w.offset);
}

void launch_sort_mixed(const ParticlesMixed &p, const ParticlesMixed &tmp, int64_t n, const Geo &g, SortWork w,
                       hipStream_t s) {
    if (n <= 0) return;
    const int32_t ncell = g.gmi * g.gmj;
    hipMemsetAsync(w.count, 0, sizeof(int32_t) * (size_t)(ncell + 1), s);
    k_key_mixed<<<grid_s(n), kBlock, 0, s>>>(p, n, g, w.key, w.count);
    scan_counts(w, ncell, s);
    k_scatter_mixed<<<grid_s(n), kBlock, 0, s>>>(p, tmp, n, w.key, w.offset);
}

}  // namespace edgpu
