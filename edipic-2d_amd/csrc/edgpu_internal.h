// Internal declarations shared by the edgpu HIP sources (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "edgpu.h"

namespace edgpu {

// uniform scale of the reference RNG, y*(1+2^-32)*2^-32 (src/WELL19937a_new.c:52);
// Philox words are mapped to [0,1) the same way
constexpr double kUniformScale = 0x1.00000001p-32;

// physical constants, pic2d_Modules.f90:161-168
constexpr double kECl = 1.602189e-19;
constexpr double kMeKg = 9.109534e-31;
constexpr double kBJK = 1.38064852e-23;
constexpr double kPi = 3.141592653589793;

constexpr int kBlock = 256;

// Waits until every vector memory operation of the wave has completed (s_waitcnt
// vmcnt(0); gfx9 encoding: vmcnt 0, expcnt 7, lgkmcnt 15).
__device__ inline void wait_vector_memory() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// ---------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), the counter-based RNG.
__host__ __device__ inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                              uint32_t k1, uint32_t out[4]) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)M0 * c0;
        uint64_t p1 = (uint64_t)M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += W0;
        k1 += W1;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Philox2x32-10 (same paper): one 32x32->64-bit product per round and two output words,
// about half the arithmetic of Philox4x32-10.
__host__ __device__ inline void philox2x32_10(uint32_t c0, uint32_t c1, uint32_t k, uint32_t out[2]) {
    const uint32_t M = 0xD256D193u, W = 0x9E3779B9u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p = (uint64_t)M * c0;
        const uint32_t n0 = (uint32_t)(p >> 32) ^ k ^ c1;
        c1 = (uint32_t)p;
        c0 = n0;
        k += W;
    }
    out[0] = c0;
    out[1] = c1;
}

// Source of uniforms for the collision kernels: the Philox streams keyed by the context
// seed and counted by (slot, step, purpose, block), or an injected stream.
struct Rng {
    const double *stream;  // non-null: u = stream[slot*stride + k]
    int64_t stream_n;
    int32_t stride;
    uint32_t key0, key1;
    uint32_t key2;  // Philox2x32 key of (seed, purpose)
    uint32_t step;
    uint32_t purpose;
};

__device__ inline void draw4(const Rng &r, int64_t slot, int block, double *u) {
    if (r.stream) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int k = block * 4 + q;
            int64_t at = slot * r.stride + k;
            u[q] = (k < r.stride && at < r.stream_n) ? r.stream[at] : 0.5;
        }
        wait_vector_memory();  // see draw2
        return;
    }
    uint32_t w[4];
    philox4x32_10((uint32_t)slot, (uint32_t)((uint64_t)slot >> 32), r.step, (r.purpose << 8) | (uint32_t)block,
                  r.key0, r.key1, w);
#pragma unroll
    for (int q = 0; q < 4; ++q) u[q] = (double)w[q] * kUniformScale;
}

// The two uniforms every particle draws for its collision test (words 0 and 1 of its slot
// in an injected stream): Philox2x32-10 counted by (slot, step).  Slots are below 2^31.
__device__ inline void draw2(const Rng &r, int64_t slot, double *u) {
    if (r.stream) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            int64_t at = slot * r.stride + q;
            u[q] = (q < r.stride && at < r.stream_n) ? r.stream[at] : 0.5;
        }
        // the injected stream's loads complete inside this branch: otherwise the registers
        // they fill carry a pending load past the join, and the wait placed there (for every
        // particle of a push, stream or not) would also drain the next particle's loads
        wait_vector_memory();
        return;
    }
    uint32_t w[2];
    philox2x32_10((uint32_t)slot, r.step, r.key2, w);
    u[0] = (double)w[0] * kUniformScale;
    u[1] = (double)w[1] * kUniformScale;
}

// The four uniforms of an e-n collision test: the two of draw2, then words 2 and 3 of the
// slot's first Philox4x32 block (words 2 and 3 of an injected stream).
__device__ inline void draw_en(const Rng &r, int64_t slot, double *u0) {
    double b[4];
    draw4(r, slot, 0, b);
    draw2(r, slot, u0);
    u0[2] = b[2];
    u0[3] = b[3];
}

// Upper bound of vz^2 over the ions of a species, kept as the bits of a non-negative double
// (whose order is that of the unsigned integers) so atomicMax raises it.
__device__ inline double vz_bound(const unsigned long long *b) { return __longlong_as_double((long long)*b); }

__device__ inline void raise_vz_bound(unsigned long long *b, double vz) {
    atomicMax(b, (unsigned long long)__double_as_longlong(vz * vz));
}

// raises the bound to max vz^2 over [begin, end): one atomic per wave
template <typename T>
__global__ void k_vz_bound(const T *vz, int64_t begin, int64_t end, unsigned long long *bound) {
    double m = 0.0;
    for (int64_t k = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < end;
         k += (int64_t)gridDim.x * blockDim.x) {
        const double v = (double)vz[k];
        m = fmax(m, v * v);
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0 && m > 0.0) atomicMax(bound, (unsigned long long)__double_as_longlong(m));
}


// ---------------------------------------------------------------------------------
// geometry of the single cluster, passed by value
struct Geo {
    int32_t gmi, gmj, nx, ny;
    int32_t periodic_x;
    double xmax, ymax, Lx;
    int32_t n_seg;
    int32_t seg_side[EDGPU_MAX_SEG], seg_start[EDGPU_MAX_SEG], seg_end[EDGPU_MAX_SEG], seg_obj[EDGPU_MAX_SEG];
    int32_t n_boundary_obj, n_inner;
    int32_t box[EDGPU_MAX_INNER][4];
    uint32_t diel_mask;  // bit o: boundary object o is a dielectric
};

// per-call device counters
struct Counters {
    unsigned long long hit[EDGPU_MAX_OBJ];
    unsigned long long emit[EDGPU_MAX_OBJ];  // electrons emitted per object (SEE, constant emission)
    unsigned long long n_removed;
    unsigned long long coll[EDGPU_MAX_PROC];
    unsigned long long n_holes, n_movers;
    unsigned long long n_cand;
    unsigned long long n_sel;  // candidates whose process selection picked a process
    int err;
    int pad;
    long long err_index;
};

// Energy bins of the collision-probability bound: 4 per octave from 2^-21 eV to 2^20 eV,
// bin of en = m 2^e (m in [0.5, 1)) is 4 (e + 20) + floor(8 (m - 0.5)); -1 outside.
constexpr int kEnBins = 164;

__host__ __device__ inline int en_bin(double en) {
    if (!(en > 0.0)) return -1;
    int e;
    const double m = frexp(en, &e);
    const int b = 4 * (e + 20) + (int)((m - 0.5) * 8.0);
    return (b >= 0 && b < kEnBins) ? b : -1;
}

// electron-neutral collision table (INITIATE_ELECTRON_NEUTRAL_COLLISIONS), by value
struct EnTable {
    int32_t n_proc;
    int32_t type[EDGPU_MAX_PROC];
    double thr[EDGPU_MAX_PROC];
    double ivf[EDGPU_MAX_PROC];
    int32_t ion_sp[EDGPU_MAX_PROC];
    double M_amu;
    int32_t n_seg;
    double seg_value[33];
    double seg_step[33];
    int32_t seg_index[33];
    int32_t n_values;
    const double *energy;  // device [n_values+1]
    const double *prob;    // device [(n_values+1)*n_proc]
    double pmax;           // max_colliding_fraction
    double V2;             // V_scale_ms^2
    // upper bound of the total (cumulative) selection probability over each energy bin:
    // a candidate whose random_r exceeds it is certainly a null collision
    double bound[kEnBins];
};

// resonant charge exchange parameters for one ion species
struct RcxPar {
    int32_t on;
    double sigma_1eV, alpha, prob_factor, pth2, factor_eV, V_scale, vfactor;
    double c2;  // (prob_factor * V_scale * sigma_1eV)^2: p^2 = c2 |v|^2 + pth2 when alpha == 0
};

// add list (new particles from ionisation), SoA, persistent count in *count
struct AddList {
    double *x, *y, *vx, *vy, *vz;
    int32_t *tag;  // tag of the new particle (0 for collision products, emitting object + 1)
    unsigned long long *count;
    int64_t cap;
};

// the add list of every ion species: an ionising process appends its ion to the list of the
// species it produces (ion_species_produced, pic2d_enCollisionsSpecificProc.f90:629)
struct IonAddLists {
    AddList s[EDGPU_MAX_ION_SPECIES];
};

}  // namespace edgpu
