// Learned warm start for the Krylov potential solve: a 6-layer, 32-channel, 3x3 BF16
// convolutional network on MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulation) with
// hand-written forward, backward and Adam kernels.  The reference has no counterpart: it
// cold-starts every KSPSolve.
//
// Activations are BF16, pixel-major with the channels contiguous (NHWC, batch 1) over the
// (gmi+1) x (gmj+1) node grid, zero padding outside.  Weights of layer l are stored
// [tap][cout][cin] (tap = 3*(dy+1) + (dx+1)), FP32 master copy plus a BF16 copy for the
// kernels.  Layer 0 reads 16 channels (4 features + 12 zeros), layers 1..4 map 32 -> 32
// with ReLU, layer 5 maps 32 -> 16 (only channel 0 is the output, the others have zero
// weights and are never trained).
//
// MFMA lane layout used (checked on the device by edgpu_mfma_probe): for D = A(16x32) B(32x16),
// lane l supplies A[l%16][8*(l/16) + i] and B[8*(l/16) + i][l%16] (i = 0..7) and receives
// D[4*(l/16) + i][l%16].
#include "cnn.h"

namespace edgpu {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// lane l supplies A[l%16][8*(l/16) + i] and B[8*(l/16) + i][l%16] (i = 0..7) and receives
// D[4*(l/16) + i][l%16]
__device__ inline f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// eight consecutive bf16 (16 bytes, 16-byte aligned)
__device__ inline bf16x8 load8(const void *p) { return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(p)); }


// ---------------------------------------------------------------------------------
// layout probe of the K = 32 form the network uses: D (16x16) = A (16x32) B (32x16) for FP32
// inputs rounded to BF16
__global__ void k_mfma_probe(const float *A, const float *B, float *D) {
    const int l = threadIdx.x, m = l & 15, kq = l >> 4;
    uint16_t a[8], b[8];
    for (int i = 0; i < 8; ++i) {
        a[i] = f2bf(A[m * 32 + kq * 8 + i]);
        b[i] = f2bf(B[(kq * 8 + i) * 16 + m]);
    }
    f32x4 d = {0.f, 0.f, 0.f, 0.f};
    d = mfma32(load8(a), load8(b), d);
    for (int i = 0; i < 4; ++i) D[(kq * 4 + i) * 16 + m] = d[i];
}

void launch_mfma_probe(const float *A, const float *B, float *D, hipStream_t s) {
    k_mfma_probe<<<1, 64, 0, s>>>(A, B, D);
}

// ---------------------------------------------------------------------------------
// Convolution as implicit GEMM over a tile staged in LDS: a workgroup computes kConvTH rows
// of kConvTW pixels for all output channels, one wave per row, 16 pixels at a time; K runs
// over 9 taps x CIN channels in steps of 32 (ConvK).  The tile's input pixels with a
// one-pixel halo (zero outside the grid) are copied to LDS once with 16-byte loads, so an
// input pixel is fetched (TH+2)(TW+2)/(TH TW) ~ 1.5 times instead of once per tap; an A
// fragment is one 16-byte LDS read, and the wave's weight fragments (STEPS x COUT/16 MFMA B
// operands) stay in registers for its whole row.
//   MODE 0: forward, out = act(conv(in) + bias) stored BF16 (RELU) or, for the last
//           layer, channel 0 stored FP32 into outf.
//   MODE 1: data gradient: in holds dL/d(out) of the layer, w the flipped/transposed
//           weights; out = conv(in) * (gate > 0) stored BF16 (gate = the layer's input
//           activation, whose ReLU the gradient passes back through).
constexpr int kConvTW = 64, kConvTH = 4;

template <int CIN>
struct ConvTile {
    static constexpr int PX = CIN * 2 + 16;  // bytes per staged pixel; the pad staggers the LDS banks
    static constexpr int TWH = kConvTW + 2, THH = kConvTH + 2;
    static constexpr int CH16 = CIN * 2 / 16;  // 16-byte pieces of one pixel
    static constexpr int BYTES = THH * TWH * PX;
};

// the tile's input pixels with a one-pixel halo, zero outside the grid, into LDS
template <int CIN>
__device__ inline void conv_stage(const uint16_t *__restrict__ in, unsigned char *tile, int x0, int y0, int H, int W) {
    using T = ConvTile<CIN>;
    for (int q = threadIdx.x; q < T::THH * T::TWH * T::CH16; q += blockDim.x) {
        const int c = q % T::CH16, p = q / T::CH16;
        const int gx = x0 - 1 + p % T::TWH, gy = y0 - 1 + p / T::TWH;
        uint4 v = {0u, 0u, 0u, 0u};
        if (gx >= 0 && gx < W && gy >= 0 && gy < H)
            v = *reinterpret_cast<const uint4 *>(in + ((int64_t)gy * W + gx) * CIN + c * 8);
        *reinterpret_cast<uint4 *>(tile + p * T::PX + c * 16) = v;
    }
}

// The K = 32 steps of a 3x3 convolution: K runs over (tap, input channel); CIN / 8 lane groups
// of 8 channels cover one tap, so a step takes 32 / CIN taps (CIN = 32: one; CIN = 16: two,
// the ninth tap paired with nothing, its fragments zero).
template <int CIN>
struct ConvK {
    static_assert(CIN == 16 || CIN == 32, "3x3 convolutions of 16 or 32 input channels");
    static constexpr int TPS = 32 / CIN;  // taps per step
    static constexpr int STEPS = (9 + TPS - 1) / TPS;
    static constexpr int LPT = CIN / 8;  // 8-channel lane groups per tap
    // tap (9: none) and first channel of the fragment of lane group kq in step st
    __device__ static int tap(int st, int kq) { return st * TPS + kq / LPT; }
    __device__ static int chan(int kq) { return (kq % LPT) * 8; }
};

// the wave's weight fragments: STEPS x COUT/16 MFMA B operands
template <int CIN, int COUT>
__device__ inline void conv_weights(const uint16_t *__restrict__ w, bf16x8 (&wr)[ConvK<CIN>::STEPS][COUT / 16]) {
    using K = ConvK<CIN>;
    const int lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int st = 0; st < K::STEPS; ++st) {
        const int tap = K::tap(st, kq), ch = K::chan(kq);
#pragma unroll
        for (int nb = 0; nb < COUT / 16; ++nb)
            wr[st][nb] = tap < 9 ? load8(w + ((int64_t)(tap * COUT + nb * 16 + m) * CIN + ch)) : bf16x8{};
    }
}

// one wave's row of the staged tile: y = y0 + row, kConvTW pixels from x0, 16 at a time
template <int CIN, int COUT, int MODE, bool RELU>
__device__ inline void conv_row(const unsigned char *tile, const bf16x8 (&wr)[ConvK<CIN>::STEPS][COUT / 16],
                                const float *__restrict__ bias, const uint16_t *__restrict__ gate,
                                uint16_t *__restrict__ out, float *__restrict__ outf, int x0, int y, int row, int W) {
    using T = ConvTile<CIN>;
    const int lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
    float bv[COUT / 16];  // forward: the bias of the lane's output channels, loaded once
#pragma unroll
    for (int nb = 0; nb < COUT / 16; ++nb) bv[nb] = MODE == 0 ? bias[nb * 16 + m] : 0.0f;
    for (int xs = 0; xs < kConvTW && x0 + xs < W; xs += 16) {
        // data gradient: the gate values of this step's outputs, loaded before the MFMAs so
        // their latency overlaps them
        uint16_t gv[MODE == 1 ? COUT / 16 : 1][4];
        if (MODE == 1) {
#pragma unroll
            for (int nb = 0; nb < COUT / 16; ++nb)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int px = min(x0 + xs + kq * 4 + i, W - 1);
                    gv[nb][i] = gate[((int64_t)y * W + px) * COUT + nb * 16 + m];
                }
        }
        f32x4 acc[COUT / 16];
#pragma unroll
        for (int nb = 0; nb < COUT / 16; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int st = 0; st < ConvK<CIN>::STEPS; ++st) {
            // input pixel (x0 + xs + m + tap % 3 - 1, y + tap / 3 - 1) in tile coordinates,
            // 8 channels from the lane group's first
            const int tap = ConvK<CIN>::tap(st, kq);
            bf16x8 a{};
            if (tap < 9)
                a = load8(tile + ((row + tap / 3) * T::TWH + xs + m + tap % 3) * T::PX + ConvK<CIN>::chan(kq) * 2);
#pragma unroll
            for (int nb = 0; nb < COUT / 16; ++nb) acc[nb] = mfma32(a, wr[st][nb], acc[nb]);
        }
#pragma unroll
        for (int nb = 0; nb < COUT / 16; ++nb) {
            const int co = nb * 16 + m;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int px = x0 + xs + kq * 4 + i;
                if (px >= W) continue;
                const int64_t p = (int64_t)y * W + px;
                float v = acc[nb][i];
                if (MODE == 0) {
                    v += bv[nb];
                    if (RELU) {
                        v = fmaxf(v, 0.0f);
                        out[p * COUT + co] = f2bf(v);
                    } else if (co == 0) {
                        outf[p] = v;
                    }
                } else {
                    if (bf2f(gv[MODE == 1 ? nb : 0][i]) <= 0.0f) v = 0.0f;
                    out[p * COUT + co] = f2bf(v);
                }
            }
        }
    }
}

// one workgroup per tile; the weights are loaded after the tile's loads are issued
template <int CIN, int COUT, int MODE, bool RELU>
__global__ __launch_bounds__(64 * kConvTH) void k_conv(const uint16_t *__restrict__ in,
                                                       const uint16_t *__restrict__ w,
                                                       const float *__restrict__ bias,
                                                       const uint16_t *__restrict__ gate, uint16_t *__restrict__ out,
                                                       float *__restrict__ outf, int H, int W) {
    __shared__ __align__(16) unsigned char tile[ConvTile<CIN>::BYTES];
    const int ntx = (W + kConvTW - 1) / kConvTW;
    const int x0 = (blockIdx.x % ntx) * kConvTW, y0 = (blockIdx.x / ntx) * kConvTH;
    conv_stage<CIN>(in, tile, x0, y0, H, W);
    __syncthreads();
    const int row = threadIdx.x >> 6, y = y0 + row;
    if (y >= H) return;  // no barrier follows
    bf16x8 wr[ConvK<CIN>::STEPS][COUT / 16];
    conv_weights<CIN, COUT>(w, wr);
    conv_row<CIN, COUT, MODE, RELU>(tile, wr, bias, gate, out, outf, x0, y, row, W);
}

// persistent workgroups (for the 32 -> 32 layers): the weights are loaded once and the
// workgroup loops over tiles, so a launch reads them 768 times instead of once per tile
template <int CIN, int COUT, int MODE, bool RELU>
__global__ __launch_bounds__(64 * kConvTH) void k_conv_persist(const uint16_t *__restrict__ in,
                                                               const uint16_t *__restrict__ w,
                                                               const float *__restrict__ bias,
                                                               const uint16_t *__restrict__ gate,
                                                               uint16_t *__restrict__ out, float *__restrict__ outf,
                                                               int H, int W) {
    __shared__ __align__(16) unsigned char tile[ConvTile<CIN>::BYTES];
    const int ntx = (W + kConvTW - 1) / kConvTW, nty = (H + kConvTH - 1) / kConvTH;
    const int row = threadIdx.x >> 6;
    bf16x8 wr[ConvK<CIN>::STEPS][COUT / 16];
    conv_weights<CIN, COUT>(w, wr);
    for (int tl = blockIdx.x; tl < ntx * nty; tl += gridDim.x) {
        const int x0 = (tl % ntx) * kConvTW, y0 = (tl / ntx) * kConvTH;
        conv_stage<CIN>(in, tile, x0, y0, H, W);
        __syncthreads();
        if (y0 + row < H) conv_row<CIN, COUT, MODE, RELU>(tile, wr, bias, gate, out, outf, x0, y0 + row, row, W);
        __syncthreads();  // the tile is restaged next
    }
}

// Weight gradient: dW[tap][co][ci] = sum_p dout[p][co] * in[p + tap][ci], db[co] =
// sum_p dout[p][co].  MFMA (K = 32) with M = co, N = ci, K = 8 consecutive pixels per lane
// group, so each operand lane needs eight pixels of ONE channel: the workgroup stages its tile
// (kConvTH rows of kConvTW pixels, plus a one-row halo and eight pixels either side for the
// input) channel-major in LDS, [row][channel][pixel], and reads a fragment with one 16-byte
// load.  The x-shifted taps reuse that load: the dx = -1 fragment is (mid << 16) | pixel
// before, the dx = +1 fragment (mid >> 16) | pixel after << 112.  Workgroups loop over tiles
// keeping the 9 x COUT/16 x CIN/16 accumulators in registers, reduce their waves in LDS (over
// the tile storage) and write one partial [9*COUT*CIN + COUT] set each.
constexpr int kWgSI = 88;  // LDS pixels per staged channel row: a multiple of 8 (16-byte
                           // fragments) and 44 dwords, 12 mod 32, so the 8 channels read by
                           // a quarter of the wave start on distinct groups of four banks
constexpr int kWgHalo = 8;  // input pixels staged on either side of the tile

__device__ inline bf16x8 u128_bf16(unsigned __int128 v) { return __builtin_bit_cast(bf16x8, v); }

template <int CIN, int COUT>
__global__ __launch_bounds__(64 * kConvTH) void k_conv_wgrad(const uint16_t *__restrict__ in,
                                                             const uint16_t *__restrict__ dout,
                                                             float *__restrict__ partial, int H, int W) {
    constexpr int NW = 9 * COUT * CIN;
    constexpr int TIN = (kConvTH + 2) * CIN * kWgSI, TDO = kConvTH * COUT * kWgSI;  // bf16 elements
    constexpr int GIN = (kConvTW + 2 * kWgHalo) / 4, GDO = kConvTW / 4;  // 4-pixel groups per row
    constexpr int LDS_BYTES = 2 * (TIN + TDO) > 4 * (NW + COUT) ? 2 * (TIN + TDO) : 4 * (NW + COUT);
    static_assert(kWgSI >= kConvTW + 2 * kWgHalo && kWgSI % 8 == 0, "staged row too short or misaligned");
    __shared__ __align__(16) unsigned char lds[LDS_BYTES];
    uint16_t *tin = reinterpret_cast<uint16_t *>(lds);  // pixel j <-> x0 - kWgHalo + j, row r <-> y0 - 1 + r
    uint16_t *tdo = tin + TIN;                          // pixel j <-> x0 + j, row r <-> y0 + r
    float *red = reinterpret_cast<float *>(lds);        // after the tile loop only
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6, m = lane & 15, kq = lane >> 4;
    const int ntx = (W + kConvTW - 1) / kConvTW, nty = (H + kConvTH - 1) / kConvTH;
    f32x4 acc[9][COUT / 16][CIN / 16];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int cb = 0; cb < COUT / 16; ++cb)
#pragma unroll
            for (int nb = 0; nb < CIN / 16; ++nb) acc[t][cb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bsum[COUT / 16];
#pragma unroll
    for (int cb = 0; cb < COUT / 16; ++cb) bsum[cb] = 0.0f;

    for (int tile = blockIdx.x; tile < ntx * nty; tile += gridDim.x) {
        const int x0 = (tile % ntx) * kConvTW, y0 = (tile / ntx) * kConvTH;
        // stage: one item = 4 pixels x 8 channels, read as four 16-byte pixel pieces and
        // written as eight 8-byte channel runs
        for (int q = threadIdx.x; q < (kConvTH + 2) * GIN * (CIN / 8) + kConvTH * GDO * (COUT / 8); q += blockDim.x) {
            const bool is_in = q < (kConvTH + 2) * GIN * (CIN / 8);
            const int qq = is_in ? q : q - (kConvTH + 2) * GIN * (CIN / 8);
            const int C = is_in ? CIN : COUT, G = is_in ? GIN : GDO;
            const int c8 = qq % (C / 8), g4 = (qq / (C / 8)) % G, r = qq / ((C / 8) * G);
            const int gy = is_in ? y0 - 1 + r : y0 + r, gx0 = (is_in ? x0 - kWgHalo : x0) + 4 * g4;
            const uint16_t *src = is_in ? in : dout;
            uint4 v[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int gx = gx0 + k;
                v[k] = uint4{0u, 0u, 0u, 0u};
                if (gy >= 0 && gy < H && gx >= 0 && gx < W)
                    v[k] = *reinterpret_cast<const uint4 *>(src + ((int64_t)gy * W + gx) * C + c8 * 8);
            }
            uint16_t *dst = (is_in ? tin : tdo) + (r * C + c8 * 8) * kWgSI + 4 * g4;
#pragma unroll
            for (int w2 = 0; w2 < 4; ++w2) {
                const uint32_t u0 = (&v[0].x)[w2], u1 = (&v[1].x)[w2], u2 = (&v[2].x)[w2], u3 = (&v[3].x)[w2];
                const uint64_t even = (uint64_t)((u0 & 0xffffu) | (u1 << 16)) |
                                      ((uint64_t)((u2 & 0xffffu) | (u3 << 16)) << 32);
                const uint64_t odd = (uint64_t)((u0 >> 16) | (u1 & 0xffff0000u)) |
                                     ((uint64_t)((u2 >> 16) | (u3 & 0xffff0000u)) << 32);
                *reinterpret_cast<uint64_t *>(dst + (2 * w2) * kWgSI) = even;
                *reinterpret_cast<uint64_t *>(dst + (2 * w2 + 1) * kWgSI) = odd;
            }
        }
        __syncthreads();
        if (y0 + row < H) {
            for (int xs = 0; xs < kConvTW && x0 + xs < W; xs += 32) {
                bf16x8 a[COUT / 16];
#pragma unroll
                for (int cb = 0; cb < COUT / 16; ++cb) {
                    const uint4 u =
                        *reinterpret_cast<const uint4 *>(tdo + (row * COUT + cb * 16 + m) * kWgSI + xs + 8 * kq);
                    a[cb] = __builtin_bit_cast(bf16x8, u);
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        bsum[cb] += bf2f((uint16_t)((&u.x)[i] & 0xffffu)) + bf2f((uint16_t)((&u.x)[i] >> 16));
                }
#pragma unroll
                for (int dy = 0; dy < 3; ++dy) {
#pragma unroll
                    for (int nb = 0; nb < CIN / 16; ++nb) {
                        const uint16_t *p =
                            tin + ((row + dy) * CIN + nb * 16 + m) * kWgSI + xs + 8 * kq + kWgHalo;
                        const unsigned __int128 mid = __builtin_bit_cast(unsigned __int128,
                                                                         *reinterpret_cast<const uint4 *>(p));
                        const unsigned __int128 lo = p[-1], hi = p[8];
                        const bf16x8 b0 = u128_bf16((mid << 16) | lo);
                        const bf16x8 b1 = u128_bf16(mid);
                        const bf16x8 b2 = u128_bf16((mid >> 16) | (hi << 112));
#pragma unroll
                        for (int cb = 0; cb < COUT / 16; ++cb) {
                            acc[dy * 3 + 0][cb][nb] = mfma32(a[cb], b0, acc[dy * 3 + 0][cb][nb]);
                            acc[dy * 3 + 1][cb][nb] = mfma32(a[cb], b1, acc[dy * 3 + 1][cb][nb]);
                            acc[dy * 3 + 2][cb][nb] = mfma32(a[cb], b2, acc[dy * 3 + 2][cb][nb]);
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    for (int q = threadIdx.x; q < NW + COUT; q += blockDim.x) red[q] = 0.0f;
    __syncthreads();
    // D[co = 4*kq + i][ci = m] of each (tap, cb, nb) block
#pragma unroll
    for (int tap = 0; tap < 9; ++tap)
#pragma unroll
        for (int cb = 0; cb < COUT / 16; ++cb)
#pragma unroll
            for (int nb = 0; nb < CIN / 16; ++nb)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int co = cb * 16 + kq * 4 + i, ci = nb * 16 + m;
                    atomicAdd(&red[(tap * COUT + co) * CIN + ci], acc[tap][cb][nb][i]);
                }
    // bias: lane (m, kq) summed pixels 8kq..8kq+7 of each segment for channel cb*16+m
#pragma unroll
    for (int cb = 0; cb < COUT / 16; ++cb) {
        float v = bsum[cb];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (kq == 0) atomicAdd(&red[NW + cb * 16 + m], v);
    }
    __syncthreads();
    float *dst = partial + (int64_t)blockIdx.x * (NW + COUT);
    for (int q = threadIdx.x; q < NW + COUT; q += blockDim.x) dst[q] = red[q];
}

// sum the per-workgroup partials: g[q] = sum_b partial[b][q].  A workgroup takes 64
// consecutive q and splits b over 16 rows of threads, then adds the 16 row sums in LDS.
__global__ __launch_bounds__(1024) void k_sum_partials(const float *partial, int nblk, int len, float *g) {
    __shared__ float sh[16][65];
    const int c = threadIdx.x & 63, r = threadIdx.x >> 6;
    const int q = blockIdx.x * 64 + c;
    float s = 0.0f;
    if (q < len) {
#pragma unroll 4
        for (int b = r; b < nblk; b += 16) s += partial[(int64_t)b * len + q];
    }
    sh[r][c] = s;
    __syncthreads();
    if (r == 0 && q < len) {
        float t = 0.0f;
        for (int k = 0; k < 16; ++k) t += sh[k][c];
        g[q] = t;
    }
}

// flipped/transposed BF16 weights for the data gradient: wt[tap][ci][co] = w[8-tap][co][ci]
__global__ void k_flip_weights(const uint16_t *w, uint16_t *wt, int cin, int cout) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 9 * cin * cout) return;
    const int tap = q / (cin * cout), r = q % (cin * cout), ci = r / cout, co = r % cout;
    wt[q] = w[((8 - tap) * cout + co) * cin + ci];
}

// ---------------------------------------------------------------------------------
// features and targets (FP64 grids in, BF16 features out)

// raw features (r, phi1 - phi2, b, free) with r = L phi1 - b on the free nodes (the residual
// of the previous potential for the new charge), zeros elsewhere
__global__ void k_cnn_features_raw(const double *b, const double *phi1, const double *phi2, const uint8_t *mask,
                                   Geo g, float *raw4) {
    const int64_t n = (int64_t)g.nx * g.ny;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    float f[4] = {0.f, 0.f, 0.f, 0.f};
    if (mask[q]) {
        const int i = (int)(q % g.nx);
        int64_t ql = q - 1, qr = q + 1;
        if (g.periodic_x) {
            if (i == 0) ql = q + (g.gmi - 2);
            if (i == g.gmi - 2) qr = q - (g.gmi - 2);
        }
        const double r = phi1[ql] + phi1[qr] + phi1[q - g.nx] + phi1[q + g.nx] - 4.0 * phi1[q] - b[q];
        f[0] = (float)r;
        f[1] = (float)(phi1[q] - phi2[q]);
        f[2] = (float)b[q];
        f[3] = 1.0f;
    }
    *reinterpret_cast<float4 *>(raw4 + q * 4) = make_float4(f[0], f[1], f[2], f[3]);
}

__global__ void k_cnn_target_raw(const double *phi, const double *phi1, const uint8_t *mask, float *t, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    t[q] = mask[q] ? (float)(phi[q] - phi1[q]) : 0.0f;
}

__global__ __launch_bounds__(256) void k_cnn_sumsq(const float *raw4, const float *t, int64_t n, double *acc) {
    double s[5] = {0, 0, 0, 0, 0};
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const float4 f = *reinterpret_cast<const float4 *>(raw4 + q * 4);
        if (f.w > 0.5f) {
            s[0] += (double)f.x * f.x;
            s[1] += (double)f.y * f.y;
            s[2] += (double)f.z * f.z;
            s[3] += (double)t[q] * t[q];
            s[4] += 1.0;
        }
    }
    __shared__ double sh[5][4];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        double v = s[c];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if ((threadIdx.x & 63) == 0) sh[c][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 5) unsafeAtomicAdd(&acc[threadIdx.x], sh[threadIdx.x][0] + sh[threadIdx.x][1] + sh[threadIdx.x][2] +
                                                                 sh[threadIdx.x][3]);
}

__global__ void k_cnn_pack(const float *raw4, CnnScales sc, uint16_t *feat, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float4 f = *reinterpret_cast<const float4 *>(raw4 + q * 4);
    uint16_t *o = feat + q * 16;
    o[0] = f2bf((float)(f.x / sc.s_r));
    o[1] = f2bf((float)(f.y / sc.s_i));
    o[2] = f2bf((float)(f.z / sc.s_b));
    o[3] = f2bf(f.w);
    for (int c = 4; c < 16; ++c) o[c] = 0;
}

__global__ void k_cnn_scale(const float *t, float inv, float *out, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = t[q] * inv;
}

// guess = phi1 + s_d * pred on the free nodes (the Dirichlet nodes keep their values)
__global__ void k_cnn_apply(const float *pred, const double *phi1, const uint8_t *mask, double s_d, double *x,
                            int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    if (mask[q]) x[q] = phi1[q] + s_d * (double)pred[q];
}

// MSE over the free nodes: loss += sum (pred - t)^2 / nfree; dout (16 channels BF16) gets
// channel 0 = 2 (pred - t)/nfree
__global__ __launch_bounds__(256) void k_cnn_loss(const float *pred, const float *t, const uint16_t *feat, double inv_n,
                                                  uint16_t *dout, double *loss, int64_t n) {
    double acc = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const bool freen = bf2f(feat[q * 16 + 3]) > 0.5f;
        const float e = freen ? pred[q] - t[q] : 0.0f;
        acc += (double)e * (double)e;
        uint16_t *o = dout + q * 16;
        o[0] = f2bf((float)(2.0 * inv_n) * e);
        for (int c = 1; c < 16; ++c) o[c] = 0;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    __shared__ double sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) unsafeAtomicAdd(loss, (sh[0] + sh[1] + sh[2] + sh[3]) * inv_n);
}

// Adam on the FP32 master weights; writes the BF16 copy.  `frozen` marks parameters that
// must stay zero (the padded output channels of the last layer).
__global__ void k_adam(float *p, const float *g, float *m1, float *m2, const uint8_t *frozen, uint16_t *pb,
                       int64_t n, float lr, float b1, float b2, float c1, float c2, float eps) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    if (frozen[q]) {
        pb[q] = 0;
        return;
    }
    const float gq = g[q];
    const float a = b1 * m1[q] + (1.0f - b1) * gq;
    const float v = b2 * m2[q] + (1.0f - b2) * gq * gq;
    m1[q] = a;
    m2[q] = v;
    const float w = p[q] - lr * (a / c1) / (sqrtf(v / c2) + eps);
    p[q] = w;
    pb[q] = f2bf(w);
}

__global__ void k_to_bf16(const float *p, uint16_t *pb, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) pb[q] = f2bf(p[q]);
}

// ---------------------------------------------------------------------------------
// host side

static int blocks_for(int64_t n, int per = 256) { return (int)((n + per - 1) / per); }

// The 32 -> 32 layers run 768 persistent workgroups (k_conv_persist: 3 resident per CU by
// registers, 256 CUs) that load their weight fragments once; the 16-channel layers,
// lighter on registers, run one workgroup per tile (k_conv) at higher occupancy.
constexpr int kConvGrid = 768;
static int conv_tiles(int H, int W) {
    return ((W + kConvTW - 1) / kConvTW) * ((H + kConvTH - 1) / kConvTH);
}
static int conv_blocks(int H, int W) { return conv_tiles(H, W); }
static int conv_blocks_32(int H, int W) { return std::min(conv_tiles(H, W), kConvGrid); }

void cnn_layer_shape(int l, int &cin, int &cout) {
    cin = l == 0 ? 16 : 32;
    cout = l == kCnnLayers - 1 ? 16 : 32;
}

void cnn_forward(const CnnNet &net, CnnActs &a, int H, int W, hipStream_t s) {
    const int nb = conv_blocks(H, W);
    k_conv<16, 32, 0, true><<<nb, 64 * kConvTH, 0, s>>>(a.feat, net.wb[0], net.b[0], nullptr, a.act[0], nullptr, H, W);
    for (int l = 1; l < kCnnLayers - 1; ++l)
        k_conv_persist<32, 32, 0, true><<<conv_blocks_32(H, W), 64 * kConvTH, 0, s>>>(a.act[l - 1], net.wb[l], net.b[l], nullptr, a.act[l], nullptr, H,
                                                  W);
    k_conv<32, 16, 0, false><<<nb, 64 * kConvTH, 0, s>>>(a.act[kCnnLayers - 2], net.wb[kCnnLayers - 1], net.b[kCnnLayers - 1],
                                               nullptr, nullptr, a.pred, H, W);
}

// gradients of the loss (whose dL/dpred sits in a.dout16) into net.grad
hipError_t cnn_backward(CnnNet &net, CnnActs &a, int H, int W, hipStream_t s) {
    const int nb = conv_blocks(H, W);
    const int L = kCnnLayers;
    uint16_t *gcur = a.dout16;  // dL/d(out of layer l)
    for (int l = L - 1; l >= 0; --l) {
        int cin, cout;
        cnn_layer_shape(l, cin, cout);
        float *part = a.partial;
        const uint16_t *lin = l == 0 ? a.feat : a.act[l - 1];
        if (cin == 16 && cout == 32)
            k_conv_wgrad<16, 32><<<kCnnWgradBlocks, 64 * kConvTH, 0, s>>>(lin, gcur, part, H, W);
        else if (cin == 32 && cout == 32)
            k_conv_wgrad<32, 32><<<kCnnWgradBlocks, 64 * kConvTH, 0, s>>>(lin, gcur, part, H, W);
        else
            k_conv_wgrad<32, 16><<<kCnnWgradBlocks, 64 * kConvTH, 0, s>>>(lin, gcur, part, H, W);
        const int len = 9 * cin * cout + cout;
        k_sum_partials<<<(len + 63) / 64, 1024, 0, s>>>(part, kCnnWgradBlocks, len, a.gsum);
        // split into weight and bias gradients
        hipError_t e = hipMemcpyAsync(net.gw[l], a.gsum, sizeof(float) * 9 * cin * cout, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(net.gb[l], a.gsum + 9 * cin * cout, sizeof(float) * cout, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
        if (l == 0) break;
        // data gradient into the previous activation (through its ReLU)
        k_flip_weights<<<blocks_for(9 * cin * cout), 256, 0, s>>>(net.wb[l], a.wflip, cin, cout);
        uint16_t *gnext = (gcur == a.g[0]) ? a.g[1] : a.g[0];
        if (cout == 16)
            k_conv<16, 32, 1, false><<<nb, 64 * kConvTH, 0, s>>>(gcur, a.wflip, nullptr, a.act[l - 1], gnext, nullptr, H, W);
        else
            k_conv_persist<32, 32, 1, false><<<conv_blocks_32(H, W), 64 * kConvTH, 0, s>>>(gcur, a.wflip, nullptr, a.act[l - 1], gnext,
                                                                               nullptr, H, W);
        gcur = gnext;
    }
    return hipSuccess;
}

hipError_t cnn_loss(CnnActs &a, const float *target, int64_t n, int64_t nfree, double *dloss, double *loss,
                    hipStream_t s) {
    hipError_t e = hipMemsetAsync(dloss, 0, sizeof(double), s);
    if (e != hipSuccess) return e;
    k_cnn_loss<<<std::min<int64_t>(blocks_for(n), 4096), 256, 0, s>>>(a.pred, target, a.feat, 1.0 / (double)nfree,
                                                                      a.dout16, dloss, n);
    e = hipMemcpyAsync(loss, dloss, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

void cnn_adam(CnnNet &net, float lr, hipStream_t s) {
    net.t += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - powf(b1, (float)net.t), c2 = 1.0f - powf(b2, (float)net.t);
    k_adam<<<blocks_for(net.nparam), 256, 0, s>>>(net.master, net.grad, net.m1, net.m2, net.frozen, net.wbf,
                                                  net.nparam, lr, b1, b2, c1, c2, eps);
}

void cnn_sync_bf16(CnnNet &net, hipStream_t s) {
    k_to_bf16<<<blocks_for(net.nparam), 256, 0, s>>>(net.master, net.wbf, net.nparam);
}

void launch_cnn_features_raw(const double *b, const double *phi1, const double *phi2, const uint8_t *mask,
                             const Geo &g, float *raw4, hipStream_t s) {
    const int64_t n = (int64_t)g.nx * g.ny;
    k_cnn_features_raw<<<blocks_for(n), 256, 0, s>>>(b, phi1, phi2, mask, g, raw4);
}

void launch_cnn_target_raw(const double *phi, const double *phi1, const uint8_t *mask, float *t, int64_t n,
                           hipStream_t s) {
    k_cnn_target_raw<<<blocks_for(n), 256, 0, s>>>(phi, phi1, mask, t, n);
}

void launch_cnn_sumsq(const float *raw4, const float *t, int64_t n, double *acc, hipStream_t s) {
    k_cnn_sumsq<<<std::min<int64_t>(blocks_for(n), 2048), 256, 0, s>>>(raw4, t, n, acc);
}

void launch_cnn_pack(const float *raw4, const CnnScales &sc, uint16_t *feat, int64_t n, hipStream_t s) {
    k_cnn_pack<<<blocks_for(n), 256, 0, s>>>(raw4, sc, feat, n);
}

void launch_cnn_scale(const float *t, float inv, float *out, int64_t n, hipStream_t s) {
    k_cnn_scale<<<blocks_for(n), 256, 0, s>>>(t, inv, out, n);
}

void launch_cnn_apply(const float *pred, const double *phi1, const uint8_t *mask, double s_d, double *x, int64_t n,
                      hipStream_t s) {
    k_cnn_apply<<<blocks_for(n), 256, 0, s>>>(pred, phi1, mask, s_d, x, n);
}


}  // namespace edgpu
