// BF16 3x3x3 convolutions on MFMA (implicit GEMM) and the volume kernels of
// the learned preconditioner: forward, backward-data, backward-weight, bias
// gradient, vertex-centred transfers, ReLU masks, Adam.
//
// Layout: channels-last volumes (D, H, W, C) with C a multiple of 16, stored
// as BF16 bit patterns (uint16).  The node grid of the FE mesh maps to
// (D, H, W) = (ny+1, nx+1, nz+1) in the mesh's node order, so a displacement
// vector (node*3 + c) is already a channels-last volume with 3 channels.
// Weights: W[tap][c_out][c_in] BF16, tap = (dy+1)*9 + (dx+1)*3 + (dz+1).
// Accumulation is FP32 on v_mfma_f32_16x16x16_bf16:
//   A (16 x 16, M x K): lane l holds A[l%16][4*(l/16) .. +3]
//   B (16 x 16, K x N): lane l holds B[4*(l/16) .. +3][l%16]
//   D (16 x 16):        lane l holds D[4*(l/16) + i][l%16], i = 0..3
// Forward: M = voxels, K = c_in (per tap), N = c_out.  Backward-weight:
// M = c_out, K = voxels, N = c_in.  Backward-data is the forward with the
// tap-flipped, channel-transposed weights.
#include <algorithm>

#include "common.hpp"

namespace fantom {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u)
        return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

constexpr int kConvBlock = 256;  // 4 waves
constexpr int kRowTiles = 2;     // 16-voxel row tiles per wave

// out = act(conv(x, w) + bias + skip); act = ReLU when relu != 0.
// out_f32 (optional) receives the FP32 pre-rounding values.
template <int CO_T>
__global__ __launch_bounds__(kConvBlock) void conv3d_fwd_k(int D, int H, int W, int Cin,
                                                         const uint16_t *__restrict__ x,
                                                         const uint16_t *__restrict__ w,
                                                         const float *__restrict__ bias,
                                                         const uint16_t *__restrict__ skip,
                                                         uint16_t *__restrict__ out, float *__restrict__ out_f32,
                                                         int relu)
{
    constexpr int Cout = CO_T * 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long N = (long)D * H * W;
    const long v0 = ((long)blockIdx.x * 4 + wave) * (16 * kRowTiles);
    if (v0 >= N)
        return;  // uniform per wave; the kernel has no block-wide barrier
    const int row = lane & 15, kg = lane >> 4;
    int pj[kRowTiles], pi[kRowTiles], pk[kRowTiles];
    bool pok[kRowTiles];
#pragma unroll
    for (int rt = 0; rt < kRowTiles; ++rt) {
        const long v = v0 + rt * 16 + row;
        pok[rt] = v < N;
        const long vv = pok[rt] ? v : 0;
        pk[rt] = (int)(vv % W);
        pi[rt] = (int)((vv / W) % H);
        pj[rt] = (int)(vv / ((long)W * H));
    }
    f4 acc[kRowTiles][CO_T];
#pragma unroll
    for (int rt = 0; rt < kRowTiles; ++rt)
#pragma unroll
        for (int ct = 0; ct < CO_T; ++ct)
            acc[rt][ct] = f4{0.f, 0.f, 0.f, 0.f};
    const int cchunks = Cin / 16;
    for (int t = 0; t < 27; ++t) {
        const int dy = t / 9 - 1, dx = (t / 3) % 3 - 1, dz = t % 3 - 1;
        long nb[kRowTiles];
        bool ok[kRowTiles];
#pragma unroll
        for (int rt = 0; rt < kRowTiles; ++rt) {
            const int j = pj[rt] + dy, i = pi[rt] + dx, k = pk[rt] + dz;
            ok[rt] = pok[rt] && j >= 0 && j < D && i >= 0 && i < H && k >= 0 && k < W;
            nb[rt] = ok[rt] ? ((long)j * H + i) * W + k : 0;
        }
        for (int c = 0; c < cchunks; ++c) {
            s4 a[kRowTiles];
#pragma unroll
            for (int rt = 0; rt < kRowTiles; ++rt) {
                a[rt] = s4{0, 0, 0, 0};
                if (ok[rt])
                    a[rt] = *reinterpret_cast<const s4 *>(x + nb[rt] * Cin + c * 16 + 4 * kg);
            }
#pragma unroll
            for (int ct = 0; ct < CO_T; ++ct) {
                const s4 b = *reinterpret_cast<const s4 *>(w + ((long)t * Cout + ct * 16 + row) * Cin + c * 16 + 4 * kg);
#pragma unroll
                for (int rt = 0; rt < kRowTiles; ++rt)
                    acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[rt], b, acc[rt][ct], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int rt = 0; rt < kRowTiles; ++rt)
#pragma unroll
        for (int ct = 0; ct < CO_T; ++ct) {
            const int co = ct * 16 + row;
            const float bco = bias ? bias[co] : 0.f;
#pragma unroll
            for (int ii = 0; ii < 4; ++ii) {
                const long v = v0 + rt * 16 + 4 * kg + ii;
                if (v >= N)
                    continue;
                float val = acc[rt][ct][ii] + bco;
                if (skip)
                    val += bf2f(skip[v * Cout + co]);
                if (relu)
                    val = fmaxf(val, 0.f);
                out[v * Cout + co] = f2bf(val);
                if (out_f32)
                    out_f32[v * Cout + co] = val;
            }
        }
}

// Forward with the weights staged in LDS: a workgroup holds the weights of
// CO_W output tiles (cout [16 CO_W blockIdx.y, +16 CO_W)) for all 27 taps and
// Cin = 16 CI_T, loaded once, rows padded by 4 entries so that the B-operand
// reads of the 16 rows fall in distinct banks.  Each wave then walks groups
// of kFwdRT consecutive 16-voxel row tiles (grid-stride), and every weight
// operand read from LDS feeds kFwdRT MFMAs.  Same output as conv3d_fwd_k.
constexpr int kFwdRT = 4;

__host__ __device__ constexpr int fwd2_row(int cin) { return cin + 4; }   // LDS row stride (entries)

template <int CI_T, int CO_W>
__global__ __launch_bounds__(kConvBlock) void conv3d_fwd2_k(int D, int H, int W, int Cout,
                                                          const uint16_t *__restrict__ x,
                                                          const uint16_t *__restrict__ w,
                                                          const float *__restrict__ bias,
                                                          const uint16_t *__restrict__ skip,
                                                          uint16_t *__restrict__ out, float *__restrict__ out_f32,
                                                          int relu)
{
    constexpr int Cin = CI_T * 16, Srow = fwd2_row(Cin), CoW = CO_W * 16;
    extern __shared__ uint16_t Wl[];   // [27][CoW][Srow]
    const int cb = blockIdx.y * CoW;
    // stage the weight slice: 8-byte pieces of w[t][cb + co][ci]
    for (int q = threadIdx.x; q < 27 * CoW * (Cin / 4); q += kConvBlock) {
        const int piece = q % (Cin / 4), rowi = q / (Cin / 4);
        const int t = rowi / CoW, co = rowi % CoW;
        *reinterpret_cast<s4 *>(Wl + (t * CoW + co) * Srow + 4 * piece) =
            *reinterpret_cast<const s4 *>(w + ((long)t * Cout + cb + co) * Cin + 4 * piece);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kg = lane >> 4;
    const long N = (long)D * H * W;
    const long ngroups = (N + 16 * kFwdRT - 1) / (16 * kFwdRT);
    for (long gi = (long)blockIdx.x * 4 + wave; gi < ngroups; gi += (long)gridDim.x * 4) {
        const long v0 = gi * (16 * kFwdRT);
        int pj[kFwdRT], pi[kFwdRT], pk[kFwdRT];
        bool pok[kFwdRT];
#pragma unroll
        for (int rt = 0; rt < kFwdRT; ++rt) {
            const long v = v0 + rt * 16 + row;
            pok[rt] = v < N;
            const long vv = pok[rt] ? v : 0;
            pk[rt] = (int)(vv % W);
            pi[rt] = (int)((vv / W) % H);
            pj[rt] = (int)(vv / ((long)W * H));
        }
        f4 acc[kFwdRT][CO_W];
#pragma unroll
        for (int rt = 0; rt < kFwdRT; ++rt)
#pragma unroll
            for (int ct = 0; ct < CO_W; ++ct)
                acc[rt][ct] = f4{0.f, 0.f, 0.f, 0.f};
        for (int t = 0; t < 27; ++t) {
            const int dy = t / 9 - 1, dx = (t / 3) % 3 - 1, dz = t % 3 - 1;
            long nb[kFwdRT];
            bool ok[kFwdRT];
#pragma unroll
            for (int rt = 0; rt < kFwdRT; ++rt) {
                const int j = pj[rt] + dy, i = pi[rt] + dx, k = pk[rt] + dz;
                ok[rt] = pok[rt] && j >= 0 && j < D && i >= 0 && i < H && k >= 0 && k < W;
                nb[rt] = ok[rt] ? ((long)j * H + i) * W + k : 0;
            }
#pragma unroll
            for (int c = 0; c < CI_T; ++c) {
                s4 a[kFwdRT];
#pragma unroll
                for (int rt = 0; rt < kFwdRT; ++rt) {
                    a[rt] = s4{0, 0, 0, 0};
                    if (ok[rt])
                        a[rt] = *reinterpret_cast<const s4 *>(x + nb[rt] * Cin + c * 16 + 4 * kg);
                }
#pragma unroll
                for (int ct = 0; ct < CO_W; ++ct) {
                    const s4 b = *reinterpret_cast<const s4 *>(Wl + (t * CoW + ct * 16 + row) * Srow + c * 16 + 4 * kg);
#pragma unroll
                    for (int rt = 0; rt < kFwdRT; ++rt)
                        acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[rt], b, acc[rt][ct], 0, 0, 0);
                }
            }
        }
#pragma unroll
        for (int rt = 0; rt < kFwdRT; ++rt)
#pragma unroll
            for (int ct = 0; ct < CO_W; ++ct) {
                const int co = cb + ct * 16 + row;
                const float bco = bias ? bias[co] : 0.f;
#pragma unroll
                for (int ii = 0; ii < 4; ++ii) {
                    const long v = v0 + rt * 16 + 4 * kg + ii;
                    if (v >= N)
                        continue;
                    float val = acc[rt][ct][ii] + bco;
                    if (skip)
                        val += bf2f(skip[v * Cout + co]);
                    if (relu)
                        val = fmaxf(val, 0.f);
                    out[v * Cout + co] = f2bf(val);
                    if (out_f32)
                        out_f32[v * Cout + co] = val;
                }
            }
    }
}

// Forward on bricks staged in LDS (level-0/1 shapes, Cin <= 32): a workgroup
// owns a brick of 16 (W) x BI (H) x BJ (D) output voxels, loads the input
// brick with its one-voxel halo -- 18 x (BI + 2) x (BJ + 2) voxels, rows padded
// by 4 channels so that the 16 voxels of an A operand hit distinct banks --
// and the weights of its CO_T output tiles once, then each wave runs the 27
// taps of its row tiles (16 voxels along W at fixed (j, i)) from LDS only.
// Every input voxel is read from memory once per brick instead of 27 times.
template <int CI_T, int CO_T, int BI, int BJ>
__global__ __launch_bounds__(kConvBlock) void conv3d_brick_k(int D, int H, int W, int Cout,
                                                           const uint16_t *__restrict__ x,
                                                           const uint16_t *__restrict__ w,
                                                           const float *__restrict__ bias,
                                                           const uint16_t *__restrict__ skip,
                                                           uint16_t *__restrict__ out, float *__restrict__ out_f32,
                                                           int relu)
{
    constexpr int Cin = CI_T * 16, Sv = Cin + 4, CoT = CO_T * 16, Sw = Cin + 4;
    constexpr int HK = 18, HI = BI + 2, HJ = BJ + 2, NHV = HK * HI * HJ;
    constexpr int RTW = BI * BJ / 4;   // row tiles per wave
    extern __shared__ uint16_t sm[];
    uint16_t *Xs = sm;                 // [NHV][Sv]
    uint16_t *Ws = sm + NHV * Sv;      // [27][CoT][Sw]
    const int nbk = (W + 15) / 16, nbi = (H + BI - 1) / BI;
    const int bk = blockIdx.x % nbk, bi = (blockIdx.x / nbk) % nbi, bj = blockIdx.x / (nbk * nbi);
    const int k0 = bk * 16, i0 = bi * BI, j0 = bj * BJ, cb = blockIdx.y * CoT;
    // input halo brick, 8-byte pieces; zero outside the volume
    for (int q = threadIdx.x; q < NHV * (Cin / 4); q += kConvBlock) {
        const int piece = q % (Cin / 4), hv = q / (Cin / 4);
        const int hk = hv % HK, hi = (hv / HK) % HI, hj = hv / (HK * HI);
        const int k = k0 - 1 + hk, i = i0 - 1 + hi, j = j0 - 1 + hj;
        s4 val = s4{0, 0, 0, 0};
        if (k >= 0 && k < W && i >= 0 && i < H && j >= 0 && j < D)
            val = *reinterpret_cast<const s4 *>(x + (((long)j * H + i) * W + k) * Cin + 4 * piece);
        *reinterpret_cast<s4 *>(Xs + hv * Sv + 4 * piece) = val;
    }
    for (int q = threadIdx.x; q < 27 * CoT * (Cin / 4); q += kConvBlock) {
        const int piece = q % (Cin / 4), rowi = q / (Cin / 4);
        const int t = rowi / CoT, co = rowi % CoT;
        *reinterpret_cast<s4 *>(Ws + (t * CoT + co) * Sw + 4 * piece) =
            *reinterpret_cast<const s4 *>(w + ((long)t * Cout + cb + co) * Cin + 4 * piece);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kg = lane >> 4;
    f4 acc[RTW][CO_T];
#pragma unroll
    for (int r = 0; r < RTW; ++r)
#pragma unroll
        for (int ct = 0; ct < CO_T; ++ct)
            acc[r][ct] = f4{0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < 27; ++t) {
        const int dy = t / 9 - 1, dx = (t / 3) % 3 - 1, dz = t % 3 - 1;
#pragma unroll
        for (int c = 0; c < CI_T; ++c) {
            s4 b[CO_T];
#pragma unroll
            for (int ct = 0; ct < CO_T; ++ct)
                b[ct] = *reinterpret_cast<const s4 *>(Ws + (t * CoT + ct * 16 + row) * Sw + c * 16 + 4 * kg);
#pragma unroll
            for (int r = 0; r < RTW; ++r) {
                const int rt = wave * RTW + r, il = rt % BI, jl = rt / BI;
                const int hv = ((jl + 1 + dy) * HI + (il + 1 + dx)) * HK + (row + 1 + dz);
                const s4 a = *reinterpret_cast<const s4 *>(Xs + hv * Sv + c * 16 + 4 * kg);
#pragma unroll
                for (int ct = 0; ct < CO_T; ++ct)
                    acc[r][ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[ct], acc[r][ct], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RTW; ++r) {
        const int rt = wave * RTW + r, il = rt % BI, jl = rt / BI;
        const int i = i0 + il, j = j0 + jl;
        if (i >= H || j >= D)
            continue;
#pragma unroll
        for (int ct = 0; ct < CO_T; ++ct) {
            const int co = cb + ct * 16 + row;
            const float bco = bias ? bias[co] : 0.f;
#pragma unroll
            for (int ii = 0; ii < 4; ++ii) {
                const int k = k0 + 4 * kg + ii;
                if (k >= W)
                    continue;
                const long v = ((long)j * H + i) * W + k;
                float val = acc[r][ct][ii] + bco;
                if (skip)
                    val += bf2f(skip[v * Cout + co]);
                if (relu)
                    val = fmaxf(val, 0.f);
                out[v * Cout + co] = f2bf(val);
                if (out_f32)
                    out_f32[v * Cout + co] = val;
            }
        }
    }
}

// Forward that walks planes along D (the level-0/1/2 shapes).  A workgroup
// owns a tile of 16 (W) x BI (H) output voxels, CO_T tiles of 16 output
// channels (blockIdx.y picks them when Cout is split), and a chunk of planes
// along D.  Its weights stay in LDS for the whole walk, and the input planes
// j - 1, j, j + 1 with their one-voxel halo (18 x (BI + 2) voxels) sit in a
// ring of three LDS slots: every input voxel is read from memory about 1.3
// times instead of 27.  Plane j + 2 is fetched into registers while plane j
// is computed and is written to the slot of plane j - 1 between two LDS-only
// barriers, so neither the fetch nor the output stores are drained by a
// barrier.
//
// The products run on v_mfma_f32_16x16x32_bf16 with the weights as A (M =
// c_out) and the voxels as B (N = 16 voxels along W).  The K axis is the
// flattened (tap, c_in) axis, f = tap Cin + c_in, in blocks of 32: lane group
// kg = lane / 16 supplies the 8 consecutive entries f = 32 b + 8 kg .. + 7 of
// block b (one tap, 8 channels: a 16-byte LDS read for A and for B), so that
// 16-channel layers pair two taps per instruction; entries past 27 Cin are
// zero.  Each lane ends with four consecutive output channels of one voxel
// and writes them as one 8-byte store (16 bytes for the FP32 copy).
//
// LDS layouts follow what ds_read_b128 was measured to read at full rate: a
// group of 16 lanes reading one aligned 256-byte block, or 16-byte pieces at
// a stride that is an odd multiple of 32 bytes (other strides, and blocks
// that straddle 256 bytes, ran 1.4-2.7x slower).  Weights are
// [tap][8-channel chunk][c_out][8]: the 16 c_out rows of a lane group are one
// aligned block.  For BI = 16 a staged plane is [chunk][W (18)][H (18)][8]:
// the 16 voxels along W of a B operand sit 288 bytes apart whatever the
// tap's shift (the 16-channel layers: 149 -> 117 us at 257 x 129 x 129).  For
// BI = 8 the plain [chunk][H][W][8] order measured faster (118 against 141 us
// for the 64-channel layer, the same for 32), so it stays there.  Same output
// as conv3d_fwd_k up to the order of the FP32 sums.
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

template <int CI_T, int CO_T, int BI>
__global__ __launch_bounds__(kConvBlock) void conv3d_roll_k(int D, int H, int W, int Cout, int chunk,
                                                          const uint16_t *__restrict__ x,
                                                          const uint16_t *__restrict__ w,
                                                          const float *__restrict__ bias,
                                                          const uint16_t *__restrict__ skip,
                                                          uint16_t *__restrict__ out, float *__restrict__ out_f32,
                                                          int relu)
{
    constexpr int Cin = CI_T * 16, CoT = CO_T * 16;
    constexpr int HK = 18, HI = BI + 2, PV = HK * HI;   // voxels of a staged plane
    constexpr int PPV = Cin / 8;                        // 16-byte pieces per voxel
    constexpr int NPIECE = PV * PPV;
    constexpr int NP = (NPIECE + kConvBlock - 1) / kConvBlock;
    constexpr int RTW = BI / 4;                         // row tiles (values of i) per wave
    constexpr int NKB = (27 * Cin + 31) / 32;           // K blocks of 32
    static_assert(BI % 4 == 0, "four waves share the BI rows");
    constexpr bool kWMajor = BI == 16;   // staged plane order (see above)
    extern __shared__ __attribute__((aligned(16))) uint16_t sm[];
    uint16_t *Ws = sm;                      // [27][PPV][CoT][8]
    uint16_t *Xs = sm + 27 * PPV * CoT * 8; // [3][PPV][HK][HI][8] (kWMajor) or [3][PPV][HI][HK][8]
    const int nbk = (W + 15) / 16, nbi = (H + BI - 1) / BI;
    const int bk = blockIdx.x % nbk, bi = (blockIdx.x / nbk) % nbi, cj = blockIdx.x / (nbk * nbi);
    const int k0 = bk * 16, i0 = bi * BI, cb = blockIdx.y * CoT;
    const int j0 = cj * chunk, j1 = min(j0 + chunk, D);   // the host sizes the grid so that j0 < D
    for (int q = threadIdx.x; q < 27 * CoT * PPV; q += kConvBlock) {
        const int piece = q % PPV, rowi = q / PPV;
        const int t = rowi / CoT, co = rowi % CoT;
        *reinterpret_cast<us8 *>(Ws + ((t * PPV + piece) * CoT + co) * 8) =
            *reinterpret_cast<const us8 *>(w + ((long)t * Cout + cb + co) * Cin + 8 * piece);
    }
    // plane j (zeros outside the volume) into registers, registers into a slot
    auto fetch = [&](int j, us8 (&pf)[NP]) {
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const int q = threadIdx.x + p * kConvBlock;
            const int hv = q / PPV, piece = q % PPV;
            const int hk = hv % HK, hi = hv / HK;
            const int k = k0 - 1 + hk, i = i0 - 1 + hi;
            us8 v = us8{0, 0, 0, 0, 0, 0, 0, 0};
            if (q < NPIECE && j >= 0 && j < D && i >= 0 && i < H && k >= 0 && k < W)
                v = *reinterpret_cast<const us8 *>(x + (((long)j * H + i) * W + k) * Cin + 8 * piece);
            pf[p] = v;
        }
    };
    auto stash = [&](int j, const us8 (&pf)[NP]) {
        const int slot = ((j % 3) + 3) % 3;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const int q = threadIdx.x + p * kConvBlock;
            if (q < NPIECE)
                *reinterpret_cast<us8 *>(
                    Xs + ((slot * PPV + q % PPV) * PV + (kWMajor ? (q / PPV) % HK * HI + (q / PPV) / HK : q / PPV)) * 8) =
                    pf[p];
        }
    };
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kg = lane >> 4;
    const int k = k0 + row;
    float bv[CO_T][4];
#pragma unroll
    for (int ct = 0; ct < CO_T; ++ct)
#pragma unroll
        for (int q = 0; q < 4; ++q)
            bv[ct][q] = bias ? bias[cb + ct * 16 + 4 * kg + q] : 0.f;
    us8 pf[NP];
    fetch(j0 - 1, pf);
    stash(j0 - 1, pf);
    fetch(j0, pf);
    stash(j0, pf);
    fetch(j0 + 1, pf);
    stash(j0 + 1, pf);
    lds_barrier();
    for (int j = j0; j < j1; ++j) {
        fetch(j + 2 <= j1 ? j + 2 : -1, pf);
        // skip values of this plane, loaded with the fetch
        s4 sk[RTW][CO_T];
#pragma unroll
        for (int r = 0; r < RTW; ++r) {
            const int i = i0 + wave * RTW + r;
            const long v = ((long)j * H + i) * W + k;
#pragma unroll
            for (int ct = 0; ct < CO_T; ++ct) {
                sk[r][ct] = s4{0, 0, 0, 0};
                if (skip && i < H && k < W)
                    sk[r][ct] = *reinterpret_cast<const s4 *>(skip + v * Cout + cb + ct * 16 + 4 * kg);
            }
        }
        f4 acc[RTW][CO_T];
#pragma unroll
        for (int r = 0; r < RTW; ++r)
#pragma unroll
            for (int ct = 0; ct < CO_T; ++ct)
                acc[r][ct] = f4{0.f, 0.f, 0.f, 0.f};
        // slots of planes j - 1, j, j + 1 (j >= 0); the block loop is
        // unrolled so that the LDS reads of later blocks are issued ahead of
        // the MFMAs
        const int s0 = (j + 2) % 3;
        const uint16_t *Xpl[3] = {Xs + s0 * PPV * PV * 8, Xs + ((s0 + 1) % 3) * PPV * PV * 8,
                                  Xs + ((s0 + 2) % 3) * PPV * PV * 8};
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb) {
            // this lane's tap and first channel; the tap depends on kg only
            // when two taps share the block (Cin = 16)
            const int f = 32 * kb + 8 * kg;
            const bool live = f < 27 * Cin;
            const int t = live ? f / Cin : 0, c0 = f % Cin;
            const int dy = t / 9 - 1, dx = (t / 3) % 3 - 1, dz = t % 3 - 1;
            const uint16_t *Xp = Xpl[dy + 1];
            bf8v a[CO_T];
#pragma unroll
            for (int ct = 0; ct < CO_T; ++ct) {
                const us8 av = live ? *reinterpret_cast<const us8 *>(Ws + ((t * PPV + c0 / 8) * CoT + ct * 16 + row) * 8)
                                    : us8{0, 0, 0, 0, 0, 0, 0, 0};
                a[ct] = __builtin_bit_cast(bf8v, av);
            }
#pragma unroll
            for (int r = 0; r < RTW; ++r) {
                const int hv = kWMajor ? (row + 1 + dz) * HI + (wave * RTW + r + 1 + dx)
                                       : (wave * RTW + r + 1 + dx) * HK + (row + 1 + dz);
                const bf8v b = __builtin_bit_cast(bf8v, *reinterpret_cast<const us8 *>(Xp + ((c0 / 8) * PV + hv) * 8));
#pragma unroll
                for (int ct = 0; ct < CO_T; ++ct)
                    acc[r][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ct], b, acc[r][ct], 0, 0, 0);
            }
        }
        lds_barrier();   // every wave is done with plane j - 1's slot
        if (j + 2 <= j1)
            stash(j + 2, pf);
#pragma unroll
        for (int r = 0; r < RTW; ++r) {
            const int i = i0 + wave * RTW + r;
            if (i >= H || k >= W)
                continue;
            const long v = ((long)j * H + i) * W + k;
#pragma unroll
            for (int ct = 0; ct < CO_T; ++ct) {
                const int co = cb + ct * 16 + 4 * kg;
                float val[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    val[q] = acc[r][ct][q] + bv[ct][q];
                    if (skip)
                        val[q] += bf2f((uint16_t)sk[r][ct][q]);
                    if (relu)
                        val[q] = fmaxf(val[q], 0.f);
                }
                s4 o;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    o[q] = (short)f2bf(val[q]);
                *reinterpret_cast<s4 *>(out + v * Cout + co) = o;
                if (out_f32)
                    *reinterpret_cast<f4 *>(out_f32 + v * Cout + co) = f4{val[0], val[1], val[2], val[3]};
            }
        }
        lds_barrier();   // plane j + 2 is in its slot
    }
}

// LDS bytes of conv3d_roll_k<CI_T, CO_T, BI>
__host__ __device__ constexpr size_t roll_lds(int ci_t, int co_t, int bi)
{
    return ((size_t)27 * co_t * 16 * ci_t * 16 + (size_t)3 * 18 * (bi + 2) * ci_t * 16) * 2;
}

// Grid of conv3d_roll_k: tiles x chunks of planes along D; the chunk count
// minimises (rounds of resident workgroups) x (planes per chunk + 2 halo).
template <int CI_T, int CO_T, int BI>
static int launch_roll(int D, int H, int W, int Cout, const uint16_t *x, const uint16_t *w, const float *bias,
                       const uint16_t *skip, uint16_t *out, float *out_f32, int relu, hipStream_t s)
{
    constexpr size_t lds = roll_lds(CI_T, CO_T, BI);
    static_assert(lds <= 160 * 1024, "LDS budget");
    static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void *>(&conv3d_roll_k<CI_T, CO_T, BI>),
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (attr != hipSuccess) {
        set_error("conv3d_roll LDS attribute: %s", hipGetErrorString(attr));
        return kStatusHipError;
    }
    const long tiles = (long)((W + 15) / 16) * ((H + BI - 1) / BI);
    const unsigned ycnt = (unsigned)(Cout / (16 * CO_T));
    // resident workgroups per CU from the kernel's registers and LDS
    static const int per_cu = [] {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void *>(&conv3d_roll_k<CI_T, CO_T, BI>),
                                                         kConvBlock, lds) != hipSuccess || nb < 1) {
            (void)hipGetLastError();
            nb = 1;
        }
        return nb;
    }();
    static const int cus = [] {
        int d = 0, c = 0;
        if (hipGetDevice(&d) != hipSuccess || hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess ||
            c < 1) {
            (void)hipGetLastError();
            c = 256;
        }
        return c;
    }();
    const long cap = (long)cus * per_cu;
    int best_n = 1;
    long best = -1;
    for (int n = 1; n <= D; ++n) {
        const int c = (D + n - 1) / n;
        if (n > 1 && c == (D + n - 2) / (n - 1))
            continue;
        const long nch = (D + c - 1) / c;
        const long cost = ((tiles * nch * ycnt + cap - 1) / cap) * (c + 2);
        if (best < 0 || cost < best) {
            best = cost;
            best_n = n;
        }
    }
    const int chunk = (D + best_n - 1) / best_n;
    const long nch = (D + chunk - 1) / chunk;
    hipLaunchKernelGGL((conv3d_roll_k<CI_T, CO_T, BI>), dim3((unsigned)(tiles * nch), ycnt), dim3(kConvBlock), lds, s,
                       D, H, W, Cout, chunk, x, w, bias, skip, out, out_f32, relu);
    return kStatusOk;
}

template <int CI_T, int CO_T, int BI, int BJ>
static void launch_brick(int D, int H, int W, int Cout, const uint16_t *x, const uint16_t *w, const float *bias,
                         const uint16_t *skip, uint16_t *out, float *out_f32, int relu, hipStream_t s)
{
    constexpr int Cin = CI_T * 16;
    const unsigned nb = (unsigned)(((W + 15) / 16) * ((H + BI - 1) / BI) * ((D + BJ - 1) / BJ));
    const size_t lds = ((size_t)18 * (BI + 2) * (BJ + 2) * (Cin + 4) + (size_t)27 * CO_T * 16 * (Cin + 4)) *
                       sizeof(uint16_t);
    hipLaunchKernelGGL((conv3d_brick_k<CI_T, CO_T, BI, BJ>), dim3(nb, (unsigned)(Cout / (16 * CO_T))),
                       dim3(kConvBlock), lds, s, D, H, W, Cout, x, w, bias, skip, out, out_f32, relu);
}

// Partial weight gradients: part[p][tap][co][ci] = sum over this wave's voxel
// chunks of dy[v][co] * x[v + off(tap)][ci].  grid = (n_parts, 27).
template <int CO_T, int CI_T>
__global__ __launch_bounds__(kConvBlock) void conv3d_wgrad_k(int D, int H, int W, const uint16_t *__restrict__ x,
                                                           const uint16_t *__restrict__ dy,
                                                           float *__restrict__ part, long chunks_per_wave)
{
    constexpr int Cout = CO_T * 16, Cin = CI_T * 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = blockIdx.y;
    const int ddy = t / 9 - 1, ddx = (t / 3) % 3 - 1, ddz = t % 3 - 1;
    const long N = (long)D * H * W;
    const long nchunks = (N + 15) / 16;
    const long wid = (long)blockIdx.x * 4 + wave;
    const int row = lane & 15, kg = lane >> 4;
    f4 acc[CO_T][CI_T];
#pragma unroll
    for (int a = 0; a < CO_T; ++a)
#pragma unroll
        for (int b = 0; b < CI_T; ++b)
            acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
    const long c0 = wid * chunks_per_wave;
    for (long ch = c0; ch < c0 + chunks_per_wave && ch < nchunks; ++ch) {
        s4 av[CO_T], bv[CI_T];
#pragma unroll
        for (int a = 0; a < CO_T; ++a)
            av[a] = s4{0, 0, 0, 0};
#pragma unroll
        for (int b = 0; b < CI_T; ++b)
            bv[b] = s4{0, 0, 0, 0};
#pragma unroll
        for (int ii = 0; ii < 4; ++ii) {
            const long v = ch * 16 + 4 * kg + ii;
            if (v >= N)
                continue;
            const int k = (int)(v % W), i = (int)((v / W) % H), j = (int)(v / ((long)W * H));
            const int jj = j + ddy, i2 = i + ddx, kk = k + ddz;
            const bool ok = jj >= 0 && jj < D && i2 >= 0 && i2 < H && kk >= 0 && kk < W;
            const long nb = ((long)jj * H + i2) * W + kk;
#pragma unroll
            for (int a = 0; a < CO_T; ++a)
                av[a][ii] = (short)dy[v * Cout + a * 16 + row];
            if (ok) {
#pragma unroll
                for (int b = 0; b < CI_T; ++b)
                    bv[b][ii] = (short)x[nb * Cin + b * 16 + row];
            }
        }
#pragma unroll
        for (int a = 0; a < CO_T; ++a)
#pragma unroll
            for (int b = 0; b < CI_T; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    float *P = part + (wid * 27 + t) * (long)(Cout * Cin);
#pragma unroll
    for (int a = 0; a < CO_T; ++a)
#pragma unroll
        for (int b = 0; b < CI_T; ++b)
#pragma unroll
            for (int ii = 0; ii < 4; ++ii)
                P[(a * 16 + 4 * kg + ii) * Cin + b * 16 + row] = acc[a][b][ii];
}

// out[i] = sum_p part[p][i], fixed order (deterministic)
__global__ void reduce_parts_k(long n, long nparts, const float *__restrict__ part, float *__restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float s = 0.f;
    for (long p = 0; p < nparts; ++p)
        s += part[p * n + i];
    out[i] = s;
}

// db[co] partials: grid-stride over voxels, per-block partial per channel
__global__ __launch_bounds__(256) void bias_grad_k(long N, int C, const uint16_t *__restrict__ dy,
                                                   float *__restrict__ part)
{
    // thread layout: channel = threadIdx.x % C (C divides 256 for C in {16, 32, 64})
    const int c = threadIdx.x % C, lanes = 256 / C, sub = threadIdx.x / C;
    float s = 0.f;
    for (long v = (long)blockIdx.x * lanes + sub; v < N; v += (long)gridDim.x * lanes)
        s += bf2f(dy[v * C + c]);
    __shared__ float sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    if (sub == 0) {
        float t = 0.f;
        for (int q = 0; q < lanes; ++q)
            t += sh[q * C + c];
        part[(long)blockIdx.x * C + c] = t;
    }
}

// wt[t][ci][co] = w[26 - t][co][ci]   (weights of the input adjoint)
__global__ void flip_transpose_k(int Cout, int Cin, const uint16_t *__restrict__ w, uint16_t *__restrict__ wt)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = 27L * Cout * Cin;
    if (i >= n)
        return;
    const int t = (int)(i / ((long)Cout * Cin));
    const int r = (int)(i % ((long)Cout * Cin));
    const int ci = r / Cout, co = r % Cout;
    wt[i] = w[((long)(26 - t) * Cout + co) * Cin + ci];
}

// g = dy where y > 0 else 0 (ReLU backward through the stored activation)
__global__ void relu_mask_k(long n, const uint16_t *__restrict__ y, const uint16_t *__restrict__ dy,
                            uint16_t *__restrict__ g)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        g[i] = bf2f(y[i]) > 0.f ? dy[i] : (uint16_t)0;
}

// a += b (bf16, FP32 add)
__global__ void add_bf16_k(long n, uint16_t *__restrict__ a, const uint16_t *__restrict__ b)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        a[i] = f2bf(bf2f(a[i]) + bf2f(b[i]));
}

// Vertex-centred transfers on channels-last BF16 volumes.
// restrict: out[c] = scale * sum_{3^3} 2^-dist x[2c + d]
__global__ void vol_restrict_k(int D, int H, int W, int C, const uint16_t *__restrict__ x,
                               uint16_t *__restrict__ out, float scale)
{
    const int Dc = (D - 1) / 2 + 1, Hc = (H - 1) / 2 + 1, Wc = (W - 1) / 2 + 1;
    const long n = (long)Dc * Hc * Wc * C;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n)
        return;
    const int c = (int)(idx % C);
    const long v = idx / C;
    const int kc = (int)(v % Wc), ic = (int)((v / Wc) % Hc), jc = (int)(v / ((long)Wc * Hc));
    float s = 0.f;
    for (int dj = -1; dj <= 1; ++dj)
        for (int di = -1; di <= 1; ++di)
            for (int dk = -1; dk <= 1; ++dk) {
                const int j = 2 * jc + dj, i = 2 * ic + di, k = 2 * kc + dk;
                if (j < 0 || j >= D || i < 0 || i >= H || k < 0 || k >= W)
                    continue;
                const float wgt = 1.f / (float)(1 << (abs(dj) + abs(di) + abs(dk)));
                s += wgt * bf2f(x[(((long)j * H + i) * W + k) * C + c]);
            }
    out[idx] = f2bf(s * scale);
}

// prolong: out[f] (+)= scale * trilinear interpolation of the coarse volume
__global__ void vol_prolong_k(int D, int H, int W, int C, const uint16_t *__restrict__ xc, uint16_t *__restrict__ out,
                              float scale, int accumulate)
{
    const int Hc = (H - 1) / 2 + 1, Wc = (W - 1) / 2 + 1;
    const long n = (long)D * H * W * C;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n)
        return;
    const int c = (int)(idx % C);
    const long v = idx / C;
    const int k = (int)(v % W), i = (int)((v / W) % H), j = (int)(v / ((long)W * H));
    const float wgt = 1.f / (float)(1 << ((j & 1) + (i & 1) + (k & 1)));
    float s = 0.f;
    for (int cj = j / 2; cj <= (j + 1) / 2; ++cj)
        for (int ci = i / 2; ci <= (i + 1) / 2; ++ci)
            for (int ck = k / 2; ck <= (k + 1) / 2; ++ck)
                s += wgt * bf2f(xc[(((long)cj * Hc + ci) * Wc + ck) * C + c]);
    s *= scale;
    out[idx] = f2bf(accumulate ? bf2f(out[idx]) + s : s);
}

// The same transfers with one thread per voxel and group of 8 channels:
// 16-byte loads and stores, 32-bit index arithmetic, and per channel the
// sums of vol_restrict_k / vol_prolong_k in the same order.
__global__ void vol_restrict8_k(int D, int H, int W, int C, const uint16_t *__restrict__ x, uint16_t *__restrict__ out,
                                float scale)
{
    const int Dc = (D - 1) / 2 + 1, Hc = (H - 1) / 2 + 1, Wc = (W - 1) / 2 + 1, G = C / 8;
    const int n = Dc * Hc * Wc * G;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n)
        return;
    const int g = idx % G, v = idx / G;
    const int kc = v % Wc, ic = (v / Wc) % Hc, jc = v / (Wc * Hc);
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int dj = -1; dj <= 1; ++dj)
        for (int di = -1; di <= 1; ++di)
            for (int dk = -1; dk <= 1; ++dk) {
                const int j = 2 * jc + dj, i = 2 * ic + di, k = 2 * kc + dk;
                if (j < 0 || j >= D || i < 0 || i >= H || k < 0 || k >= W)
                    continue;
                const float wgt = 1.f / (float)(1 << (abs(dj) + abs(di) + abs(dk)));
                const us8 a = *reinterpret_cast<const us8 *>(x + ((long)(j * H + i) * W + k) * C + 8 * g);
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    s[q] += wgt * bf2f(a[q]);
            }
    us8 o;
#pragma unroll
    for (int q = 0; q < 8; ++q)
        o[q] = f2bf(s[q] * scale);
    *reinterpret_cast<us8 *>(out + (long)idx * 8) = o;
}

__global__ void vol_prolong8_k(int D, int H, int W, int C, const uint16_t *__restrict__ xc, uint16_t *__restrict__ out,
                               float scale, int accumulate)
{
    const int Hc = (H - 1) / 2 + 1, Wc = (W - 1) / 2 + 1, G = C / 8;
    const int n = D * H * W * G;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n)
        return;
    const int g = idx % G, v = idx / G;
    const int k = v % W, i = (v / W) % H, j = v / (W * H);
    const float wgt = 1.f / (float)(1 << ((j & 1) + (i & 1) + (k & 1)));
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int cj = j / 2; cj <= (j + 1) / 2; ++cj)
        for (int ci = i / 2; ci <= (i + 1) / 2; ++ci)
            for (int ck = k / 2; ck <= (k + 1) / 2; ++ck) {
                const us8 a = *reinterpret_cast<const us8 *>(xc + ((long)(cj * Hc + ci) * Wc + ck) * C + 8 * g);
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    s[q] += wgt * bf2f(a[q]);
            }
    uint16_t *o = out + (long)idx * 8;
    us8 prev = accumulate ? *reinterpret_cast<const us8 *>(o) : us8{0, 0, 0, 0, 0, 0, 0, 0};
    us8 r;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const float t = s[q] * scale;
        r[q] = f2bf(accumulate ? bf2f(prev[q]) + t : t);
    }
    *reinterpret_cast<us8 *>(o) = r;
}

// Adam on FP32 master weights; also refreshes the BF16 copy used by the kernels.
__global__ void adam_k(long n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, uint16_t *__restrict__ pbf, float lr, float b1, float b2, float eps,
                       float bc1, float bc2)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float upd = lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    const float pi = p[i] - upd;
    p[i] = pi;
    if (pbf)
        pbf[i] = f2bf(pi);
}

__global__ void f32_to_bf16_k(long n, const float *__restrict__ a, uint16_t *__restrict__ b)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        b[i] = f2bf(a[i]);
}

// Network input: x[v][0..2] = r[3v + c] * inv_s, x[v][3] = rho_node[v],
// channels 4..15 zero (C = 16).
__global__ void net_input_k(long nv, const double *__restrict__ r, const float *__restrict__ rho_node,
                            const double *__restrict__ scal, int slot, uint16_t *__restrict__ x)
{
    // one voxel per thread: its 16 channels as two 16-byte stores
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv)
        return;
    const double inv_s = 1.0 / sqrt(fmax(scal[slot], 1e-300) / (double)(3 * nv));
    us8 lo = us8{0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 3; ++c)
        lo[c] = f2bf((float)(r[3 * v + c] * inv_s));
    lo[3] = f2bf(rho_node ? rho_node[v] : 0.f);
    *reinterpret_cast<us8 *>(x + 16 * v) = lo;
    *reinterpret_cast<us8 *>(x + 16 * v + 8) = us8{0, 0, 0, 0, 0, 0, 0, 0};
}

// z[3v + c] = s * y[v][c] (c < 3, y has C = 16 channels), s = RMS of r
__global__ void net_output_k(long nv, const float *__restrict__ y, const double *__restrict__ scal, int slot,
                             double *__restrict__ z, int accumulate)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nv * 3)
        return;
    const long v = idx / 3;
    const int c = (int)(idx % 3);
    const double s = sqrt(fmax(scal[slot], 0.0) / (double)(3 * nv));
    const double val = s * (double)y[v * 16 + c];
    z[idx] = accumulate ? z[idx] + val : val;
}

// MSE loss against the target t[3v + c] scaled by 1/s: dy[v][c] = 2 (y - t/s) / (3 nv)
// (c < 3; channels 3..15 get 0); part[block] = sum of squared errors
__global__ __launch_bounds__(256) void mse_k(long nv, const float *__restrict__ y, const double *__restrict__ t,
                                            const double *__restrict__ scal, int slot, uint16_t *__restrict__ dy,
                                            double *__restrict__ part)
{
    const double inv_s = 1.0 / sqrt(fmax(scal[slot], 1e-300) / (double)(3 * nv));
    double acc = 0.0;
    for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < nv * 16; idx += (long)gridDim.x * 256) {
        const long v = idx / 16;
        const int c = (int)(idx % 16);
        float g = 0.f;
        if (c < 3) {
            const double e = (double)y[idx] - t[3 * v + c] * inv_s;
            acc += e * e;
            g = (float)(2.0 * e / (double)(3 * nv));
        }
        dy[idx] = f2bf(g);
    }
    acc = block_sum<256>(acc);
    if (threadIdx.x == 0)
        part[blockIdx.x] = acc;
}

// nodal density: mean of the moduli-free design values of the elements
// around each node (hex8 grid, node order of the mesh)
__global__ void node_density_k(Grid g, const double *__restrict__ x, float *__restrict__ rn)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.n_nodes)
        return;
    const int k = (int)(n % g.nz);
    const long t = n / g.nz;
    const int i = (int)(t % g.nx), j = (int)(t / g.nx);
    double s = 0.0;
    int m = 0;
    for (int oy = 0; oy < 2; ++oy)
        for (int ox = 0; ox < 2; ++ox)
            for (int oz = 0; oz < 2; ++oz) {
                const int ei = i - ox, ej = j - oy, ek = k - oz;
                if (ei < 0 || ei >= g.nelx || ej < 0 || ej >= g.nely || ek < 0 || ek >= g.nelz)
                    continue;
                s += x[((long)ei * g.nely + ej) * g.nelz + ek];
                m++;
            }
    rn[n] = (float)(s / m);
}

static unsigned blk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace fantom

using namespace fantom;

extern "C" {

int fantom_conv3d_fwd_bf16(int D, int H, int W, int Cin, int Cout, const uint16_t *x, const uint16_t *w,
                           const float *bias, const uint16_t *skip, uint16_t *out, float *out_f32, int relu,
                           void *stream)
{
    FANTOM_CHECK_ARG(D > 0 && H > 0 && W > 0 && x && w && out, "bad operands");
    FANTOM_CHECK_ARG(Cin % 16 == 0 && Cin >= 16 && Cin <= 256, "Cin must be a multiple of 16 in [16, 256]");
    const long N = (long)D * H * W;
    hipStream_t s = as_stream(stream);
    // planes walked along D with an LDS ring (large volumes)
    if (N >= (1L << 16)) {
        int st = -1;
        if (Cin == 16 && Cout == 16)
            st = launch_roll<1, 1, 16>(D, H, W, Cout, x, w, bias, skip, out, out_f32, relu, s);
        else if (Cin == 16 && Cout == 32)
            st = launch_roll<1, 2, 8>(D, H, W, Cout, x, w, bias, skip, out, out_f32, relu, s);
        else if (Cin == 32 && Cout == 16)
            st = launch_roll<2, 1, 8>(D, H, W, Cout, x, w, bias, skip, out, out_f32, relu, s);
        else if (Cin == 32 && (Cout == 32 || Cout == 64))
            st = launch_roll<2, 1, 8>(D, H, W, Cout, x, w, bias, skip, out, out_f32, relu, s);
        else if (Cin == 64 && Cout % 16 == 0 && Cout <= 64)
            st = launch_roll<4, 1, 8>(D, H, W, Cout, x, w, bias, skip, out, out_f32, relu, s);
        if (st >= 0) {
            if (st != kStatusOk)
                return st;
            FANTOM_LAUNCH_CHECK("conv3d_fwd");
            return kStatusOk;
        }
    }
    // input bricks and weights in LDS (<= 64 KB) for Cin = 16 (the 32-channel
    // brick measured slower than the weight-staged kernel below)
    if (Cin == 16 && (Cout == 16 || Cout == 32)) {
        if (Cout == 16)
            launch_brick<1, 1, 8, 2>(D, H, W, Cout, x, w, bias, skip, out, out_f32, relu, s);
        else
            launch_brick<1, 2, 8, 2>(D, H, W, Cout, x, w, bias, skip, out, out_f32, relu, s);
        FANTOM_LAUNCH_CHECK("conv3d_fwd");
        return kStatusOk;
    }
    // LDS-staged weights for Cin <= 64: (CI_T, CO_W) with the slice <= 64 KB
    {
        const int ci_t = Cin / 16;
        const int co_w = (ci_t == 4 || Cout == 16) ? 1 : 2;
        if (ci_t <= 4 && ci_t != 3 && (Cout == 16 || Cout == 32 || Cout == 64)) {
            const long groups = (N + 16 * kFwdRT - 1) / (16 * kFwdRT);
            const unsigned gx = (unsigned)std::min<long>((groups + 3) / 4, 2048);
            const dim3 grid(gx, (unsigned)(Cout / (16 * co_w)));
            const size_t lds = (size_t)27 * 16 * co_w * fwd2_row(Cin) * sizeof(uint16_t);
#define FWD2(CI, CO)                                                                                       \
    if (ci_t == CI && co_w == CO) {                                                                        \
        hipLaunchKernelGGL((conv3d_fwd2_k<CI, CO>), grid, dim3(kConvBlock), lds, s, D, H, W, Cout, x, w, bias, \
                           skip, out, out_f32, relu);                                                      \
        FANTOM_LAUNCH_CHECK("conv3d_fwd");                                                                 \
        return kStatusOk;                                                                                  \
    }
            FWD2(1, 1) FWD2(1, 2) FWD2(2, 1) FWD2(2, 2) FWD2(4, 1)
#undef FWD2
        }
    }
    const unsigned blocks = (unsigned)((N + 4 * 16 * kRowTiles - 1) / (4 * 16 * kRowTiles));
    switch (Cout) {
    case 16:
        hipLaunchKernelGGL(conv3d_fwd_k<1>, dim3(blocks), dim3(kConvBlock), 0, s, D, H, W, Cin, x, w, bias, skip, out,
                           out_f32, relu);
        break;
    case 32:
        hipLaunchKernelGGL(conv3d_fwd_k<2>, dim3(blocks), dim3(kConvBlock), 0, s, D, H, W, Cin, x, w, bias, skip, out,
                           out_f32, relu);
        break;
    case 64:
        hipLaunchKernelGGL(conv3d_fwd_k<4>, dim3(blocks), dim3(kConvBlock), 0, s, D, H, W, Cin, x, w, bias, skip, out,
                           out_f32, relu);
        break;
    default:
        set_error("Cout must be 16, 32 or 64 (got %d)", Cout);
        return kStatusBadArg;
    }
    FANTOM_LAUNCH_CHECK("conv3d_fwd");
    return kStatusOk;
}

long fantom_conv3d_wgrad_parts(int D, int H, int W, int Cin, int Cout, long *chunks_per_wave)
{
    const long N = (long)D * H * W;
    const long nchunks = (N + 15) / 16;
    long waves = 256;  // 64 blocks of 4 waves per tap
    long cpw = (nchunks + waves - 1) / waves;
    if (chunks_per_wave)
        *chunks_per_wave = cpw;
    return waves;  // partial slices, each 27 * Cout * Cin floats
}

int fantom_conv3d_wgrad_bf16(int D, int H, int W, int Cin, int Cout, const uint16_t *x, const uint16_t *dy,
                             float *part, float *dw, void *stream)
{
    FANTOM_CHECK_ARG(D > 0 && H > 0 && W > 0 && x && dy && part && dw, "bad operands");
    long cpw = 0;
    const long waves = fantom_conv3d_wgrad_parts(D, H, W, Cin, Cout, &cpw);
    const dim3 grid((unsigned)(waves / 4), 27);
    hipStream_t s = as_stream(stream);
#define WG(CO, CI)                                                                                         \
    if (Cout == CO * 16 && Cin == CI * 16) {                                                               \
        hipLaunchKernelGGL((conv3d_wgrad_k<CO, CI>), grid, dim3(kConvBlock), 0, s, D, H, W, x, dy, part, cpw); \
    } else
    WG(1, 1) WG(2, 1) WG(1, 2) WG(2, 2) WG(4, 2) WG(2, 4) WG(4, 4) WG(1, 4) WG(4, 1)
    {
        set_error("unsupported channel pair %d -> %d", Cin, Cout);
        return kStatusBadArg;
    }
#undef WG
    const long n = 27L * Cout * Cin;
    hipLaunchKernelGGL(reduce_parts_k, dim3(blk(n)), dim3(256), 0, s, n, waves, part, dw);
    FANTOM_LAUNCH_CHECK("conv3d_wgrad");
    return kStatusOk;
}

int fantom_bias_grad_bf16(long N, int C, const uint16_t *dy, float *part, float *db, void *stream)
{
    FANTOM_CHECK_ARG(N > 0 && (C == 16 || C == 32 || C == 64) && dy && part && db, "bad operands");
    const int nb = 256;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(bias_grad_k, dim3(nb), dim3(256), 0, s, N, C, dy, part);
    hipLaunchKernelGGL(reduce_parts_k, dim3(1), dim3(64), 0, s, (long)C, (long)nb, part, db);
    FANTOM_LAUNCH_CHECK("bias_grad");
    return kStatusOk;
}

int fantom_flip_transpose_bf16(int Cout, int Cin, const uint16_t *w, uint16_t *wt, void *stream)
{
    FANTOM_CHECK_ARG(w && wt && w != wt, "bad operands");
    const long n = 27L * Cout * Cin;
    hipLaunchKernelGGL(flip_transpose_k, dim3(blk(n)), dim3(256), 0, as_stream(stream), Cout, Cin, w, wt);
    FANTOM_LAUNCH_CHECK("flip_transpose");
    return kStatusOk;
}

int fantom_relu_mask_bf16(long n, const uint16_t *y, const uint16_t *dy, uint16_t *g, void *stream)
{
    FANTOM_CHECK_ARG(n > 0 && y && dy && g, "bad operands");
    hipLaunchKernelGGL(relu_mask_k, dim3(blk(n)), dim3(256), 0, as_stream(stream), n, y, dy, g);
    FANTOM_LAUNCH_CHECK("relu_mask");
    return kStatusOk;
}

int fantom_add_bf16(long n, uint16_t *a, const uint16_t *b, void *stream)
{
    FANTOM_CHECK_ARG(n > 0 && a && b, "bad operands");
    hipLaunchKernelGGL(add_bf16_k, dim3(blk(n)), dim3(256), 0, as_stream(stream), n, a, b);
    FANTOM_LAUNCH_CHECK("add_bf16");
    return kStatusOk;
}

int fantom_vol_restrict_bf16(int D, int H, int W, int C, const uint16_t *x, uint16_t *out, float scale, void *stream)
{
    FANTOM_CHECK_ARG(D % 2 == 1 && H % 2 == 1 && W % 2 == 1 && x && out, "volume sizes must be odd");
    const long n = (long)((D - 1) / 2 + 1) * ((H - 1) / 2 + 1) * ((W - 1) / 2 + 1) * C;
    if (C % 8 == 0 && (long)D * H * W * C < (1L << 31))
        hipLaunchKernelGGL(vol_restrict8_k, dim3(blk(n / 8)), dim3(256), 0, as_stream(stream), D, H, W, C, x, out,
                           scale);
    else
        hipLaunchKernelGGL(vol_restrict_k, dim3(blk(n)), dim3(256), 0, as_stream(stream), D, H, W, C, x, out, scale);
    FANTOM_LAUNCH_CHECK("vol_restrict");
    return kStatusOk;
}

int fantom_vol_prolong_bf16(int D, int H, int W, int C, const uint16_t *xc, uint16_t *out, float scale, int accumulate,
                            void *stream)
{
    FANTOM_CHECK_ARG(D % 2 == 1 && H % 2 == 1 && W % 2 == 1 && xc && out, "volume sizes must be odd");
    const long n = (long)D * H * W * C;
    if (C % 8 == 0 && n < (1L << 31))
        hipLaunchKernelGGL(vol_prolong8_k, dim3(blk(n / 8)), dim3(256), 0, as_stream(stream), D, H, W, C, xc, out,
                           scale, accumulate);
    else
        hipLaunchKernelGGL(vol_prolong_k, dim3(blk(n)), dim3(256), 0, as_stream(stream), D, H, W, C, xc, out, scale,
                           accumulate);
    FANTOM_LAUNCH_CHECK("vol_prolong");
    return kStatusOk;
}

int fantom_adam_f32(long n, float *p, const float *g, float *m, float *v, uint16_t *pbf, float lr, float b1, float b2,
                    float eps, int step, void *stream)
{
    FANTOM_CHECK_ARG(n > 0 && p && g && m && v && step >= 1, "bad operands");
    const float bc1 = 1.f - powf(b1, (float)step), bc2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_k, dim3(blk(n)), dim3(256), 0, as_stream(stream), n, p, g, m, v, pbf, lr, b1, b2, eps, bc1,
                       bc2);
    FANTOM_LAUNCH_CHECK("adam");
    return kStatusOk;
}

int fantom_f32_to_bf16(long n, const float *a, uint16_t *b, void *stream)
{
    FANTOM_CHECK_ARG(n > 0 && a && b, "bad operands");
    hipLaunchKernelGGL(f32_to_bf16_k, dim3(blk(n)), dim3(256), 0, as_stream(stream), n, a, b);
    FANTOM_LAUNCH_CHECK("f32_to_bf16");
    return kStatusOk;
}

int fantom_net_input(long nv, const double *r, const float *rho_node, const double *scal, int slot, uint16_t *x,
                     void *stream)
{
    FANTOM_CHECK_ARG(nv > 0 && r && scal && x && slot >= 0 && slot < 16, "bad operands");
    hipLaunchKernelGGL(net_input_k, dim3(blk(nv)), dim3(256), 0, as_stream(stream), nv, r, rho_node, scal, slot, x);
    FANTOM_LAUNCH_CHECK("net_input");
    return kStatusOk;
}

int fantom_net_output(long nv, const float *y, const double *scal, int slot, double *z, int accumulate, void *stream)
{
    FANTOM_CHECK_ARG(nv > 0 && y && scal && z && slot >= 0 && slot < 16, "bad operands");
    hipLaunchKernelGGL(net_output_k, dim3(blk(nv * 3)), dim3(256), 0, as_stream(stream), nv, y, scal, slot, z,
                       accumulate);
    FANTOM_LAUNCH_CHECK("net_output");
    return kStatusOk;
}

int fantom_mse_bf16(long nv, const float *y, const double *t, const double *scal, int slot, uint16_t *dy,
                    double *part, void *stream)
{
    FANTOM_CHECK_ARG(nv > 0 && y && t && scal && dy && part, "bad operands");
    hipLaunchKernelGGL(mse_k, dim3(256), dim3(256), 0, as_stream(stream), nv, y, t, scal, slot, dy, part);
    FANTOM_LAUNCH_CHECK("mse");
    return kStatusOk;
}

int fantom_node_density(int nelx, int nely, int nelz, const double *x, float *rn, void *stream)
{
    FANTOM_CHECK_ARG(grid_ok(nelx, nely, nelz) && nelz > 0 && x && rn, "bad operands");
    const Grid g = make_grid(nelx, nely, nelz);
    hipLaunchKernelGGL(node_density_k, dim3(blk(g.n_nodes)), dim3(256), 0, as_stream(stream), g, x, rn);
    FANTOM_LAUNCH_CHECK("node_density");
    return kStatusOk;
}

}  // extern "C"
