// Element-tiled matrix-free K(E)u for structured hex8 grids in the
// Walsh-Hadamard block form (FP64).
//
// Why this form: the dense 24x24 contraction costs 1152 flop per element
// against ~56 B of compulsory HBM traffic, which is FP64-compute-bound on
// MI355X (ridge ~9.8 flop/B).  For a box hex8 the reflection group of the
// element commutes with K0, so in the basis
//   v_{s,c}[corner b, comp c] = (-1)^{popc(b & (s ^ (1<<c)))}
// (b = bx + 2by + 4bz) K0 is block-diagonal with eight 3x3 blocks B_s.
// K0 u_e = WHT -> eight 3x3 products -> WHT: 144 add/sub + 72 FMA + 24 mul
// per element instead of 576 FMA.  The blocks are derived and checked on the
// host (fantom_hex8_blocks).
//
// Tiling: a workgroup of 256 threads owns a 16 (x) x 16 (z) column of
// elements and walks it along y (the slowest node axis, so each node plane is
// contiguous in HBM).  Per element layer:
//   1. every thread computes one element's 24 outputs from the u node plane
//      pair held in LDS (structure-of-arrays, conflict-free),
//   2. outputs go to LDS; after a barrier each thread gathers the four
//      in-plane elements of its node in a fixed order (deterministic, no
//      atomics), adds the upper-plane contributions of the previous layer kept
//      in registers, applies the Dirichlet/epilogue rules and writes the node,
//   3. the next node plane, prefetched into registers at the start of the
//      step, is stored into the LDS ring slot that just went dead.
// Tiles overlap by one element so that 15 x 15 nodes per tile are complete;
// y is split into chunks of CHUNK node planes to give >> 256 workgroups, and
// the blockIdx -> tile map keeps neighbouring tiles on the same XCD so their
// shared halo rows hit the same L2.
#include "common.hpp"
#include "fantom_internal.hpp"

namespace fantom {

constexpr int TB = 16;            // elements per tile edge (x and z)
constexpr int TN = TB + 1;        // nodes per tile edge
constexpr int TOUT = TB - 1;      // output nodes per tile edge
constexpr int PLANE = TN * TN;    // nodes per tile plane
constexpr int CHUNK = 16;         // node planes per workgroup
constexpr int kTileBlock = TB * TB;
constexpr int kPrefetch = (3 * PLANE + kTileBlock - 1) / kTileBlock;   // doubles per thread
constexpr int kPrefetchM = (PLANE + kTileBlock - 1) / kTileBlock;      // mask bytes per thread

template <typename T>
struct HexBlocks {
    T b[72];  // B_s[c][d] at b[9*s + 3*c + d], scaled by 1/8
};

template <typename T>
__device__ __forceinline__ void wht8(T *w)
{
#pragma unroll
    for (int h = 1; h < 8; h <<= 1) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if ((i & h) == 0) {
                const T a = w[i], b = w[i | h];
                w[i] = a + b;
                w[i | h] = a - b;
            }
        }
    }
}

template <typename T, int MODE>
__global__ __launch_bounds__(kTileBlock, 2) void ku3d_tile(Grid g, HexBlocks<T> B, const T *__restrict__ E,
                                                         const uint8_t *__restrict__ mask, const T *__restrict__ u,
                                                         T *__restrict__ y, const T *__restrict__ bvec,
                                                         const T *__restrict__ dinv, T omega,
                                                         double *__restrict__ partials, int tiles_x, int tiles_z,
                                                         int chunks_y)
{
    __shared__ T us[2][3][PLANE];        // u node planes, SoA, ring of 2
    __shared__ uint8_t ms[2][PLANE];     // constraint bits of those planes
    __shared__ T outs[24][kTileBlock];   // element outputs, SoA

    // XCD-aware tile order: the hardware deals workgroups to the 8 XCDs
    // round-robin, so give each XCD a contiguous range of logical tiles.  The
    // grid is a multiple of 8 blocks, which makes this map a bijection;
    // logical ids past the last tile do no work.
    const int per_xcd = gridDim.x / 8;
    const int bid = blockIdx.x;
    int logical = (bid % 8) * per_xcd + bid / 8;
    double part = 0.0;
    const int tid = threadIdx.x;
    const int lx = tid / TB, lz = tid % TB;
    const bool valid_tile = logical < tiles_x * tiles_z * chunks_y;
    if (!valid_tile)
        logical = 0;
    const int tz = logical % tiles_z;
    const int tx = (logical / tiles_z) % tiles_x;
    const int cy = logical / (tiles_z * tiles_x);
    const int ex0 = tx * TOUT - 1, ez0 = tz * TOUT - 1;   // element (and node) origin of the tile
    const int jc0 = cy * CHUNK;
    const int jc1 = min(jc0 + CHUNK, g.ny);               // output node planes [jc0, jc1)
    const long nplane = (long)g.nx * g.nz;

    auto load_plane = [&](int j, T *vals, uint8_t *mv) {
#pragma unroll
        for (int q = 0; q < kPrefetch; ++q) {
            const int idx = tid + q * kTileBlock;
            T v = T(0);
            if (idx < 3 * PLANE && j >= 0 && j < g.ny) {
                const int a = idx / (3 * TN), rem = idx % (3 * TN);
                const int gx = ex0 + a, gz = ez0 + rem / 3;
                if (gx >= 0 && gx < g.nx && gz >= 0 && gz < g.nz)
                    v = u[3 * ((long)j * nplane + (long)gx * g.nz + gz) + rem % 3];
            }
            vals[q] = v;
        }
#pragma unroll
        for (int q = 0; q < kPrefetchM; ++q) {
            const int idx = tid + q * kTileBlock;
            uint8_t m = 0;
            if (mask && idx < PLANE && j >= 0 && j < g.ny) {
                const int a = idx / TN, c = idx % TN;
                const int gx = ex0 + a, gz = ez0 + c;
                if (gx >= 0 && gx < g.nx && gz >= 0 && gz < g.nz)
                    m = mask[(long)j * nplane + (long)gx * g.nz + gz];
            }
            mv[q] = m;
        }
    };
    auto store_plane = [&](int slot, const T *vals, const uint8_t *mv) {
#pragma unroll
        for (int q = 0; q < kPrefetch; ++q) {
            const int idx = tid + q * kTileBlock;
            if (idx < 3 * PLANE) {
                const int a = idx / (3 * TN), rem = idx % (3 * TN);
                us[slot][rem % 3][a * TN + rem / 3] = vals[q];
            }
        }
#pragma unroll
        for (int q = 0; q < kPrefetchM; ++q) {
            const int idx = tid + q * kTileBlock;
            if (idx < PLANE)
                ms[slot][idx] = mv[q];
        }
    };

    T pre[kPrefetch];
    uint8_t prem[kPrefetchM];
    if (valid_tile) {
        load_plane(jc0 - 1, pre, prem);
        store_plane((jc0 + 1) & 1, pre, prem);   // slot of plane p is p & 1 (p >= -1)
        load_plane(jc0, pre, prem);
        store_plane(jc0 & 1, pre, prem);
    }
    __syncthreads();

    T hi[3] = {T(0), T(0), T(0)};    // upper-plane contributions of the previous layer
    const int gxe = ex0 + lx, gze = ez0 + lz;  // this thread's element (x, z)
    const bool el_in = gxe >= 0 && gxe < g.nelx && gze >= 0 && gze < g.nelz;

    for (int L = jc0 - 1; L < jc1 && valid_tile; ++L) {
        // element layer L lies between node planes L and L+1
        const int sl = (L + 2) & 1, su = (L + 3) & 1;   // slots of planes L, L+1
        if (L + 2 < jc1 + 1)
            load_plane(L + 2, pre, prem);

        // own node (lx, lz) of plane L, unmasked, for the epilogue
        T uown[3] = {T(0), T(0), T(0)};
        int mown = 0;
        if (L >= jc0) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
                uown[c] = us[sl][c][lx * TN + lz];
            mown = ms[sl][lx * TN + lz];
        }

        // ---- 1. element outputs
        T w[3][8];
        const bool active = el_in && L >= 0 && L < g.nely;
        if (active) {
            const T Ee = E[((long)gxe * g.nely + L) * g.nelz + gze];
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const int bx = b & 1, by = (b >> 1) & 1, bz = b >> 2;
                const int sidx = by ? su : sl;
                const int p = (lx + bx) * TN + (lz + bz);
                const int m = ms[sidx][p];
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    w[c][b] = ((m >> c) & 1) ? T(0) : us[sidx][c][p];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c)
                wht8(w[c]);
            T z[3][8];
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const T h0 = w[0][s ^ 1], h1 = w[1][s ^ 2], h2 = w[2][s ^ 4];
                const T *Bs = B.b + 9 * s;
                z[0][s ^ 1] = Ee * (Bs[0] * h0 + Bs[1] * h1 + Bs[2] * h2);
                z[1][s ^ 2] = Ee * (Bs[3] * h0 + Bs[4] * h1 + Bs[5] * h2);
                z[2][s ^ 4] = Ee * (Bs[6] * h0 + Bs[7] * h1 + Bs[8] * h2);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                wht8(z[c]);
#pragma unroll
                for (int b = 0; b < 8; ++b)
                    outs[b * 3 + c][tid] = z[c][b];
            }
        } else {
#pragma unroll
            for (int v = 0; v < 24; ++v)
                outs[v][tid] = T(0);
        }
        __syncthreads();

        // ---- 2. node gather for plane L (lower corners) and L+1 (upper)
        const bool node_thread = lx >= 1 && lz >= 1;
        if (node_thread) {
            T lo[3] = {T(0), T(0), T(0)}, up[3] = {T(0), T(0), T(0)};
#pragma unroll
            for (int bx = 0; bx < 2; ++bx)
#pragma unroll
                for (int bz = 0; bz < 2; ++bz) {
                    const int t = (lx - bx) * TB + (lz - bz);
                    const int bl = bx + 4 * bz, bu = bl + 2;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        lo[c] += outs[bl * 3 + c][t];
                        up[c] += outs[bu * 3 + c][t];
                    }
                }
            if (L >= jc0) {
                const int gx = ex0 + lx, gz = ez0 + lz;
                if (gx < g.nx && gz < g.nz) {
                    const long n = (long)L * nplane + (long)gx * g.nz + gz;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const long r = 3 * n + c;
                        const T ku = ((mown >> c) & 1) ? uown[c] : hi[c] + lo[c];
                        T o;
                        if (MODE == MODE_JAC)
                            o = uown[c] + omega * dinv[r] * (bvec[r] - ku);
                        else if (MODE == MODE_RES)
                            o = bvec[r] - ku;
                        else
                            o = ku;
                        y[r] = o;
                        if (MODE == MODE_KUDOT)
                            part += (double)uown[c] * (double)o;
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 3; ++c)
                hi[c] = up[c];
        }
        // ---- 3. the slot of plane L is dead: store plane L+2 there
        if (L + 2 < jc1 + 1)
            store_plane(sl, pre, prem);
        __syncthreads();
    }
    if (MODE == MODE_KUDOT) {
        const double s = block_sum<kTileBlock>(part);
        if (threadIdx.x == 0)
            partials[blockIdx.x] = s;
    }
}

static void tile_grid(const Grid &g, int &tx, int &tz, int &cy)
{
    tx = (g.nx + TOUT - 1) / TOUT;
    tz = (g.nz + TOUT - 1) / TOUT;
    cy = (g.ny + CHUNK - 1) / CHUNK;
}

// launch size: the tile count rounded up to a multiple of the 8 XCDs
static unsigned tile_blocks(const Grid &g)
{
    int tx, tz, cy;
    tile_grid(g, tx, tz, cy);
    const long n = (long)tx * tz * cy;
    return (unsigned)((n + 7) / 8 * 8);
}

long ku_tile_partials(const Grid &g) { return (long)tile_blocks(g); }

template <typename T>
int launch_ku_tile_t(int mode, const Grid &g, const double *Kb, const T *E, const uint8_t *mask, const T *u, T *y,
                     const T *b, const T *dinv, double omega, double *partials, hipStream_t s)
{
    HexBlocks<T> B;
    // Kb is a host pointer: the 72 block entries travel as kernel arguments
    for (int i = 0; i < 72; ++i)
        B.b[i] = (T)Kb[i];
    int tx, tz, cy;
    tile_grid(g, tx, tz, cy);
    const unsigned blocks = tile_blocks(g);
    const T om = (T)omega;
#define FANTOM_TILE_CASE(M)                                                                                \
    case M:                                                                                                \
        hipLaunchKernelGGL((ku3d_tile<T, M>), dim3(blocks), dim3(kTileBlock), 0, s, g, B, E, mask, u, y, b, dinv, \
                           om, partials, tx, tz, cy);                                                      \
        break;
    switch (mode) {
        FANTOM_TILE_CASE(MODE_KU)
        FANTOM_TILE_CASE(MODE_JAC)
        FANTOM_TILE_CASE(MODE_RES)
        FANTOM_TILE_CASE(MODE_KUDOT)
    default:
        set_error("unknown matvec mode %d", mode);
        return kStatusBadArg;
    }
#undef FANTOM_TILE_CASE
    FANTOM_LAUNCH_CHECK("ku_tile");
    return kStatusOk;
}

int launch_ku_tile(int mode, const Grid &g, const double *Kb, const double *E, const uint8_t *mask,
                   const double *u, double *y, const double *b, const double *dinv, double omega,
                   double *partials, hipStream_t s)
{
    return launch_ku_tile_t<double>(mode, g, Kb, E, mask, u, y, b, dinv, omega, partials, s);
}

int launch_ku_tile_f32(int mode, const Grid &g, const double *Kb, const float *E, const uint8_t *mask,
                       const float *u, float *y, const float *b, const float *dinv, double omega, double *partials,
                       hipStream_t s)
{
    return launch_ku_tile_t<float>(mode, g, Kb, E, mask, u, y, b, dinv, omega, partials, s);
}

}  // namespace fantom
