// Matrix-free K(E)u for structured hex8 grids, element-tiled, in the
// Walsh-Hadamard block form (FP64, and FP32 for the inner solver of the
// mixed-precision path).
//
// Why this form: the dense 24x24 contraction costs 1152 flop per element
// against ~56 B of compulsory HBM traffic, which is FP64-compute-bound on
// MI355X (ridge ~9.8 flop/B).  For a box hex8 the reflection group of the
// element commutes with K0, so in the basis
//   v_{s,c}[corner b, comp c] = (-1)^{popc(b & (s ^ (1<<c)))}
// (b = bx + 2by + 4bz) K0 is block-diagonal with eight symmetric 3x3 blocks
// B_s, and the blocks have a fixed zero pattern for every box shape and
// Poisson ratio: 33 distinct entries, 45 products per element instead of 576
// (the three zero rows are the rigid translations).  The blocks are derived
// on the host (fantom_hex8_blocks) and the pattern is checked before packing.
// Gather/scatter and the shared WHT stages live in tile_skel.hpp.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>

#include "tile_skel.hpp"

namespace fantom {

// 33 packed block entries (scaled by 1/8), order:
//   s0: 00 01 02 11 12 22 | s1: 11 12 22 | s2: 00 02 22 | s3: 00 01 11 22
//   s4: 00 01 11 | s5: 00 02 11 22 | s6: 00 11 12 22 | s7: 00 01 02 11 12 22
constexpr int kPackedBlocks = 33;
static const int kPackS[kPackedBlocks] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 3, 4,
                                          4, 4, 5, 5, 5, 5, 6, 6, 6, 6, 7, 7, 7, 7, 7, 7};
static const int kPackRC[kPackedBlocks] = {0, 1, 2, 4, 5, 8, 4, 5, 8, 0, 2, 8, 0, 1, 4, 8, 0,
                                           1, 4, 0, 2, 4, 8, 0, 4, 5, 8, 0, 1, 2, 4, 5, 8};

template <typename T>
struct FineOp {
    // resident waves per SIMD the kernels are built for (FP64 152-169 VGPRs:
    // 3; FP32 76-95: 5)
    static constexpr int kWavesPerEU = sizeof(T) == 8 ? 3 : 5;
    // one wave fewer where the budget above would spill inside the layer
    // loop: the prolongating sweep (10-19 spilled values; 127 / 67 us
    // against 180 / 94 us, FP64 / FP32, 256x128x128) and the FP32 zero-guess
    // sweep with staged stores (60.3 us against 64.7 unstaged at five waves).
    // Not the sweeps that sum partials: K.u + p.Kp shares their chunk shape.
    static constexpr int waves_for(int mode)
    {
        return (mode == MODE_JACPRO || (sizeof(T) == 4 && mode == MODE_JAC0OWN)) ? kWavesPerEU - 1 : kWavesPerEU;
    }
    // staged row stores where they pay (measured per mode, 256x128x128,
    // HBM-resident): not in the Jacobi + b.x sweeps (slower by 2-3 us with
    // them, and by 5-22 us at one wave fewer without spills), nor in the FP32
    // Jacobi sweep (no gain), nor in the prolongating sweep (registers).  The
    // FP64 zero-guess sweep gains 7 us with them (its spill lies outside the
    // loop), the FP32 one 4 us at one wave fewer.
    static constexpr bool staged_store(int mode)
    {
        return mode != MODE_JACDOT && mode != MODE_JACPRO && (sizeof(T) == 8 || mode != MODE_JACOWN);
    }
    // apply leaves out the element modulus; the skeleton multiplies it in
    static constexpr bool kScaled = true;
    // every corner of a box hex8 has the same diagonal entry per component,
    // d_c = K0[3b + c][3b + c]; rd holds 1 / d_c
    static constexpr bool kDiagFromModuli = true;
    T c[kPackedBlocks];
    T rd[3];
    const T *E;
    struct Data {
        T e;
    };
    __device__ __forceinline__ void load(Data &d, const Grid &g, bool active, int ex, int ey, int ez) const
    {
        d.e = active ? E[((long)ex * g.nely + ey) * g.nelz + ez] : T(0);
    }
    __device__ __forceinline__ T scale(const Data &d) const { return d.e; }
    __device__ __forceinline__ T inv_diag(int comp, T inv_sum) const { return rd[comp] * inv_sum; }
    // block s reads h_k = w[k][s ^ (1 << k)] and writes z[k][s ^ (1 << k)]
    // (without the factor d.e)
    __device__ __forceinline__ void apply(const Data &d, const T (&w)[3][8], T (&z)[3][8]) const
    {
        {   // s = 0
            const T h0 = w[0][1], h1 = w[1][2], h2 = w[2][4];
            z[0][1] = c[0] * h0 + c[1] * h1 + c[2] * h2;
            z[1][2] = c[1] * h0 + c[3] * h1 + c[4] * h2;
            z[2][4] = c[2] * h0 + c[4] * h1 + c[5] * h2;
        }
        {   // s = 1 (row/column 0 vanish)
            const T h1 = w[1][3], h2 = w[2][5];
            z[0][0] = T(0);
            z[1][3] = c[6] * h1 + c[7] * h2;
            z[2][5] = c[7] * h1 + c[8] * h2;
        }
        {   // s = 2
            const T h0 = w[0][3], h2 = w[2][6];
            z[0][3] = c[9] * h0 + c[10] * h2;
            z[1][0] = T(0);
            z[2][6] = c[10] * h0 + c[11] * h2;
        }
        {   // s = 3
            const T h0 = w[0][2], h1 = w[1][1], h2 = w[2][7];
            z[0][2] = c[12] * h0 + c[13] * h1;
            z[1][1] = c[13] * h0 + c[14] * h1;
            z[2][7] = c[15] * h2;
        }
        {   // s = 4
            const T h0 = w[0][5], h1 = w[1][6];
            z[0][5] = c[16] * h0 + c[17] * h1;
            z[1][6] = c[17] * h0 + c[18] * h1;
            z[2][0] = T(0);
        }
        {   // s = 5
            const T h0 = w[0][4], h1 = w[1][7], h2 = w[2][1];
            z[0][4] = c[19] * h0 + c[20] * h2;
            z[1][7] = c[21] * h1;
            z[2][1] = c[20] * h0 + c[22] * h2;
        }
        {   // s = 6
            const T h0 = w[0][7], h1 = w[1][4], h2 = w[2][2];
            z[0][7] = c[23] * h0;
            z[1][4] = c[24] * h1 + c[25] * h2;
            z[2][2] = c[25] * h1 + c[26] * h2;
        }
        {   // s = 7
            const T h0 = w[0][6], h1 = w[1][5], h2 = w[2][3];
            z[0][6] = c[27] * h0 + c[28] * h1 + c[29] * h2;
            z[1][5] = c[28] * h0 + c[30] * h1 + c[31] * h2;
            z[2][3] = c[29] * h0 + c[31] * h1 + c[32] * h2;
        }
    }
};

// Pack the 72 block entries, checking symmetry and the zero pattern.
static int pack_blocks(const double *Kb, double *packed)
{
    double mx = 0.0;
    for (int i = 0; i < 72; ++i)
        mx = fmax(mx, fabs(Kb[i]));
    const double tol = 1e-10 * mx;
    bool used[72] = {};
    for (int k = 0; k < kPackedBlocks; ++k) {
        const int s = kPackS[k], rc = kPackRC[k], r = rc / 3, c = rc % 3;
        packed[k] = Kb[9 * s + rc];
        used[9 * s + rc] = used[9 * s + 3 * c + r] = true;
        if (fabs(Kb[9 * s + rc] - Kb[9 * s + 3 * c + r]) > tol) {
            set_error("hex8 WHT block %d is not symmetric", s);
            return kStatusBadArg;
        }
    }
    for (int i = 0; i < 72; ++i)
        if (!used[i] && fabs(Kb[i]) > tol) {
            set_error("element matrix does not have the box-hex8 block pattern (entry %d); use the node kernel", i);
            return kStatusBadArg;
        }
    return kStatusOk;
}

static int device_cus()
{
    static const int n = [] {
        int d = 0, c = 0;
        if (hipGetDevice(&d) == hipSuccess && hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, d) == hipSuccess &&
            c > 0)
            return c;
        (void)hipGetLastError();
        return 256;
    }();
    return n;
}

TileShape tile_shape(const Grid &g, int per_cu)
{
    TileShape ts;
    ts.tiles_x = (g.nx + kTOut - 1) / kTOut;
    ts.tiles_z = (g.nz + kTOut - 1) / kTOut;
    static const int forced = [] {
        const char *e = getenv("FANTOM_TILE_CHUNK");
        return e ? atoi(e) : 0;
    }();
    int chunk = forced > 0 ? (forced < g.ny ? forced : g.ny) : 0;
    if (chunk == 0) {
        const long cols = (long)ts.tiles_x * ts.tiles_z;
        const long cap = (long)device_cus() * per_cu;
        long best = -1;
        for (int n = 1; n <= g.ny; ++n) {
            const int c = (g.ny + n - 1) / n;
            if (n > 1 && c == (g.ny + n - 2) / (n - 1))
                continue;   // same chunk length as n - 1
            const long rounds = (cols * ((g.ny + c - 1) / c) + cap - 1) / cap;
            const long cost = rounds * (c + 1);
            if (best < 0 || cost < best) {
                best = cost;
                chunk = c;
            }
        }
    }
    ts.chunk = chunk;
    ts.chunks_y = (g.ny + chunk - 1) / chunk;
    return ts;
}

unsigned tile_blocks(const TileShape &ts)
{
    const long n = (long)ts.tiles_x * ts.tiles_z * ts.chunks_y;
    return (unsigned)((n + 7) / 8 * 8);
}

// Workgroups per CU the chunk choice assumes for each mode of the fine
// operator: the occupancy of that mode's instantiation (register use differs
// by mode: FP32 84-96 VGPRs, 5-6 waves per SIMD), the register budget
// without a device.  The two modes that write per-block partial sums share
// one value, so that one partial count serves both.
template <typename T>
static int fine_per_cu(int mode)
{
    static const std::array<int, 9> pc = [] {
        std::array<int, 9> a = {tile_occupancy<T, 0, FineOp<T>>(), tile_occupancy<T, 1, FineOp<T>>(),
                                tile_occupancy<T, 2, FineOp<T>>(), tile_occupancy<T, 3, FineOp<T>>(),
                                tile_occupancy<T, 4, FineOp<T>>(), tile_occupancy<T, 5, FineOp<T>>(),
                                tile_occupancy<T, 6, FineOp<T>>(), tile_occupancy<T, 7, FineOp<T>>(),
                                tile_occupancy<T, 8, FineOp<T>>()};
        for (int m = 0; m < (int)a.size(); ++m)
            if (a[m] <= 0)
                a[m] = FineOp<T>::waves_for(m);   // a workgroup is one wave per SIMD
        a[MODE_KUDOT] = a[MODE_JACDOT] = std::min(a[MODE_KUDOT], a[MODE_JACDOT]);
        return a;
    }();
    return pc[mode];
}

long ku_tile_partials(const Grid &g) { return (long)tile_blocks(tile_shape(g, fine_per_cu<double>(MODE_KUDOT))); }
long ku_tile_partials_f32(const Grid &g) { return (long)tile_blocks(tile_shape(g, fine_per_cu<float>(MODE_KUDOT))); }

// ec: the coarse correction of MODE_JACPRO (coarse node c on fine node 2c),
// unused by the other modes
template <typename T>
static int launch_ku_tile_t(int mode, const Grid &g, const double *Kb, const T *E, const uint8_t *mask, const T *u,
                            T *y, const T *b, const T *dinv, double omega, double *partials, hipStream_t s,
                            const T *ec = nullptr)
{
    double packed[kPackedBlocks];
    const int st = pack_blocks(Kb, packed);
    if (st != kStatusOk)
        return st;
    FineOp<T> op;
    for (int i = 0; i < kPackedBlocks; ++i)
        op.c[i] = (T)packed[i];
    // K0[3b + c][3b + c] = sum over the blocks s of their (c, c) entries
    for (int c = 0; c < 3; ++c) {
        double d = 0.0;
        for (int sb = 0; sb < 8; ++sb)
            d += Kb[9 * sb + 4 * c];
        op.rd[c] = (T)(1.0 / d);
    }
    op.E = E;
    if (mode == MODE_JAC && dinv == nullptr)
        mode = MODE_JACOWN;
    const TileShape ts = tile_shape(g, fine_per_cu<T>(mode >= 0 && mode <= MODE_JACPRO ? mode : 0));   // bad modes fail below
    const unsigned blocks = tile_blocks(ts);
    const T om = (T)omega;
#define FANTOM_TILE_CASE(M)                                                                                   \
    case M:                                                                                                   \
        hipLaunchKernelGGL((ku_tile_kernel<T, M, FineOp<T>>), dim3(blocks), dim3(kTThreads), 0, s, g, op, mask, u, \
                           y, b, dinv, om, partials, ts, ec);                                                 \
        break;
    switch (mode) {
        FANTOM_TILE_CASE(MODE_KU)
        FANTOM_TILE_CASE(MODE_JAC)
        FANTOM_TILE_CASE(MODE_RES)
        FANTOM_TILE_CASE(MODE_KUDOT)
        FANTOM_TILE_CASE(MODE_JAC0)
        FANTOM_TILE_CASE(MODE_JACDOT)
        FANTOM_TILE_CASE(MODE_JACOWN)
        FANTOM_TILE_CASE(MODE_JAC0OWN)
        FANTOM_TILE_CASE(MODE_JACPRO)
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

int launch_ku_tile_pro(const Grid &g, const double *Kb, const double *E, const uint8_t *mask, const double *u,
                       const double *ec, double *y, const double *b, double omega, hipStream_t s)
{
    return launch_ku_tile_t<double>(MODE_JACPRO, g, Kb, E, mask, u, y, b, nullptr, omega, nullptr, s, ec);
}

int launch_ku_tile_pro_f32(const Grid &g, const double *Kb, const float *E, const uint8_t *mask, const float *u,
                           const float *ec, float *y, const float *b, double omega, hipStream_t s)
{
    return launch_ku_tile_t<float>(MODE_JACPRO, g, Kb, E, mask, u, y, b, nullptr, omega, nullptr, s, ec);
}

}  // namespace fantom
