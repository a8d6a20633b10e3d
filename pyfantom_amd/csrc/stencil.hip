// Assembled 27-point node stencils for the coarse Galerkin levels (>= 2).
//
// A coarse level with explicit 24x24 element matrices is applied far faster
// from its assembled rows than element by element: a node's three rows hold
// 27 neighbour blocks of 3x3, against 8 x 72 element-matrix entries that a
// node-centric gather would read.  The operator is symmetric, so only the
// centre block and the 13 forward neighbour blocks (t = 13 ... 26 below) are
// stored; the block of backward neighbour t < 13 is the transpose of the
// forward block 26 - t of that neighbour, read at the neighbour's index as
// its u value is.  That halves the coefficient bytes, which bound the sweep.
// Layout: tiles of 64 nodes, each a structure of arrays over the 126
// coefficients, S[n / 64][k][n % 64] with k = (t - 13) * 9 + 3 c + d,
// neighbour t = (dy+1)*9 + (dx+1)*3 + (dz+1) (y slowest, as the node
// numbering), row component c, column component d: consecutive threads
// (nodes) read consecutive words, and a workgroup's own coefficients are
// one contiguous 64 x 126-word block instead of 126 runs far apart.  The Dirichlet rule (constrained
// rows are identity rows, constrained columns are skipped) is folded into the
// coefficients when the stencil is built, so the apply reads no mask.
#include "../../include/fantom.h"
#include "common.hpp"
#include "fantom_internal.hpp"

namespace fantom {

static_assert(FANTOM_STENCIL_COEFS == 14 * 9, "stencil: centre block and 13 forward blocks");
constexpr int kStTile = 64;   // nodes per layout tile

// word offset of coefficient k of node n
__device__ __forceinline__ long st_at(int k, long n)
{
    return ((n / kStTile) * FANTOM_STENCIL_COEFS + k) * kStTile + (n % kStTile);
}

// one thread per (node, forward neighbour t): the 9 coefficients of block t,
// summed over the (up to 8) elements around the node in a fixed order
template <typename TS>
__global__ void stencil_from_elem_k(Grid g, const double *__restrict__ Ke, const uint8_t *__restrict__ mask,
                                    TS *__restrict__ S)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.n_nodes * 14)
        return;
    const long n = idx % g.n_nodes;
    const int t = 13 + (int)(idx / g.n_nodes);
    const int dy = t / 9 - 1, dx = (t / 3) % 3 - 1, dz = t % 3 - 1;
    const int k = (int)(n % g.nz);
    const long q = n / g.nz;
    const int i = (int)(q % g.nx), j = (int)(q / g.nx);
    const int mi = i + dx, mj = j + dy, mk = k + dz;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const bool nb_in = mi >= 0 && mi < g.nx && mj >= 0 && mj < g.ny && mk >= 0 && mk < g.nz;
    if (nb_in) {
        for (int ox = 1; ox >= 0; --ox)
            for (int oy = 1; oy >= 0; --oy)
                for (int oz = 1; oz >= 0; --oz) {
                    const int ei = i - ox, ej = j - oy, ek = k - oz;
                    if (ei < 0 || ei >= g.nelx || ej < 0 || ej >= g.nely || ek < 0 || ek >= g.nelz)
                        continue;
                    const int bx = ox + dx, by = oy + dy, bz = oz + dz;   // neighbour's corner in this element
                    if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1)
                        continue;
                    const long e = ((long)ei * g.nely + ej) * g.nelz + ek;
                    const int a = ox + 2 * oy + 4 * oz, b = bx + 2 * by + 4 * bz;
                    const double *K = Ke + e * 576;
                    for (int c = 0; c < 3; ++c)
                        for (int d = 0; d < 3; ++d)
                            acc[3 * c + d] += K[(a * 3 + c) * 24 + b * 3 + d];
                }
    }
    const int mn = mask ? mask[n] : 0;
    const long m = nb_in ? ((long)mj * g.nx + mi) * g.nz + mk : n;
    const int mm = (mask && nb_in) ? mask[m] : 0;
    for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) {
            double v = acc[3 * c + d];
            if ((mn >> c) & 1)
                v = (t == 13 && c == d) ? 1.0 : 0.0;
            else if ((mm >> d) & 1)
                v = 0.0;
            S[st_at((t - 13) * 9 + 3 * c + d, n)] = (TS)v;
        }
}

// y = K u from the stencil with the solver epilogues.  A block of 576
// threads covers 64 nodes; wave p (p = 0 ... 8) takes the 3 neighbours
// (dy, dx) = (p / 3 - 1, p % 3 - 1), dz = -1, 0, 1 of each node, so every
// thread has one round of 36 independent loads (coarse levels are small and
// latency-bound: with 9 neighbours per thread the loads did not fit in
// flight), and wave 0 adds the nine partial sums in a fixed order.
constexpr int kStNodes = 64;
constexpr int kStWaves = 9;
constexpr int kStBlock = kStWaves * kStNodes;

template <typename T, int MODE>
__global__ __launch_bounds__(kStBlock) void ku_stencil_k(Grid g, const T *__restrict__ S, const T *__restrict__ u,
                                                       T *__restrict__ y, const T *__restrict__ bvec,
                                                       const T *__restrict__ dinv, T omega,
                                                       double *__restrict__ partials)
{
    if (MODE == MODE_JAC || MODE == MODE_JAC0)
        omega = jacobi_weight(g, omega);
    __shared__ T psum[kStWaves][3][kStNodes];
    const int ln = threadIdx.x % kStNodes, ty = threadIdx.x / kStNodes;
    const long n = (long)blockIdx.x * kStNodes + ln;
    const bool valid = n < g.n_nodes;
    const long nn = valid ? n : 0;
    const int k = (int)(nn % g.nz);
    const long q = nn / g.nz;
    const int i = (int)(q % g.nx), j = (int)(q / g.nx);
    T acc[3] = {T(0), T(0), T(0)};
    if (valid) {
        const int dy = ty / 3 - 1, dx = ty % 3 - 1;
#pragma unroll
        for (int tq = 0; tq < 3; ++tq) {
            const int t = ty * 3 + tq;   // wave-uniform
            const int dz = tq - 1;
            const int mi = i + dx, mj = j + dy, mk = k + dz;
            // outside the grid the neighbour value is 0; read the node itself
            // so that the loads issue unconditionally
            const bool in = mi >= 0 && mi < g.nx && mj >= 0 && mj < g.ny && mk >= 0 && mk < g.nz;
            const long m = in ? ((long)mj * g.nx + mi) * g.nz + mk : nn;
            T u0, u1, u2;
            if (MODE == MODE_JAC0) {
                u0 = omega * dinv[3 * m] * bvec[3 * m];
                u1 = omega * dinv[3 * m + 1] * bvec[3 * m + 1];
                u2 = omega * dinv[3 * m + 2] * bvec[3 * m + 2];
            } else {
                u0 = u[3 * m];
                u1 = u[3 * m + 1];
                u2 = u[3 * m + 2];
            }
            if (!in)
                u0 = u1 = u2 = T(0);
            // forward block t of this node, or forward block 26 - t of the
            // neighbour, transposed
            const bool own = t >= 13;
            const T *St = S + st_at((own ? t - 13 : 13 - t) * 9, own ? nn : m);
            T s[9];
#pragma unroll
            for (int e = 0; e < 9; ++e)
                s[e] = St[e * kStTile];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T a0 = own ? s[3 * c] : s[c];
                const T a1 = own ? s[3 * c + 1] : s[3 + c];
                const T a2 = own ? s[3 * c + 2] : s[6 + c];
                acc[c] += a0 * u0 + a1 * u1 + a2 * u2;
            }
        }
    }
    // the epilogue's own-node operands, loaded by wave 0 before the barrier
    constexpr bool kNeedU = MODE == MODE_JAC || MODE == MODE_KUDOT;
    constexpr bool kNeedB = MODE == MODE_JAC || MODE == MODE_JAC0 || MODE == MODE_RES;
    constexpr bool kNeedD = MODE == MODE_JAC || MODE == MODE_JAC0;
    T ur[3] = {T(0), T(0), T(0)}, br[3] = {T(0), T(0), T(0)}, dr[3] = {T(0), T(0), T(0)};
    if (ty == 0 && valid) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (kNeedU)
                ur[c] = u[3 * n + c];
            if (kNeedB)
                br[c] = bvec[3 * n + c];
            if (kNeedD)
                dr[c] = dinv[3 * n + c];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
        psum[ty][c][ln] = acc[c];
    __syncthreads();
    double part = 0.0;
    if (ty == 0 && valid) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T kuv = psum[0][c][ln];
#pragma unroll
            for (int w = 1; w < kStWaves; ++w)
                kuv += psum[w][c][ln];
            T o;
            if (MODE == MODE_JAC)
                o = ur[c] + omega * dr[c] * (br[c] - kuv);
            else if (MODE == MODE_JAC0) {   // x0 = omega D^-1 b, then one sweep of weight omega * w2_ratio
                const T x0 = omega * dr[c] * br[c];
                o = x0 + (omega * (T)g.w2_ratio) * dr[c] * (br[c] - kuv);
            }
            else if (MODE == MODE_RES)
                o = br[c] - kuv;
            else
                o = kuv;
            y[3 * n + c] = o;
            if (MODE == MODE_KUDOT)
                part += (double)ur[c] * (double)o;
        }
    }
    if (MODE == MODE_KUDOT) {
        const double s = block_sum<kStBlock>(part);
        if (threadIdx.x == 0)
            partials[blockIdx.x] = s;
    }
}

// Jacobi weights from the stencil's centre block (identity rows give 1)
template <typename T>
__global__ void dinv_stencil_k(Grid g, const T *__restrict__ S, T *__restrict__ d)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.n_nodes)
        return;
    for (int c = 0; c < 3; ++c)
        d[3 * n + c] = T(1) / S[st_at(4 * c, n)];
}

long stencil_partials(const Grid &g) { return (g.n_nodes + kStNodes - 1) / kStNodes; }

long stencil_size(const Grid &g) { return (g.n_nodes + kStTile - 1) / kStTile * kStTile * FANTOM_STENCIL_COEFS; }

template <typename TS>
static int stencil_build_t(const Grid &g, const double *Ke, const uint8_t *mask, TS *S, hipStream_t s)
{
    const long n = g.n_nodes * 14;
    hipLaunchKernelGGL(stencil_from_elem_k<TS>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, Ke, mask, S);
    FANTOM_LAUNCH_CHECK("stencil_from_elem");
    return kStatusOk;
}

int launch_stencil_build(const Grid &g, const double *Ke, const uint8_t *mask, double *S, hipStream_t s)
{
    return stencil_build_t<double>(g, Ke, mask, S, s);
}

int launch_stencil_build_f32(const Grid &g, const double *Ke, const uint8_t *mask, float *S, hipStream_t s)
{
    return stencil_build_t<float>(g, Ke, mask, S, s);
}

template <typename T>
static int stencil_ku_t(int mode, const Grid &g, const T *S, const T *u, T *y, const T *b, const T *dinv,
                        double omega, double *partials, hipStream_t s)
{
    const unsigned blocks = (unsigned)stencil_partials(g);
    const T om = (T)omega;
#define ST_CASE(M)                                                                                        \
    case M:                                                                                               \
        hipLaunchKernelGGL((ku_stencil_k<T, M>), dim3(blocks), dim3(kStBlock), 0, s, g, S, u, y, b, dinv, om,    \
                           partials);                                                                   \
        break;
    switch (mode) {
        ST_CASE(MODE_KU)
        ST_CASE(MODE_JAC)
        ST_CASE(MODE_RES)
        ST_CASE(MODE_KUDOT)
        ST_CASE(MODE_JAC0)
    default:
        set_error("unknown mode %d", mode);
        return kStatusBadArg;
    }
#undef ST_CASE
    FANTOM_LAUNCH_CHECK("ku_stencil");
    return kStatusOk;
}

int launch_stencil_ku(int mode, const Grid &g, const double *S, const double *u, double *y, const double *b,
                      const double *dinv, double omega, double *partials, hipStream_t s)
{
    return stencil_ku_t<double>(mode, g, S, u, y, b, dinv, omega, partials, s);
}

int launch_stencil_ku_f32(int mode, const Grid &g, const float *S, const float *u, float *y, const float *b,
                          const float *dinv, double omega, double *partials, hipStream_t s)
{
    return stencil_ku_t<float>(mode, g, S, u, y, b, dinv, omega, partials, s);
}

template <typename T>
static int dinv_stencil_t(const Grid &g, const T *S, T *d, hipStream_t s)
{
    hipLaunchKernelGGL(dinv_stencil_k<T>, dim3((unsigned)((g.n_nodes + 255) / 256)), dim3(256), 0, s, g, S, d);
    FANTOM_LAUNCH_CHECK("dinv_stencil");
    return kStatusOk;
}

int launch_dinv_stencil(const Grid &g, const double *S, double *d, hipStream_t s) { return dinv_stencil_t(g, S, d, s); }
int launch_dinv_stencil_f32(const Grid &g, const float *S, float *d, hipStream_t s) { return dinv_stencil_t(g, S, d, s); }

}  // namespace fantom

using namespace fantom;

extern "C" int fantom_stencil_build_f64(int nelx, int nely, int nelz, const double *Ke, const uint8_t *mask,
                                        double *S, void *stream)
{
    FANTOM_CHECK_ARG(grid_ok(nelx, nely, nelz) && nelz > 0, "bad 3D grid");
    FANTOM_CHECK_ARG(Ke && S, "null operand");
    return launch_stencil_build(make_grid(nelx, nely, nelz), Ke, mask, S, as_stream(stream));
}

extern "C" int fantom_stencil_ku_f64(int mode, int nelx, int nely, int nelz, const double *S, const double *u,
                                     double *y, const double *b, const double *dinv, double omega, double *partials,
                                     void *stream)
{
    FANTOM_CHECK_ARG(grid_ok(nelx, nely, nelz) && nelz > 0, "bad 3D grid");
    FANTOM_CHECK_ARG(S && (u || mode == MODE_JAC0) && y, "null operand");
    FANTOM_CHECK_ARG(mode >= MODE_KU && mode <= MODE_JAC0, "unknown mode %d", mode);
    FANTOM_CHECK_ARG((mode != MODE_JAC && mode != MODE_JAC0 && mode != MODE_RES) || b, "mode needs b");
    FANTOM_CHECK_ARG((mode != MODE_JAC && mode != MODE_JAC0) || dinv, "mode needs dinv");
    FANTOM_CHECK_ARG(mode != MODE_KUDOT || partials, "mode needs partials");
    return launch_stencil_ku(mode, make_grid(nelx, nely, nelz), S, u, y, b, dinv, omega, partials,
                             as_stream(stream));
}

extern "C" long fantom_stencil_size(int nelx, int nely, int nelz)
{
    if (!grid_ok(nelx, nely, nelz) || nelz <= 0)
        return -1;
    return stencil_size(make_grid(nelx, nely, nelz));
}

extern "C" long fantom_stencil_partials(int nelx, int nely, int nelz)
{
    if (!grid_ok(nelx, nely, nelz) || nelz <= 0)
        return -1;
    return stencil_partials(make_grid(nelx, nely, nelz));
}
