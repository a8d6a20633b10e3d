// Assembled 27-point node stencils for the coarse Galerkin levels (>= 2).
//
// A coarse level with explicit 24x24 element matrices is applied far faster
// from its assembled rows than element by element: a node's three rows hold
// 27 neighbour blocks of 3x3 (243 coefficients), against 8 x 72 element-matrix
// entries that a node-centric gather would read.  Layout: structure of
// arrays S[k][n] with k = t * 9 + 3 c + d, neighbour t = (dy+1)*9 + (dx+1)*3 +
// (dz+1) (y slowest, as the node numbering), row component c, column
// component d, so that consecutive threads (nodes) read consecutive words.
// The Dirichlet rule (constrained rows are identity rows, constrained columns
// are skipped) is folded into the coefficients when the stencil is built, so
// the apply reads no mask.
#include "../../include/fantom.h"
#include "common.hpp"
#include "fantom_internal.hpp"

namespace fantom {

constexpr int kStencilK = 243;

// one thread per (node, neighbour t): the 9 coefficients of block t, summed
// over the (up to 8) elements around the node in a fixed order
template <typename TS>
__global__ void stencil_from_elem_k(Grid g, const double *__restrict__ Ke, const uint8_t *__restrict__ mask,
                                    TS *__restrict__ S)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.n_nodes * 27)
        return;
    const long n = idx % g.n_nodes;
    const int t = (int)(idx / g.n_nodes);
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
            S[(long)(t * 9 + 3 * c + d) * g.n_nodes + n] = (TS)v;
        }
}

// y = K u from the stencil with the solver epilogues.  A block of 192
// threads covers 64 nodes; wave p (p = 0, 1, 2) takes the 9 neighbours in
// node plane y + p - 1 of each node, so every thread has one round of
// independent loads (coarse levels are small and latency-bound), and wave 0
// adds the three plane sums in a fixed order.
constexpr int kStNodes = 64;
constexpr int kStBlock = 3 * kStNodes;

template <typename T, int MODE>
__global__ __launch_bounds__(kStBlock) void ku_stencil_k(Grid g, const T *__restrict__ S, const T *__restrict__ u,
                                                       T *__restrict__ y, const T *__restrict__ bvec,
                                                       const T *__restrict__ dinv, T omega,
                                                       double *__restrict__ partials)
{
    if (MODE == MODE_JAC || MODE == MODE_JAC0)
        omega = jacobi_weight(g, omega);
    __shared__ T psum[3][3][kStNodes];
    const int ln = threadIdx.x % kStNodes, ty = threadIdx.x / kStNodes;
    const long n = (long)blockIdx.x * kStNodes + ln;
    const bool valid = n < g.n_nodes;
    const long nn = valid ? n : 0;
    const int k = (int)(nn % g.nz);
    const long q = nn / g.nz;
    const int i = (int)(q % g.nx), j = (int)(q / g.nx);
    T acc[3] = {T(0), T(0), T(0)};
    if (valid) {
        const int dy = ty - 1;
#pragma unroll
        for (int tq = 0; tq < 9; ++tq) {
            const int t = ty * 9 + tq;
            const int dx = tq / 3 - 1, dz = tq % 3 - 1;
            const int mi = i + dx, mj = j + dy, mk = k + dz;
            // outside the grid the coefficients are 0 by construction; read
            // the node itself so that the loads issue unconditionally
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
            const T *St = S + (long)t * 9 * g.n_nodes + nn;
#pragma unroll
            for (int c = 0; c < 3; ++c)
                acc[c] += St[(3 * c) * g.n_nodes] * u0 + St[(3 * c + 1) * g.n_nodes] * u1 +
                          St[(3 * c + 2) * g.n_nodes] * u2;
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
            const T kuv = psum[0][c][ln] + psum[1][c][ln] + psum[2][c][ln];
            const long r = 3 * n + c;
            T o;
            if (MODE == MODE_JAC)
                o = u[r] + omega * dinv[r] * (bvec[r] - kuv);
            else if (MODE == MODE_JAC0) {   // x0 = omega D^-1 b, then one sweep of weight omega * w2_ratio
                const T x0 = omega * dinv[r] * bvec[r];
                o = x0 + (omega * (T)g.w2_ratio) * dinv[r] * (bvec[r] - kuv);
            }
            else if (MODE == MODE_RES)
                o = bvec[r] - kuv;
            else
                o = kuv;
            y[r] = o;
            if (MODE == MODE_KUDOT)
                part += (double)u[r] * (double)o;
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
        d[3 * n + c] = T(1) / S[(long)(13 * 9 + 4 * c) * g.n_nodes + n];
}

long stencil_partials(const Grid &g) { return (g.n_nodes + kStNodes - 1) / kStNodes; }

template <typename TS>
static int stencil_build_t(const Grid &g, const double *Ke, const uint8_t *mask, TS *S, hipStream_t s)
{
    const long n = g.n_nodes * 27;
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

extern "C" long fantom_stencil_partials(int nelx, int nely, int nelz)
{
    if (!grid_ok(nelx, nely, nelz) || nelz <= 0)
        return -1;
    return stencil_partials(make_grid(nelx, nely, nelz));
}
