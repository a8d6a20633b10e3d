// Node-centric matrix-free K(E)u for structured hex8 / quad4 grids (FP64),
// plus diag(K(E)) and the per-element sensitivity -u_e^T K0 u_e.
//
// One thread per node gathers over the (up to) 8 / 4 incident elements, like
// the reference's node-basis kernel (core/CUDA/Kernels/matvec.cu:45-98), but
// with implicit structured indexing: no connectivity, incidence or sorter
// streams are read, so the only HBM traffic is u, y, E and a one-byte-per-node
// constraint mask.  The sum over elements is a fixed-order gather, so results
// are deterministic (no atomics).
//
// Dirichlet convention (matvec.cu:61-68,88-92): a constrained row returns
// y_i = u_i and constrained columns are skipped (u treated as 0).
//
// The epilogue is selected at compile time so that smoother sweeps, residuals
// and the PCG p.Kp dot need no extra pass over the vectors:
//   MODE_KU    y = K u
//   MODE_JAC   y = u + omega * dinv * (b - K u)
//   MODE_RES   y = b - K u
//   MODE_KUDOT y = K u and partials[block] = sum_i u_i y_i
#include "common.hpp"
#include "fantom_internal.hpp"

namespace fantom {

// local corner index of offset (ox, oy, oz) inside a hex8, reference order
// (core/CPU/_geom.py:49-58): 0 (0,0,0) 1 (1,0,0) 2 (1,1,0) 3 (0,1,0) then +z.
__device__ __forceinline__ int hex_corner(int ox, int oy, int oz)
{
    return oz * 4 + (oy ? (ox ? 2 : 3) : (ox ? 1 : 0));
}
__device__ __forceinline__ int quad_corner(int ox, int oy) { return oy ? (ox ? 2 : 3) : (ox ? 1 : 0); }

constexpr int kNodeBlock = 256;

template <int MODE>
__global__ __launch_bounds__(kNodeBlock) void ku3d_node_f64(Grid g, const double *__restrict__ K0g,
                                                          const double *__restrict__ E,
                                                          const uint8_t *__restrict__ mask,
                                                          const double *__restrict__ u, double *__restrict__ y,
                                                          const double *__restrict__ b,
                                                          const double *__restrict__ dinv, double omega,
                                                          double *__restrict__ partials)
{
    if (MODE == MODE_JAC)
        omega = jacobi_weight(g, omega);
    __shared__ double K0[576];
    for (int t = threadIdx.x; t < 576; t += kNodeBlock)
        K0[t] = K0g[t];
    __syncthreads();

    const long n = (long)blockIdx.x * kNodeBlock + threadIdx.x;
    double part = 0.0;
    if (n < g.n_nodes) {
        const int k = (int)(n % g.nz);
        const long t = n / g.nz;
        const int i = (int)(t % g.nx), j = (int)(t / g.nx);
        const long plane = (long)g.nx * g.nz;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int oy = 0; oy < 2; ++oy) {
            const int ej = j - oy;
            if (ej < 0 || ej >= g.nely)
                continue;
            for (int ox = 0; ox < 2; ++ox) {
                const int ei = i - ox;
                if (ei < 0 || ei >= g.nelx)
                    continue;
                for (int oz = 0; oz < 2; ++oz) {
                    const int ek = k - oz;
                    if (ek < 0 || ek >= g.nelz)
                        continue;
                    const double Ee = E[((long)ei * g.nely + ej) * g.nelz + ek];
                    const double *Kr = K0 + hex_corner(ox, oy, oz) * 72;
                    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        const int cx = (c == 1 || c == 2 || c == 5 || c == 6);
                        const int cy = (c == 2 || c == 3 || c == 6 || c == 7);
                        const int cz = c >= 4;
                        const long m = (long)(ej + cy) * plane + (long)(ei + cx) * g.nz + (ek + cz);
                        const int mm = mask ? mask[m] : 0;
                        const double u0 = (mm & 1) ? 0.0 : u[3 * m];
                        const double u1 = (mm & 2) ? 0.0 : u[3 * m + 1];
                        const double u2 = (mm & 4) ? 0.0 : u[3 * m + 2];
                        s0 += Kr[3 * c] * u0 + Kr[3 * c + 1] * u1 + Kr[3 * c + 2] * u2;
                        s1 += Kr[24 + 3 * c] * u0 + Kr[24 + 3 * c + 1] * u1 + Kr[24 + 3 * c + 2] * u2;
                        s2 += Kr[48 + 3 * c] * u0 + Kr[48 + 3 * c + 1] * u1 + Kr[48 + 3 * c + 2] * u2;
                    }
                    a0 += Ee * s0;
                    a1 += Ee * s1;
                    a2 += Ee * s2;
                }
            }
        }
        const int mn = mask ? mask[n] : 0;
        double acc[3] = {a0, a1, a2};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const long r = 3 * n + c;
            const double ur = u[r];
            const double ku = ((mn >> c) & 1) ? ur : acc[c];
            double out;
            if (MODE == MODE_JAC)
                out = ur + omega * dinv[r] * (b[r] - ku);
            else if (MODE == MODE_RES)
                out = b[r] - ku;
            else
                out = ku;
            y[r] = out;
            if (MODE == MODE_KUDOT)
                part += ur * out;
        }
    }
    if (MODE == MODE_KUDOT) {
        const double s = block_sum<kNodeBlock>(part);
        if (threadIdx.x == 0)
            partials[blockIdx.x] = s;
    }
}

template <int MODE>
__global__ __launch_bounds__(kNodeBlock) void ku2d_node_f64(Grid g, const double *__restrict__ K0g,
                                                          const double *__restrict__ E,
                                                          const uint8_t *__restrict__ mask,
                                                          const double *__restrict__ u, double *__restrict__ y,
                                                          const double *__restrict__ b,
                                                          const double *__restrict__ dinv, double omega,
                                                          double *__restrict__ partials)
{
    if (MODE == MODE_JAC)
        omega = jacobi_weight(g, omega);
    __shared__ double K0[64];
    if (threadIdx.x < 64)
        K0[threadIdx.x] = K0g[threadIdx.x];
    __syncthreads();

    const long n = (long)blockIdx.x * kNodeBlock + threadIdx.x;
    double part = 0.0;
    if (n < g.n_nodes) {
        const int i = (int)(n % g.nx), j = (int)(n / g.nx);
        double a0 = 0.0, a1 = 0.0;
        for (int oy = 0; oy < 2; ++oy) {
            const int ej = j - oy;
            if (ej < 0 || ej >= g.nely)
                continue;
            for (int ox = 0; ox < 2; ++ox) {
                const int ei = i - ox;
                if (ei < 0 || ei >= g.nelx)
                    continue;
                const double Ee = E[(long)ei * g.nely + ej];
                const double *Kr = K0 + quad_corner(ox, oy) * 16;
                double s0 = 0.0, s1 = 0.0;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int cx = (c == 1 || c == 2), cy = (c >= 2);
                    const long m = (long)(ej + cy) * g.nx + (ei + cx);
                    const int mm = mask ? mask[m] : 0;
                    const double u0 = (mm & 1) ? 0.0 : u[2 * m];
                    const double u1 = (mm & 2) ? 0.0 : u[2 * m + 1];
                    s0 += Kr[2 * c] * u0 + Kr[2 * c + 1] * u1;
                    s1 += Kr[8 + 2 * c] * u0 + Kr[8 + 2 * c + 1] * u1;
                }
                a0 += Ee * s0;
                a1 += Ee * s1;
            }
        }
        const int mn = mask ? mask[n] : 0;
        double acc[2] = {a0, a1};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const long r = 2 * n + c;
            const double ur = u[r];
            const double ku = ((mn >> c) & 1) ? ur : acc[c];
            double out;
            if (MODE == MODE_JAC)
                out = ur + omega * dinv[r] * (b[r] - ku);
            else if (MODE == MODE_RES)
                out = b[r] - ku;
            else
                out = ku;
            y[r] = out;
            if (MODE == MODE_KUDOT)
                part += ur * out;
        }
    }
    if (MODE == MODE_KUDOT) {
        const double s = block_sum<kNodeBlock>(part);
        if (threadIdx.x == 0)
            partials[blockIdx.x] = s;
    }
}

// diag(K(E)); constrained DOFs -> 1 (core/CUDA/Kernels/diag.cu:1-47).  invert
// stores 1/diag instead (the Jacobi weights).
__global__ __launch_bounds__(kNodeBlock) void diag_node_f64(Grid g, int dim, const double *__restrict__ K0g,
                                                          const double *__restrict__ E,
                                                          const uint8_t *__restrict__ mask, double *__restrict__ d,
                                                          int invert)
{
    const long n = (long)blockIdx.x * kNodeBlock + threadIdx.x;
    if (n >= g.n_nodes)
        return;
    const int dof = dim == 3 ? 3 : 2, nl = dim == 3 ? 24 : 8;
    double acc[3] = {0.0, 0.0, 0.0};
    if (dim == 3) {
        const int k = (int)(n % g.nz);
        const long t = n / g.nz;
        const int i = (int)(t % g.nx), j = (int)(t / g.nx);
        for (int oy = 0; oy < 2; ++oy)
            for (int ox = 0; ox < 2; ++ox)
                for (int oz = 0; oz < 2; ++oz) {
                    const int ei = i - ox, ej = j - oy, ek = k - oz;
                    if (ei < 0 || ei >= g.nelx || ej < 0 || ej >= g.nely || ek < 0 || ek >= g.nelz)
                        continue;
                    const double Ee = E[((long)ei * g.nely + ej) * g.nelz + ek];
                    const int a = hex_corner(ox, oy, oz);
                    for (int c = 0; c < 3; ++c)
                        acc[c] += K0g[(a * 3 + c) * nl + a * 3 + c] * Ee;
                }
    } else {
        const int i = (int)(n % g.nx), j = (int)(n / g.nx);
        for (int oy = 0; oy < 2; ++oy)
            for (int ox = 0; ox < 2; ++ox) {
                const int ei = i - ox, ej = j - oy;
                if (ei < 0 || ei >= g.nelx || ej < 0 || ej >= g.nely)
                    continue;
                const double Ee = E[(long)ei * g.nely + ej];
                const int a = quad_corner(ox, oy);
                for (int c = 0; c < 2; ++c)
                    acc[c] += K0g[(a * 2 + c) * nl + a * 2 + c] * Ee;
            }
    }
    const int mn = mask ? mask[n] : 0;
    for (int c = 0; c < dof; ++c) {
        double v = ((mn >> c) & 1) ? 1.0 : acc[c];
        d[dof * n + c] = invert ? 1.0 / v : v;
    }
}

// dc_e = -u_e^T K0 u_e, one thread per element (core/CUDA/Kernels/grad.cu:1-16).
template <int DIM>
__global__ __launch_bounds__(kNodeBlock) void dc_elem_f64(Grid g, const double *__restrict__ K0g,
                                                        const double *__restrict__ u, double *__restrict__ dc)
{
    constexpr int NL = DIM == 3 ? 24 : 8;
    constexpr int NPE = DIM == 3 ? 8 : 4;
    __shared__ double K0[NL * NL];
    for (int t = threadIdx.x; t < NL * NL; t += kNodeBlock)
        K0[t] = K0g[t];
    __syncthreads();
    const long e = (long)blockIdx.x * kNodeBlock + threadIdx.x;
    if (e >= g.n_el)
        return;
    double ue[NL];
    if (DIM == 3) {
        const int ei = (int)(e / ((long)g.nely * g.nelz));
        const int r = (int)(e % ((long)g.nely * g.nelz));
        const int ej = r / g.nelz, ek = r % g.nelz;
        const long plane = (long)g.nx * g.nz;
#pragma unroll
        for (int c = 0; c < NPE; ++c) {
            const int cx = (c == 1 || c == 2 || c == 5 || c == 6);
            const int cy = (c == 2 || c == 3 || c == 6 || c == 7);
            const int cz = c >= 4;
            const long m = (long)(ej + cy) * plane + (long)(ei + cx) * g.nz + (ek + cz);
            ue[3 * c] = u[3 * m];
            ue[3 * c + 1] = u[3 * m + 1];
            ue[3 * c + 2] = u[3 * m + 2];
        }
    } else {
        const int ei = (int)(e / g.nely), ej = (int)(e % g.nely);
#pragma unroll
        for (int c = 0; c < NPE; ++c) {
            const int cx = (c == 1 || c == 2), cy = (c >= 2);
            const long m = (long)(ej + cy) * g.nx + (ei + cx);
            ue[2 * c] = u[2 * m];
            ue[2 * c + 1] = u[2 * m + 1];
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < NL; ++r) {
        double t = 0.0;
#pragma unroll
        for (int l = 0; l < NL; ++l)
            t += K0[r * NL + l] * ue[l];
        acc += t * ue[r];
    }
    dc[e] = -acc;
}

// FP32 Jacobi weights dinv = 1 / diag(K(E)) for a hex8 grid (constrained -> 1),
// accumulated in FP64 from the FP64 element matrix.
__global__ __launch_bounds__(kNodeBlock) void dinv3d_f32(Grid g, const double *__restrict__ K0g,
                                                       const float *__restrict__ E,
                                                       const uint8_t *__restrict__ mask, float *__restrict__ d)
{
    const long n = (long)blockIdx.x * kNodeBlock + threadIdx.x;
    if (n >= g.n_nodes)
        return;
    double acc[3] = {0.0, 0.0, 0.0};
    const int k = (int)(n % g.nz);
    const long t = n / g.nz;
    const int i = (int)(t % g.nx), j = (int)(t / g.nx);
    for (int oy = 0; oy < 2; ++oy)
        for (int ox = 0; ox < 2; ++ox)
            for (int oz = 0; oz < 2; ++oz) {
                const int ei = i - ox, ej = j - oy, ek = k - oz;
                if (ei < 0 || ei >= g.nelx || ej < 0 || ej >= g.nely || ek < 0 || ek >= g.nelz)
                    continue;
                const double Ee = E[((long)ei * g.nely + ej) * g.nelz + ek];
                const int a = hex_corner(ox, oy, oz);
                for (int c = 0; c < 3; ++c)
                    acc[c] += K0g[(a * 3 + c) * 24 + a * 3 + c] * Ee;
            }
    const int mn = mask ? mask[n] : 0;
    for (int c = 0; c < 3; ++c)
        d[3 * n + c] = ((mn >> c) & 1) ? 1.0f : (float)(1.0 / acc[c]);
}

int launch_diag_f32(const Grid &g, const double *K0, const float *E, const uint8_t *mask, float *d, hipStream_t s)
{
    const unsigned blocks = (unsigned)((g.n_nodes + kNodeBlock - 1) / kNodeBlock);
    hipLaunchKernelGGL(dinv3d_f32, dim3(blocks), dim3(kNodeBlock), 0, s, g, K0, E, mask, d);
    FANTOM_LAUNCH_CHECK("dinv_f32");
    return kStatusOk;
}

int launch_ku_node(int mode, const Grid &g, int dim, const double *K0, const double *E, const uint8_t *mask,
                   const double *u, double *y, const double *b, const double *dinv, double omega,
                   double *partials, hipStream_t s)
{
    const unsigned blocks = (unsigned)((g.n_nodes + kNodeBlock - 1) / kNodeBlock);
#define FANTOM_KU_CASE(M)                                                                                  \
    case M:                                                                                                \
        if (dim == 3)                                                                                      \
            hipLaunchKernelGGL(ku3d_node_f64<M>, dim3(blocks), dim3(kNodeBlock), 0, s, g, K0, E, mask, u, y,  \
                               b, dinv, omega, partials);                                                   \
        else                                                                                               \
            hipLaunchKernelGGL(ku2d_node_f64<M>, dim3(blocks), dim3(kNodeBlock), 0, s, g, K0, E, mask, u, y,  \
                               b, dinv, omega, partials);                                                   \
        break;
    switch (mode) {
        FANTOM_KU_CASE(MODE_KU)
        FANTOM_KU_CASE(MODE_JAC)
        FANTOM_KU_CASE(MODE_RES)
        FANTOM_KU_CASE(MODE_KUDOT)
    default:
        set_error("unknown matvec mode %d", mode);
        return kStatusBadArg;
    }
#undef FANTOM_KU_CASE
    FANTOM_LAUNCH_CHECK("ku_node");
    return kStatusOk;
}

long ku_node_partials(const Grid &g) { return (g.n_nodes + kNodeBlock - 1) / kNodeBlock; }

int launch_diag(const Grid &g, int dim, const double *K0, const double *E, const uint8_t *mask, double *d,
                int invert, hipStream_t s)
{
    const unsigned blocks = (unsigned)((g.n_nodes + kNodeBlock - 1) / kNodeBlock);
    hipLaunchKernelGGL(diag_node_f64, dim3(blocks), dim3(kNodeBlock), 0, s, g, dim, K0, E, mask, d, invert);
    FANTOM_LAUNCH_CHECK("diag");
    return kStatusOk;
}

int launch_dc(const Grid &g, int dim, const double *K0, const double *u, double *dc, hipStream_t s)
{
    const unsigned blocks = (unsigned)((g.n_el + kNodeBlock - 1) / kNodeBlock);
    if (dim == 3)
        hipLaunchKernelGGL(dc_elem_f64<3>, dim3(blocks), dim3(kNodeBlock), 0, s, g, K0, u, dc);
    else
        hipLaunchKernelGGL(dc_elem_f64<2>, dim3(blocks), dim3(kNodeBlock), 0, s, g, K0, u, dc);
    FANTOM_LAUNCH_CHECK("dc");
    return kStatusOk;
}

}  // namespace fantom
