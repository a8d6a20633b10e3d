// Vector kernels for PCG (FP64): deterministic dot products and the fused
// PCG updates, with every scalar (alpha, beta, r.z, p.Kp, r.r) kept on the
// device so that an iteration needs no host round trip except the optional
// convergence read.
//
// Reductions are two-stage and fixed-order: each block writes one partial
// (wave64 butterfly + ordered per-wave sum), then a single-block finalize sums
// the partials in a fixed strided order.  Results are therefore identical run
// to run for the same vector length.
#include "common.hpp"
#include "fantom_internal.hpp"
#include "vec.hpp"

namespace fantom {

constexpr int kVecBlock = 256;

__global__ __launch_bounds__(kVecBlock) void dot_partials_f64(long n, const double *__restrict__ a,
                                                            const double *__restrict__ b,
                                                            double *__restrict__ partials)
{
    double s = 0.0;
    for (long i = (long)blockIdx.x * kVecBlock + threadIdx.x; i < n; i += (long)gridDim.x * kVecBlock)
        s += a[i] * b[i];
    s = block_sum<kVecBlock>(s);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = s;
}

// partials of a.b into partials[0:nb] and of a.c into partials[nb:2 nb]
// (nb = gridDim.x): one pass over a for two dot products
__global__ __launch_bounds__(kVecBlock) void dot2_partials_f64(long n, const double *__restrict__ a,
                                                             const double *__restrict__ b,
                                                             const double *__restrict__ c,
                                                             double *__restrict__ partials)
{
    double s = 0.0, t = 0.0;
    for (long i = (long)blockIdx.x * kVecBlock + threadIdx.x; i < n; i += (long)gridDim.x * kVecBlock) {
        const double ai = a[i];
        s += ai * b[i];
        t += ai * c[i];
    }
    s = block_sum<kVecBlock>(s);
    t = block_sum<kVecBlock>(t);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s;
        partials[gridDim.x + blockIdx.x] = t;
    }
}

// Single block: S = sum(partials[0:np]) in fixed order, then apply op.
__global__ __launch_bounds__(1024) void finalize_f64(const double *__restrict__ partials, long np,
                                                   double *__restrict__ scal, int op, int slot)
{
    double s = 0.0;
    for (long i = threadIdx.x; i < np; i += 1024)
        s += partials[i];
    s = block_sum<1024>(s);
    if (threadIdx.x == 0) {
        switch (op) {
        case FIN_STORE:
            scal[slot] = s;
            break;
        case FIN_ALPHA:  // p.Kp -> alpha = rz / pq (0 when p.Kp is not positive and finite: no update)
            scal[S_PQ] = s;
            scal[S_ALPHA] = (s > 0.0 && isfinite(s)) ? scal[S_RZ] / s : 0.0;
            break;
        case FIN_BETA:  // new r.z -> beta = rz_new / rz_old
            scal[S_BETA] = s / scal[S_RZ];
            scal[S_RZ] = s;
            break;
        case FIN_RZ0:  // first r.z
            scal[S_RZ] = s;
            break;
        default:
            break;
        }
    }
}

// x += alpha p ; r -= alpha q ; partials of r.r
__global__ __launch_bounds__(kVecBlock) void pcg_xr_f64(long n, double *__restrict__ x, double *__restrict__ r,
                                                      const double *__restrict__ p, const double *__restrict__ q,
                                                      const double *__restrict__ scal, double *__restrict__ partials)
{
    const double alpha = scal[S_ALPHA];
    double s = 0.0;
    for (long i = (long)blockIdx.x * kVecBlock + threadIdx.x; i < n; i += (long)gridDim.x * kVecBlock) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        r[i] = ri;
        s += ri * ri;
    }
    s = block_sum<kVecBlock>(s);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = s;
}

// p = z + beta p  (beta_is_zero: p = z)
__global__ __launch_bounds__(kVecBlock) void pcg_p_f64(long n, double *__restrict__ p, const double *__restrict__ z,
                                                     const double *__restrict__ scal, int first)
{
    const double beta = first ? 0.0 : scal[S_BETA];
    for (long i = (long)blockIdx.x * kVecBlock + threadIdx.x; i < n; i += (long)gridDim.x * kVecBlock)
        p[i] = first ? z[i] : z[i] + beta * p[i];
}

// x = omega * dinv * b  (first Jacobi sweep from a zero guess)
__global__ __launch_bounds__(kVecBlock) void scale_f64(long n, double omega, const double *__restrict__ omega_scale,
                                                     const double *__restrict__ dinv, const double *__restrict__ b,
                                                     double *__restrict__ x)
{
    if (omega_scale)
        omega *= omega_scale[0];
    for (long i = (long)blockIdx.x * kVecBlock + threadIdx.x; i < n; i += (long)gridDim.x * kVecBlock)
        x[i] = omega * dinv[i] * b[i];
}

// y = a x + b y
__global__ __launch_bounds__(kVecBlock) void axpby_f64(long n, double a, const double *__restrict__ x, double b,
                                                     double *__restrict__ y)
{
    for (long i = (long)blockIdx.x * kVecBlock + threadIdx.x; i < n; i += (long)gridDim.x * kVecBlock)
        y[i] = a * x[i] + b * y[i];
}

unsigned vec_grid(long n)
{
    long b = (n + kVecBlock - 1) / kVecBlock;
    if (b > kVecMaxBlocks)
        b = kVecMaxBlocks;
    if (b < 1)
        b = 1;
    return (unsigned)b;
}

int launch_dot_partials(long n, const double *a, const double *b, double *partials, hipStream_t s)
{
    hipLaunchKernelGGL(dot_partials_f64, dim3(vec_grid(n)), dim3(kVecBlock), 0, s, n, a, b, partials);
    FANTOM_LAUNCH_CHECK("dot_partials");
    return kStatusOk;
}

int launch_dot2_partials(long n, const double *a, const double *b, const double *c, double *partials, hipStream_t s)
{
    hipLaunchKernelGGL(dot2_partials_f64, dim3(vec_grid(n)), dim3(kVecBlock), 0, s, n, a, b, c, partials);
    FANTOM_LAUNCH_CHECK("dot2_partials");
    return kStatusOk;
}

int launch_finalize(const double *partials, long np, double *scal, int op, int slot, hipStream_t s)
{
    hipLaunchKernelGGL(finalize_f64, dim3(1), dim3(1024), 0, s, partials, np, scal, op, slot);
    FANTOM_LAUNCH_CHECK("finalize");
    return kStatusOk;
}

int launch_pcg_xr(long n, double *x, double *r, const double *p, const double *q, const double *scal,
                  double *partials, hipStream_t s)
{
    hipLaunchKernelGGL(pcg_xr_f64, dim3(vec_grid(n)), dim3(kVecBlock), 0, s, n, x, r, p, q, scal, partials);
    FANTOM_LAUNCH_CHECK("pcg_xr");
    return kStatusOk;
}

int launch_pcg_p(long n, double *p, const double *z, const double *scal, int first, hipStream_t s)
{
    hipLaunchKernelGGL(pcg_p_f64, dim3(vec_grid(n)), dim3(kVecBlock), 0, s, n, p, z, scal, first);
    FANTOM_LAUNCH_CHECK("pcg_p");
    return kStatusOk;
}

int launch_scale(long n, double omega, const double *dinv, const double *b, double *x, hipStream_t s,
                 const double *omega_scale)
{
    hipLaunchKernelGGL(scale_f64, dim3(vec_grid(n)), dim3(kVecBlock), 0, s, n, omega, omega_scale, dinv, b, x);
    FANTOM_LAUNCH_CHECK("scale");
    return kStatusOk;
}

int launch_axpby(long n, double a, const double *x, double b, double *y, hipStream_t s)
{
    hipLaunchKernelGGL(axpby_f64, dim3(vec_grid(n)), dim3(kVecBlock), 0, s, n, a, x, b, y);
    FANTOM_LAUNCH_CHECK("axpby");
    return kStatusOk;
}

}  // namespace fantom
