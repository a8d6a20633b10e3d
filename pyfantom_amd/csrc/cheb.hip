// Chebyshev (fourth-kind) smoothing: weights and the spectral-radius estimate.
//
// Reference: the Chebyshev smoother and power method of the CPU MultiGrid
// (solvers/CPU/_solvers.py:884-954, called from :1045/:1066/:1088 where the
// reference leaves the calls commented out) -- Lottes' fourth-kind
// polynomial, z_1 = 4/3 rho^-1 D^-1 r_0, z_k = (2k-3)/(2k+1) z_{k-1} +
// (8k-4)/(2k+1) rho^-1 D^-1 r_{k-1}, x_k = x_{k-1} + z_k, with rho the
// spectral radius of D^-1 A from a power method.
//
// The error after n such steps is p_n(D^-1 A / rho) e with p_n(0) = 1 and
// roots t_i = (1 - cos(2 i pi / (2n + 1))) / 2, i = 1..n, so the same
// smoother is n Jacobi sweeps with weights 1 / (rho t_i): the drivers run
// their Jacobi kernels with those weights (smallest first) and rho read from
// device memory, so graph-captured iterations follow each new estimate.
//
// rho: a few power iterations per setup on D^-1 A in the D inner product,
// v <- D^-1 A v / |v|_D, rho = v.Av / v.Dv (the Rayleigh quotient, which
// rises monotonically to the top eigenvalue), warm-started from the previous
// setup's vector.
#include <cmath>

#include "common.hpp"
#include "fantom_internal.hpp"
#include "vec.hpp"

namespace fantom {

constexpr int kChBlock = 256;

void chebyshev4_weights(int n, double *w)
{
    // 1 / t_i, ascending: the sweep with the smallest weight runs first, so
    // no intermediate iterate amplifies a mode by more than the final error
    for (int i = 0; i < n; ++i)
        w[i] = 2.0 / (1.0 - cos(2.0 * M_PI * (double)(n - i) / (2.0 * n + 1.0)));
}

// deterministic start vector in [-1, 1)
__device__ __forceinline__ double ch_hash(long i, int salt)
{
    unsigned x = (unsigned)i * 2654435761u ^ (unsigned)(salt + 1) * 0x9E3779B9u;
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return (double)(x >> 8) * (2.0 / 16777216.0) - 1.0;
}

// fill (optional) and partial sums of v.D v = sum v^2 / dinv
template <typename T>
__global__ __launch_bounds__(kChBlock) void ch_init_k(long n, T *__restrict__ v, const T *__restrict__ dinv, int fill,
                                                    int salt, double *__restrict__ partials)
{
    double s = 0.0;
    for (long i = (long)blockIdx.x * kChBlock + threadIdx.x; i < n; i += (long)gridDim.x * kChBlock) {
        T x;
        if (fill) {
            x = (T)ch_hash(i, salt);
            v[i] = x;
        } else {
            x = v[i];
        }
        s += (double)x * (double)x / (double)dinv[i];
    }
    s = block_sum<kChBlock>(s);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = s;
}

// with y = A v: partials of v.y (first G) and of |D^-1 y|_D^2 = sum dinv y^2
// (second G); v <- D^-1 y / |v|_D (st[0] = |v|_D^2 of the incoming v)
template <typename T>
__global__ __launch_bounds__(kChBlock) void ch_pass_k(long n, T *__restrict__ v, const T *__restrict__ y,
                                                    const T *__restrict__ dinv, const double *__restrict__ st,
                                                    double *__restrict__ partials)
{
    const double sig = 1.0 / sqrt(st[0]);
    double s1 = 0.0, s3 = 0.0;
    for (long i = (long)blockIdx.x * kChBlock + threadIdx.x; i < n; i += (long)gridDim.x * kChBlock) {
        const double yi = (double)y[i], di = (double)dinv[i];
        s1 += (double)v[i] * yi;
        s3 += di * yi * yi;
        v[i] = (T)(sig * di * yi);
    }
    s1 = block_sum<kChBlock>(s1);
    s3 = block_sum<kChBlock>(s3);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s1;
        partials[gridDim.x + blockIdx.x] = s3;
    }
}

// st = {|v|_D^2, rho, 1 / (rho_scale rho)}; op 0: |v|_D^2 from the init sums;
// op 1: rho = v.Av / |v|_D^2 and the new |v|_D^2 from a pass
__global__ __launch_bounds__(1024) void ch_fin_k(const double *__restrict__ partials, long np, int op,
                                                double rho_scale, double *__restrict__ st)
{
    double a = 0.0, c = 0.0;
    for (long i = threadIdx.x; i < np; i += 1024) {
        a += partials[i];
        if (op == 1)
            c += partials[np + i];
    }
    a = block_sum<1024>(a);
    c = block_sum<1024>(c);
    if (threadIdx.x == 0) {
        if (op == 0) {
            st[0] = a > 0.0 ? a : 1.0;
        } else {
            const double vn = st[0];
            const double rho = a / vn;
            st[1] = rho;
            st[2] = rho > 0.0 ? 1.0 / (rho_scale * rho) : 1.0;
            st[0] = c / vn > 0.0 ? c / vn : 1.0;
        }
    }
}

static unsigned ch_grid(long n)
{
    long b = (n + kChBlock - 1) / kChBlock;
    return (unsigned)std::max(1L, std::min(b, kVecMaxBlocks));
}

long chebyshev_partials(long n) { return 2 * (long)ch_grid(n); }

template <typename T>
int launch_ch_init_t(long n, T *v, const T *dinv, int fill, int salt, double *partials, double *st, hipStream_t s)
{
    const unsigned G = ch_grid(n);
    hipLaunchKernelGGL(ch_init_k<T>, dim3(G), dim3(kChBlock), 0, s, n, v, dinv, fill, salt, partials);
    hipLaunchKernelGGL(ch_fin_k, dim3(1), dim3(1024), 0, s, partials, (long)G, 0, 1.0, st);
    FANTOM_LAUNCH_CHECK("chebyshev init");
    return kStatusOk;
}

template <typename T>
int launch_ch_pass_t(long n, T *v, const T *y, const T *dinv, double rho_scale, double *partials, double *st,
                     hipStream_t s)
{
    const unsigned G = ch_grid(n);
    hipLaunchKernelGGL(ch_pass_k<T>, dim3(G), dim3(kChBlock), 0, s, n, v, y, dinv, st, partials);
    hipLaunchKernelGGL(ch_fin_k, dim3(1), dim3(1024), 0, s, partials, (long)G, 1, rho_scale, st);
    FANTOM_LAUNCH_CHECK("chebyshev pass");
    return kStatusOk;
}

int launch_ch_init(long n, double *v, const double *dinv, int fill, int salt, double *partials, double *st,
                   hipStream_t s)
{
    return launch_ch_init_t<double>(n, v, dinv, fill, salt, partials, st, s);
}
int launch_ch_init_f32(long n, float *v, const float *dinv, int fill, int salt, double *partials, double *st,
                       hipStream_t s)
{
    return launch_ch_init_t<float>(n, v, dinv, fill, salt, partials, st, s);
}
int launch_ch_pass(long n, double *v, const double *y, const double *dinv, double rho_scale, double *partials,
                   double *st, hipStream_t s)
{
    return launch_ch_pass_t<double>(n, v, y, dinv, rho_scale, partials, st, s);
}
int launch_ch_pass_f32(long n, float *v, const float *y, const float *dinv, double rho_scale, double *partials,
                       double *st, hipStream_t s)
{
    return launch_ch_pass_t<float>(n, v, y, dinv, rho_scale, partials, st, s);
}

}  // namespace fantom

extern "C" int fantom_chebyshev4_weights(int n, double *w)
{
    FANTOM_CHECK_ARG(n >= 1 && n <= 16 && w, "n must be in [1, 16] and w non-null");
    fantom::chebyshev4_weights(n, w);
    return fantom::kStatusOk;
}
