// Optimality-criteria update on the device (FP64).
//
// Reference: Optimizers/CPU/OC.py:82-140 (single material).
//   ocP   = x * sqrt(max(-df/dg, 0))  (NaN -> 0); if sum|ocP| == 0, ocP = 1e-3
//   bisection on l in [1e-9, 1e9] while (l2 - l1)/(l2 + l1) > 1e-8:
//     l_mid = (l1 + l2)/2
//     x_new = max(0, max(xL, min(1, min(xU, ocP / l_mid))))
//     if sum(dg * x_new) - vf <= 0: l2 = l_mid else l1 = l_mid
//   with xU = clip(x + move, 0, 1), xL = clip(x - move, 0, 1); the result is
//   x_new of the last evaluated l_mid.
//
// The device version evaluates the constraint for the 2^L - 1 midpoints of the
// next L bisection levels in one pass over the data, then a single block walks
// those L levels exactly as the sequential loop would (same midpoints, same
// comparisons), so the sequence of l_mid values is the reference's while
// taking L times fewer passes and no host synchronisation.
#include <cfloat>
#include <mutex>

#include "../../include/fantom.h"
#include "common.hpp"
#include "fantom_internal.hpp"
#include "vec.hpp"

namespace fantom {

constexpr int kOcBlock = 256;
constexpr int kOcLevels = 4;
constexpr int kOcCand = (1 << kOcLevels) - 1;
// state slots
enum { OC_L1 = 0, OC_L2 = 1, OC_LMID = 2, OC_DONE = 3, OC_ZERO = 4, OC_G = 5 };

__device__ __forceinline__ double oc_x(double x, double ocp, double move, double lam)
{
    const double xu = fmin(fmax(x + move, 0.0), 1.0);
    const double xl = fmin(fmax(x - move, 0.0), 1.0);
    return fmax(0.0, fmax(xl, fmin(1.0, fmin(xu, ocp / lam))));
}

__global__ __launch_bounds__(kOcBlock) void oc_prepare_k(long n, const double *__restrict__ x,
                                                       const double *__restrict__ df,
                                                       const double *__restrict__ dg, double *__restrict__ ocp,
                                                       double *__restrict__ partials)
{
    double s = 0.0;
    for (long i = (long)blockIdx.x * kOcBlock + threadIdx.x; i < n; i += (long)gridDim.x * kOcBlock) {
        double q = -df[i] / dg[i];
        double r = sqrt(fmax(q, 0.0));
        if (isnan(q) || isnan(r))
            r = 0.0;
        else if (isinf(r))
            r = DBL_MAX;
        const double v = x[i] * r;
        ocp[i] = v;
        s += fabs(v);
    }
    s = block_sum<kOcBlock>(s);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(1024) void oc_init_k(const double *__restrict__ partials, long np,
                                                double *__restrict__ state)
{
    double s = 0.0;
    for (long i = threadIdx.x; i < np; i += 1024)
        s += partials[i];
    s = block_sum<1024>(s);
    if (threadIdx.x == 0) {
        state[OC_L1] = 1e-9;
        state[OC_L2] = 1e9;
        state[OC_LMID] = 0.0;
        state[OC_DONE] = 0.0;
        state[OC_ZERO] = (s == 0.0) ? 1.0 : 0.0;
        state[OC_G] = 0.0;
    }
}

// midpoint of heap node h (1-based) below the bracket [l1, l2]
__device__ __forceinline__ double heap_mid(double a, double b, int h)
{
    int depth = 31 - __clz(h);
    for (int d = depth - 1; d >= 0; --d) {
        const double m = (a + b) / 2;
        if ((h >> d) & 1)
            a = m;
        else
            b = m;
    }
    return (a + b) / 2;
}

__global__ __launch_bounds__(kOcBlock) void oc_candidates_k(long n, const double *__restrict__ x,
                                                          const double *__restrict__ ocp,
                                                          const double *__restrict__ dg, double move,
                                                          const double *__restrict__ state,
                                                          double *__restrict__ partials)
{
    if (state[OC_DONE] != 0.0)
        return;
    const double l1 = state[OC_L1], l2 = state[OC_L2];
    const bool zero = state[OC_ZERO] != 0.0;
    double lam[kOcCand], acc[kOcCand];
#pragma unroll
    for (int c = 0; c < kOcCand; ++c) {
        lam[c] = heap_mid(l1, l2, c + 1);
        acc[c] = 0.0;
    }
    for (long i = (long)blockIdx.x * kOcBlock + threadIdx.x; i < n; i += (long)gridDim.x * kOcBlock) {
        const double xi = x[i], oi = zero ? 1e-3 : ocp[i], gi = dg[i];
#pragma unroll
        for (int c = 0; c < kOcCand; ++c)
            acc[c] += gi * oc_x(xi, oi, move, lam[c]);
    }
#pragma unroll
    for (int c = 0; c < kOcCand; ++c) {
        const double s = block_sum<kOcBlock>(acc[c]);
        if (threadIdx.x == 0)
            partials[(long)c * gridDim.x + blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(1024) void oc_walk_k(const double *__restrict__ partials, long np, double vf,
                                                double *__restrict__ state)
{
    __shared__ double S[kOcCand];
    if (state[OC_DONE] != 0.0)
        return;
    for (int c = 0; c < kOcCand; ++c) {
        double s = 0.0;
        for (long i = threadIdx.x; i < np; i += 1024)
            s += partials[(long)c * np + i];
        s = block_sum<1024>(s);
        if (threadIdx.x == 0)
            S[c] = s;
    }
    if (threadIdx.x != 0)
        return;
    double l1 = state[OC_L1], l2 = state[OC_L2], lmid = state[OC_LMID];
    int h = 1;
    bool done = false;
    for (int lev = 0; lev < kOcLevels; ++lev) {
        if (!((l2 - l1) / (l2 + l1) > 1e-8)) {
            done = true;
            break;
        }
        lmid = (l1 + l2) / 2;
        const double g = S[h - 1] - vf;
        state[OC_G] = g;
        if (g <= 0.0) {
            l2 = lmid;
            h = 2 * h;
        } else {
            l1 = lmid;
            h = 2 * h + 1;
        }
    }
    if (!done && !((l2 - l1) / (l2 + l1) > 1e-8))
        done = true;
    state[OC_L1] = l1;
    state[OC_L2] = l2;
    state[OC_LMID] = lmid;
    state[OC_DONE] = done ? 1.0 : 0.0;
}

__global__ __launch_bounds__(kOcBlock) void oc_apply_k(long n, const double *__restrict__ x,
                                                     const double *__restrict__ ocp, double move,
                                                     const double *__restrict__ state, double *__restrict__ xnew,
                                                     double *__restrict__ partials)
{
    const double lam = state[OC_LMID];
    const bool zero = state[OC_ZERO] != 0.0;
    double s = 0.0;
    for (long i = (long)blockIdx.x * kOcBlock + threadIdx.x; i < n; i += (long)gridDim.x * kOcBlock) {
        const double xi = x[i];
        const double v = oc_x(xi, zero ? 1e-3 : ocp[i], move, lam);
        xnew[i] = v;
        s += (xi - v) * (xi - v);
    }
    s = block_sum<kOcBlock>(s);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = s;
}

static unsigned oc_grid(long n)
{
    long b = (n + kOcBlock - 1) / kOcBlock;
    if (b > 1024)
        b = 1024;
    return (unsigned)(b < 1 ? 1 : b);
}

long oc_partials(long n) { return (long)oc_grid(n); }

int launch_oc_prepare(long n, const double *x, const double *df, const double *dg, double *ocp, double *state,
                      double *partials, hipStream_t s)
{
    const unsigned G = oc_grid(n);
    hipLaunchKernelGGL(oc_prepare_k, dim3(G), dim3(kOcBlock), 0, s, n, x, df, dg, ocp, partials);
    hipLaunchKernelGGL(oc_init_k, dim3(1), dim3(1024), 0, s, partials, (long)G, state);
    FANTOM_LAUNCH_CHECK("oc_prepare");
    return kStatusOk;
}

int launch_oc_bisect(long n, const double *x, const double *ocp, const double *dg, double move, double vf,
                     double *state, double *partials, int max_passes, hipStream_t s)
{
    const unsigned G = oc_grid(n);
    for (int p = 0; p < max_passes; ++p) {
        hipLaunchKernelGGL(oc_candidates_k, dim3(G), dim3(kOcBlock), 0, s, n, x, ocp, dg, move, state, partials);
        hipLaunchKernelGGL(oc_walk_k, dim3(1), dim3(1024), 0, s, partials, (long)G, vf, state);
    }
    FANTOM_LAUNCH_CHECK("oc_bisect");
    return kStatusOk;
}

int launch_oc_apply(long n, const double *x, const double *ocp, double move, const double *state, double *xnew,
                    double *partials, hipStream_t s)
{
    const unsigned G = oc_grid(n);
    hipLaunchKernelGGL(oc_apply_k, dim3(G), dim3(kOcBlock), 0, s, n, x, ocp, move, state, xnew, partials);
    FANTOM_LAUNCH_CHECK("oc_apply");
    return kStatusOk;
}

}  // namespace fantom

// ---- allocation-free variant -------------------------------------------------
// fantom_oc_update_noalloc_f64 runs the same kernels as fantom_oc_update_f64
// with its three scratch buffers (bisection state, per-block partial sums,
// the ocP vector) in one file-scope device array instead of caller memory.
// Each device holds its own copy of the array for the life of the process;
// host code only ever uses the address hipGetSymbolAddress reports for the
// current device, never the array's name.  Every call on a device shares that
// device's array, so calls that overlap on one device (two streams, two host
// threads) race and must be ordered by the caller.
namespace fantom {

constexpr long kOcNoAllocMaxN = 1L << 22;                 // 4,194,304 design variables
constexpr long kOcScrState = 0;                           // double[16]
constexpr long kOcScrPartials = 16;                       // double[16 * 1024]
constexpr long kOcScrOcp = kOcScrPartials + 16 * 1024;    // double[kOcNoAllocMaxN]
constexpr long kOcScrDoubles = kOcScrOcp + kOcNoAllocMaxN;  // 33.7 MB in all

static __device__ __align__(256) double scratch[kOcScrDoubles];

constexpr int kOcMaxDevices = 64;
static double *g_oc_scratch[kOcMaxDevices];
static std::mutex g_oc_scratch_mu;

// Device address of this device's scratch array, looked up once per device
// with that device current (the caller's current device).
static int oc_scratch_ptr(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error("oc_update_noalloc: hipGetDevice: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    FANTOM_CHECK_ARG(dev >= 0 && dev < kOcMaxDevices, "oc_update_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_oc_scratch_mu);
    if (!g_oc_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("oc_update_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return kStatusHipError;
        }
        g_oc_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_oc_scratch[dev];
    return kStatusOk;
}

__global__ void oc_lambda_k(const double *__restrict__ state, double *__restrict__ result)
{
    if (threadIdx.x == 0)
        result[0] = state[OC_LMID];
}

}  // namespace fantom

extern "C" int fantom_oc_update_noalloc_f64(long n, const double *x, const double *df, const double *dg, double vf,
                                            double move, double *x_new, double *result, void *stream)
{
    using namespace fantom;
    FANTOM_CHECK_ARG(n > 0 && x && df && dg && x_new && result, "bad operands");
    FANTOM_CHECK_ARG(n <= kOcNoAllocMaxN, "oc_update_noalloc: n = %ld exceeds the scratch capacity of %ld", n,
                     kOcNoAllocMaxN);
    FANTOM_CHECK_ARG(x != x_new, "x and x_new must not alias");
    double *scr = nullptr;
    int st = oc_scratch_ptr(&scr);
    if (st)
        return st;
    double *state = scr + kOcScrState, *partials = scr + kOcScrPartials, *ocp = scr + kOcScrOcp;
    hipStream_t s = as_stream(stream);
    // oc_prepare/oc_init write every state slot and partial they later read
    st = launch_oc_prepare(n, x, df, dg, ocp, state, partials, s);
    if (st)
        return st;
    st = launch_oc_bisect(n, x, ocp, dg, move, vf, state, partials, 24, s);
    if (st)
        return st;
    st = launch_oc_apply(n, x, ocp, move, state, x_new, partials, s);
    if (st)
        return st;
    st = launch_finalize(partials, oc_partials(n), result, FIN_STORE, 1, s);
    if (st)
        return st;
    hipLaunchKernelGGL(oc_lambda_k, dim3(1), dim3(64), 0, s, state, result);
    FANTOM_LAUNCH_CHECK("oc_update_noalloc");
    return kStatusOk;
}
