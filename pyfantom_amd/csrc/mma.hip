// Method of Moving Asymptotes update on the device (Svanberg's MMA with the
// primal-dual interior-point subproblem solve), as pyFANTOM runs it
// (Optimizers/CPU/MMA.py:103-180 -> mma/CPU/_mma.py:32-165 mmasub,
// :288-528 subsolv) for m <= kMmaMaxM constraints with uniform bounds.
//
// Every per-variable quantity (asymptotes, the convex approximation, the
// Newton right-hand sides, the search direction, the line-search trial point
// and its residual) is computed by fused kernels over the n design variables;
// the m x m dual system, the scalar variables (y, z, lam, mu, zet, s) and the
// interior-point control flow stay on the host, which reads back one record of
// block partial sums per kernel and adds them in block order (deterministic).
// The subproblem takes the m < n branch of the Newton system (topology
// problems have n >> m).
#include <cmath>
#include <vector>

#include "../../include/fantom.h"
#include "common.hpp"
#include "fantom_internal.hpp"

namespace fantom {

constexpr int kMmaMaxM = 8;
constexpr int kMmaBlock = 256;
constexpr int kMmaMaxBlocks = 1024;

// asymptote / approximation constants of mmasub (mma/CPU/_mma.py:32-36)
constexpr double kAsyInit = 0.5, kAsyDecr = 0.7, kAsyIncr = 1.2, kAsyMin = 0.01, kAsyMax = 10.0;
constexpr double kRaa0 = 1e-5, kAlbefa = 0.1;

struct MmaDev {
    long n;
    int m;
    double xmin, xmax;
    const double *x, *xold1, *xold2, *df0dx, *dfdx;   // dfdx: m x n row-major
    double *low, *upp, *alfa, *beta, *p0, *q0, *P, *Q;
    double *xs, *xsi, *eta;      // current subproblem point (n part)
    double *xt, *xsit, *etat;    // line-search trial point
    double *partials;
};

// record layouts of the block partials
//   prep:   b-sums (m)
//   resid:  sumsq, maxabs, gvec (m)
//   newton: gvec (m), GG.(delx/diagx) (m), GG diag(1/diagx) GG^T (m*m)
//   dir:    max step ratio (1)
__device__ inline void block_reduce(double *vals, int nv, bool is_max, double *out)
{
    __shared__ double red[kMmaBlock / 64][kMmaMaxM * kMmaMaxM + 2 * kMmaMaxM + 2];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int k = 0; k < nv; ++k) {
        double v = vals[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(v, off, 64);
            v = is_max ? fmax(v, o) : v + o;
        }
        if (lane == 0)
            red[wid][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < nv) {
        double v = red[0][threadIdx.x];
        for (int w = 1; w < kMmaBlock / 64; ++w)
            v = is_max ? fmax(v, red[w][threadIdx.x]) : v + red[w][threadIdx.x];
        out[(long)blockIdx.x * nv + threadIdx.x] = v;
    }
}

// asymptotes, move bounds, approximation terms, the starting point of the
// subproblem, and the sums b_i = sum_j P_ij / ux_j + Q_ij / xl_j
template <int M>
__global__ __launch_bounds__(kMmaBlock) void mma_prep_k(MmaDev d, int iter, double move)
{
    double bsum[M];
#pragma unroll
    for (int i = 0; i < M; ++i)
        bsum[i] = 0.0;
    const double span = d.xmax - d.xmin;
    const double xmami_inv = 1.0 / fmax(span, 1e-5);
    for (long j = (long)blockIdx.x * kMmaBlock + threadIdx.x; j < d.n; j += (long)gridDim.x * kMmaBlock) {
        const double x = d.x[j];
        double low, upp;
        if (iter <= 2) {
            low = x - kAsyInit * span;
            upp = x + kAsyInit * span;
        } else {
            const double x1 = d.xold1[j], x2 = d.xold2[j];
            const double sgn = (x - x1) * (x1 - x2);
            const double fac = sgn > 0 ? kAsyIncr : (sgn < 0 ? kAsyDecr : 1.0);
            low = x - fac * (x1 - d.low[j]);
            upp = x + fac * (d.upp[j] - x1);
            low = fmin(fmax(low, x - kAsyMax * span), x - kAsyMin * span);
            upp = fmax(fmin(upp, x + kAsyMax * span), x + kAsyMin * span);
        }
        d.low[j] = low;
        d.upp[j] = upp;
        const double alfa = fmax(fmax(low + kAlbefa * (x - low), x - move * span), d.xmin);
        const double beta = fmin(fmin(upp - kAlbefa * (upp - x), x + move * span), d.xmax);
        d.alfa[j] = alfa;
        d.beta[j] = beta;
        const double ux = upp - x, xl = x - low;
        const double g0 = d.df0dx[j];
        double p0 = fmax(g0, 0.0), q0 = fmax(-g0, 0.0);
        const double pq0 = 0.001 * (p0 + q0) + kRaa0 * xmami_inv;
        d.p0[j] = (p0 + pq0) * ux * ux;
        d.q0[j] = (q0 + pq0) * xl * xl;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double g = d.dfdx[(long)i * d.n + j];
            double P = fmax(g, 0.0), Q = fmax(-g, 0.0);
            const double pq = 0.001 * (P + Q) + kRaa0 * xmami_inv;
            P = (P + pq) * ux * ux;
            Q = (Q + pq) * xl * xl;
            d.P[(long)i * d.n + j] = P;
            d.Q[(long)i * d.n + j] = Q;
            bsum[i] += P / ux + Q / xl;
        }
        const double xs = 0.5 * (alfa + beta);
        d.xs[j] = xs;
        d.xsi[j] = fmax(1.0 / (xs - alfa), 1.0);
        d.eta[j] = fmax(1.0 / (beta - xs), 1.0);
    }
    block_reduce(bsum, M, false, d.partials);
}

struct MmaScal {
    double lam[kMmaMaxM];
};

// residual of the subproblem optimality conditions at (x, xsi, eta): the n
// part's sum of squares and max |.|, and gvec_i for the m part
template <int M>
__device__ inline void resid_point(const MmaDev &d, const MmaScal &sc, double epsi, const double *X,
                                   const double *XSI, const double *ETA, double *out)
{
    double sums[1 + M];
#pragma unroll
    for (int i = 0; i < 1 + M; ++i)
        sums[i] = 0.0;
    double mx = 0.0;
    for (long j = (long)blockIdx.x * kMmaBlock + threadIdx.x; j < d.n; j += (long)gridDim.x * kMmaBlock) {
        const double x = X[j], xsi = XSI[j], eta = ETA[j];
        const double ux = d.upp[j] - x, xl = x - d.low[j];
        double plam = d.p0[j], qlam = d.q0[j];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double P = d.P[(long)i * d.n + j], Q = d.Q[(long)i * d.n + j];
            plam += P * sc.lam[i];
            qlam += Q * sc.lam[i];
            sums[1 + i] += P / ux + Q / xl;
        }
        const double rex = plam / (ux * ux) - qlam / (xl * xl) - xsi + eta;
        const double rexsi = xsi * (x - d.alfa[j]) - epsi;
        const double reeta = eta * (d.beta[j] - x) - epsi;
        sums[0] += rex * rex + rexsi * rexsi + reeta * reeta;
        mx = fmax(mx, fmax(fabs(rex), fmax(fabs(rexsi), fabs(reeta))));
    }
    block_reduce(sums, 1 + M, false, out);
    __syncthreads();   // block_reduce's LDS is reused
    block_reduce(&mx, 1, true, out + (long)gridDim.x * (1 + M));
}

template <int M>
__global__ __launch_bounds__(kMmaBlock) void mma_resid_k(MmaDev d, MmaScal sc, double epsi)
{
    resid_point<M>(d, sc, epsi, d.xs, d.xsi, d.eta, d.partials);
}

// Newton right-hand side pieces: gvec, GG (delx / diagx), GG diag(1/diagx) GG^T
template <int M>
__global__ __launch_bounds__(kMmaBlock) void mma_newton_k(MmaDev d, MmaScal sc, double epsi)
{
    constexpr int nv = 2 * M + M * M;
    double v[nv];
#pragma unroll
    for (int k = 0; k < nv; ++k)
        v[k] = 0.0;
    for (long j = (long)blockIdx.x * kMmaBlock + threadIdx.x; j < d.n; j += (long)gridDim.x * kMmaBlock) {
        const double x = d.xs[j], xsi = d.xsi[j], eta = d.eta[j];
        const double alfa = d.alfa[j], beta = d.beta[j];
        const double ux = d.upp[j] - x, xl = x - d.low[j];
        const double ux2 = ux * ux, xl2 = xl * xl;
        double plam = d.p0[j], qlam = d.q0[j];
        double GG[M];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double P = d.P[(long)i * d.n + j], Q = d.Q[(long)i * d.n + j];
            plam += P * sc.lam[i];
            qlam += Q * sc.lam[i];
            v[i] += P / ux + Q / xl;
            GG[i] = P / ux2 - Q / xl2;
        }
        const double delx = plam / ux2 - qlam / xl2 - epsi / (x - alfa) + epsi / (beta - x);
        const double diagx = 2.0 * (plam / (ux2 * ux) + qlam / (xl2 * xl)) + xsi / (x - alfa) + eta / (beta - x);
        const double dd = delx / diagx;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            v[M + i] += GG[i] * dd;
#pragma unroll
            for (int k = 0; k < M; ++k)
                v[2 * M + i * M + k] += GG[i] * GG[k] / diagx;
        }
    }
    block_reduce(v, nv, false, d.partials);
}

// search direction (n part) at element j of the current point (x, xsi, eta)
template <int M>
__device__ __forceinline__ void mma_direction(const MmaDev &d, const MmaScal &sc, const MmaScal &dl, double epsi,
                                              long j, double x, double xsi, double eta, double alfa, double beta,
                                              double &dx, double &dxsi, double &deta)
{
    const double ux = d.upp[j] - x, xl = x - d.low[j];
    const double ux2 = ux * ux, xl2 = xl * xl;
    double plam = d.p0[j], qlam = d.q0[j], gdl = 0.0;
#pragma unroll
    for (int i = 0; i < M; ++i) {
        const double P = d.P[(long)i * d.n + j], Q = d.Q[(long)i * d.n + j];
        plam += P * sc.lam[i];
        qlam += Q * sc.lam[i];
        gdl += (P / ux2 - Q / xl2) * dl.lam[i];
    }
    const double delx = plam / ux2 - qlam / xl2 - epsi / (x - alfa) + epsi / (beta - x);
    const double diagx = 2.0 * (plam / (ux2 * ux) + qlam / (xl2 * xl)) + xsi / (x - alfa) + eta / (beta - x);
    dx = -delx / diagx - gdl / diagx;
    dxsi = -xsi + epsi / (x - alfa) - (xsi * dx) / (x - alfa);
    deta = -eta + epsi / (beta - x) + (eta * dx) / (beta - x);
}

// largest step ratio over the search direction; the direction itself is not
// stored: the line-search trial forms it again where it needs it, which
// costs arithmetic instead of writing and reading back three n-vectors
template <int M>
__global__ __launch_bounds__(kMmaBlock) void mma_dir_k(MmaDev d, MmaScal sc, MmaScal dl, double epsi)
{
    double mx = -INFINITY;
    for (long j = (long)blockIdx.x * kMmaBlock + threadIdx.x; j < d.n; j += (long)gridDim.x * kMmaBlock) {
        const double x = d.xs[j], xsi = d.xsi[j], eta = d.eta[j];
        const double alfa = d.alfa[j], beta = d.beta[j];
        double dx, dxsi, deta;
        mma_direction<M>(d, sc, dl, epsi, j, x, xsi, eta, alfa, beta, dx, dxsi, deta);
        mx = fmax(mx, fmax(fmax(-1.01 * dxsi / xsi, -1.01 * deta / eta),
                           fmax(-1.01 * dx / (x - alfa), 1.01 * dx / (beta - x))));
    }
    block_reduce(&mx, 1, true, d.partials);
}

// line-search trial point with its residual and, in the same pass, the Newton
// sums of mma_newton_k at that point (same arithmetic, same block order): when
// the trial is accepted and the barrier parameter stays, the next Newton step
// takes them from here instead of another pass over the n arrays.  Partials:
// the residual record (sums of squares and gvec per block, then the block
// maxima), then the Newton record at gridDim.x * (2 + M).
template <int M>
__global__ __launch_bounds__(kMmaBlock) void mma_trial_newton_k(MmaDev d, MmaScal sc, MmaScal base, MmaScal dl,
                                                               double steg, double epsi)
{
    constexpr int nv = 2 * M + M * M;
    double sums[1 + M], v[nv];
#pragma unroll
    for (int i = 0; i < 1 + M; ++i)
        sums[i] = 0.0;
#pragma unroll
    for (int k = 0; k < nv; ++k)
        v[k] = 0.0;
    double mx = 0.0;
    for (long j = (long)blockIdx.x * kMmaBlock + threadIdx.x; j < d.n; j += (long)gridDim.x * kMmaBlock) {
        const double alfa = d.alfa[j], beta = d.beta[j];
        const double x0 = d.xs[j], xsi0 = d.xsi[j], eta0 = d.eta[j];
        double dx, dxsi, deta;   // the direction of mma_dir_k at the current point
        mma_direction<M>(d, base, dl, epsi, j, x0, xsi0, eta0, alfa, beta, dx, dxsi, deta);
        const double x = x0 + steg * dx;
        const double xsi = xsi0 + steg * dxsi;
        const double eta = eta0 + steg * deta;
        d.xt[j] = x;
        d.xsit[j] = xsi;
        d.etat[j] = eta;
        const double ux = d.upp[j] - x, xl = x - d.low[j];
        const double ux2 = ux * ux, xl2 = xl * xl;
        double plam = d.p0[j], qlam = d.q0[j];
        double GG[M];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double P = d.P[(long)i * d.n + j], Q = d.Q[(long)i * d.n + j];
            plam += P * sc.lam[i];
            qlam += Q * sc.lam[i];
            const double g = P / ux + Q / xl;
            sums[1 + i] += g;
            v[i] += g;
            GG[i] = P / ux2 - Q / xl2;
        }
        const double rex = plam / ux2 - qlam / xl2 - xsi + eta;
        const double rexsi = xsi * (x - alfa) - epsi;
        const double reeta = eta * (beta - x) - epsi;
        sums[0] += rex * rex + rexsi * rexsi + reeta * reeta;
        mx = fmax(mx, fmax(fabs(rex), fmax(fabs(rexsi), fabs(reeta))));
        const double delx = plam / ux2 - qlam / xl2 - epsi / (x - alfa) + epsi / (beta - x);
        const double diagx = 2.0 * (plam / (ux2 * ux) + qlam / (xl2 * xl)) + xsi / (x - alfa) + eta / (beta - x);
        const double dd = delx / diagx;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            v[M + i] += GG[i] * dd;
#pragma unroll
            for (int k = 0; k < M; ++k)
                v[2 * M + i * M + k] += GG[i] * GG[k] / diagx;
        }
    }
    block_reduce(sums, 1 + M, false, d.partials);
    __syncthreads();   // block_reduce's LDS is reused
    block_reduce(&mx, 1, true, d.partials + (long)gridDim.x * (1 + M));
    __syncthreads();
    block_reduce(v, nv, false, d.partials + (long)gridDim.x * (2 + M));
}

// launch a kernel template instantiated for the constraint count m
#define MMA_DISPATCH(m, KERNEL, grid, ...)                                                                 \
    switch (m) {                                                                                          \
    case 1: hipLaunchKernelGGL(KERNEL<1>, grid, dim3(kMmaBlock), 0, s, __VA_ARGS__); break;               \
    case 2: hipLaunchKernelGGL(KERNEL<2>, grid, dim3(kMmaBlock), 0, s, __VA_ARGS__); break;               \
    case 3: hipLaunchKernelGGL(KERNEL<3>, grid, dim3(kMmaBlock), 0, s, __VA_ARGS__); break;               \
    case 4: hipLaunchKernelGGL(KERNEL<4>, grid, dim3(kMmaBlock), 0, s, __VA_ARGS__); break;               \
    case 5: hipLaunchKernelGGL(KERNEL<5>, grid, dim3(kMmaBlock), 0, s, __VA_ARGS__); break;               \
    case 6: hipLaunchKernelGGL(KERNEL<6>, grid, dim3(kMmaBlock), 0, s, __VA_ARGS__); break;               \
    case 7: hipLaunchKernelGGL(KERNEL<7>, grid, dim3(kMmaBlock), 0, s, __VA_ARGS__); break;               \
    default: hipLaunchKernelGGL(KERNEL<8>, grid, dim3(kMmaBlock), 0, s, __VA_ARGS__); break;              \
    }

static unsigned mma_blocks(long n)
{
    long b = (n + kMmaBlock - 1) / kMmaBlock;
    return (unsigned)(b < kMmaMaxBlocks ? (b > 0 ? b : 1) : kMmaMaxBlocks);
}

}  // namespace fantom

using namespace fantom;

namespace {

struct Reader {
    hipStream_t s;
    std::vector<double> buf;
    int fetch(const double *dev, size_t count)
    {
        buf.resize(count);
        if (hipMemcpyAsync(buf.data(), dev, count * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            set_error("mma readback: %s", hipGetErrorString(hipGetLastError()));
            return kStatusHipError;
        }
        return kStatusOk;
    }
    // sum of record entry k over blocks, in block order
    double sum(unsigned nb, int rec, int k) const
    {
        double v = 0.0;
        for (unsigned b = 0; b < nb; ++b)
            v += buf[(size_t)b * rec + k];
        return v;
    }
    double max(unsigned nb, size_t off, int rec, int k) const
    {
        double v = -INFINITY;
        for (unsigned b = 0; b < nb; ++b)
            v = fmax(v, buf[off + (size_t)b * rec + k]);
        return v;
    }
};

// small dense solve (partial pivoting), n <= kMmaMaxM + 1
bool solve_small(int n, double *A, double *b)
{
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (fabs(A[i * n + k]) > fabs(A[p * n + k]))
                p = i;
        if (A[p * n + k] == 0.0)
            return false;
        if (p != k) {
            for (int j = 0; j < n; ++j)
                std::swap(A[k * n + j], A[p * n + j]);
            std::swap(b[k], b[p]);
        }
        for (int i = k + 1; i < n; ++i) {
            const double f = A[i * n + k] / A[k * n + k];
            for (int j = k; j < n; ++j)
                A[i * n + j] -= f * A[k * n + j];
            b[i] -= f * b[k];
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        double v = b[k];
        for (int j = k + 1; j < n; ++j)
            v -= A[k * n + j] * b[j];
        b[k] = v / A[k * n + k];
    }
    return true;
}

}  // namespace

extern "C" long fantom_mma_workspace_size(long n, int m)
{
    if (n <= 0 || m < 1 || m > kMmaMaxM)
        return -1;
    const long rec = 2 * m + m * m + 2 * (1 + m);
    return (long)sizeof(double) * ((long)(11 + 2 * m) * n + (long)kMmaMaxBlocks * (rec + 2) + 64);
}

extern "C" int fantom_mma_update_f64(long n, int m, int iter, const double *x, const double *xold1,
                                     const double *xold2, double xmin, double xmax, const double *df0dx,
                                     const double *fval, const double *dfdx, double *low, double *upp, double move,
                                     double a0, const double *a, const double *c, const double *d, double sub_tol,
                                     int sub_maxiter, double *x_new, void *work, long work_bytes, int *newton_steps,
                                     void *stream)
{
    FANTOM_CHECK_ARG(n > 0 && m >= 1 && m <= kMmaMaxM, "need n > 0 and 1 <= m <= %d", kMmaMaxM);
    FANTOM_CHECK_ARG(x && df0dx && fval && dfdx && low && upp && a && c && d && x_new && work, "null operand");
    FANTOM_CHECK_ARG(iter <= 2 || (xold1 && xold2), "iterations after the second need xold1 and xold2");
    FANTOM_CHECK_ARG(work_bytes >= fantom_mma_workspace_size(n, m), "workspace too small");
    FANTOM_CHECK_ARG(xmax > xmin, "need xmax > xmin");
    hipStream_t s = as_stream(stream);
    double *w = reinterpret_cast<double *>(work);
    MmaDev dv;
    dv.n = n;
    dv.m = m;
    dv.xmin = xmin;
    dv.xmax = xmax;
    dv.x = x;
    dv.xold1 = xold1 ? xold1 : x;
    dv.xold2 = xold2 ? xold2 : x;
    dv.df0dx = df0dx;
    dv.dfdx = dfdx;
    dv.low = low;
    dv.upp = upp;
    double *cur = w;
    auto take = [&](long cnt) {
        double *p = cur;
        cur += cnt;
        return p;
    };
    dv.alfa = take(n);
    dv.beta = take(n);
    dv.p0 = take(n);
    dv.q0 = take(n);
    dv.P = take((long)m * n);
    dv.Q = take((long)m * n);
    dv.xs = take(n);
    dv.xsi = take(n);
    dv.eta = take(n);
    dv.xt = take(n);
    dv.xsit = take(n);
    dv.etat = take(n);
    dv.partials = take(0);
    const unsigned nb = mma_blocks(n);
    Reader rd{s, {}};

    MMA_DISPATCH(m, mma_prep_k, dim3(nb), dv, iter, move);
    FANTOM_LAUNCH_CHECK("mma_prep");
    int st = rd.fetch(dv.partials, (size_t)nb * m);
    if (st)
        return st;
    double b[kMmaMaxM];
    for (int i = 0; i < m; ++i)
        b[i] = rd.sum(nb, m, i) - fval[i];

    // subproblem state (m part) on the host, mirroring subsolv's initial point
    double y[kMmaMaxM], lam[kMmaMaxM], mu[kMmaMaxM], sl[kMmaMaxM];
    double z = 1.0, zet = 1.0;
    for (int i = 0; i < m; ++i) {
        y[i] = 1.0;
        lam[i] = 1.0;
        mu[i] = fmax(1.0, 0.5 * c[i]);
        sl[i] = 1.0;
    }
    const int rrec = 1 + m;   // residual record: sumsq, gvec[m]; then nb maxabs
    // the current point's multipliers and the dual direction of the Newton
    // step whose line search is running (the trial forms the n direction)
    MmaScal trial_base, trial_dl;
    auto residual = [&](double epsi, bool trial, double steg, double *gvec, double &rnorm, double &rmax,
                        const double *yv, double zv, const double *lamv, const double *muv, double zetv,
                        const double *sv) -> int {
        MmaScal sc;
        for (int i = 0; i < m; ++i)
            sc.lam[i] = lamv[i];
        if (trial) {
            MMA_DISPATCH(m, mma_trial_newton_k, dim3(nb), dv, sc, trial_base, trial_dl, steg, epsi);
        } else {
            MMA_DISPATCH(m, mma_resid_k, dim3(nb), dv, sc, epsi);
        }
        FANTOM_LAUNCH_CHECK("mma_resid");
        int r = rd.fetch(dv.partials, (size_t)nb * (rrec + 1));
        if (r)
            return r;
        double ss = rd.sum(nb, rrec, 0);
        rmax = rd.max(nb, (size_t)nb * rrec, 1, 0);
        double alam = 0.0;
        for (int i = 0; i < m; ++i)
            alam += a[i] * lamv[i];
        for (int i = 0; i < m; ++i) {
            gvec[i] = rd.sum(nb, rrec, 1 + i);
            const double rey = c[i] + d[i] * yv[i] - muv[i] - lamv[i];
            const double relam = gvec[i] - a[i] * zv - yv[i] + sv[i] - b[i];
            const double remu = muv[i] * yv[i] - epsi;
            const double res = lamv[i] * sv[i] - epsi;
            ss += rey * rey + relam * relam + remu * remu + res * res;
            rmax = fmax(rmax, fmax(fmax(fabs(rey), fabs(relam)), fmax(fabs(remu), fabs(res))));
        }
        const double rez = a0 - zetv - alam, rezet = zetv * zv - epsi;
        ss += rez * rez + rezet * rezet;
        rmax = fmax(rmax, fmax(fabs(rez), fabs(rezet)));
        rnorm = sqrt(ss);
        return kStatusOk;
    };

    double epsi = 1.0;
    int outer = 0, steps = 0;
    double gvec[kMmaMaxM];
    while (epsi > sub_tol && outer < sub_maxiter) {
        ++outer;
        double rnorm, rmax;
        st = residual(epsi, false, 0.0, gvec, rnorm, rmax, y, z, lam, mu, zet, sl);
        if (st)
            return st;
        int ittt = 0;
        // the accepted trial of the previous step left the Newton sums of
        // this point (same barrier parameter) behind its residual record
        bool have_sums = false;
        while (rmax > 0.9 * epsi && ittt < 200) {
            ++ittt;
            ++steps;
            MmaScal sc;
            for (int i = 0; i < m; ++i)
                sc.lam[i] = lam[i];
            const int nrec = 2 * m + m * m;
            if (have_sums) {
                st = rd.fetch(dv.partials + (size_t)nb * (2 + m), (size_t)nb * nrec);
            } else {
                MMA_DISPATCH(m, mma_newton_k, dim3(nb), dv, sc, epsi);
                FANTOM_LAUNCH_CHECK("mma_newton");
                st = rd.fetch(dv.partials, (size_t)nb * nrec);
            }
            if (st)
                return st;
            double gv[kMmaMaxM], sgd[kMmaMaxM], AA[(kMmaMaxM + 1) * (kMmaMaxM + 1)], rhs[kMmaMaxM + 1];
            for (int i = 0; i < m; ++i) {
                gv[i] = rd.sum(nb, nrec, i);
                sgd[i] = rd.sum(nb, nrec, m + i);
            }
            double dely[kMmaMaxM], diagy[kMmaMaxM];
            double alam = 0.0;
            for (int i = 0; i < m; ++i)
                alam += a[i] * lam[i];
            const double delz = a0 - alam - epsi / z;
            const int M1 = m + 1;
            for (int i = 0; i < m; ++i) {
                dely[i] = c[i] + d[i] * y[i] - lam[i] - epsi / y[i];
                const double dellam = gv[i] - a[i] * z - y[i] - b[i] + epsi / lam[i];
                diagy[i] = d[i] + mu[i] / y[i];
                const double diaglamyi = sl[i] / lam[i] + 1.0 / diagy[i];
                rhs[i] = dellam + dely[i] / diagy[i] - sgd[i];
                for (int k = 0; k < m; ++k)
                    AA[i * M1 + k] = rd.sum(nb, nrec, 2 * m + i * m + k) + (i == k ? diaglamyi : 0.0);
                AA[i * M1 + m] = a[i];
                AA[m * M1 + i] = a[i];
            }
            AA[m * M1 + m] = -zet / z;
            rhs[m] = delz;
            if (!solve_small(M1, AA, rhs)) {
                set_error("MMA Newton system is singular");
                return kStatusHipError;
            }
            MmaScal dl;
            for (int i = 0; i < m; ++i)
                dl.lam[i] = rhs[i];
            trial_base = sc;
            trial_dl = dl;
            const double dz = rhs[m];
            MMA_DISPATCH(m, mma_dir_k, dim3(nb), dv, sc, dl, epsi);
            FANTOM_LAUNCH_CHECK("mma_dir");
            st = rd.fetch(dv.partials, (size_t)nb);
            if (st)
                return st;
            double stm = rd.max(nb, 0, 1, 0);
            double dy[kMmaMaxM], dmu[kMmaMaxM], ds[kMmaMaxM];
            for (int i = 0; i < m; ++i) {
                dy[i] = -dely[i] / diagy[i] + dl.lam[i] / diagy[i];
                dmu[i] = -mu[i] + epsi / y[i] - (mu[i] * dy[i]) / y[i];
                ds[i] = -sl[i] + epsi / lam[i] - (sl[i] * dl.lam[i]) / lam[i];
                stm = fmax(stm, fmax(fmax(-1.01 * dy[i] / y[i], -1.01 * dl.lam[i] / lam[i]),
                                     fmax(-1.01 * dmu[i] / mu[i], -1.01 * ds[i] / sl[i])));
            }
            const double dzet = -zet + epsi / z - zet * dz / z;
            stm = fmax(stm, fmax(-1.01 * dz / z, -1.01 * dzet / zet));
            double steg = 1.0 / fmax(stm, 1.0);
            double yo[kMmaMaxM], lamo[kMmaMaxM], muo[kMmaMaxM], so[kMmaMaxM];
            for (int i = 0; i < m; ++i) {
                yo[i] = y[i];
                lamo[i] = lam[i];
                muo[i] = mu[i];
                so[i] = sl[i];
            }
            const double zo = z, zeto = zet;
            // a non-finite residual would skip the line search below and leave
            // the trial arrays unwritten: stop with an error instead
            if (!std::isfinite(rnorm) || !std::isfinite(rmax)) {
                set_error("MMA subproblem residual is not finite (%g)", rnorm);
                return kStatusHipError;
            }
            double resinew = 2.0 * rnorm, rmx = rmax;
            int itto = 0;
            while (resinew > rnorm && itto < 50) {
                ++itto;
                for (int i = 0; i < m; ++i) {
                    y[i] = yo[i] + steg * dy[i];
                    lam[i] = lamo[i] + steg * dl.lam[i];
                    mu[i] = muo[i] + steg * dmu[i];
                    sl[i] = so[i] + steg * ds[i];
                }
                z = zo + steg * dz;
                zet = zeto + steg * dzet;
                st = residual(epsi, true, steg, gvec, resinew, rmx, y, z, lam, mu, zet, sl);
                if (st)
                    return st;
                steg /= 2.0;
            }
            // accept the last trial point (as the reference does, also after 50
            // halvings); the trial arrays hold a point only once a trial ran
            if (itto > 0) {
                std::swap(dv.xs, dv.xt);
                std::swap(dv.xsi, dv.xsit);
                std::swap(dv.eta, dv.etat);
                rnorm = resinew;
                rmax = rmx;
                have_sums = true;
            }
        }
        epsi *= 0.1;
    }
    if (hipMemcpyAsync(x_new, dv.xs, sizeof(double) * n, hipMemcpyDeviceToDevice, s) != hipSuccess) {
        set_error("mma copy: %s", hipGetErrorString(hipGetLastError()));
        return kStatusHipError;
    }
    if (newton_steps)
        *newton_steps = steps;
    return kStatusOk;
}
