// Method of Moving Asymptotes update on the device (Svanberg's MMA with the
// primal-dual interior-point subproblem solve), as pyFANTOM runs it
// (Optimizers/CPU/MMA.py:103-180 -> mma/CPU/_mma.py:32-165 mmasub,
// :288-528 subsolv) for m <= kMmaMaxM constraints with uniform bounds.
//
// Every per-variable quantity (asymptotes, the convex approximation, the
// Newton right-hand sides, the search direction, the line-search trial point
// and its residual) is computed by fused passes over the n design variables.
// The m x m dual system, the scalar variables (y, z, lam, mu, zet, s) and the
// interior-point control flow run on the device too: subsolv's loops become a
// state machine whose state (MmaCtl) lives in the workspace.  One unit of
// work is two launches: mma_pass_k runs the n pass the current phase needs
// (prep, residual, Newton sums, direction, line-search trial), and
// mma_ctl_k adds its block partials in a fixed order (deterministic) and takes
// the next control step (the dual solve, step lengths, acceptance, barrier
// updates), choosing the next phase.  The host enqueues units in batches and
// reads one small status record per batch; the first batch is sized from the
// previous update's unit count, so an update normally needs a single read, at
// its end.  Units after the end are no-ops.  The subproblem takes the m < n
// branch of the Newton system (topology problems have n >> m).
#include <cmath>
#include <algorithm>
#include <atomic>

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
__device__ inline void mma_prep(const MmaDev &d, int iter, double move)
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

// Newton right-hand side pieces: gvec, GG (delx / diagx), GG diag(1/diagx) GG^T
template <int M>
__device__ inline void mma_newton(const MmaDev &d, const MmaScal &sc, double epsi)
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
__device__ inline void mma_dir(const MmaDev &d, const MmaScal &sc, const MmaScal &dl, double epsi)
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
// sums of mma_newton at that point (same arithmetic, same block order): when
// the trial is accepted and the barrier parameter stays, the next Newton step
// takes them from here instead of another pass over the n arrays.  Partials:
// the residual record (sums of squares and gvec per block, then the block
// maxima), then the Newton record at gridDim.x * (2 + M).
template <int M>
__device__ inline void mma_trial_newton(const MmaDev &d, const MmaScal &sc, const MmaScal &base, const MmaScal &dl,
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
        double dx, dxsi, deta;   // the direction of mma_dir at the current point
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


// ---------------------------------------------------------------------------
// Device control of the subproblem solve
// ---------------------------------------------------------------------------

// which n pass runs next
enum MmaPhase : int { kPhPrep = 0, kPhResid = 1, kPhNewton = 2, kPhDir = 3, kPhTrial = 4, kPhDone = 5 };
enum MmaCtlStatus : int { kCtlOk = 0, kCtlSingular = 1, kCtlNonFinite = 2 };

// the m part of subsolv's state and its loop counters; the first ints are the
// status record the host reads
struct MmaCtl {
    int phase, status, steps, units;
    int par;         // 1: the current n point sits in the trial arrays
    int outer, ittt, itto, have_sums, pad;
    double epsi, steg, rnorm, rmax, z, zet, zo, zeto, dz, dzet;
    double b[kMmaMaxM], y[kMmaMaxM], lam[kMmaMaxM], mu[kMmaMaxM], sl[kMmaMaxM];
    double yo[kMmaMaxM], lamo[kMmaMaxM], muo[kMmaMaxM], so[kMmaMaxM];
    double dy[kMmaMaxM], dl[kMmaMaxM], dmu[kMmaMaxM], ds[kMmaMaxM], base[kMmaMaxM];
};
constexpr int kMmaCtlHeaderInts = 4;

// the problem's m-part constants (host arrays of the entry point)
struct MmaConst {
    int m, sub_maxiter;
    double a0, sub_tol;
    double a[kMmaMaxM], c[kMmaMaxM], d[kMmaMaxM], fval[kMmaMaxM];
};

// the n pass of the current phase; the current point is (xs, xsi, eta), or
// the trial arrays when par is set (an accepted trial swaps the two sets)
template <int M>
__global__ __launch_bounds__(kMmaBlock) void mma_pass_k(MmaDev d, const MmaCtl *__restrict__ ctl, int iter,
                                                       double move)
{
    const int phase = ctl->phase;
    if (phase == kPhDone)
        return;
    if (phase == kPhPrep) {
        mma_prep<M>(d, iter, move);
        return;
    }
    MmaDev cur = d;
    if (ctl->par) {
        cur.xs = d.xt;
        cur.xsi = d.xsit;
        cur.eta = d.etat;
        cur.xt = d.xs;
        cur.xsit = d.xsi;
        cur.etat = d.eta;
    }
    MmaScal sc, base, dl;
#pragma unroll
    for (int i = 0; i < M; ++i) {
        sc.lam[i] = ctl->lam[i];
        base.lam[i] = ctl->base[i];
        dl.lam[i] = ctl->dl[i];
    }
    const double epsi = ctl->epsi;
    switch (phase) {
    case kPhResid: resid_point<M>(cur, sc, epsi, cur.xs, cur.xsi, cur.eta, cur.partials); break;
    case kPhNewton: mma_newton<M>(cur, sc, epsi); break;
    case kPhDir: mma_dir<M>(cur, base, dl, epsi); break;
    default: mma_trial_newton<M>(cur, sc, base, dl, ctl->steg, epsi); break;
    }
}

// small dense solve (partial pivoting), n <= kMmaMaxM + 1
__device__ bool mma_solve_small(int n, double *A, double *b)
{
#pragma clang fp contract(off)
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (fabs(A[i * n + k]) > fabs(A[p * n + k]))
                p = i;
        if (A[p * n + k] == 0.0)
            return false;
        if (p != k) {
            for (int j = 0; j < n; ++j) {
                const double t = A[k * n + j];
                A[k * n + j] = A[p * n + j];
                A[p * n + j] = t;
            }
            const double t = b[k];
            b[k] = b[p];
            b[p] = t;
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

// the full residual norm and max from the n part's sums (red[0] = sum of
// squares, red[1 + i] = gvec_i, red[1 + m] = max) and the m part at (y, z,
// lam, mu, zet, s)
__device__ void mma_m_residual(const MmaConst &k, const MmaCtl &c, const double *red, double epsi, double &rnorm,
                               double &rmax)
{
#pragma clang fp contract(off)
    const int m = k.m;
    double ss = red[0];
    rmax = red[1 + m];
    double alam = 0.0;
    for (int i = 0; i < m; ++i)
        alam += k.a[i] * c.lam[i];
    for (int i = 0; i < m; ++i) {
        const double gvec = red[1 + i];
        const double rey = k.c[i] + k.d[i] * c.y[i] - c.mu[i] - c.lam[i];
        const double relam = gvec - k.a[i] * c.z - c.y[i] + c.sl[i] - c.b[i];
        const double remu = c.mu[i] * c.y[i] - epsi;
        const double res = c.lam[i] * c.sl[i] - epsi;
        ss += rey * rey + relam * relam + remu * remu + res * res;
        rmax = fmax(rmax, fmax(fmax(fabs(rey), fabs(relam)), fmax(fabs(remu), fabs(res))));
    }
    const double rez = k.a0 - c.zet - alam, rezet = c.zet * c.z - epsi;
    ss += rez * rez + rezet * rezet;
    rmax = fmax(rmax, fmax(fabs(rez), fabs(rezet)));
    rnorm = sqrt(ss);
}

// one Newton step's dual system from the Newton record nr (gvec m, GG.dd m,
// the m x m product); sets the dual direction and the next phase (DIR)
__device__ void mma_newton_solve(const MmaConst &k, MmaCtl &c, const double *nr)
{
#pragma clang fp contract(off)
    const int m = k.m, M1 = m + 1;
    double AA[(kMmaMaxM + 1) * (kMmaMaxM + 1)], rhs[kMmaMaxM + 1];
    const double epsi = c.epsi;
    double alam = 0.0;
    for (int i = 0; i < m; ++i)
        alam += k.a[i] * c.lam[i];
    const double delz = k.a0 - alam - epsi / c.z;
    double dely[kMmaMaxM], diagy[kMmaMaxM];
    for (int i = 0; i < m; ++i) {
        dely[i] = k.c[i] + k.d[i] * c.y[i] - c.lam[i] - epsi / c.y[i];
        const double dellam = nr[i] - k.a[i] * c.z - c.y[i] - c.b[i] + epsi / c.lam[i];
        diagy[i] = k.d[i] + c.mu[i] / c.y[i];
        const double diaglamyi = c.sl[i] / c.lam[i] + 1.0 / diagy[i];
        rhs[i] = dellam + dely[i] / diagy[i] - nr[m + i];
        for (int q = 0; q < m; ++q)
            AA[i * M1 + q] = nr[2 * m + i * m + q] + (i == q ? diaglamyi : 0.0);
        AA[i * M1 + m] = k.a[i];
        AA[m * M1 + i] = k.a[i];
    }
    AA[m * M1 + m] = -c.zet / c.z;
    rhs[m] = delz;
    if (!mma_solve_small(M1, AA, rhs)) {
        c.status = kCtlSingular;
        c.phase = kPhDone;
        return;
    }
    for (int i = 0; i < m; ++i) {
        c.dl[i] = rhs[i];
        c.base[i] = c.lam[i];
        c.dy[i] = -dely[i] / diagy[i] + rhs[i] / diagy[i];
    }
    c.dz = rhs[m];
    c.phase = kPhDir;
}

// subsolv's inner-loop test: another Newton step at this barrier level, or
// the next level (or the end)
__device__ void mma_next_step(const MmaConst &k, MmaCtl &c, const double *trial_newton)
{
    if (c.rmax > 0.9 * c.epsi && c.ittt < 200) {
        ++c.ittt;
        ++c.steps;
        if (c.have_sums) {
            mma_newton_solve(k, c, trial_newton);
        } else {
            c.phase = kPhNewton;
        }
        return;
    }
    c.epsi *= 0.1;
    if (c.epsi > k.sub_tol && c.outer < k.sub_maxiter) {
        ++c.outer;
        c.phase = kPhResid;
    } else {
        c.phase = kPhDone;
    }
}

// set the trial m part for step length steg and request the trial pass
__device__ void mma_trial_point(const MmaConst &k, MmaCtl &c)
{
#pragma clang fp contract(off)
    ++c.itto;
    const double steg = c.steg;
    for (int i = 0; i < k.m; ++i) {
        c.y[i] = c.yo[i] + steg * c.dy[i];
        c.lam[i] = c.lamo[i] + steg * c.dl[i];
        c.mu[i] = c.muo[i] + steg * c.dmu[i];
        c.sl[i] = c.so[i] + steg * c.ds[i];
    }
    c.z = c.zo + steg * c.dz;
    c.zet = c.zeto + steg * c.dzet;
    c.phase = kPhTrial;
}

constexpr int kCtlThreads = 1024;
constexpr int kCtlWaves = kCtlThreads / 64;

// reduce the record the last pass left and take the next control step.
// Each record entry is reduced by one wave: lane l adds blocks l, l + 64, ...
// in that order, then a fixed butterfly combines the lanes, so the sums are
// the same on every run (a single thread walking the blocks serialised 1024
// dependent loads per entry).
__global__ __launch_bounds__(kCtlThreads) void mma_ctl_k(MmaCtl *__restrict__ ctl,
                                                         const double *__restrict__ partials, unsigned nb,
                                                         MmaConst k)
{
#pragma clang fp contract(off)
    __shared__ double red[4 + 3 * kMmaMaxM + kMmaMaxM * kMmaMaxM];
    const int phase = ctl->phase;
    if (phase == kPhDone)
        return;
    const int m = k.m, nrec = 2 * m + m * m, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    // entries: [0, 1 + m) residual sums, 1 + m residual max, then the Newton
    // record (newton pass: at 0; trial pass: behind the residual record)
    const int ne = phase == kPhPrep ? m : phase == kPhNewton ? nrec : phase == kPhDir ? 1
                 : phase == kPhResid ? 2 + m : 2 + m + nrec;
    for (int e = wv; e < ne; e += kCtlWaves) {
        long off;
        int rec, idx, slot;
        bool is_max = false;
        if (phase == kPhPrep) {
            off = 0, rec = m, idx = e, slot = e;
        } else if (phase == kPhNewton) {
            off = 0, rec = nrec, idx = e, slot = 2 + m + e;
        } else if (phase == kPhDir) {
            off = 0, rec = 1, idx = 0, slot = 0, is_max = true;
        } else if (e < 1 + m) {
            off = 0, rec = 1 + m, idx = e, slot = e;
        } else if (e == 1 + m) {
            off = (long)nb * (1 + m), rec = 1, idx = 0, slot = e, is_max = true;
        } else {
            off = (long)nb * (2 + m), rec = nrec, idx = e - 2 - m, slot = e;
        }
        double v = is_max ? -INFINITY : 0.0;
#pragma unroll
        for (int q = 0; q < kMmaMaxBlocks / 64; ++q) {
            const unsigned b = (unsigned)(lane + 64 * q);
            if (b < nb) {
                const double x = partials[off + (long)b * rec + idx];
                v = is_max ? fmax(v, x) : v + x;
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double w = __shfl_xor(v, o, 64);
            v = is_max ? fmax(v, w) : v + w;
        }
        if (lane == 0)
            red[slot] = v;
    }
    __syncthreads();
    if (t != 0)
        return;
    MmaCtl c = *ctl;
    ++c.units;
    const double *newton_rec = red + 2 + m;
    switch (phase) {
    case kPhPrep: {
        for (int i = 0; i < m; ++i) {
            c.b[i] = red[i] - k.fval[i];
            c.y[i] = 1.0;
            c.lam[i] = 1.0;
            c.mu[i] = fmax(1.0, 0.5 * k.c[i]);
            c.sl[i] = 1.0;
        }
        c.z = 1.0;
        c.zet = 1.0;
        c.epsi = 1.0;
        c.outer = 0;
        if (c.epsi > k.sub_tol && c.outer < k.sub_maxiter) {
            ++c.outer;
            c.phase = kPhResid;
        } else {
            c.phase = kPhDone;
        }
        break;
    }
    case kPhResid: {
        mma_m_residual(k, c, red, c.epsi, c.rnorm, c.rmax);
        c.ittt = 0;
        c.have_sums = 0;
        mma_next_step(k, c, newton_rec);
        break;
    }
    case kPhNewton: mma_newton_solve(k, c, newton_rec); break;
    case kPhDir: {
        const double epsi = c.epsi;
        double stm = red[0];
        for (int i = 0; i < m; ++i) {
            c.dmu[i] = -c.mu[i] + epsi / c.y[i] - (c.mu[i] * c.dy[i]) / c.y[i];
            c.ds[i] = -c.sl[i] + epsi / c.lam[i] - (c.sl[i] * c.dl[i]) / c.lam[i];
            stm = fmax(stm, fmax(fmax(-1.01 * c.dy[i] / c.y[i], -1.01 * c.dl[i] / c.lam[i]),
                                 fmax(-1.01 * c.dmu[i] / c.mu[i], -1.01 * c.ds[i] / c.sl[i])));
        }
        c.dzet = -c.zet + epsi / c.z - c.zet * c.dz / c.z;
        stm = fmax(stm, fmax(-1.01 * c.dz / c.z, -1.01 * c.dzet / c.zet));
        c.steg = 1.0 / fmax(stm, 1.0);
        for (int i = 0; i < m; ++i) {
            c.yo[i] = c.y[i];
            c.lamo[i] = c.lam[i];
            c.muo[i] = c.mu[i];
            c.so[i] = c.sl[i];
        }
        c.zo = c.z;
        c.zeto = c.zet;
        // a non-finite residual would end the line search before any trial
        if (!isfinite(c.rnorm) || !isfinite(c.rmax)) {
            c.status = kCtlNonFinite;
            c.phase = kPhDone;
            break;
        }
        c.itto = 0;
        if (2.0 * c.rnorm > c.rnorm) {   // resinew = 2 rnorm > rnorm
            mma_trial_point(k, c);
        } else {   // rnorm == 0: no trial, keep the point; its Newton sums
                   // are not in this record, so the next step recomputes them
            c.have_sums = 0;
            mma_next_step(k, c, newton_rec);
        }
        break;
    }
    default: {   // trial
        double resinew, rmx;
        mma_m_residual(k, c, red, c.epsi, resinew, rmx);
        c.steg /= 2.0;
        if (resinew > c.rnorm && c.itto < 50) {
            mma_trial_point(k, c);
            break;
        }
        // accept the last trial point (as the reference does, also after 50
        // halvings)
        c.par ^= 1;
        c.rnorm = resinew;
        c.rmax = rmx;
        c.have_sums = 1;
        mma_next_step(k, c, newton_rec);
        break;
    }
    }
    *ctl = c;
}

__global__ void mma_ctl_init_k(MmaCtl *ctl)
{
    if (threadIdx.x == 0) {
        MmaCtl c = {};
        c.phase = kPhPrep;
        *ctl = c;
    }
}

// x_new = the current point's x
__global__ __launch_bounds__(kMmaBlock) void mma_result_k(const MmaCtl *__restrict__ ctl, long n,
                                                         const double *__restrict__ xs,
                                                         const double *__restrict__ xt, double *__restrict__ x_new)
{
    const double *src = ctl->par ? xt : xs;
    for (long j = (long)blockIdx.x * kMmaBlock + threadIdx.x; j < n; j += (long)gridDim.x * kMmaBlock)
        x_new[j] = src[j];
}

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

// units of the previous update (first batch size of the next one)
static std::atomic<int> g_mma_last_units{0};

}  // namespace fantom

using namespace fantom;

static long mma_ctl_offset(long n, int m)
{
    const long rec = 2 * m + m * m + 2 * (1 + m);
    return (long)(11 + 2 * m) * n + (long)kMmaMaxBlocks * (rec + 2);
}

extern "C" long fantom_mma_workspace_size(long n, int m)
{
    if (n <= 0 || m < 1 || m > kMmaMaxM)
        return -1;
    const long ctl_doubles = ((long)sizeof(MmaCtl) + 7) / 8;
    return (long)sizeof(double) * (mma_ctl_offset(n, m) + ctl_doubles + 64);
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
    MmaCtl *ctl = reinterpret_cast<MmaCtl *>(w + mma_ctl_offset(n, m));
    const unsigned nb = mma_blocks(n);
    MmaConst kc = {};
    kc.m = m;
    kc.sub_maxiter = sub_maxiter;
    kc.a0 = a0;
    kc.sub_tol = sub_tol;
    for (int i = 0; i < m; ++i) {
        kc.a[i] = a[i];
        kc.c[i] = c[i];
        kc.d[i] = d[i];
        kc.fval[i] = fval[i];
    }

    hipLaunchKernelGGL(mma_ctl_init_k, dim3(1), dim3(64), 0, s, ctl);
    FANTOM_LAUNCH_CHECK("mma_ctl_init");
    int hdr[kMmaCtlHeaderInts];
    int batch = std::max(8, g_mma_last_units.load() + 2);
    for (;;) {
        for (int u = 0; u < batch; ++u) {
            MMA_DISPATCH(m, mma_pass_k, dim3(nb), dv, ctl, iter, move);
            hipLaunchKernelGGL(mma_ctl_k, dim3(1), dim3(kCtlThreads), 0, s, ctl, dv.partials, nb, kc);
        }
        FANTOM_LAUNCH_CHECK("mma units");
        if (hipMemcpyAsync(hdr, ctl, sizeof(hdr), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            set_error("mma status read: %s", hipGetErrorString(hipGetLastError()));
            return kStatusHipError;
        }
        if (hdr[0] == kPhDone)
            break;
        batch = 16;
    }
    g_mma_last_units.store(hdr[3]);
    if (hdr[1] == kCtlSingular) {
        set_error("MMA Newton system is singular");
        return kStatusHipError;
    }
    if (hdr[1] == kCtlNonFinite) {
        set_error("MMA subproblem residual is not finite");
        return kStatusHipError;
    }
    hipLaunchKernelGGL(mma_result_k, dim3(nb), dim3(kMmaBlock), 0, s, ctl, n, dv.xs, dv.xt, x_new);
    FANTOM_LAUNCH_CHECK("mma_result");
    if (newton_steps)
        *newton_steps = hdr[2];
    return kStatusOk;
}
