// Mixed-precision solve: FP64 iterative refinement around an FP32 GMG-PCG.
//
//   x = x0 (FP64)
//   loop: r = f - K x            FP64 operator, FP64 true residual
//         stop if ||r||/||f|| < tol
//         s = ||r|| ; r32 = r / s
//         K d = r32 approximately: PCG in FP32 storage (vectors, moduli,
//           Walsh-Hadamard element contraction with FP32 accumulation, FP32
//           Jacobi/transfers) with FP64 dot products and step lengths, and the
//           coarsest solve with the FP64 inverse (FP64 accumulation)
//         x += s d
//
// The inner PCG is the FP64 driver's algorithm on FP32 data (see mgpcg.hip):
// scalars on the device, each iteration one HIP graph, device-side stop at
// the inner tolerance with the update frozen afterwards.  The inner solve
// always starts from zero, so one graph serves every refinement step.
//
// FP64 traffic per inner iteration is replaced by FP32 traffic: the fine K.u
// moves 8 B/DOF + 4 B/element instead of 16 B/DOF + 8 B/element, and every
// vector kernel half the bytes.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/fantom.h"
#include "common.hpp"
#include "fantom_internal.hpp"
#include "vec.hpp"

namespace fantom {

enum { X_DONE = 8, X_ITER = 9, X_RES = 10, X_TOL = 11, X_NB = 12, X_K = 13, X_OMEGA = 14, X_XUPD = 15 };

struct MxLevel {
    Grid g;
    long n_dof;
    double scale;
    float *x, *x2, *b, *r, *dinv, *E;
    uint8_t *mask;
    double K0dense[576];
    double *K0dense_d;
    double Kb[72];
    double *Ke;   // explicit Galerkin element matrices (levels >= 2), FP64
    float *S;     // assembled FP32 stencil of an element-matrix level that is smoothed
    int kind;     // 0 fine, 1 rediscretised, 2 Galerkin level 1, 3 element matrices
    double om;    // weight argument of the smoothing kernels (see OmegaSchedule)
    float *ev;    // Chebyshev: power-iteration vector (kept between setups)
};

struct MxHandle {
    int nlev, n_smooth, galerkin;
    double *Gtab, *Btab, *Gdiag, *Ke1;
    double *G2tab;   // 64 x 576 level-2 tables
    float *Btab32;   // the level-1 table in the FP32 operator's precision
    double gal_cls[20];   // class values of Btab (cubic elements), if gal_cls_ok
    bool gal_cls_ok = false;
    double omega;
    OmegaSchedule sched;   // adaptive weight of the FP64-PCG scheme
    std::vector<MxLevel> lev;
    // FP64 fine-level data
    Grid g;
    long n;
    double *E64, *x64, *r64, *f64;
    uint8_t *mask;
    double Kb64[72];
    // FP32 inner PCG vectors
    float *d, *r, *z, *p, *q;
    // FP64 PCG vectors of the FP32-preconditioned solve
    double *p64, *q64;
    double *partials, *scal, *Ainv, *gj_work;
    float *Ainv32;   // FP32 copy of the coarsest inverse, applied by the FP32 V-cycle
    int nc;
    hipStream_t stream;
    hipGraphExec_t iter_exec, pc_exec;
    bool have_graph, have_pc_graph;
    int last_inner, last_pc;
    // fourth-kind Chebyshev smoothing (see fantom_mg_set_smoother)
    int cheb;
    double cw[16];
    double cheb_scale;
    int cheb_iters_first, cheb_iters, cheb_setups;
    double *chs;   // device double[4 nlev]
    // operator of the fine and level-1 sweeps (fantom_mx_set_operator): bit
    // 0 / 1 the fine / level-1 sweeps on the matrix cores (BF16 operands,
    // else FP32 in the WHT form), bits 2-3 the split forms of the operands
    int op_kind;
    uint16_t *Amf1, *Amf8;   // MFMA operand tables: K0, the 8 level-1 child matrices
    float rd32[3];           // 1 / K0 diagonal per component
    // flexible PCG (fantom_mx_set_flexible): Polak-Ribiere beta with the
    // previous preconditioned residual
    int flexible;
    float *zprev;
};

__global__ void mx_copy_f32(long n, const float *__restrict__ a, float *__restrict__ b)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        b[i] = a[i];
}

__global__ void mx_copy_f64(long n, const double *__restrict__ a, double *__restrict__ b)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// r32 = r64 / ||r64|| (scal[S_TMP] holds ||r64||^2)
__global__ void mx_down_k(long n, const double *__restrict__ r64, float *__restrict__ r32,
                          const double *__restrict__ scal)
{
    const double inv = 1.0 / sqrt(scal[S_TMP]);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        r32[i] = (float)(r64[i] * inv);
}

// x64 += ||r64|| * d32
__global__ void mx_up_k(long n, const float *__restrict__ d, double *__restrict__ x, const double *__restrict__ scal)
{
    const double s = sqrt(scal[S_TMP]);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] += s * (double)d[i];
}

__global__ void mx_scale_k(long n, float omega, const double *__restrict__ omega_scale, const float *__restrict__ dinv,
                           const float *__restrict__ b, float *__restrict__ x)
{
    if (omega_scale)
        omega = (float)((double)omega * omega_scale[0]);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = omega * dinv[i] * b[i];
}

__global__ __launch_bounds__(256) void mx_dot_k(long n, const float *__restrict__ a, const float *__restrict__ b,
                                               double *__restrict__ partials)
{
    double s = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        s += (double)a[i] * (double)b[i];
    s = block_sum<256>(s);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void mx_xr_k(long n, float *__restrict__ x, float *__restrict__ r,
                                              const float *__restrict__ p, const float *__restrict__ q,
                                              const double *__restrict__ scal, double *__restrict__ partials)
{
    const bool frozen = scal[X_DONE] != 0.0;
    const double alpha = scal[S_ALPHA];
    double s = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        float ri = r[i];
        if (!frozen) {
            x[i] = (float)((double)x[i] + alpha * (double)p[i]);
            ri = (float)((double)ri - alpha * (double)q[i]);
            r[i] = ri;
        }
        s += (double)ri * (double)ri;
    }
    s = block_sum<256>(s);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = s;
}

__global__ void mx_p_k(long n, float *__restrict__ p, const float *__restrict__ z, const double *__restrict__ scal,
                       int first)
{
    const double beta = first ? 0.0 : scal[S_BETA];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = first ? z[i] : (float)((double)z[i] + beta * (double)p[i]);
}

__global__ void mx_conv_k(const double *__restrict__ partials, long np, double *__restrict__ scal, OmegaSchedule om)
{
    double s = 0.0;
    for (long i = threadIdx.x; i < np; i += 1024)
        s += partials[i];
    s = block_sum<1024>(s);
    if (threadIdx.x == 0) {
        scal[X_K] += 1.0;
        if (scal[X_DONE] == 0.0) {
            // the inner right-hand side has unit norm
            const double R = sqrt(fmax(s, 0.0));
            scal[X_RES] = R;
            scal[X_ITER] = scal[X_K];
            if (R < scal[X_TOL])
                scal[X_DONE] = 1.0;
            else if (om.on)
                scal[X_OMEGA] = schedule_omega(om, R, scal[X_TOL]);
        }
    }
}

__global__ void mx_init_k(double *__restrict__ scal, double tol, double omega0)
{
    if (threadIdx.x == 0) {
        scal[X_OMEGA] = omega0;
        scal[X_DONE] = 0.0;
        scal[X_ITER] = 0.0;
        scal[X_RES] = 0.0;
        scal[X_TOL] = tol;
        scal[X_K] = 0.0;
    }
}

__global__ void mx_zero_f32(long n, float *__restrict__ a)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        a[i] = 0.0f;
}

// ---- FP64 PCG with the FP32 V-cycle as preconditioner ---------------------------
// r -= alpha q (frozen after convergence), r32 = (float) r, and the partial
// sums of r.r; scal[X_XUPD] records whether this iteration updated, for
// mxd_px_k, which applies the matching x += alpha p with the next direction
__global__ __launch_bounds__(256) void mxd_r_k(long n, double *__restrict__ r, float *__restrict__ r32,
                                              const double *__restrict__ q, double *__restrict__ scal,
                                              double *__restrict__ partials)
{
    const bool frozen = scal[X_DONE] != 0.0;
    const double alpha = scal[S_ALPHA];
    double s = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        double ri = r[i];
        if (!frozen) {
            ri -= alpha * q[i];
            r[i] = ri;
        }
        r32[i] = (float)ri;
        s += ri * ri;
    }
    s = block_sum<256>(s);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s;
        if (blockIdx.x == 0)
            scal[X_XUPD] = frozen ? 0.0 : 1.0;
    }
}

// x += alpha p (when mxd_r_k updated r in this iteration), then p = z + beta p;
// zprev (flexible PCG, optional) = z
__global__ __launch_bounds__(256) void mxd_px_k(long n, double *__restrict__ x, double *__restrict__ p,
                                               const float *__restrict__ z, const double *__restrict__ scal,
                                               float *__restrict__ zprev)
{
    const bool upd = scal[X_XUPD] != 0.0;
    const double alpha = scal[S_ALPHA], beta = scal[S_BETA];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double pi = p[i];
        if (upd)
            x[i] += alpha * pi;
        const float zi = z[i];
        p[i] = (double)zi + beta * pi;
        if (zprev)
            zprev[i] = zi;
    }
}

// convergence test on ||r|| / ||f|| (scal[X_NB] = ||f||)
__global__ void mxd_conv_k(const double *__restrict__ partials, long np, double *__restrict__ scal, OmegaSchedule om)
{
    double s = 0.0;
    for (long i = threadIdx.x; i < np; i += 1024)
        s += partials[i];
    s = block_sum<1024>(s);
    if (threadIdx.x == 0) {
        scal[X_K] += 1.0;
        if (scal[X_DONE] == 0.0) {
            const double R = sqrt(fmax(s, 0.0)) / scal[X_NB];
            scal[X_RES] = R;
            scal[X_ITER] = scal[X_K];
            if (R < scal[X_TOL])
                scal[X_DONE] = 1.0;
            else if (om.on)
                scal[X_OMEGA] = schedule_omega(om, R, scal[X_TOL]);
        }
    }
}

// partial sums of r . z with z the FP32 preconditioned residual, and (flexible
// PCG, zprev given) of r . zprev behind them
__global__ __launch_bounds__(256) void mxd_dot_k(long n, const double *__restrict__ r, const float *__restrict__ z,
                                                const float *__restrict__ zprev, double *__restrict__ partials)
{
    double s = 0.0, s2 = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double ri = r[i];
        s += ri * (double)z[i];
        if (zprev)
            s2 += ri * (double)zprev[i];
    }
    s = block_sum<256>(s);
    if (zprev)
        s2 = block_sum<256>(s2);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s;
        if (zprev)
            partials[gridDim.x + blockIdx.x] = s2;
    }
}

// beta of the flexible PCG (Polak-Ribiere): (r.z - r.zprev) / rz_old
__global__ __launch_bounds__(1024) void mxd_beta_pr_k(const double *__restrict__ partials, long np,
                                                     double *__restrict__ scal)
{
    double a = 0.0, c = 0.0;
    for (long i = threadIdx.x; i < np; i += 1024) {
        a += partials[i];
        c += partials[np + i];
    }
    a = block_sum<1024>(a);
    c = block_sum<1024>(c);
    if (threadIdx.x == 0) {
        scal[S_BETA] = (a - c) / scal[S_RZ];
        scal[S_RZ] = a;
    }
}

// p = z + beta p (first: p = z)
__global__ void mxd_p_k(long n, double *__restrict__ p, const float *__restrict__ z, const double *__restrict__ scal,
                        int first)
{
    const double beta = first ? 0.0 : scal[S_BETA];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = first ? (double)z[i] : (double)z[i] + beta * p[i];
}

__global__ void mxd_down_k(long n, const double *__restrict__ r, float *__restrict__ r32)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        r32[i] = (float)r[i];
}

// ||f|| into scal[X_NB] (1 when f = 0)
__global__ void mxd_norm_k(const double *__restrict__ partials, long np, double *__restrict__ scal)
{
    double s = 0.0;
    for (long i = threadIdx.x; i < np; i += 1024)
        s += partials[i];
    s = block_sum<1024>(s);
    if (threadIdx.x == 0) {
        const double nb = sqrt(fmax(s, 0.0));
        scal[X_NB] = nb > 0.0 ? nb : 1.0;
    }
}

#define MX_TRY(expr)                                                                                       \
    do {                                                                                                   \
        int _s = (expr);                                                                                   \
        if (_s)                                                                                            \
            return _s;                                                                                     \
    } while (0)

// om: the weight argument of a smoothing mode (Jacobi's omega or a Chebyshev weight)
static int mx_ku(MxHandle *h, int l, int mode, const float *u, float *y, const float *b, const float *dinv,
                 double *partials, double om = 0.0)
{
    MxLevel &L = h->lev[l];
    if (h->op_kind && mode != MODE_KUDOT && mode != MODE_JACDOT) {
        const int split = (h->op_kind >> 2) & 3;
        if (L.kind == 0 && (h->op_kind & 1))   // the fine sweeps form D^-1 from the moduli (x0 reads it)
            return launch_mf_sweep(mode, 1, L.g, h->Amf1, L.E, L.mask, u, y, b, mode == MODE_JAC ? nullptr : dinv,
                                   om, h->rd32, h->stream, split);
        if (L.kind == 2 && (h->op_kind & 2))
            return launch_mf_sweep(mode, 8, L.g, h->Amf8, h->lev[0].E, L.mask, u, y, b, dinv, om, nullptr,
                                   h->stream, split & 1);
    }
    if (L.kind == 2)
        return launch_gal1_f32(mode, L.g, h->Btab32, h->lev[0].E, L.mask, u, y, b, dinv, om, partials, h->stream,
                               h->gal_cls_ok ? h->gal_cls : nullptr);
    if (L.kind == 3)
        return L.S ? launch_stencil_ku_f32(mode, L.g, L.S, u, y, b, dinv, om, partials, h->stream)
                   : launch_elem_f32(mode, L.g, L.Ke, L.mask, u, y, b, dinv, om, partials, h->stream);
    // the hex8 tile sweeps form D^-1 from the moduli (dinv is the level's
    // own D^-1: the zero-guess sweep reads it once, for x0)
    return launch_ku_tile_f32(mode == MODE_JAC0 ? MODE_JAC0OWN : mode, L.g, L.Kb, L.E, L.mask, u, y, b,
                              mode == MODE_JAC ? nullptr : dinv, om, partials, h->stream);
}

// n_smooth sweeps on x (*px); final (optional) receives the last sweep.
// rz (optional, fine level): the last sweep also leaves the partial sums of
// b.x (the PCG's r.z, b being the FP32 residual) and *rz is set.
// pro (optional): a coarse correction to prolongate onto x first, which the
// first sweep (not the last) does as it loads x
static int mx_smooth(MxHandle *h, int l, const float *b, float **px, bool from_zero, float *final = nullptr,
                     bool *rz = nullptr, const float *pro = nullptr)
{
    MxLevel &L = h->lev[l];
    float *x = *px;
    float *other = (x == L.x) ? L.x2 : L.x;
    // weight argument of sweep s: Jacobi's omega, or the Chebyshev weight
    // (the kernels multiply it by 1 / rho through the level's omega_scale)
    auto w = [&](int s) { return h->cheb ? h->cw[s] : L.om; };
    int s0 = 0;
    // every kind of the FP32 hierarchy fuses the zero-guess scaling except
    // element-matrix levels without a stencil
    if (from_zero && h->n_smooth >= 2 && (L.kind != 3 || L.S != nullptr)) {
        MX_TRY(mx_ku(h, l, MODE_JAC0, nullptr, x, b, L.dinv, nullptr, w(0)));
        s0 = 2;
    }
    for (int s = s0; s < h->n_smooth; ++s) {
        if (s == 0 && from_zero) {
            hipLaunchKernelGGL(mx_scale_k, dim3(vec_grid(L.n_dof)), dim3(256), 0, h->stream, L.n_dof,
                               (float)w(0), L.g.omega_scale, L.dinv, b, x);
            continue;
        }
        const bool last = final && s == h->n_smooth - 1;
        float *dst = last ? final : other;
        if (pro && s == 0) {
            MX_TRY(launch_ku_tile_pro_f32(L.g, L.Kb, L.E, L.mask, x, pro, dst, b, w(0), h->stream));
        } else if (last && rz && l == 0 && !(h->op_kind & 1)) {
            MX_TRY(mx_ku(h, l, MODE_JACDOT, x, dst, b, L.dinv, h->partials, w(s)));
            *rz = true;
        } else {
            MX_TRY(mx_ku(h, l, MODE_JAC, x, dst, b, L.dinv, nullptr, w(s)));
        }
        other = x;
        x = dst;
    }
    *px = x;
    FANTOM_LAUNCH_CHECK("mx smooth");
    return kStatusOk;
}

static int mx_cycle(MxHandle *h, int l, const float *b, float **out, float *final = nullptr, bool *rz = nullptr)
{
    MxLevel &L = h->lev[l];
    if (l == h->nlev - 1) {
        float *dst = final ? final : L.x;
        MX_TRY(launch_gemv_f32(h->nc, h->Ainv32, b, dst, h->stream));
        *out = dst;
        return kStatusOk;
    }
    float *x = L.x;
    MX_TRY(mx_smooth(h, l, b, &x, true));
    MxLevel &C = h->lev[l + 1];
    MX_TRY(mx_ku(h, l, MODE_RES, x, L.r, b, nullptr, nullptr));
    MX_TRY(launch_restrict_f32(L.g, 3, 3, 0, C.mask, L.r, C.b, h->stream));
    float *e = nullptr;
    MX_TRY(mx_cycle(h, l + 1, C.b, &e));
    // on the FP32 tile level the first post-smoothing sweep prolongates e as
    // it loads x (x + P e is never stored); not on the matrix-core sweeps
    const bool fuse = l == 0 && !(h->op_kind & 1) && h->n_smooth >= 2;
    if (!fuse)
        MX_TRY(launch_prolong_f32(L.g, 3, 3, 0, L.mask, e, x, 1, h->stream));
    MX_TRY(mx_smooth(h, l, b, &x, false, final, rz, fuse ? e : nullptr));
    *out = x;
    return kStatusOk;
}

static int mx_precond(MxHandle *h, const float *r, float *z, bool *rz = nullptr)
{
    float *out = nullptr;
    MX_TRY(mx_cycle(h, 0, r, &out, z, rz));
    if (out != z)   // no post-smoothing sweep
        hipLaunchKernelGGL(mx_copy_f32, dim3(vec_grid(h->n)), dim3(256), 0, h->stream, h->n, out, z);
    FANTOM_LAUNCH_CHECK("mx precond");
    return kStatusOk;
}

static int mx_iteration(MxHandle *h)
{
    const long n = h->n;
    const unsigned G = vec_grid(n);
    MxLevel &L0 = h->lev[0];
    MX_TRY(mx_ku(h, 0, MODE_KUDOT, h->p, h->q, nullptr, nullptr, h->partials));
    MX_TRY(launch_finalize(h->partials, ku_tile_partials_f32(L0.g), h->scal, FIN_ALPHA, 0, h->stream));
    hipLaunchKernelGGL(mx_xr_k, dim3(G), dim3(256), 0, h->stream, n, h->d, h->r, h->p, h->q, h->scal, h->partials);
    hipLaunchKernelGGL(mx_conv_k, dim3(1), dim3(1024), 0, h->stream, h->partials, (long)G, h->scal, h->sched);
    MX_TRY(mx_precond(h, h->r, h->z));
    hipLaunchKernelGGL(mx_dot_k, dim3(G), dim3(256), 0, h->stream, n, h->r, h->z, h->partials);
    MX_TRY(launch_finalize(h->partials, (long)G, h->scal, FIN_BETA, 0, h->stream));
    hipLaunchKernelGGL(mx_p_k, dim3(G), dim3(256), 0, h->stream, n, h->p, h->z, h->scal, 0);
    FANTOM_LAUNCH_CHECK("mx iteration");
    return kStatusOk;
}

static int mxd_iteration(MxHandle *h)
{
    const long n = h->n;
    const unsigned G = vec_grid(n);
    hipStream_t s = h->stream;
    MX_TRY(launch_ku_tile(MODE_KUDOT, h->g, h->Kb64, h->E64, h->mask, h->p64, h->q64, nullptr, nullptr, 0.0,
                          h->partials, s));
    MX_TRY(launch_finalize(h->partials, ku_tile_partials(h->g), h->scal, FIN_ALPHA, 0, s));
    hipLaunchKernelGGL(mxd_r_k, dim3(G), dim3(256), 0, s, n, h->r64, h->r, h->q64, h->scal, h->partials);
    hipLaunchKernelGGL(mxd_conv_k, dim3(1), dim3(1024), 0, s, h->partials, (long)G, h->scal, h->sched);
    float *zp = h->flexible ? h->zprev : nullptr;
    // standard PCG: the last fine sweep of the V-cycle forms r.z (with r
    // rounded to FP32, a relative change of about 1e-7 in beta)
    bool rz = false;
    MX_TRY(mx_precond(h, h->r, h->z, zp ? nullptr : &rz));
    if (rz) {
        MX_TRY(launch_finalize(h->partials, ku_tile_partials_f32(h->lev[0].g), h->scal, FIN_BETA, 0, s));
    } else {
        hipLaunchKernelGGL(mxd_dot_k, dim3(G), dim3(256), 0, s, n, h->r64, h->z, zp, h->partials);
        if (zp)
            hipLaunchKernelGGL(mxd_beta_pr_k, dim3(1), dim3(1024), 0, s, h->partials, (long)G, h->scal);
        else
            MX_TRY(launch_finalize(h->partials, (long)G, h->scal, FIN_BETA, 0, s));
    }
    hipLaunchKernelGGL(mxd_px_k, dim3(G), dim3(256), 0, s, n, h->x64, h->p64, h->z, h->scal, zp);
    FANTOM_LAUNCH_CHECK("mxd iteration");
    return kStatusOk;
}

struct MxCarve {
    char *base;
    size_t off;
    template <class T>
    T *take(size_t count)
    {
        off = (off + 255) & ~(size_t)255;
        T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
        off += count * sizeof(T);
        return p;
    }
};

static void mx_layout(MxHandle *h, int nelx, int nely, int nelz, int n_coarse, int galerkin, MxCarve &cv)
{
    h->nlev = n_coarse + 1;
    h->galerkin = galerkin ? 1 : 0;
    h->lev.resize(h->nlev);
    h->g = make_grid(nelx, nely, nelz);
    h->n = h->g.n_nodes * 3;
    for (int l = 0; l < h->nlev; ++l) {
        MxLevel &L = h->lev[l];
        L.g = make_grid(nelx >> l, nely >> l, nelz >> l);
        L.n_dof = L.g.n_nodes * 3;
        L.scale = (double)(1L << l);
        L.x = cv.take<float>(L.n_dof);
        L.x2 = cv.take<float>(L.n_dof);
        L.b = cv.take<float>(L.n_dof);
        L.r = cv.take<float>(L.n_dof);
        L.dinv = cv.take<float>(L.n_dof);
        L.E = cv.take<float>(L.g.n_el);
        L.mask = cv.take<uint8_t>(L.g.n_nodes);
        L.K0dense_d = cv.take<double>(576);
        L.ev = cv.take<float>(L.n_dof);
        L.kind = l == 0 ? 0 : (!h->galerkin ? 1 : (l == 1 ? 2 : 3));
        L.Ke = L.kind == 3 ? cv.take<double>((size_t)L.g.n_el * 576) : nullptr;
        L.S = (L.kind == 3 && l < h->nlev - 1) ? cv.take<float>((size_t)stencil_size(L.g)) : nullptr;
    }
    h->Gtab = h->Btab = h->Gdiag = h->Ke1 = h->G2tab = nullptr;
    h->Btab32 = nullptr;
    if (h->galerkin) {
        h->Gtab = cv.take<double>(8 * 576);
        h->G2tab = cv.take<double>(64 * 576);
        h->Btab = cv.take<double>(576);
        h->Btab32 = cv.take<float>(576);
        h->Gdiag = cv.take<double>(8 * 24);
        if (h->nlev == 2)
            h->Ke1 = cv.take<double>((size_t)h->lev[1].g.n_el * 576);
    }
    h->zprev = cv.take<float>(h->n);
    h->Amf1 = cv.take<uint16_t>(mf_table_size(1));
    h->Amf8 = h->galerkin ? cv.take<uint16_t>(mf_table_size(8)) : nullptr;
    h->E64 = cv.take<double>(h->g.n_el);
    h->x64 = cv.take<double>(h->n);
    h->r64 = cv.take<double>(h->n);
    h->f64 = cv.take<double>(h->n);
    h->mask = h->lev[0].mask;
    h->d = cv.take<float>(h->n);
    h->r = cv.take<float>(h->n);
    h->z = cv.take<float>(h->n);
    h->p = cv.take<float>(h->n);
    h->q = cv.take<float>(h->n);
    h->p64 = cv.take<double>(h->n);
    h->q64 = cv.take<double>(h->n);
    // at least two sums per block for the Chebyshev power pass
    long np = std::max(2 * (long)kVecMaxBlocks, std::max(ku_tile_partials(h->g), ku_tile_partials_f32(h->g)));
    for (int l = 1; l < h->nlev; ++l)
        np = std::max(np, std::max(std::max(elem_partials(h->lev[l].g), gal1_partials(h->lev[l].g)),
                                   stencil_partials(h->lev[l].g)));
    h->partials = cv.take<double>(np);
    h->scal = cv.take<double>(FANTOM_SCALARS);
    h->chs = cv.take<double>(4 * (size_t)h->nlev);
    h->nc = (int)h->lev[h->nlev - 1].n_dof;
    h->Ainv = cv.take<double>((size_t)h->nc * h->nc);
    h->Ainv32 = cv.take<float>((size_t)h->nc * h->nc);
    h->gj_work = cv.take<double>((size_t)h->nc * h->nc + 1024);
}

static int mx_join(hipStream_t from, hipStream_t to)
{
    hipEvent_t ev;
    hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess)
        e = hipEventRecord(ev, from);
    if (e == hipSuccess)
        e = hipStreamWaitEvent(to, ev, 0);
    hipEventDestroy(ev);
    if (e != hipSuccess) {
        set_error("stream join: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    return kStatusOk;
}

static int mx_read(MxHandle *h, double *hs)
{
    hipError_t e = hipMemcpyAsync(hs, h->scal, sizeof(double) * FANTOM_SCALARS, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        set_error("scalar read: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    return kStatusOk;
}

}  // namespace fantom

using namespace fantom;

extern "C" {

long fantom_mx_workspace_size(int nelx, int nely, int nelz, int n_coarse, int galerkin)
{
    if (!grid_ok(nelx, nely, nelz) || nelz <= 0 || n_coarse < 0 || n_coarse > 10)
        return -1;
    const int m = 1 << n_coarse;
    if (nelx % m || nely % m || nelz % m)
        return -1;
    MxHandle h;
    MxCarve cv{nullptr, 0};
    mx_layout(&h, nelx, nely, nelz, n_coarse, galerkin, cv);
    return (long)cv.off + 256;
}

int fantom_mx_create(int nelx, int nely, int nelz, int n_coarse, int n_smooth, double omega, int galerkin,
                     const double *K0_host, void *workspace, long workspace_bytes, void **handle)
{
    const long need = fantom_mx_workspace_size(nelx, nely, nelz, n_coarse, galerkin);
    FANTOM_CHECK_ARG(need > 0, "grid %d x %d x %d (3D only) does not halve %d times", nelx, nely, nelz, n_coarse);
    FANTOM_CHECK_ARG(K0_host && workspace && handle, "null operand");
    FANTOM_CHECK_ARG(workspace_bytes >= need, "workspace of %ld bytes is smaller than the %ld needed",
                     workspace_bytes, need);
    FANTOM_CHECK_ARG(n_smooth >= 1 && n_smooth <= 16, "n_smooth must be in [1, 16]");
    MxHandle *h = new MxHandle();
    MxCarve cv{reinterpret_cast<char *>(workspace), 0};
    cv.off = (256 - (reinterpret_cast<uintptr_t>(workspace) & 255)) & 255;
    mx_layout(h, nelx, nely, nelz, n_coarse, galerkin, cv);
    if (h->nc > 20000) {
        delete h;
        set_error("coarsest grid too large for the dense solve");
        return kStatusBadArg;
    }
    h->n_smooth = n_smooth;
    h->omega = omega;
    h->sched = OmegaSchedule{0, omega, 0.0};
    for (auto &L : h->lev) {
        L.om = omega;
        L.g.omega_scale = nullptr;
    }
    h->have_graph = false;
    h->have_pc_graph = false;
    h->last_inner = 0;
    h->last_pc = 0;
    h->cheb = 0;
    h->cheb_scale = 1.0;
    h->cheb_iters_first = 12;
    h->cheb_iters = 4;
    h->cheb_setups = 0;
    int st = fantom_hex8_blocks(K0_host, h->Kb64, nullptr);
    if (st) {
        delete h;
        return st;
    }
    for (int l = 0; l < h->nlev; ++l) {
        MxLevel &L = h->lev[l];
        for (int i = 0; i < 576; ++i)
            L.K0dense[i] = K0_host[i] * L.scale;
        for (int i = 0; i < 72; ++i)
            L.Kb[i] = h->Kb64[i] * L.scale;
    }
    std::vector<double> tG(8 * 576), tB(576), tD(8 * 24);
    if (h->galerkin) {
        st = fantom_galerkin_tables(K0_host, tG.data(), tB.data(), tD.data(), nullptr);
        if (st) {
            delete h;
            return st;
        }
    }
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    for (int l = 0; l < h->nlev && e == hipSuccess; ++l)
        e = hipMemcpyAsync(h->lev[l].K0dense_d, h->lev[l].K0dense, sizeof(double) * 576, hipMemcpyHostToDevice,
                           h->stream);
    if (h->galerkin && e == hipSuccess)
        e = hipMemcpyAsync(h->Gtab, tG.data(), sizeof(double) * tG.size(), hipMemcpyHostToDevice, h->stream);
    std::vector<double> tG2(64 * 576);
    if (h->galerkin)
        galerkin_tables2(tG.data(), tG2.data());
    if (h->galerkin && e == hipSuccess)
        e = hipMemcpyAsync(h->G2tab, tG2.data(), sizeof(double) * tG2.size(), hipMemcpyHostToDevice, h->stream);
    if (h->galerkin && e == hipSuccess)
        e = hipMemcpyAsync(h->Btab, tB.data(), sizeof(double) * tB.size(), hipMemcpyHostToDevice, h->stream);
    h->gal_cls_ok = h->galerkin && gal1_class_values(tB.data(), h->gal_cls);
    const std::vector<float> tB32(tB.begin(), tB.end());
    if (h->galerkin && e == hipSuccess)
        e = hipMemcpyAsync(h->Btab32, tB32.data(), sizeof(float) * tB32.size(), hipMemcpyHostToDevice, h->stream);
    if (h->galerkin && e == hipSuccess)
        e = hipMemcpyAsync(h->Gdiag, tD.data(), sizeof(double) * tD.size(), hipMemcpyHostToDevice, h->stream);
    // matrix-core operand tables: K0 in natural corner order (b = bx + 2 by +
    // 4 bz), the level-1 child matrices G_ch (already natural)
    static const int loc[8] = {0, 1, 3, 2, 4, 5, 7, 6};   // natural corner -> reference local corner
    std::vector<double> Kn(576);
    for (int b = 0; b < 8; ++b)
        for (int c = 0; c < 3; ++c)
            for (int b2 = 0; b2 < 8; ++b2)
                for (int d = 0; d < 3; ++d)
                    Kn[(b * 3 + c) * 24 + b2 * 3 + d] = K0_host[(loc[b] * 3 + c) * 24 + loc[b2] * 3 + d];
    for (int c = 0; c < 3; ++c)
        h->rd32[c] = (float)(1.0 / Kn[c * 24 + c]);
    h->op_kind = 0;
    h->flexible = 0;
    std::vector<uint16_t> tA1(mf_table_size(1)), tA8(h->galerkin ? mf_table_size(8) : 0);
    mf_tables(Kn.data(), 1, tA1.data());
    if (h->galerkin)
        mf_tables(tG.data(), 8, tA8.data());
    if (e == hipSuccess)
        e = hipMemcpyAsync(h->Amf1, tA1.data(), sizeof(uint16_t) * tA1.size(), hipMemcpyHostToDevice, h->stream);
    if (h->galerkin && e == hipSuccess)
        e = hipMemcpyAsync(h->Amf8, tA8.data(), sizeof(uint16_t) * tA8.size(), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        delete h;
        set_error("mixed create: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    *handle = h;
    return kStatusOk;
}

int fantom_mx_set_operator(void *handle, int kind)
{
    MxHandle *h = reinterpret_cast<MxHandle *>(handle);
    FANTOM_CHECK_ARG(h, "null handle");
    FANTOM_CHECK_ARG(kind >= 0 && kind <= 15, "operator kind must be in [0, 15]");
    h->op_kind = kind;
    if (h->have_pc_graph) {   // the captured iterations hold the old kernels
        hipGraphExecDestroy(h->pc_exec);
        h->have_pc_graph = false;
    }
    if (h->have_graph) {
        hipGraphExecDestroy(h->iter_exec);
        h->have_graph = false;
    }
    return kStatusOk;
}

int fantom_mx_set_flexible(void *handle, int on)
{
    MxHandle *h = reinterpret_cast<MxHandle *>(handle);
    FANTOM_CHECK_ARG(h, "null handle");
    h->flexible = on ? 1 : 0;
    if (h->have_pc_graph) {   // the captured iterations hold the old beta
        hipGraphExecDestroy(h->pc_exec);
        h->have_pc_graph = false;
    }
    return kStatusOk;
}

int fantom_mx_set_omega_schedule(void *handle, int adaptive, double min_omega, double level_boost)
{
    MxHandle *h = reinterpret_cast<MxHandle *>(handle);
    FANTOM_CHECK_ARG(h, "null handle");
    FANTOM_CHECK_ARG(!adaptive || (min_omega > 0.0 && level_boost > 0.0),
                     "adaptive schedule needs min_omega > 0 and level_boost > 0");
    h->sched = adaptive ? OmegaSchedule{1, min_omega, h->omega - min_omega} : OmegaSchedule{0, h->omega, 0.0};
    h->cheb = 0;   // the weight schedule replaces a Chebyshev smoother
    double boost = 1.0;
    for (auto &L : h->lev) {
        L.om = adaptive ? boost : h->omega;
        L.g.omega_scale = adaptive ? h->scal + X_OMEGA : nullptr;
        L.g.w2_ratio = 1.0;
        boost *= level_boost;
    }
    if (adaptive) {   // the weight of a V-cycle applied before any solve
        const hipError_t e = hipMemcpy(h->scal + X_OMEGA, &min_omega, sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            set_error("omega schedule: %s", hipGetErrorString(e));
            return kStatusHipError;
        }
    }
    if (h->have_pc_graph) {   // the captured iterations hold the old weights
        hipGraphExecDestroy(h->pc_exec);
        h->have_pc_graph = false;
    }
    if (h->have_graph) {
        hipGraphExecDestroy(h->iter_exec);
        h->have_graph = false;
    }
    return kStatusOk;
}

int fantom_mx_set_smoother(void *handle, int kind, double rho_scale, int iters_first, int iters)
{
    MxHandle *h = reinterpret_cast<MxHandle *>(handle);
    FANTOM_CHECK_ARG(h, "null handle");
    FANTOM_CHECK_ARG(kind == 0 || kind == 1, "smoother kind must be 0 (Jacobi) or 1 (Chebyshev)");
    FANTOM_CHECK_ARG(kind == 0 || (rho_scale >= 0.5 && rho_scale <= 4.0 && iters_first >= 1 &&
                                   iters_first <= 100 && iters >= 0 && iters <= 100),
                     "Chebyshev needs rho_scale in [0.5, 4], iters_first in [1, 100], iters in [0, 100]");
    h->cheb = kind;
    h->sched = OmegaSchedule{0, h->omega, 0.0};
    if (kind == 1) {
        chebyshev4_weights(h->n_smooth, h->cw);
        h->cheb_scale = rho_scale;
        h->cheb_iters_first = iters_first;
        h->cheb_iters = iters;
        h->cheb_setups = 0;
    }
    for (int l = 0; l < h->nlev; ++l) {
        MxLevel &L = h->lev[l];
        L.om = h->omega;
        L.g.omega_scale = kind == 1 ? h->chs + 4 * l + 2 : nullptr;
        L.g.w2_ratio = (kind == 1 && h->n_smooth >= 2) ? h->cw[1] / h->cw[0] : 1.0;
    }
    if (h->have_pc_graph) {   // the captured iterations hold the old weights
        hipGraphExecDestroy(h->pc_exec);
        h->have_pc_graph = false;
    }
    if (h->have_graph) {
        hipGraphExecDestroy(h->iter_exec);
        h->have_graph = false;
    }
    return kStatusOk;
}

int fantom_mx_smoother_rho(void *handle, double *rho, int n)
{
    MxHandle *h = reinterpret_cast<MxHandle *>(handle);
    FANTOM_CHECK_ARG(h && rho && n >= h->nlev - 1, "null operand or n below the number of smoothed levels");
    std::vector<double> st(4 * (size_t)h->nlev);
    hipError_t e = hipMemcpyAsync(st.data(), h->chs, sizeof(double) * st.size(), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        set_error("smoother rho read: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    for (int l = 0; l < h->nlev - 1; ++l)
        rho[l] = st[4 * l + 1];
    return kStatusOk;
}

int fantom_mx_destroy(void *handle)
{
    MxHandle *h = reinterpret_cast<MxHandle *>(handle);
    if (!h)
        return kStatusOk;
    hipStreamSynchronize(h->stream);
    if (h->have_graph)
        hipGraphExecDestroy(h->iter_exec);
    if (h->have_pc_graph)
        hipGraphExecDestroy(h->pc_exec);
    hipStreamDestroy(h->stream);
    delete h;
    return kStatusOk;
}

__global__ void mx_f2d_k(long n, const float *__restrict__ a, double *__restrict__ b)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        b[i] = (double)a[i];
}

int fantom_mx_precondition(void *handle, const double *r, double *z, void *stream)
{
    MxHandle *h = reinterpret_cast<MxHandle *>(handle);
    FANTOM_CHECK_ARG(h && r && z, "null operand");
    MX_TRY(mx_join(as_stream(stream), h->stream));
    const long n = h->n;
    const unsigned G = vec_grid(n);
    hipLaunchKernelGGL(mxd_down_k, dim3(G), dim3(256), 0, h->stream, n, r, h->r);
    MX_TRY(mx_precond(h, h->r, h->z));
    hipLaunchKernelGGL(mx_f2d_k, dim3(G), dim3(256), 0, h->stream, n, h->z, z);
    FANTOM_LAUNCH_CHECK("mx precondition");
    return mx_join(h->stream, as_stream(stream));
}

__global__ void mx_e_down_k(long n, const double *__restrict__ a, float *__restrict__ b)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

int fantom_mx_setup(void *handle, const double *E, const uint8_t *mask, void *stream)
{
    MxHandle *h = reinterpret_cast<MxHandle *>(handle);
    FANTOM_CHECK_ARG(h && E, "null operand");
    MX_TRY(mx_join(as_stream(stream), h->stream));
    hipStream_t s = h->stream;
    const long ne = h->g.n_el;
    hipLaunchKernelGGL(mx_copy_f64, dim3(vec_grid(ne)), dim3(256), 0, s, ne, E, h->E64);
    hipLaunchKernelGGL(mx_e_down_k, dim3(vec_grid(ne)), dim3(256), 0, s, ne, E, h->lev[0].E);
    hipError_t e = mask ? hipMemcpyAsync(h->mask, mask, h->g.n_nodes, hipMemcpyDeviceToDevice, s)
                        : hipMemsetAsync(h->mask, 0, h->g.n_nodes, s);
    if (e != hipSuccess) {
        set_error("mixed setup copy: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    for (int l = 1; l < h->nlev; ++l) {
        MxLevel &F = h->lev[l - 1], &C = h->lev[l];
        if (!h->galerkin)
            MX_TRY(launch_coarsen_E_f32(F.g, 3, F.E, C.E, s));
        MX_TRY(launch_coarsen_mask(F.g, 3, F.mask, C.mask, s));
    }
    if (h->galerkin) {
        for (int l = 2; l < h->nlev; ++l) {
            MxLevel &C = h->lev[l];
            if (l == 2)
                MX_TRY(launch_gal2_build(C.g, h->G2tab, h->E64, C.Ke, s));
            else
                MX_TRY(launch_gal_build(C.g, 1, h->Gtab, nullptr, h->lev[l - 1].Ke, C.Ke, s));
        }
        if (h->nlev == 2)
            MX_TRY(launch_elem_from_fine(h->lev[1].g, h->Gtab, h->E64, h->Ke1, s));
    }
    for (int l = 0; l < h->nlev; ++l) {
        MxLevel &L = h->lev[l];
        if (L.kind == 2)
            MX_TRY(launch_dinv_gal1_f32(L.g, h->Gdiag, h->lev[0].E, L.mask, L.dinv, s));
        else if (L.kind == 3 && L.S) {
            MX_TRY(launch_stencil_build_f32(L.g, L.Ke, L.mask, L.S, s));
            MX_TRY(launch_dinv_stencil_f32(L.g, L.S, L.dinv, s));
        } else if (L.kind == 3)
            MX_TRY(launch_dinv_elem_f32(L.g, L.Ke, L.mask, L.dinv, s));
        else
            MX_TRY(launch_diag_f32(L.g, L.K0dense_d, L.E, L.mask, L.dinv, s));
    }
    if (h->cheb) {   // rho of D^-1 A per smoothed level (see fantom_mg_setup)
        const bool first = h->cheb_setups == 0;
        const int iters = first ? h->cheb_iters_first : h->cheb_iters;
        for (int l = 0; l < h->nlev - 1 && iters > 0; ++l) {
            MxLevel &L = h->lev[l];
            double *st = h->chs + 4 * l;
            MX_TRY(launch_ch_init_f32(L.n_dof, L.ev, L.dinv, first ? 1 : 0, l, h->partials, st, s));
            for (int it = 0; it < iters; ++it) {
                MX_TRY(mx_ku(h, l, MODE_KU, L.ev, L.x2, nullptr, nullptr, nullptr));
                MX_TRY(launch_ch_pass_f32(L.n_dof, L.ev, L.x2, L.dinv, h->cheb_scale, h->partials, st, s));
            }
        }
        ++h->cheb_setups;
    }
    // the coarsest operator and its inverse are formed in FP64; the V-cycle
    // applies an FP32 copy of the inverse (half the bytes of its one read per
    // cycle), with the products summed in FP64
    MxLevel &C = h->lev[h->nlev - 1];
    if (C.kind == 3)
        MX_TRY(launch_dense_from_elem(C.g, C.Ke, C.mask, h->Ainv, s));
    else if (C.kind == 2)
        MX_TRY(launch_dense_from_elem(C.g, h->Ke1, C.mask, h->Ainv, s));
    else
        MX_TRY(launch_dense_assemble_f32E(C.g, C.K0dense_d, C.E, C.mask, h->Ainv, s));
    MX_TRY(launch_spd_inverse(h->nc, h->Ainv, h->gj_work, s));
    MX_TRY(launch_to_f32((long)h->nc * h->nc, h->Ainv, h->Ainv32, s));
    FANTOM_LAUNCH_CHECK("mixed setup");
    return mx_join(s, as_stream(stream));
}

static int mx_capture(MxHandle *h, bool fp64_pcg = false)
{
    hipGraph_t graph;
    hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        set_error("begin capture: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    int st = fp64_pcg ? mxd_iteration(h) : mx_iteration(h);
    e = hipStreamEndCapture(h->stream, &graph);
    if (st)
        return st;
    if (e != hipSuccess) {
        set_error("end capture: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    e = hipGraphInstantiate(fp64_pcg ? &h->pc_exec : &h->iter_exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) {
        set_error("graph instantiate: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    (fp64_pcg ? h->have_pc_graph : h->have_graph) = true;
    return kStatusOk;
}

// inner FP32 PCG on K d = r32 from d = 0; returns the iteration count
static int mx_inner(MxHandle *h, double inner_tol, int maxiter, int use_graph, int *iters)
{
    const long n = h->n;
    const unsigned G = vec_grid(n);
    hipStream_t s = h->stream;
    hipLaunchKernelGGL(mx_zero_f32, dim3(G), dim3(256), 0, s, n, h->d);
    hipLaunchKernelGGL(mx_init_k, dim3(1), dim3(64), 0, s, h->scal, inner_tol, h->sched.min);
    MX_TRY(mx_precond(h, h->r, h->z));
    hipLaunchKernelGGL(mx_dot_k, dim3(G), dim3(256), 0, s, n, h->r, h->z, h->partials);
    MX_TRY(launch_finalize(h->partials, (long)G, h->scal, FIN_RZ0, 0, s));
    hipLaunchKernelGGL(mx_p_k, dim3(G), dim3(256), 0, s, n, h->p, h->z, h->scal, 1);
    FANTOM_LAUNCH_CHECK("mx inner init");
    if (use_graph && !h->have_graph)
        MX_TRY(mx_capture(h));
    double hs[FANTOM_SCALARS];
    int done = 0, batch = h->last_inner > 3 ? h->last_inner - 2 : 1;
    while (done < maxiter) {
        batch = std::min(batch, maxiter - done);
        for (int b = 0; b < batch; ++b) {
            if (use_graph) {
                hipError_t e = hipGraphLaunch(h->iter_exec, s);
                if (e != hipSuccess) {
                    set_error("graph launch: %s", hipGetErrorString(e));
                    return kStatusHipError;
                }
            } else {
                MX_TRY(mx_iteration(h));
            }
        }
        done += batch;
        MX_TRY(mx_read(h, hs));
        if (hs[X_DONE] != 0.0) {
            *iters = (int)hs[X_ITER];
            h->last_inner = *iters;
            return kStatusOk;
        }
        batch = 1;
    }
    *iters = done;
    h->last_inner = done;
    return kStatusOk;
}

int fantom_mx_solve(void *handle, const double *f, double *x, double tol, double inner_tol, int max_outer,
                    int max_inner, int use_x0, int use_graph, int *outer_iters, int *inner_iters, double *rel_res,
                    void *stream)
{
    MxHandle *h = reinterpret_cast<MxHandle *>(handle);
    FANTOM_CHECK_ARG(h && f && x && outer_iters && inner_iters && rel_res, "null operand");
    FANTOM_CHECK_ARG(max_outer >= 1 && max_inner >= 1, "iteration limits must be positive");
    MX_TRY(mx_join(as_stream(stream), h->stream));
    hipStream_t s = h->stream;
    const long n = h->n;
    const unsigned G = vec_grid(n);
    hipLaunchKernelGGL(mx_copy_f64, dim3(G), dim3(256), 0, s, n, f, h->f64);
    if (use_x0)
        hipLaunchKernelGGL(mx_copy_f64, dim3(G), dim3(256), 0, s, n, x, h->x64);
    else
        hipMemsetAsync(h->x64, 0, sizeof(double) * n, s);
    MX_TRY(launch_dot_partials(n, h->f64, h->f64, h->partials, s));
    MX_TRY(launch_finalize(h->partials, (long)G, h->scal, FIN_STORE, 7, s));
    double hs[FANTOM_SCALARS];
    MX_TRY(mx_read(h, hs));
    double nf = sqrt(fmax(hs[7], 0.0));
    if (nf == 0.0)
        nf = 1.0;
    int total_inner = 0, outer = 0;
    double R = 0.0;
    for (outer = 0; outer <= max_outer; ++outer) {
        // FP64 true residual
        MX_TRY(launch_ku_tile(MODE_RES, h->g, h->Kb64, h->E64, h->mask, h->x64, h->r64, h->f64, nullptr, 0.0,
                              nullptr, s));
        MX_TRY(launch_dot_partials(n, h->r64, h->r64, h->partials, s));
        MX_TRY(launch_finalize(h->partials, (long)G, h->scal, FIN_STORE, S_TMP, s));
        MX_TRY(mx_read(h, hs));
        R = sqrt(fmax(hs[S_TMP], 0.0)) / nf;
        if (R < tol || outer == max_outer || hs[S_TMP] == 0.0)
            break;
        hipLaunchKernelGGL(mx_down_k, dim3(G), dim3(256), 0, s, n, h->r64, h->r, h->scal);
        int it = 0;
        MX_TRY(mx_inner(h, inner_tol, max_inner, use_graph, &it));
        total_inner += it;
        hipLaunchKernelGGL(mx_up_k, dim3(G), dim3(256), 0, s, n, h->d, h->x64, h->scal);
        FANTOM_LAUNCH_CHECK("mx refine");
    }
    hipLaunchKernelGGL(mx_copy_f64, dim3(G), dim3(256), 0, s, n, h->x64, x);
    FANTOM_LAUNCH_CHECK("mx solve");
    *outer_iters = outer;
    *inner_iters = total_inner;
    *rel_res = R;
    return mx_join(s, as_stream(stream));
}

// FP64 PCG preconditioned by one FP32 V-cycle per iteration: FP64 fine K.u,
// FP64 vectors, dots and step lengths; the V-cycle (smoothing, residuals,
// transfers and coarse operators) runs on FP32 copies.  Stops on the
// recursive FP64 residual; returns the true relative residual.
int fantom_mx_solve_pcg(void *handle, const double *f, double *x, double tol, int maxiter, int use_x0, int use_graph,
                        int *iters, double *rel_res, void *stream)
{
    MxHandle *h = reinterpret_cast<MxHandle *>(handle);
    FANTOM_CHECK_ARG(h && f && x && iters && rel_res, "null operand");
    FANTOM_CHECK_ARG(maxiter >= 1, "maxiter must be positive");
    MX_TRY(mx_join(as_stream(stream), h->stream));
    hipStream_t s = h->stream;
    const long n = h->n;
    const unsigned G = vec_grid(n);
    hipLaunchKernelGGL(mx_copy_f64, dim3(G), dim3(256), 0, s, n, f, h->f64);
    if (use_x0)
        hipLaunchKernelGGL(mx_copy_f64, dim3(G), dim3(256), 0, s, n, x, h->x64);
    else
        hipMemsetAsync(h->x64, 0, sizeof(double) * n, s);
    MX_TRY(launch_dot_partials(n, h->f64, h->f64, h->partials, s));
    hipLaunchKernelGGL(mxd_norm_k, dim3(1), dim3(1024), 0, s, h->partials, (long)G, h->scal);
    hipLaunchKernelGGL(mx_init_k, dim3(1), dim3(64), 0, s, h->scal, tol, h->sched.min);
    MX_TRY(launch_ku_tile(MODE_RES, h->g, h->Kb64, h->E64, h->mask, h->x64, h->r64, h->f64, nullptr, 0.0, nullptr,
                          s));
    MX_TRY(launch_dot_partials(n, h->r64, h->r64, h->partials, s));
    MX_TRY(launch_finalize(h->partials, (long)G, h->scal, FIN_STORE, S_TMP, s));
    double hs[FANTOM_SCALARS];
    MX_TRY(mx_read(h, hs));
    int done = 0;
    if (sqrt(fmax(hs[S_TMP], 0.0)) / hs[X_NB] >= tol) {
        hipLaunchKernelGGL(mxd_down_k, dim3(G), dim3(256), 0, s, n, h->r64, h->r);
        MX_TRY(mx_precond(h, h->r, h->z));
        hipLaunchKernelGGL(mxd_dot_k, dim3(G), dim3(256), 0, s, n, h->r64, h->z, nullptr, h->partials);
        MX_TRY(launch_finalize(h->partials, (long)G, h->scal, FIN_RZ0, 0, s));
        hipLaunchKernelGGL(mxd_p_k, dim3(G), dim3(256), 0, s, n, h->p64, h->z, h->scal, 1);
        if (h->flexible)
            hipLaunchKernelGGL(mx_copy_f32, dim3(G), dim3(256), 0, s, n, h->z, h->zprev);
        FANTOM_LAUNCH_CHECK("mxd init");
        if (use_graph && !h->have_pc_graph)
            MX_TRY(mx_capture(h, true));
        int batch = h->last_pc > 3 ? h->last_pc - 2 : 1;
        bool conv = false;
        // stagnation: the recursive residual has not dropped by 10% over
        // kStallWindow iterations -> stop (the caller sees rel_res >= tol)
        constexpr int kStallWindow = 20;
        int w0 = 0;
        double r0 = INFINITY;
        while (done < maxiter && !conv) {
            batch = std::min(batch, maxiter - done);
            for (int b = 0; b < batch; ++b) {
                if (use_graph) {
                    const hipError_t e = hipGraphLaunch(h->pc_exec, s);
                    if (e != hipSuccess) {
                        set_error("graph launch: %s", hipGetErrorString(e));
                        return kStatusHipError;
                    }
                } else {
                    MX_TRY(mxd_iteration(h));
                }
            }
            done += batch;
            MX_TRY(mx_read(h, hs));
            if (hs[X_DONE] != 0.0) {
                done = (int)hs[X_ITER];
                conv = true;
            } else if (done - w0 >= kStallWindow) {
                if (!(hs[X_RES] < 0.9 * r0) && r0 != INFINITY)
                    break;
                w0 = done;
                r0 = hs[X_RES];
            } else if (r0 == INFINITY) {
                w0 = done;
                r0 = hs[X_RES];
            }
            batch = 1;
        }
        h->last_pc = done;
    }
    // true FP64 residual of the result
    MX_TRY(launch_ku_tile(MODE_RES, h->g, h->Kb64, h->E64, h->mask, h->x64, h->r64, h->f64, nullptr, 0.0, nullptr,
                          s));
    MX_TRY(launch_dot_partials(n, h->r64, h->r64, h->partials, s));
    MX_TRY(launch_finalize(h->partials, (long)G, h->scal, FIN_STORE, S_TMP, s));
    MX_TRY(mx_read(h, hs));
    hipLaunchKernelGGL(mx_copy_f64, dim3(G), dim3(256), 0, s, n, h->x64, x);
    FANTOM_LAUNCH_CHECK("mxd solve");
    *iters = done;
    *rel_res = sqrt(fmax(hs[S_TMP], 0.0)) / hs[X_NB];
    return mx_join(s, as_stream(stream));
}

}  // extern "C"
