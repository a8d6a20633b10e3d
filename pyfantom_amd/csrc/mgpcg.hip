// Native GMG-PCG driver (FP64): the whole solve runs from C++ with all PCG
// scalars on the device and each PCG iteration captured once into a HIP
// graph, so an iteration costs one graph launch instead of ~40 kernel
// launches from Python.
//
// Algorithm (reference: solvers/CUDA/_solvers.py:518-660, with the
// deliberate differences stated in pyfantom_amd/solvers.py):
//   r = f - K x ; z = V(r) ; p = z ; rz = r.z
//   repeat: q = K p (fused p.q) ; alpha = rz / p.q ; x += alpha p ;
//           r -= alpha q (fused r.r) ; stop when sqrt(r.r)/||f|| < tol ;
//           z = V(r) ; beta = r.z_new / rz ; p = z + beta p
// V(b) on level l: n_smooth Jacobi sweeps from zero, r_l = b - A_l x,
// restrict, recurse, prolongate-and-add, n_smooth sweeps; the coarsest level
// is an exact dense solve with A_c^-1 formed at setup.
//
// Convergence is detected on the device: the r.r finalize sets a done flag
// and records the iteration; the x/r update kernel is a no-op once the flag
// is set, so iterations launched past convergence leave x unchanged and the
// result equals stopping at the first converged iteration.  The host reads
// the flag after batches of iterations sized from the previous solve.
//
// Memory: the caller provides one device workspace of
// fantom_mg_workspace_size() bytes; the handle is host memory.  Nothing is
// allocated on the device by the library.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "../../include/fantom.h"
#include "common.hpp"
#include "fantom_internal.hpp"
#include "vec.hpp"

namespace fantom {

enum { D_DONE = 8, D_ITER = 9, D_RES = 10, D_TOL = 11, D_NB = 12, D_K = 13, D_OMEGA = 14, D_XUPD = 15 };


struct MgLevel {
    Grid g;
    int dim, dof;
    long n_dof;
    double scale;
    double *x, *x2, *b, *r, *dinv, *E;
    uint8_t *mask;
    double K0dense[576];   // host copy of the scaled dense element matrix
    double *K0dense_d;     // device copy (diag / dense assembly / node kernel)
    double Kb[72];         // host block form (tile kernel)
    double *Ke;            // explicit Galerkin element matrices (levels >= 2)
    double *S;             // assembled stencil of an element-matrix level that is smoothed
    int kind;              // LK_* operator of this level
    double om;             // weight argument of the smoothing kernels (see OmegaSchedule)
    double *ev;            // Chebyshev: power-iteration vector (kept between setups)
};

// operator kind of a level
enum { LK_FINE = 0, LK_REDISC = 1, LK_GAL1 = 2, LK_ELEM = 3 };

struct MgHandle {
    int dim, dof, nlev, n_smooth, kernel, galerkin;
    double *Gtab, *Btab, *Gdiag;     // device Galerkin tables (hex8)
    double gal_cls[20];              // class values of Btab (cubic elements), if gal_cls_ok
    bool gal_cls_ok = false;
    double *G2tab;                   // 64 x 576 level-2 tables (levels >= 2 built from the fine moduli)
    double *Ke1;                     // level-1 element matrices when level 1 is the coarsest
    double omega;
    OmegaSchedule sched;
    std::vector<MgLevel> lev;
    double *xs, *r, *z, *p, *q, *f;   // fine-level PCG vectors
    double *partials;
    long n_partials;
    double *scal;                      // FANTOM_SCALARS doubles of PCG state + flags
    double *Ainv, *gj_work;
    int nc;
    hipStream_t stream;
    hipGraphExec_t iter_exec;
    bool have_graph;
    int last_iters;
    // fourth-kind Chebyshev smoothing (off: weighted Jacobi / the schedule)
    int cheb;
    double cw[16];         // sweep weights in units of 1 / rho, ascending
    double cheb_scale;     // rho is multiplied by this before use
    int cheb_iters_first, cheb_iters;
    int cheb_setups;
    double *chs;           // device double[4 nlev]: per level {|v|_D^2, rho, 1 / (scale rho), -}
};

// ---- small device kernels specific to the driver --------------------------------
__global__ void mg_conv_k(const double *__restrict__ partials, long np, double *__restrict__ scal, OmegaSchedule om)
{
    // r.r finalize + convergence test on the device
    double s = 0.0;
    for (long i = threadIdx.x; i < np; i += 1024)
        s += partials[i];
    s = block_sum<1024>(s);
    if (threadIdx.x == 0) {
        scal[D_K] += 1.0;
        if (scal[D_DONE] == 0.0) {
            scal[S_RR] = s;
            const double R = sqrt(fmax(s, 0.0)) / scal[D_NB];
            scal[D_RES] = R;
            scal[D_ITER] = scal[D_K];
            if (R < scal[D_TOL])
                scal[D_DONE] = 1.0;
            else if (om.on)
                scal[D_OMEGA] = schedule_omega(om, R, scal[D_TOL]);
        }
    }
}

// r -= alpha q (frozen after convergence) and the partial sums of r.r;
// scal[D_XUPD] records whether this iteration updated, for mg_px_k, which
// applies the matching x += alpha p together with the next direction
__global__ void mg_r_k(long n, double *__restrict__ r, const double *__restrict__ q, double *__restrict__ scal,
                       double *__restrict__ partials)
{
    const bool frozen = scal[D_DONE] != 0.0;
    const double alpha = scal[S_ALPHA];
    double s = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double ri = r[i];
        if (!frozen) {
            ri -= alpha * q[i];
            r[i] = ri;
        }
        s += ri * ri;
    }
    s = block_sum<256>(s);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s;
        if (blockIdx.x == 0)
            scal[D_XUPD] = frozen ? 0.0 : 1.0;
    }
}

// x += alpha p (when mg_r_k updated r in this iteration), then p = z + beta p:
// the iterate update rides on the direction update's pass over p
__global__ __launch_bounds__(256) void mg_px_k(long n, double *__restrict__ x, double *__restrict__ p,
                                               const double *__restrict__ z, const double *__restrict__ scal)
{
    const bool upd = scal[D_XUPD] != 0.0;
    const double alpha = scal[S_ALPHA], beta = scal[S_BETA];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double pi = p[i];
        if (upd)
            x[i] += alpha * pi;
        p[i] = z[i] + beta * pi;
    }
}

__global__ void mg_init_scal_k(double *__restrict__ scal, double tol, double omega0)
{
    if (threadIdx.x == 0) {
        scal[D_OMEGA] = omega0;
        scal[D_DONE] = 0.0;
        scal[D_ITER] = 0.0;
        scal[D_RES] = 0.0;
        scal[D_TOL] = tol;
        scal[D_K] = 0.0;
    }
}

__global__ void mg_norm_k(const double *__restrict__ partials, long np, double *__restrict__ scal)
{
    double s = 0.0;
    for (long i = threadIdx.x; i < np; i += 1024)
        s += partials[i];
    s = block_sum<1024>(s);
    if (threadIdx.x == 0) {
        const double nb = sqrt(fmax(s, 0.0));
        scal[D_NB] = nb > 0.0 ? nb : 1.0;
    }
}

__global__ void mg_copy_k(long n, const double *__restrict__ a, double *__restrict__ b)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// ---- helpers ----------------------------------------------------------------------
static int lev_ku(MgHandle *h, int l, int mode, const double *u, double *y, const double *b, const double *dinv,
                  double omega, double *partials)
{
    MgLevel &L = h->lev[l];
    if (L.kind == LK_GAL1)
        return launch_gal1(mode, L.g, h->Btab, h->lev[0].E, L.mask, u, y, b, dinv, omega, partials, h->stream,
                           h->gal_cls_ok ? h->gal_cls : nullptr);
    if (L.kind == LK_ELEM)
        return L.S ? launch_stencil_ku(mode, L.g, L.S, u, y, b, dinv, omega, partials, h->stream)
                   : launch_elem(mode, L.g, L.Ke, L.mask, u, y, b, dinv, omega, partials, h->stream);
    if (h->kernel == FANTOM_KU_TILE && L.dim == 3) {
        // the hex8 tile sweeps form D^-1 from the moduli (dinv is the level's
        // own D^-1: the zero-guess sweep reads it once, for x0)
        const double *d = (mode == MODE_JAC || mode == MODE_JACDOT) ? nullptr : dinv;
        return launch_ku_tile(mode == MODE_JAC0 ? MODE_JAC0OWN : mode, L.g, L.Kb, L.E, L.mask, u, y, b, d, omega,
                              partials, h->stream);
    }
    return launch_ku_node(mode, L.g, L.dim, L.K0dense_d, L.E, L.mask, u, y, b, dinv, omega, partials, h->stream);
}

#define MG_TRY(expr)                                                                                       \
    do {                                                                                                   \
        int _s = (expr);                                                                                   \
        if (_s)                                                                                            \
            return _s;                                                                                     \
    } while (0)

// n Jacobi sweeps; *px holds the input (ignored when from_zero) and gets the output
// operators whose kernels fuse the zero-guess scaling into the first sweep
static bool lev_has_jac0(const MgHandle *h, int l)
{
    const MgLevel &L = h->lev[l];
    if (L.kind == LK_GAL1)
        return true;
    if (L.kind == LK_ELEM)
        return L.S != nullptr;
    return h->kernel == FANTOM_KU_TILE && L.dim == 3;
}

// n_smooth sweeps on x (*px); the result is returned in *px.  final
// (optional) receives the last sweep instead of the ping-pong buffer.  With
// rz (level 0 on the tiled 3D kernel only) that last sweep also leaves the
// block partials of b.final in h->partials and *rz is set.
// pro (optional): a coarse correction to prolongate onto x first, which the
// first sweep (not the last) does as it loads x (lev_pro_ok)
static int mg_smooth(MgHandle *h, int l, const double *b, double **px, bool from_zero, double *final = nullptr,
                     bool *rz = nullptr, const double *pro = nullptr)
{
    MgLevel &L = h->lev[l];
    double *x = *px;
    double *other = (x == L.x) ? L.x2 : L.x;
    // weight argument of sweep s: Jacobi's omega, or the Chebyshev weight
    // (the kernels multiply it by 1 / rho through the level's omega_scale)
    auto w = [&](int s) { return h->cheb ? h->cw[s] : L.om; };
    int s0 = 0;
    if (from_zero && h->n_smooth >= 2 && lev_has_jac0(h, l)) {
        // x = w0 D^-1 b followed by one sweep of weight w1 (L.g.w2_ratio), in one pass
        MG_TRY(lev_ku(h, l, MODE_JAC0, nullptr, x, b, L.dinv, w(0), nullptr));
        s0 = 2;
    }
    for (int s = s0; s < h->n_smooth; ++s) {
        if (s == 0 && from_zero) {
            MG_TRY(launch_scale(L.n_dof, w(0), L.dinv, b, x, h->stream, L.g.omega_scale));
            continue;
        }
        const bool last = final && s == h->n_smooth - 1;
        double *dst = last ? final : other;
        if (pro && s == 0) {
            MG_TRY(launch_ku_tile_pro(L.g, L.Kb, L.E, L.mask, x, pro, dst, b, w(0), h->stream));
        } else if (last && rz && l == 0 && h->kernel == FANTOM_KU_TILE && L.dim == 3) {
            MG_TRY(lev_ku(h, l, MODE_JACDOT, x, dst, b, L.dinv, w(s), h->partials));
            *rz = true;
        } else {
            MG_TRY(lev_ku(h, l, MODE_JAC, x, dst, b, L.dinv, w(s), nullptr));
        }
        other = x;
        x = dst;
    }
    *px = x;
    return kStatusOk;
}

// one V-cycle on level l for right-hand side b; final (optional) is where
// the result should land (the last post-smoothing sweep writes it there)
static int mg_cycle(MgHandle *h, int l, const double *b, double **out, double *final = nullptr, bool *rz = nullptr)
{
    MgLevel &L = h->lev[l];
    if (l == h->nlev - 1) {
        double *dst = final ? final : L.x;
        MG_TRY(launch_gemv(h->nc, h->Ainv, b, dst, h->stream));
        *out = dst;
        return kStatusOk;
    }
    double *x = L.x;
    MG_TRY(mg_smooth(h, l, b, &x, true));
    MgLevel &C = h->lev[l + 1];
    MG_TRY(lev_ku(h, l, MODE_RES, x, L.r, b, nullptr, 0.0, nullptr));
    MG_TRY(launch_restrict(L.g, L.dim, L.dof, 0, C.mask, L.r, C.b, h->stream));
    double *e = nullptr;
    MG_TRY(mg_cycle(h, l + 1, C.b, &e));
    // on the hex8 tile level the first post-smoothing sweep prolongates e
    // as it loads x (x + P e is never stored); it must not be the last sweep,
    // which may be the one that also sums r.z
    const bool fuse = l == 0 && h->kernel == FANTOM_KU_TILE && L.dim == 3 && h->n_smooth >= 2;
    if (!fuse)
        MG_TRY(launch_prolong(L.g, L.dim, L.dof, 0, L.mask, e, x, 1, h->stream));
    MG_TRY(mg_smooth(h, l, b, &x, false, final, rz, fuse ? e : nullptr));
    *out = x;
    return kStatusOk;
}

// z = V(r).  With np_rz, also the block partials of r.z in h->partials
// (fused into the last smoothing sweep when the fine level allows it) and
// their count in *np_rz.
static int mg_precond(MgHandle *h, const double *r, double *z, long *np_rz = nullptr)
{
    double *out = nullptr;
    bool rz = false;
    MG_TRY(mg_cycle(h, 0, r, &out, z, np_rz ? &rz : nullptr));
    if (out != z) {   // no post-smoothing sweep
        const long n = h->lev[0].n_dof;
        hipLaunchKernelGGL(mg_copy_k, dim3(vec_grid(n)), dim3(256), 0, h->stream, n, out, z);
        FANTOM_LAUNCH_CHECK("mg copy");
    }
    if (!np_rz)
        return kStatusOk;
    if (rz) {
        *np_rz = ku_tile_partials(h->lev[0].g);
    } else {
        const long n = h->lev[0].n_dof;
        MG_TRY(launch_dot_partials(n, r, z, h->partials, h->stream));
        *np_rz = (long)vec_grid(n);
    }
    return kStatusOk;
}

static int mg_iteration(MgHandle *h)
{
    const long n = h->lev[0].n_dof;
    const long npk = h->kernel == FANTOM_KU_TILE && h->dim == 3 ? ku_tile_partials(h->lev[0].g)
                                                                : ku_node_partials(h->lev[0].g);
    MG_TRY(lev_ku(h, 0, MODE_KUDOT, h->p, h->q, nullptr, nullptr, 0.0, h->partials));
    MG_TRY(launch_finalize(h->partials, npk, h->scal, FIN_ALPHA, 0, h->stream));
    hipLaunchKernelGGL(mg_r_k, dim3(vec_grid(n)), dim3(256), 0, h->stream, n, h->r, h->q, h->scal, h->partials);
    hipLaunchKernelGGL(mg_conv_k, dim3(1), dim3(1024), 0, h->stream, h->partials, (long)vec_grid(n), h->scal,
                       h->sched);
    long nrz = 0;
    MG_TRY(mg_precond(h, h->r, h->z, &nrz));
    MG_TRY(launch_finalize(h->partials, nrz, h->scal, FIN_BETA, 0, h->stream));
    hipLaunchKernelGGL(mg_px_k, dim3(vec_grid(n)), dim3(256), 0, h->stream, n, h->xs, h->p, h->z, h->scal);
    FANTOM_LAUNCH_CHECK("mg iteration");
    return kStatusOk;
}

// ---- workspace layout ----------------------------------------------------------------
struct Carve {
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

static void layout(MgHandle *h, int nelx, int nely, int nelz, int n_coarse, int galerkin, Carve &cv)
{
    h->dim = nelz > 0 ? 3 : 2;
    h->dof = h->dim;
    h->nlev = n_coarse + 1;
    h->galerkin = (galerkin && h->dim == 3) ? 1 : 0;
    h->lev.resize(h->nlev);
    for (int l = 0; l < h->nlev; ++l) {
        MgLevel &L = h->lev[l];
        L.kind = l == 0 ? LK_FINE : (!h->galerkin ? LK_REDISC : (l == 1 ? LK_GAL1 : LK_ELEM));
        L.Ke = nullptr;
        L.g = make_grid(nelx >> l, nely >> l, nelz > 0 ? (nelz >> l) : 0);
        L.dim = h->dim;
        L.dof = h->dof;
        L.n_dof = L.g.n_nodes * L.dof;
        L.scale = h->dim == 3 ? (double)(1L << l) : 1.0;
        L.x = cv.take<double>(L.n_dof);
        L.x2 = cv.take<double>(L.n_dof);
        L.b = cv.take<double>(L.n_dof);
        L.r = cv.take<double>(L.n_dof);
        L.dinv = cv.take<double>(L.n_dof);
        L.E = cv.take<double>(L.g.n_el);
        L.mask = cv.take<uint8_t>(L.g.n_nodes);
        L.K0dense_d = cv.take<double>(576);
        L.ev = cv.take<double>(L.n_dof);
        L.S = nullptr;
        if (L.kind == LK_ELEM) {
            L.Ke = cv.take<double>((size_t)L.g.n_el * 576);
            if (l < h->nlev - 1)   // the coarsest level is solved directly
                L.S = cv.take<double>((size_t)stencil_size(L.g));
        }
    }
    h->Gtab = h->Btab = h->Gdiag = h->Ke1 = h->G2tab = nullptr;
    if (h->galerkin) {
        h->Gtab = cv.take<double>(8 * 576);
        h->G2tab = cv.take<double>(64 * 576);
        h->Btab = cv.take<double>(576);
        h->Gdiag = cv.take<double>(8 * 24);
        if (h->nlev == 2)
            h->Ke1 = cv.take<double>((size_t)h->lev[1].g.n_el * 576);
    }
    const long n = h->lev[0].n_dof;
    h->xs = cv.take<double>(n);
    h->r = cv.take<double>(n);
    h->z = cv.take<double>(n);
    h->p = cv.take<double>(n);
    h->q = cv.take<double>(n);
    h->f = cv.take<double>(n);
    long np = 2 * kVecMaxBlocks;   // the Chebyshev power pass keeps two sums per block
    np = std::max(np, ku_node_partials(h->lev[0].g));
    if (h->dim == 3)
        np = std::max(np, ku_tile_partials(h->lev[0].g));
    for (int l = 1; l < h->nlev; ++l) {
        np = std::max(np, elem_partials(h->lev[l].g));
        if (h->dim == 3) {
            np = std::max(np, gal1_partials(h->lev[l].g));
            np = std::max(np, stencil_partials(h->lev[l].g));
        }
    }
    h->n_partials = np;
    h->partials = cv.take<double>(np);
    h->scal = cv.take<double>(FANTOM_SCALARS);
    h->chs = cv.take<double>(4 * (size_t)h->nlev);
    h->nc = (int)h->lev[h->nlev - 1].n_dof;
    h->Ainv = cv.take<double>((size_t)h->nc * h->nc);
    h->gj_work = cv.take<double>((size_t)h->nc * h->nc + 1024);
}

static bool levels_ok(int nelx, int nely, int nelz, int n_coarse)
{
    const int m = 1 << n_coarse;
    if (n_coarse < 0 || n_coarse > 10)
        return false;
    if (nelx % m || nely % m || (nelz > 0 && nelz % m))
        return false;
    return true;
}

}  // namespace fantom

using namespace fantom;

extern "C" {

long fantom_mg_workspace_size(int nelx, int nely, int nelz, int n_coarse, int galerkin)
{
    if (!grid_ok(nelx, nely, nelz) || !levels_ok(nelx, nely, nelz, n_coarse))
        return -1;
    MgHandle h;
    Carve cv{nullptr, 0};
    layout(&h, nelx, nely, nelz, n_coarse, galerkin, cv);
    return (long)cv.off + 256;
}

int fantom_mg_create(int nelx, int nely, int nelz, int n_coarse, int n_smooth, double omega, int kernel,
                     int galerkin, const double *K0_host, void *workspace, long workspace_bytes, void **handle)
{
    FANTOM_CHECK_ARG(grid_ok(nelx, nely, nelz), "invalid grid %d x %d x %d", nelx, nely, nelz);
    FANTOM_CHECK_ARG(levels_ok(nelx, nely, nelz, n_coarse), "grid %d x %d x %d does not halve %d times", nelx, nely,
                     nelz, n_coarse);
    FANTOM_CHECK_ARG(K0_host && workspace && handle, "null operand");
    FANTOM_CHECK_ARG(n_smooth >= 1 && n_smooth <= 16, "n_smooth must be in [1, 16]");
    FANTOM_CHECK_ARG(kernel == FANTOM_KU_NODE || kernel == FANTOM_KU_TILE, "unknown kernel %d", kernel);
    const long need = fantom_mg_workspace_size(nelx, nely, nelz, n_coarse, galerkin);
    FANTOM_CHECK_ARG(workspace_bytes >= need, "workspace of %ld bytes is smaller than the %ld needed",
                     workspace_bytes, need);
    MgHandle *h = new MgHandle();
    Carve cv{reinterpret_cast<char *>(workspace), 0};
    // align the base
    cv.off = (256 - (reinterpret_cast<uintptr_t>(workspace) & 255)) & 255;
    layout(h, nelx, nely, nelz, n_coarse, galerkin, cv);
    FANTOM_CHECK_ARG(h->nc <= 20000, "coarsest grid of %d DOF is too large for the dense solve", h->nc);
    h->n_smooth = n_smooth;
    h->omega = omega;
    h->sched = OmegaSchedule{0, omega, 0.0};
    for (auto &L : h->lev) {
        L.om = omega;
        L.g.omega_scale = nullptr;
    }
    h->kernel = (h->dim == 3) ? kernel : FANTOM_KU_NODE;
    h->have_graph = false;
    h->last_iters = 0;
    h->cheb = 0;
    h->cheb_scale = 1.0;
    h->cheb_iters_first = 12;
    h->cheb_iters = 4;
    h->cheb_setups = 0;
    const int nl = h->dim == 3 ? 24 : 8;
    for (int l = 0; l < h->nlev; ++l) {
        MgLevel &L = h->lev[l];
        for (int i = 0; i < nl * nl; ++i)
            L.K0dense[i] = K0_host[i] * L.scale;
        if (h->dim == 3) {
            int st = fantom_hex8_blocks(L.K0dense, L.Kb, nullptr);
            if (st && h->kernel == FANTOM_KU_TILE) {
                delete h;
                return st;
            }
        }
    }
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete h;
        set_error("stream create: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    for (int l = 0; l < h->nlev; ++l) {
        MgLevel &L = h->lev[l];
        e = hipMemcpyAsync(L.K0dense_d, L.K0dense, sizeof(double) * nl * nl, hipMemcpyHostToDevice, h->stream);
        if (e != hipSuccess)
            break;
    }
    std::vector<double> tG(8 * 576), tB(576), tD(8 * 24), tG2(64 * 576);
    if (h->galerkin) {
        int st = fantom_galerkin_tables(K0_host, tG.data(), tB.data(), tD.data(), nullptr);
        if (st) {
            hipStreamDestroy(h->stream);
            delete h;
            return st;
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(h->Gtab, tG.data(), sizeof(double) * tG.size(), hipMemcpyHostToDevice, h->stream);
        galerkin_tables2(tG.data(), tG2.data());
        if (e == hipSuccess)
            e = hipMemcpyAsync(h->G2tab, tG2.data(), sizeof(double) * tG2.size(), hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(h->Btab, tB.data(), sizeof(double) * tB.size(), hipMemcpyHostToDevice, h->stream);
        h->gal_cls_ok = gal1_class_values(tB.data(), h->gal_cls);
        if (e == hipSuccess)
            e = hipMemcpyAsync(h->Gdiag, tD.data(), sizeof(double) * tD.size(), hipMemcpyHostToDevice, h->stream);
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        hipStreamDestroy(h->stream);
        delete h;
        set_error("K0 upload: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    *handle = h;
    return kStatusOk;
}

int fantom_mg_set_omega_schedule(void *handle, int adaptive, double min_omega, double level_boost)
{
    MgHandle *h = reinterpret_cast<MgHandle *>(handle);
    FANTOM_CHECK_ARG(h, "null handle");
    FANTOM_CHECK_ARG(!adaptive || (min_omega > 0.0 && level_boost > 0.0),
                     "adaptive schedule needs min_omega > 0 and level_boost > 0");
    h->sched = adaptive ? OmegaSchedule{1, min_omega, h->omega - min_omega} : OmegaSchedule{0, h->omega, 0.0};
    h->cheb = 0;   // the weight schedule replaces a Chebyshev smoother
    double boost = 1.0;
    for (auto &L : h->lev) {
        L.om = adaptive ? boost : h->omega;
        L.g.omega_scale = adaptive ? h->scal + D_OMEGA : nullptr;
        L.g.w2_ratio = 1.0;
        boost *= level_boost;
    }
    if (adaptive) {   // the weight of a V-cycle applied before any solve
        const hipError_t e = hipMemcpy(h->scal + D_OMEGA, &min_omega, sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            set_error("omega schedule: %s", hipGetErrorString(e));
            return kStatusHipError;
        }
    }
    if (h->have_graph) {   // the captured iteration holds the old weights
        hipGraphExecDestroy(h->iter_exec);
        h->have_graph = false;
    }
    return kStatusOk;
}

int fantom_mg_set_smoother(void *handle, int kind, double rho_scale, int iters_first, int iters)
{
    MgHandle *h = reinterpret_cast<MgHandle *>(handle);
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
        MgLevel &L = h->lev[l];
        L.om = h->omega;
        L.g.omega_scale = kind == 1 ? h->chs + 4 * l + 2 : nullptr;
        L.g.w2_ratio = (kind == 1 && h->n_smooth >= 2) ? h->cw[1] / h->cw[0] : 1.0;
    }
    if (h->have_graph) {   // the captured iteration holds the old weights
        hipGraphExecDestroy(h->iter_exec);
        h->have_graph = false;
    }
    return kStatusOk;
}

int fantom_mg_smoother_rho(void *handle, double *rho, int n)
{
    MgHandle *h = reinterpret_cast<MgHandle *>(handle);
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

int fantom_mg_destroy(void *handle)
{
    MgHandle *h = reinterpret_cast<MgHandle *>(handle);
    if (!h)
        return kStatusOk;
    hipStreamSynchronize(h->stream);
    if (h->have_graph)
        hipGraphExecDestroy(h->iter_exec);
    hipStreamDestroy(h->stream);
    delete h;
    return kStatusOk;
}

// Make the handle's stream wait for the caller's stream (and vice versa at the end).
static int join(hipStream_t from, hipStream_t to)
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

int fantom_mg_setup(void *handle, const double *E, const uint8_t *mask, void *stream)
{
    MgHandle *h = reinterpret_cast<MgHandle *>(handle);
    FANTOM_CHECK_ARG(h && E, "null operand");
    MG_TRY(join(as_stream(stream), h->stream));
    MgLevel &L0 = h->lev[0];
    hipError_t e = hipMemcpyAsync(L0.E, E, sizeof(double) * L0.g.n_el, hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) {
        if (mask)
            e = hipMemcpyAsync(L0.mask, mask, L0.g.n_nodes, hipMemcpyDeviceToDevice, h->stream);
        else
            e = hipMemsetAsync(L0.mask, 0, L0.g.n_nodes, h->stream);
    }
    if (e != hipSuccess) {
        set_error("setup copy: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    for (int l = 1; l < h->nlev; ++l) {
        MgLevel &F = h->lev[l - 1], &C = h->lev[l];
        if (!h->galerkin)
            MG_TRY(launch_coarsen_E(F.g, h->dim, F.E, C.E, h->stream));
        MG_TRY(launch_coarsen_mask(F.g, h->dim, F.mask, C.mask, h->stream));
    }
    if (h->galerkin) {
        // explicit element matrices from level 2 down
        for (int l = 2; l < h->nlev; ++l) {
            MgLevel &C = h->lev[l];
            if (l == 2)
                MG_TRY(launch_gal2_build(C.g, h->G2tab, L0.E, C.Ke, h->stream));
            else
                MG_TRY(launch_gal_build(C.g, 1, h->Gtab, nullptr, h->lev[l - 1].Ke, C.Ke, h->stream));
        }
        if (h->nlev == 2)
            MG_TRY(launch_elem_from_fine(h->lev[1].g, h->Gtab, L0.E, h->Ke1, h->stream));
    }
    for (int l = 0; l < h->nlev; ++l) {
        MgLevel &L = h->lev[l];
        if (L.kind == LK_GAL1)
            MG_TRY(launch_dinv_gal1(L.g, h->Gdiag, L0.E, L.mask, L.dinv, h->stream));
        else if (L.kind == LK_ELEM && L.S) {
            MG_TRY(launch_stencil_build(L.g, L.Ke, L.mask, L.S, h->stream));
            MG_TRY(launch_dinv_stencil(L.g, L.S, L.dinv, h->stream));
        } else if (L.kind == LK_ELEM)
            MG_TRY(launch_dinv_elem(L.g, L.Ke, L.mask, L.dinv, h->stream));
        else
            MG_TRY(launch_diag(L.g, h->dim, L.K0dense_d, L.E, L.mask, L.dinv, 1, h->stream));
    }
    if (h->cheb) {
        // rho of D^-1 A per smoothed level; the first setup starts from a
        // fixed pseudo-random vector, later ones from the previous vector
        const bool first = h->cheb_setups == 0;
        const int iters = first ? h->cheb_iters_first : h->cheb_iters;
        for (int l = 0; l < h->nlev - 1 && iters > 0; ++l) {
            MgLevel &L = h->lev[l];
            double *st = h->chs + 4 * l;
            MG_TRY(launch_ch_init(L.n_dof, L.ev, L.dinv, first ? 1 : 0, l, h->partials, st, h->stream));
            for (int it = 0; it < iters; ++it) {
                MG_TRY(lev_ku(h, l, MODE_KU, L.ev, L.x2, nullptr, nullptr, 0.0, nullptr));
                MG_TRY(launch_ch_pass(L.n_dof, L.ev, L.x2, L.dinv, h->cheb_scale, h->partials, st, h->stream));
            }
        }
        ++h->cheb_setups;
    }
    MgLevel &C = h->lev[h->nlev - 1];
    if (C.kind == LK_ELEM)
        MG_TRY(launch_dense_from_elem(C.g, C.Ke, C.mask, h->Ainv, h->stream));
    else if (C.kind == LK_GAL1)
        MG_TRY(launch_dense_from_elem(C.g, h->Ke1, C.mask, h->Ainv, h->stream));
    else
        MG_TRY(launch_dense_assemble(C.g, h->dim, C.K0dense_d, C.E, C.mask, h->Ainv, h->stream));
    MG_TRY(launch_spd_inverse(h->nc, h->Ainv, h->gj_work, h->stream));
    return join(h->stream, as_stream(stream));
}

int fantom_mg_precondition(void *handle, const double *r, double *z, void *stream)
{
    MgHandle *h = reinterpret_cast<MgHandle *>(handle);
    FANTOM_CHECK_ARG(h && r && z, "null operand");
    MG_TRY(join(as_stream(stream), h->stream));
    MG_TRY(mg_precond(h, r, z));
    return join(h->stream, as_stream(stream));
}

static int capture_iteration(MgHandle *h)
{
    hipGraph_t graph;
    hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        set_error("begin capture: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    int st = mg_iteration(h);
    e = hipStreamEndCapture(h->stream, &graph);
    if (st)
        return st;
    if (e != hipSuccess) {
        set_error("end capture: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    e = hipGraphInstantiate(&h->iter_exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) {
        set_error("graph instantiate: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    h->have_graph = true;
    return kStatusOk;
}

int fantom_mg_solve(void *handle, const double *f, double *x, double tol, int maxiter, int use_x0,
                    int use_graph, int *iters, double *rel_res, void *stream)
{
    MgHandle *h = reinterpret_cast<MgHandle *>(handle);
    FANTOM_CHECK_ARG(h && f && x && iters && rel_res, "null operand");
    FANTOM_CHECK_ARG(maxiter >= 1, "maxiter must be positive");
    const long n = h->lev[0].n_dof;
    MG_TRY(join(as_stream(stream), h->stream));
    hipStream_t s = h->stream;
    const unsigned G = vec_grid(n);
    // x0, ||f||, r = f - K x0, z = V(r), p = z, rz = r.z
    hipLaunchKernelGGL(mg_copy_k, dim3(G), dim3(256), 0, s, n, f, h->f);
    if (use_x0)
        hipLaunchKernelGGL(mg_copy_k, dim3(G), dim3(256), 0, s, n, x, h->xs);
    else
        hipMemsetAsync(h->xs, 0, sizeof(double) * n, s);
    hipLaunchKernelGGL(mg_init_scal_k, dim3(1), dim3(64), 0, s, h->scal, tol, h->sched.min);
    MG_TRY(launch_dot_partials(n, h->f, h->f, h->partials, s));
    hipLaunchKernelGGL(mg_norm_k, dim3(1), dim3(1024), 0, s, h->partials, (long)G, h->scal);
    MG_TRY(lev_ku(h, 0, MODE_RES, h->xs, h->r, h->f, nullptr, 0.0, nullptr));
    long nrz = 0;
    MG_TRY(mg_precond(h, h->r, h->z, &nrz));
    MG_TRY(launch_finalize(h->partials, nrz, h->scal, FIN_RZ0, 0, s));
    MG_TRY(launch_pcg_p(n, h->p, h->z, h->scal, 1, s));
    FANTOM_LAUNCH_CHECK("mg solve init");
    if (use_graph && !h->have_graph)
        MG_TRY(capture_iteration(h));
    double hs[FANTOM_SCALARS];
    int done_iters = 0;
    // first batch: close to the previous solve's count, then single steps
    int batch = h->last_iters > 3 ? h->last_iters - 2 : 1;
    bool converged = false;
    while (done_iters < maxiter) {
        if (batch > maxiter - done_iters)
            batch = maxiter - done_iters;
        for (int b = 0; b < batch; ++b) {
            if (use_graph) {
                hipError_t e = hipGraphLaunch(h->iter_exec, s);
                if (e != hipSuccess) {
                    set_error("graph launch: %s", hipGetErrorString(e));
                    return kStatusHipError;
                }
            } else {
                MG_TRY(mg_iteration(h));
            }
        }
        done_iters += batch;
        hipError_t e = hipMemcpyAsync(hs, h->scal, sizeof(hs), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
            e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            set_error("convergence read: %s", hipGetErrorString(e));
            return kStatusHipError;
        }
        if (hs[D_DONE] != 0.0) {
            converged = true;
            break;
        }
        batch = 1;
    }
    *iters = converged ? (int)hs[D_ITER] : done_iters;
    h->last_iters = *iters;
    // true residual ||f - K x|| / ||f||
    MG_TRY(lev_ku(h, 0, MODE_RES, h->xs, h->q, h->f, nullptr, 0.0, nullptr));
    MG_TRY(launch_dot_partials(n, h->q, h->q, h->partials, s));
    MG_TRY(launch_finalize(h->partials, (long)G, h->scal, FIN_STORE, S_TMP, s));
    hipLaunchKernelGGL(mg_copy_k, dim3(G), dim3(256), 0, s, n, h->xs, x);
    hipError_t e = hipMemcpyAsync(hs, h->scal, sizeof(hs), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        set_error("final read: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    *rel_res = sqrt(fmax(hs[S_TMP], 0.0)) / hs[D_NB];
    return join(h->stream, as_stream(stream));
}

}  // extern "C"
