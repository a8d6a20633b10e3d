// extern "C" layer of libfantom_hip.so: argument checks, status codes and
// dispatch to the kernel launchers.  Declarations and the reference interface
// each entry point replaces are in include/fantom.h.
#include <cmath>
#include <cstdarg>

#include "../../include/fantom.h"
#include "common.hpp"
#include "fantom_internal.hpp"
#include "vec.hpp"

namespace fantom {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char *get_error() { return g_err; }

static int check_grid(int nelx, int nely, int nelz)
{
    FANTOM_CHECK_ARG(grid_ok(nelx, nely, nelz), "invalid grid %d x %d x %d", nelx, nely, nelz);
    return kStatusOk;
}

}  // namespace fantom

using namespace fantom;

#define GRID_OR_RETURN(nelx, nely, nelz)                                                                   \
    do {                                                                                                   \
        int _st = check_grid(nelx, nely, nelz);                                                            \
        if (_st)                                                                                           \
            return _st;                                                                                    \
    } while (0)

extern "C" {

int fantom_abi_version(void) { return FANTOM_ABI_VERSION; }

const char *fantom_last_error(void) { return get_error(); }

int fantom_stream_sync(void *stream)
{
    hipError_t e = hipStreamSynchronize(as_stream(stream));
    if (e != hipSuccess) {
        set_error("stream sync: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    return kStatusOk;
}

static int ku_dispatch(int kernel, int mode, int nelx, int nely, int nelz, const double *K0, const double *E,
                       const uint8_t *mask, const double *u, double *y, const double *b, const double *dinv,
                       double omega, double *partials, void *stream)
{
    GRID_OR_RETURN(nelx, nely, nelz);
    FANTOM_CHECK_ARG(K0 && E && (u || mode == MODE_JAC0) && y, "null operand");
    const Grid g = make_grid(nelx, nely, nelz);
    const int dim = nelz > 0 ? 3 : 2;
    if (kernel == FANTOM_KU_TILE) {
        FANTOM_CHECK_ARG(dim == 3, "the tiled kernel is 3D (hex8) only");
        return launch_ku_tile(mode, g, K0, E, mask, u, y, b, dinv, omega, partials, as_stream(stream));
    }
    FANTOM_CHECK_ARG(kernel == FANTOM_KU_NODE, "unknown kernel %d", kernel);
    return launch_ku_node(mode, g, dim, K0, E, mask, u, y, b, dinv, omega, partials, as_stream(stream));
}

int fantom_ku_f64(int kernel, int nelx, int nely, int nelz, const double *K0, const double *E, const uint8_t *mask,
                  const double *u, double *y, void *stream)
{
    FANTOM_CHECK_ARG(u != y, "u and y must not alias");
    return ku_dispatch(kernel, MODE_KU, nelx, nely, nelz, K0, E, mask, u, y, nullptr, nullptr, 0.0, nullptr, stream);
}

int fantom_ku_f32(int nelx, int nely, int nelz, const double *Kb, const float *E, const uint8_t *mask,
                  const float *u, float *y, void *stream)
{
    GRID_OR_RETURN(nelx, nely, nelz);
    if (nelz <= 0) {
        set_error("the FP32 operator is implemented for hex8 grids (nelz > 0)");
        return kStatusNotSupported;
    }
    FANTOM_CHECK_ARG(Kb && E && u && y, "null operand");
    FANTOM_CHECK_ARG(u != y, "u and y must not alias");
    return launch_ku_tile_f32(MODE_KU, make_grid(nelx, nely, nelz), Kb, E, mask, u, y, nullptr, nullptr, 0.0, nullptr,
                              as_stream(stream));
}

int fantom_ku_jacobi_f64(int kernel, int nelx, int nely, int nelz, const double *K0, const double *E,
                         const uint8_t *mask, const double *b, const double *dinv, double omega, const double *x_in,
                         double *x_out, void *stream)
{
    FANTOM_CHECK_ARG(b && (dinv || kernel == FANTOM_KU_TILE), "null operand");
    FANTOM_CHECK_ARG(x_in != x_out, "x_in and x_out must not alias");
    return ku_dispatch(kernel, MODE_JAC, nelx, nely, nelz, K0, E, mask, x_in, x_out, b, dinv, omega, nullptr, stream);
}

int fantom_ku_jacobi_prolong_f64(int nelx, int nely, int nelz, const double *Kb, const double *E,
                                 const uint8_t *mask, const double *b, double omega, const double *x_in,
                                 const double *ec, double *x_out, void *stream)
{
    GRID_OR_RETURN(nelx, nely, nelz);
    FANTOM_CHECK_ARG(nelz > 0, "the prolongating sweep is implemented for hex8 grids (nelz > 0)");
    FANTOM_CHECK_ARG(Kb && E && b && x_in && ec && x_out, "null operand");
    FANTOM_CHECK_ARG(x_in != x_out && ec != x_out, "x_out must not alias x_in or ec");
    return launch_ku_tile_pro(make_grid(nelx, nely, nelz), Kb, E, mask, x_in, ec, x_out, b, omega,
                              as_stream(stream));
}

int fantom_ku_jacobi_prolong_f32(int nelx, int nely, int nelz, const double *Kb, const float *E,
                                 const uint8_t *mask, const float *b, double omega, const float *x_in,
                                 const float *ec, float *x_out, void *stream)
{
    GRID_OR_RETURN(nelx, nely, nelz);
    FANTOM_CHECK_ARG(nelz > 0, "the prolongating sweep is implemented for hex8 grids (nelz > 0)");
    FANTOM_CHECK_ARG(Kb && E && b && x_in && ec && x_out, "null operand");
    FANTOM_CHECK_ARG(x_in != x_out && ec != x_out, "x_out must not alias x_in or ec");
    return launch_ku_tile_pro_f32(make_grid(nelx, nely, nelz), Kb, E, mask, x_in, ec, x_out, b, omega,
                                  as_stream(stream));
}

int fantom_ku_jacobi0_f64(int kernel, int nelx, int nely, int nelz, const double *K0, const double *E,
                          const uint8_t *mask, const double *b, const double *dinv, double omega, double *x_out,
                          void *stream)
{
    FANTOM_CHECK_ARG(b && dinv && x_out, "null operand");
    FANTOM_CHECK_ARG(kernel == FANTOM_KU_TILE && nelz > 0, "the fused first sweep is implemented by the tiled kernel");
    return ku_dispatch(kernel, MODE_JAC0, nelx, nely, nelz, K0, E, mask, nullptr, x_out, b, dinv, omega, nullptr,
                       stream);
}

int fantom_ku_residual_f64(int kernel, int nelx, int nely, int nelz, const double *K0, const double *E,
                           const uint8_t *mask, const double *b, const double *x, double *r, void *stream)
{
    FANTOM_CHECK_ARG(b, "null operand");
    FANTOM_CHECK_ARG(x != r, "x and r must not alias");
    return ku_dispatch(kernel, MODE_RES, nelx, nely, nelz, K0, E, mask, x, r, b, nullptr, 0.0, nullptr, stream);
}

int fantom_ku_dot_f64(int kernel, int nelx, int nely, int nelz, const double *K0, const double *E,
                      const uint8_t *mask, const double *u, double *y, double *partials, void *stream)
{
    FANTOM_CHECK_ARG(partials, "null partials");
    FANTOM_CHECK_ARG(u != y, "u and y must not alias");
    return ku_dispatch(kernel, MODE_KUDOT, nelx, nely, nelz, K0, E, mask, u, y, nullptr, nullptr, 0.0, partials,
                       stream);
}

long fantom_ku_partials(int kernel, int nelx, int nely, int nelz)
{
    if (!grid_ok(nelx, nely, nelz))
        return -1;
    const Grid g = make_grid(nelx, nely, nelz);
    return kernel == FANTOM_KU_TILE ? ku_tile_partials(g) : ku_node_partials(g);
}

int fantom_hex8_blocks(const double *K0, double *blocks, double *max_offblock)
{
    FANTOM_CHECK_ARG(K0 && blocks, "null operand");
    // natural corner index b = bx + 2by + 4bz -> reference local index
    static const int loc[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    double Kn[24][24];
    for (int b = 0; b < 8; ++b)
        for (int c = 0; c < 3; ++c)
            for (int g2 = 0; g2 < 8; ++g2)
                for (int d = 0; d < 3; ++d)
                    Kn[b * 3 + c][g2 * 3 + d] = K0[(loc[b] * 3 + c) * 24 + loc[g2] * 3 + d];
    auto H = [](int b, int t) { return (__builtin_popcount(b & t) & 1) ? -1.0 : 1.0; };
    // Kh[(s,c),(s',d)] = 1/64 sum_{b,b'} H[b][s^e_c] Kn[(b,c),(b',d)] H[b'][s'^e_d]
    double maxoff = 0.0, maxabs = 0.0;
    for (int s = 0; s < 8; ++s)
        for (int c = 0; c < 3; ++c)
            for (int s2 = 0; s2 < 8; ++s2)
                for (int d = 0; d < 3; ++d) {
                    const int tc = s ^ (1 << c), td = s2 ^ (1 << d);
                    double v = 0.0;
                    for (int b = 0; b < 8; ++b)
                        for (int b2 = 0; b2 < 8; ++b2)
                            v += H(b, tc) * Kn[b * 3 + c][b2 * 3 + d] * H(b2, td);
                    v /= 64.0;
                    maxabs = fmax(maxabs, fabs(v));
                    if (s == s2)
                        blocks[9 * s + 3 * c + d] = v;
                    else
                        maxoff = fmax(maxoff, fabs(v));
                }
    if (max_offblock)
        *max_offblock = maxabs > 0 ? maxoff / maxabs : 0.0;
    if (maxoff > 1e-12 * maxabs) {
        set_error("K0 is not block-diagonal in the Walsh-Hadamard basis (relative off-block %.3e)",
                  maxabs > 0 ? maxoff / maxabs : 0.0);
        return kStatusBadArg;
    }
    return kStatusOk;
}

int fantom_diag_f64(int nelx, int nely, int nelz, const double *K0, const double *E, const uint8_t *mask, double *d,
                    int invert, void *stream)
{
    GRID_OR_RETURN(nelx, nely, nelz);
    FANTOM_CHECK_ARG(K0 && E && d, "null operand");
    const Grid g = make_grid(nelx, nely, nelz);
    return launch_diag(g, nelz > 0 ? 3 : 2, K0, E, mask, d, invert, as_stream(stream));
}

int fantom_dc_f64(int nelx, int nely, int nelz, const double *K0, const double *u, double *dc, void *stream)
{
    GRID_OR_RETURN(nelx, nely, nelz);
    FANTOM_CHECK_ARG(K0 && u && dc, "null operand");
    const Grid g = make_grid(nelx, nely, nelz);
    return launch_dc(g, nelz > 0 ? 3 : 2, K0, u, dc, as_stream(stream));
}

int fantom_restrict_f64(int nelx, int nely, int nelz, int normalize, const uint8_t *mask_c, const double *rf,
                        double *rc, void *stream)
{
    GRID_OR_RETURN(nelx, nely, nelz);
    FANTOM_CHECK_ARG(rf && rc && rf != rc, "bad operands");
    const Grid g = make_grid(nelx, nely, nelz);
    const int dim = nelz > 0 ? 3 : 2;
    return launch_restrict(g, dim, dim, normalize, mask_c, rf, rc, as_stream(stream));
}

int fantom_prolong_f64(int nelx, int nely, int nelz, int normalize, const uint8_t *mask_f, const double *ec,
                       double *xf, int accumulate, void *stream)
{
    GRID_OR_RETURN(nelx, nely, nelz);
    FANTOM_CHECK_ARG(ec && xf && ec != xf, "bad operands");
    const Grid g = make_grid(nelx, nely, nelz);
    const int dim = nelz > 0 ? 3 : 2;
    return launch_prolong(g, dim, dim, normalize, mask_f, ec, xf, accumulate, as_stream(stream));
}

int fantom_coarsen_E_f64(int nelx, int nely, int nelz, const double *Ef, double *Ec, void *stream)
{
    GRID_OR_RETURN(nelx, nely, nelz);
    FANTOM_CHECK_ARG(Ef && Ec, "null operand");
    FANTOM_CHECK_ARG(nelx % 2 == 0 && nely % 2 == 0 && (nelz % 2 == 0), "grid %d x %d x %d is not even", nelx, nely,
                     nelz);
    return launch_coarsen_E(make_grid(nelx, nely, nelz), nelz > 0 ? 3 : 2, Ef, Ec, as_stream(stream));
}

int fantom_coarsen_mask(int nelx, int nely, int nelz, const uint8_t *mf, uint8_t *mc, void *stream)
{
    GRID_OR_RETURN(nelx, nely, nelz);
    FANTOM_CHECK_ARG(mf && mc, "null operand");
    return launch_coarsen_mask(make_grid(nelx, nely, nelz), nelz > 0 ? 3 : 2, mf, mc, as_stream(stream));
}

int fantom_dense_assemble_f64(int nelx, int nely, int nelz, const double *K0, const double *E, const uint8_t *mask,
                              double *A, void *stream)
{
    GRID_OR_RETURN(nelx, nely, nelz);
    FANTOM_CHECK_ARG(K0 && E && A, "null operand");
    const Grid g = make_grid(nelx, nely, nelz);
    const int dof = nelz > 0 ? 3 : 2;
    FANTOM_CHECK_ARG(g.n_nodes * dof <= 20000, "coarsest grid too large for a dense solve (%ld DOF)",
                     g.n_nodes * dof);
    return launch_dense_assemble(g, nelz > 0 ? 3 : 2, K0, E, mask, A, as_stream(stream));
}

int fantom_spd_inverse_f64(int n, double *A, double *work, void *stream)
{
    FANTOM_CHECK_ARG(n > 0 && n <= 20000 && A && work, "bad operands");
    return launch_spd_inverse(n, A, work, as_stream(stream));
}

int fantom_gemv_f64(int n, const double *A, const double *x, double *y, void *stream)
{
    FANTOM_CHECK_ARG(n > 0 && A && x && y && x != y, "bad operands");
    return launch_gemv(n, A, x, y, as_stream(stream));
}

long fantom_vec_partials(long n) { return (long)vec_grid(n); }

long fantom_oc_partials(long n) { return oc_partials(n); }

int fantom_dot_partials_f64(long n, const double *a, const double *b, double *partials, void *stream)
{
    FANTOM_CHECK_ARG(n > 0 && a && b && partials, "bad operands");
    return launch_dot_partials(n, a, b, partials, as_stream(stream));
}

int fantom_dot2_partials_f64(long n, const double *a, const double *b, const double *c, double *partials,
                             void *stream)
{
    FANTOM_CHECK_ARG(n > 0 && a && b && c && partials, "bad operands");
    return launch_dot2_partials(n, a, b, c, partials, as_stream(stream));
}

int fantom_finalize_f64(const double *partials, long np, double *scal, int op, int slot, void *stream)
{
    FANTOM_CHECK_ARG(np > 0 && partials && scal, "bad operands");
    FANTOM_CHECK_ARG(slot >= 0 && slot < FANTOM_SCALARS, "bad slot %d", slot);
    FANTOM_CHECK_ARG(op >= FANTOM_FIN_STORE && op <= FANTOM_FIN_RZ0, "bad op %d", op);
    return launch_finalize(partials, np, scal, op, slot, as_stream(stream));
}

int fantom_pcg_xr_f64(long n, double *x, double *r, const double *p, const double *q, const double *scal,
                      double *partials, void *stream)
{
    FANTOM_CHECK_ARG(n > 0 && x && r && p && q && scal && partials, "bad operands");
    return launch_pcg_xr(n, x, r, p, q, scal, partials, as_stream(stream));
}

int fantom_pcg_p_f64(long n, double *p, const double *z, const double *scal, int first, void *stream)
{
    FANTOM_CHECK_ARG(n > 0 && p && z && scal, "bad operands");
    return launch_pcg_p(n, p, z, scal, first, as_stream(stream));
}

int fantom_scale_f64(long n, double omega, const double *dinv, const double *b, double *x, void *stream)
{
    FANTOM_CHECK_ARG(n > 0 && dinv && b && x, "bad operands");
    return launch_scale(n, omega, dinv, b, x, as_stream(stream));
}

int fantom_axpby_f64(long n, double a, const double *x, double b, double *y, void *stream)
{
    FANTOM_CHECK_ARG(n > 0 && x && y, "bad operands");
    return launch_axpby(n, a, x, b, y, as_stream(stream));
}

int fantom_filter_normalizer_f64(int nelx, int nely, int nelz, int ntap, const int32_t *offsets, const double *weights,
                                 double *norm, void *stream)
{
    GRID_OR_RETURN(nelx, nely, nelz);
    FANTOM_CHECK_ARG(offsets && weights && norm, "null operand");
    return launch_filter_normalizer(nelz > 0 ? 3 : 2, nelx, nely, nelz, ntap, offsets, weights, norm,
                                    as_stream(stream));
}

int fantom_filter_fwd_f64(int nelx, int nely, int nelz, int ntap, const int32_t *offsets, const double *weights,
                          const double *x, double *y, void *stream)
{
    GRID_OR_RETURN(nelx, nely, nelz);
    FANTOM_CHECK_ARG(offsets && weights && x && y && x != y, "bad operands");
    return launch_filter_fwd(nelz > 0 ? 3 : 2, nelx, nely, nelz, ntap, offsets, weights, x, y, as_stream(stream));
}

int fantom_filter_adj_f64(int nelx, int nely, int nelz, int ntap, const int32_t *offsets, const double *weights,
                          const double *norm, const double *y, double *x, void *stream)
{
    GRID_OR_RETURN(nelx, nely, nelz);
    FANTOM_CHECK_ARG(offsets && weights && norm && x && y && x != y, "bad operands");
    return launch_filter_adj(nelz > 0 ? 3 : 2, nelx, nely, nelz, ntap, offsets, weights, norm, y, x,
                             as_stream(stream));
}

int fantom_oc_update_f64(long n, const double *x, const double *df, const double *dg, double vf, double move,
                         double *ocp, double *state, double *partials, double *x_new, void *stream)
{
    FANTOM_CHECK_ARG(n > 0 && x && df && dg && ocp && state && partials && x_new, "bad operands");
    FANTOM_CHECK_ARG(x != x_new, "x and x_new must not alias");
    hipStream_t s = as_stream(stream);
    int st = launch_oc_prepare(n, x, df, dg, ocp, state, partials, s);
    if (st)
        return st;
    // 4 bisection levels per pass; from [1e-9, 1e9] down to a relative width of
    // 1e-8 takes at most ~87 levels (multiplier near 1e-9); finished passes exit
    // at once on the device-side done flag
    st = launch_oc_bisect(n, x, ocp, dg, move, vf, state, partials, 24, s);
    if (st)
        return st;
    return launch_oc_apply(n, x, ocp, move, state, x_new, partials, s);
}

}  // extern "C"
