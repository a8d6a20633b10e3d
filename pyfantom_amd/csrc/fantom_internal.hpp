// Launchers shared between the kernel files and the extern "C" layer.
#pragma once
#include "common.hpp"

namespace fantom {

// matvec epilogues: y = K u | y = u + omega D^-1 (b - K u) | y = b - K u |
// y = K u with partial sums of u.(K u) | MODE_JAC0: the Jacobi sweep from
// x0 = omega D^-1 b, with x0 formed on load (u unused); equals a scaling
// followed by MODE_JAC, in one pass | MODE_JACDOT: MODE_JAC with the partial
// sums of b.y (the PCG r.z when the sweep ends a V-cycle on r), tiled kernel
// | MODE_JACOWN: MODE_JAC with D^-1 formed from the moduli (fine hex8 tile
// operator; launch_ku_tile picks it for MODE_JAC with dinv = NULL).  On the
// fine hex8 tile operator MODE_JACDOT forms D^-1 the same way | MODE_JAC0OWN:
// MODE_JAC0 for a dinv that is the operator's own D^-1 (the drivers' fine
// level): the sweep reuses x0 = omega D^-1 b, y = 2 x0 - omega D^-1 K x0 with
// D^-1 formed from the moduli, so b and dinv are read once, not twice.
enum {
    MODE_KU = 0,
    MODE_JAC = 1,
    MODE_RES = 2,
    MODE_KUDOT = 3,
    MODE_JAC0 = 4,
    MODE_JACDOT = 5,
    MODE_JACOWN = 6,
    MODE_JAC0OWN = 7,
    // MODE_JACOWN on u + P e (hex8 tile kernels only): the first
    // post-smoothing sweep of the fine level with the prolongation of the
    // coarse correction e (the transfers' trilinear P, normalize 0) formed
    // as the node values load, so that u + P e is never stored
    MODE_JACPRO = 8
};

// Adaptive Jacobi weight of the reference's MultiGrid.solve
// (solvers/CUDA/_solvers.py:626-648): omega = min_omega before the first
// cycle and, after each iteration's convergence test on R = ||r|| / ||f||,
// omega = min_omega + (max_omega - min_omega) / 2 * exp(500 (tol - clip(R, tol, 0.1))),
// times 1.06^level on each level (:524-526).
struct OmegaSchedule {
    int on;
    double min, delta;   // min_omega, max_omega - min_omega
};

__device__ __forceinline__ double schedule_omega(const OmegaSchedule &o, double R, double tol)
{
    const double Rc = fmin(fmax(R, tol), 0.1);
    return o.min + o.delta / 2 * exp((-Rc + tol) * 500.0);
}


// cheb.hip (fourth-kind Chebyshev smoothing).  st: device double[4] per
// level = {|v|_D^2, rho, 1 / (rho_scale rho), -}; partials: 2 * kVecMaxBlocks
void chebyshev4_weights(int n, double *w);
// BF16 matrix-core sweeps (mfma_sweep.hip): nterm 1 = K0 scaled by the
// level's moduli, 8 = the level-1 Galerkin operator from the fine moduli
// split bit 0: node values as hi + lo BF16 pairs; bit 1 (nterm 1): the
// element matrix as hi + lo pairs (mf_tables writes both tables)
void mf_tables(const double *M, int nterm, uint16_t *out);
long mf_table_halves(int nterm);
long mf_table_size(int nterm);
int launch_mf_sweep(int mode, int nterm, const Grid &g, const uint16_t *Atab, const float *E, const uint8_t *mask,
                    const float *u, float *y, const float *b, const float *dinv, double omega, const float *rd,
                    hipStream_t s, int split = 0);

int launch_ch_init(long n, double *v, const double *dinv, int fill, int salt, double *partials, double *st,
                   hipStream_t s);
int launch_ch_init_f32(long n, float *v, const float *dinv, int fill, int salt, double *partials, double *st,
                       hipStream_t s);
int launch_ch_pass(long n, double *v, const double *y, const double *dinv, double rho_scale, double *partials,
                   double *st, hipStream_t s);
int launch_ch_pass_f32(long n, float *v, const float *y, const float *dinv, double rho_scale, double *partials,
                       double *st, hipStream_t s);

// ku_node.hip
int launch_ku_node(int mode, const Grid &g, int dim, const double *K0, const double *E, const uint8_t *mask,
                   const double *u, double *y, const double *b, const double *dinv, double omega,
                   double *partials, hipStream_t s);
long ku_node_partials(const Grid &g);
int launch_diag(const Grid &g, int dim, const double *K0, const double *E, const uint8_t *mask, double *d,
                int invert, hipStream_t s);
int launch_dc(const Grid &g, int dim, const double *K0, const double *u, double *dc, hipStream_t s);

// ku_tile.hip (hex8, element-tiled Walsh-Hadamard form)
int launch_ku_tile(int mode, const Grid &g, const double *Kb, const double *E, const uint8_t *mask,
                   const double *u, double *y, const double *b, const double *dinv, double omega,
                   double *partials, hipStream_t s);
int launch_ku_tile_f32(int mode, const Grid &g, const double *Kb, const float *E, const uint8_t *mask,
                       const float *u, float *y, const float *b, const float *dinv, double omega, double *partials,
                       hipStream_t s);
// MODE_JACPRO: y = one Jacobi sweep (D^-1 from the moduli) on u + P ec
int launch_ku_tile_pro(const Grid &g, const double *Kb, const double *E, const uint8_t *mask, const double *u,
                       const double *ec, double *y, const double *b, double omega, hipStream_t s);
int launch_ku_tile_pro_f32(const Grid &g, const double *Kb, const float *E, const uint8_t *mask, const float *u,
                           const float *ec, float *y, const float *b, double omega, hipStream_t s);
long ku_tile_partials(const Grid &g);       // FP64 operator
long ku_tile_partials_f32(const Grid &g);   // FP32 operator

// mg.hip
int launch_restrict(const Grid &gf, int dim, int dof, int normalize, const uint8_t *mask_c, const double *rf,
                    double *rc, hipStream_t s);
int launch_prolong(const Grid &gf, int dim, int dof, int normalize, const uint8_t *mask_f, const double *ec,
                   double *xf, int accumulate, hipStream_t s);
int launch_coarsen_E(const Grid &gf, int dim, const double *Ef, double *Ec, hipStream_t s);
int launch_restrict_f32(const Grid &gf, int dim, int dof, int normalize, const uint8_t *mask_c, const float *rf,
                        float *rc, hipStream_t s);
int launch_prolong_f32(const Grid &gf, int dim, int dof, int normalize, const uint8_t *mask_f, const float *ec,
                       float *xf, int accumulate, hipStream_t s);
int launch_coarsen_E_f32(const Grid &gf, int dim, const float *Ef, float *Ec, hipStream_t s);
int launch_gemv_f32(int n, const float *A, const float *x, float *y, hipStream_t s);
int launch_to_f32(long n, const double *a, float *b, hipStream_t s);
int launch_diag_f32(const Grid &g, const double *K0, const float *E, const uint8_t *mask, float *d, hipStream_t s);
int launch_coarsen_mask(const Grid &gf, int dim, const uint8_t *mf, uint8_t *mc, hipStream_t s);
int launch_dense_assemble(const Grid &g, int dim, const double *K0, const double *E, const uint8_t *mask,
                          double *A, hipStream_t s);
int launch_dense_assemble_f32E(const Grid &g, const double *K0, const float *E, const uint8_t *mask, double *A,
                               hipStream_t s);
int launch_spd_inverse(int n, double *A, double *work, hipStream_t s);
int launch_gemv(int n, const double *A, const double *x, double *y, hipStream_t s);

// galerkin.hip
// cls (optional, host): the 20 class values of a cubic element's table
// (gal1_class_values); the operator then takes them as kernel arguments
int launch_gal1(int mode, const Grid &g, const double *Btab, const double *Ef, const uint8_t *mask, const double *u,
                double *y, const double *b, const double *dinv, double omega, double *partials, hipStream_t s,
                const double *cls = nullptr);
int launch_gal1_f32(int mode, const Grid &g, const float *Btab, const float *Ef, const uint8_t *mask, const float *u,
                    float *y, const float *b, const float *dinv, double omega, double *partials, hipStream_t s,
                    const double *cls = nullptr);
// the class values of a level-1 table B (576 entries) if it has the cubic
// element's structure
bool gal1_class_values(const double *B, double *cls);
long gal1_partials(const Grid &g);
int launch_elem(int mode, const Grid &g, const double *Ke, const uint8_t *mask, const double *u, double *y,
                const double *b, const double *dinv, double omega, double *partials, hipStream_t s);
int launch_elem_f32(int mode, const Grid &g, const double *Ke, const uint8_t *mask, const float *u, float *y,
                    const float *b, const float *dinv, double omega, double *partials, hipStream_t s);
long elem_partials(const Grid &g);
int launch_dinv_gal1(const Grid &g, const double *Gdiag, const double *Ef, const uint8_t *mask, double *d,
                     hipStream_t s);
int launch_dinv_gal1_f32(const Grid &g, const double *Gdiag, const float *Ef, const uint8_t *mask, float *d,
                         hipStream_t s);
int launch_dinv_elem(const Grid &g, const double *Ke, const uint8_t *mask, double *d, hipStream_t s);
int launch_dinv_elem_f32(const Grid &g, const double *Ke, const uint8_t *mask, float *d, hipStream_t s);
int launch_gal_build(const Grid &gc, int src_kind, const double *G, const double *Ef, const double *Kf, double *Kc,
                     hipStream_t s);
int launch_gal_build_f32E(const Grid &gc, int src_kind, const double *G, const float *Ef, const double *Kf,
                          double *Kc, hipStream_t s);
int launch_elem_from_fine(const Grid &gc, const double *G, const double *Ef, double *Ke, hipStream_t s);
int launch_elem_from_fine_f32E(const Grid &gc, const double *G, const float *Ef, double *Ke, hipStream_t s);
int galerkin_tables2(const double *G, double *G2);   // host: 64 x 576 level-2 tables from the 8 child tables
int launch_gal2_build(const Grid &gc, const double *G2, const double *Ef, double *Kc, hipStream_t s);
int launch_gal2_build_f32E(const Grid &gc, const double *G2, const float *Ef, double *Kc, hipStream_t s);
int launch_dense_from_elem(const Grid &g, const double *Ke, const uint8_t *mask, double *A, hipStream_t s);

// stencil.hip (assembled 27-point stencils of Galerkin levels >= 2)
int launch_stencil_build(const Grid &g, const double *Ke, const uint8_t *mask, double *S, hipStream_t s);
int launch_stencil_build_f32(const Grid &g, const double *Ke, const uint8_t *mask, float *S, hipStream_t s);
int launch_stencil_ku(int mode, const Grid &g, const double *S, const double *u, double *y, const double *b,
                      const double *dinv, double omega, double *partials, hipStream_t s);
int launch_stencil_ku_f32(int mode, const Grid &g, const float *S, const float *u, float *y, const float *b,
                          const float *dinv, double omega, double *partials, hipStream_t s);
int launch_dinv_stencil(const Grid &g, const double *S, double *d, hipStream_t s);
int launch_dinv_stencil_f32(const Grid &g, const float *S, float *d, hipStream_t s);
long stencil_partials(const Grid &g);
long stencil_size(const Grid &g);   // coefficients (words) of a stencil

}  // namespace fantom
