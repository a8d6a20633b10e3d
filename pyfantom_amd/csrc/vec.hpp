// Device-scalar slots and vector launchers shared by the C ABI.
#pragma once
#include "common.hpp"

namespace fantom {

// Slots of the per-solve device scalar array (double[FANTOM_SCALARS]).
enum { S_RZ = 0, S_PQ = 1, S_ALPHA = 2, S_RR = 3, S_BETA = 4, S_BB = 5, S_TMP = 6 };
enum { FIN_STORE = 0, FIN_ALPHA = 1, FIN_BETA = 2, FIN_RZ0 = 3 };

constexpr long kVecMaxBlocks = 2048;

unsigned vec_grid(long n);
int launch_dot_partials(long n, const double *a, const double *b, double *partials, hipStream_t s);
int launch_dot2_partials(long n, const double *a, const double *b, const double *c, double *partials, hipStream_t s);
int launch_finalize(const double *partials, long np, double *scal, int op, int slot, hipStream_t s);
int launch_pcg_xr(long n, double *x, double *r, const double *p, const double *q, const double *scal,
                  double *partials, hipStream_t s);
int launch_pcg_p(long n, double *p, const double *z, const double *scal, int first, hipStream_t s);
// x = omega * dinv * b; omega_scale (optional, device): factor on omega
int launch_scale(long n, double omega, const double *dinv, const double *b, double *x, hipStream_t s,
                 const double *omega_scale = nullptr);
int launch_axpby(long n, double a, const double *x, double b, double *y, hipStream_t s);

// filter.hip
int launch_filter_normalizer(int dim, int nelx, int nely, int nelz, int ntap, const int32_t *offsets,
                             const double *weights, double *norm, hipStream_t s);
int launch_filter_fwd(int dim, int nelx, int nely, int nelz, int ntap, const int32_t *offsets,
                      const double *weights, const double *x, double *y, hipStream_t s);
int launch_filter_adj(int dim, int nelx, int nely, int nelz, int ntap, const int32_t *offsets,
                      const double *weights, const double *norm, const double *y, double *x, hipStream_t s);

// oc.hip
int launch_oc_prepare(long n, const double *x, const double *df, const double *dg, double *ocp, double *state,
                      double *partials, hipStream_t s);
int launch_oc_bisect(long n, const double *x, const double *ocp, const double *dg, double move, double vf,
                     double *state, double *partials, int max_passes, hipStream_t s);
long oc_partials(long n);
int launch_oc_apply(long n, const double *x, const double *ocp, double move, const double *state, double *xnew,
                    double *partials, hipStream_t s);

}  // namespace fantom
