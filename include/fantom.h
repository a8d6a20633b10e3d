/*
 * fantom.h -- C ABI of libfantom_hip.so, the gfx950 (MI355X) kernels behind
 * the structured SIMP finite-element solve path.
 *
 * Conventions
 *   - Grids are structured: nelx x nely x nelz elements; nelz == 0 means 2D
 *     (4-node quads, 2 DOF per node), nelz > 0 means 3D (8-node hexes, 3 DOF).
 *   - Layouts are the reference's (pyFANTOM core/CPU/_geom.py:4-98):
 *       3D node  n(i,j,k) = k + nz*i + nz*nx*j       (nx = nelx+1, nz = nelz+1)
 *       3D elem  e(i,j,k) = (i*nely + j)*nelz + k
 *       2D node  n(i,j)   = j*nx + i;  2D elem e(i,j) = i*nely + j
 *       DOF      node*dof + component (interleaved)
 *   - Every pointer except K0/taps metadata is a device pointer owned by the
 *     caller; no call allocates device memory.  Calls are asynchronous on the
 *     given stream (NULL = default stream) unless stated otherwise.
 *   - mask: optional (NULL = no constraints) one byte per NODE, bit c set when
 *     DOF c of that node is Dirichlet-constrained.  Constrained rows act as
 *     identity and constrained columns are skipped, the reference's convention
 *     (core/CUDA/Kernels/matvec.cu:45-98).
 *   - Return value: 0 on success; 1 = invalid argument (ValueError in the
 *     reference's Python API), 2 = HIP runtime error (RuntimeError),
 *     3 = unsupported.  fantom_last_error() returns a thread-local message.
 *
 * The reference launches CuPy RawModule kernels with raw pointers and no
 * status codes (pyFANTOM core/CUDA/_ops.py:59-100, core/CUDA/_mgm.py:173-264,
 * core/CUDA/_filter.py:11-63); each entry point below names the one it
 * replaces.
 */
#ifndef FANTOM_H
#define FANTOM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FANTOM_ABI_VERSION 3

/* Number of double slots of the PCG device-scalar array. */
#define FANTOM_SCALARS 16
/* Slots of that array. */
#define FANTOM_S_RZ 0
#define FANTOM_S_PQ 1
#define FANTOM_S_ALPHA 2
#define FANTOM_S_RR 3
#define FANTOM_S_BETA 4
/* finalize ops */
#define FANTOM_FIN_STORE 0
#define FANTOM_FIN_ALPHA 1
#define FANTOM_FIN_BETA 2
#define FANTOM_FIN_RZ0 3
/* matvec kernels */
#define FANTOM_KU_NODE 0 /* node-centric gather, dense K0 rows (2D and 3D) */
#define FANTOM_KU_TILE 1 /* element-tiled Walsh-Hadamard block form (3D only) */

int fantom_abi_version(void);
const char *fantom_last_error(void);
/* Blocks until all work on the stream is complete. */
int fantom_stream_sync(void *stream);

/* ---- A5: y = K(E) u -------------------------------------------------------
 * Replaces mat_vec_node_basis_parallel(_wcon)_cuda (core/CUDA/_ops.py:59-72,
 * kernel core/CUDA/Kernels/matvec.cu:1-98) as used by
 * StructuredStiffnessKernel._matvec (stiffness/CUDA/_FEA.py:518-549).
 * K0: for FANTOM_KU_NODE a DEVICE pointer to the (npe*dof)^2 row-major
 * element matrix for E = 1; for FANTOM_KU_TILE a HOST pointer to the 72-double
 * Walsh-Hadamard block form from fantom_hex8_blocks (passed by value to the
 * kernel).
 */
int fantom_ku_f64(int kernel, int nelx, int nely, int nelz, const double *K0, const double *E,
                  const uint8_t *mask, const double *u, double *y, void *stream);
/* FP32 y = K(E) u on hex8 grids (the dtype = float32 dispatch of
 * core/CUDA/Kernels/_matvec.py:25-31 and core/CPU/_ops.py:212-245): the
 * tiled Walsh-Hadamard operator in FP32; Kb is the HOST 72-double block form
 * (fantom_hex8_blocks).  2D grids return status 3. */
int fantom_ku_f32(int nelx, int nely, int nelz, const double *Kb, const float *E, const uint8_t *mask,
                  const float *u, float *y, void *stream);
/* Jacobi sweep fused into the operator: x_out = x_in + omega*dinv*(b - K x_in)
 * (solvers/CUDA/_solvers.py:437-440). x_out must not alias x_in.  With the
 * tiled kernel dinv may be NULL: D^-1 is then formed in the sweep from E
 * (diagonal d_c * sum of E over the node's elements, 1 on constrained DOFs),
 * which saves reading a vector. */
int fantom_ku_jacobi_f64(int kernel, int nelx, int nely, int nelz, const double *K0, const double *E,
                         const uint8_t *mask, const double *b, const double *dinv, double omega,
                         const double *x_in, double *x_out, void *stream);
/* The first Jacobi sweep from the zero guess, x_out = x0 + omega D^-1 (b - K x0)
 * with x0 = omega D^-1 b formed while loading (the scaling and the sweep of
 * the reference's smoother, solvers/CUDA/_solvers.py:437-440, in one pass);
 * tiled hex8 kernel only. */
int fantom_ku_jacobi0_f64(int kernel, int nelx, int nely, int nelz, const double *K0, const double *E,
                          const uint8_t *mask, const double *b, const double *dinv, double omega, double *x_out,
                          void *stream);
/* No single counterpart in the reference: its prolongation
 * (apply_prolongation_cuda, core/CUDA/_mgm.py:220, called at
 * solvers/CUDA/_solvers.py:541) and the first post-smoothing sweep
 * (solvers/CUDA/_solvers.py:437-440, called at :545) in one pass, for
 * callers that drive a V-cycle through this ABI.  x_out = v + omega D^-1 (b - K v) with v = x_in + P ec: P is the
 * prolongation of fantom_prolong_f64 with normalize 0 (coarse node c on fine
 * node 2c, constrained fine DOFs receive nothing), D^-1 is formed from E as
 * in fantom_ku_jacobi_f64 with dinv NULL, and v is never stored.  Kb is the
 * HOST block form (fantom_hex8_blocks); hex8 grids only (status 1 for 2D);
 * x_out must not alias x_in or ec.  The sums of P may round differently from
 * fantom_prolong_f64 in the last bit. */
int fantom_ku_jacobi_prolong_f64(int nelx, int nely, int nelz, const double *Kb, const double *E,
                                 const uint8_t *mask, const double *b, double omega, const double *x_in,
                                 const double *ec, double *x_out, void *stream);
int fantom_ku_jacobi_prolong_f32(int nelx, int nely, int nelz, const double *Kb, const float *E,
                                 const uint8_t *mask, const float *b, double omega, const float *x_in,
                                 const float *ec, float *x_out, void *stream);
/* Residual fused into the operator: r = b - K x (solvers/CUDA/_solvers.py:538). */
int fantom_ku_residual_f64(int kernel, int nelx, int nely, int nelz, const double *K0, const double *E,
                           const uint8_t *mask, const double *b, const double *x, double *r, void *stream);
/* y = K u plus per-block partial sums of u.y (PCG's p.Kp,
 * solvers/CUDA/_solvers.py:636-637).  partials needs
 * fantom_ku_partials(kernel, ...) doubles; reduce with fantom_finalize_f64. */
int fantom_ku_dot_f64(int kernel, int nelx, int nely, int nelz, const double *K0, const double *E,
                      const uint8_t *mask, const double *u, double *y, double *partials, void *stream);
long fantom_ku_partials(int kernel, int nelx, int nely, int nelz);
/* Host: derive the Walsh-Hadamard block form (8 blocks of 3x3, 72 doubles)
 * of a box hex8 K0 (24x24, host pointer).  Returns 1 if K0 does not have the
 * box-element reflection symmetry that the form needs. */
int fantom_hex8_blocks(const double *K0_host, double *blocks_host, double *max_offblock);

/* ---- A6: diag(K(E)), constrained -> 1; invert != 0 stores 1/diag ----------
 * Replaces get_diagonal_node_basis_cuda (core/CUDA/_ops.py:430-447,
 * core/CUDA/Kernels/diag.cu:1-47). */
int fantom_diag_f64(int nelx, int nely, int nelz, const double *K0, const double *E, const uint8_t *mask,
                    double *d, int invert, void *stream);

/* ---- A14: dc_e = -u_e^T K0 u_e ---------------------------------------------
 * Replaces process_dk_kernel_cuda (core/CUDA/_ops.py:74-81,
 * core/CUDA/Kernels/grad.cu:1-16). */
int fantom_dc_f64(int nelx, int nely, int nelz, const double *K0, const double *u, double *dc, void *stream);

/* ---- element post-processing of a solution ------------------------------------
 * Replaces FEA_locals_node_basis_parallel (core/CPU/_ops.py:1030-1058,
 * FEA.cu:1-39) and the von Mises formulas of MinimumCompliance.FEA
 * (Problem/CPU/MinimumCompliance.py:659-736) on structured grids: per element
 * strain = B u_e (6 or 3 components), stress = D strain, energy =
 * 1/2 u_e^T (w_e K0) u_e, von Mises of the stress.  K0 (24x24 / 8x8), D (6x6 /
 * 3x3) and B (6x24 / 3x8, the Gauss-point-averaged operator) are device
 * arrays in the reference's local order; w holds the element weights. */
int fantom_fea_locals_f64(int nelx, int nely, int nelz, const double *K0, const double *D, const double *B,
                          const double *w, const double *u, double *strain, double *stress, double *energy,
                          double *von_mises, void *stream);

/* ---- structured mesh arrays (A1 numbering) -------------------------------------
 * Replace generate_elements_2d / generate_elements_3d and the linspace /
 * meshgrid node positions of generate_structured_mesh_cuda
 * (core/CUDA/Kernels/_geom.py:3-118; numbering core/CPU/_geom.py:4-98).
 * elements: int32 [n_el][8] (3D) or [n_el][4] (2D), 16-byte aligned;
 * nodes: [n_nodes][3] or [n_nodes][2] with coordinate i (L / n) per axis and
 * the last one exactly L, in the output type (numpy's linspace of a length of
 * that type).  nelz = 0 selects 2D (lz ignored). */
int fantom_mesh_elements(int nelx, int nely, int nelz, int32_t *elements, void *stream);
int fantom_mesh_nodes_f64(int nelx, int nely, int nelz, double lx, double ly, double lz, double *nodes, void *stream);
int fantom_mesh_nodes_f32(int nelx, int nely, int nelz, double lx, double ly, double lz, float *nodes, void *stream);

/* ---- explicit CSR of K(E) for direct-solver users ------------------------------
 * Replaces StructuredStiffnessKernel.construct (stiffness/CUDA/_FEA.py:484-516)
 * with its two kernels, matmat_node_basis_nnz_per_row_wcon_kernel
 * (core/CUDA/Kernels/_matmat.py:258-300) and
 * matmat_node_basis_parallel_wcon_kernel (core/CUDA/Kernels/matmat.cu:54-126),
 * for the product K(E) @ I.  Dirichlet convention of the matrix-free operator:
 * a constrained DOF's row is the unit row and its column is dropped from the
 * free rows; explicit zeros of K0 stay in the pattern, as in the reference.
 * csr_row_nnz writes the length of every row (n_dof ints); the caller forms
 * indptr[0] = 0, indptr[r + 1] = indptr[r] + counts[r] (n_dof + 1 int64) and
 * sizes indices (int32) and data to indptr[n_dof]; assemble expects exactly
 * that indptr for the same grid and mask.  Columns come out in
 * ascending order within each row (the reference keeps the order of first
 * visit: the same matrix); values are summed in a fixed order. */
int fantom_csr_row_nnz(int nelx, int nely, int nelz, const uint8_t *mask, int *counts, void *stream);
int fantom_csr_assemble_f64(int nelx, int nely, int nelz, const double *K0, const double *E, const uint8_t *mask,
                            const long long *indptr, int *indices, double *data, void *stream);

/* ---- A7: grid transfers (fine grid nelx x nely x nelz) ---------------------
 * Replace apply_restriction_cuda / apply_prolongation_cuda
 * (core/CUDA/_mgm.py:173-264, core/CUDA/Kernels/mgm.cu:1-244).
 * normalize = 1: the reference's TW-normalised operators; 0: plain trilinear
 * P_lin^T / P_lin.  mask_c / mask_f (optional, per node) zero constrained
 * DOFs of the output.  accumulate: xf += P ec instead of xf = P ec. */
int fantom_restrict_f64(int nelx, int nely, int nelz, int normalize, const uint8_t *mask_c, const double *rf,
                        double *rc, void *stream);
int fantom_prolong_f64(int nelx, int nely, int nelz, int normalize, const uint8_t *mask_f, const double *ec,
                       double *xf, int accumulate, void *stream);
/* Rediscretised hierarchy (no reference counterpart: the reference builds
 * Galerkin CSR operators, core/CUDA/_mgm.py:19-171): coarse E = mean of the
 * 2^d children; coarse mask = mask of the coincident fine node. */
int fantom_coarsen_E_f64(int nelx, int nely, int nelz, const double *Ef, double *Ec, void *stream);
int fantom_coarsen_mask(int nelx, int nely, int nelz, const uint8_t *mf, uint8_t *mc, void *stream);

/* ---- A10: coarsest direct solve ---------------------------------------------
 * Replaces the host CHOLMOD factor/solve of solvers/CUDA/_solvers.py:498-516.
 * dense_assemble writes the dense n_dof x n_dof K(E) with the Dirichlet
 * convention; spd_inverse inverts an SPD matrix in place (blocked
 * Gauss-Jordan without pivoting, 2 ceil(n/32) launches; work = n*n + 1024
 * doubles);
 * gemv computes y = A x. */
int fantom_dense_assemble_f64(int nelx, int nely, int nelz, const double *K0, const double *E,
                              const uint8_t *mask, double *A, void *stream);
int fantom_spd_inverse_f64(int n, double *A, double *work, void *stream);
int fantom_gemv_f64(int n, const double *A, const double *x, double *y, void *stream);

/* ---- A11: PCG vector kernels (solvers/CUDA/_solvers.py:629-650) -----------
 * dot_partials: per-block partial sums of a.b into partials
 * (fantom_vec_partials(n) doubles).  finalize reduces np partials in a fixed
 * order and applies op to the device scalars (FANTOM_FIN_*; STORE writes
 * scal[slot]; ALPHA sets alpha = r.z / p.Kp, or 0 when p.Kp is not positive
 * and finite).  pcg_xr: x += alpha p; r -= alpha q; partials of r.r.
 * pcg_p: p = z + beta p (first != 0: p = z).  scale: x = omega*dinv*b.
 * axpby: y = a x + b y. */
long fantom_vec_partials(long n);
int fantom_dot_partials_f64(long n, const double *a, const double *b, double *partials, void *stream);
/* a.b and a.c in one pass: partials[0:nb] and partials[nb:2 nb], nb =
 * fantom_vec_partials(n) (no reference counterpart: the flexible PCG of the
 * learned preconditioner needs r.z_new and r.z_old per iteration). */
int fantom_dot2_partials_f64(long n, const double *a, const double *b, const double *c, double *partials,
                             void *stream);
int fantom_finalize_f64(const double *partials, long np, double *scal, int op, int slot, void *stream);
int fantom_pcg_xr_f64(long n, double *x, double *r, const double *p, const double *q, const double *scal,
                      double *partials, void *stream);
int fantom_pcg_p_f64(long n, double *p, const double *z, const double *scal, int first, void *stream);
int fantom_scale_f64(long n, double omega, const double *dinv, const double *b, double *x, void *stream);
int fantom_axpby_f64(long n, double a, const double *x, double b, double *y, void *stream);

/* ---- A15: structured density filter ----------------------------------------
 * Replace apply_filter_2D/3D_cuda, get_filter_2D/3D_weights_cuda and the
 * transpose kernels (core/CUDA/_filter.py:11-63, core/CUDA/Kernels/filters.cu).
 * offsets: device int32[ntap*dim]; weights: device double[ntap]; ntap <= 2048.
 * Elements in the reference's order (x slowest). */
int fantom_filter_normalizer_f64(int nelx, int nely, int nelz, int ntap, const int32_t *offsets,
                                 const double *weights, double *norm, void *stream);
int fantom_filter_fwd_f64(int nelx, int nely, int nelz, int ntap, const int32_t *offsets, const double *weights,
                          const double *x, double *y, void *stream);
int fantom_filter_adj_f64(int nelx, int nely, int nelz, int ntap, const int32_t *offsets, const double *weights,
                          const double *norm, const double *y, double *x, void *stream);

/* ---- A16: OC update (Optimizers/CUDA/OC.py:67-140) -------------------------
 * state: device double[16] (slot 2 = final Lagrange multiplier); partials:
 * device double[16*1024]; ocp: device double[n].  x_new gets the update and
 * partials[0:fantom_oc_partials(n)] the partial sums of (x - x_new)^2. */
long fantom_oc_partials(long n);
int fantom_oc_update_f64(long n, const double *x, const double *df, const double *dg, double vf, double move,
                         double *ocp, double *state, double *partials, double *x_new, void *stream);
/* Allocation-free OC update.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same kernels and
 * results as fantom_oc_update_f64, with ocp, state and partials held in a
 * library-owned device array (33.7 MB per device, kept until the process
 * ends); n <= 4,194,304, larger n returns status 1.  result: device double[2]
 * = {final Lagrange multiplier, sum (x - x_new)^2}.  Calls that overlap on
 * one device share that array and race: order them on one stream. */
int fantom_oc_update_noalloc_f64(long n, const double *x, const double *df, const double *dg, double vf,
                                 double move, double *x_new, double *result, void *stream);

/* ---- A17: MMA update ----------------------------------------------------------
 * Replaces mmasub/subsolv (mma/CPU/_mma.py:32-165, 288-528) as driven by
 * Optimizers/CPU/MMA.py:103-180: asymptotes, convex approximation and the
 * primal-dual interior-point solve of the subproblem, for 1 <= m <= 8
 * constraints, uniform bounds [xmin, xmax], constants a0, a[m], c[m], d[m]
 * (host arrays), fval[m] on the host.  Device pointers: x, xold1, xold2
 * (NULL allowed while iter <= 2), df0dx (n), dfdx (m x n row-major), low/upp
 * (n, updated in place), x_new (n), work (workspace_size bytes).  iter counts
 * from 1 as the reference's.  Synchronises the stream once per Newton step
 * (scalar read-back); newton_steps (optional) gets the number of steps. */
long fantom_mma_workspace_size(long n, int m);
int fantom_mma_update_f64(long n, int m, int iter, const double *x, const double *xold1, const double *xold2,
                          double xmin, double xmax, const double *df0dx, const double *fval, const double *dfdx,
                          double *low, double *upp, double move, double a0, const double *a, const double *c,
                          const double *d, double sub_tol, int sub_maxiter, double *x_new, void *work,
                          long work_bytes, int *newton_steps, void *stream);

/* ---- A8: Galerkin coarse operators (hex8) ------------------------------------
 * Replace get_restricted_l0/l1p (core/CUDA/_mgm.py:19-171, mgm.cu:246-590):
 * A_{l+1} = P^T A_l P assembled element by element (trilinear P).
 * galerkin_tables (host): from the fine K0 (24x24, reference corner order)
 * the 8 child matrices G (8 x 576, natural corner order b = bx+2by+4bz),
 * their diagonals Gdiag (8 x 24) and the Walsh-Hadamard block table Btab
 * (576) of the level-1 operator; check[2] gets the symmetry residuals.
 * gal1_ku: y = A_1 u on the level-1 grid nelx x nely x nelz from the fine
 * moduli (grid 2x), matrix-free.  gal_build: explicit level-(l+1) element
 * matrices Kc (n_el x 576) from level-1 moduli (src_kind 0) or level-l
 * element matrices (src_kind 1).  elem_ku: y = A u from element matrices.
 * elem_from_fine: explicit level-1 element matrices.  Device pointers except
 * the table outputs of galerkin_tables (host). */
int fantom_galerkin_tables(const double *K0, double *G, double *Btab, double *Gdiag, double *check);
int fantom_gal1_ku_f64(int nelx, int nely, int nelz, const double *Btab, const double *Efine, const uint8_t *mask,
                       const double *u, double *y, void *stream);
/* No counterpart in the reference (its level-1 operator is an assembled CSR
 * product, core/CUDA/_mgm.py:19-171): for cubic elements the 279 nonzero
 * entries of Btab take 20 distinct values, the same classes for every
 * Poisson ratio.  gal1_classes (host arrays): *is_cubic = 1 and cls[20] the
 * class values if Btab has that structure, else 0.  gal1_ku_cls: gal1_ku
 * with the operator's coefficients passed as those 20 values (cls is a HOST
 * array, read at launch) instead of read from a device table. */
int fantom_gal1_classes(const double *Btab, double *cls, int *is_cubic);
int fantom_gal1_ku_cls_f64(int nelx, int nely, int nelz, const double *cls, const double *Efine, const uint8_t *mask,
                           const double *u, double *y, void *stream);
int fantom_gal_build_f64(int nelx, int nely, int nelz, int src_kind, const double *G, const double *Efine,
                         const double *Kfine, double *Kc, void *stream);
int fantom_elem_ku_f64(int nelx, int nely, int nelz, const double *Ke, const uint8_t *mask, const double *u,
                       double *y, void *stream);
int fantom_elem_from_fine_f64(int nelx, int nely, int nelz, const double *G, const double *Efine, double *Ke,
                              void *stream);
/* galerkin_tables2 (host): the 64 level-2 tables G2 (64 x 576, entry
 * 8 c + g = P_c^T G_g P_c for level-1 child c and fine child g) from the 8
 * child tables G; gal2_build: level-2 element matrices Kc (n_el x 576) as
 * sum_f E_f G2[f] over the 64 fine moduli of each element (grid = level 2,
 * the fine grid is 4x in every direction). */
int fantom_galerkin_tables2(const double *G, double *G2);
int fantom_gal2_build_f64(int nelx, int nely, int nelz, const double *G2, const double *Efine, double *Kc,
                          void *stream);

/* Assembled 27-point node stencils of a Galerkin level (hex8), the form the
 * native drivers apply on levels >= 2 (the reference applies its coarse CSR
 * matrices with a generic SpMV, core/CUDA/_mgm.py:19-171).  stencil_build:
 * S (fantom_stencil_size doubles: the operator is symmetric, so only the
 * centre block and the 13 forward neighbour blocks are stored,
 * FANTOM_STENCIL_COEFS = 126 coefficients k = (t - 13)*9 + 3c + d per node,
 * neighbour t = (dy+1)*9 + (dx+1)*3 + (dz+1) >= 13, in tiles of 64 nodes,
 * S[n/64][k][n%64]; the apply reads block t < 13 as the transpose of the
 * neighbour's block 26 - t) from element matrices Ke, with constrained rows
 * set to identity rows and constrained columns zeroed.  stencil_ku: mode 0 y = A u, 1 one Jacobi sweep
 * y = u + omega dinv (b - A u), 2 residual y = b - A u, 3 y = A u with the
 * per-block partial sums of u.(A u) in partials[0:stencil_partials], 4 the
 * first sweep from u = omega dinv b (u unused) (as fantom_ku_f64 / _jacobi /
 * _residual / _dot / _jacobi0; unused operands may be NULL). */
#define FANTOM_STENCIL_COEFS 126
int fantom_stencil_build_f64(int nelx, int nely, int nelz, const double *Ke, const uint8_t *mask, double *S,
                             void *stream);
int fantom_stencil_ku_f64(int mode, int nelx, int nely, int nelz, const double *S, const double *u, double *y,
                          const double *b, const double *dinv, double omega, double *partials, void *stream);
long fantom_stencil_partials(int nelx, int nely, int nelz);
long fantom_stencil_size(int nelx, int nely, int nelz);

/* ---- A9-A11: native GMG-PCG driver ------------------------------------------
 * Replaces MultiGrid._setup / _multi_grid / solve
 * (solvers/CUDA/_solvers.py:442-660): rediscretised matrix-free hierarchy of
 * n_coarse coarsenings (the grid must halve evenly that many times); coarse
 * operators are Galerkin products (galerkin != 0, hex8: level 1 matrix-free
 * from the fine moduli, deeper levels explicit element matrices) or
 * rediscretised from 2^d-averaged moduli (galerkin == 0); Jacobi
 * smoothing (n_smooth sweeps, weight omega), exact dense coarsest solve, PCG
 * with all scalars on the device and each iteration replayed from a HIP
 * graph (use_graph != 0).
 * workspace_size: bytes of device workspace the caller allocates and passes
 * to create (the library allocates no device memory).  K0_host: host dense
 * element matrix for E = 1 on the fine grid.  The handle owns one stream of
 * default priority and joins the caller's stream on entry and exit.
 * setup: copies E and mask (per-node bits, may be NULL) and builds the
 * hierarchy and the coarsest inverse.  precondition: z = V(r).
 * solve: x = K^-1 f to relative recursive residual tol (x is the initial
 * guess when use_x0 != 0); *iters gets the PCG iteration count and *rel_res
 * the true residual ||f - K x|| / ||f||. */
long fantom_mg_workspace_size(int nelx, int nely, int nelz, int n_coarse, int galerkin);
int fantom_mg_create(int nelx, int nely, int nelz, int n_coarse, int n_smooth, double omega, int kernel,
                     int galerkin, const double *K0_host, void *workspace, long workspace_bytes, void **handle);
int fantom_mg_destroy(void *handle);
/* Jacobi weight schedule of the solve (replaces the omega update of
 * MultiGrid.solve, solvers/CUDA/_solvers.py:626-648, and the per-level
 * factor 1.06^level of _multi_grid, :524-526).  adaptive = 0: the fixed
 * weight omega of create on every level (the default).  adaptive = 1: the
 * reference's schedule -- min_omega for the first cycle, then after each
 * iteration's test on R = ||r|| / ||f||
 *   omega = min_omega + (omega - min_omega) / 2 * exp(500 (tol - clip(R, tol, 0.1))),
 * times level_boost^level on each level.  The weight lives in a device
 * scalar, so batched and graph-replayed iterations follow it. */
int fantom_mg_set_omega_schedule(void *handle, int adaptive, double min_omega, double level_boost);
/* Smoother of the V-cycle: kind 0 = the weighted Jacobi sweeps above; kind 1
 * = the fourth-kind Chebyshev smoother of the reference's CPU MultiGrid
 * (_chebyshev_smoother / _estimate_rho, solvers/CPU/_solvers.py:884-954,
 * left commented out there at :1045, :1066, :1088): n_smooth Jacobi sweeps
 * with the weights fantom_chebyshev4_weights(n_smooth) / rho, rho the
 * spectral radius of D^-1 A on each level from power iterations at setup
 * (iters_first at the first setup, iters warm-started after; 0 keeps the
 * previous estimate), multiplied by rho_scale.  Replaces the weight
 * schedule; setting a schedule switches back to Jacobi.  smoother_rho reads
 * the current estimates of the n_coarse smoothed levels. */
int fantom_mg_set_smoother(void *handle, int kind, double rho_scale, int iters_first, int iters);
int fantom_mg_smoother_rho(void *handle, double *rho, int n);
/* Host helper: the n weights 1 / t_i (ascending) of the fourth-kind
 * Chebyshev smoother of degree n, t_i = (1 - cos(2 i pi / (2n + 1))) / 2. */
int fantom_chebyshev4_weights(int n, double *w);
int fantom_mg_setup(void *handle, const double *E, const uint8_t *mask, void *stream);
int fantom_mg_precondition(void *handle, const double *r, double *z, void *stream);
int fantom_mg_solve(void *handle, const double *f, double *x, double tol, int maxiter, int use_x0, int use_graph,
                    int *iters, double *rel_res, void *stream);

/* ---- A19/A21: mixed-precision solve (no reference counterpart) --------------
 * FP64 iterative refinement around an FP32 GMG-PCG on a hex8 grid: the true
 * residual f - K x and the update x += ||r|| d stay FP64; the inner PCG on
 * K d = r/||r|| (from d = 0, to relative residual inner_tol) stores vectors,
 * moduli and transfers in FP32 and runs the Walsh-Hadamard element
 * contraction in FP32, with FP64 dot products, step lengths and coarsest
 * solve.  *outer_iters gets the refinement steps, *inner_iters the total
 * inner PCG iterations, *rel_res the final FP64 ||f - K x||/||f||.
 * Workspace / handle / stream rules as for the fantom_mg_* driver. */
long fantom_mx_workspace_size(int nelx, int nely, int nelz, int n_coarse, int galerkin);
int fantom_mx_create(int nelx, int nely, int nelz, int n_coarse, int n_smooth, double omega, int galerkin,
                     const double *K0_host, void *workspace, long workspace_bytes, void **handle);
int fantom_mx_destroy(void *handle);
/* As fantom_mg_set_omega_schedule, for the FP32 V-cycle; the adaptive
 * schedule follows the residual of the PCG that the V-cycle preconditions
 * (the FP64 PCG of fantom_mx_solve_pcg, the inner FP32 PCG of
 * fantom_mx_solve, restarted at min_omega on each refinement step). */
int fantom_mx_set_omega_schedule(void *handle, int adaptive, double min_omega, double level_boost);
/* As fantom_mg_set_smoother / fantom_mg_smoother_rho, for the FP32 V-cycle
 * (power iterations on FP32 vectors with FP64 sums). */
int fantom_mx_set_smoother(void *handle, int kind, double rho_scale, int iters_first, int iters);
int fantom_mx_smoother_rho(void *handle, double *rho, int n);
/* Operator of the V-cycle's fine and level-1 sweeps (replaces, inside the
 * preconditioner, the element contraction of core/CUDA/Kernels/matvec.cu:1-98
 * and the Galerkin product of core/CPU/_mgm.py:308-942): kind 0 = FP32 in
 * the Walsh-Hadamard form (default); bit 0 / bit 1: the fine / level-1
 * sweeps with BF16 operands (node values and the element / child matrices)
 * and FP32 accumulation on the matrix cores (v_mfma_f32_16x16x16_bf16),
 * moduli and vectors FP32; bits 2-3: the split forms of fantom_sweep_bf16.
 * The outer PCG, the coarser levels and the coarsest solve are unchanged. */
int fantom_mx_set_operator(void *handle, int kind);
/* Flexible PCG in fantom_mx_solve_pcg (on != 0): beta = r_new.(z_new -
 * z_old) / r_old.z_old (Polak-Ribiere) instead of the reference's
 * r_new.z_new / r_old.z_old (solvers/CUDA/_solvers.py:646-652), which keeps
 * the PCG converging when the preconditioner is not an exactly fixed linear
 * operator (rounding to BF16 in the V-cycle); one more FP32 vector. */
int fantom_mx_set_flexible(void *handle, int on);
/* The BF16 matrix-core sweep of fantom_mx_set_operator kind 1, on its own
 * (same element contraction as core/CUDA/Kernels/matvec.cu:1-98, BF16
 * operands, FP32 sums and vectors).  sweep_bf16_tables (host): the operand
 * table (table_size uint16 values) of nterm 1 = K0 (24 x 24, reference corner
 * order) or nterm 8 = the 8 level-1 child matrices G of galerkin_tables.
 * sweep_bf16 on the grid nelx x nely x nelz (the level-1 grid for nterm 8,
 * whose moduli E are then the fine ones, grid 2x): mode 0 y = A u, 1 one
 * Jacobi sweep y = u + omega D^-1 (b - A u), 2 residual y = b - A u, 4 two
 * sweeps from x0 = omega dinv b (u unused).  nterm 1 forms D^-1 from the
 * moduli when rd (host float[3], 1 / K0 diagonal per component) is given
 * (mode 1 with dinv NULL, and the second sweep of mode 4); Dirichlet nodes
 * (mask bit c) keep identity rows.  split bit 0: the node values enter as
 * hi + lo BF16 pairs (x - bf16(x) rounded again, two products), bit 1
 * (nterm 1): the element matrix likewise; the table holds both parts.
 * Device pointers except tab's source and rd. */
long fantom_sweep_bf16_table_size(int nterm);
int fantom_sweep_bf16_tables(int nterm, const double *M, uint16_t *out);
int fantom_sweep_bf16(int mode, int nterm, int split, int nelx, int nely, int nelz, const uint16_t *tab, const float *E,
                      const uint8_t *mask, const float *u, float *y, const float *b, const float *dinv, double omega,
                      const float *rd, void *stream);
int fantom_mx_setup(void *handle, const double *E, const uint8_t *mask, void *stream);
/* z = V(r) with the FP32 V-cycle: r (FP64) is rounded to FP32, the cycle runs
 * as in fantom_mx_solve_pcg, z is returned in FP64 (a preconditioner for a
 * caller's own Krylov loop, e.g. flexible PCG). */
int fantom_mx_precondition(void *handle, const double *r, double *z, void *stream);
int fantom_mx_solve(void *handle, const double *f, double *x, double tol, double inner_tol, int max_outer,
                    int max_inner, int use_x0, int use_graph, int *outer_iters, int *inner_iters, double *rel_res,
                    void *stream);
/* FP64 PCG preconditioned by the handle's FP32 V-cycle (one per iteration):
 * the fine K.u, the vectors, dot products and step lengths stay FP64, the
 * V-cycle runs on FP32 copies.  Stops when the recursive relative residual is
 * below tol, or early when it has not dropped by 10% over 20 iterations
 * (stagnation: the caller sees rel_res >= tol and falls back to the FP64
 * solver); rel_res gets the true FP64 relative residual ||f - K x|| / ||f||.
 * No counterpart in the reference (its solver is FP64 throughout,
 * solvers/CUDA/_solvers.py:573-660). */
int fantom_mx_solve_pcg(void *handle, const double *f, double *x, double tol, int maxiter, int use_x0,
                        int use_graph, int *iters, double *rel_res, void *stream);

/* ---- A20: learned preconditioner kernels (no reference counterpart) ----------
 * Channels-last BF16 volumes (D, H, W, C), C a multiple of 16; weights
 * W[tap][c_out][c_in] BF16, tap = (dy+1)*9 + (dx+1)*3 + (dz+1); FP32
 * accumulation on MFMA.  conv3d_fwd: out = act(conv(x, W) + bias + skip)
 * (bias, skip, out_f32 optional; act = ReLU when relu != 0; Cout in
 * {16, 32, 64}).  wgrad: dW = sum_v dy[v] x[v + off]^T through
 * wgrad_parts() partial slices of 27*Cout*Cin floats (part), reduced in a
 * fixed order into dw.  bias_grad: db = sum_v dy[v] (part: 256*C floats).
 * flip_transpose: W'[t][ci][co] = W[26-t][co][ci] (input adjoint weights).
 * relu_mask: g = dy where y > 0.  add: a += b.  vol_restrict / vol_prolong:
 * vertex-centred full weighting / trilinear interpolation (odd sizes),
 * times scale.  adam: Adam step on FP32 master weights with BF16 copy.
 * net_input / net_output / mse: map a nodal vector (3 DOF per node) to the
 * 16-channel network input normalised by the RMS stored in scal[slot] (a
 * sum of squares), back, and the MSE loss gradient against a target.
 * node_density: nodal mean of an element field (hex8). */
int fantom_conv3d_fwd_bf16(int D, int H, int W, int Cin, int Cout, const uint16_t *x, const uint16_t *w,
                           const float *bias, const uint16_t *skip, uint16_t *out, float *out_f32, int relu,
                           void *stream);
long fantom_conv3d_wgrad_parts(int D, int H, int W, int Cin, int Cout, long *chunks_per_wave);
int fantom_conv3d_wgrad_bf16(int D, int H, int W, int Cin, int Cout, const uint16_t *x, const uint16_t *dy,
                             float *part, float *dw, void *stream);
int fantom_bias_grad_bf16(long N, int C, const uint16_t *dy, float *part, float *db, void *stream);
int fantom_flip_transpose_bf16(int Cout, int Cin, const uint16_t *w, uint16_t *wt, void *stream);
int fantom_relu_mask_bf16(long n, const uint16_t *y, const uint16_t *dy, uint16_t *g, void *stream);
int fantom_add_bf16(long n, uint16_t *a, const uint16_t *b, void *stream);
int fantom_vol_restrict_bf16(int D, int H, int W, int C, const uint16_t *x, uint16_t *out, float scale, void *stream);
int fantom_vol_prolong_bf16(int D, int H, int W, int C, const uint16_t *xc, uint16_t *out, float scale, int accumulate,
                            void *stream);
int fantom_adam_f32(long n, float *p, const float *g, float *m, float *v, uint16_t *pbf, float lr, float b1, float b2,
                    float eps, int step, void *stream);
int fantom_f32_to_bf16(long n, const float *a, uint16_t *b, void *stream);
int fantom_net_input(long nv, const double *r, const float *rho_node, const double *scal, int slot, uint16_t *x,
                     void *stream);
int fantom_net_output(long nv, const float *y, const double *scal, int slot, double *z, int accumulate, void *stream);
int fantom_mse_bf16(long nv, const float *y, const double *t, const double *scal, int slot, uint16_t *dy,
                    double *part, void *stream);
int fantom_node_density(int nelx, int nely, int nelz, const double *x, float *rn, void *stream);

/* ---- A4: SIMP interpolation ------------------------------------------------
 * Replaces MinimumCompliance.penalize / penalize_grad
 * (Problem/CUDA/MinimumCompliance.py:305-351, CPU :358-464), single material.
 * law 0: E = max(clip(h^p, lo, 1)*hi, lo) with lo = void, hi = E_mul (the
 * reference's law); law 1: E = lo + h^p (hi - lo) with lo = Emin, hi = E0.
 * h is the Heaviside projection of rho when heaviside != 0, else rho.
 * E or dE may be NULL.  mul: out = a*b elementwise (the chain rule). */
int fantom_simp_f64(long n, const double *rho, int law, double p, double lo, double hi, int heaviside,
                    double beta, double eta, double *E, double *dE, void *stream);
int fantom_mul_f64(long n, const double *a, const double *b, double *out, void *stream);

/* ---- Unstructured meshes: node-basis operators ----------------------------
 * Replace the "full" and "flat" node-basis kernels of UniformStiffnessKernel /
 * GeneralStiffnessKernel (core/CUDA/Kernels/matvec.cu:145-300 with their
 * _wcon variants, diag.cu:98-147, grad.cu:18-52; CPU forms
 * core/CPU/_ops.py:16-330; classes stiffness/CUDA/_FEA.py:675-1630).
 * Elements are grouped by node count (at most 8 groups); a group's arrays are
 * on the device, structure-of-arrays so that consecutive threads read
 * consecutive addresses: node j of element l at nodes[j n_el + l], entry
 * (i, k) of its m x m matrix (m = npe dof, DOF order (node, component)) at
 * K[(i m + k) n_el + l]; el_ids[l] is the element's global id (the index into
 * E and dk).  The node-sorted incidence list gives node n's pairs for j in
 * [node_ptr[n], node_ptr[n+1]): inc_l[j] the element's index in its group,
 * inc_ga[j] = group * 256 + local node index; each row is summed in that
 * order (deterministic, no atomics).  con (n_nodes dof bytes, may be NULL):
 * Dirichlet DOFs keep identity rows and are skipped as columns.
 * ku: mode 0 y = K(E) u, 2 y = b - K(E) u, 3 y = K(E) u and the per-block
 * partial sums of u.y in partials[0:ku_general_partials]; Y is a device
 * workspace of ku_general_workspace doubles (per-element products E_e K_e u_e,
 * laid out like K).
 * diag: diag K(E), 1 on Dirichlet DOFs.  dc: dk[el_ids[l]] = -u_e^T K_e u_e.
 * Limits: 1 <= dof <= 3, npe dof <= 96. */
typedef struct fantom_elem_group {
    int n_el;          /* elements in the group */
    int npe;           /* nodes per element */
    const int *el_ids; /* [n_el] global element ids */
    const int *nodes;  /* [npe][n_el] node ids */
    const double *K;   /* [m m][n_el] element matrices, m = npe dof */
} fantom_elem_group;

long fantom_ku_general_partials(int n_nodes, int dof);
long fantom_ku_general_workspace(int dof, int n_groups, const fantom_elem_group *groups);
int fantom_ku_general_f64(int mode, int n_nodes, int dof, int n_groups, const fantom_elem_group *groups,
                          const int *node_ptr, const int *inc_l, const int *inc_ga, const double *E,
                          const uint8_t *con, const double *u, double *y, const double *b, double *Y,
                          double *partials, void *stream);
int fantom_diag_general_f64(int n_nodes, int dof, int n_groups, const fantom_elem_group *groups,
                            const int *node_ptr, const int *inc_l, const int *inc_ga, const double *E,
                            const uint8_t *con, double *diag, void *stream);
int fantom_dc_general_f64(int dof, int n_groups, const fantom_elem_group *groups, const double *u, double *dk,
                          void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FANTOM_H */
