// Unstructured node-basis operators: K(E) u, diag K(E) and the compliance
// sensitivity for meshes of any element type (triangles, quads, tetrahedra,
// hexes, mixed), with per-element stiffness matrices.
//
// Reference: the "full" (one element size) and "flat" (mixed sizes) node-basis
// kernels core/CUDA/Kernels/matvec.cu:145-300 (with their _wcon Dirichlet
// variants), diag.cu:98-147 and grad.cu:18-52.  The reference gathers, per
// DOF row, the rows K_e[(a dof + k), :] of the node's incident elements;
// every thread walks a different element matrix, so its loads do not
// coalesce.  Here the elements are grouped by size and each group's matrices
// are stored structure-of-arrays (entry (i, j) of element l at
// K[(i m + j) n_el + l]); the product is two passes:
//   1. one thread per element (per slice of its rows): Y_e = E_e (K_e u_e),
//      K read once, coalesced across the wave, u_e gathered with Dirichlet
//      columns zeroed;
//   2. one thread per DOF row: y_r = sum over the node's incident (element,
//      local node) pairs of Y_e[a dof + k], in a fixed order (the elements in
//      ascending id), Dirichlet rows y_r = u_r.
// Per row that is the reference's arithmetic: val = sum_l K u_l in column
// order, then acc += val * E_e element by element; deterministic, no atomics.
#include <cmath>

#include "../../include/fantom.h"
#include "common.hpp"
#include "fantom_internal.hpp"

namespace fantom {

constexpr int kKgBlock = 256;
constexpr int kKgMaxGroups = 8;
constexpr int kKgMaxM = 96;   // element matrix order limit (32 nodes x 3 DOF)

struct KgGroups {
    int n, dof;
    int n_el[kKgMaxGroups], m[kKgMaxGroups];
    long yoff[kKgMaxGroups];
    const double *K[kKgMaxGroups];
    const int *el_ids[kKgMaxGroups];
    const int *nodes[kKgMaxGroups];
};

// pass 1, element matrix order M known at compile time: u_e in registers.
// blockIdx.y selects a slice of M / PARTS rows, so that large elements still
// give the chip enough threads.  Y has the layout of K (row i of element l at
// Y[i n_el + l]): the stores coalesce like the loads.  K is read once per
// product, so its loads are nontemporal and leave u and Y in the caches.
template <int NPE, int DOF, int PARTS>
__global__ __launch_bounds__(kKgBlock) void kg_elem_k(int n_el, const int *__restrict__ el_ids,
                                                    const int *__restrict__ nodes, const double *__restrict__ K,
                                                    const double *__restrict__ E, const uint8_t *__restrict__ con,
                                                    const double *__restrict__ u, double *__restrict__ Y)
{
    constexpr int M = NPE * DOF;
    constexpr int R = M / PARTS;
    static_assert(M % PARTS == 0, "rows must split evenly");
    const int l = blockIdx.x * kKgBlock + threadIdx.x;
    if (l >= n_el)
        return;
    double ue[M];
#pragma unroll
    for (int a = 0; a < NPE; ++a) {
        const long g = (long)nodes[(long)a * n_el + l] * DOF;
#pragma unroll
        for (int c = 0; c < DOF; ++c)
            ue[a * DOF + c] = (con && con[g + c]) ? 0.0 : u[g + c];
    }
    const double Ee = E[el_ids[l]];
    const int i0 = blockIdx.y * R;
    const double *Kl = K + (long)i0 * M * n_el + l;
    double *Yl = Y + (long)i0 * n_el + l;
#pragma unroll
    for (int i = 0; i < R; ++i) {
        double val = 0.0;
#pragma unroll
        for (int j = 0; j < M; ++j)
            val += __builtin_nontemporal_load(Kl + (long)(i * M + j) * n_el) * ue[j];
        Yl[(long)i * n_el] = val * Ee;
    }
}

// pass 1, any element size up to kKgMaxM: u gathered on the fly
__global__ __launch_bounds__(kKgBlock) void kg_elem_any_k(int n_el, int npe, int dof, const int *__restrict__ el_ids,
                                                        const int *__restrict__ nodes, const double *__restrict__ K,
                                                        const double *__restrict__ E, const uint8_t *__restrict__ con,
                                                        const double *__restrict__ u, double *__restrict__ Y)
{
    const int l = blockIdx.x * kKgBlock + threadIdx.x;
    if (l >= n_el)
        return;
    const int M = npe * dof;
    const double Ee = E[el_ids[l]];
    for (int i = 0; i < M; ++i) {
        double val = 0.0;
        for (int b = 0; b < npe; ++b) {
            const long g = (long)nodes[(long)b * n_el + l] * dof;
            for (int c = 0; c < dof; ++c) {
                if (con && con[g + c])
                    continue;
                val += K[(long)(i * M + b * dof + c) * n_el + l] * u[g + c];
            }
        }
        Y[(long)i * n_el + l] = val * Ee;
    }
}

// pass 2.  MODE_KU: y = K u; MODE_RES: y = b - K u; MODE_KUDOT: y = K u and
// the block partial sums of u . y
template <int MODE>
__global__ __launch_bounds__(kKgBlock) void kg_gather_k(int n_nodes, KgGroups G, const int *__restrict__ node_ptr,
                                                      const int *__restrict__ inc_l, const int *__restrict__ inc_ga,
                                                      const uint8_t *__restrict__ con, const double *__restrict__ u,
                                                      const double *__restrict__ Y, double *__restrict__ y,
                                                      const double *__restrict__ b, double *__restrict__ partials)
{
    // The hardware deals workgroups to the 8 XCDs round-robin; give each XCD
    // a contiguous range of nodes so that the element outputs shared by
    // neighbouring nodes are fetched into one L2, not into eight.  The grid
    // is 8 * per_xcd blocks; blocks past the last node exit whole.
    const int per_xcd = gridDim.x / 8;
    const int blk = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    const int dof = G.dof;
    const long n_rows = (long)n_nodes * dof;
    if ((long)blk * kKgBlock >= n_rows)
        return;
    const long r = (long)blk * kKgBlock + threadIdx.x;
    double part = 0.0;
    if (r < n_rows) {
        double kuv;
        if (con && con[r]) {
            kuv = u[r];
        } else {
            const int n = (int)(r / dof), k = (int)(r % dof);
            kuv = 0.0;
            // unrolled so that several incidences' loads are in flight at once
#pragma unroll 4
            for (int j = node_ptr[n]; j < node_ptr[n + 1]; ++j) {
                const int ga = inc_ga[j], g = ga >> 8, a = ga & 255;
                kuv += Y[G.yoff[g] + (long)(a * dof + k) * G.n_el[g] + inc_l[j]];
            }
        }
        const double o = MODE == MODE_RES ? b[r] - kuv : kuv;
        y[r] = o;
        if (MODE == MODE_KUDOT)
            part = u[r] * o;
    }
    if (MODE == MODE_KUDOT) {
        part = block_sum<kKgBlock>(part);
        if (threadIdx.x == 0)
            partials[blk] = part;
    }
}

// diag K(E); Dirichlet DOFs 1
__global__ __launch_bounds__(kKgBlock) void kg_diag_k(int n_nodes, KgGroups G, const int *__restrict__ node_ptr,
                                                    const int *__restrict__ inc_l, const int *__restrict__ inc_ga,
                                                    const double *__restrict__ E, const uint8_t *__restrict__ con,
                                                    double *__restrict__ diag)
{
    const int dof = G.dof;
    const long r = (long)blockIdx.x * kKgBlock + threadIdx.x;
    if (r >= (long)n_nodes * dof)
        return;
    if (con && con[r]) {
        diag[r] = 1.0;
        return;
    }
    const int n = (int)(r / dof), k = (int)(r % dof);
    double acc = 0.0;
    for (int j = node_ptr[n]; j < node_ptr[n + 1]; ++j) {
        const int ga = inc_ga[j], g = ga >> 8, a = ga & 255, l = inc_l[j];
        const int q = a * dof + k, m = G.m[g];
        acc += G.K[g][(long)(q * m + q) * G.n_el[g] + l] * E[G.el_ids[g][l]];
    }
    diag[r] = acc;
}

// dk_e = -u_e^T K_e u_e (no modulus, no Dirichlet rule: grad.cu:18-52)
template <int NPE, int DOF>
__global__ __launch_bounds__(kKgBlock) void kg_dk_k(int n_el, const int *__restrict__ el_ids,
                                                  const int *__restrict__ nodes, const double *__restrict__ K,
                                                  const double *__restrict__ u, double *__restrict__ dk)
{
    constexpr int M = NPE * DOF;
    const int l = blockIdx.x * kKgBlock + threadIdx.x;
    if (l >= n_el)
        return;
    double ue[M];
#pragma unroll
    for (int a = 0; a < NPE; ++a) {
        const long g = (long)nodes[(long)a * n_el + l] * DOF;
#pragma unroll
        for (int c = 0; c < DOF; ++c)
            ue[a * DOF + c] = u[g + c];
    }
    const double *Kl = K + l;
    double tot = 0.0;
#pragma unroll
    for (int i = 0; i < M; ++i) {
        double row = 0.0;
#pragma unroll
        for (int j = 0; j < M; ++j)
            row += __builtin_nontemporal_load(Kl + (long)(i * M + j) * n_el) * ue[j];
        tot += ue[i] * row;
    }
    dk[el_ids[l]] = -tot;
}

__global__ __launch_bounds__(kKgBlock) void kg_dk_any_k(int n_el, int npe, int dof, const int *__restrict__ el_ids,
                                                      const int *__restrict__ nodes, const double *__restrict__ K,
                                                      const double *__restrict__ u, double *__restrict__ dk)
{
    const int l = blockIdx.x * kKgBlock + threadIdx.x;
    if (l >= n_el)
        return;
    const int M = npe * dof;
    double tot = 0.0;
    for (int i = 0; i < M; ++i) {
        double row = 0.0;
        for (int j = 0; j < M; ++j)
            row += K[(long)(i * M + j) * n_el + l] * u[(long)nodes[(long)(j / dof) * n_el + l] * dof + j % dof];
        tot += u[(long)nodes[(long)(i / dof) * n_el + l] * dof + i % dof] * row;
    }
    dk[el_ids[l]] = -tot;
}

static unsigned kg_blocks(long n) { return (unsigned)((n + kKgBlock - 1) / kKgBlock); }

// fills the kernel-argument table; returns the element-output size (doubles) or -1
static long kg_table(int dof, int n_groups, const fantom_elem_group *groups, KgGroups &G)
{
    if (dof < 1 || dof > 3 || n_groups < 1 || n_groups > kKgMaxGroups || !groups)
        return -1;
    G.n = n_groups;
    G.dof = dof;
    long off = 0;
    for (int g = 0; g < kKgMaxGroups; ++g) {
        G.n_el[g] = 0;
        G.m[g] = 0;
        G.yoff[g] = 0;
        G.K[g] = nullptr;
        G.el_ids[g] = nullptr;
        G.nodes[g] = nullptr;
    }
    for (int g = 0; g < n_groups; ++g) {
        const fantom_elem_group &q = groups[g];
        if (q.n_el < 0 || q.npe < 1 || q.npe > 255 || q.npe * dof > kKgMaxM)
            return -1;
        if (q.n_el > 0 && (!q.K || !q.el_ids || !q.nodes))
            return -1;
        G.n_el[g] = q.n_el;
        G.m[g] = q.npe * dof;
        G.yoff[g] = off;
        G.K[g] = q.K;
        G.el_ids[g] = q.el_ids;
        G.nodes[g] = q.nodes;
        off += (long)q.n_el * q.npe * dof;
    }
    return off;
}

template <int NPE, int DOF, int PARTS>
static void kg_launch_elem(const fantom_elem_group &q, const double *E, const uint8_t *con, const double *u,
                           double *Yg, hipStream_t s)
{
    hipLaunchKernelGGL((kg_elem_k<NPE, DOF, PARTS>), dim3(kg_blocks(q.n_el), PARTS), dim3(kKgBlock), 0, s, q.n_el,
                       q.el_ids, q.nodes, q.K, E, con, u, Yg);
}

template <int NPE, int DOF>
static void kg_launch_dk(const fantom_elem_group &q, const double *u, double *dk, hipStream_t s)
{
    hipLaunchKernelGGL((kg_dk_k<NPE, DOF>), dim3(kg_blocks(q.n_el)), dim3(kKgBlock), 0, s, q.n_el, q.el_ids,
                       q.nodes, q.K, u, dk);
}

}  // namespace fantom

using namespace fantom;

extern "C" long fantom_ku_general_partials(int n_nodes, int dof)
{
    if (n_nodes <= 0 || dof < 1 || dof > 3)
        return -1;
    return kg_blocks((long)n_nodes * dof);
}

extern "C" long fantom_ku_general_workspace(int dof, int n_groups, const fantom_elem_group *groups)
{
    KgGroups G;
    return kg_table(dof, n_groups, groups, G);
}

extern "C" int fantom_ku_general_f64(int mode, int n_nodes, int dof, int n_groups, const fantom_elem_group *groups,
                                     const int *node_ptr, const int *inc_l, const int *inc_ga, const double *E,
                                     const uint8_t *con, const double *u, double *y, const double *b, double *Y,
                                     double *partials, void *stream)
{
    KgGroups G;
    FANTOM_CHECK_ARG(n_nodes > 0, "need n_nodes > 0");
    FANTOM_CHECK_ARG(kg_table(dof, n_groups, groups, G) >= 0,
                     "need 1 <= dof <= 3, 1..8 element groups, npe*dof <= 96 and non-null group arrays");
    FANTOM_CHECK_ARG(node_ptr && inc_l && inc_ga && E && u && y && Y, "null operand");
    FANTOM_CHECK_ARG(mode == MODE_KU || mode == MODE_RES || mode == MODE_KUDOT,
                     "mode must be 0 (K u), 2 (residual) or 3 (K u with u.Ku partials)");
    FANTOM_CHECK_ARG(mode != MODE_RES || b, "mode 2 needs b");
    FANTOM_CHECK_ARG(mode != MODE_KUDOT || partials, "mode 3 needs partials");
    hipStream_t s = as_stream(stream);
    for (int g = 0; g < n_groups; ++g) {
        const fantom_elem_group &q = groups[g];
        if (q.n_el == 0)
            continue;
        double *Yg = Y + G.yoff[g];
        if (dof == 2 && q.npe == 3)
            kg_launch_elem<3, 2, 1>(q, E, con, u, Yg, s);
        else if (dof == 2 && q.npe == 4)
            kg_launch_elem<4, 2, 1>(q, E, con, u, Yg, s);
        else if (dof == 3 && q.npe == 4)
            kg_launch_elem<4, 3, 1>(q, E, con, u, Yg, s);
        else if (dof == 3 && q.npe == 8)
            kg_launch_elem<8, 3, 2>(q, E, con, u, Yg, s);
        else
            hipLaunchKernelGGL(kg_elem_any_k, dim3(kg_blocks(q.n_el)), dim3(kKgBlock), 0, s, q.n_el, q.npe, dof,
                               q.el_ids, q.nodes, q.K, E, con, u, Yg);
    }
    const dim3 grid(8 * ((kg_blocks((long)n_nodes * dof) + 7) / 8));
    if (mode == MODE_KU)
        hipLaunchKernelGGL(kg_gather_k<MODE_KU>, grid, dim3(kKgBlock), 0, s, n_nodes, G, node_ptr, inc_l, inc_ga, con,
                           u, Y, y, b, partials);
    else if (mode == MODE_RES)
        hipLaunchKernelGGL(kg_gather_k<MODE_RES>, grid, dim3(kKgBlock), 0, s, n_nodes, G, node_ptr, inc_l, inc_ga,
                           con, u, Y, y, b, partials);
    else
        hipLaunchKernelGGL(kg_gather_k<MODE_KUDOT>, grid, dim3(kKgBlock), 0, s, n_nodes, G, node_ptr, inc_l, inc_ga,
                           con, u, Y, y, b, partials);
    FANTOM_LAUNCH_CHECK("ku_general");
    return kStatusOk;
}

extern "C" int fantom_diag_general_f64(int n_nodes, int dof, int n_groups, const fantom_elem_group *groups,
                                       const int *node_ptr, const int *inc_l, const int *inc_ga, const double *E,
                                       const uint8_t *con, double *diag, void *stream)
{
    KgGroups G;
    FANTOM_CHECK_ARG(n_nodes > 0, "need n_nodes > 0");
    FANTOM_CHECK_ARG(kg_table(dof, n_groups, groups, G) >= 0,
                     "need 1 <= dof <= 3, 1..8 element groups, npe*dof <= 96 and non-null group arrays");
    FANTOM_CHECK_ARG(node_ptr && inc_l && inc_ga && E && diag, "null operand");
    hipLaunchKernelGGL(kg_diag_k, dim3(kg_blocks((long)n_nodes * dof)), dim3(kKgBlock), 0, as_stream(stream), n_nodes,
                       G, node_ptr, inc_l, inc_ga, E, con, diag);
    FANTOM_LAUNCH_CHECK("diag_general");
    return kStatusOk;
}

extern "C" int fantom_dc_general_f64(int dof, int n_groups, const fantom_elem_group *groups, const double *u,
                                     double *dk, void *stream)
{
    KgGroups G;
    FANTOM_CHECK_ARG(kg_table(dof, n_groups, groups, G) >= 0,
                     "need 1 <= dof <= 3, 1..8 element groups, npe*dof <= 96 and non-null group arrays");
    FANTOM_CHECK_ARG(u && dk, "null operand");
    hipStream_t s = as_stream(stream);
    for (int g = 0; g < n_groups; ++g) {
        const fantom_elem_group &q = groups[g];
        if (q.n_el == 0)
            continue;
        if (dof == 2 && q.npe == 3)
            kg_launch_dk<3, 2>(q, u, dk, s);
        else if (dof == 2 && q.npe == 4)
            kg_launch_dk<4, 2>(q, u, dk, s);
        else if (dof == 3 && q.npe == 4)
            kg_launch_dk<4, 3>(q, u, dk, s);
        else if (dof == 3 && q.npe == 8)
            kg_launch_dk<8, 3>(q, u, dk, s);
        else
            hipLaunchKernelGGL(kg_dk_any_k, dim3(kg_blocks(q.n_el)), dim3(kKgBlock), 0, s, q.n_el, q.npe, dof,
                               q.el_ids, q.nodes, q.K, u, dk);
    }
    FANTOM_LAUNCH_CHECK("dc_general");
    return kStatusOk;
}
