// Galerkin coarse operators for the hex8 multigrid hierarchy.
//
// The reference forms A_{l+1} = R A_l R^T as padded CSR on every solve
// (core/CUDA/_mgm.py:19-171, mgm.cu:246-590).  Here the same product is
// formed element by element: with trilinear prolongation P (coarse node c on
// fine node 2c), every fine element lies in one coarse element E, so
//   A_{l+1} = sum_E P_E^T (sum_{children c of E} K_c) P_E,
// and coarse element matrices assemble the coarse operator exactly.  (The
// reference's R = D^-1 P^T differs from P^T by a diagonal scaling that
// cancels in the V-cycle.)  Constrained coarse DOFs are identity rows and
// skipped columns, the coincident-node rule.
//
// Level 1 is applied without storing matrices.  With child moduli E_c and the
// constant child matrices G_c = P_c^T K0 P_c, K_E = sum_c E_c G_c.  The
// reflections of the coarse element permute the children and commute with
// K0, so G_{c^g} = rho(g) G_c rho(g)^T, and in the Walsh-Hadamard basis
// (v_{s,d}[corner b, comp d] = (-1)^{popc(b & (s ^ e_d))})
//   Khat_E[s, s'] = Ehat_{s ^ s'} * Ghat_0[s, s']     (3x3 blocks)
// where Ehat_t = sum_c (-1)^{popc(c & t)} E_c.  One element costs a WHT of u,
// a WHT of the 8 child moduli, 64 3x3 block products and a WHT back; only the
// fine moduli are read.
// Levels >= 2 store explicit 24x24 element matrices (natural corner order
// b = bx + 2by + 4bz), built once per setup from the level below.
#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

#include "../../include/fantom.h"
#include "tile_skel.hpp"

namespace fantom {

// Nonzero pattern of the level-1 WHT blocks Ghat_0[s][s'] (bit 3 d + d' of
// entry s * 8 + s'): the same for every box shape and Poisson ratio (the
// zeros come from the reflection symmetry and the rigid translations), and
// checked by fantom_galerkin_tables.  279 of 576 entries are nonzero.
constexpr unsigned short kGalNZ[64] = {
    0x1ff, 0x1b6, 0x16d, 0x124, 0x0db, 0x092, 0x049, 0x000, 0x1f8, 0x1b0, 0x168, 0x138, 0x0d8, 0x1d0, 0x048, 0x158,
    0x1c7, 0x186, 0x145, 0x107, 0x0c3, 0x082, 0x1c1, 0x183, 0x1c0, 0x192, 0x149, 0x11b, 0x0c0, 0x1d2, 0x1c9, 0x1db,
    0x03f, 0x036, 0x02d, 0x024, 0x01b, 0x017, 0x039, 0x035, 0x038, 0x134, 0x028, 0x13c, 0x059, 0x155, 0x079, 0x17d,
    0x007, 0x006, 0x125, 0x127, 0x093, 0x097, 0x1b1, 0x1b7, 0x000, 0x116, 0x129, 0x13f, 0x0d1, 0x1d7, 0x1f9, 0x1ff};

// Classes of equal entries of the level-1 table for cubic elements: entry
// (s * 8 + s') * 9 + 3 d + d' holds class value kGalClass[...] (-1: zero).
// The partition into 20 classes comes from the symmetry of the cube under
// axis permutations and is the same for every Poisson ratio (generic ratios
// give 20 distinct values, special ones fewer); gal1_class_values checks a
// given table against it on the host, and other boxes (71 distinct values
// for 1 x 1.5 x 0.7) keep the table path.  With the classes the operator
// takes its 20 coefficients as kernel arguments, held in scalar registers,
// instead of reloading the 279 table entries through the scalar cache on
// every layer (the loads issued a few instructions before each use left the
// two waves per SIMD waiting on them).
constexpr int kGalClasses = 20;
constexpr signed char kGalClass[576] = {
    0, 1, 1, 1, 0, 1, 1, 1, 0, -1, 2, 2, -1, 3, 2, -1, 2, 3, 3, -1, 2, 2, -1, 2,
    2, -1, 3, -1, -1, 4, -1, -1, 4, -1, -1, 5, 3, 2, -1, 2, 3, -1, 2, 2, -1, -1, 4, -1,
    -1, 5, -1, -1, 4, -1, 5, -1, -1, 4, -1, -1, 4, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1,
    -1, -1, -1, 2, 3, 2, 2, 2, 3, -1, -1, -1, -1, 6, 7, -1, 7, 6, -1, -1, -1, 8, -1, 4,
    4, -1, 5, -1, -1, -1, 9, 9, 10, -1, -1, 11, -1, -1, -1, 4, 5, -1, 8, 4, -1, -1, -1, -1,
    -1, 11, -1, 9, 10, 9, -1, -1, -1, 12, -1, -1, 12, -1, -1, -1, -1, -1, 13, 13, -1, 13, -1, 13,
    3, 2, 2, -1, -1, -1, 2, 2, 3, -1, 8, 4, -1, -1, -1, -1, 4, 5, 6, -1, 7, -1, -1, -1,
    7, -1, 6, 9, 9, 10, -1, -1, -1, -1, -1, 11, 5, 4, -1, -1, -1, -1, 4, 8, -1, -1, 12, -1,
    -1, -1, -1, -1, 12, -1, 11, -1, -1, -1, -1, -1, 10, 9, 9, 13, 13, -1, -1, -1, -1, -1, 13, 13,
    -1, -1, -1, -1, -1, -1, 4, 4, 5, -1, 9, -1, -1, 9, -1, -1, 10, 11, 9, -1, -1, 9, -1, -1,
    10, -1, 11, 14, 14, -1, 14, 14, -1, -1, -1, 15, -1, -1, -1, -1, -1, -1, 12, 12, -1, -1, 13, -1,
    -1, 13, -1, 13, 16, 13, 13, -1, -1, 13, -1, -1, 16, 13, 13, 9, 9, -1, 9, 9, -1, 17, 17, 18,
    3, 2, 2, 2, 3, 2, -1, -1, -1, -1, 4, 8, -1, 5, 4, -1, -1, -1, 5, -1, 4, 4, -1, 8,
    -1, -1, -1, -1, -1, 12, -1, -1, 12, -1, -1, -1, 6, 7, -1, 7, 6, -1, -1, -1, -1, 9, 10, 9,
    -1, 11, -1, -1, -1, -1, 11, -1, -1, 10, 9, 9, -1, -1, -1, 13, -1, 13, -1, 13, 13, -1, -1, -1,
    -1, -1, -1, 4, 5, 4, -1, -1, -1, -1, -1, 9, -1, 11, 10, -1, -1, 9, -1, -1, -1, 12, -1, 12,
    -1, -1, -1, -1, -1, 13, 13, 13, 16, -1, -1, 13, 9, -1, -1, 10, 11, -1, 9, -1, -1, 14, -1, 14,
    -1, 15, -1, 14, -1, 14, 13, -1, -1, 16, 13, 13, 13, -1, -1, 9, -1, 9, 17, 18, 17, 9, -1, 9,
    5, 4, 4, -1, -1, -1, -1, -1, -1, -1, 12, 12, -1, -1, -1, -1, -1, -1, 11, -1, 10, -1, -1, 9,
    -1, -1, 9, 13, 13, 16, -1, -1, 13, -1, -1, 13, 11, 10, -1, -1, 9, -1, -1, 9, -1, 13, 16, 13,
    -1, 13, -1, -1, 13, -1, 15, -1, -1, -1, 14, 14, -1, 14, 14, 18, 17, 17, -1, 9, 9, -1, 9, 9,
    -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, 13, 13, -1, 13, -1, -1, -1, 13, 13, -1, -1, 13, -1, 13,
    -1, -1, 13, 9, 9, 17, 9, 9, 17, -1, -1, 18, 13, -1, -1, -1, 13, -1, 13, 13, -1, 9, 17, 9,
    -1, 18, -1, 9, 17, 9, 18, -1, -1, 17, 9, 9, 17, 9, 9, 19, 18, 18, 18, 19, 18, 18, 18, 19,
};

template <typename T>
__device__ __forceinline__ void gwht8(T *w)
{
#pragma unroll
    for (int h = 1; h < 8; h <<= 1) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if ((i & h) == 0) {
                const T a = w[i], b = w[i | h];
                w[i] = a + b;
                w[i | h] = a - b;
            }
        }
    }
}

// Level-1 Galerkin operator on the level-1 grid g (its fine grid is 2x in
// every direction), run on the shared tile skeleton.  Btab: 576 entries in
// the operator's precision, block (s, s') at Btab[(s*8 + s')*9 + 3*d + d']
// (scaled by 1/8); uniform
// reads, served by the scalar cache.
template <typename T, bool kCls = false>
struct Gal1Op {
    // resident waves per SIMD the kernels are built for (FP64 192-224 VGPRs:
    // 2; FP32 104-120: 4)
    static constexpr int kWavesPerEU = sizeof(T) == 8 ? 2 : 4;
    // the FP32 class form holds its 20 values in VGPRs and spills at four
    // waves; three without spills measured 5.6 us faster per V-cycle
    static constexpr int waves_for(int) { return (kCls && sizeof(T) == 4) ? kWavesPerEU - 1 : kWavesPerEU; }
    static constexpr bool staged_store(int mode) { return true; }
    static constexpr bool kScaled = false;
    static constexpr bool kDiagFromModuli = false;   // the Galerkin diagonal is read (dinv_gal1)
    __device__ __forceinline__ T inv_diag(int, T) const { return T(0); }
    const T *Btab;
    const T *Ef;
    int zero;   // 0, set by the host
    T cls[kCls ? kGalClasses : 1];   // kCls: the class values (kGalClass)
    struct Data {
        T e[8];   // child moduli, child bits cx + 2 cy + 4 cz
        int k;    // ey & zero
    };
    __device__ __forceinline__ void load(Data &d, const Grid &g, bool active, int ex, int ey, int ez) const
    {
        const int fy = 2 * g.nely, fz = 2 * g.nelz;
        d.k = ey & zero;
#pragma unroll
        for (int c = 0; c < 8; ++c)
            d.e[c] = active ? Ef[((long)(2 * ex + (c & 1)) * fy + (2 * ey + ((c >> 1) & 1))) * fz + (2 * ez + (c >> 2))]
                            : T(0);
    }
    __device__ __forceinline__ void apply(const Data &d, const T (&w)[3][8], T (&z)[3][8]) const
    {
        T eh[8];
#pragma unroll
        for (int c = 0; c < 8; ++c)
            eh[c] = d.e[c];
        gwht8(eh);
        // d.k is 0 at run time but depends on the layer, which keeps the 279
        // uniform table reads inside the layer loop as scalar loads; hoisted
        // out of the loop they would occupy VGPRs and cut occupancy.
        typedef const __attribute__((address_space(4))) T ConstT;   // read-only for the whole kernel
        ConstT *tab = (ConstT *)(Btab + __builtin_amdgcn_readfirstlane(d.k));
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            T a[3] = {T(0), T(0), T(0)};
#pragma unroll
            for (int s2 = 0; s2 < 8; ++s2) {
                const unsigned nz = kGalNZ[s * 8 + s2];
                if (nz == 0)
                    continue;
                const T h[3] = {w[0][s2 ^ 1], w[1][s2 ^ 2], w[2][s2 ^ 4]};
                ConstT *Bs = tab + (s * 8 + s2) * 9;
#pragma unroll
                for (int dr = 0; dr < 3; ++dr) {
                    if (((nz >> (3 * dr)) & 7) == 0)
                        continue;
                        T t = T(0);
#pragma unroll
                    for (int dc = 0; dc < 3; ++dc)
                        if ((nz >> (3 * dr + dc)) & 1) {
                            if constexpr (kCls)
                                t += cls[kGalClass[(s * 8 + s2) * 9 + 3 * dr + dc]] * h[dc];
                            else
                                t += Bs[3 * dr + dc] * h[dc];
                        }
                    a[dr] += eh[s ^ s2] * t;
                }
            }
            z[0][s ^ 1] = a[0];
            z[1][s ^ 2] = a[1];
            z[2][s ^ 4] = a[2];
        }
    }
};

// ---- explicit element matrices (levels >= 2) ----------------------------------------
// corner b of an element sits at node offset (b & 1, (b >> 1) & 1, b >> 2)
__device__ __forceinline__ long node_of(const Grid &g, int ei, int ej, int ek, int b)
{
    return (long)(ej + ((b >> 1) & 1)) * g.nx * g.nz + (long)(ei + (b & 1)) * g.nz + (ek + (b >> 2));
}

// y = K u with explicit element matrices Ke[e][24][24] (natural order),
// node-centric gather in a fixed element order; Dirichlet rule as above.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void ku_elem(Grid g, const double *__restrict__ Ke,
                                             const uint8_t *__restrict__ mask, const T *__restrict__ u,
                                             T *__restrict__ y, const T *__restrict__ bvec,
                                             const T *__restrict__ dinv, T omega, double *__restrict__ partials)
{
    if (MODE == MODE_JAC)
        omega = jacobi_weight(g, omega);
    const long n = (long)blockIdx.x * 256 + threadIdx.x;
    double part = 0.0;
    if (n < g.n_nodes) {
        const int k = (int)(n % g.nz);
        const long t = n / g.nz;
        const int i = (int)(t % g.nx), j = (int)(t / g.nx);
        double acc[3] = {0.0, 0.0, 0.0};
        for (int ox = 1; ox >= 0; --ox)
            for (int oy = 1; oy >= 0; --oy)
                for (int oz = 1; oz >= 0; --oz) {
                    const int ei = i - ox, ej = j - oy, ek = k - oz;
                    if (ei < 0 || ei >= g.nelx || ej < 0 || ej >= g.nely || ek < 0 || ek >= g.nelz)
                        continue;
                    const long e = ((long)ei * g.nely + ej) * g.nelz + ek;
                    const int a = ox + 2 * oy + 4 * oz;
                    const double *K = Ke + e * 576 + a * 72;
                    for (int b = 0; b < 8; ++b) {
                        const long m = node_of(g, ei, ej, ek, b);
                        const int mm = mask ? mask[m] : 0;
                        for (int d = 0; d < 3; ++d) {
                            if ((mm >> d) & 1)
                                continue;
                            const double ud = (double)u[3 * m + d];
                            acc[0] += K[b * 3 + d] * ud;
                            acc[1] += K[24 + b * 3 + d] * ud;
                            acc[2] += K[48 + b * 3 + d] * ud;
                        }
                    }
                }
        const int mn = mask ? mask[n] : 0;
        for (int c = 0; c < 3; ++c) {
            const long r = 3 * n + c;
            const T ur = u[r];
            const T ku = ((mn >> c) & 1) ? ur : (T)acc[c];
            T o;
            if (MODE == MODE_JAC)
                o = ur + omega * dinv[r] * (bvec[r] - ku);
            else if (MODE == MODE_RES)
                o = bvec[r] - ku;
            else
                o = ku;
            y[r] = o;
            if (MODE == MODE_KUDOT)
                part += (double)ur * (double)o;
        }
    }
    if (MODE == MODE_KUDOT) {
        const double s = block_sum<256>(part);
        if (threadIdx.x == 0)
            partials[blockIdx.x] = s;
    }
}

// Jacobi weights 1/diag from explicit element matrices (constrained -> 1)
template <typename T>
__global__ void dinv_elem(Grid g, const double *__restrict__ Ke, const uint8_t *__restrict__ mask,
                          T *__restrict__ d)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.n_nodes)
        return;
    const int k = (int)(n % g.nz);
    const long t = n / g.nz;
    const int i = (int)(t % g.nx), j = (int)(t / g.nx);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int ox = 1; ox >= 0; --ox)
        for (int oy = 1; oy >= 0; --oy)
            for (int oz = 1; oz >= 0; --oz) {
                const int ei = i - ox, ej = j - oy, ek = k - oz;
                if (ei < 0 || ei >= g.nelx || ej < 0 || ej >= g.nely || ek < 0 || ek >= g.nelz)
                    continue;
                const long e = ((long)ei * g.nely + ej) * g.nelz + ek;
                const int a = ox + 2 * oy + 4 * oz;
                for (int c = 0; c < 3; ++c)
                    acc[c] += Ke[e * 576 + (a * 3 + c) * 24 + a * 3 + c];
            }
    const int mn = mask ? mask[n] : 0;
    for (int c = 0; c < 3; ++c)
        d[3 * n + c] = ((mn >> c) & 1) ? T(1) : (T)(1.0 / acc[c]);
}

// Jacobi weights of the level-1 Galerkin operator from the fine moduli:
// diag K_E[(a,c),(a,c)] = sum_child E_child * Gdiag[child][a*3+c]
template <typename T>
__global__ void dinv_gal1(Grid g, const double *__restrict__ Gdiag, const T *__restrict__ Ef,
                          const uint8_t *__restrict__ mask, T *__restrict__ d)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.n_nodes)
        return;
    const int fy = 2 * g.nely, fz = 2 * g.nelz;
    const int k = (int)(n % g.nz);
    const long t = n / g.nz;
    const int i = (int)(t % g.nx), j = (int)(t / g.nx);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int ox = 1; ox >= 0; --ox)
        for (int oy = 1; oy >= 0; --oy)
            for (int oz = 1; oz >= 0; --oz) {
                const int ei = i - ox, ej = j - oy, ek = k - oz;
                if (ei < 0 || ei >= g.nelx || ej < 0 || ej >= g.nely || ek < 0 || ek >= g.nelz)
                    continue;
                const int a = ox + 2 * oy + 4 * oz;
                for (int ch = 0; ch < 8; ++ch) {
                    const double Ec =
                        (double)Ef[((long)(2 * ei + (ch & 1)) * fy + (2 * ej + ((ch >> 1) & 1))) * fz + (2 * ek + (ch >> 2))];
                    for (int c = 0; c < 3; ++c)
                        acc[c] += Ec * Gdiag[ch * 24 + a * 3 + c];
                }
            }
    const int mn = mask ? mask[n] : 0;
    for (int c = 0; c < 3; ++c)
        d[3 * n + c] = ((mn >> c) & 1) ? T(1) : (T)(1.0 / acc[c]);
}

// Build the coarse element matrices of level l+1 (grid gc) from level l:
//   Kc = sum_{child} P_child^T K_child P_child
// src_kind 0: children are level-1 elements given by fine moduli Ef and the
// child tables G (8 x 576, K_child = sum_g E_g G_g); src_kind 1: children are
// explicit element matrices Kf of the level below (grid 2x gc).
// One block of 256 threads per coarse element; P weights are computed from
// the trilinear rule.
template <typename TE>
__global__ __launch_bounds__(256) void gal_build(Grid gc, int src_kind, const double *__restrict__ G,
                                               const TE *__restrict__ Ef, const double *__restrict__ Kf,
                                               double *__restrict__ Kc)
{
    __shared__ double Kch[576];
    __shared__ double Tm[576];
    const long E = blockIdx.x;
    if (E >= gc.n_el)
        return;
    const int ei = (int)(E / ((long)gc.nely * gc.nelz));
    const int r0 = (int)(E % ((long)gc.nely * gc.nelz));
    const int ej = r0 / gc.nelz, ek = r0 % gc.nelz;
    // children live on the grid below: 2 nelx x 2 nely x 2 nelz elements
    const int fnx = 2 * gc.nelx, fny = 2 * gc.nely, fnz = 2 * gc.nelz;
    (void)fnx;
    double acc[3] = {0.0, 0.0, 0.0};  // each thread owns up to 3 output entries (576 / 256)
    for (int ch = 0; ch < 8; ++ch) {
        const int cx = ch & 1, cy = (ch >> 1) & 1, cz = ch >> 2;
        const long fe = ((long)(2 * ei + cx) * fny + (2 * ej + cy)) * fnz + (2 * ek + cz);
        __syncthreads();
        if (src_kind == 0) {
            // level-1 child: K = sum over its 8 fine children g of E_g G_g
            double Eg[8];
            for (int gg = 0; gg < 8; ++gg) {
                const int gx = gg & 1, gy = (gg >> 1) & 1, gz = gg >> 2;
                Eg[gg] = (double)Ef[((long)(2 * (2 * ei + cx) + gx) * (2 * fny) + (2 * (2 * ej + cy) + gy)) * (2 * fnz) +
                                    (2 * (2 * ek + cz) + gz)];
            }
            for (int q = threadIdx.x; q < 576; q += 256) {
                double s = 0.0;
                for (int gg = 0; gg < 8; ++gg)
                    s += Eg[gg] * G[gg * 576 + q];
                Kch[q] = s;
            }
        } else {
            for (int q = threadIdx.x; q < 576; q += 256)
                Kch[q] = Kf[fe * 576 + q];
        }
        __syncthreads();
        // T = Kch * P_ch  (P_ch[(gamma, d), (beta, d')] = delta_dd' w(beta, ch + gamma))
        for (int q = threadIdx.x; q < 576; q += 256) {
            const int row = q / 24, col = q % 24;          // row: fine corner dof, col: coarse corner dof
            const int beta = col / 3, dcol = col % 3;
            double s = 0.0;
            for (int gam = 0; gam < 8; ++gam) {
                // weight of coarse corner beta at fine node ch + gam (per axis)
                double wgt = 1.0;
                for (int ax = 0; ax < 3; ++ax) {
                    const int p = ((ch >> ax) & 1) + ((gam >> ax) & 1);  // 0, 1, 2
                    const int bb = (beta >> ax) & 1;
                    wgt *= (p == 1) ? 0.5 : ((p == 0) == (bb == 0) ? 1.0 : 0.0);
                }
                if (wgt != 0.0)
                    s += Kch[row * 24 + gam * 3 + dcol] * wgt;
            }
            Tm[q] = s;
        }
        __syncthreads();
        // acc += P_ch^T T
        for (int qi = 0; qi < 3; ++qi) {
            const int q = threadIdx.x + qi * 256;
            if (q >= 576)
                break;
            const int row = q / 24, col = q % 24;          // row: coarse corner dof
            const int beta = row / 3, drow = row % 3;
            double s = 0.0;
            for (int gam = 0; gam < 8; ++gam) {
                double wgt = 1.0;
                for (int ax = 0; ax < 3; ++ax) {
                    const int p = ((ch >> ax) & 1) + ((gam >> ax) & 1);
                    const int bb = (beta >> ax) & 1;
                    wgt *= (p == 1) ? 0.5 : ((p == 0) == (bb == 0) ? 1.0 : 0.0);
                }
                if (wgt != 0.0)
                    s += wgt * Tm[(gam * 3 + drow) * 24 + col];
            }
            acc[qi] += s;
        }
    }
    for (int qi = 0; qi < 3; ++qi) {
        const int q = threadIdx.x + qi * 256;
        if (q < 576)
            Kc[E * 576 + q] = acc[qi];
    }
}

// Level-2 element matrices straight from the fine moduli: with the 64 fine
// elements f = 8 c + g of a level-2 element (c its level-1 child, g that
// child's fine child), K = sum_f E_f G2[f], G2[8c + g] = P_c^T G_g P_c
// (galerkin_tables2).  A block of 256 threads builds 8 coarse elements: the
// 8 x 64 moduli wait in LDS, each thread owns up to 3 of the 576 entries for
// all 8 elements and streams the tables from L2.
constexpr int kG2Elems = 8;

template <typename TE>
__global__ __launch_bounds__(256) void gal2_from_fine(Grid gc, const double *__restrict__ G2,
                                                    const TE *__restrict__ Ef, double *__restrict__ Kc)
{
    __shared__ double Es[kG2Elems][64];
    const long e0 = (long)blockIdx.x * kG2Elems;
    const int fny = 4 * gc.nely, fnz = 4 * gc.nelz;   // fine element counts along y and z
    for (int q = threadIdx.x; q < kG2Elems * 64; q += 256) {
        const int el = q / 64, f = q % 64, c = f >> 3, gch = f & 7;
        const long E = e0 + el;
        double v = 0.0;
        if (E < gc.n_el) {
            const int ei = (int)(E / ((long)gc.nely * gc.nelz));
            const int r0 = (int)(E % ((long)gc.nely * gc.nelz));
            const int ej = r0 / gc.nelz, ek = r0 % gc.nelz;
            const int fx = 2 * (2 * ei + (c & 1)) + (gch & 1);
            const int fy = 2 * (2 * ej + ((c >> 1) & 1)) + ((gch >> 1) & 1);
            const int fz = 2 * (2 * ek + (c >> 2)) + (gch >> 2);
            v = (double)Ef[((long)fx * fny + fy) * fnz + fz];
        }
        Es[el][f] = v;
    }
    __syncthreads();
    double acc[3][kG2Elems];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int el = 0; el < kG2Elems; ++el)
            acc[k][el] = 0.0;
    const int q0 = threadIdx.x, q1 = threadIdx.x + 256, q2 = threadIdx.x + 512;
    for (int f = 0; f < 64; ++f) {
        const double g0 = G2[f * 576 + q0], g1 = G2[f * 576 + q1];
        const double g2 = q2 < 576 ? G2[f * 576 + q2] : 0.0;
#pragma unroll
        for (int el = 0; el < kG2Elems; ++el) {
            const double e = Es[el][f];
            acc[0][el] += e * g0;
            acc[1][el] += e * g1;
            acc[2][el] += e * g2;
        }
    }
#pragma unroll
    for (int el = 0; el < kG2Elems; ++el) {
        const long E = e0 + el;
        if (E >= gc.n_el)
            break;
        Kc[E * 576 + q0] = acc[0][el];
        Kc[E * 576 + q1] = acc[1][el];
        if (q2 < 576)
            Kc[E * 576 + q2] = acc[2][el];
    }
}

// Level-1 element matrices materialised from the fine moduli (used when the
// coarsest level is level 1): K_E = sum_g E_g G_g
template <typename TE>
__global__ void elem_from_fine(Grid gc, const double *__restrict__ G, const TE *__restrict__ Ef,
                               double *__restrict__ Ke)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= gc.n_el * 576)
        return;
    const long E = idx / 576;
    const int q = (int)(idx % 576);
    const int ei = (int)(E / ((long)gc.nely * gc.nelz));
    const int r0 = (int)(E % ((long)gc.nely * gc.nelz));
    const int ej = r0 / gc.nelz, ek = r0 % gc.nelz;
    const int fy = 2 * gc.nely, fz = 2 * gc.nelz;
    double s = 0.0;
    for (int gg = 0; gg < 8; ++gg) {
        const double Eg = (double)Ef[((long)(2 * ei + (gg & 1)) * fy + (2 * ej + ((gg >> 1) & 1))) * fz + (2 * ek + (gg >> 2))];
        s += Eg * G[gg * 576 + q];
    }
    Ke[idx] = s;
}

// dense operator from explicit element matrices (one thread per row)
__global__ void dense_from_elem(Grid g, const double *__restrict__ Ke, const uint8_t *__restrict__ mask,
                                double *__restrict__ A)
{
    const long ndof = g.n_nodes * 3;
    const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= ndof)
        return;
    const long n = row / 3;
    const int c = (int)(row % 3);
    double *Ar = A + row * ndof;
    const int mn = mask ? mask[n] : 0;
    if ((mn >> c) & 1) {
        Ar[row] = 1.0;
        return;
    }
    const int k = (int)(n % g.nz);
    const long t = n / g.nz;
    const int i = (int)(t % g.nx), j = (int)(t / g.nx);
    for (int ox = 1; ox >= 0; --ox)
        for (int oy = 1; oy >= 0; --oy)
            for (int oz = 1; oz >= 0; --oz) {
                const int ei = i - ox, ej = j - oy, ek = k - oz;
                if (ei < 0 || ei >= g.nelx || ej < 0 || ej >= g.nely || ek < 0 || ek >= g.nelz)
                    continue;
                const long e = ((long)ei * g.nely + ej) * g.nelz + ek;
                const int a = ox + 2 * oy + 4 * oz;
                for (int b = 0; b < 8; ++b) {
                    const long m = node_of(g, ei, ej, ek, b);
                    const int mm = mask ? mask[m] : 0;
                    for (int d = 0; d < 3; ++d) {
                        if ((mm >> d) & 1)
                            continue;
                        Ar[m * 3 + d] += Ke[e * 576 + (a * 3 + c) * 24 + b * 3 + d];
                    }
                }
            }
}

// Workgroups per CU the chunk choice assumes for each mode of the level-1
// operator: the occupancy of that mode's instantiation (FP64 2, FP32 4 at
// 116-123 VGPRs), the register budget without a device.
template <typename T, bool kCls>
static int gal1_per_cu(int mode)
{
    using Op = Gal1Op<T, kCls>;
    static const std::array<int, 5> pc = [] {
        std::array<int, 5> a = {tile_occupancy<T, 0, Op>(), tile_occupancy<T, 1, Op>(), tile_occupancy<T, 2, Op>(),
                                tile_occupancy<T, 3, Op>(), tile_occupancy<T, 4, Op>()};
        for (int m = 0; m < (int)a.size(); ++m)
            if (a[m] <= 0)
                a[m] = Op::waves_for(m);   // a workgroup is one wave per SIMD
        return a;
    }();
    return pc[mode];
}

// partial sums of the one mode that writes them (K.u + u.Ku), any variant
long gal1_partials(const Grid &g)
{
    unsigned n = 0;
    for (int v : {gal1_per_cu<double, false>(MODE_KUDOT), gal1_per_cu<float, false>(MODE_KUDOT),
                  gal1_per_cu<double, true>(MODE_KUDOT), gal1_per_cu<float, true>(MODE_KUDOT)})
        n = std::max(n, tile_blocks(tile_shape(g, v)));
    return (long)n;
}

bool gal1_class_values(const double *B, double *cls)
{
    double mx = 0.0;
    for (int i = 0; i < 576; ++i)
        mx = std::fmax(mx, std::fabs(B[i]));
    if (!(mx > 0.0))
        return false;
    bool seen[kGalClasses] = {};
    for (int i = 0; i < 576; ++i) {
        const int c = kGalClass[i];
        if (c < 0) {
            if (std::fabs(B[i]) > 1e-12 * mx)
                return false;
            continue;
        }
        if (!seen[c]) {
            cls[c] = B[i];
            seen[c] = true;
        } else if (std::fabs(B[i] - cls[c]) > 1e-12 * mx) {
            return false;
        }
    }
    for (bool v : seen)
        if (!v)
            return false;
    return true;
}

template <typename T, bool kCls>
static int gal1_launch(int mode, const Grid &g, const T *Btab, const T *Ef, const uint8_t *mask, const T *u,
                       T *y, const T *b, const T *dinv, double omega, double *partials, hipStream_t s,
                       const double *cls)
{
    const TileShape ts = tile_shape(g, gal1_per_cu<T, kCls>(mode >= 0 && mode < 5 ? mode : 0));   // bad modes fail below
    const unsigned blocks = tile_blocks(ts);
    const T om = (T)omega;
    Gal1Op<T, kCls> op;
    op.Btab = Btab;
    op.Ef = Ef;
    op.zero = 0;
    if constexpr (kCls)
        for (int c = 0; c < kGalClasses; ++c)
            op.cls[c] = (T)cls[c];
#define GAL_CASE(M)                                                                                           \
    case M:                                                                                                   \
        hipLaunchKernelGGL((ku_tile_kernel<T, M, Gal1Op<T, kCls>>), dim3(blocks), dim3(kTThreads), 0, s, g, op, mask, u, \
                           y, b, dinv, om, partials, ts, nullptr);                                            \
        break;
    switch (mode) {
        GAL_CASE(MODE_KU)
        GAL_CASE(MODE_JAC)
        GAL_CASE(MODE_RES)
        GAL_CASE(MODE_KUDOT)
        GAL_CASE(MODE_JAC0)
    default:
        set_error("unknown mode %d", mode);
        return kStatusBadArg;
    }
#undef GAL_CASE
    FANTOM_LAUNCH_CHECK("ku_gal1");
    return kStatusOk;
}

int launch_gal1(int mode, const Grid &g, const double *Btab, const double *Ef, const uint8_t *mask, const double *u,
                double *y, const double *b, const double *dinv, double omega, double *partials, hipStream_t s,
                const double *cls)
{
    return cls ? gal1_launch<double, true>(mode, g, Btab, Ef, mask, u, y, b, dinv, omega, partials, s, cls)
               : gal1_launch<double, false>(mode, g, Btab, Ef, mask, u, y, b, dinv, omega, partials, s, nullptr);
}

int launch_gal1_f32(int mode, const Grid &g, const float *Btab, const float *Ef, const uint8_t *mask, const float *u,
                    float *y, const float *b, const float *dinv, double omega, double *partials, hipStream_t s,
                    const double *cls)
{
    return cls ? gal1_launch<float, true>(mode, g, Btab, Ef, mask, u, y, b, dinv, omega, partials, s, cls)
               : gal1_launch<float, false>(mode, g, Btab, Ef, mask, u, y, b, dinv, omega, partials, s, nullptr);
}

template <typename T>
static int elem_launch(int mode, const Grid &g, const double *Ke, const uint8_t *mask, const T *u, T *y, const T *b,
                       const T *dinv, double omega, double *partials, hipStream_t s)
{
    const unsigned blocks = (unsigned)((g.n_nodes + 255) / 256);
    const T om = (T)omega;
    switch (mode) {
    case MODE_KU:
        hipLaunchKernelGGL((ku_elem<T, MODE_KU>), dim3(blocks), dim3(256), 0, s, g, Ke, mask, u, y, b, dinv, om,
                           partials);
        break;
    case MODE_JAC:
        hipLaunchKernelGGL((ku_elem<T, MODE_JAC>), dim3(blocks), dim3(256), 0, s, g, Ke, mask, u, y, b, dinv, om,
                           partials);
        break;
    case MODE_RES:
        hipLaunchKernelGGL((ku_elem<T, MODE_RES>), dim3(blocks), dim3(256), 0, s, g, Ke, mask, u, y, b, dinv, om,
                           partials);
        break;
    case MODE_KUDOT:
        hipLaunchKernelGGL((ku_elem<T, MODE_KUDOT>), dim3(blocks), dim3(256), 0, s, g, Ke, mask, u, y, b, dinv, om,
                           partials);
        break;
    default:
        set_error("unknown mode %d", mode);
        return kStatusBadArg;
    }
    FANTOM_LAUNCH_CHECK("ku_elem");
    return kStatusOk;
}

int launch_elem(int mode, const Grid &g, const double *Ke, const uint8_t *mask, const double *u, double *y,
                const double *b, const double *dinv, double omega, double *partials, hipStream_t s)
{
    return elem_launch<double>(mode, g, Ke, mask, u, y, b, dinv, omega, partials, s);
}

int launch_elem_f32(int mode, const Grid &g, const double *Ke, const uint8_t *mask, const float *u, float *y,
                    const float *b, const float *dinv, double omega, double *partials, hipStream_t s)
{
    return elem_launch<float>(mode, g, Ke, mask, u, y, b, dinv, omega, partials, s);
}

long elem_partials(const Grid &g) { return (g.n_nodes + 255) / 256; }

template <typename T>
static int dinv_gal1_t(const Grid &g, const double *Gdiag, const T *Ef, const uint8_t *mask, T *d, hipStream_t s)
{
    hipLaunchKernelGGL(dinv_gal1<T>, dim3((unsigned)((g.n_nodes + 255) / 256)), dim3(256), 0, s, g, Gdiag, Ef, mask, d);
    FANTOM_LAUNCH_CHECK("dinv_gal1");
    return kStatusOk;
}

int launch_dinv_gal1(const Grid &g, const double *Gdiag, const double *Ef, const uint8_t *mask, double *d,
                     hipStream_t s)
{
    return dinv_gal1_t<double>(g, Gdiag, Ef, mask, d, s);
}

int launch_dinv_gal1_f32(const Grid &g, const double *Gdiag, const float *Ef, const uint8_t *mask, float *d,
                         hipStream_t s)
{
    return dinv_gal1_t<float>(g, Gdiag, Ef, mask, d, s);
}

int launch_dinv_elem(const Grid &g, const double *Ke, const uint8_t *mask, double *d, hipStream_t s)
{
    hipLaunchKernelGGL(dinv_elem<double>, dim3((unsigned)((g.n_nodes + 255) / 256)), dim3(256), 0, s, g, Ke, mask, d);
    FANTOM_LAUNCH_CHECK("dinv_elem");
    return kStatusOk;
}

int launch_dinv_elem_f32(const Grid &g, const double *Ke, const uint8_t *mask, float *d, hipStream_t s)
{
    hipLaunchKernelGGL(dinv_elem<float>, dim3((unsigned)((g.n_nodes + 255) / 256)), dim3(256), 0, s, g, Ke, mask, d);
    FANTOM_LAUNCH_CHECK("dinv_elem");
    return kStatusOk;
}

template <typename TE>
static int gal_build_t(const Grid &gc, int src_kind, const double *G, const TE *Ef, const double *Kf, double *Kc,
                       hipStream_t s)
{
    hipLaunchKernelGGL(gal_build<TE>, dim3((unsigned)gc.n_el), dim3(256), 0, s, gc, src_kind, G, Ef, Kf, Kc);
    FANTOM_LAUNCH_CHECK("gal_build");
    return kStatusOk;
}

template <typename TE>
static int gal2_launch(const Grid &gc, const double *G2, const TE *Ef, double *Kc, hipStream_t s)
{
    const unsigned blocks = (unsigned)((gc.n_el + kG2Elems - 1) / kG2Elems);
    hipLaunchKernelGGL(gal2_from_fine<TE>, dim3(blocks), dim3(256), 0, s, gc, G2, Ef, Kc);
    FANTOM_LAUNCH_CHECK("gal2_from_fine");
    return kStatusOk;
}

int launch_gal2_build(const Grid &gc, const double *G2, const double *Ef, double *Kc, hipStream_t s)
{
    return gal2_launch<double>(gc, G2, Ef, Kc, s);
}

int launch_gal2_build_f32E(const Grid &gc, const double *G2, const float *Ef, double *Kc, hipStream_t s)
{
    return gal2_launch<float>(gc, G2, Ef, Kc, s);
}

int launch_gal_build(const Grid &gc, int src_kind, const double *G, const double *Ef, const double *Kf, double *Kc,
                     hipStream_t s)
{
    return gal_build_t<double>(gc, src_kind, G, Ef, Kf, Kc, s);
}

int launch_gal_build_f32E(const Grid &gc, int src_kind, const double *G, const float *Ef, const double *Kf,
                          double *Kc, hipStream_t s)
{
    return gal_build_t<float>(gc, src_kind, G, Ef, Kf, Kc, s);
}

int launch_elem_from_fine(const Grid &gc, const double *G, const double *Ef, double *Ke, hipStream_t s)
{
    const long n = gc.n_el * 576;
    hipLaunchKernelGGL(elem_from_fine<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, gc, G, Ef, Ke);
    FANTOM_LAUNCH_CHECK("elem_from_fine");
    return kStatusOk;
}

int launch_elem_from_fine_f32E(const Grid &gc, const double *G, const float *Ef, double *Ke, hipStream_t s)
{
    const long n = gc.n_el * 576;
    hipLaunchKernelGGL(elem_from_fine<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, gc, G, Ef, Ke);
    FANTOM_LAUNCH_CHECK("elem_from_fine");
    return kStatusOk;
}

int launch_dense_from_elem(const Grid &g, const double *Ke, const uint8_t *mask, double *A, hipStream_t s)
{
    const long ndof = g.n_nodes * 3;
    hipError_t e = hipMemsetAsync(A, 0, sizeof(double) * ndof * ndof, s);
    if (e != hipSuccess) {
        set_error("dense_from_elem memset: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    hipLaunchKernelGGL(dense_from_elem, dim3((unsigned)((ndof + 255) / 256)), dim3(256), 0, s, g, Ke, mask, A);
    FANTOM_LAUNCH_CHECK("dense_from_elem");
    return kStatusOk;
}

}  // namespace fantom

using namespace fantom;

// ---- host tables -----------------------------------------------------------------------
// out[ch] = P_ch^T Kn P_ch (24 x 24 each, natural corner order) for the 8
// children ch of a coarse element, P_ch the trilinear prolongation from the
// coarse element's corners to child ch's corners.
static void child_products(const double *Kn, double *out)
{
    for (int ch = 0; ch < 8; ++ch) {
        double P[24][24] = {};
        for (int gam = 0; gam < 8; ++gam)
            for (int beta = 0; beta < 8; ++beta) {
                double wgt = 1.0;
                for (int ax = 0; ax < 3; ++ax) {
                    const int p = ((ch >> ax) & 1) + ((gam >> ax) & 1);
                    const int bb = (beta >> ax) & 1;
                    wgt *= (p == 1) ? 0.5 : ((p == 0) == (bb == 0) ? 1.0 : 0.0);
                }
                for (int d = 0; d < 3; ++d)
                    P[gam * 3 + d][beta * 3 + d] = wgt;
            }
        double T[24][24];
        for (int i = 0; i < 24; ++i)
            for (int j = 0; j < 24; ++j) {
                double s = 0.0;
                for (int k = 0; k < 24; ++k)
                    s += Kn[i * 24 + k] * P[k][j];
                T[i][j] = s;
            }
        for (int i = 0; i < 24; ++i)
            for (int j = 0; j < 24; ++j) {
                double s = 0.0;
                for (int k = 0; k < 24; ++k)
                    s += P[k][i] * T[k][j];
                out[ch * 576 + i * 24 + j] = s;
            }
    }
}

int fantom::galerkin_tables2(const double *G, double *G2)
{
    std::vector<double> tmp(8 * 576);
    for (int g = 0; g < 8; ++g) {
        child_products(G + g * 576, tmp.data());
        for (int c = 0; c < 8; ++c)
            for (int q = 0; q < 576; ++q)
                G2[(c * 8 + g) * 576 + q] = tmp[c * 576 + q];
    }
    return kStatusOk;
}

extern "C" int fantom_galerkin_tables(const double *K0, double *G, double *Btab, double *Gdiag, double *check)
{
    FANTOM_CHECK_ARG(K0 && G && Btab && Gdiag, "null operand");
    static const int loc[8] = {0, 1, 3, 2, 4, 5, 7, 6};  // natural corner -> reference local corner
    double Kn[24][24];
    for (int b = 0; b < 8; ++b)
        for (int c = 0; c < 3; ++c)
            for (int b2 = 0; b2 < 8; ++b2)
                for (int d = 0; d < 3; ++d)
                    Kn[b * 3 + c][b2 * 3 + d] = K0[(loc[b] * 3 + c) * 24 + loc[b2] * 3 + d];
    // child matrices G_ch = P_ch^T Kn P_ch
    child_products(&Kn[0][0], G);
    for (int ch = 0; ch < 8; ++ch)
        for (int i = 0; i < 24; ++i)
            Gdiag[ch * 24 + i] = G[ch * 576 + i * 24 + i];
    // Ghat_0 = V^T G_0 V with orthonormal V, blocks scaled by 1/8
    auto H = [](int b, int t) { return (__builtin_popcount(b & t) & 1) ? -1.0 : 1.0; };
    double maxabs = 0.0, sym_err = 0.0, sum_err = 0.0;
    for (int s = 0; s < 8; ++s)
        for (int d = 0; d < 3; ++d)
            for (int s2 = 0; s2 < 8; ++s2)
                for (int d2 = 0; d2 < 3; ++d2) {
                    const int t1 = s ^ (1 << d), t2 = s2 ^ (1 << d2);
                    double acc = 0.0;
                    for (int b = 0; b < 8; ++b)
                        for (int b2 = 0; b2 < 8; ++b2)
                            acc += H(b, t1) * G[(b * 3 + d) * 24 + b2 * 3 + d2] * H(b2, t2);
                    acc /= 8.0;  // orthonormal V: (1/sqrt8)^2
                    Btab[(s * 8 + s2) * 9 + d * 3 + d2] = acc / 8.0;
                    maxabs = fmax(maxabs, fabs(acc));
                }
    // consistency: the child matrices are reflections of G_0
    for (int ch = 0; ch < 8; ++ch)
        for (int b = 0; b < 8; ++b)
            for (int c = 0; c < 3; ++c)
                for (int b2 = 0; b2 < 8; ++b2)
                    for (int d = 0; d < 3; ++d) {
                        const double sgn = (((ch >> c) & 1) ? -1.0 : 1.0) * (((ch >> d) & 1) ? -1.0 : 1.0);
                        const double ref = sgn * G[((b ^ ch) * 3 + c) * 24 + (b2 ^ ch) * 3 + d];
                        sym_err = fmax(sym_err, fabs(G[ch * 576 + (b * 3 + c) * 24 + b2 * 3 + d] - ref));
                    }
    // consistency: the children sum to the doubled element (3D hex scaling)
    for (int i = 0; i < 576; ++i) {
        double s = 0.0;
        for (int ch = 0; ch < 8; ++ch)
            s += G[ch * 576 + i];
        sum_err = fmax(sum_err, fabs(s - 2.0 * Kn[i / 24][i % 24]));
    }
    double pat_err = 0.0;   // entries outside the fixed nonzero pattern
    for (int q = 0; q < 64; ++q)
        for (int i = 0; i < 9; ++i)
            if (!((kGalNZ[q] >> i) & 1))
                pat_err = fmax(pat_err, fabs(Btab[q * 9 + i]) * 8.0);
    if (pat_err > 1e-12 * maxabs) {
        set_error("level-1 Galerkin blocks leave the box-hex8 nonzero pattern (%.2e)", pat_err);
        return kStatusBadArg;
    }
    if (check) {
        check[0] = sym_err / fmax(maxabs, 1e-300);
        check[1] = sum_err / fmax(maxabs, 1e-300);
    }
    if (sym_err > 1e-12 * maxabs || sum_err > 1e-12 * maxabs) {
        set_error("K0 lacks the box-element symmetry the Galerkin tables need (%.2e, %.2e)", sym_err, sum_err);
        return kStatusBadArg;
    }
    return kStatusOk;
}

extern "C" int fantom_galerkin_tables2(const double *G, double *G2)
{
    FANTOM_CHECK_ARG(G && G2, "null operand");
    return galerkin_tables2(G, G2);
}

extern "C" int fantom_gal2_build_f64(int nelx, int nely, int nelz, const double *G2, const double *Efine, double *Kc,
                                     void *stream)
{
    FANTOM_CHECK_ARG(grid_ok(nelx, nely, nelz) && nelz > 0 && G2 && Efine && Kc, "bad operands");
    return launch_gal2_build(make_grid(nelx, nely, nelz), G2, Efine, Kc, as_stream(stream));
}

extern "C" int fantom_gal1_ku_f64(int nelx, int nely, int nelz, const double *Btab, const double *Efine,
                                  const uint8_t *mask, const double *u, double *y, void *stream)
{
    FANTOM_CHECK_ARG(grid_ok(nelx, nely, nelz) && nelz > 0 && Btab && Efine && u && y && u != y, "bad operands");
    return launch_gal1(MODE_KU, make_grid(nelx, nely, nelz), Btab, Efine, mask, u, y, nullptr, nullptr, 0.0, nullptr,
                       as_stream(stream));
}

extern "C" int fantom_gal1_classes(const double *B, double *cls, int *is_cubic)
{
    FANTOM_CHECK_ARG(B && cls && is_cubic, "null operand");
    *is_cubic = gal1_class_values(B, cls) ? 1 : 0;
    return kStatusOk;
}

extern "C" int fantom_gal1_ku_cls_f64(int nelx, int nely, int nelz, const double *cls, const double *Efine,
                                      const uint8_t *mask, const double *u, double *y, void *stream)
{
    FANTOM_CHECK_ARG(grid_ok(nelx, nely, nelz) && nelz > 0, "bad 3D grid");
    FANTOM_CHECK_ARG(cls && Efine && u && y && u != y, "bad operands");
    return launch_gal1(MODE_KU, make_grid(nelx, nely, nelz), nullptr, Efine, mask, u, y, nullptr, nullptr, 0.0,
                       nullptr, as_stream(stream), cls);
}

extern "C" int fantom_gal_build_f64(int nelx, int nely, int nelz, int src_kind, const double *G, const double *Efine,
                                    const double *Kfine, double *Kc, void *stream)
{
    FANTOM_CHECK_ARG(grid_ok(nelx, nely, nelz) && nelz > 0 && G && Kc, "bad operands");
    FANTOM_CHECK_ARG((src_kind == 0 && Efine) || (src_kind == 1 && Kfine), "bad source");
    return launch_gal_build(make_grid(nelx, nely, nelz), src_kind, G, Efine, Kfine, Kc, as_stream(stream));
}

extern "C" int fantom_elem_ku_f64(int nelx, int nely, int nelz, const double *Ke, const uint8_t *mask, const double *u,
                                  double *y, void *stream)
{
    FANTOM_CHECK_ARG(grid_ok(nelx, nely, nelz) && nelz > 0 && Ke && u && y && u != y, "bad operands");
    return launch_elem(MODE_KU, make_grid(nelx, nely, nelz), Ke, mask, u, y, nullptr, nullptr, 0.0, nullptr,
                       as_stream(stream));
}

extern "C" int fantom_elem_from_fine_f64(int nelx, int nely, int nelz, const double *G, const double *Efine,
                                         double *Ke, void *stream)
{
    FANTOM_CHECK_ARG(grid_ok(nelx, nely, nelz) && nelz > 0 && G && Efine && Ke, "bad operands");
    return launch_elem_from_fine(make_grid(nelx, nely, nelz), G, Efine, Ke, as_stream(stream));
}
