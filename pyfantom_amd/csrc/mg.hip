// Geometric multigrid transfers and the coarsest-level direct solve (FP64).
//
// Restriction / prolongation follow the reference's structured transfer
// (core/CUDA/Kernels/mgm.cu:1-244, core/CPU/_mgm.py:10-297): coarse node c sits
// on fine node 2c; full weighting with weights 2^-dist over the 3^d fine
// neighbourhood.  normalize=1 divides by the boundary-dependent total weight TW
// exactly as the reference does; normalize=0 gives the plain trilinear pair
// P_lin^T / P_lin used by the rediscretised hierarchy (the V-cycle with an
// exact coarse solve is invariant to that diagonal scaling).
#include <rocblas/rocblas.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "fantom_internal.hpp"

namespace fantom {

constexpr int kBlock = 256;

__device__ __forceinline__ double tw3(int b) { return b == 3 ? 3.375 : (b == 2 ? 4.5 : (b == 1 ? 6.0 : 8.0)); }
__device__ __forceinline__ double tw2(int b) { return b == 2 ? 2.25 : (b == 1 ? 3.0 : 4.0); }

// One thread per coarse node; blockIdx.y is the coarse node plane (y), so
// that the in-plane index arithmetic stays 32-bit.  mask_c (optional):
// constrained coarse DOFs get 0.  The 27 fine neighbours are read
// unconditionally (an index outside the grid is clamped to the centre node
// and weighted 0, which leaves every sum unchanged), so that the loads issue
// together instead of one latency after another; the summation order is the
// reference's.
template <typename T, int DOF>
__global__ __launch_bounds__(kBlock) void restrict3d(Grid gf, int normalize, const uint8_t *__restrict__ mask_c,
                                                   const T *__restrict__ rf, T *__restrict__ rc)
{
    const int nxc = gf.nelx / 2 + 1, nyc = gf.nely / 2 + 1, nzc = gf.nelz / 2 + 1;
    const int pc = blockIdx.x * kBlock + threadIdx.x;
    if (pc >= nxc * nzc)
        return;
    const int jc = blockIdx.y;
    const int icx = pc / nzc;
    const int kc = pc - icx * nzc;
    const long ic = (long)jc * nxc * nzc + pc;
    const int bnd = (icx == 0 || icx == nxc - 1) + (jc == 0 || jc == nyc - 1) + (kc == 0 || kc == nzc - 1);
    const long fc = 2 * kc + (long)gf.nz * (2 * icx) + (long)gf.nz * gf.nx * (2 * jc);   // coincident fine node
    double acc[DOF];
#pragma unroll
    for (int d = 0; d < DOF; ++d)
        acc[d] = 0.0;
    // x outer, y, z inner: the reference's summation order (mgm.cu:28-60)
#pragma unroll
    for (int di = -1; di <= 1; ++di) {
        const int ii = 2 * icx + di;
        const bool vi = ii >= 0 && ii < gf.nx;
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj) {
            const int jj = 2 * jc + dj;
            const bool vj = jj >= 0 && jj < gf.ny;
#pragma unroll
            for (int dk = -1; dk <= 1; ++dk) {
                const int kk = 2 * kc + dk;
                const bool v = vi && vj && kk >= 0 && kk < gf.nz;
                const int dist = (di != 0) + (dj != 0) + (dk != 0);
                const double w0 = dist == 0 ? 1.0 : (dist == 1 ? 0.5 : (dist == 2 ? 0.25 : 0.125));
                const double w = v ? w0 : 0.0;
                const long f = v ? kk + (long)gf.nz * ii + (long)gf.nz * gf.nx * jj : fc;
#pragma unroll
                for (int d = 0; d < DOF; ++d)
                    acc[d] += w * (double)rf[f * DOF + d];
            }
        }
    }
    const int mc = mask_c ? mask_c[ic] : 0;
#pragma unroll
    for (int d = 0; d < DOF; ++d) {
        // divide (not multiply by 1/TW) to round like the reference (mgm.cu:66)
        const double v = normalize ? acc[d] / tw3(bnd) : acc[d];
        rc[ic * DOF + d] = ((mc >> d) & 1) ? T(0) : (T)v;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void restrict2d(Grid gf, int dof, int normalize,
                                                   const uint8_t *__restrict__ mask_c, const T *__restrict__ rf,
                                                   T *__restrict__ rc)
{
    const int nxc = gf.nelx / 2 + 1, nyc = gf.nely / 2 + 1;
    const long ncn = (long)nxc * nyc;
    const long ic = (long)blockIdx.x * kBlock + threadIdx.x;
    if (ic >= ncn)
        return;
    const int icx = (int)(ic % nxc), jc = (int)(ic / nxc);
    const int bnd = (icx == 0 || icx == nxc - 1) + (jc == 0 || jc == nyc - 1);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int di = -1; di <= 1; ++di) {
        const int ii = 2 * icx + di;
        if (ii < 0 || ii >= gf.nx)
            continue;
        for (int dj = -1; dj <= 1; ++dj) {
            const int jj = 2 * jc + dj;
            if (jj < 0 || jj >= gf.ny)
                continue;
            const int dist = abs(di) + abs(dj);
            const double w = dist == 0 ? 1.0 : (dist == 1 ? 0.5 : 0.25);
            const long f = (long)jj * gf.nx + ii;
            for (int d = 0; d < dof; ++d)
                acc[d] += w * rf[f * dof + d];
        }
    }
    const int mc = mask_c ? mask_c[ic] : 0;
    for (int d = 0; d < dof; ++d) {
        const double v = normalize ? acc[d] / tw2(bnd) : acc[d];
        rc[ic * dof + d] = ((mc >> d) & 1) ? T(0) : (T)v;
    }
}

// Fine value = average of its parent coarse values (each divided by its TW
// when normalize); accumulate: xf += P ec, else xf = P ec; mask_f (optional):
// constrained fine DOFs receive 0.  prolong_row: the fine nodes 2c + (ox, OY,
// oz) of coarse node c, whose parents are c + o' with o' <= o, from the
// coarse nodes loaded once per thread; each wave writes contiguous runs of
// z-node pairs.  Every fine value sums its parents in the order x outer, y,
// z inner, as one thread per fine node would, so results are unchanged bit
// for bit.
template <typename T, int DOF, int OY>
__device__ __forceinline__ void prolong_row(const Grid &gf, int normalize, const uint8_t *__restrict__ mask_f,
                                            const T *__restrict__ ec, T *__restrict__ xf, int accumulate, int ic,
                                            int jc, int kc)
{
    const int nxc = gf.nelx / 2 + 1, nyc = gf.nely / 2 + 1, nzc = gf.nelz / 2 + 1;
    // parent values, clamped to the grid (a clamped copy is never used)
    double ev[2][OY + 1][2][DOF];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b <= OY; ++b)
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                const int ci = min(ic + a, nxc - 1), cj = min(jc + b, nyc - 1), ck = min(kc + d, nzc - 1);
                const long c = ck + (long)nzc * ci + (long)nzc * nxc * cj;
                const int bnd = (ci == 0 || ci == nxc - 1) + (cj == 0 || cj == nyc - 1) + (ck == 0 || ck == nzc - 1);
                const double den = normalize ? tw3(bnd) : 1.0;
#pragma unroll
                for (int q = 0; q < DOF; ++q) {
                    const double e = (double)ec[c * DOF + q];
                    ev[a][b][d][q] = normalize ? e / den : e;
                }
            }
    // the fine values being accumulated into and the masks, all loaded
    // before the first store (stores to xf would otherwise order every later
    // load of xf behind them)
    const long plane = (long)gf.nx * gf.nz;
    constexpr int oy = OY;
    T xo[2][2][DOF];
    int mo[2][2];
#pragma unroll
    for (int ox = 0; ox < 2; ++ox)
#pragma unroll
        for (int oz = 0; oz < 2; ++oz) {
            const int i = min(2 * ic + ox, gf.nx - 1), j = min(2 * jc + oy, gf.ny - 1), k = min(2 * kc + oz, gf.nz - 1);
            const long f = k + (long)gf.nz * i + plane * j;
#pragma unroll
            for (int q = 0; q < DOF; ++q)
                xo[ox][oz][q] = accumulate ? xf[f * DOF + q] : T(0);
            mo[ox][oz] = mask_f ? mask_f[f] : 0;
        }
#pragma unroll
    for (int ox = 0; ox < 2; ++ox)
#pragma unroll
        for (int oz = 0; oz < 2; ++oz) {
            {
                const int i = 2 * ic + ox, j = 2 * jc + oy, k = 2 * kc + oz;
                if (i >= gf.nx || j >= gf.ny || k >= gf.nz)
                    continue;
                const long f = k + (long)gf.nz * i + plane * j;
                double acc[DOF];
#pragma unroll
                for (int q = 0; q < DOF; ++q)
                    acc[q] = 0.0;
                int count = 0;
#pragma unroll
                for (int a = 0; a <= ox; ++a)
#pragma unroll
                    for (int b = 0; b <= oy; ++b)
#pragma unroll
                        for (int d = 0; d <= oz; ++d) {
                            // a parent past the last coarse node exists only
                            // on grids with an odd element count
                            const bool v = ic + a < nxc && jc + b < nyc && kc + d < nzc;
#pragma unroll
                            for (int q = 0; q < DOF; ++q)
                                acc[q] += v ? ev[a][b][d][q] : 0.0;
                            count += v;
                        }
                // count is 1, 2, 4 or 8: its reciprocal is exact, so the
                // product equals the reference's division
                const double inv_count =
                    count == 8 ? 0.125 : count == 4 ? 0.25 : count == 2 ? 0.5 : count == 1 ? 1.0 : 0.0;
                const int mf = mo[ox][oz];
#pragma unroll
                for (int q = 0; q < DOF; ++q) {
                    double v = acc[q] * inv_count;
                    if ((mf >> q) & 1)
                        v = 0.0;
                    xf[f * DOF + q] = accumulate ? (T)((double)xo[ox][oz][q] + v) : (T)v;
                }
            }
}
}

// One thread per coarse node c and fine plane 2 c_y + oy (blockIdx.y = 2 c_y
// + oy): the plane's up to 4 fine nodes 2c + (ox, oy, oz) from the coarse
// nodes they need (2x1x2 for oy = 0, 2x2x2 for oy = 1); half the registers of
// one thread per whole 2x2x2 block, so more waves cover the memory latency.
template <typename T, int DOF>
__global__ __launch_bounds__(kBlock) void prolong3d(Grid gf, int normalize, const uint8_t *__restrict__ mask_f,
                                                  const T *__restrict__ ec, T *__restrict__ xf, int accumulate)
{
    const int nxc = gf.nelx / 2 + 1, nzc = gf.nelz / 2 + 1;
    const int pc = blockIdx.x * kBlock + threadIdx.x;
    const int jc = blockIdx.y >> 1, oy = blockIdx.y & 1;
    if (pc >= nxc * nzc || 2 * jc + oy >= gf.ny)
        return;
    const int ic = pc / nzc;
    const int kc = pc - ic * nzc;
    if (oy == 0)
        prolong_row<T, DOF, 0>(gf, normalize, mask_f, ec, xf, accumulate, ic, jc, kc);
    else
        prolong_row<T, DOF, 1>(gf, normalize, mask_f, ec, xf, accumulate, ic, jc, kc);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void prolong2d(Grid gf, int dof, int normalize,
                                                  const uint8_t *__restrict__ mask_f, const T *__restrict__ ec,
                                                  T *__restrict__ xf, int accumulate)
{
    const int nxc = gf.nelx / 2 + 1, nyc = gf.nely / 2 + 1;
    const long f = (long)blockIdx.x * kBlock + threadIdx.x;
    if (f >= gf.n_nodes)
        return;
    const int i = (int)(f % gf.nx), j = (int)(f / gf.nx);
    double acc[3] = {0.0, 0.0, 0.0};
    int count = 0;
    for (int ci = i / 2; ci <= i / 2 + (i & 1); ++ci) {
        if (ci >= nxc)
            continue;
        for (int cj = j / 2; cj <= j / 2 + (j & 1); ++cj) {
            if (cj >= nyc)
                continue;
            const int bnd = (ci == 0 || ci == nxc - 1) + (cj == 0 || cj == nyc - 1);
            const long c = (long)cj * nxc + ci;
            for (int d = 0; d < dof; ++d)
                acc[d] += normalize ? (double)ec[c * dof + d] / tw2(bnd) : (double)ec[c * dof + d];
            count++;
        }
    }
    const int mf = mask_f ? mask_f[f] : 0;
    for (int d = 0; d < dof; ++d) {
        double v = count > 0 ? acc[d] / count : 0.0;
        if ((mf >> d) & 1)
            v = 0.0;
        const long r = f * dof + d;
        xf[r] = accumulate ? (T)((double)xf[r] + v) : (T)v;
    }
}

// Coarse element modulus = mean of its 2^d children.
template <typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void coarsen_E(Grid gf, int dim, const TI *__restrict__ Ef, TO *__restrict__ Ec)
{
    const int ncx = gf.nelx / 2, ncy = gf.nely / 2, ncz = dim == 3 ? gf.nelz / 2 : 1;
    const long n = (long)ncx * ncy * ncz;
    const long c = (long)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n)
        return;
    if (dim == 3) {
        const int ci = (int)(c / ((long)ncy * ncz));
        const int r = (int)(c % ((long)ncy * ncz));
        const int cj = r / ncz, ck = r % ncz;
        double s = 0.0;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int d = 0; d < 2; ++d)
                    s += (double)Ef[((long)(2 * ci + a) * gf.nely + (2 * cj + b)) * gf.nelz + (2 * ck + d)];
        Ec[c] = (TO)(s * 0.125);
    } else {
        const int ci = (int)(c / ncy), cj = (int)(c % ncy);
        double s = 0.0;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                s += (double)Ef[(long)(2 * ci + a) * gf.nely + (2 * cj + b)];
        Ec[c] = (TO)(s * 0.25);
    }
}

// A coarse DOF is constrained iff its coincident fine DOF is.
__global__ __launch_bounds__(kBlock) void coarsen_mask_u8(Grid gf, int dim, const uint8_t *__restrict__ mf,
                                                        uint8_t *__restrict__ mc)
{
    const int nxc = gf.nelx / 2 + 1, nyc = gf.nely / 2 + 1, nzc = dim == 3 ? gf.nelz / 2 + 1 : 1;
    const long n = (long)nxc * nyc * nzc;
    const long c = (long)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n)
        return;
    long f;
    if (dim == 3) {
        const int k = (int)(c % nzc);
        const long t = c / nzc;
        const int i = (int)(t % nxc), j = (int)(t / nxc);
        f = 2 * k + (long)gf.nz * 2 * i + (long)gf.nz * gf.nx * 2 * j;
    } else {
        const int i = (int)(c % nxc), j = (int)(c / nxc);
        f = (long)2 * j * gf.nx + 2 * i;
    }
    mc[c] = mf[f];
}

// Dense K(E) with the Dirichlet convention, one thread per row.  A must be
// zero on entry (row-major n_dof x n_dof).
template <typename TE>
__global__ __launch_bounds__(kBlock) void dense_assemble(Grid g, int dim, const double *__restrict__ K0,
                                                       const TE *__restrict__ E, const uint8_t *__restrict__ mask,
                                                       double *__restrict__ A)
{
    const int dof = dim == 3 ? 3 : 2, nl = dim == 3 ? 24 : 8;
    const long ndof = g.n_nodes * dof;
    const long row = (long)blockIdx.x * kBlock + threadIdx.x;
    if (row >= ndof)
        return;
    const long n = row / dof;
    const int c = (int)(row % dof);
    double *Ar = A + row * ndof;
    const int mn = mask ? mask[n] : 0;
    if ((mn >> c) & 1) {
        Ar[row] = 1.0;
        return;
    }
    if (dim == 3) {
        const int k = (int)(n % g.nz);
        const long t = n / g.nz;
        const int i = (int)(t % g.nx), j = (int)(t / g.nx);
        const long plane = (long)g.nx * g.nz;
        for (int oy = 0; oy < 2; ++oy)
            for (int ox = 0; ox < 2; ++ox)
                for (int oz = 0; oz < 2; ++oz) {
                    const int ei = i - ox, ej = j - oy, ek = k - oz;
                    if (ei < 0 || ei >= g.nelx || ej < 0 || ej >= g.nely || ek < 0 || ek >= g.nelz)
                        continue;
                    const double Ee = (double)E[((long)ei * g.nely + ej) * g.nelz + ek];
                    const int a = oz * 4 + (oy ? (ox ? 2 : 3) : (ox ? 1 : 0));
                    for (int cc = 0; cc < 8; ++cc) {
                        const int cx = (cc == 1 || cc == 2 || cc == 5 || cc == 6);
                        const int cy = (cc == 2 || cc == 3 || cc == 6 || cc == 7);
                        const int cz = cc >= 4;
                        const long m = (long)(ej + cy) * plane + (long)(ei + cx) * g.nz + (ek + cz);
                        const int mm = mask ? mask[m] : 0;
                        for (int d = 0; d < 3; ++d) {
                            if ((mm >> d) & 1)
                                continue;
                            Ar[m * 3 + d] += Ee * K0[(a * 3 + c) * nl + cc * 3 + d];
                        }
                    }
                }
    } else {
        const int i = (int)(n % g.nx), j = (int)(n / g.nx);
        for (int oy = 0; oy < 2; ++oy)
            for (int ox = 0; ox < 2; ++ox) {
                const int ei = i - ox, ej = j - oy;
                if (ei < 0 || ei >= g.nelx || ej < 0 || ej >= g.nely)
                    continue;
                const double Ee = (double)E[(long)ei * g.nely + ej];
                const int a = oy ? (ox ? 2 : 3) : (ox ? 1 : 0);
                for (int cc = 0; cc < 4; ++cc) {
                    const int cx = (cc == 1 || cc == 2), cy = cc >= 2;
                    const long m = (long)(ej + cy) * g.nx + (ei + cx);
                    const int mm = mask ? mask[m] : 0;
                    for (int d = 0; d < 2; ++d) {
                        if ((mm >> d) & 1)
                            continue;
                        Ar[m * 2 + d] += Ee * K0[(a * 2 + c) * nl + cc * 2 + d];
                    }
                }
            }
    }
}

// Blocked Gauss-Jordan inversion of an SPD matrix (no pivoting needed).
// Per block of kGJ pivots K, with R the remaining indices,
//   A_KK <- P = A_KK^-1,  A_KR <- P A_KR,  A_RK <- -A_RK P,
//   A_RR <- A_RR - A_RK P A_KR.
// gj_pivot (one workgroup) inverts A_KK in LDS and stores P; gj_update (one
// workgroup per kGJ x kGJ output tile, aligned with the pivot blocks) reads
// one buffer and writes the other, because every workgroup reads the tiles of
// row block K and column block K.
constexpr int kGJ = 32;
constexpr int kGJPer = kGJ * kGJ / 256;

// One wave inverts the pivot block in registers: lane l holds row l % 32,
// columns [16 (l / 32), 16 (l / 32) + 16); the pivot row, pivot column and
// pivot come from other lanes by shuffles.  The step loop is unrolled so that
// every register index is a constant.
__global__ __launch_bounds__(64) void gj_pivot(int n, int k0, const double *__restrict__ A, double *__restrict__ Pout)
{
    const int lane = threadIdx.x, r = lane & 31, h = lane >> 5;
    const int b = min(kGJ, n - k0);
    double v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int c = 16 * h + j;
        v[j] = (r < b && c < b) ? A[(long)(k0 + r) * n + k0 + c] : (r == c ? 1.0 : 0.0);
    }
#pragma unroll
    for (int k = 0; k < kGJ; ++k) {
        if (k < b) {   // uniform: b is the same for every lane
            const int hk = k >> 4, jk = k & 15;
            double rowk[16];
#pragma unroll
            for (int j = 0; j < 16; ++j)
                rowk[j] = __shfl(v[j], k + 32 * h, 64);           // P[k][16h + j]
            const double rk = __shfl(v[jk], r + 32 * hk, 64);     // P[r][k]
            const double pinv = 1.0 / __shfl(v[jk], k + 32 * hk, 64);
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int c = 16 * h + j;
                double o;
                if (r == k)
                    o = c == k ? pinv : rowk[j] * pinv;
                else
                    o = c == k ? -rk * pinv : v[j] - rk * rowk[j] * pinv;
                v[j] = o;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 16; ++j)
        Pout[r * kGJ + 16 * h + j] = v[j];
}

__global__ __launch_bounds__(256) void gj_update(int n, int k0, const double *__restrict__ A,
                                               const double *__restrict__ Pin, double *__restrict__ Out)
{
    __shared__ double P[kGJ][kGJ + 1];
    __shared__ double T[kGJ][kGJ + 1];    // A[K, J]
    __shared__ double X[kGJ][kGJ + 1];    // P A[K, J]
    __shared__ double Rk[kGJ][kGJ + 1];   // A[I, K]
    const int b = min(kGJ, n - k0);
    const int i0 = blockIdx.y * kGJ, j0 = blockIdx.x * kGJ;
    const bool rowK = i0 == k0, colK = j0 == k0;
    const int tid = threadIdx.x;
    for (int q = tid; q < kGJ * kGJ; q += 256) {
        const int r = q / kGJ, c = q % kGJ;
        P[r][c] = Pin[q];
        T[r][c] = (!colK && r < b && j0 + c < n) ? A[(long)(k0 + r) * n + j0 + c] : 0.0;
        Rk[r][c] = (!rowK && i0 + r < n && c < b) ? A[(long)(i0 + r) * n + k0 + c] : 0.0;
    }
    __syncthreads();
    if (!colK) {
#pragma unroll
        for (int m = 0; m < kGJPer; ++m) {
            const int q = tid + 256 * m, r = q / kGJ, c = q % kGJ;
            double x = 0.0;
            for (int t = 0; t < b; ++t)
                x += P[r][t] * T[t][c];
            X[r][c] = x;
        }
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < kGJPer; ++m) {
        const int q = tid + 256 * m, r = q / kGJ, c = q % kGJ;
        const int i = i0 + r, j = j0 + c;
        if (i >= n || j >= n)
            continue;
        double o;
        if (rowK && colK) {
            o = P[r][c];
        } else if (rowK) {
            o = X[r][c];
        } else if (colK) {
            double acc = 0.0;
            for (int t = 0; t < b; ++t)
                acc += Rk[r][t] * P[t][c];
            o = -acc;
        } else {
            double acc = 0.0;
            for (int t = 0; t < b; ++t)
                acc += Rk[r][t] * X[t][c];
            o = A[(long)i * n + j] - acc;
        }
        Out[(long)i * n + j] = o;
    }
}

// y = A x for a dense row-major n x n matrix, A in FP64 or FP32, sums in
// FP64: one workgroup per row, eight partial sums per thread, so that a row
// is read in two trips of eight loads per thread at n = 4131.  (One wave per
// row made eight dependent trips per wave with one round of waves on the
// chip, and read the matrix at 3.5 TB/s.)  Deterministic: fixed partial-sum
// and block-sum order.
template <typename TA, typename T>
__global__ __launch_bounds__(kBlock) void gemv(int n, const TA *__restrict__ A, const T *__restrict__ x,
                                             T *__restrict__ y)
{
    const int row = blockIdx.x;
    const TA *Ar = A + (long)row * n;
    double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int j = threadIdx.x;
    for (; j + 7 * kBlock < n; j += 8 * kBlock) {
#pragma unroll
        for (int u = 0; u < 8; ++u)
            s[u] += (double)Ar[j + kBlock * u] * (double)x[j + kBlock * u];
    }
#pragma unroll
    for (int u = 0; u < 7; ++u)
        if (j + kBlock * u < n)
            s[u] += (double)Ar[j + kBlock * u] * (double)x[j + kBlock * u];
    const double t = block_sum<kBlock>(((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7])));
    if (threadIdx.x == 0)
        y[row] = (T)t;
}

// FP32 copy of a dense matrix (the coarsest inverse applied by the FP32 V-cycle)
__global__ __launch_bounds__(kBlock) void to_f32_k(long n, const double *__restrict__ a, float *__restrict__ b)
{
    for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock)
        b[i] = (float)a[i];
}

static unsigned nblk(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// the 3D transfers launch one workgroup row per node plane (grid y)
static int check_planes(const Grid &gf)
{
    if (gf.ny > 65534 || (long)gf.nx * gf.nz > (1L << 30)) {   // grid.y: two per coarse plane
        set_error("3D transfer: %d node planes of %d x %d nodes exceed the launch limits", gf.ny, gf.nx, gf.nz);
        return kStatusBadArg;
    }
    return kStatusOk;
}

template <typename T>
static int restrict_t(const Grid &gf, int dim, int dof, int normalize, const uint8_t *mask_c, const T *rf, T *rc,
                      hipStream_t s)
{
    if (dim == 3) {
        if (int st = check_planes(gf))
            return st;
        const dim3 grid(nblk((long)(gf.nelx / 2 + 1) * (gf.nelz / 2 + 1)), gf.nely / 2 + 1);
        if (dof == 3)
            hipLaunchKernelGGL((restrict3d<T, 3>), grid, dim3(kBlock), 0, s, gf, normalize, mask_c, rf, rc);
        else if (dof == 1)
            hipLaunchKernelGGL((restrict3d<T, 1>), grid, dim3(kBlock), 0, s, gf, normalize, mask_c, rf, rc);
        else {
            set_error("3D transfers take dof 1 or 3, not %d", dof);
            return kStatusBadArg;
        }
    } else {
        long ncn = (long)(gf.nelx / 2 + 1) * (gf.nely / 2 + 1);
        hipLaunchKernelGGL(restrict2d<T>, dim3(nblk(ncn)), dim3(kBlock), 0, s, gf, dof, normalize, mask_c, rf, rc);
    }
    FANTOM_LAUNCH_CHECK("restrict");
    return kStatusOk;
}

template <typename T>
static int prolong_t(const Grid &gf, int dim, int dof, int normalize, const uint8_t *mask_f, const T *ec, T *xf,
                     int accumulate, hipStream_t s)
{
    if (dim == 3) {
        if (int st = check_planes(gf))
            return st;
        const dim3 grid(nblk((long)(gf.nelx / 2 + 1) * (gf.nelz / 2 + 1)), 2 * (gf.nely / 2 + 1));
        if (dof == 3)
            hipLaunchKernelGGL((prolong3d<T, 3>), grid, dim3(kBlock), 0, s, gf, normalize, mask_f, ec, xf, accumulate);
        else if (dof == 1)
            hipLaunchKernelGGL((prolong3d<T, 1>), grid, dim3(kBlock), 0, s, gf, normalize, mask_f, ec, xf, accumulate);
        else {
            set_error("3D transfers take dof 1 or 3, not %d", dof);
            return kStatusBadArg;
        }
    } else {
        hipLaunchKernelGGL(prolong2d<T>, dim3(nblk(gf.n_nodes)), dim3(kBlock), 0, s, gf, dof, normalize, mask_f, ec,
                           xf, accumulate);
    }
    FANTOM_LAUNCH_CHECK("prolong");
    return kStatusOk;
}

int launch_restrict(const Grid &gf, int dim, int dof, int normalize, const uint8_t *mask_c, const double *rf,
                    double *rc, hipStream_t s)
{
    return restrict_t<double>(gf, dim, dof, normalize, mask_c, rf, rc, s);
}

int launch_restrict_f32(const Grid &gf, int dim, int dof, int normalize, const uint8_t *mask_c, const float *rf,
                        float *rc, hipStream_t s)
{
    return restrict_t<float>(gf, dim, dof, normalize, mask_c, rf, rc, s);
}

int launch_prolong(const Grid &gf, int dim, int dof, int normalize, const uint8_t *mask_f, const double *ec,
                   double *xf, int accumulate, hipStream_t s)
{
    return prolong_t<double>(gf, dim, dof, normalize, mask_f, ec, xf, accumulate, s);
}

int launch_prolong_f32(const Grid &gf, int dim, int dof, int normalize, const uint8_t *mask_f, const float *ec,
                       float *xf, int accumulate, hipStream_t s)
{
    return prolong_t<float>(gf, dim, dof, normalize, mask_f, ec, xf, accumulate, s);
}

template <typename TI, typename TO>
static int coarsen_E_t(const Grid &gf, int dim, const TI *Ef, TO *Ec, hipStream_t s)
{
    long n = (long)(gf.nelx / 2) * (gf.nely / 2) * (dim == 3 ? gf.nelz / 2 : 1);
    hipLaunchKernelGGL((coarsen_E<TI, TO>), dim3(nblk(n)), dim3(kBlock), 0, s, gf, dim, Ef, Ec);
    FANTOM_LAUNCH_CHECK("coarsen_E");
    return kStatusOk;
}

int launch_coarsen_E(const Grid &gf, int dim, const double *Ef, double *Ec, hipStream_t s)
{
    return coarsen_E_t<double, double>(gf, dim, Ef, Ec, s);
}

int launch_coarsen_E_f32(const Grid &gf, int dim, const float *Ef, float *Ec, hipStream_t s)
{
    return coarsen_E_t<float, float>(gf, dim, Ef, Ec, s);
}

int launch_coarsen_mask(const Grid &gf, int dim, const uint8_t *mf, uint8_t *mc, hipStream_t s)
{
    long n = (long)(gf.nelx / 2 + 1) * (gf.nely / 2 + 1) * (dim == 3 ? gf.nelz / 2 + 1 : 1);
    hipLaunchKernelGGL(coarsen_mask_u8, dim3(nblk(n)), dim3(kBlock), 0, s, gf, dim, mf, mc);
    FANTOM_LAUNCH_CHECK("coarsen_mask");
    return kStatusOk;
}

template <typename TE>
static int dense_assemble_t(const Grid &g, int dim, const double *K0, const TE *E, const uint8_t *mask, double *A,
                            hipStream_t s)
{
    const int dof = dim == 3 ? 3 : 2;
    const long ndof = g.n_nodes * dof;
    hipError_t e = hipMemsetAsync(A, 0, sizeof(double) * ndof * ndof, s);
    if (e != hipSuccess) {
        set_error("dense_assemble memset: %s", hipGetErrorString(e));
        return kStatusHipError;
    }
    hipLaunchKernelGGL(dense_assemble<TE>, dim3(nblk(ndof)), dim3(kBlock), 0, s, g, dim, K0, E, mask, A);
    FANTOM_LAUNCH_CHECK("dense_assemble");
    return kStatusOk;
}

int launch_dense_assemble(const Grid &g, int dim, const double *K0, const double *E, const uint8_t *mask,
                          double *A, hipStream_t s)
{
    return dense_assemble_t<double>(g, dim, K0, E, mask, A, s);
}

int launch_dense_assemble_f32E(const Grid &g, const double *K0, const float *E, const uint8_t *mask, double *A,
                               hipStream_t s)
{
    return dense_assemble_t<float>(g, 3, K0, E, mask, A, s);
}

// ---- large-n inverse: 128-wide pivot blocks, trailing update as one GEMM ----------
// With K the pivot block and P = A_KK^-1, one Gauss-Jordan block step is
//   A <- Ahat - Xhat Rhat,
// Ahat = A with block row K and block column K set to zero, Xhat (n x b) =
// A_{:,K} with rows K replaced by -I, Rhat (b x n) = P A_{K,:} with columns K
// replaced by P.  Rows K then hold P A_KR (and P), column K holds -A_RK P and
// the rest A_RR - A_RK P A_KR, the same step as gj_pivot/gj_update above; the
// product is a plain dense GEMM, left to rocBLAS.  Each step streams the
// matrix once, so 128-wide blocks cost a quarter of the 32-wide passes.
constexpr int kGJB = 128;

// One 1024-thread workgroup inverts the b x b pivot block (b <= 128) in
// registers: thread t holds row t / 8, columns c = t % 8 + 8 j (j < 16; the
// interleave keeps the LDS reads of a pivot row free of bank conflicts).
// Each step eliminates four pivots K4 at once: with Rt the pivot rows
// A_{K4,:} whose K4 columns are replaced by I, D = A_{K4,K4} (SPD: a
// principal block of a Schur complement of an SPD matrix) and the working
// values of the pivot rows and columns set to zero, every entry takes
// v += g . Rt_c (four FMAs), with g = (D^-1)_{r - k, :} in a pivot row and
// -A_{r,K4} D^-1 elsewhere.  Every thread inverts the 4 x 4 block itself; Rt
// and A_{:,K4} go through double-buffered LDS, one barrier per step.
constexpr int kPivThreads = 1024;
constexpr int kPivGroups = 8;                          // threads per row
constexpr int kPivCols = kGJB / kPivGroups;            // 16 columns per thread

__global__ __launch_bounds__(kPivThreads) void gj_pivot128(int n, int k0, int b, const double *__restrict__ A,
                                                          double *__restrict__ P)
{
    __shared__ double rowb[2][4][kGJB];
    __shared__ double colb[2][kGJB][4];
    const int t = threadIdx.x, r = t / kPivGroups, cg = t % kPivGroups;
    double v[kPivCols];
#pragma unroll
    for (int j = 0; j < kPivCols; ++j) {
        const int c = cg + kPivGroups * j;
        v[j] = (r < b && c < b) ? A[(long)(k0 + r) * n + k0 + c] : (r == c ? 1.0 : 0.0);
    }
#pragma unroll 1
    for (int q = 0; q < kGJB / 4; ++q) {
        const int k = 4 * q, buf = q & 1;
        // pivot column k + i is column cg = k % 8 + i at slot jq of its threads
        const int jq = k / kPivGroups, ic = cg - k % kPivGroups;
        const bool pcol = ic >= 0 && ic < 4;
        const bool prow = (r >> 2) == q;
        if (pcol) {
            double x = 0.0;
#pragma unroll
            for (int j = 0; j < kPivCols; ++j)
                x = j == jq ? v[j] : x;
            colb[buf][r][ic] = x;
#pragma unroll
            for (int j = 0; j < kPivCols; ++j)
                v[j] = j == jq ? 0.0 : v[j];
        }
        if (prow) {
            const int i = r - k;
#pragma unroll
            for (int j = 0; j < kPivCols; ++j) {
                // the K4 columns of Rt are the identity
                rowb[buf][i][cg + kPivGroups * j] = v[j] + ((pcol && j == jq && ic == i) ? 1.0 : 0.0);
                v[j] = 0.0;
            }
        }
        __syncthreads();
        double d[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int i2 = 0; i2 < 4; ++i2)
                d[i][i2] = colb[buf][k + i][i2];
#pragma unroll
        for (int m = 0; m < 4; ++m) {   // in-place Gauss-Jordan inverse of D
            const double pm = 1.0 / d[m][m];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (j != m)
                    d[m][j] *= pm;
            d[m][m] = pm;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (i == m)
                    continue;
                const double f = d[i][m];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (j != m)
                        d[i][j] -= f * d[m][j];
                d[i][m] = -f * pm;
            }
        }
        double g[4];
        if (prow) {
#pragma unroll
            for (int i2 = 0; i2 < 4; ++i2)
                g[i2] = (r - k == 0) ? d[0][i2] : (r - k == 1) ? d[1][i2] : (r - k == 2) ? d[2][i2] : d[3][i2];
        } else {
            double a[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                a[i] = colb[buf][r][i];
#pragma unroll
            for (int i2 = 0; i2 < 4; ++i2)
                g[i2] = -(a[0] * d[0][i2] + a[1] * d[1][i2] + a[2] * d[2][i2] + a[3] * d[3][i2]);
        }
        // compiler-only fences keep the row loads of one rank-1 part at a
        // time in registers (hoisting all of them spills)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            asm volatile("" ::: "memory");
#pragma unroll
            for (int j = 0; j < kPivCols; ++j)
                v[j] = fma(g[i], rowb[buf][i][cg + kPivGroups * j], v[j]);
        }
        asm volatile("" ::: "memory");
    }
    if (r < b) {
#pragma unroll
        for (int j = 0; j < kPivCols; ++j) {
            const int c = cg + kPivGroups * j;
            if (c < b)
                P[r * kGJB + c] = v[j];
        }
    }
}

// Xhat (n x b, row stride b) and the P columns of Rhat (after the GEMM that
// forms P A_{K,:})
__global__ __launch_bounds__(256) void gj_panel_x(int n, int k0, int b, const double *__restrict__ A,
                                                 const double *__restrict__ P, double *__restrict__ X,
                                                 double *__restrict__ R)
{
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q < (long)n * b) {
        const int i = (int)(q / b), c = (int)(q % b);
        const bool inK = i >= k0 && i < k0 + b;
        X[q] = inK ? (i - k0 == c ? -1.0 : 0.0) : A[(long)i * n + k0 + c];
    }
    if (q < (long)b * b) {   // Rhat[t][k0 + c] = P[t][c]
        const int t = (int)(q / b), c = (int)(q % b);
        R[(long)t * n + k0 + c] = P[t * kGJB + c];
    }
}

// gj_panel_x, gj_zero_k and the copy of the next pivot block (rows and
// columns K1, b1 wide, into T with row stride kGJB) in one pass.  Thread q
// reads entry (q / b, k0 + q % b) of the column block for Xhat before it
// zeroes that same entry, and the K1 block lies outside block row and column
// K, so the three parts do not race.
__global__ __launch_bounds__(256) void gj_panel_zero_copy(int n, int k0, int b, int k1, int b1,
                                                         double *__restrict__ A, const double *__restrict__ P,
                                                         double *__restrict__ X, double *__restrict__ R,
                                                         double *__restrict__ T)
{
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q < (long)n * b) {
        const int i = (int)(q / b), c = (int)(q % b);
        const bool inK = i >= k0 && i < k0 + b;
        const long ac = (long)i * n + k0 + c;
        X[q] = inK ? (i - k0 == c ? -1.0 : 0.0) : A[ac];
        A[ac] = 0.0;
        A[(long)k0 * n + q] = 0.0;
    }
    if (q < (long)b * b) {
        const int t = (int)(q / b), c = (int)(q % b);
        R[(long)t * n + k0 + c] = P[t * kGJB + c];
    }
    if (q < (long)b1 * b1) {
        const int r = (int)(q / b1), c = (int)(q % b1);
        T[r * kGJB + c] = A[(long)(k1 + r) * n + k1 + c];
    }
}

// Ahat: block row K and block column K of A set to zero
__global__ __launch_bounds__(256) void gj_zero_k(int n, int k0, int b, double *__restrict__ A)
{
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q < (long)n * b) {
        A[(q / b) * n + k0 + q % b] = 0.0;   // column block K: b consecutive entries per row
        A[(long)k0 * n + q] = 0.0;           // row block K: b contiguous rows
    }
}

static std::mutex g_blas_mu;
static rocblas_handle g_blas[64];

// Look-ahead streams of the blocked inverse, per device: the pivot-block
// inversions run on CUs of their own (a one-workgroup, latency-bound kernel
// that slows to twice its time when it shares a CU with the GEMM's
// workgroups), the GEMMs and the small kernels on all other CUs.  Default
// priority, like every stream here.
struct GjStreams {
    bool tried = false, ok = false;
    hipStream_t gemm = nullptr, piv = nullptr;
    hipEvent_t fork = nullptr, block = nullptr, pivot = nullptr, done = nullptr;
};
static GjStreams g_gj[64];

static bool gj_streams(int dev, GjStreams &gs)
{
    if (gs.tried)
        return gs.ok;
    gs.tried = true;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 2)
        return false;
    // logical CUs 0-7 are one CU on each of the 8 XCDs (a queue needs a CU
    // on every XCD, which dispatches round-robin); the pivot stream gets
    // those, the GEMM stream the rest
    constexpr int kPivCUs = 8;
    if (ncu <= 2 * kPivCUs)
        return false;
    const int words = (ncu + 31) / 32;
    std::vector<uint32_t> rest(words, 0u), piv(words, 0u);
    for (int c = 0; c < ncu; ++c)
        (c < kPivCUs ? piv : rest)[c / 32] |= 1u << (c % 32);
    gs.ok = hipExtStreamCreateWithCUMask(&gs.gemm, (uint32_t)words, rest.data()) == hipSuccess &&
            hipExtStreamCreateWithCUMask(&gs.piv, (uint32_t)words, piv.data()) == hipSuccess &&
            hipEventCreateWithFlags(&gs.fork, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&gs.block, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&gs.pivot, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&gs.done, hipEventDisableTiming) == hipSuccess;
    (void)hipGetLastError();   // a refused CU mask leaves the single-stream path
    return gs.ok;
}

#define GJ_HIP(expr, what)                                                                                 \
    do {                                                                                                  \
        const hipError_t _e = (expr);                                                                     \
        if (_e != hipSuccess) {                                                                           \
            set_error("spd_inverse: %s: %s", what, hipGetErrorString(_e));                                \
            return kStatusHipError;                                                                       \
        }                                                                                                 \
    } while (0)
#define GJ_BLAS(expr, what)                                                                                \
    do {                                                                                                  \
        if ((expr) != rocblas_status_success) {                                                           \
            set_error("spd_inverse: %s failed", what);                                                    \
            return kStatusHipError;                                                                       \
        }                                                                                                 \
    } while (0)

// The blocked inverse with look-ahead.  The next pivot inversion reads only
// the next pivot block, and A keeps no other use of that block (the next
// step replaces it by its inverse), so each step copies that block to T
// before the full rank-b update, and the pivot stream forms its update (a
// b x b x b product) and inverts it while the full update runs on the other
// CUs.
static int spd_inverse_lookahead(int n, double *A, double *work, hipStream_t s, rocblas_handle hb,
                                 const GjStreams &gs)
{
    double *X = work, *R = work + (size_t)n * kGJB, *P = R + (size_t)n * kGJB, *T = P + (size_t)kGJB * kGJB;
    const double one = 1.0, minus_one = -1.0, zero = 0.0;
    hipStream_t sg = gs.gemm, sp = gs.piv;
    GJ_BLAS(rocblas_set_stream(hb, sg), "rocblas stream");
    GJ_BLAS(rocblas_set_pointer_mode(hb, rocblas_pointer_mode_host), "rocblas pointer mode");
    GJ_HIP(hipEventRecord(gs.fork, s), "fork");
    GJ_HIP(hipStreamWaitEvent(sg, gs.fork, 0), "fork");
    GJ_HIP(hipStreamWaitEvent(sp, gs.fork, 0), "fork");
    hipLaunchKernelGGL(gj_pivot128, dim3(1), dim3(kPivThreads), 0, sp, n, 0, std::min(kGJB, n), A, P);
    FANTOM_LAUNCH_CHECK("spd_inverse pivot");
    GJ_HIP(hipEventRecord(gs.pivot, sp), "pivot event");
    for (int k0 = 0; k0 < n; k0 += kGJB) {
        const int b = std::min(kGJB, n - k0);
        const long nb = (long)n * b;
        GJ_HIP(hipStreamWaitEvent(sg, gs.pivot, 0), "wait for the pivot");
        // Rhat = P A_{K,:} (see spd_inverse_blocked for the column-major mapping)
        GJ_BLAS(rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, n, b, b, &one, A + (size_t)k0 * n, n,
                              P, kGJB, &zero, R, n),
                "rocblas_dgemm (panel)");
        const int k1 = k0 + b, b1 = std::max(0, std::min(kGJB, n - k1));
        hipLaunchKernelGGL(gj_panel_zero_copy, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, sg, n, k0, b, k1, b1,
                           A, P, X, R, T);
        FANTOM_LAUNCH_CHECK("spd_inverse block step");
        if (b1 > 0) {
            // on the pivot stream: T -= Xhat_{K1,:} Rhat_{:,K1} (T held
            // A_{K1,K1}, copied before the full update below changes it), then
            // the inversion of T
            GJ_HIP(hipEventRecord(gs.block, sg), "block event");
            GJ_HIP(hipStreamWaitEvent(sp, gs.block, 0), "wait for the block");
            GJ_BLAS(rocblas_set_stream(hb, sp), "rocblas stream");
            GJ_BLAS(rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, b1, b1, b, &minus_one, R + k1, n,
                                  X + (size_t)k1 * b, b, &one, T, kGJB),
                    "rocblas_dgemm (pivot block)");
            GJ_BLAS(rocblas_set_stream(hb, sg), "rocblas stream");
            hipLaunchKernelGGL(gj_pivot128, dim3(1), dim3(kPivThreads), 0, sp, kGJB, 0, b1, T, P);
            FANTOM_LAUNCH_CHECK("spd_inverse pivot");
            GJ_HIP(hipEventRecord(gs.pivot, sp), "pivot event");
        }
        GJ_BLAS(rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, n, n, b, &minus_one, R, n, X, b, &one,
                              A, n),
                "rocblas_dgemm");
    }
    GJ_HIP(hipEventRecord(gs.done, sg), "join");
    GJ_HIP(hipStreamWaitEvent(s, gs.done, 0), "join");
    return kStatusOk;
}

static int spd_inverse_blocked(int n, double *A, double *work, hipStream_t s)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        set_error("spd_inverse: no current device");
        return kStatusHipError;
    }
    std::lock_guard<std::mutex> lock(g_blas_mu);
    if (!g_blas[dev] && rocblas_create_handle(&g_blas[dev]) != rocblas_status_success) {
        g_blas[dev] = nullptr;
        set_error("spd_inverse: rocblas_create_handle failed");
        return kStatusHipError;
    }
    rocblas_handle hb = g_blas[dev];
    GjStreams &gs = g_gj[dev];
    // the look-ahead pays once a rank-128 update outlasts a pivot inversion
    if (n > 16 * kGJB && gj_streams(dev, gs))
        return spd_inverse_lookahead(n, A, work, s, hb, gs);
    if (rocblas_set_stream(hb, s) != rocblas_status_success ||
        rocblas_set_pointer_mode(hb, rocblas_pointer_mode_host) != rocblas_status_success) {
        set_error("spd_inverse: rocblas stream / pointer mode");
        return kStatusHipError;
    }
    double *X = work, *R = work + (size_t)n * kGJB, *P = R + (size_t)n * kGJB;
    const double one = 1.0, minus_one = -1.0;
    for (int k0 = 0; k0 < n; k0 += kGJB) {
        const int b = std::min(kGJB, n - k0);
        const long nb = (long)n * b;
        hipLaunchKernelGGL(gj_pivot128, dim3(1), dim3(kPivThreads), 0, s, n, k0, b, A, P);
        FANTOM_LAUNCH_CHECK("spd_inverse pivot");
        // Rhat = P A_{K,:}: row-major b x n, i.e. Rhat^T = A_{K,:}^T P^T with
        // A_{K,:}^T (n x b, ld n) and P^T (b x b, ld kGJB) column-major
        static const double zero = 0.0;
        if (rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, n, b, b, &one, A + (size_t)k0 * n, n,
                          P, kGJB, &zero, R, n) != rocblas_status_success) {
            set_error("spd_inverse: rocblas_dgemm (panel) failed");
            return kStatusHipError;
        }
        hipLaunchKernelGGL(gj_panel_x, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, n, k0, b, A, P, X, R);
        hipLaunchKernelGGL(gj_zero_k, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, n, k0, b, A);
        FANTOM_LAUNCH_CHECK("spd_inverse block step");
        // row-major A (n x n) -= Xhat (n x b) Rhat (b x n), i.e. in rocBLAS's
        // column-major terms A^T -= Rhat^T Xhat^T with Rhat^T (n x b, ld n)
        // and Xhat^T (b x n, ld b)
        if (rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, n, n, b, &minus_one, R, n, X, b, &one,
                          A, n) != rocblas_status_success) {
            set_error("spd_inverse: rocblas_dgemm failed");
            return kStatusHipError;
        }
    }
    return kStatusOk;
}

int launch_spd_inverse(int n, double *A, double *work, hipStream_t s)
{
    // work: n*n doubles + kGJ*kGJ; the blocked path uses 2 n kGJB + 2 kGJB^2 of it
    if (n > 4 * kGJB)
        return spd_inverse_blocked(n, A, work, s);
    const unsigned nt = (unsigned)((n + kGJ - 1) / kGJ);
    double *src = A, *dst = work, *pinv = work + (size_t)n * n;
    for (unsigned kb = 0; kb < nt; ++kb) {
        hipLaunchKernelGGL(gj_pivot, dim3(1), dim3(64), 0, s, n, (int)(kb * kGJ), src, pinv);
        hipLaunchKernelGGL(gj_update, dim3(nt, nt), dim3(256), 0, s, n, (int)(kb * kGJ), src, pinv, dst);
        double *t = src;
        src = dst;
        dst = t;
    }
    FANTOM_LAUNCH_CHECK("spd_inverse");
    if (src != A) {
        const hipError_t e = hipMemcpyAsync(A, src, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) {
            set_error("spd_inverse copy: %s", hipGetErrorString(e));
            return kStatusHipError;
        }
    }
    return kStatusOk;
}

int launch_gemv(int n, const double *A, const double *x, double *y, hipStream_t s)
{
    hipLaunchKernelGGL((gemv<double, double>), dim3((unsigned)n), dim3(kBlock), 0, s, n, A, x, y);
    FANTOM_LAUNCH_CHECK("gemv");
    return kStatusOk;
}

int launch_gemv_f32(int n, const float *A, const float *x, float *y, hipStream_t s)
{
    hipLaunchKernelGGL((gemv<float, float>), dim3((unsigned)n), dim3(kBlock), 0, s, n, A, x, y);
    FANTOM_LAUNCH_CHECK("gemv");
    return kStatusOk;
}

int launch_to_f32(long n, const double *a, float *b, hipStream_t s)
{
    hipLaunchKernelGGL(to_f32_k, dim3((unsigned)std::min<long>(nblk(n), 4096)), dim3(kBlock), 0, s, n, a, b);
    FANTOM_LAUNCH_CHECK("to_f32");
    return kStatusOk;
}

}  // namespace fantom
