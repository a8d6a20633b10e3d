// Explicit CSR assembly of K(E) on a structured grid, with the Dirichlet
// convention of the matrix-free operator: a constrained DOF's row is the unit
// row, its column is dropped from every free row.  This is K(E) @ I as the
// reference forms it for direct solvers (StructuredStiffnessKernel.construct,
// stiffness/CUDA/_FEA.py:484-516, with matmat_node_basis_nnz_per_row_wcon and
// matmat_node_basis_parallel_wcon_kernel, core/CUDA/Kernels/_matmat.py:258-300
// and core/CUDA/Kernels/matmat.cu:54-126).
//
// Two passes, as the reference: the row lengths (free DOFs of the up to 27 / 9
// neighbour nodes that share an element with the row's node; 1 for a
// constrained row), an exclusive scan by the caller, then the fill.  The fill
// gathers instead of scattering: each (row, neighbour node) pair sums E_e K0
// over the elements the two nodes share, and the row's free columns come out
// in ascending order (the reference keeps the order of first visit; the
// matrix is the same).  No atomics, so the values are deterministic.
#include "../../include/fantom.h"
#include "common.hpp"

namespace fantom {

constexpr int kCsrBlock = 256;

// local node number of the corner at offset (cx, cy, cz) of its element, the
// reference's hex8 / quad4 order (core/CPU/_geom.py:4-98)
__device__ __forceinline__ int local_node(int cx, int cy, int cz)
{
    return 4 * cz + (cy ? 3 - cx : cx);
}

template <int DIM>
__device__ __forceinline__ void node_coords(const Grid &g, long p, int &i, int &j, int &k)
{
    if (DIM == 3) {
        const long plane = (long)g.nx * g.nz;
        j = (int)(p / plane);
        const long r = p - (long)j * plane;
        i = (int)(r / g.nz);
        k = (int)(r - (long)i * g.nz);
    } else {
        j = (int)(p / g.nx);
        i = (int)(p - (long)j * g.nx);
        k = 0;
    }
}

template <int DIM>
__global__ __launch_bounds__(kCsrBlock) void csr_row_nnz_k(Grid g, const unsigned char *__restrict__ mask,
                                                         int *__restrict__ counts)
{
    constexpr unsigned kAll = (1u << DIM) - 1;
    const long p = (long)blockIdx.x * kCsrBlock + threadIdx.x;
    if (p >= g.n_nodes)
        return;
    int i, j, k;
    node_coords<DIM>(g, p, i, j, k);
    int cnt = 0;
    for (int dj = -1; dj <= 1; ++dj) {
        const int jj = j + dj;
        if (jj < 0 || jj >= g.ny)
            continue;
        for (int di = -1; di <= 1; ++di) {
            const int ii = i + di;
            if (ii < 0 || ii >= g.nx)
                continue;
            for (int dk = (DIM == 3 ? -1 : 0); dk <= (DIM == 3 ? 1 : 0); ++dk) {
                const int kk = k + dk;
                if (kk < 0 || kk >= g.nz)
                    continue;
                const long q = DIM == 3 ? (long)jj * g.nx * g.nz + (long)ii * g.nz + kk : (long)jj * g.nx + ii;
                const unsigned mq = mask ? mask[q] : 0u;
                cnt += __popc(kAll & ~mq);
            }
        }
    }
    const unsigned mp = mask ? mask[p] : 0u;
#pragma unroll
    for (int d = 0; d < DIM; ++d)
        counts[DIM * p + d] = (mp >> d) & 1u ? 1 : cnt;
}

// Fill: a block takes kCsrRows consecutive rows; one thread per (row,
// neighbour node) pair computes that neighbour's DIM column values from the
// elements the two nodes share and stages them in LDS at their place in the
// row.  The group's rows occupy one contiguous range of the CSR arrays, which
// the block then writes with unit-stride streaming stores.
constexpr int kCsrRows = 16;   // 3D: 432 pairs on 448 threads; 2D: 144 on 192
constexpr int kCsrGroups = 8;  // row groups per block (one K0 staging each)

template <int DIM>
__global__ __launch_bounds__(448) void csr_fill_k(
    Grid g, const double *__restrict__ K0g, const double *__restrict__ E, const unsigned char *__restrict__ mask,
    const long long *__restrict__ indptr, int *__restrict__ indices, double *__restrict__ data)
{
    constexpr int NL = DIM == 3 ? 24 : 8;     // element DOFs
    constexpr int NB = DIM == 3 ? 27 : 9;     // neighbour nodes
    constexpr int ITEMS = kCsrRows * NB;      // (row, neighbour) pairs per block
    constexpr unsigned kAll = (1u << DIM) - 1;
    constexpr int kStage = kCsrRows * NB * DIM;   // staging entries: the longest possible rows
    __shared__ double K0[NL * NL];
    __shared__ double sdata[kCsrRows * NB * DIM];
    __shared__ int sidx[kCsrRows * NB * DIM];
    __shared__ unsigned char scnt[ITEMS];
    const int tid = threadIdx.x;
    for (int t = tid; t < NL * NL; t += blockDim.x)
        K0[t] = K0g[t];
    const long n_rows = (long)DIM * g.n_nodes;
    for (int grp = 0; grp < kCsrGroups; ++grp) {
        const long row0 = ((long)blockIdx.x * kCsrGroups + grp) * kCsrRows;
        if (row0 >= n_rows)
            break;
        const long row_end = min(row0 + kCsrRows, n_rows);
        const long long base = indptr[row0];
        // this thread's pair
        const int r_loc = tid / NB, nb = tid - r_loc * NB;
        const long row = row0 + r_loc;
        const bool active = tid < ITEMS && row < row_end;
        long p = 0, q = 0;
        int d = 0, i = 0, j = 0, k = 0, ii = 0, jj = 0, kk = 0;
        unsigned mp = 0, mq = kAll;  // out-of-grid neighbour: nothing free
        if (active) {
            p = row / DIM;
            d = (int)(row - p * DIM);
            node_coords<DIM>(g, p, i, j, k);
            const int dj = nb / (DIM == 3 ? 9 : 3) - 1;
            const int di = (DIM == 3 ? (nb / 3) % 3 : nb % 3) - 1;
            const int dk = DIM == 3 ? nb % 3 - 1 : 0;
            jj = j + dj;
            ii = i + di;
            kk = k + dk;
            mp = mask ? mask[p] : 0u;
            if (jj >= 0 && jj < g.ny && ii >= 0 && ii < g.nx && kk >= 0 && kk < g.nz) {
                q = DIM == 3 ? (long)jj * g.nx * g.nz + (long)ii * g.nz + kk : (long)jj * g.nx + ii;
                mq = mask ? mask[q] : 0u;
            }
        }
        const bool row_free = !((mp >> d) & 1u);
        if (tid < ITEMS)
            scnt[tid] = active && row_free ? (unsigned char)__popc(kAll & ~mq) : 0;
        __syncthreads();
        if (active) {
            // place of this pair in the block's range
            long long off = indptr[row] - base;
            if (!row_free) {
                if (nb == NB / 2 && (unsigned long long)off < (unsigned long long)kStage) {
                    sidx[off] = (int)row;
                    sdata[off] = 1.0;
                }
            } else if (scnt[tid]) {
                for (int t = r_loc * NB; t < tid; ++t)
                    off += scnt[t];
                double v[DIM];
#pragma unroll
                for (int c = 0; c < DIM; ++c)
                    v[c] = 0.0;
                // elements holding both p and q: lower corner ex in {i-1, i} with
                // ex <= ii <= ex + 1 and 0 <= ex < nelx, likewise in y and z
                for (int ey = max(j, jj) - 1; ey <= min(j, jj); ++ey) {
                    if (ey < 0 || ey >= g.nely)
                        continue;
                    for (int ex = max(i, ii) - 1; ex <= min(i, ii); ++ex) {
                        if (ex < 0 || ex >= g.nelx)
                            continue;
                        for (int ez = (DIM == 3 ? max(k, kk) - 1 : 0); ez <= (DIM == 3 ? min(k, kk) : 0); ++ez) {
                            if (DIM == 3 && (ez < 0 || ez >= g.nelz))
                                continue;
                            const long e = DIM == 3 ? ((long)ex * g.nely + ey) * g.nelz + ez : (long)ex * g.nely + ey;
                            const int lp = local_node(i - ex, j - ey, DIM == 3 ? k - ez : 0);
                            const int lq = local_node(ii - ex, jj - ey, DIM == 3 ? kk - ez : 0);
                            const double w = E[e];
                            const double *kr = K0 + (DIM * lp + d) * NL + DIM * lq;
#pragma unroll
                            for (int c = 0; c < DIM; ++c)
                                v[c] += w * kr[c];
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < DIM; ++c) {
                    if (!((mq >> c) & 1u) && (unsigned long long)off < (unsigned long long)kStage) {
                        sidx[off] = (int)(DIM * q + c);
                        sdata[off] = v[c];
                        ++off;
                    }
                }
            }
        }
        __syncthreads();
        // a row pointer array that does not match the mask cannot push the
        // staging past its end
        const int total = (int)min(indptr[row_end] - base, (long long)kStage);
        for (int t = tid; t < total; t += blockDim.x) {
            __builtin_nontemporal_store(sidx[t], indices + base + t);
            __builtin_nontemporal_store(sdata[t], data + base + t);
        }
        __syncthreads();  // the next group reuses the staging buffers
    }
}

}  // namespace fantom

using namespace fantom;

extern "C" {

int fantom_csr_row_nnz(int nelx, int nely, int nelz, const uint8_t *mask, int *counts, void *stream)
{
    FANTOM_CHECK_ARG(grid_ok(nelx, nely, nelz), "invalid grid %d x %d x %d", nelx, nely, nelz);
    FANTOM_CHECK_ARG(counts, "null operand");
    const Grid g = make_grid(nelx, nely, nelz);
    const unsigned blocks = (unsigned)((g.n_nodes + kCsrBlock - 1) / kCsrBlock);
    if (nelz > 0)
        hipLaunchKernelGGL(csr_row_nnz_k<3>, dim3(blocks), dim3(kCsrBlock), 0, as_stream(stream), g, mask, counts);
    else
        hipLaunchKernelGGL(csr_row_nnz_k<2>, dim3(blocks), dim3(kCsrBlock), 0, as_stream(stream), g, mask, counts);
    FANTOM_LAUNCH_CHECK("csr_row_nnz");
    return kStatusOk;
}

int fantom_csr_assemble_f64(int nelx, int nely, int nelz, const double *K0, const double *E, const uint8_t *mask,
                            const long long *indptr, int *indices, double *data, void *stream)
{
    FANTOM_CHECK_ARG(grid_ok(nelx, nely, nelz), "invalid grid %d x %d x %d", nelx, nely, nelz);
    FANTOM_CHECK_ARG(K0 && E && indptr && indices && data, "null operand");
    const Grid g = make_grid(nelx, nely, nelz);
    const int dim = nelz > 0 ? 3 : 2;
    const long rows_per_block = (long)kCsrRows * kCsrGroups;
    const unsigned blocks = (unsigned)((dim * g.n_nodes + rows_per_block - 1) / rows_per_block);
    if (dim == 3)
        hipLaunchKernelGGL(csr_fill_k<3>, dim3(blocks), dim3(448), 0, as_stream(stream), g, K0, E, mask,
                           indptr, indices, data);
    else
        hipLaunchKernelGGL(csr_fill_k<2>, dim3(blocks), dim3(192), 0, as_stream(stream), g, K0, E, mask,
                           indptr, indices, data);
    FANTOM_LAUNCH_CHECK("csr_assemble");
    return kStatusOk;
}

}  // extern "C"
