// Structured density filter (FP64), one thread per element.
//
// Forward (core/CUDA/Kernels/filters.cu:29-58, core/CPU/_filter.py:96-124):
//   (Hx)_i = sum_k w_k x_{i+o_k} / sum_k w_k   over in-bounds taps
// Adjoint (filters.cu:130-154):
//   (H^T y)_i = sum_k (w_k / norm_{i+o_k}) y_{i+o_k}
// with norm_j the in-bounds tap sum of element j (filters.cu:60-105).  The
// taps are symmetric (o and -o carry the same weight), so the adjoint is the
// exact transpose.  Taps are summed in the order given, as the reference does.
// Element order: 3D e = (ix*nely + iy)*nelz + iz, 2D e = ix*nely + iy.
#include "common.hpp"
#include "fantom_internal.hpp"
#include "vec.hpp"

namespace fantom {

constexpr int kFiltBlock = 256;
constexpr int kMaxTaps = 2048;

template <int DIM>
__device__ __forceinline__ void el_pos(long e, int nely, int nelz, int *p)
{
    if (DIM == 3) {
        p[0] = (int)(e / ((long)nely * nelz));
        const int r = (int)(e % ((long)nely * nelz));
        p[1] = r / nelz;
        p[2] = r % nelz;
    } else {
        p[0] = (int)(e / nely);
        p[1] = (int)(e % nely);
        p[2] = 0;
    }
}

// MODE 0: normalizer, 1: forward, 2: adjoint
template <int DIM, int MODE>
__global__ __launch_bounds__(kFiltBlock) void filter_kernel(int nelx, int nely, int nelz, int ntap,
                                                          const int32_t *__restrict__ offsets_g,
                                                          const double *__restrict__ weights_g,
                                                          const double *__restrict__ norm,
                                                          const double *__restrict__ vin,
                                                          double *__restrict__ vout)
{
    __shared__ int32_t off[kMaxTaps * 3];
    __shared__ double wts[kMaxTaps];
    for (int t = threadIdx.x; t < ntap * DIM; t += kFiltBlock)
        off[t] = offsets_g[t];
    for (int t = threadIdx.x; t < ntap; t += kFiltBlock)
        wts[t] = weights_g[t];
    __syncthreads();
    const long n_el = (long)nelx * nely * (DIM == 3 ? nelz : 1);
    const long e = (long)blockIdx.x * kFiltBlock + threadIdx.x;
    if (e >= n_el)
        return;
    int p[3];
    el_pos<DIM>(e, nely, nelz, p);
    double sv = 0.0, sw = 0.0;
    for (int t = 0; t < ntap; ++t) {
        const int qx = p[0] + off[t * DIM], qy = p[1] + off[t * DIM + 1];
        const int qz = DIM == 3 ? p[2] + off[t * DIM + 2] : 0;
        if (qx < 0 || qx >= nelx || qy < 0 || qy >= nely)
            continue;
        if (DIM == 3 && (qz < 0 || qz >= nelz))
            continue;
        const long j = DIM == 3 ? ((long)qx * nely + qy) * nelz + qz : (long)qx * nely + qy;
        const double w = wts[t];
        if (MODE == 0) {
            sw += w;
        } else if (MODE == 1) {
            sv += w * vin[j];
            sw += w;
        } else {
            sv += (w / norm[j]) * vin[j];
        }
    }
    if (MODE == 0)
        vout[e] = sw;
    else if (MODE == 1)
        vout[e] = sw > 0 ? sv / sw : 0.0;
    else
        vout[e] = sv;
}

template <int MODE>
static int launch_filter(int dim, int nelx, int nely, int nelz, int ntap, const int32_t *offsets,
                         const double *weights, const double *norm, const double *vin, double *vout, hipStream_t s)
{
    if (ntap <= 0 || ntap > kMaxTaps) {
        set_error("filter: tap count %d outside [1, %d]", ntap, kMaxTaps);
        return kStatusBadArg;
    }
    const long n_el = (long)nelx * nely * (dim == 3 ? nelz : 1);
    const unsigned blocks = (unsigned)((n_el + kFiltBlock - 1) / kFiltBlock);
    if (dim == 3)
        hipLaunchKernelGGL((filter_kernel<3, MODE>), dim3(blocks), dim3(kFiltBlock), 0, s, nelx, nely, nelz, ntap,
                           offsets, weights, norm, vin, vout);
    else
        hipLaunchKernelGGL((filter_kernel<2, MODE>), dim3(blocks), dim3(kFiltBlock), 0, s, nelx, nely, nelz, ntap,
                           offsets, weights, norm, vin, vout);
    FANTOM_LAUNCH_CHECK("filter");
    return kStatusOk;
}

int launch_filter_normalizer(int dim, int nelx, int nely, int nelz, int ntap, const int32_t *offsets,
                             const double *weights, double *norm, hipStream_t s)
{
    return launch_filter<0>(dim, nelx, nely, nelz, ntap, offsets, weights, nullptr, nullptr, norm, s);
}

int launch_filter_fwd(int dim, int nelx, int nely, int nelz, int ntap, const int32_t *offsets,
                      const double *weights, const double *x, double *y, hipStream_t s)
{
    return launch_filter<1>(dim, nelx, nely, nelz, ntap, offsets, weights, nullptr, x, y, s);
}

int launch_filter_adj(int dim, int nelx, int nely, int nelz, int ntap, const int32_t *offsets,
                      const double *weights, const double *norm, const double *y, double *x, hipStream_t s)
{
    return launch_filter<2>(dim, nelx, nely, nelz, ntap, offsets, weights, norm, y, x, s);
}

}  // namespace fantom
