// Element post-processing of a solved displacement field (strain, stress,
// strain energy, von Mises), the structured-mesh case of
// FEA_locals_node_basis_parallel (core/CPU/_ops.py:1030-1058) and of the
// von Mises formulas of MinimumCompliance.FEA
// (Problem/CPU/MinimumCompliance.py:659-736):
//   strain_e = B u_e            (B: the Gauss-point-averaged B operator)
//   stress_e = D strain_e       (D for the unit modulus, as the reference)
//   energy_e = 1/2 u_e^T (w_e K0) u_e
// One thread per element; K0, D and B are staged in LDS.
#include "../../include/fantom.h"
#include "common.hpp"

namespace fantom {

constexpr int kFeaBlock = 256;

template <int DIM>
__global__ __launch_bounds__(kFeaBlock) void fea_locals_k(Grid g, const double *__restrict__ K0g,
                                                        const double *__restrict__ Dg, const double *__restrict__ Bg,
                                                        const double *__restrict__ w, const double *__restrict__ u,
                                                        double *__restrict__ strain, double *__restrict__ stress,
                                                        double *__restrict__ energy, double *__restrict__ von_mises)
{
    constexpr int NL = DIM == 3 ? 24 : 8;    // element DOFs
    constexpr int NS = DIM == 3 ? 6 : 3;     // strain components
    constexpr int NPE = DIM == 3 ? 8 : 4;
    __shared__ double K0[NL * NL], D[NS * NS], B[NS * NL];
    for (int t = threadIdx.x; t < NL * NL; t += kFeaBlock)
        K0[t] = K0g[t];
    for (int t = threadIdx.x; t < NS * NS; t += kFeaBlock)
        D[t] = Dg[t];
    for (int t = threadIdx.x; t < NS * NL; t += kFeaBlock)
        B[t] = Bg[t];
    __syncthreads();
    const long e = (long)blockIdx.x * kFeaBlock + threadIdx.x;
    if (e >= g.n_el)
        return;
    // element DOFs in the reference's local node order (A1 connectivity)
    double ue[NL];
    if (DIM == 3) {
        const int ei = (int)(e / ((long)g.nely * g.nelz));
        const int r = (int)(e % ((long)g.nely * g.nelz));
        const int ej = r / g.nelz, ek = r % g.nelz;
        const long plane = (long)g.nx * g.nz;
#pragma unroll
        for (int c = 0; c < NPE; ++c) {
            const int cx = (c == 1 || c == 2 || c == 5 || c == 6);
            const int cy = (c == 2 || c == 3 || c == 6 || c == 7);
            const int cz = c >= 4;
            const long m = (long)(ej + cy) * plane + (long)(ei + cx) * g.nz + (ek + cz);
#pragma unroll
            for (int d = 0; d < 3; ++d)
                ue[3 * c + d] = u[3 * m + d];
        }
    } else {
        const int ei = (int)(e / g.nely), ej = (int)(e % g.nely);
#pragma unroll
        for (int c = 0; c < NPE; ++c) {
            const int cx = (c == 1 || c == 2), cy = (c >= 2);
            const long m = (long)(ej + cy) * g.nx + (ei + cx);
            ue[2 * c] = u[2 * m];
            ue[2 * c + 1] = u[2 * m + 1];
        }
    }
    double eps[NS], sig[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < NL; ++j)
            s += B[k * NL + j] * ue[j];
        eps[k] = s;
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < NS; ++j)
            s += D[k * NS + j] * eps[j];
        sig[k] = s;
    }
    double en = 0.0;
    const double we = w[e];
#pragma unroll
    for (int r = 0; r < NL; ++r) {
        double t = 0.0;
#pragma unroll
        for (int l = 0; l < NL; ++l)
            t += we * K0[r * NL + l] * ue[l];
        en += 0.5 * ue[r] * t;
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        strain[e * NS + k] = eps[k];
        stress[e * NS + k] = sig[k];
    }
    energy[e] = en;
    double vm;
    if (DIM == 3)
        vm = sqrt(0.5 * ((sig[0] - sig[1]) * (sig[0] - sig[1]) + (sig[1] - sig[2]) * (sig[1] - sig[2]) +
                         (sig[2] - sig[0]) * (sig[2] - sig[0]) +
                         6.0 * (sig[3] * sig[3] + sig[4] * sig[4] + sig[5] * sig[5])));
    else
        vm = sqrt(sig[0] * sig[0] + sig[1] * sig[1] - sig[0] * sig[1] + 3.0 * sig[2] * sig[2]);
    von_mises[e] = vm;
}

}  // namespace fantom

using namespace fantom;

extern "C" {

int fantom_fea_locals_f64(int nelx, int nely, int nelz, const double *K0, const double *D, const double *B,
                          const double *w, const double *u, double *strain, double *stress, double *energy,
                          double *von_mises, void *stream)
{
    FANTOM_CHECK_ARG(grid_ok(nelx, nely, nelz), "invalid grid %d x %d x %d", nelx, nely, nelz);
    FANTOM_CHECK_ARG(K0 && D && B && w && u && strain && stress && energy && von_mises, "null operand");
    const Grid g = make_grid(nelx, nely, nelz);
    const unsigned blocks = (unsigned)((g.n_el + kFeaBlock - 1) / kFeaBlock);
    if (nelz > 0)
        hipLaunchKernelGGL(fea_locals_k<3>, dim3(blocks), dim3(kFeaBlock), 0, as_stream(stream), g, K0, D, B, w, u,
                           strain, stress, energy, von_mises);
    else
        hipLaunchKernelGGL(fea_locals_k<2>, dim3(blocks), dim3(kFeaBlock), 0, as_stream(stream), g, K0, D, B, w, u,
                           strain, stress, energy, von_mises);
    FANTOM_LAUNCH_CHECK("fea_locals");
    return kStatusOk;
}

}  // extern "C"
