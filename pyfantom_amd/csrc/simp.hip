// SIMP material interpolation and the elementwise products of the
// sensitivity chain (FP64, one thread per element).
//
// law 0 -- the reference's law (Problem/CPU/MinimumCompliance.py:358-402):
//   h  = heaviside ? (tanh(b*eta) + tanh(b*(rho-eta))) / (tanh(b*eta) + tanh(b*(1-eta))) : rho
//   E  = max(clip(h^p, lo, 1) * hi, lo)          lo = void, hi = E_mul
//   dE = p * h^(p-1) * h' * hi                   (:404-464; the clip is ignored)
// law 1 -- the modified SIMP law E = Emin + h^p (E0 - Emin):
//   E  = lo + h^p * (hi - lo)                    lo = Emin, hi = E0
//   dE = p * h^(p-1) * h' * (hi - lo)
#include "common.hpp"

namespace fantom {

constexpr int kSimpBlock = 256;

__global__ __launch_bounds__(kSimpBlock) void simp_k(long n, const double *__restrict__ rho, int law, double p,
                                                   double lo, double hi, int heaviside, double beta, double eta,
                                                   double *__restrict__ E, double *__restrict__ dE)
{
    const long i = (long)blockIdx.x * kSimpBlock + threadIdx.x;
    if (i >= n)
        return;
    const double r = rho[i];
    double h = r, dh = 1.0;
    if (heaviside) {
        const double den = tanh(beta * eta) + tanh(beta * (1.0 - eta));
        const double t = tanh(beta * (r - eta));
        h = (tanh(beta * eta) + t) / den;
        dh = beta * (1.0 - t * t) / den;
    }
    const double hp = pow(h, p);
    if (E) {
        if (law == 0)
            E[i] = fmax(fmin(fmax(hp, lo), 1.0) * hi, lo);
        else
            E[i] = lo + hp * (hi - lo);
    }
    if (dE) {
        const double g = p * pow(h, p - 1.0) * dh;
        dE[i] = law == 0 ? g * hi : g * (hi - lo);
    }
}

__global__ __launch_bounds__(kSimpBlock) void mul_k(long n, const double *__restrict__ a,
                                                  const double *__restrict__ b, double *__restrict__ out)
{
    const long i = (long)blockIdx.x * kSimpBlock + threadIdx.x;
    if (i < n)
        out[i] = a[i] * b[i];
}

}  // namespace fantom

using namespace fantom;

extern "C" int fantom_simp_f64(long n, const double *rho, int law, double p, double lo, double hi, int heaviside,
                               double beta, double eta, double *E, double *dE, void *stream)
{
    FANTOM_CHECK_ARG(n > 0 && rho && (E || dE), "bad operands");
    FANTOM_CHECK_ARG(law == 0 || law == 1, "unknown SIMP law %d", law);
    const unsigned blocks = (unsigned)((n + kSimpBlock - 1) / kSimpBlock);
    hipLaunchKernelGGL(simp_k, dim3(blocks), dim3(kSimpBlock), 0, as_stream(stream), n, rho, law, p, lo, hi,
                       heaviside, beta, eta, E, dE);
    FANTOM_LAUNCH_CHECK("simp");
    return kStatusOk;
}

extern "C" int fantom_mul_f64(long n, const double *a, const double *b, double *out, void *stream)
{
    FANTOM_CHECK_ARG(n > 0 && a && b && out, "bad operands");
    const unsigned blocks = (unsigned)((n + kSimpBlock - 1) / kSimpBlock);
    hipLaunchKernelGGL(mul_k, dim3(blocks), dim3(kSimpBlock), 0, as_stream(stream), n, a, b, out);
    FANTOM_LAUNCH_CHECK("mul");
    return kStatusOk;
}
