// Shared helpers for the gfx950 kernels behind the fantom C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <cstring>

namespace fantom {

// Thread-local message of the last failed call, returned by fantom_last_error().
void set_error(const char *fmt, ...);
const char *get_error();

constexpr int kStatusOk = 0;
constexpr int kStatusBadArg = 1;     // maps to ValueError on the Python side
constexpr int kStatusHipError = 2;   // maps to RuntimeError
constexpr int kStatusNotSupported = 3;

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

#define FANTOM_CHECK_ARG(cond, ...)                                                                        \
    do {                                                                                                   \
        if (!(cond)) {                                                                                     \
            ::fantom::set_error(__VA_ARGS__);                                                              \
            return ::fantom::kStatusBadArg;                                                                \
        }                                                                                                  \
    } while (0)

#define FANTOM_LAUNCH_CHECK(where)                                                                         \
    do {                                                                                                   \
        hipError_t _e = hipGetLastError();                                                                 \
        if (_e != hipSuccess) {                                                                            \
            ::fantom::set_error("%s: %s", where, hipGetErrorString(_e));                                  \
            return ::fantom::kStatusHipError;                                                              \
        }                                                                                                  \
    } while (0)

// Structured grid of nelx x nely x nelz elements (nelz = 0 marks 2D).
// omega_scale (optional, device memory): the current value of an adaptive
// Jacobi weight; the smoothing kernels multiply their weight argument by it.
struct Grid {
    int nelx, nely, nelz;
    int nx, ny, nz;  // nodes per axis (nz = 1 in 2D)
    long n_nodes, n_el;
    const double *omega_scale = nullptr;
    // fused zero-guess sweep (MODE_JAC0*): x0 = w D^-1 b followed by one
    // sweep of weight w * w2_ratio (1: both sweeps share the weight)
    double w2_ratio = 1.0;
};

// Jacobi weight of a smoothing kernel: the argument, times the device scalar
// of an adaptive schedule when the grid carries one (a uniform scalar load).
template <typename T>
__device__ __forceinline__ T jacobi_weight(const Grid &g, T omega)
{
    return g.omega_scale ? (T)((double)omega * g.omega_scale[0]) : omega;
}

inline Grid make_grid(int nelx, int nely, int nelz)
{
    Grid g;
    g.nelx = nelx;
    g.nely = nely;
    g.nelz = nelz;
    g.nx = nelx + 1;
    g.ny = nely + 1;
    g.nz = nelz > 0 ? nelz + 1 : 1;
    g.n_nodes = (long)g.nx * g.ny * g.nz;
    g.n_el = (long)nelx * nely * (nelz > 0 ? nelz : 1);
    return g;
}

inline bool grid_ok(int nelx, int nely, int nelz)
{
    if (nelx <= 0 || nely <= 0 || nelz < 0)
        return false;
    // 32-bit DOF indices are used on the device
    long nn = (long)(nelx + 1) * (nely + 1) * (nelz > 0 ? nelz + 1 : 1);
    return nn * 3 < (1L << 31);
}

// Workgroup barrier that orders LDS only.  __syncthreads() is a workgroup
// fence on every address space, which on gfx9 makes each wave wait for all
// of its outstanding global loads and stores (s_waitcnt vmcnt(0)) before the
// barrier, so no prefetch survives it.  Fences restricted to the local
// address space wait for the LDS counter only.  For kernels whose waves
// exchange data through LDS and never through global memory.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Fused multiply-add in the operands' own precision.  __builtin_fma is the
// double routine: called on floats it converts all three operands to double,
// runs the FP64 FMA and rounds back, three extra instructions per FMA.
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Deterministic block sum: wave64 butterfly, then the first wave sums the
// per-wave results in a fixed order.  Returns the sum in thread 0.
template <int BLOCK>
__device__ inline double block_sum(double v)
{
    static_assert(BLOCK % 64 == 0 && BLOCK <= 1024, "block must be a multiple of 64");
    __shared__ double wsum[BLOCK / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        wsum[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w)
            s += wsum[w];
    }
    __syncthreads();
    return s;
}

}  // namespace fantom
