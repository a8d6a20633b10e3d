// Structured mesh generation on the device: element connectivity and node
// coordinates in the reference's numbering, for callers that want the
// explicit arrays (the operators index the grid implicitly and never read
// them).  Replaces generate_elements_2d / generate_elements_3d and the
// linspace / meshgrid node positions of generate_structured_mesh_cuda
// (core/CUDA/Kernels/_geom.py:3-118, numbering of core/CPU/_geom.py:4-98):
//   3D element e = (i nely + j) nelz + k, corners
//     n(i,j,k), n(i+1,j,k), n(i+1,j+1,k), n(i,j+1,k), then the same at k+1,
//     with node n(i,j,k) = k + nz i + nz nx j;
//   2D element e = i nely + j, corners n(i,j), n(i+1,j), n(i+1,j+1), n(i,j+1)
//     with node n(i,j) = j nx + i;
//   node coordinates (x_i, y_j[, z_k]) with x_i = i (L / (n - 1)) and the last
//   one exactly L, in the array's type, as numpy's (and CuPy's) linspace forms
//   them from a length of that type.
#include "../../include/fantom.h"
#include "common.hpp"

namespace fantom {

constexpr int kMeshBlock = 256;

template <int DIM>
__global__ __launch_bounds__(kMeshBlock) void mesh_elements_k(Grid g, int *__restrict__ el)
{
    const long e = (long)blockIdx.x * kMeshBlock + threadIdx.x;
    if (e >= g.n_el)
        return;
    if (DIM == 3) {
        const long yz = (long)g.nely * g.nelz;
        const int i = (int)(e / yz);
        const long r = e - (long)i * yz;
        const int j = (int)(r / g.nelz), k = (int)(r - (long)j * g.nelz);
        const int plane = g.nx * g.nz;
        const int b = k + g.nz * i + plane * j;
        const int4 lo = make_int4(b, b + g.nz, b + g.nz + plane, b + plane);
        const int4 hi = make_int4(lo.x + 1, lo.y + 1, lo.z + 1, lo.w + 1);
        int4 *o = reinterpret_cast<int4 *>(el + 8 * e);
        o[0] = lo;
        o[1] = hi;
    } else {
        const int i = (int)(e / g.nely), j = (int)(e - (long)i * g.nely);
        const int b = j * g.nx + i;
        *reinterpret_cast<int4 *>(el + 4 * e) = make_int4(b, b + 1, b + 1 + g.nx, b + g.nx);
    }
}

// linspace(0, L, n)[i] in the array's type: i * step, the last one L
template <typename T>
__device__ __forceinline__ T lin(int i, int n, T L, T step)
{
    return i == n - 1 ? L : (T)i * step;
}

template <typename T, int DIM>
__global__ __launch_bounds__(kMeshBlock) void mesh_nodes_k(Grid g, T lx, T ly, T lz, T sx, T sy, T sz,
                                                         T *__restrict__ X)
{
    const long n = (long)blockIdx.x * kMeshBlock + threadIdx.x;
    if (n >= g.n_nodes)
        return;
    if (DIM == 3) {
        const long plane = (long)g.nx * g.nz;
        const int j = (int)(n / plane);
        const long r = n - (long)j * plane;
        const int i = (int)(r / g.nz), k = (int)(r - (long)i * g.nz);
        X[3 * n] = lin(i, g.nx, lx, sx);
        X[3 * n + 1] = lin(j, g.ny, ly, sy);
        X[3 * n + 2] = lin(k, g.nz, lz, sz);
    } else {
        const int j = (int)(n / g.nx), i = (int)(n - (long)j * g.nx);
        X[2 * n] = lin(i, g.nx, lx, sx);
        X[2 * n + 1] = lin(j, g.ny, ly, sy);
    }
}

template <typename T>
static int mesh_nodes_t(int nelx, int nely, int nelz, double lx, double ly, double lz, T *nodes, void *stream)
{
    FANTOM_CHECK_ARG(grid_ok(nelx, nely, nelz), "invalid grid %d x %d x %d", nelx, nely, nelz);
    FANTOM_CHECK_ARG(nodes, "null operand");
    const Grid g = make_grid(nelx, nely, nelz);
    const unsigned blocks = (unsigned)((g.n_nodes + kMeshBlock - 1) / kMeshBlock);
    // linspace's step (stop - start) / (num - 1), in the array's type
    const T Lx = (T)lx, Ly = (T)ly, Lz = (T)lz;
    const T sx = Lx / (T)nelx, sy = Ly / (T)nely, sz = nelz > 0 ? Lz / (T)nelz : T(0);
    if (nelz > 0)
        hipLaunchKernelGGL((mesh_nodes_k<T, 3>), dim3(blocks), dim3(kMeshBlock), 0, as_stream(stream), g, Lx, Ly, Lz,
                           sx, sy, sz, nodes);
    else
        hipLaunchKernelGGL((mesh_nodes_k<T, 2>), dim3(blocks), dim3(kMeshBlock), 0, as_stream(stream), g, Lx, Ly, Lz,
                           sx, sy, sz, nodes);
    FANTOM_LAUNCH_CHECK("mesh_nodes");
    return kStatusOk;
}

}  // namespace fantom

using namespace fantom;

extern "C" {

int fantom_mesh_elements(int nelx, int nely, int nelz, int32_t *elements, void *stream)
{
    FANTOM_CHECK_ARG(grid_ok(nelx, nely, nelz), "invalid grid %d x %d x %d", nelx, nely, nelz);
    FANTOM_CHECK_ARG(elements, "null operand");
    FANTOM_CHECK_ARG(((uintptr_t)elements & 15) == 0, "elements must be 16-byte aligned");
    const Grid g = make_grid(nelx, nely, nelz);
    const unsigned blocks = (unsigned)((g.n_el + kMeshBlock - 1) / kMeshBlock);
    if (nelz > 0)
        hipLaunchKernelGGL(mesh_elements_k<3>, dim3(blocks), dim3(kMeshBlock), 0, as_stream(stream), g, elements);
    else
        hipLaunchKernelGGL(mesh_elements_k<2>, dim3(blocks), dim3(kMeshBlock), 0, as_stream(stream), g, elements);
    FANTOM_LAUNCH_CHECK("mesh_elements");
    return kStatusOk;
}

int fantom_mesh_nodes_f64(int nelx, int nely, int nelz, double lx, double ly, double lz, double *nodes, void *stream)
{
    return mesh_nodes_t<double>(nelx, nely, nelz, lx, ly, lz, nodes, stream);
}

int fantom_mesh_nodes_f32(int nelx, int nely, int nelz, double lx, double ly, double lz, float *nodes, void *stream)
{
    return mesh_nodes_t<float>(nelx, nely, nelz, lx, ly, lz, nodes, stream);
}

}  // extern "C"
