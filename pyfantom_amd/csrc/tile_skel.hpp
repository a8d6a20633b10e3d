// Tile skeleton shared by the matrix-free hex8 operators (fine K(E)u in
// ku_tile.hip, level-1 Galerkin operator in galerkin.hip).
//
// Both operators act on an element in the Walsh-Hadamard (WHT) basis of its
// eight corners: w[c][s] = sum_b (-1)^popc(b & s) u[b][c] with the natural
// corner index b = bx + 2 by + 4 bz.  The transform is a tensor product of
// 2-point butterflies along z, x and y, and the skeleton exploits that the
// first two stages are shared between neighbouring elements:
//
//   * a workgroup of 16 x 16 threads owns a 16 (x) x 16 (z) block of node
//     columns and walks a chunk of node planes along y (the slowest node
//     axis).  Lane layout: lz = lane % 16 (one DPP row), lx = tid / 16.
//   * per plane, each thread loads its own node (3 doubles + mask byte),
//     takes the z-butterfly with the node at lz + 1 by a DPP row shift and the
//     x-butterfly with the column lx + 1 through LDS: the "face transform"
//     F[c][sx + 2 sz] of the x-z face whose low corner is the thread's node.
//     A face transform is used by the element layers below and above it.
//   * per element layer the y-butterfly of the two face transforms gives the
//     element's 24 coefficients, the operator maps them to 24 output
//     coefficients, and the inverse y-butterfly splits them into the lower
//     (this plane) and upper (next plane) face parts; the upper part waits in
//     registers for the next layer.
//   * the inverse x/z butterflies and the sum over the four faces around a
//     node use the mirror-image moves: DPP row shift from lz - 1, LDS from
//     lx - 1.  Every node sum is taken in a fixed order (deterministic, no
//     atomics).
// Thread columns lx, lz in [1, 14] own output nodes; columns 0 and 15 are halo
// that neighbouring tiles also compute.  Per thread and layer 9 doubles are
// written to LDS and 15 read back; all other data stays in registers.  The
// LDS buffers alternate between layers, so one barrier per layer orders both
// exchanges: the z-butterflies of the plane two above are written at the end
// of a layer, together with its x-sums.  Inactive
// elements (outside the grid or in the halo) get a zero modulus so that their
// contribution vanishes without divergent code around the cross-lane moves.
#pragma once
#include "common.hpp"
#include "fantom_internal.hpp"

namespace fantom {

constexpr int kTW = 16;                // node columns per tile edge
constexpr int kTOut = kTW - 2;         // output node columns per tile edge
constexpr int kTThreads = kTW * kTW;   // threads per workgroup

struct TileShape {
    int tiles_x, tiles_z, chunks_y, chunk;
};

// Host: tile counts for grid g and an operator with per_cu resident
// workgroups per CU.  A workgroup walks chunk planes in sequence (chunk + 1
// element layers, one of them halo), so with n chunks along y the run takes
// about ceil(cols n / capacity) rounds of (ceil(ny / n) + 1) layer steps;
// the chunk count minimises that product (fewest workgroups on ties).
// FANTOM_TILE_CHUNK=<planes> overrides the choice (tuning sweeps).
TileShape tile_shape(const Grid &g, int per_cu);
unsigned tile_blocks(const TileShape &ts);


// DPP moves inside rows of 16 lanes; a lane whose source lies outside its row
// reads 0 (bound_ctrl).
constexpr int kDppRowShl1 = 0x101;   // lane l <- lane l + 1
constexpr int kDppRowShr1 = 0x111;   // lane l <- lane l - 1

template <int CTRL>
__device__ __forceinline__ float dpp_row(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}

template <int CTRL>
__device__ __forceinline__ double dpp_row(double v)
{
    const unsigned long long bits = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)bits, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(bits >> 32), CTRL, 0xF, 0xF, true);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo);
}

// The operator policy Op provides
//   struct Data;                                         per-element inputs
//   void load(Data &, const Grid &, bool active, int ex, int ey, int ez) const;
//   void apply(const Data &, const T (&w)[3][8], T (&z)[3][8]) const;
//   static constexpr bool kScaled;  T scale(const Data &) const;
//   static constexpr bool kDiagFromModuli;  T inv_diag(int c, T inv_sum) const;
//   static constexpr int kWavesPerEU;  static constexpr int waves_for(int mode);
//   static constexpr bool staged_store(int mode);
// where load fills Data for element (ex, ey, ez) (zero modulus when !active)
// and apply maps WHT coefficients to output coefficients (the 1/8 of the
// inverse transform is folded into the operator's tables).  With kScaled the
// outputs of apply still lack one per-element factor, scale(d), which the
// inverse y-butterfly folds in (one FMA and one product per coefficient pair
// instead of a product per coefficient and an add).  With kDiagFromModuli
// the operator's diagonal at a node is d_c * (sum of scale() over the node's
// elements), and the Jacobi sweeps (not the zero-guess one, whose first half
// needs D^-1 two planes ahead) form D^-1 = inv_diag(c, 1 / sum) from the
// moduli they already hold instead of reading a D^-1 vector.  kWavesPerEU
// is the occupancy the register budget is held to, waves_for(MODE) the one a
// mode is built for (lower where the budget would spill inside the layer
// loop: one wave fewer measured faster than spilling).  staged_store(MODE):
// the outputs of plane L go to an LDS row buffer in layer L and are stored in
// layer L + 1 (after its barrier) as contiguous words per tile row, 14 x 3
// values of one x-column, instead of three 8-byte stores at a 24-byte stride
// per node.  That cut the HBM-resident 256x128x128 K.u from 79.7 to 73.2 us
// and the residual from 97.4 to 87.6 us; where the extra registers spill (the
// Jacobi modes) the operator keeps the direct stores.
template <typename T, int MODE, class Op>
__global__ __launch_bounds__(kTThreads) __attribute__((amdgpu_waves_per_eu(Op::waves_for(MODE), 8))) void ku_tile_kernel(Grid g, Op op, const uint8_t *__restrict__ mask,
                                                          const T *__restrict__ u, T *__restrict__ y,
                                                          const T *__restrict__ bvec, const T *__restrict__ dinv,
                                                          T omega, double *__restrict__ partials, TileShape ts,
                                                          const T *__restrict__ ec)
{
    constexpr bool kZeroGuess = MODE == MODE_JAC0 || MODE == MODE_JAC0OWN;
    constexpr bool kPro = MODE == MODE_JACPRO;   // node values u + P ec
    constexpr bool kJacobi =
        kZeroGuess || MODE == MODE_JAC || MODE == MODE_JACOWN || MODE == MODE_JACDOT || kPro;
    constexpr bool kOwnDiag = Op::kDiagFromModuli && (MODE == MODE_JACOWN || MODE == MODE_JACDOT ||
                                                      MODE == MODE_JAC0OWN || kPro);
    static_assert(!kPro || Op::kDiagFromModuli, "the prolongating sweep forms D^-1 from the moduli");
    // right-hand side (and D^-1 unless formed here) read for the epilogue
    constexpr bool kEpiB = (kJacobi || MODE == MODE_RES) && MODE != MODE_JAC0OWN;
    if (kJacobi)
        omega = jacobi_weight(g, omega);
    // weight of the sweep that follows x0 in the zero-guess modes
    const T ratio = (T)g.w2_ratio;
    const T omega2 = omega * ratio;
    // double-buffered by layer parity, so that one barrier per layer orders
    // both exchanges
    __shared__ T xf[2][6][kTThreads];   // z-butterflies, exchanged with column lx - 1
    __shared__ T xr[2][3][kTThreads];   // upper-x node sums, exchanged with column lx + 1
    __shared__ T xe[2][kOwnDiag ? kTThreads : 1];   // z-pair sums of the moduli, exchanged with column lx + 1
    constexpr bool kStaged = Op::staged_store(MODE);
    __shared__ T ys[2][kStaged ? kTOut : 1][3 * kTOut];   // staged outputs of a plane, one row per output x-column
    // kPro: the coarse correction of a fine node plane, combined in y (the
    // mean of the two coarse planes around an odd fine plane), over the 9 x 9
    // coarse nodes under the tile, 3 components; three planes in a ring
    constexpr int kProNodes = kTOut / 2 + 2;
    __shared__ T cs[kPro ? 3 : 1][kPro ? kTThreads : 1];   // 243 of the 256 slots used

    // The hardware deals workgroups to the 8 XCDs round-robin; give each XCD
    // a contiguous range of logical tiles so that tiles sharing halo planes
    // and columns meet in the same L2.  The grid is a multiple of 8 blocks.
    const int per_xcd = gridDim.x / 8;
    const int bid = blockIdx.x;
    const int logical = (bid % 8) * per_xcd + bid / 8;
    double part = 0.0;
    if (logical < ts.tiles_x * ts.tiles_z * ts.chunks_y) {   // uniform per workgroup
        const int tz = logical % ts.tiles_z;
        const int tx = (logical / ts.tiles_z) % ts.tiles_x;
        const int cyk = logical / (ts.tiles_z * ts.tiles_x);
        const int tid = threadIdx.x, lx = tid >> 4, lz = tid & 15;
        const int gx = tx * kTOut - 1 + lx, gz = tz * kTOut - 1 + lz;
        const bool col_in = gx >= 0 && gx < g.nx && gz >= 0 && gz < g.nz;
        const bool el_col = lx < kTW - 1 && lz < kTW - 1 && gx >= 0 && gx < g.nelx && gz >= 0 && gz < g.nelz;
        const bool out_col = lx >= 1 && lx < kTW - 1 && lz >= 1 && lz < kTW - 1 && gx < g.nx && gz < g.nz;
        const long nplane = (long)g.nx * g.nz;
        const long col = (long)gx * g.nz + gz;
        const int nb_up = tid + kTW < kTThreads ? tid + kTW : tid;   // column lx + 1 (lx = 15: any valid slot)
        const int nb_dn = tid >= kTW ? tid - kTW : tid;              // column lx - 1 (lx = 0: any valid slot)
        const int jc0 = cyk * ts.chunk;
        const int gz0 = tz * kTOut;                 // first output node of the tile's rows
        const int nrow = min(kTOut, g.nz - gz0);    // output nodes per row
        const bool row_out = lx >= 1 && lx < kTW - 1 && gx < g.nx;
        const int jc1 = min(jc0 + ts.chunk, g.ny);   // output planes [jc0, jc1)

        // zero-guess sweeps: the node value is x0 = omega D^-1 b.  load_node
        // leaves b in v and D^-1 in vd and x0 is formed when the value is
        // first used (form_x0), so that the loads stay in flight meanwhile;
        // the fine operator's sweep with a read D^-1 (register-bound) forms
        // it on load instead.
        constexpr bool kDeferX0 = kZeroGuess && (MODE == MODE_JAC0OWN || !Op::kDiagFromModuli);
        T vd_cur[3] = {T(0), T(0), T(0)}, vd_nx[3] = {T(0), T(0), T(0)}, vd_pf[3] = {T(0), T(0), T(0)};
        auto load_node = [&](int j, T (&v)[3], int &m, T (&vd)[3]) {
            v[0] = v[1] = v[2] = T(0);
            m = 0;
            if (col_in && j >= 0 && j < g.ny) {
                const long n = (long)j * nplane + col;
                if (kDeferX0) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        v[c] = bvec[3 * n + c];
                        vd[c] = dinv[3 * n + c];
                    }
                } else if (kZeroGuess) {
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        v[c] = omega * dinv[3 * n + c] * bvec[3 * n + c];
                } else {
                    v[0] = u[3 * n];
                    v[1] = u[3 * n + 1];
                    v[2] = u[3 * n + 2];
                }
                if (mask)
                    m = mask[n];
            }
        };
        auto form_x0 = [&](T (&v)[3], const T (&vd)[3]) {
            if (kDeferX0) {
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    v[c] = omega * vd[c] * v[c];
            }
        };
        // face transform of a plane in two halves: face_write takes the
        // z-butterflies (DPP) of the thread's (unmasked) node values into LDS,
        // face_read (after a barrier) adds the x-butterflies with column lx + 1
        auto face_write = [&](const T (&v)[3], int m, int buf) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T a = ((m >> c) & 1) ? T(0) : v[c];
                const T b = dpp_row<kDppRowShl1>(a);
                xf[buf][2 * c][tid] = a + b;
                xf[buf][2 * c + 1][tid] = a - b;
            }
        };
        auto face_read = [&](int buf, T (&F)[3][4]) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int sz = 0; sz < 2; ++sz) {
                    const T own = xf[buf][2 * c + sz][tid];
                    const T nb = xf[buf][2 * c + sz][nb_up];
                    F[c][2 * sz] = own + nb;
                    F[c][2 * sz + 1] = own - nb;
                }
        };
        auto load_elem = [&](typename Op::Data &d, int L) {
            op.load(d, g, el_col && L >= 0 && L < g.nely, gx, L, gz);
        };
        // kPro.  Coarse node c sits on fine node 2c; the tile's fine columns
        // tx*14 - 1 ... tx*14 + 14 lie over coarse columns cx0 ... cx0 + 8
        // (cz0 ... cz0 + 8 in z).  Staging thread tid < 243 holds component
        // tid % 3 of coarse column (tid / 3) / 9, row (tid / 3) % 9; its
        // offset within a coarse plane is fixed, -1 outside the grid.  Every
        // access below is in range for every thread, so that none of it
        // branches.
        const int nxc = g.nelx / 2 + 1, nyc = g.nely / 2 + 1, nzc = g.nelz / 2 + 1;
        const int cx0 = tx * (kTOut / 2) - 1, cz0 = tz * (kTOut / 2) - 1;
        [[maybe_unused]] int pro_off = -1;
        if constexpr (kPro) {
            const int cc = tid / 3, ic = cx0 + cc / kProNodes, kc = cz0 + cc % kProNodes;
            if (tid < kProNodes * kProNodes * 3 && ic >= 0 && ic < nxc && kc >= 0 && kc < nzc)
                pro_off = 3 * (ic * nzc + kc) + tid % 3;
        }
        // pro_fetch: this thread's share of fine plane j's coarse values (the
        // mean of coarse planes j / 2 and j / 2 + 1 for odd j), loaded now and
        // written to LDS by pro_put later
        auto pro_fetch = [&](int j) -> T {
            const bool ok = pro_off >= 0 && j >= 0 && j < g.ny;
            const int jc = ok ? j >> 1 : 0;
            const bool two = ok && (j & 1) && jc + 1 < nyc;
            const long plane = 3L * nxc * nzc;
            const long i = ok ? jc * plane + pro_off : 0;
            const T a = ec[i], b = ec[two ? i + plane : i];
            return ok ? (two ? T(0.5) * (a + b) : a) : T(0);
        };
        auto pro_put = [&](int buf, T v) { cs[buf][tid] = v; };   // slots >= 243: unused
        // v += (P e) at this thread's node of plane j, from ring buffer buf
        // (bilinear in x and z over the staged plane; constrained DOFs get 0)
        auto pro_add = [&](int j, T (&v)[3], int m, int buf) {
            const bool ok = col_in && j >= 0 && j < g.ny;
            const int ic0 = gx >> 1, kc0 = gz >> 1;
            const bool two_x = (gx & 1) && ic0 + 1 < nxc, two_z = (gz & 1) && kc0 + 1 < nzc;
            const T wx1 = two_x ? T(0.5) : T(0), wx0 = T(1) - wx1;
            const T wz1 = two_z ? T(0.5) : T(0), wz0 = T(1) - wz1;
            const int o00 = 3 * ((ic0 - cx0) * kProNodes + (kc0 - cz0));
            const int o01 = o00 + (two_z ? 3 : 0), o10 = o00 + (two_x ? 3 * kProNodes : 0);
            const int o11 = o10 + (two_z ? 3 : 0);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T a = wz0 * cs[buf][o00 + c] + wz1 * cs[buf][o01 + c];
                const T b = wz0 * cs[buf][o10 + c] + wz1 * cs[buf][o11 + c];
                v[c] += (ok && !((m >> c) & 1)) ? wx0 * a + wx1 * b : T(0);
            }
        };
        T FpA[3][4], FpB[3][4], ucur[3], unx[3];
        int mcur, mnx;
        typename Op::Data dcur, dnx;
        load_node(jc0 - 1, ucur, mcur, vd_cur);
        load_node(jc0, unx, mnx, vd_nx);
        form_x0(ucur, vd_cur);
        form_x0(unx, vd_nx);
        if constexpr (kPro) {
            // ring buffer of plane P: (P - (jc0 - 1)) % 3
            const T c0 = pro_fetch(jc0 - 1), c1 = pro_fetch(jc0), c2 = pro_fetch(jc0 + 1);
            pro_put(0, c0);
            pro_put(1, c1);
            pro_put(2, c2);
            __syncthreads();
            pro_add(jc0 - 1, ucur, mcur, 0);
            pro_add(jc0, unx, mnx, 1);
        }
        load_elem(dcur, jc0 - 1);
        face_write(ucur, mcur, 0);
        __syncthreads();
        face_read(0, FpA);
        face_write(unx, mnx, 1);
        __syncthreads();
        int pb = 1, rb = 0;   // LDS buffers of plane L + 1's z-butterflies and of this layer's x-sums
        T bcur[3] = {T(0), T(0), T(0)}, dgcur[3] = {T(0), T(0), T(0)};
        T es_below = T(0);   // kOwnDiag: moduli around the node in the layer below
        T Zu[3][4];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                Zu[c][q] = T(0);

        // one element layer: Fp holds the face transform of plane L, the
        // layer leaves that of plane L + 1 in Fc.  The loop alternates the two
        // face buffers instead of copying one into the other.
        auto layer = [&](const int L, const T (&Fp)[3][4], T (&Fc)[3][4]) {
            // element layer L lies between node planes L and L + 1; plane L is
            // written at the end of this step (L >= jc0)
            T upf[3];
            int mpf;
            load_node(L + 2 <= jc1 ? L + 2 : -1, upf, mpf, vd_pf);
            load_elem(dnx, L + 1 < jc1 ? L + 1 : -1);
            // kPro: plane L + 3's coarse values go to ring buffer r % 3 before
            // this layer's barrier; plane L + 2's were staged by the layer before
            [[maybe_unused]] int pro_r = 0;
            [[maybe_unused]] T pro_next = T(0);
            if constexpr (kPro) {
                pro_r = L - (jc0 - 1);
                pro_next = pro_fetch(L + 3 <= jc1 ? L + 3 : -1);
            }
            // right-hand side (and D^-1) of plane L, used by this layer's
            // epilogue after the barrier: loaded here rather than a layer
            // ahead, so that they occupy registers for one layer only
            if (kEpiB && out_col && L >= jc0) {
                const long r0 = 3 * ((long)L * nplane + col);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    bcur[c] = bvec[r0 + c];
                    if (kJacobi && !kOwnDiag)
                        dgcur[c] = dinv[r0 + c];
                }
            }

            face_read(pb, Fc);

            T w[3][8], z[3][8];
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int sx = 0; sx < 2; ++sx)
#pragma unroll
                    for (int sz = 0; sz < 2; ++sz) {
                        const int q = sx + 2 * sz;
                        w[c][sx + 4 * sz] = Fp[c][q] + Fc[c][q];
                        w[c][sx + 2 + 4 * sz] = Fp[c][q] - Fc[c][q];
                    }
            op.apply(dcur, w, z);

            T val[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                T A[4];
#pragma unroll
                for (int sx = 0; sx < 2; ++sx)
#pragma unroll
                    for (int sz = 0; sz < 2; ++sz) {
                        const int q = sx + 2 * sz;
                        const T lo = z[c][sx + 4 * sz] + z[c][sx + 2 + 4 * sz];
                        const T hi = z[c][sx + 4 * sz] - z[c][sx + 2 + 4 * sz];
                        if constexpr (Op::kScaled) {
                            const T e = op.scale(dcur);
                            A[q] = fma_t(e, lo, Zu[c][q]);
                            Zu[c][q] = e * hi;
                        } else {
                            A[q] = lo + Zu[c][q];
                            Zu[c][q] = hi;
                        }
                    }
                // inverse x butterfly: P[bx][sz]; inverse z butterfly: corner bz
                const T p00 = A[0] + A[1], p10 = A[0] - A[1];   // sz = 0
                const T p01 = A[2] + A[3], p11 = A[2] - A[3];   // sz = 1
                const T q00 = p00 + p01, q01 = p00 - p01;       // bx = 0: corners bz = 0, 1
                const T q10 = p10 + p11, q11 = p10 - p11;       // bx = 1
                const T r0 = q00 + dpp_row<kDppRowShr1>(q01);   // node (lx, lz) from this column
                const T r1 = q10 + dpp_row<kDppRowShr1>(q11);   // node (lx + 1, lz)
                val[c] = r0;
                xr[rb][c][tid] = r1;
            }
            form_x0(upf, vd_pf);
            if constexpr (kPro) {
                pro_add(L + 2 <= jc1 ? L + 2 : -1, upf, mpf, (pro_r + 2) % 3);
                pro_put(pro_r % 3, pro_next);
            }
            face_write(upf, mpf, pb ^ 1);   // plane L + 2, read by the next layer
            if constexpr (kOwnDiag) {
                const T e = op.scale(dcur);
                xe[rb][tid] = e + dpp_row<kDppRowShr1>(e);   // element columns lz - 1, lz
            }
            __syncthreads();
            if constexpr (kOwnDiag) {
                // node (lx, lz) of plane L: its four element columns in layers
                // L - 1 and L
                const T es = xe[rb][tid] + xe[rb][nb_dn];
                const T inv_sum = T(1) / (es_below + es);
                es_below = es;
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    dgcur[c] = ((mcur >> c) & 1) ? T(1) : op.inv_diag(c, inv_sum);
            }
            if (L >= jc0 && out_col) {
                const long r0 = 3 * ((long)L * nplane + col);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const T kuv = ((mcur >> c) & 1) ? ucur[c] : val[c] + xr[rb][c][nb_dn];
                    T o;
                    if (MODE == MODE_JAC0OWN)   // x0 = omega D^-1 b, so x0 + omega2 D^-1 (b - K x0) is:
                        o = (ucur[c] + ratio * ucur[c]) - omega2 * dgcur[c] * kuv;
                    else if (MODE == MODE_JAC0)
                        o = ucur[c] + omega2 * dgcur[c] * (bcur[c] - kuv);
                    else if (kJacobi)
                        o = ucur[c] + omega * dgcur[c] * (bcur[c] - kuv);
                    else if (MODE == MODE_RES)
                        o = bcur[c] - kuv;
                    else
                        o = kuv;
                    if (kStaged)
                        ys[rb][lx - 1][3 * (lz - 1) + c] = o;
                    else
                        y[r0 + c] = o;
                    if (MODE == MODE_KUDOT)
                        part += (double)ucur[c] * (double)o;
                    if (MODE == MODE_JACDOT)
                        part += (double)bcur[c] * (double)o;
                }
            }
            // kStaged: the outputs of plane L - 1, staged by the layer before,
            // as contiguous words per tile row
            if (kStaged && L - 1 >= jc0 && row_out) {
                const long base = 3 * ((long)(L - 1) * nplane + (long)gx * g.nz + gz0);
                for (int wi = lz; wi < 3 * nrow; wi += kTW)
                    y[base + wi] = ys[rb ^ 1][lx - 1][wi];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                ucur[c] = unx[c];
                unx[c] = upf[c];
            }
            mcur = mnx;
            mnx = mpf;
            dcur = dnx;
            pb ^= 1;
            rb ^= 1;
        };
        int L = jc0 - 1;
        if constexpr (!(MODE == MODE_JAC || MODE == MODE_JAC0)) {
            for (; L + 1 < jc1; L += 2) {
                layer(L, FpA, FpB);
                layer(L + 1, FpB, FpA);
            }
            if (L < jc1)
                layer(L, FpA, FpB);
        } else {
            // the sweeps that read D^-1 keep one layer per trip: two copies of
            // the layer would not fit the register budget there
            for (; L < jc1; ++L) {
                layer(L, FpA, FpB);
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        FpA[c][q] = FpB[c][q];
            }
        }
        if (kStaged) {   // the chunk's last plane
            __syncthreads();
            if (jc1 - 1 >= jc0 && row_out) {
                const long base = 3 * ((long)(jc1 - 1) * nplane + (long)gx * g.nz + gz0);
                for (int wi = lz; wi < 3 * nrow; wi += kTW)
                    y[base + wi] = ys[rb ^ 1][lx - 1][wi];
            }
        }
    }
    if (MODE == MODE_KUDOT || MODE == MODE_JACDOT) {
        const double s = block_sum<kTThreads>(part);
        if (threadIdx.x == 0)
            partials[blockIdx.x] = s;
    }
}

// Host: resident workgroups per CU of ku_tile_kernel<T, MODE, Op> from the
// occupancy API (registers and LDS of that instantiation), 0 without a
// device.  The chunk choice (tile_shape) depends on it: with a count above
// what fits, the last round of workgroups runs nearly empty; below, the chip
// is filled with fewer waves than it could hold.
template <typename T, int MODE, class Op>
int tile_occupancy()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void *>(&ku_tile_kernel<T, MODE, Op>),
                                                     kTThreads, 0) != hipSuccess) {
        (void)hipGetLastError();   // leave no error for the next launch check
        return 0;
    }
    return nb;
}

}  // namespace fantom
