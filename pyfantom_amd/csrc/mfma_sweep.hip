// BF16 matrix-core sweeps for the mixed-precision V-cycle: the fine hex8
// operator (K0 scaled by the element modulus) and the level-1 Galerkin
// operator (sum over the 8 children of E_child * P_child^T K0 P_child), with
// BF16 operands and FP32 accumulation on v_mfma_f32_16x16x16_bf16.
//
// Replaces, inside the preconditioner only, the element contraction of
// core/CUDA/Kernels/matvec.cu:1-98 (and the Galerkin product the reference
// assembles at core/CPU/_mgm.py:308-942); the outer PCG keeps FP64.
//
// Layout of one workgroup (256 threads, 4 waves): a 16 x 16 column of
// elements in (x, z) walking a chunk of element layers along y.  Per layer:
//   * the node plane above the layer is staged in LDS as (x, y, z, 0) BF16
//     quadruples (one 8-byte word per node, Dirichlet components zeroed), the
//     moduli of the layer as FP32;
//   * wave w takes element rows x = 4w .. 4w + 3; a row of 16 elements along z
//     is one MFMA column block: B[k][n] = the element's input vector with
//     k = 16 kh + 4 g + comp, corner (bx, by, bz) = (g & 1, g >> 1, kh), so
//     that a lane's four values per K-half are one node (one 8-byte LDS
//     read); the output rows m = 4 q + comp of M-block t are corner
//     (q >> 1, t, q & 1), so a lane's accumulator holds one output node;
//   * the products are scaled by the modulus in FP32 and written to LDS per
//     (output corner, element); each node then adds its four contributions
//     per layer in a fixed order (deterministic), the lower layer's from
//     registers.
// A tile computes 16 x 16 elements and writes the 15 x 15 nodes whose four
// element columns it holds; tiles overlap by one element column.  Loads run
// one layer ahead.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.hpp"
#include "fantom_internal.hpp"

namespace fantom {

typedef short mf_s4 __attribute__((ext_vector_type(4)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));

constexpr int kMfE = 16;        // elements per tile edge (x and z)
constexpr int kMfN = kMfE + 1;  // staged nodes per tile edge
constexpr int kMfOut = kMfE - 1;  // output nodes per tile edge
constexpr int kMfThreads = 256;

__device__ __forceinline__ uint32_t mf_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u)
        return 0x7fc0u;
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}

__device__ __forceinline__ uint2 mf_pack(float a, float b, float c)
{
    return make_uint2(mf_bf16(a) | (mf_bf16(b) << 16), mf_bf16(c));
}

struct MfArgs {
    Grid g;                  // the level's grid
    const uint8_t *mask;
    const float *u;
    float *y;
    const float *b;
    const float *dinv;
    const float *E;          // NTERM 1: the level's moduli; 8: the fine moduli (grid 2x)
    const mf_s4 *A;          // [term][t][kh][lane]
    float omega;
    float rd[3];             // 1 / K0 diagonal per component (diagonal from the moduli)
    int tiles_x, tiles_z, chunks, chunk;
};

// SPLIT bit 0: the node values as hi + lo BF16 pairs (two products per
// operand); bit 1 (NTERM 1): the element matrix as hi + lo pairs
template <int NTERM, int MODE, int SPLIT>
__global__ __launch_bounds__(kMfThreads) void mf_sweep_k(MfArgs a)
{
    constexpr bool kSplitX = SPLIT & 1;
    constexpr bool kSplitK = (SPLIT & 2) && NTERM == 1;
    constexpr bool kZero = MODE == MODE_JAC0 || MODE == MODE_JAC0OWN;
    constexpr bool kOwn = MODE == MODE_JACOWN || MODE == MODE_JAC0OWN;
    constexpr bool kJac = kZero || MODE == MODE_JAC || MODE == MODE_JACOWN;
    constexpr bool kReadB = kJac || MODE == MODE_RES;
    constexpr bool kReadD = kZero || MODE == MODE_JAC;   // D^-1 read (x0 or the sweep)
    static_assert(NTERM == 1 || !kOwn, "the diagonal from the moduli is the fine operator's");

    __shared__ uint2 P[2][kMfN * kMfN];
    __shared__ uint2 Plo[kSplitX ? 2 : 1][kSplitX ? kMfN * kMfN : 1];
    __shared__ mf_f4 S[2][4][kMfE][kMfE];
    __shared__ float Et[2][NTERM][kMfE * kMfE];

    const Grid &g = a.g;
    const int per_xcd = gridDim.x / 8;
    const int logical = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (logical >= a.tiles_x * a.tiles_z * a.chunks)   // uniform per workgroup
        return;
    const int tz = logical % a.tiles_z;
    const int tx = (logical / a.tiles_z) % a.tiles_x;
    const int cy = logical / (a.tiles_z * a.tiles_x);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, lg = lane >> 4, ln = lane & 15;
    const int ox = tx * kMfOut - 1, oz = tz * kMfOut - 1;   // global node / element of staged index 0
    const int jc0 = cy * a.chunk, jc1 = min(jc0 + a.chunk, g.ny);
    const long nplane = (long)g.nx * g.nz;

    float omega = a.omega;
    if (kJac)
        omega = jacobi_weight(g, omega);
    const float ratio = (float)g.w2_ratio;
    const float omega2 = omega * ratio;

    // A operands (constant per kernel)
    mf_s4 A[NTERM][2][2];
#pragma unroll
    for (int c = 0; c < NTERM; ++c)
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int kh = 0; kh < 2; ++kh)
                A[c][t][kh] = a.A[((c * 2 + t) * 2 + kh) * 64 + lane];
    mf_s4 Alo[kSplitK ? 2 : 1][2];
    if (kSplitK) {
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int kh = 0; kh < 2; ++kh)
                Alo[t][kh] = a.A[NTERM * 256 + (t * 2 + kh) * 64 + lane];
    }

    // the thread's nodes: primary (x, z) = (tid / 16, tid % 16); threads
    // below 33 also stage the x = 16 row or the z = 16 column
    const int px = tid >> 4, pz = tid & 15;
    const int sx = tid < kMfN ? kMfE : tid - kMfN, sz = tid < kMfN ? tid : kMfE;
    const bool has_sec = tid < kMfN + kMfE;
    const int pgx = ox + px, pgz = oz + pz, sgx = ox + sx, sgz = oz + sz;
    const bool p_in = pgx >= 0 && pgx < g.nx && pgz >= 0 && pgz < g.nz;
    const bool s_in = has_sec && sgx >= 0 && sgx < g.nx && sgz >= 0 && sgz < g.nz;
    const bool out = px >= 1 && pz >= 1 && pgx < g.nx && pgz < g.nz;
    const long pcol = (long)pgx * g.nz + pgz, scol = (long)sgx * g.nz + sgz;

    struct Node {
        float v[3], bb[3], dd[3];
        int m;
    };
    auto load_primary = [&](int j, Node &nd) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
            nd.v[c] = nd.bb[c] = nd.dd[c] = 0.f;
        nd.m = 0;
        if (p_in && j >= 0 && j < g.ny) {
            const long n = (long)j * nplane + pcol;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (!kZero)
                    nd.v[c] = a.u[3 * n + c];
                if (kReadB)
                    nd.bb[c] = a.b[3 * n + c];
                if (kReadD)
                    nd.dd[c] = a.dinv[3 * n + c];
            }
            if (a.mask)
                nd.m = a.mask[n];
        }
    };
    // staged value of the primary node: u, or x0 = omega D^-1 b
    auto value = [&](Node &nd) {
        if (kZero) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
                nd.v[c] = omega * nd.dd[c] * nd.bb[c];
        }
    };
    auto load_secondary = [&](int j, float (&v)[3], int &m) {
        v[0] = v[1] = v[2] = 0.f;
        m = 0;
        if (s_in && j >= 0 && j < g.ny) {
            const long n = (long)j * nplane + scol;
            if (kZero) {
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    v[c] = omega * a.dinv[3 * n + c] * a.b[3 * n + c];
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    v[c] = a.u[3 * n + c];
            }
            if (a.mask)
                m = a.mask[n];
        }
    };
    // moduli of layer L for element (tid / 16, tid % 16)
    auto load_e = [&](int L, float (&e)[NTERM]) {
        const int gx = ox + px, gz = oz + pz;
        const bool in = L >= 0 && L < g.nely && gx >= 0 && gx < g.nelx && gz >= 0 && gz < g.nelz;
#pragma unroll
        for (int c = 0; c < NTERM; ++c) {
            if (!in) {
                e[c] = 0.f;
            } else if (NTERM == 1) {
                e[c] = a.E[((long)gx * g.nely + L) * g.nelz + gz];
            } else {   // child c = cx + 2 cy + 4 cz of the fine grid
                const int fy = 2 * g.nely, fz = 2 * g.nelz;
                e[c] = a.E[((long)(2 * gx + (c & 1)) * fy + (2 * L + ((c >> 1) & 1))) * fz + (2 * gz + (c >> 2))];
            }
        }
    };
    auto masked = [](const float (&v)[3], int m, int c) { return ((m >> c) & 1) ? 0.f : v[c]; };
    auto put = [&](int buf, int idx, float x0, float x1, float x2) {
        P[buf][idx] = mf_pack(x0, x1, x2);
        if (kSplitX) {   // the remainders, rounded again
            const uint2 h = P[buf][idx];
            const float h0 = __uint_as_float(h.x << 16), h1 = __uint_as_float(h.x & 0xffff0000u);
            const float h2 = __uint_as_float(h.y << 16);
            Plo[buf][idx] = mf_pack(x0 - h0, x1 - h1, x2 - h2);
        }
    };
    auto stage = [&](int buf, const Node &pn, const float (&sv)[3], int sm) {
        put(buf, px * kMfN + pz, masked(pn.v, pn.m, 0), masked(pn.v, pn.m, 1), masked(pn.v, pn.m, 2));
        if (has_sec)
            put(buf, sx * kMfN + sz, masked(sv, sm, 0), masked(sv, sm, 1), masked(sv, sm, 2));
    };

    // pipeline: cur = plane L (for the epilogue), nx = plane L + 1 (staged at
    // the top of layer L), pf = plane L + 2 (in flight; a second layer in
    // flight costs a wave per SIMD and measured slower)
    Node cur, nx, pf;
    float snx[3], spf[3];
    int smnx, smpf;
    float enx[NTERM], epf[NTERM];
    int L = jc0 - 1;
    load_primary(L, cur);
    value(cur);
    {
        float sv[3];
        int sm;
        load_secondary(L, sv, sm);
        stage(L & 1, cur, sv, sm);
    }
    load_primary(L + 1, nx);
    load_secondary(L + 1, snx, smnx);
    load_e(L, enx);
    float top[3] = {0.f, 0.f, 0.f};
    float es_below = 0.f;
    for (; L < jc1; ++L) {
        value(nx);
        stage((L + 1) & 1, nx, snx, smnx);
#pragma unroll
        for (int c = 0; c < NTERM; ++c)
            Et[L & 1][c][tid] = enx[c];
        load_primary(L + 2, pf);
        load_secondary(L + 2, spf, smpf);
        load_e(L + 1, epf);
        __syncthreads();
        // products of the wave's four element rows
        const int bx = lg & 1, by = lg >> 1;
        const uint2 *Pin = P[(L + by) & 1];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int ex = 4 * w + r;
            const int i0 = (ex + bx) * kMfN + ln;
            auto operand = [](uint2 wv) {
                mf_s4 o;
                o[0] = (short)(wv.x & 0xffff);
                o[1] = (short)(wv.x >> 16);
                o[2] = (short)(wv.y & 0xffff);
                o[3] = 0;
                return o;
            };
            const mf_s4 b0 = operand(Pin[i0]), b1 = operand(Pin[i0 + 1]);
            mf_s4 l0 = b0, l1 = b1;
            if (kSplitX) {
                const uint2 *Pl = Plo[(L + by) & 1];
                l0 = operand(Pl[i0]);
                l1 = operand(Pl[i0 + 1]);
            }
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                mf_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int c = 0; c < NTERM; ++c) {
                    mf_f4 d = {0.f, 0.f, 0.f, 0.f};
                    if (kSplitK) {   // small terms first
                        d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Alo[t][0], b0, d, 0, 0, 0);
                        d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Alo[t][1], b1, d, 0, 0, 0);
                    }
                    if (kSplitX) {
                        d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[c][t][0], l0, d, 0, 0, 0);
                        d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[c][t][1], l1, d, 0, 0, 0);
                    }
                    d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[c][t][0], b0, d, 0, 0, 0);
                    d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[c][t][1], b1, d, 0, 0, 0);
                    const float e = Et[L & 1][c][ex * kMfE + ln];
                    if (NTERM == 1)
                        acc = d * e;
                    else
                        acc += d * e;
                }
                S[t][lg][ex][ln] = acc;
            }
        }
        __syncthreads();
        if (out) {
            float bot[3] = {0.f, 0.f, 0.f}, up[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int ex = px - (q >> 1), ez = pz - (q & 1);
                const mf_f4 s0 = S[0][q][ex][ez], s1 = S[1][q][ex][ez];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    bot[c] += s0[c];
                    up[c] += s1[c];
                }
            }
            float es = 0.f;
            if (kOwn) {
                const float *e0 = Et[L & 1][0];
                es = (e0[(px - 1) * kMfE + pz - 1] + e0[(px - 1) * kMfE + pz]) +
                     (e0[px * kMfE + pz - 1] + e0[px * kMfE + pz]);
            }
            if (L >= jc0) {   // plane L is an output plane (L < jc1 <= ny)
                const long r0 = 3 * ((long)L * nplane + pcol);
                const float inv_sum = kOwn ? 1.f / (es_below + es) : 0.f;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const bool mc = (cur.m >> c) & 1;
                    const float kuv = mc ? cur.v[c] : top[c] + bot[c];
                    const float dg = kOwn ? (mc ? 1.f : a.rd[c] * inv_sum) : cur.dd[c];
                    float o;
                    if (MODE == MODE_JAC0OWN)
                        o = (cur.v[c] + ratio * cur.v[c]) - omega2 * dg * kuv;
                    else if (MODE == MODE_JAC0)
                        o = cur.v[c] + omega2 * dg * (cur.bb[c] - kuv);
                    else if (kJac)
                        o = cur.v[c] + omega * dg * (cur.bb[c] - kuv);
                    else if (MODE == MODE_RES)
                        o = cur.bb[c] - kuv;
                    else
                        o = kuv;
                    a.y[r0 + c] = o;
                }
            }
#pragma unroll
            for (int c = 0; c < 3; ++c)
                top[c] = up[c];
            es_below = es;
        }
        cur = nx;
        nx = pf;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            snx[c] = spf[c];
        smnx = smpf;
#pragma unroll
        for (int c = 0; c < NTERM; ++c)
            enx[c] = epf[c];
    }
}

// ---- host side --------------------------------------------------------------

static uint16_t host_bf16(double v)
{
    float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// out: the BF16 table of M (mf_table_halves(nterm) values), then the table
// of the remainders M - bf16(M) rounded to BF16 (the same count)
void mf_tables(const double *M, int nterm, uint16_t *out)
{
    const long half = mf_table_halves(nterm);
    for (int c = 0; c < nterm; ++c)
        for (int t = 0; t < 2; ++t)
            for (int kh = 0; kh < 2; ++kh)
                for (int l = 0; l < 64; ++l)
                    for (int i = 0; i < 4; ++i) {
                        const int m = l & 15, q = m >> 2, oc = m & 3;
                        const int gi = l >> 4, ic = i;
                        double v = 0.0;
                        if (oc < 3 && ic < 3) {
                            const int nat_out = (q >> 1) + 2 * t + 4 * (q & 1);
                            const int nat_in = (gi & 1) + 2 * (gi >> 1) + 4 * kh;
                            v = M[(long)c * 576 + (nat_out * 3 + oc) * 24 + nat_in * 3 + ic];
                        }
                        const long idx = ((((long)c * 2 + t) * 2 + kh) * 64 + l) * 4 + i;
                        const uint16_t hi = host_bf16(v);
                        uint32_t hb = (uint32_t)hi << 16;
                        float hf;
                        std::memcpy(&hf, &hb, 4);
                        out[idx] = hi;
                        out[half + idx] = host_bf16(v - (double)hf);
                    }
}

long mf_table_halves(int nterm) { return (long)nterm * 2 * 2 * 64 * 4; }
long mf_table_size(int nterm) { return 2 * mf_table_halves(nterm); }

static int mf_chunk(long tiles, int ny, int per_cu)
{
    const long cap = 256L * per_cu;
    long best_cost = -1;
    int best = ny;
    for (int n = 1; n <= ny; ++n) {
        const int len = (ny + n - 1) / n;
        const long cost = ((tiles * n + cap - 1) / cap) * (len + 1);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = len;
        }
    }
    return best;
}

#define MF_LAUNCH(NT, M, SP) hipLaunchKernelGGL((mf_sweep_k<NT, M, SP>), dim3(grid), dim3(kMfThreads), 0, s, a)
#define MF_CASE(NT, M)                                                                                     \
    case M:                                                                                               \
        if constexpr (NT == 8) {                                                                          \
            if (split & 1)                                                                                \
                MF_LAUNCH(NT, M, 1);                                                                      \
            else                                                                                          \
                MF_LAUNCH(NT, M, 0);                                                                      \
        } else {                                                                                          \
            if (split == 0)                                                                               \
                MF_LAUNCH(NT, M, 0);                                                                      \
            else if (split == 1)                                                                          \
                MF_LAUNCH(NT, M, 1);                                                                      \
            else if (split == 2)                                                                          \
                MF_LAUNCH(NT, M, 2);                                                                      \
            else                                                                                          \
                MF_LAUNCH(NT, M, 3);                                                                      \
        }                                                                                                 \
        break;

int launch_mf_sweep(int mode, int nterm, const Grid &g, const uint16_t *Atab, const float *E, const uint8_t *mask,
                    const float *u, float *y, const float *b, const float *dinv, double omega, const float *rd,
                    hipStream_t s, int split)
{
    if (g.nelz <= 0 || (nterm != 1 && nterm != 8)) {
        set_error("the BF16 matrix-core sweep needs a 3D grid and 1 or 8 terms");
        return kStatusNotSupported;
    }
    if (mode == MODE_JAC && dinv == nullptr)
        mode = MODE_JACOWN;
    if (mode == MODE_JAC0 && nterm == 1 && rd != nullptr)
        mode = MODE_JAC0OWN;
    MfArgs a;
    a.g = g;
    a.mask = mask;
    a.u = u;
    a.y = y;
    a.b = b;
    a.dinv = dinv;
    a.E = E;
    a.A = reinterpret_cast<const mf_s4 *>(Atab);
    a.omega = (float)omega;
    for (int c = 0; c < 3; ++c)
        a.rd[c] = rd ? rd[c] : 0.f;
    a.tiles_x = (g.nx + kMfOut - 1) / kMfOut;
    a.tiles_z = (g.nz + kMfOut - 1) / kMfOut;
    const long cols = (long)a.tiles_x * a.tiles_z;
    a.chunk = mf_chunk(cols, g.ny, nterm == 1 ? 4 : 3);
    a.chunks = (g.ny + a.chunk - 1) / a.chunk;
    const long tiles = cols * a.chunks;
    const unsigned grid = (unsigned)((tiles + 7) / 8 * 8);
    if (nterm == 1) {
        switch (mode) {
            MF_CASE(1, MODE_KU)
            MF_CASE(1, MODE_RES)
            MF_CASE(1, MODE_JAC)
            MF_CASE(1, MODE_JACOWN)
            MF_CASE(1, MODE_JAC0)
            MF_CASE(1, MODE_JAC0OWN)
        default:
            set_error("mode %d not supported by the BF16 sweep", mode);
            return kStatusNotSupported;
        }
    } else {
        switch (mode) {
            MF_CASE(8, MODE_KU)
            MF_CASE(8, MODE_RES)
            MF_CASE(8, MODE_JAC)
            MF_CASE(8, MODE_JAC0)
        default:
            set_error("mode %d not supported by the BF16 Galerkin sweep", mode);
            return kStatusNotSupported;
        }
    }
    FANTOM_LAUNCH_CHECK("mf_sweep");
    return kStatusOk;
}

}  // namespace fantom

using namespace fantom;

extern "C" long fantom_sweep_bf16_table_size(int nterm)
{
    return (nterm == 1 || nterm == 8) ? mf_table_size(nterm) : -1;
}

extern "C" int fantom_sweep_bf16_tables(int nterm, const double *M, uint16_t *out)
{
    FANTOM_CHECK_ARG((nterm == 1 || nterm == 8) && M && out, "need nterm 1 or 8 and non-null operands");
    if (nterm == 8) {
        mf_tables(M, 8, out);
        return kStatusOk;
    }
    static const int loc[8] = {0, 1, 3, 2, 4, 5, 7, 6};   // natural corner -> reference local corner
    double Kn[576];
    for (int b = 0; b < 8; ++b)
        for (int c = 0; c < 3; ++c)
            for (int b2 = 0; b2 < 8; ++b2)
                for (int d = 0; d < 3; ++d)
                    Kn[(b * 3 + c) * 24 + b2 * 3 + d] = M[(loc[b] * 3 + c) * 24 + loc[b2] * 3 + d];
    mf_tables(Kn, 1, out);
    return kStatusOk;
}

extern "C" int fantom_sweep_bf16(int mode, int nterm, int split, int nelx, int nely, int nelz, const uint16_t *tab,
                                 const float *E, const uint8_t *mask, const float *u, float *y, const float *b,
                                 const float *dinv, double omega, const float *rd, void *stream)
{
    FANTOM_CHECK_ARG(split >= 0 && split <= 3 && (nterm == 1 || split <= 1), "split must be 0..3 (0..1 for nterm 8)");
    FANTOM_CHECK_ARG(grid_ok(nelx, nely, nelz) && nelz > 0, "bad 3D grid");
    FANTOM_CHECK_ARG(nterm == 1 || nterm == 8, "nterm must be 1 or 8");
    FANTOM_CHECK_ARG(mode == MODE_KU || mode == MODE_JAC || mode == MODE_RES || mode == MODE_JAC0,
                     "mode must be 0 (K u), 1 (Jacobi), 2 (residual) or 4 (two sweeps from zero)");
    FANTOM_CHECK_ARG(tab && E && y && (u || mode == MODE_JAC0), "null operand");
    FANTOM_CHECK_ARG(mode == MODE_KU || b, "mode needs b");
    FANTOM_CHECK_ARG(mode != MODE_JAC0 || dinv, "mode 4 needs dinv for x0");
    FANTOM_CHECK_ARG(nterm == 1 || mode == MODE_KU || mode == MODE_RES || dinv, "the Galerkin sweeps need dinv");
    FANTOM_CHECK_ARG(mode != MODE_JAC || dinv || rd, "mode 1 needs dinv or rd");
    return launch_mf_sweep(mode, nterm, make_grid(nelx, nely, nelz), tab, E, mask, u, y, b, dinv, omega,
                           nterm == 1 ? rd : nullptr, as_stream(stream), split);
}
