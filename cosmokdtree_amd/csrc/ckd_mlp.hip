// ckd_mlp.hip -- the learned search-radius proposal: features, a 32-64-64-1
// GELU network (sigmoid form) on bf16 MFMA (v_mfma_f32_16x16x32_bf16, fp32
// accumulation),
// its backward pass and Adam.  No reference counterpart: the network only
// seeds the kNN pruning bound; a query that finds fewer than k points inside
// the seed falls back to the plain search (ckd_query_impl.h), so results never
// depend on it.
//
// Fragment layout of v_mfma_f32_16x16x32_bf16 in a wave of 64 lanes
// (lane l, r = l % 16, g = l / 16):
//   A (16 x 32):  a[j] = A[r][8g + j]          j < 8
//   B (32 x 16):  b[j] = B[8g + j][r]          j < 8
//   C (16 x 16):  c[j] = C[4g + j][r]          j < 4
// Parameters are one fp32 vector: W1[32][64] (row = input feature), b1[64],
// W2[64][64] (row = input unit), b2[64], w3[64], b3.
#include <cmath>

#include "ckd_features.h"

namespace ckd {

namespace {

constexpr int NIN = 32, NH = 64;
constexpr int OW1 = 0, OB1 = OW1 + NIN * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + NH;
constexpr int NPARAM = OB3 + 1;
constexpr int MT = 256;             // threads per block
constexpr int WPB = MT / WAVE;      // waves per block
constexpr int TR = 32;              // rows per wave tile

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));


// GELU in its sigmoid form, z * sigmoid(1.702 z) (Hendrycks & Gimpel,
// "Gaussian Error Linear Units", 2016): one exp and one reciprocal
// instead of erf.
constexpr float GELU_A = 1.702f;
__device__ __forceinline__ float sigm(float x) { return __builtin_amdgcn_rcpf(1.0f + __expf(-x)); }
__device__ __forceinline__ float gelu(float z) { return z * sigm(GELU_A * z); }
__device__ __forceinline__ float gelu_grad(float z) {
    const float sg = sigm(GELU_A * z);
    return sg + GELU_A * z * sg * (1.0f - sg);
}

// 8 consecutive bf16 (16 bytes) from LDS as an MFMA operand
__device__ __forceinline__ bf16x8 ld8(const uint16_t *p) {
    const uint4 v = *(const uint4 *)p;
    return __builtin_bit_cast(bf16x8, v);
}
// 8 bf16 with a stride (column of a row-major tile)
__device__ __forceinline__ bf16x8 ld8s(const uint16_t *p, int stride) {
    uint16_t t[8];
#pragma unroll
    for (int j = 0; j < 8; j++) t[j] = p[j * stride];
    uint4 v;
    v.x = t[0] | ((uint32_t)t[1] << 16);
    v.y = t[2] | ((uint32_t)t[3] << 16);
    v.z = t[4] | ((uint32_t)t[5] << 16);
    v.w = t[6] | ((uint32_t)t[7] << 16);
    return __builtin_bit_cast(bf16x8, v);
}
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// LDS of the forward-only kernel (37 KB: 4 blocks per CU) ...
struct MlpLdsF {
    uint16_t w1t[NH][NIN];   // W1 transposed: w1t[n][k] = W1[k][n]
    uint16_t w2t[NH][NH];    // w2t[o][i] = W2[i][o]
    float b1[NH], b2[NH], w3[NH], b3;
    uint16_t x[WPB][TR][NIN];
    uint16_t h1[WPB][TR][NH];
};
// ... and of the training kernel, which also needs W2 untransposed and a
// backward buffer.
struct MlpLds {
    uint16_t w1t[NH][NIN];
    uint16_t w2t[NH][NH];
    uint16_t w2r[NH][NH];    // w2r[i][o] = W2[i][o]
    float b1[NH], b2[NH], w3[NH], b3;
    uint16_t x[WPB][TR][NIN];
    uint16_t h1[WPB][TR][NH];
    uint16_t h2[WPB][TR][NH];  // backward: dz2, then dz1
};

template <typename S>
__device__ void load_params(S &s, const float *P) {
    constexpr bool need_w2r = sizeof(S) == sizeof(MlpLds);
    for (int e = threadIdx.x; e < NIN * NH; e += MT) {
        const int k = e / NH, n = e % NH;
        s.w1t[n][k] = f2bf(P[OW1 + e]);
    }
    for (int e = threadIdx.x; e < NH * NH; e += MT) {
        const int i = e / NH, o = e % NH;
        const uint16_t v = f2bf(P[OW2 + e]);
        s.w2t[o][i] = v;
        if constexpr (need_w2r) s.w2r[i][o] = v;
    }
    for (int e = threadIdx.x; e < NH; e += MT) {
        s.b1[e] = P[OB1 + e];
        s.b2[e] = P[OB2 + e];
        s.w3[e] = bf2f(f2bf(P[OW3 + e]));
    }
    if (threadIdx.x == 0) s.b3 = P[OB3];
    __syncthreads();
}

// Load a 32 x 32 bf16 input tile (rows >= n are zero).
__device__ __forceinline__ void load_x(uint16_t (*xs)[NIN], const uint16_t *X, int64_t row0, int64_t n) {
    const int l = threadIdx.x & (WAVE - 1);
    const int r = l >> 1, half = l & 1;
    uint4 v0 = make_uint4(0, 0, 0, 0), v1 = v0;
    if (row0 + r < n) {
        const uint4 *src = (const uint4 *)(X + (row0 + r) * NIN + half * 16);
        v0 = src[0];
        v1 = src[1];
    }
    *(uint4 *)&xs[r][half * 16] = v0;
    *(uint4 *)&xs[r][half * 16 + 8] = v1;
}

// Forward of one 32-row tile held in LDS (s.x[w]); z1/z2 pre-activations
// kept in C layout for the backward pass; y per (half, j) in every lane of
// the row's 16-lane group.
struct FwdState {
    f32x4 z1[2][4], z2[2][4];
    float y[2][4];
};

template <typename S>
__device__ __forceinline__ void forward_tile(S &s, int w, FwdState &st) {
    const int l = threadIdx.x & (WAVE - 1), r = l & 15, g = l >> 4;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const bf16x8 a = ld8(&s.x[w][h * 16 + r][g * 8]);
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            f32x4 c = {0, 0, 0, 0};
            c = mfma(a, ld8(&s.w1t[nt * 16 + r][g * 8]), c);
            const float bb = s.b1[nt * 16 + r];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                c[j] += bb;
                s.h1[w][h * 16 + g * 4 + j][nt * 16 + r] = f2bf(gelu(c[j]));
            }
            st.z1[h][nt] = c;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int h = 0; h < 2; h++) {
        float ys[4] = {0, 0, 0, 0};
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            f32x4 c = {0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < 2; ks++)
                c = mfma(ld8(&s.h1[w][h * 16 + r][ks * 32 + g * 8]), ld8(&s.w2t[nt * 16 + r][ks * 32 + g * 8]), c);
            const float bb = s.b2[nt * 16 + r], w3 = s.w3[nt * 16 + r];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                c[j] += bb;
                ys[j] += bf2f(f2bf(gelu(c[j]))) * w3;
            }
            st.z2[h][nt] = c;
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            float v = ys[j];
#pragma unroll
            for (int m = 1; m < 16; m <<= 1) v += __shfl_xor(v, m, WAVE);
            st.y[h][j] = v + s.b3;
        }
    }
}

// mode 0: out = y (log r); mode 1: out = scale * exp(y) (a seed radius)
__global__ __launch_bounds__(MT) void k_mlp_forward(const float *P, const uint16_t *X, int64_t n, float scale,
                                                    int mode, float *out) {
    __shared__ MlpLdsF s;
    load_params(s, P);
    const int w = threadIdx.x / WAVE, l = threadIdx.x & (WAVE - 1), r = l & 15, g = l >> 4;
    const int64_t ntile = ceil_div(n, TR);
    const int64_t iters = ceil_div(ntile, (int64_t)gridDim.x * WPB);
    for (int64_t it = 0; it < iters; it++) {
        const int64_t tile = (it * gridDim.x + blockIdx.x) * WPB + w;
        const int64_t row0 = tile * TR;
        load_x(s.x[w], X, row0, n);
        __syncthreads();
        FwdState st;
        forward_tile(s, w, st);
        if (r == 0) {
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int64_t row = row0 + h * 16 + g * 4 + j;
                    if (row < n) out[row] = mode == 1 ? scale * expf(st.y[h][j]) : st.y[h][j];
                }
        }
        __syncthreads();
    }
}

// Forward + backward of the MSE loss over B rows; gradients (and the loss
// sum) are added atomically into grad[NPARAM] / loss.
__global__ __launch_bounds__(MT) void k_mlp_grad(const float *P, const uint16_t *X, const float *t, int64_t B,
                                                 float *grad, float *loss) {
    __shared__ MlpLds s;
    load_params(s, P);
    const int w = threadIdx.x / WAVE, l = threadIdx.x & (WAVE - 1), r = l & 15, g = l >> 4;
    f32x4 gw2[4][4], gw1[2][4];
    float gb1[4], gb2[4], gw3[4], gb3 = 0, lsum = 0;
#pragma unroll
    for (int a = 0; a < 4; a++) {
        gb1[a] = gb2[a] = gw3[a] = 0;
#pragma unroll
        for (int b = 0; b < 4; b++) gw2[a][b] = f32x4{0, 0, 0, 0};
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) gw1[a][b] = f32x4{0, 0, 0, 0};
    const float invB = 2.0f / (float)B;
    const int64_t ntile = ceil_div(B, TR);
    const int64_t iters = ceil_div(ntile, (int64_t)gridDim.x * WPB);
    for (int64_t it = 0; it < iters; it++) {
        const int64_t tile = (it * gridDim.x + blockIdx.x) * WPB + w;
        const int64_t row0 = tile * TR;
        load_x(s.x[w], X, row0, B);
        __syncthreads();
        FwdState st;
        forward_tile(s, w, st);
        // output layer: dy, gw3, gb3, dz2
#pragma unroll
        for (int h = 0; h < 2; h++) {
            float dy[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int64_t row = row0 + h * 16 + g * 4 + j;
                const float e = row < B ? st.y[h][j] - t[row] : 0.0f;
                dy[j] = e * invB;
                if (r == 0) lsum += e * e;
                if (r == 0) gb3 += dy[j];
            }
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                const float w3 = s.w3[nt * 16 + r];
                float pw3 = 0, pb2 = 0;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const float z = st.z2[h][nt][j];
                    pw3 += bf2f(f2bf(gelu(z))) * dy[j];
                    const float dz = dy[j] * w3 * gelu_grad(z);
                    pb2 += dz;
                    s.h2[w][h * 16 + g * 4 + j][nt * 16 + r] = f2bf(dz);
                }
                gw3[nt] += pw3;
                gb2[nt] += pb2;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // gW2[i][o] += sum_r h1[r][i] dz2[r][o]   (K = the 32 rows)
#pragma unroll
        for (int it2 = 0; it2 < 4; it2++) {
            const bf16x8 a = ld8s(&s.h1[w][g * 8][it2 * 16 + r], NH);
#pragma unroll
            for (int ot = 0; ot < 4; ot++) gw2[it2][ot] = mfma(a, ld8s(&s.h2[w][g * 8][ot * 16 + r], NH), gw2[it2][ot]);
        }
        // dz1 = (dz2 W2^T) * gelu'(z1)
        f32x4 dz1[2][4];
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int it2 = 0; it2 < 4; it2++) {
                f32x4 c = {0, 0, 0, 0};
#pragma unroll
                for (int ks = 0; ks < 2; ks++)
                    c = mfma(ld8(&s.h2[w][h * 16 + r][ks * 32 + g * 8]), ld8(&s.w2r[it2 * 16 + r][ks * 32 + g * 8]), c);
#pragma unroll
                for (int j = 0; j < 4; j++) c[j] *= gelu_grad(st.z1[h][it2][j]);
                dz1[h][it2] = c;
            }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int it2 = 0; it2 < 4; it2++) {
                float pb1 = 0;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    pb1 += dz1[h][it2][j];
                    s.h2[w][h * 16 + g * 4 + j][it2 * 16 + r] = f2bf(dz1[h][it2][j]);
                }
                gb1[it2] += pb1;
            }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // gW1[f][i] += sum_r x[r][f] dz1[r][i]
#pragma unroll
        for (int ft = 0; ft < 2; ft++) {
            const bf16x8 a = ld8s(&s.x[w][g * 8][ft * 16 + r], NIN);
#pragma unroll
            for (int it2 = 0; it2 < 4; it2++) gw1[ft][it2] = mfma(a, ld8s(&s.h2[w][g * 8][it2 * 16 + r], NH), gw1[ft][it2]);
        }
        __syncthreads();
    }
    // reduce the per-lane partial vectors over the 4 lane groups, then add
#pragma unroll
    for (int a = 0; a < 4; a++) {
#pragma unroll
        for (int m = 16; m < WAVE; m <<= 1) {
            gb1[a] += __shfl_xor(gb1[a], m, WAVE);
            gb2[a] += __shfl_xor(gb2[a], m, WAVE);
            gw3[a] += __shfl_xor(gw3[a], m, WAVE);
        }
    }
#pragma unroll
    for (int m = 16; m < WAVE; m <<= 1) {
        gb3 += __shfl_xor(gb3, m, WAVE);
        lsum += __shfl_xor(lsum, m, WAVE);
    }
    if (g == 0) {
#pragma unroll
        for (int a = 0; a < 4; a++) {
            atomicAdd(&grad[OB1 + a * 16 + r], gb1[a]);
            atomicAdd(&grad[OB2 + a * 16 + r], gb2[a]);
            atomicAdd(&grad[OW3 + a * 16 + r], gw3[a]);
        }
    }
    if (l == 0) {
        atomicAdd(&grad[OB3], gb3);
        atomicAdd(loss, lsum);
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                atomicAdd(&grad[OW2 + (a * 16 + g * 4 + j) * NH + b * 16 + r], gw2[a][b][j]);
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                atomicAdd(&grad[OW1 + (a * 16 + g * 4 + j) * NH + b * 16 + r], gw1[a][b][j]);
}

__global__ void k_adam(float *P, const float *G, float *M, float *V, int n, float lr, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = G[i];
    const float m = 0.9f * M[i] + 0.1f * g;
    const float v = 0.999f * V[i] + 0.001f * g * g;
    M[i] = m;
    V[i] = v;
    P[i] -= lr * (m / c1) / (sqrtf(v / c2) + 1e-8f);
}

__global__ void k_zero_i32(int32_t *p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_zero_f32(float *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// Occupancy grid of the tree's points: G^3 cells over [-L/2, L/2)^3.
template <typename T>
__global__ void k_occupancy(const T *pts, int64_t n, int ps, double L0, double L1, double L2, int G, int32_t *grid) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double Ls[3] = {L0, L1, L2};
    int c[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double x = (double)pts[i * ps + d];
        int v = (int)floor((x + Ls[d] / 2) / Ls[d] * G);
        c[d] = v < 0 ? 0 : (v > G - 1 ? G - 1 : v);
    }
    atomicAdd(&grid[((int64_t)c[2] * G + c[1]) * G + c[0]], 1);
}

// bf16 log1p of every cell count (the 27 neighbourhood features are
// lookups in it).
__global__ void k_logtab(const int32_t *grid, int64_t cells, uint16_t *tab) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cells) tab[i] = f2bf((float)log1p((double)grid[i]));
}

// The 32 bf16 features of each query (query_features).
template <typename T>
__global__ void k_features(TreeView<T> tv, const T *q, int64_t nq, double L0, double L1, double L2, int G,
                           const uint16_t *tab, const int32_t *leaf_of, uint16_t *X) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const double Ls[3] = {L0, L1, L2};
    T qq[3];
#pragma unroll
    for (int d = 0; d < 3; d++) qq[d] = q[i * 3 + d];
    const int32_t leaf = leaf_of ? leaf_of[i] : find_leaf<T, 3>(tv, qq, 0);
    query_features<T>(tv, qq, leaf, Ls, G, tab, X + i * NIN);
}

unsigned grid_blocks(int64_t rows) {
    const int64_t tiles = ceil_div(rows, TR * WPB);
    return (unsigned)(tiles < 1024 ? (tiles > 0 ? tiles : 1) : 1024);
}

}  // namespace
}  // namespace ckd

using namespace ckd;

extern "C" {

int ckd_mlp_nparam(void) { return NPARAM; }

int ckd_occupancy(const ckd_tree *t, int32_t G, int32_t *grid, void *stream) {
    if (!t || !grid || G < 1 || G > 1024 || t->ndim != 3 || !t->periodic) {
        set_error("ckd_occupancy needs a periodic 3-D tree and 1 <= G <= 1024");
        return CKD_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    const int64_t cells = (int64_t)G * G * G;
    k_zero_i32<<<(unsigned)ceil_div(cells, 256), 256, 0, s>>>(grid, cells);
    const unsigned g = (unsigned)ceil_div(t->n, 256);
    if (t->prec == 4)
        k_occupancy<float><<<g, 256, 0, s>>>((const float *)t->pts, t->n, pstride(3), t->L[0], t->L[1], t->L[2], G, grid);
    else
        k_occupancy<double><<<g, 256, 0, s>>>((const double *)t->pts, t->n, pstride(3), t->L[0], t->L[1], t->L[2], G,
                                              grid);
    CKD_HIP(hipGetLastError());
    return CKD_OK;
}

int ckd_mlp_table(const int32_t *grid, int32_t G, uint16_t *table, void *stream) {
    if (!grid || !table || G < 1 || G > 1024) {
        set_error("invalid ckd_mlp_table arguments");
        return CKD_ERR_ARG;
    }
    const int64_t cells = (int64_t)G * G * G;
    k_logtab<<<(unsigned)ceil_div(cells, 256), 256, 0, (hipStream_t)stream>>>(grid, cells, table);
    CKD_HIP(hipGetLastError());
    return CKD_OK;
}

int ckd_mlp_features(const ckd_tree *t, const void *queries, int64_t nq, int32_t G, const uint16_t *table,
                     const int32_t *leaf_of, uint16_t *X, void *stream) {
    if (!t || t->ndim != 3 || !t->periodic || nq < 0 || !table || G < 1 || G > 1024 ||
        (nq > 0 && (!queries || !X))) {
        set_error("invalid ckd_mlp_features arguments (a periodic 3-D tree is required)");
        return CKD_ERR_ARG;
    }
    if (nq == 0) return CKD_OK;
    hipStream_t s = (hipStream_t)stream;
    const unsigned g = (unsigned)ceil_div(nq, 256);
    if (t->prec == 4)
        k_features<float><<<g, 256, 0, s>>>(view<float>(t), (const float *)queries, nq, t->L[0], t->L[1], t->L[2], G,
                                            table, leaf_of, X);
    else
        k_features<double><<<g, 256, 0, s>>>(view<double>(t), (const double *)queries, nq, t->L[0], t->L[1], t->L[2],
                                             G, table, leaf_of, X);
    CKD_HIP(hipGetLastError());
    return CKD_OK;
}

int ckd_mlp_forward(const float *params, const uint16_t *X, int64_t n, float scale, int32_t mode, float *out,
                    void *stream) {
    if (!params || n < 0 || (n > 0 && (!X || !out))) {
        set_error("invalid ckd_mlp_forward arguments");
        return CKD_ERR_ARG;
    }
    if (n == 0) return CKD_OK;
    k_mlp_forward<<<grid_blocks(n), MT, 0, (hipStream_t)stream>>>(params, X, n, scale, mode, out);
    CKD_HIP(hipGetLastError());
    return CKD_OK;
}

int ckd_mlp_grad(const float *params, const uint16_t *X, const float *target, int64_t B, float *grad, float *loss,
                 void *stream) {
    if (!params || !X || !target || !grad || !loss || B < 1) {
        set_error("invalid ckd_mlp_grad arguments");
        return CKD_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    k_zero_f32<<<(NPARAM + 1 + 255) / 256, 256, 0, s>>>(grad, NPARAM);
    k_zero_f32<<<1, 32, 0, s>>>(loss, 1);
    const int64_t tiles = ceil_div(B, TR * WPB * 4);  // about four tiles per wave
    const unsigned g = (unsigned)(tiles < 1024 ? tiles : 1024);
    k_mlp_grad<<<g, MT, 0, s>>>(params, X, target, B, grad, loss);
    CKD_HIP(hipGetLastError());
    return CKD_OK;
}

int ckd_mlp_adam(float *params, const float *grad, float *m, float *v, int64_t step, float lr, void *stream) {
    if (!params || !grad || !m || !v || step < 1) {
        set_error("invalid ckd_mlp_adam arguments");
        return CKD_ERR_ARG;
    }
    const float c1 = 1.0f - powf(0.9f, (float)step), c2 = 1.0f - powf(0.999f, (float)step);
    k_adam<<<(NPARAM + 255) / 256, 256, 0, (hipStream_t)stream>>>(params, grad, m, v, NPARAM, lr, c1, c2);
    CKD_HIP(hipGetLastError());
    return CKD_OK;
}

}  // extern "C"
