// ckd_features.h -- the 32 bf16 input features of the learned search-radius
// proposal, shared by the feature kernel (ckd_mlp.hip) and the query
// ordering kernel of the learned kNN (ckd_query_impl.h), which computes them
// while it holds each query's leaf.
#pragma once
#include "ckd_traverse.h"

namespace ckd {

__device__ __forceinline__ uint16_t f2bf(float x) {  // round to nearest even
    uint32_t u = __float_as_uint(x);
    u = u + 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// (1/3) ln(count / volume) of a node's tight point box, each extent at
// least 1e-6 L: the log inverse length scale of the local point density.
template <typename T>
__device__ __forceinline__ float density_feature(const TreeView<T> &tv, int32_t v, const double *Ls) {
    const T *b = tv.dbox + (int64_t)v * 6;
    double V = 1.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double e = (double)b[3 + d] - (double)b[d], m = 1e-6 * Ls[d];
        V = V * (e > m ? e : m);
    }
    return (float)(log((double)tv.range[v].count / V) / 3.0);
}

// Features of query q (3-D, periodic tree) whose leaf is `leaf`, written as
// 32 bf16 to out (16-byte aligned): 3 positions (q + L/2) / L, the 27 log1p
// counts of its 3x3x3 periodic cell neighbourhood in the G^3 grid (dz-major,
// looked up in `tab`), and the density feature of the smallest ancestor of
// the leaf (leaf included) holding at least 16 points, then at least 64.
template <typename T>
__device__ __forceinline__ void query_features(const TreeView<T> &tv, const T (&q)[3], int32_t leaf, const double *Ls,
                                               int G, const uint16_t *tab, uint16_t *out) {
    uint16_t f[32];
    int c[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double u = ((double)q[d] + Ls[d] / 2) / Ls[d];
        f[d] = f2bf((float)u);
        int v = (int)floor(u * G);
        c[d] = v < 0 ? 0 : (v > G - 1 ? G - 1 : v);
    }
    int j = 3;
#pragma unroll
    for (int dz = -1; dz <= 1; dz++)
#pragma unroll
        for (int dy = -1; dy <= 1; dy++)
#pragma unroll
            for (int dx = -1; dx <= 1; dx++) {
                const int x = (c[0] + dx + G) % G, y = (c[1] + dy + G) % G, z = (c[2] + dz + G) % G;
                f[j++] = tab[((int64_t)z * G + y) * G + x];
            }
    int32_t v = leaf;
    while (tv.range[v].count < 16 && tv.nodes[v].parent >= 0) v = tv.nodes[v].parent;
    f[30] = f2bf(density_feature<T>(tv, v, Ls));
    while (tv.range[v].count < 64 && tv.nodes[v].parent >= 0) v = tv.nodes[v].parent;
    f[31] = f2bf(density_feature<T>(tv, v, Ls));
    uint4 *dst = (uint4 *)out;
#pragma unroll
    for (int w = 0; w < 4; w++)
        dst[w] = make_uint4(f[8 * w] | ((uint32_t)f[8 * w + 1] << 16), f[8 * w + 2] | ((uint32_t)f[8 * w + 3] << 16),
                            f[8 * w + 4] | ((uint32_t)f[8 * w + 5] << 16), f[8 * w + 6] | ((uint32_t)f[8 * w + 7] << 16));
}

}  // namespace ckd
