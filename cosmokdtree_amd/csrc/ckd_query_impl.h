// ckd_query_impl.h -- single-tree kNN, ball and box kernels and their host
// dispatch, templated on the coordinate type; the traversals themselves are
// in ckd_traverse.h.  Included once per precision (ckd_query_f32.hip,
// ckd_query_f64.hip) so the two halves compile in parallel.
#pragma once
#include <cstdlib>

#include "ckd_features.h"

namespace ckd {

namespace {

constexpr int QT = 256;  // threads per block
constexpr int64_t ORDER_MIN_QUERIES = 4096;  // below this, queries run in input order

// Scratch bytes of a kNN call: the leaf ordering of the queries (leaf_of,
// rank, order as int32 x nq; a flag set when the queries are not the
// tree's points; leaf counts and offsets as int64 x nnodes+1; scan scratch), taken
// when the batch is large enough to be ordered, and
// the re-search lists of a seeded call (flag and offsets int64 x nq+1,
// two lists int32 x nq, next seeds float x nq, a counter, scan scratch).
inline size_t order_scratch_bytes(const ckd_tree *t, int64_t nq) {
    const size_t a = 3 * sizeof(int32_t) * (size_t)nq, b = sizeof(int64_t) * (size_t)(t->nnodes + 1);
    return a + 2 * b + scan_tmp_bytes(t->nnodes) + 512;
}
inline bool orders_queries(const ckd_tree *t, int64_t nq) {
    return nq >= ORDER_MIN_QUERIES && nq < INT32_MAX && t->nnodes > 1;
}
inline size_t retry_scratch_bytes(int64_t nq) {
    const size_t nb = (size_t)nq;
    return 2 * 8 * (nb + 1) + 2 * 4 * nb + 4 * nb + 64 + scan_tmp_bytes(nq) + 8 * 256;  // 7 aligned takes
}

// Query ordering: each query's leaf, and its rank among the queries of
// that leaf; a counting sort then lists the queries leaf by leaf so a wave
// works on neighbouring queries.  `differ` is the flag k_self_order left
// when the queries may be the tree's own points: 0 means the ordering is
// already made and these passes return at once (nullptr: they always run).
template <typename T, int D>
__global__ __launch_bounds__(QT) void k_locate(TreeView<T> tv, const T *queries, int64_t nq, int32_t *leaf_of,
                                               int32_t *rank, int64_t *leaf_cnt, const int32_t *differ) {
    if (differ && *differ == 0) return;
    const int64_t qi = (int64_t)blockIdx.x * QT + threadIdx.x;
    if (qi >= nq) return;
    T q[D];
#pragma unroll
    for (int d = 0; d < D; d++) q[d] = queries[qi * D + d];
    const int32_t leaf = find_leaf<T, D>(tv, q, 0);
    leaf_of[qi] = leaf;
    rank[qi] = (int32_t)atomicAdd((unsigned long long *)&leaf_cnt[leaf], 1ull);
}

// k_locate plus the learned features of each query (3-D periodic trees),
// computed while the leaf is at hand; with the ordering already made the
// features are still written, from the leaves k_self_order found.
template <typename T>
__global__ __launch_bounds__(QT) void k_locate_feat(TreeView<T> tv, const T *queries, int64_t nq, int32_t *leaf_of,
                                                    int32_t *rank, int64_t *leaf_cnt, double L0, double L1,
                                                    double L2, int G, const uint16_t *tab, uint16_t *X,
                                                    const int32_t *differ) {
    const bool made = differ && *differ == 0;
    const int64_t qi = (int64_t)blockIdx.x * QT + threadIdx.x;
    if (qi >= nq) return;
    T q[3];
#pragma unroll
    for (int d = 0; d < 3; d++) q[d] = queries[qi * 3 + d];
    int32_t leaf;
    if (made) {
        leaf = leaf_of[qi];
    } else {
        leaf = find_leaf<T, 3>(tv, q, 0);
        leaf_of[qi] = leaf;
        rank[qi] = (int32_t)atomicAdd((unsigned long long *)&leaf_cnt[leaf], 1ull);
    }
    const double Ls[3] = {L0, L1, L2};
    query_features<T>(tv, q, leaf, Ls, G, tab, X + qi * 32);
}

__global__ void k_order(const int32_t *leaf_of, const int32_t *rank, const int64_t *leaf_off, int64_t nq,
                        int32_t *order, const int32_t *differ) {
    if (differ && *differ == 0) return;
    const int64_t qi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    order[leaf_off[leaf_of[qi]] + rank[qi]] = (int32_t)qi;
}

// Self-query ordering (nq == n, one point set).  The tree stores its rows
// leaf by leaf, so taking query idx[r] at position r lists the queries leaf
// by leaf with no locate pass and no sort, and a query's leaf is the leaf
// holding its row -- if query idx[r] is row r's point.  Two passes: each
// leaf writes its id over its rows (k_row_leaf), then one thread per row
// writes the order and the query's leaf, and every 16th row compares its
// query's coordinates with the row's; a wave that finds one differing sets
// *differ, which makes k_locate and k_order redo the ordering from the
// coordinates.  Results never depend on the ordering (it only groups the
// walks and picks the rows that prefill the lists), so the check only has
// to tell the tree's own points from other points, and a sample does: a
// batch that differs in a few rows is still ordered well by the rows.
template <typename T>
__global__ __launch_bounds__(QT) void k_row_leaf(TreeView<T> tv, int32_t *row_leaf) {
    const int64_t v = (int64_t)blockIdx.x * QT + threadIdx.x;
    if (v >= tv.nnodes || tv.nodes[v].axis >= 0) return;
    const Range r = tv.range[v];
    for (int64_t p = r.start; p < r.start + r.count; p++) row_leaf[p] = (int32_t)v;
}

template <typename T, int D>
__global__ __launch_bounds__(QT) void k_self_order(TreeView<T> tv, const T *queries, const int32_t *row_leaf,
                                                   int32_t *leaf_of, int32_t *order, int32_t *differ) {
    constexpr int PS = pstride(D);
    const int64_t p = (int64_t)blockIdx.x * QT + threadIdx.x;
    bool same = true;
    if (p < tv.nrows) {
        const int64_t qi = tv.idx64 ? ((const int64_t *)tv.idx)[p] : (int64_t)((const int32_t *)tv.idx)[p];
        if ((p & 15) == 0) {
#pragma unroll
            for (int d = 0; d < D; d++) same = same && queries[qi * D + d] == tv.pts[p * PS + d];
        }
        order[p] = (int32_t)qi;
        leaf_of[qi] = row_leaf[p];
    }
    if (wave_any(!same) && (threadIdx.x & (WAVE - 1)) == 0) *differ = 1;
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_i64(int64_t *p, int64_t n, int64_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Per-wave reduction of the work counters, one atomic per counter.
__device__ __forceinline__ void flush_stats(Counters c, int64_t nqw, int64_t *stats) {
#pragma unroll
    for (int m = WAVE / 2; m >= 1; m >>= 1) {
        c.visited += __shfl_xor(c.visited, m, WAVE);
        c.tested += __shfl_xor(c.tested, m, WAVE);
        c.reruns += __shfl_xor(c.reruns, m, WAVE);
        c.inserted += __shfl_xor(c.inserted, m, WAVE);
        c.climbed += __shfl_xor(c.climbed, m, WAVE);
        nqw += __shfl_xor(nqw, m, WAVE);
    }
    if ((threadIdx.x & (WAVE - 1)) == 0 && (nqw > 0 || c.visited > 0)) {
        atomicAdd((unsigned long long *)&stats[0], (unsigned long long)c.visited);
        atomicAdd((unsigned long long *)&stats[1], (unsigned long long)c.tested);
        atomicAdd((unsigned long long *)&stats[2], (unsigned long long)c.reruns);
        atomicAdd((unsigned long long *)&stats[3], (unsigned long long)nqw);
        atomicAdd((unsigned long long *)&stats[4], (unsigned long long)c.inserted);
        atomicAdd((unsigned long long *)&stats[5], (unsigned long long)c.climbed);
    }
}

// One traversal per query (knn_one).  With nq_dev the query count is read
// on the device (the last retry round of the wave kernels; those queries
// are not counted again in stats[3]).
template <typename T, int D, bool PER, int KMAX>
__global__ __launch_bounds__(QT) void k_knn(TreeView<T> tv, const T *queries, int64_t nq, const int64_t *nq_dev,
                                            int k, int sorted, const float *seed, const int32_t *order,
                                            const int32_t *leaf_of, T *dout, int64_t *iout, int64_t *stats) {
    const int64_t t = (int64_t)blockIdx.x * QT + threadIdx.x;
    const int64_t n = nq_dev ? *nq_dev : nq;
    Counters c;
    if (t < n) {
        const int64_t qi = order ? (int64_t)order[t] : t;
        T q[D];
#pragma unroll
        for (int d = 0; d < D; d++) q[d] = queries[qi * D + d];
        const int32_t leaf = leaf_of ? leaf_of[qi] : find_leaf<T, D>(tv, q, 0);
        T seed2 = T(FLT_MAX);
        if (seed) {
            const T sr = (T)seed[qi];
            seed2 = sr * sr;
        }
        knn_one<T, D, PER, KMAX>(tv, q, leaf, k, sorted, seed2, dout + qi * k, iout + qi * k, 0, c);
    }
    if (stats) flush_stats(c, (t < n && !nq_dev) ? 1 : 0, stats);
}

// Where the wave kernels send a seeded query whose seed held fewer than k
// points: with `flag`, round one marks its position (flag[t] = 1, compacted
// in leaf order afterwards) and proposes a larger seed from the m points it
// found, s' = 1.25 s ((k + 1) / (m + 1))^(1/3), the radius that would hold
// k + 1 points at the density seen inside s; otherwise (round two) the query
// index is appended to `list` for the unseeded per-query kernel.
struct Retry {
    int64_t *flag = nullptr;
    float *seed_next = nullptr;
    int32_t *list = nullptr;
    int64_t *count = nullptr;
};

__device__ __forceinline__ void send_retry(const Retry &rt, int64_t t, int64_t qi, float seed, int found, int k) {
    if (rt.flag) {
        rt.flag[t] = 1;
        rt.seed_next[qi] = 1.25f * seed * cbrtf((float)(k + 1) / (float)(found + 1));
    } else {
        rt.list[atomicAdd((unsigned long long *)rt.count, 1ull)] = (int32_t)qi;
    }
}

// Wave-uniform variant (knn_wave): queries in leaf order, one traversal per
// wave, parent-pointer walk (any tree depth).  nq_dev as for k_knn.
template <typename T, int D, bool PER, int KMAX>
__global__ __launch_bounds__(QT, 3) void k_knn_w(TreeView<T> tv, const T *queries, int64_t nq,
                                              const int64_t *nq_dev, int k, int sorted, const float *seed,
                                              const int32_t *order, T *dout, int64_t *iout, int64_t *stats,
                                              Retry rt) {
    __shared__ T buf_d[CBUF * QT];
    __shared__ int32_t buf_s[CBUF * QT];
    const int64_t blk = xcd_block();
    const int64_t t = blk * QT + threadIdx.x;
    const int64_t n = nq_dev ? *nq_dev : nq;
    if (blk * QT + (threadIdx.x & ~(WAVE - 1)) >= n) return;  // whole wave idle
    Counters c;
    const bool active = t < n;
    const int64_t qi = active ? (order ? (int64_t)order[t] : t) : 0;
    T q[D];
#pragma unroll
    for (int d = 0; d < D; d++) q[d] = active ? queries[qi * D + d] : T(0);
    T seed2 = T(FLT_MAX);
    float sr = 0.0f;
    if (seed && active) {
        sr = seed[qi];
        seed2 = (T)sr * (T)sr;
    }
    int found = 0;
    const bool done = knn_wave<T, D, PER, KMAX>(tv, q, active, 0, k, sorted, seed2, dout + qi * k, iout + qi * k,
                                                0, c, buf_d, buf_s, threadIdx.x, QT, found);
    if (!done) send_retry(rt, t, qi, sr, found, k);
    if (stats) {
        if (!active) c = Counters();
        flush_stats(c, active && !nq_dev ? 1 : 0, stats);
    }
}

// Stack variant (knn_stack) over the wide records; SD >= tree depth.
// Candidate buffering pays from k = 16 on (below, the short list update is
// cheaper than the LDS round trip and the occupancy the buffer costs); 8
// slots per lane in double precision keep the LDS within 4 blocks per CU.
// At k <= 32 in float32 with one lane per query (91 registers) the LDS sets
// the occupancy: 12 slots per lane (24 KB) and the 4 KB stack fit 5 blocks
// per CU, 5 waves per SIMD; 16 slots fitted 4, and the self-query of the
// clustered 2^24 box took 32.6 ms instead of 28.5.  In double precision at
// k <= 32 the launch bound asks for 4 waves per SIMD (128 registers, no
// scratch): the same self-query in float64 took 47.3 ms instead of 54.4 at
// 3 waves (143 registers); 5 waves spilled (720 ms).
// G lanes per query (1, 2, 4, ..., 64; 64 / G queries per wave): the grid
// covers nq * G lanes.  GC = 1 compiles the one-lane-per-query walk with G
// folded away (the group scans cost registers and time even when G is 1 at
// run time); GC = 0 takes G from the argument.  ST = false compiles the
// work counters away (calls without `stats`): live, their 64-bit sums held
// ten registers and were spilled to scratch inside the walk.
template <typename T, int D, bool PER, int KMAX, int SD, int GC, bool ST>
__global__ __launch_bounds__(QT, KMAX <= 32 && GC == 1 ? (sizeof(T) == 4 ? 5 : 4) : (sizeof(T) == 4 ? 4 : 3)) void k_knn_s(TreeView<T> tv, int2 root, const T *queries, int64_t nq,
                                              const int64_t *nq_dev, int k, int sorted, const float *seed,
                                              const int32_t *order, const int32_t *leaf_of, T *dout,
                                              int64_t *iout, int64_t *stats, Retry rt, int Grt) {
    const int G = GC ? GC : Grt;
    constexpr int CB =
        KMAX < 16 ? 0 : (sizeof(T) == 4 ? (GC == 1 ? (KMAX == 16 ? 8 : (KMAX == 32 ? 12 : 16)) : 16) : 8);
    __shared__ T buf_d[CB > 0 ? CB * QT : 1];
    __shared__ int32_t buf_s[CB > 0 ? CB * QT : 1];
    __shared__ StackEntry<T, D> stk[QT / WAVE][SD];
    const int64_t blk = xcd_block();
    const int lane = threadIdx.x & (WAVE - 1), sub = lane & (G - 1), gbase = lane & ~(G - 1);
    const int64_t t = (blk * QT + threadIdx.x) / G;
    const int64_t n = nq_dev ? *nq_dev : nq;
    if ((blk * QT + (threadIdx.x & ~(WAVE - 1))) / G >= n) return;  // whole wave idle
    Counters c;
    const bool active = t < n;
    const int64_t qi = active ? (order ? (int64_t)order[t] : t) : 0;
    T q[D];
#pragma unroll
    for (int d = 0; d < D; d++) q[d] = active ? queries[qi * D + d] : T(0);
    T seed2 = T(FLT_MAX);
    float sr = 0.0f;
    if (seed && active) {
        sr = seed[qi];
        seed2 = (T)sr * (T)sr;
    }
    int found = 0;
    int32_t leaf = -1;  // prefill from the leaf's ancestor (one lane per query, inner root)
    if (G == 1 && active && root.y < 0) leaf = leaf_of ? leaf_of[qi] : find_leaf<T, D>(tv, q, 0);
    const bool done = knn_stack<T, D, PER, KMAX, SD, CB>(tv, root, q, active, k, sorted, seed2, dout + qi * k,
                                                          iout + qi * k, 0, c, buf_d, buf_s, threadIdx.x, QT,
                                                          stk[threadIdx.x / WAVE], found, G, sub, gbase, leaf);
    if (!done && sub == 0) send_retry(rt, t, qi, sr, found, k);
    if (ST && stats) {
        if (!active || sub != 0) c = Counters();
        flush_stats(c, active && sub == 0 && !nq_dev ? 1 : 0, stats);
    }
}

// SPH smoothing length and density of each query, fused after the stack
// walk (sph_row): writes h[qi], rho[qi] instead of the k-row.
template <typename T, bool PER, int KMAX, int SD, bool NOIDX>
__global__ __launch_bounds__(QT, 3) void k_sph_s(TreeView<T> tv, int2 root, const T *queries, int64_t nq, int k,
                                              const int32_t *order, const int32_t *leaf_of, const T *mass,
                                              T *hout, T *rhoout) {
    constexpr int D = 3;
    // 12 slots with rows, as k_knn_s; without rows 16 (4 bytes a slot: the
    // LDS does not limit the occupancy, and fewer flushes pay)
    constexpr int CB = KMAX < 16 ? 0 : (sizeof(T) == 4 ? (KMAX == 32 && !NOIDX ? 12 : 16) : 8);
    __shared__ T buf_d[CB > 0 ? CB * QT : 1];
    __shared__ int32_t buf_s[CB > 0 ? CB * QT : 1];
    __shared__ StackEntry<T, D> stk[QT / WAVE][SD];
    const int64_t blk = xcd_block();
    const int64_t t = blk * QT + threadIdx.x;
    if (blk * QT + (threadIdx.x & ~(WAVE - 1)) >= nq) return;  // whole wave idle
    Counters c;
    const bool active = t < nq;
    const int64_t qi = active ? (order ? (int64_t)order[t] : t) : 0;
    T q[D];
#pragma unroll
    for (int d = 0; d < D; d++) q[d] = active ? queries[qi * D + d] : T(0);
    int found = 0;
    int32_t leaf = -1;
    if (active && root.y < 0) leaf = leaf_of ? leaf_of[qi] : find_leaf<T, D>(tv, q, 0);
    knn_stack<T, D, PER, KMAX, SD, CB, 1, NOIDX>(tv, root, q, active, k, 1, T(FLT_MAX), hout + qi, nullptr, 0, c, buf_d,
                                         buf_s, threadIdx.x, QT, stk[threadIdx.x / WAVE], found, 1, 0, 0, leaf, mass,
                                         rhoout + qi);
}

// Positions flagged by round one, compacted in order: list[off[t]] = query.
__global__ void k_compact_retry(const int64_t *flag, const int64_t *off, int64_t n, const int32_t *order,
                                int32_t *list) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n && flag[t]) list[off[t]] = order ? order[t] : (int32_t)t;
}

// kNN for k beyond the register list: the max-heap of the reference
// (max_heap_insert, cosmokdtree.f90:859-905) kept in the output row, leaf
// rows stored in the index row until the end.
template <typename T, int D, bool PER>
__global__ __launch_bounds__(QT) void k_knn_heap(TreeView<T> tv, const T *queries, int64_t nq, int k, T *dout,
                                                 int64_t *iout) {
    const int64_t qi = (int64_t)blockIdx.x * QT + threadIdx.x;
    if (qi >= nq) return;
    T q[D];
#pragma unroll
    for (int d = 0; d < D; d++) q[d] = queries[qi * D + d];
    T *hd = dout + qi * k;
    int64_t *hi = iout + qi * k;
    for (int j = 0; j < k; j++) {
        hd[j] = T(FLT_MAX);
        hi[j] = -1;
    }
    auto scan = [&](int32_t leaf) {
        const Range r = tv.range[leaf];
        T far = hd[0];
        for (int64_t p = r.start; p < r.start + r.count; p++) {
            T x[D];
            load_pt<T, D>(tv.pts, p, x);
            const T dc = dist2<T, D, PER>(x, q, tv.L);
            if (dc < far) {
                hd[0] = dc;
                hi[0] = p;
                int i = 0;
                for (;;) {
                    const int l = 2 * i + 1, rr = 2 * i + 2;
                    int lg = i;
                    if (l < k && hd[l] > hd[lg]) lg = l;
                    if (rr < k && hd[rr] > hd[lg]) lg = rr;
                    if (lg == i) break;
                    const T td = hd[i];
                    hd[i] = hd[lg];
                    hd[lg] = td;
                    const int64_t ti = hi[i];
                    hi[i] = hi[lg];
                    hi[lg] = ti;
                    i = lg;
                }
                far = hd[0];
            }
        }
    };
    const int32_t leaf = find_leaf<T, D>(tv, q, 0);
    scan(leaf);
    int32_t from = leaf, cur = tv.nodes[leaf].parent;
    while (cur >= 0) {
        const NodeRec<T> nr = tv.nodes[cur];
        const int32_t nearc = nr.child + (q[nr.axis] < nr.split ? 0 : 1);
        bool moved = false;
        if (from == nearc) {
            const int32_t farc = nearc == nr.child ? nr.child + 1 : nr.child;
            if (node_mindist<T, D, PER>(tv, farc, q) < hd[0]) {
                int32_t n2 = farc;
                for (;;) {
                    const NodeRec<T> r2 = tv.nodes[n2];
                    if (r2.axis < 0) {
                        scan(n2);
                        from = n2;
                        cur = r2.parent;
                        break;
                    }
                    const int32_t nc2 = r2.child + (q[r2.axis] < r2.split ? 0 : 1);
                    if (node_mindist<T, D, PER>(tv, nc2, q) < hd[0]) {
                        n2 = nc2;
                        continue;
                    }
                    from = nc2;
                    cur = n2;
                    break;
                }
                moved = true;
            }
        }
        if (!moved) {
            from = cur;
            cur = nr.parent;
        }
    }
    for (int j = 0; j < k; j++) {
        hd[j] = sqrt_rn<T>(hd[j]);
        hi[j] = hi[j] >= 0 ? load_idx(tv.idx, tv.idx64, hi[j]) : (int64_t)CKD_EMPTY_SLOT;
    }
}

// ---------------------------------------------------------------- ball / box
// Ball search with one walk per wave (ball_stack), queries in leaf order.
template <typename T, int D, bool PER, bool FILL, int SD>
__global__ __launch_bounds__(QT) void k_ball_s(TreeView<T> tv, int2 root, int2 rootrows, const T *queries, int64_t nq,
                                               const int32_t *order, T r2, int64_t *counts, const int64_t *offsets,
                                               T *dout, int64_t *iout) {
    __shared__ BallEntry stk[QT / WAVE][SD];
    __shared__ EmitEntry eq[QT / WAVE][FILL ? EMIT_CAP : 1];
    const int64_t blk = xcd_block();
    const int64_t t = blk * QT + threadIdx.x;
    if (blk * QT + (threadIdx.x & ~(WAVE - 1)) >= nq) return;  // whole wave idle
    const bool active = t < nq;
    const int64_t qi = active ? (order ? (int64_t)order[t] : t) : 0;
    T q[D];
#pragma unroll
    for (int d = 0; d < D; d++) q[d] = active ? queries[qi * D + d] : T(0);
    const int64_t w0 = FILL && active ? offsets[qi] : 0;
    const int wv = threadIdx.x / WAVE;
    const int64_t c = ball_stack<T, D, PER, FILL, SD>(tv, root, tv.dbox, rootrows, q, active, r2, dout, iout, w0, 0,
                                                      stk[wv], eq[wv]);
    if (!FILL && active) counts[qi] = c;
}

template <typename T, int D, bool PER, bool FILL>
__global__ __launch_bounds__(QT) void k_ball(TreeView<T> tv, const T *queries, int64_t nq, T r2, int64_t *counts,
                                             const int64_t *offsets, T *dout, int64_t *iout) {
    const int64_t qi = (int64_t)blockIdx.x * QT + threadIdx.x;
    if (qi >= nq) return;
    T q[D];
#pragma unroll
    for (int d = 0; d < D; d++) q[d] = queries[qi * D + d];
    const int64_t w = FILL ? offsets[qi] : 0;
    const int64_t cnt = ball_one<T, D, PER, FILL>(tv, q, 0, r2, FILL ? dout + w : nullptr, FILL ? iout + w : nullptr, 0);
    if (!FILL) counts[qi] = cnt;
}

// One-walk ball search (ball_stage): counts per query and staging records.
template <typename T, int D, bool PER, int SD>
__global__ __launch_bounds__(QT) void k_ball_stage(TreeView<T> tv, int2 root, int2 rootrows, const T *queries,
                                                   int64_t nq, const int32_t *order, T r2, int64_t *counts,
                                                   StageOut so) {
    __shared__ BallEntry stk[QT / WAVE][SD];
    const int64_t blk = xcd_block();
    const int64_t t = blk * QT + threadIdx.x;
    if (blk * QT + (threadIdx.x & ~(WAVE - 1)) >= nq) return;  // whole wave idle
    const bool active = t < nq;
    const int64_t qi = active ? (order ? (int64_t)order[t] : t) : 0;
    T q[D];
#pragma unroll
    for (int d = 0; d < D; d++) q[d] = active ? queries[qi * D + d] : T(0);
    const int64_t c = ball_stage<T, D, PER, SD>(tv, root, tv.dbox, rootrows, q, active, r2, (int32_t)qi, so,
                                                stk[threadIdx.x / WAVE]);
    if (active) counts[qi] = c;
}

// The staging records to their CSR places.  The blocks stride over the
// taken chunks; in a chunk, a thread per batch record (its accepted rows in
// row order); the subtrees taken whole are spread over the block
// afterwards.  The distances are computed again in the walk's formula (same
// bits), their square roots taken here.  Nothing is written when the
// staging area overflowed or the total exceeds `cap` (the caller then fills
// the rows with the second walk).
constexpr unsigned SCATTER_BLOCKS = 4096;

template <typename T, int D, bool PER>
__global__ __launch_bounds__(256) void k_ball_scatter(TreeView<T> tv, const T *queries, StageOut so, const int64_t *off,
                                                      int64_t nq, int64_t cap, T *dout, int64_t *iout) {
    if (*so.overflow || off[nq] > cap) return;
    const uint32_t taken = *so.next_chunk, used = taken < (uint32_t)so.nchunks ? taken : (uint32_t)so.nchunks;
    __shared__ int4 ranges[STAGE_CHUNK];
    __shared__ int nr;
#pragma unroll 1
    for (uint32_t c = blockIdx.x; c < used; c += gridDim.x) {
        if (threadIdx.x == 0) nr = 0;
        __syncthreads();
        const int fill = so.chunk_fill[c];
        CKD_DASSERT(fill >= 0 && fill <= STAGE_CHUNK);
#pragma unroll 1
        for (int r = threadIdx.x; r < fill; r += blockDim.x) {
            const int4 e = so.rec[(int64_t)c * STAGE_CHUNK + r];
            if ((uint32_t)e.x & STAGE_RANGE) {
                ranges[atomicAdd(&nr, 1)] = e;
                continue;
            }
            CKD_DASSERT(e.x >= 0 && e.x < nq && e.z >= 0 && e.z < tv.nrows);
            int64_t o = off[e.x] + e.y;
            T q[D];
#pragma unroll
            for (int d = 0; d < D; d++) q[d] = queries[(int64_t)e.x * D + d];
            uint32_t mk = (uint32_t)e.w;
            while (mk) {
                const int j = __builtin_ctz(mk);
                mk &= mk - 1;
                CKD_DASSERT(o < off[e.x + 1]);
                T x[D];
                load_pt<T, D>(tv.pts, e.z + j, x);
                dout[o] = sqrt_rn<T>(dist2<T, D, PER>(x, q, tv.L));
                iout[o] = load_idx(tv.idx, tv.idx64, e.z + j);
                o++;
            }
        }
        __syncthreads();
#pragma unroll 1
        for (int k = 0; k < nr; k++) {
            const int4 e = ranges[k];
            const int32_t qi = (int32_t)((uint32_t)e.x & ~STAGE_RANGE);
            CKD_DASSERT(qi >= 0 && qi < nq && e.z >= 0 && e.z + e.w <= tv.nrows);
            const int64_t o = off[qi] + e.y;
            CKD_DASSERT(o >= off[qi] && o + e.w <= off[qi + 1]);
            T q[D];
#pragma unroll
            for (int d = 0; d < D; d++) q[d] = queries[(int64_t)qi * D + d];
#pragma unroll 1
            for (int j = threadIdx.x; j < e.w; j += blockDim.x) {
                T x[D];
                load_pt<T, D>(tv.pts, e.z + j, x);
                dout[o + j] = sqrt_rn<T>(dist2<T, D, PER>(x, q, tv.L));
                iout[o + j] = load_idx(tv.idx, tv.idx64, e.z + j);
            }
        }
        __syncthreads();  // nr and ranges are reused by the next chunk
    }
}

template <typename T, int D, bool FILL>
__global__ __launch_bounds__(QT) void k_box(TreeView<T> tv, const T *boxes, int64_t nb, int64_t *counts,
                                            const int64_t *offsets, int64_t *iout) {
    const int64_t bi = (int64_t)blockIdx.x * QT + threadIdx.x;
    if (bi >= nb) return;
    T bx[2 * D];
#pragma unroll
    for (int d = 0; d < 2 * D; d++) bx[d] = boxes[bi * 2 * D + d];
    const int64_t cnt = box_one<T, D, FILL>(tv, bx, 0, FILL ? iout + offsets[bi] : nullptr, 0);
    if (!FILL) counts[bi] = cnt;
}

// Boxes taken modulo a periodic tree's box (wrap_pieces): every sub-box of
// the product of the axes' pieces (axis 0 fastest) is searched with the
// reference's rules (box_one), its rows written after the previous ones.
template <typename T, int D, bool FILL>
__global__ __launch_bounds__(QT) void k_box_wrap(TreeView<T> tv, const T *boxes, int64_t nb, int64_t *counts,
                                                 const int64_t *offsets, int64_t *iout) {
    const int64_t bi = (int64_t)blockIdx.x * QT + threadIdx.x;
    if (bi >= nb) return;
    T pa[D][3], pb[D][3];
    int np[D];
    int total = 1;
#pragma unroll
    for (int d = 0; d < D; d++) {
        np[d] = wrap_pieces<T>(boxes[bi * 2 * D + 2 * d], boxes[bi * 2 * D + 2 * d + 1], tv.L[d], pa[d], pb[d]);
        total *= np[d];
    }
    int64_t cnt = 0;
    int64_t *out = FILL ? iout + offsets[bi] : nullptr;
    for (int c = 0; c < total; c++) {
        T sb[2 * D];
        int r = c;
#pragma unroll
        for (int d = 0; d < D; d++) {
            const int j = r % np[d];
            r /= np[d];
            sb[2 * d] = j == 0 ? pa[d][0] : (j == 1 ? pa[d][1] : pa[d][2]);
            sb[2 * d + 1] = j == 0 ? pb[d][0] : (j == 1 ? pb[d][1] : pb[d][2]);
        }
        cnt += box_one<T, D, FILL>(tv, sb, 0, FILL ? out + cnt : nullptr, 0);
    }
    if (!FILL) counts[bi] = cnt;
}

// ------------------------------------------------------------ dispatch

// kNN kernel: CKD_KNN_LANE (one traversal per query, k_knn),
// CKD_KNN_WAVE (one traversal per wave, parent pointers, k_knn_w) or
// CKD_KNN_STACK (one traversal per wave over the wide records, k_knn_s).
// The tree's setting (ckd_knn_config) wins over the environment variable
// CKD_KNN_KERNEL=lane|wave|stack, which wins over the automatic choice.
int knn_kernel_env() {
    static int choice = -2;
    if (choice == -2) {
        const char *e = getenv("CKD_KNN_KERNEL");
        choice = !e ? -1 : (e[0] == 'l' ? CKD_KNN_LANE : e[0] == 'w' ? CKD_KNN_WAVE : e[0] == 's' ? CKD_KNN_STACK : -1);
    }
    return choice;
}

template <typename T, int D>
bool stack_ok(const ckd_tree *t, int sd) {
    return D <= 3 && t->wide != nullptr && t->depth <= sd;
}

// What ckd_knn_learned adds: the network that proposes the seeds.
struct Learned {
    const float *params;
    float margin;
    int G;
    const uint16_t *table;
};

template <typename T, int D, bool PER, int KMAX>
int launch_wave(const ckd_tree *t, const TreeView<T> &tv, int kind, int sd, int qpw, unsigned g, const T *q,
                int64_t nq, const int64_t *nq_dev, int k, int sorted, const float *seed, const int32_t *order,
                const int32_t *leaf_of, T *dist, int64_t *idx, int64_t *stats, const Retry &rt, hipStream_t s) {
    if (kind == CKD_KNN_STACK) {
        if constexpr (D <= 3) {
            const int2 root = make_int2(0, t->nnodes > 1 ? -1 : (int32_t)t->n);
            const int G = WAVE / qpw;
            const unsigned gs = (unsigned)ceil_div(nq * G, QT);
#define CKD_KNN_S(SD_, GC_)                                                                                    \
    do {                                                                                                   \
        if (stats)                                                                                         \
            k_knn_s<T, D, PER, KMAX, SD_, GC_, true><<<gs, QT, 0, s>>>(tv, root, q, nq, nq_dev, k, sorted, seed,  \
                                                                       order, leaf_of, dist, idx, stats, rt, G); \
        else                                                                                               \
            k_knn_s<T, D, PER, KMAX, SD_, GC_, false><<<gs, QT, 0, s>>>(tv, root, q, nq, nq_dev, k, sorted, seed, \
                                                                        order, leaf_of, dist, idx, stats, rt, G); \
    } while (0)
            if (sd <= 32) {
                if (G == 1) CKD_KNN_S(32, 1);
                else CKD_KNN_S(32, 0);
            } else {
                if (G == 1) CKD_KNN_S(64, 1);
                else CKD_KNN_S(64, 0);
            }
#undef CKD_KNN_S
        }
    } else if constexpr (D <= 3) {
        k_knn_w<T, D, PER, KMAX><<<g, QT, 0, s>>>(tv, q, nq, nq_dev, k, sorted, seed, order, dist, idx, stats, rt);
    }
    CKD_HIP(hipGetLastError());
    return CKD_OK;
}

// Events around the main traversal launch, when the tree records them.
struct LaunchTimer {
    const ckd_tree *t;
    bool allowed = true;  // false for the entry points that create nothing
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int start(hipStream_t s) {
        if (!t->timing || !allowed) return CKD_OK;
        CKD_HIP(hipEventCreate(&e0));
        CKD_HIP(hipEventCreate(&e1));
        CKD_HIP(hipEventRecord(e0, s));
        return CKD_OK;
    }
    int stop(hipStream_t s) {
        if (!t->timing || !allowed) return CKD_OK;
        CKD_HIP(hipEventRecord(e1, s));
        ckd_tree *tw = const_cast<ckd_tree *>(t);  // the event list only; the tree itself stays read-only
        std::lock_guard<std::mutex> lock(tw->timing_mu);
        tw->knn_events.emplace_back(e0, e1);
        return CKD_OK;
    }
};

template <typename T, int D, bool PER, int KMAX>
int knn_launch_k(const ckd_tree *t, const TreeView<T> &tv, int kind, int sd, int qpw, const T *q, int64_t nq, int k,
                 int sorted, const float *seed, const int32_t *order, const int32_t *leaf_of, T *dist, int64_t *idx,
                 int64_t *stats, Scratch &sc, hipStream_t s) {
    const unsigned g = (unsigned)ceil_div(nq, QT);
    LaunchTimer tm{t, sc.arena == nullptr};
    int rc;
    if (kind == CKD_KNN_LANE) {
        if ((rc = tm.start(s))) return rc;
        k_knn<T, D, PER, KMAX><<<g, QT, 0, s>>>(tv, q, nq, nullptr, k, sorted, seed, order, leaf_of, dist, idx, stats);
        CKD_HIP(hipGetLastError());
        return tm.stop(s);
    }
    if (!seed) {
        if ((rc = tm.start(s))) return rc;
        if ((rc = launch_wave<T, D, PER, KMAX>(t, tv, kind, sd, qpw, g, q, nq, nullptr, k, sorted, nullptr, order, leaf_of, dist, idx,
                                               stats, Retry(), s)))
            return rc;
        return tm.stop(s);
    }
    // Seeded: round one with the given seeds; round two, in leaf order, for
    // the queries whose seed held fewer than k points, with larger seeds;
    // round three, unseeded and one traversal per query, for what is left.
    // workspace: flag, off (int64 x nq+1), list, list2 (int32 x nq),
    // seed_next (float x nq), count2, scan scratch
    const size_t nb = (size_t)nq;
    void *ws = nullptr;
    if ((rc = sc.take(retry_scratch_bytes(nq), &ws))) return rc;
    char *p = (char *)ws;
    auto take = [&](size_t b) {
        char *r = p;
        p += (b + 255) & ~(size_t)255;
        return (void *)r;
    };
    int64_t *flag = (int64_t *)take(8 * (nb + 1)), *off = (int64_t *)take(8 * (nb + 1));
    int32_t *list = (int32_t *)take(4 * nb), *list2 = (int32_t *)take(4 * nb);
    float *seed_next = (float *)take(4 * nb);
    int64_t *count2 = (int64_t *)take(64);
    void *scr = take(scan_tmp_bytes(nq));
    k_fill_i64<<<(unsigned)ceil_div(nq, 256), 256, 0, s>>>(flag, nq, 0);
    k_fill_i64<<<1, 64, 0, s>>>(count2, 1, 0);
    Retry r1;
    r1.flag = flag;
    r1.seed_next = seed_next;
    if ((rc = tm.start(s))) return rc;
    if ((rc = launch_wave<T, D, PER, KMAX>(t, tv, kind, sd, qpw, g, q, nq, nullptr, k, sorted, seed, order, leaf_of, dist, idx,
                                           stats, r1, s)))
        return rc;
    if ((rc = tm.stop(s))) return rc;
    if ((rc = scan_i64(flag, off, nq, true, scr, s))) return rc;
    k_compact_retry<<<(unsigned)ceil_div(nq, 256), 256, 0, s>>>(flag, off, nq, order, list);
    Retry r2;
    r2.list = list2;
    r2.count = count2;
    if ((rc = launch_wave<T, D, PER, KMAX>(t, tv, kind, sd, qpw, g, q, nq, off + nq, k, sorted, seed_next, list, leaf_of, dist, idx,
                                           stats, r2, s)))
        return rc;
    k_knn<T, D, PER, KMAX><<<g, QT, 0, s>>>(tv, q, nq, count2, k, sorted, nullptr, list2, nullptr, dist, idx, stats);
    CKD_HIP(hipGetLastError());
    return CKD_OK;
}

// Scratch of the learned seeds: the features (nq x 32 bf16) and seeds (nq floats).
inline int learned_buffers(Scratch &sc, int64_t nq, uint16_t **X, float **seeds) {
    void *ws = nullptr;
    const int rc = sc.take((size_t)nq * (64 + 4) + 256, &ws);
    if (rc) return rc;
    *X = (uint16_t *)ws;
    *seeds = (float *)((char *)ws + (size_t)nq * 64);
    return CKD_OK;
}

// The leaf ordering of a batch (orders_queries) in scratch: order (the
// query at each position) and leaf_of (each query's leaf).  With X (learned
// seeds, 3-D trees) the same pass writes each query's features.  A batch as
// large as the tree's one point set first tries the tree's own row order
// (k_self_order); the locate pass and the counting sort then do their work
// only if the queries are not the tree's points, as a device flag decides,
// so the host does not wait.  The leaf of a row need not be the leaf the
// descent by split values reaches (a point on a split plane may have been
// placed on either side): only callers that use leaf_of as a hint (the
// stack walks' prefill, the features) may take it (self_ok); the walk of
// k_knn climbs from its leaf and needs the descent's leaf.
template <typename T, int D>
int order_queries(const ckd_tree *t, const TreeView<T> &tv, const T *q, int64_t nq, const int32_t **order,
                  const int32_t **leaf_of, Scratch &sc, hipStream_t s, const Learned *ln = nullptr,
                  uint16_t *X = nullptr, bool self_ok = true) {
    const unsigned g = (unsigned)ceil_div(nq, QT);
    const size_t a = 3 * sizeof(int32_t) * (size_t)nq;
    void *ws = nullptr;
    int rc = sc.take(order_scratch_bytes(t, nq), &ws);
    if (rc) return rc;
    int32_t *lo = (int32_t *)ws, *rk = lo + nq, *ord = rk + nq;
    int32_t *differ = (int32_t *)((char *)ws + ((a + 255) & ~(size_t)255));
    int64_t *cnt = (int64_t *)((char *)differ + 256);
    int64_t *off = cnt + t->nnodes + 1;
    void *scr = off + t->nnodes + 1;
    const int32_t *self = nullptr;
    if (self_ok && nq == t->n && t->nsets == 1) {
        k_fill_i32<<<1, 64, 0, s>>>(differ, 1, 0);
        k_row_leaf<T><<<(unsigned)ceil_div(t->nnodes, QT), QT, 0, s>>>(tv, rk);
        k_self_order<T, D><<<g, QT, 0, s>>>(tv, q, rk, lo, ord, differ);
        self = differ;
    }
    k_fill_i64<<<(unsigned)ceil_div(t->nnodes, 256), 256, 0, s>>>(cnt, t->nnodes, 0);
    bool feats = false;
    if constexpr (D == 3) {
        if (X) {
            k_locate_feat<T><<<g, QT, 0, s>>>(tv, q, nq, lo, rk, cnt, t->L[0], t->L[1], t->L[2], ln->G, ln->table,
                                              X, self);
            feats = true;
        }
    }
    if (!feats) k_locate<T, D><<<g, QT, 0, s>>>(tv, q, nq, lo, rk, cnt, self);
    if ((rc = scan_i64(cnt, off, t->nnodes, false, scr, s))) return rc;
    k_order<<<(unsigned)ceil_div(nq, 256), 256, 0, s>>>(lo, rk, off, nq, ord, self);
    CKD_HIP(hipGetLastError());
    *order = ord;
    if (leaf_of) *leaf_of = lo;
    return CKD_OK;
}

template <typename T, int D, bool PER>
int knn_launch(const ckd_tree *t, const T *q, int64_t nq, int k, int sorted, const float *seed, const Learned *ln,
               T *dist, int64_t *idx, int64_t *stats, Scratch &sc, hipStream_t s) {
    const TreeView<T> tv = view<T>(t);
    const unsigned g = (unsigned)ceil_div(nq, QT);
    if (k <= 64 && t->n < INT32_MAX) {
        // kernel and stack depth
        int kind = t->knn_kernel >= 0 ? t->knn_kernel : knn_kernel_env();
        int sd = t->knn_stack > 0 ? t->knn_stack : (t->depth <= 32 ? 32 : 64);
        if (sd < t->depth) sd = 64;
        int qpw = t->knn_qpw > 0 ? t->knn_qpw : 64;
        if (kind < 0) {
            // Automatic choice (measured on MI355X, 2^18 and 2^22 clustered
            // points, k = 8, 32, 64): a batch of at most 65536 queries with
            // k <= 32 and fewer queries than a quarter of the points runs one
            // query per wave (its 64 lanes split the leaf rows); otherwise
            // queries at least 1/20 of the points, or k > 32, share walks 64
            // per wave; a large sparse batch with k <= 32 runs one walk per
            // query.
            if (!stack_ok<T, D>(t, 64)) kind = k > 8 ? CKD_KNN_WAVE : CKD_KNN_LANE;
            else if (nq <= 65536 && k <= 32 && nq * 4 < t->n) {
                kind = CKD_KNN_STACK;
                if (t->knn_qpw == 0) qpw = 1;
            } else if (nq * 20 >= t->n || k > 32) kind = CKD_KNN_STACK;
            else kind = CKD_KNN_LANE;
        }
        if (kind == CKD_KNN_STACK && !stack_ok<T, D>(t, sd)) kind = CKD_KNN_WAVE;  // same results, any depth
        if (D > 3) kind = CKD_KNN_LANE;  // the wave kernels are built for ndim <= 3
        const int32_t *order = nullptr, *leaf_of = nullptr;
        uint16_t *X = nullptr;
        float *seeds = nullptr;
        bool feats_done = false;
        if (orders_queries(t, nq)) {
            int rc;
            if constexpr (D == 3) {
                if (ln) {  // learned: the features come with the leaves
                    if ((rc = learned_buffers(sc, nq, &X, &seeds))) return rc;
                    feats_done = true;
                }
            }
            // the self-query ordering's leaves serve the stack walk's prefill;
            // the per-query walk (k_knn) needs the descent's leaves
            if ((rc = order_queries<T, D>(t, tv, q, nq, &order, &leaf_of, sc, s, ln, feats_done ? X : nullptr,
                                          kind == CKD_KNN_STACK)))
                return rc;
        }
        if (ln) {
            // learned seeds: the features (unless the ordering pass made
            // them), then the network
            int rc;
            if (!feats_done) {
                if ((rc = learned_buffers(sc, nq, &X, &seeds))) return rc;
                if ((rc = ckd_mlp_features(t, q, nq, ln->G, ln->table, leaf_of, X, s))) return rc;
            }
            if ((rc = ckd_mlp_forward(ln->params, X, nq, ln->margin, 1, seeds, s))) return rc;
            seed = seeds;
        }
        if (k <= 8) return knn_launch_k<T, D, PER, 8>(t, tv, kind, sd, qpw, q, nq, k, sorted, seed, order, leaf_of, dist, idx, stats, sc, s);
        if (k <= 16) return knn_launch_k<T, D, PER, 16>(t, tv, kind, sd, qpw, q, nq, k, sorted, seed, order, leaf_of, dist, idx, stats, sc, s);
        if (k <= 32) return knn_launch_k<T, D, PER, 32>(t, tv, kind, sd, qpw, q, nq, k, sorted, seed, order, leaf_of, dist, idx, stats, sc, s);
        return knn_launch_k<T, D, PER, 64>(t, tv, kind, sd, qpw, q, nq, k, sorted, seed, order, leaf_of, dist, idx, stats, sc, s);
    }
    if (ln) {
        set_error("learned seeds need k <= 64");
        return CKD_ERR_ARG;
    }
    k_knn_heap<T, D, PER><<<g, QT, 0, s>>>(tv, q, nq, k, dist, idx);
    CKD_HIP(hipGetLastError());
    if (sorted) return segmented_sort<T>(dist, idx, nullptr, nq, k, s);
    return CKD_OK;
}

template <typename T, bool PER>
int sph_launch(const ckd_tree *t, const T *q, int64_t nq, int k, const T *mass, T *h, T *rho, Scratch &sc,
               hipStream_t s) {
    const TreeView<T> tv = view<T>(t);
    if (!stack_ok<T, 3>(t, 64) || t->n >= INT32_MAX) {
        set_error("the SPH step needs a 3-D tree of depth <= 64 and fewer than 2^31 points");
        return CKD_ERR_LIMIT;
    }
    const int32_t *order = nullptr, *leaf_of = nullptr;
    if (nq >= ORDER_MIN_QUERIES && nq < INT32_MAX && t->nnodes > 1) {
        const int rc = order_queries<T, 3>(t, tv, q, nq, &order, &leaf_of, sc, s);
        if (rc) return rc;
    }
    const int2 root = make_int2(0, t->nnodes > 1 ? -1 : (int32_t)t->n);
    const unsigned g = (unsigned)ceil_div(nq, QT);
    LaunchTimer tm{t};
    int rc = tm.start(s);
    if (rc) return rc;
#define CKD_SPH(KM)                                                                                          \
    do {                                                                                                    \
        if (mass) {                                                                                         \
            if (t->depth <= 32) k_sph_s<T, PER, KM, 32, false><<<g, QT, 0, s>>>(tv, root, q, nq, k, order, leaf_of, mass, h, rho); \
            else k_sph_s<T, PER, KM, 64, false><<<g, QT, 0, s>>>(tv, root, q, nq, k, order, leaf_of, mass, h, rho);          \
        } else {  /* unit masses: the lists carry no rows */                                                \
            if (t->depth <= 32) k_sph_s<T, PER, KM, 32, true><<<g, QT, 0, s>>>(tv, root, q, nq, k, order, leaf_of, mass, h, rho); \
            else k_sph_s<T, PER, KM, 64, true><<<g, QT, 0, s>>>(tv, root, q, nq, k, order, leaf_of, mass, h, rho);           \
        }                                                                                                   \
    } while (0)
    if (k <= 8) CKD_SPH(8);
    else if (k <= 16) CKD_SPH(16);
    else if (k <= 32) CKD_SPH(32);
    else CKD_SPH(64);
#undef CKD_SPH
    CKD_HIP(hipGetLastError());
    return tm.stop(s);
}

// CKD_BALL_KERNEL=lane|stack overrides the automatic choice of the ball
// kernel (the walk per wave where the wide records allow it); read at every
// call.
int ball_kernel_env() {
    const char *e = getenv("CKD_BALL_KERNEL");
    return !e ? -1 : (e[0] == 'l' ? CKD_KNN_LANE : e[0] == 's' ? CKD_KNN_STACK : -1);
}

template <typename T, int D>
int ball_launch(const ckd_tree *t, const T *q, int64_t nq, T r2, int64_t *counts, const int64_t *off, T *dist,
                int64_t *idx, Scratch &sc, hipStream_t s) {
    const TreeView<T> tv = view<T>(t);
    const unsigned g = (unsigned)ceil_div(nq, QT);
    const bool fill = off != nullptr;
    int kind = ball_kernel_env();
    if (kind < 0) kind = CKD_KNN_STACK;
    if constexpr (D <= 3) {
        if (kind == CKD_KNN_STACK && t->wide && t->depth <= 64 && t->n < INT32_MAX) {
            const int32_t *order = nullptr;
            if (nq >= ORDER_MIN_QUERIES && nq < INT32_MAX && t->nnodes > 1) {
                const int rc = order_queries<T, D>(t, tv, q, nq, &order, nullptr, sc, s);
                if (rc) return rc;
            }
            const int2 root = make_int2(0, t->nnodes > 1 ? -1 : (int32_t)t->n);
            const int2 rows = make_int2(0, (int32_t)t->n);
            if (t->depth <= 32) {
                if (t->periodic) {
                    if (fill) k_ball_s<T, D, true, true, 32><<<g, QT, 0, s>>>(tv, root, rows, q, nq, order, r2, counts, off, dist, idx);
                    else k_ball_s<T, D, true, false, 32><<<g, QT, 0, s>>>(tv, root, rows, q, nq, order, r2, counts, off, dist, idx);
                } else {
                    if (fill) k_ball_s<T, D, false, true, 32><<<g, QT, 0, s>>>(tv, root, rows, q, nq, order, r2, counts, off, dist, idx);
                    else k_ball_s<T, D, false, false, 32><<<g, QT, 0, s>>>(tv, root, rows, q, nq, order, r2, counts, off, dist, idx);
                }
            } else {
                if (t->periodic) {
                    if (fill) k_ball_s<T, D, true, true, 64><<<g, QT, 0, s>>>(tv, root, rows, q, nq, order, r2, counts, off, dist, idx);
                    else k_ball_s<T, D, true, false, 64><<<g, QT, 0, s>>>(tv, root, rows, q, nq, order, r2, counts, off, dist, idx);
                } else {
                    if (fill) k_ball_s<T, D, false, true, 64><<<g, QT, 0, s>>>(tv, root, rows, q, nq, order, r2, counts, off, dist, idx);
                    else k_ball_s<T, D, false, false, 64><<<g, QT, 0, s>>>(tv, root, rows, q, nq, order, r2, counts, off, dist, idx);
                }
            }
            if (fill) k_sqrt_rows<T><<<SQRT_BLOCKS, 256, 0, s>>>(dist, off + nq, INT64_MAX);
            CKD_HIP(hipGetLastError());
            return CKD_OK;
        }
    }
    if (t->periodic) {
        if (fill) k_ball<T, D, true, true><<<g, QT, 0, s>>>(tv, q, nq, r2, counts, off, dist, idx);
        else k_ball<T, D, true, false><<<g, QT, 0, s>>>(tv, q, nq, r2, counts, off, dist, idx);
    } else {
        if (fill) k_ball<T, D, false, true><<<g, QT, 0, s>>>(tv, q, nq, r2, counts, off, dist, idx);
        else k_ball<T, D, false, false><<<g, QT, 0, s>>>(tv, q, nq, r2, counts, off, dist, idx);
    }
    if (fill) k_sqrt_rows<T><<<SQRT_BLOCKS, 256, 0, s>>>(dist, off + nq, INT64_MAX);
    CKD_HIP(hipGetLastError());
    return CKD_OK;
}

// Whether the one-walk ball search runs on this tree and batch (the stack
// walk's conditions); otherwise ckd_ball_search counts and fills in two.
inline bool ball_one_walk_ok(const ckd_tree *t, int64_t nq) {
    return t->ndim <= 3 && t->wide && t->depth <= 64 && t->n < INT32_MAX && nq < INT32_MAX;
}

// One-walk ball search: order the queries, walk once writing counts and
// staging records, scan the counts into `off`, move the records to their
// places.  The staging area holds cap records plus one chunk per wave (each
// wave's last chunk is partly filled), so it overflows only when the total
// exceeds cap.  *total and *overflow [host] are read back at the end (one
// synchronisation of the stream).
template <typename T, int D>
int ball_search_launch(const ckd_tree *t, const T *q, int64_t nq, T r2, int64_t *off, int64_t cap, T *dist,
                       int64_t *idx, int64_t *total, int32_t *overflow, Scratch &sc, hipStream_t s) {
    if constexpr (D > 3) {
        return CKD_ERR_ARG;
    } else {
        const TreeView<T> tv = view<T>(t);
        const int32_t *order = nullptr;
        if (nq >= ORDER_MIN_QUERIES && t->nnodes > 1) {
            const int rc = order_queries<T, D>(t, tv, q, nq, &order, nullptr, sc, s);
            if (rc) return rc;
        }
        const int64_t nwaves = ceil_div(nq, WAVE);
        // slots are 32-bit: a larger capacity is staged up to the limit (the
        // rows beyond it are then filled by the second walk)
        const int64_t max_chunks = (INT32_MAX / STAGE_CHUNK) - 1;
        int64_t nchunks = ceil_div(cap > 0 ? cap : 0, STAGE_CHUNK) + nwaves + 1;
        if (nchunks > max_chunks) nchunks = max_chunks;
        // scratch: counts (nq), chunk fills (nchunks), counter + flag, scan, records
        const size_t b_cnt = ((size_t)nq * 8 + 255) & ~(size_t)255;
        const size_t b_fill = ((size_t)nchunks * 4 + 255) & ~(size_t)255;
        const size_t b_scan = (scan_tmp_bytes(nq) + 255) & ~(size_t)255;
        const size_t b_rec = (size_t)nchunks * STAGE_CHUNK * sizeof(int4);
        void *ws = nullptr;
        int rc = sc.take(b_cnt + b_fill + 256 + b_scan + b_rec, &ws);
        if (rc) return rc;
        char *p = (char *)ws;
        int64_t *counts = (int64_t *)p;
        StageOut so;
        so.chunk_fill = (int32_t *)(p + b_cnt);
        so.next_chunk = (uint32_t *)(p + b_cnt + b_fill);
        so.overflow = (int32_t *)(p + b_cnt + b_fill + 8);
        void *scr = p + b_cnt + b_fill + 256;
        so.rec = (int4 *)(p + b_cnt + b_fill + 256 + b_scan);
        so.nchunks = (int32_t)nchunks;
        k_fill_i64<<<1, 64, 0, s>>>((int64_t *)so.next_chunk, 2, 0);  // the counter and the flag: no memset node
        const int2 root = make_int2(0, t->nnodes > 1 ? -1 : (int32_t)t->n);
        const int2 rows = make_int2(0, (int32_t)t->n);
        const unsigned g = (unsigned)ceil_div(nq, QT);
        if (t->periodic) {
            if (t->depth <= 32) k_ball_stage<T, D, true, 32><<<g, QT, 0, s>>>(tv, root, rows, q, nq, order, r2, counts, so);
            else k_ball_stage<T, D, true, 64><<<g, QT, 0, s>>>(tv, root, rows, q, nq, order, r2, counts, so);
        } else {
            if (t->depth <= 32) k_ball_stage<T, D, false, 32><<<g, QT, 0, s>>>(tv, root, rows, q, nq, order, r2, counts, so);
            else k_ball_stage<T, D, false, 64><<<g, QT, 0, s>>>(tv, root, rows, q, nq, order, r2, counts, so);
        }
        CKD_HIP(hipGetLastError());
        if ((rc = scan_i64(counts, off, nq, true, scr, s))) return rc;
        const unsigned gs = nchunks < SCATTER_BLOCKS ? (unsigned)nchunks : SCATTER_BLOCKS;
        if (t->periodic) k_ball_scatter<T, D, true><<<gs, 256, 0, s>>>(tv, q, so, off, nq, cap, dist, idx);
        else k_ball_scatter<T, D, false><<<gs, 256, 0, s>>>(tv, q, so, off, nq, cap, dist, idx);
        CKD_HIP(hipGetLastError());
        CKD_HIP(hipMemcpyAsync(total, off + nq, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        CKD_HIP(hipMemcpyAsync(overflow, so.overflow, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        CKD_HIP(hipStreamSynchronize(s));
        return CKD_OK;
    }
}

template <typename T, int D>
int box_launch(const ckd_tree *t, const T *b, int64_t nb, int64_t *counts, const int64_t *off, int64_t *idx,
               int wrap, hipStream_t s) {
    const TreeView<T> tv = view<T>(t);
    const unsigned g = (unsigned)ceil_div(nb, QT);
    if (wrap) {
        if (off) k_box_wrap<T, D, true><<<g, QT, 0, s>>>(tv, b, nb, counts, off, idx);
        else k_box_wrap<T, D, false><<<g, QT, 0, s>>>(tv, b, nb, counts, off, idx);
    } else if (off) k_box<T, D, true><<<g, QT, 0, s>>>(tv, b, nb, counts, off, idx);
    else k_box<T, D, false><<<g, QT, 0, s>>>(tv, b, nb, counts, off, idx);
    CKD_HIP(hipGetLastError());
    return CKD_OK;
}

#define CKD_DIM_SWITCH(ndim, CALL)            \
    switch (ndim) {                           \
        case 1: { constexpr int D_ = 1; return CALL; } \
        case 2: { constexpr int D_ = 2; return CALL; } \
        case 3: { constexpr int D_ = 3; return CALL; } \
        case 4: { constexpr int D_ = 4; return CALL; } \
        case 5: { constexpr int D_ = 5; return CALL; } \
        case 6: { constexpr int D_ = 6; return CALL; } \
        case 7: { constexpr int D_ = 7; return CALL; } \
        case 8: { constexpr int D_ = 8; return CALL; } \
    }

// kNN of one precision and periodicity (each pair is its own translation
// unit, ckd_knn_<prec>_<per|open>.hip, so they compile in parallel);
// params != nullptr asks for learned seeds.
template <typename T, bool PER>
int knn_per(const ckd_tree *t, const void *q, int64_t nq, int k, int sorted, const float *seed, const float *params,
            float margin, int G, const uint16_t *table, void *dist, int64_t *idx, int64_t *stats, Scratch &sc,
            hipStream_t s) {
    const Learned ln{params, margin, G, table};
    const Learned *lp = params ? &ln : nullptr;
    CKD_DIM_SWITCH(t->ndim, (knn_launch<T, D_, PER>(t, (const T *)q, nq, k, sorted, seed, lp, (T *)dist, idx, stats, sc, s)))
    return CKD_ERR_ARG;
}

template <typename T>
int ball_t(const ckd_tree *t, const void *q, int64_t nq, double r, int64_t *counts, const int64_t *off, void *dist,
           int64_t *idx, Scratch &sc, hipStream_t s) {
    const T rr = (T)r;   // radius is a real(prec) dummy (cosmokdtree.f90:1019)
    const T r2 = rr * rr;
    CKD_DIM_SWITCH(t->ndim, (ball_launch<T, D_>(t, (const T *)q, nq, r2, counts, off, (T *)dist, idx, sc, s)))
    return CKD_ERR_ARG;
}

template <typename T>
int ball_search_t(const ckd_tree *t, const void *q, int64_t nq, double r, int64_t *off, int64_t cap, void *dist,
                  int64_t *idx, int64_t *total, int32_t *overflow, Scratch &sc, hipStream_t s) {
    const T rr = (T)r;  // radius is a real(prec) dummy (cosmokdtree.f90:1019)
    const T r2 = rr * rr;
    CKD_DIM_SWITCH(t->ndim, (ball_search_launch<T, D_>(t, (const T *)q, nq, r2, off, cap, (T *)dist, idx, total,
                                                       overflow, sc, s)))
    return CKD_ERR_ARG;
}

template <typename T>
int box_t(const ckd_tree *t, const void *b, int64_t nb, int64_t *counts, const int64_t *off, int64_t *idx, int wrap,
          hipStream_t s) {
    CKD_DIM_SWITCH(t->ndim, (box_launch<T, D_>(t, (const T *)b, nb, counts, off, idx, wrap, s)))
    return CKD_ERR_ARG;
}

}  // namespace

}  // namespace ckd
