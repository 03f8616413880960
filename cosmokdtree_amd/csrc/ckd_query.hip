// ckd_query.hip -- kNN, ball and box queries over a built tree.
//
// One thread per query.  Traversals are stack-free: a node's parent
// pointer replaces the recursion stack of the reference, so any tree depth
// (sliding-midpoint trees can be very deep on duplicates) is handled
// without a stack bound.
//
// kNN keeps the k best candidates in registers as an ascending list
// (distance, leaf row) of static length KMAX >= k: the k live entries are
// the top k slots, the slots below are -inf and inert, so the pruning bound
// is always slot KMAX-1.  A candidate is accepted only when strictly closer
// than that bound, as in cosmokdtree.f90:658,808; the list starts filled
// with FLT_MAX (HUGE(0.0), cosmokdtree.f90:517) or with the squared seed
// radius.  The visit order is depth-first, nearer child first (the order
// of knn_search_recursive_hyperp, cosmokdtree.f90:820-846), with every
// subtree pruned by its cell's minimum distance (bbox_mindist_kd,
// cosmokdtree.f90:1525-1546; minimum image when periodic).  Pruning never
// drops a candidate that could be accepted, so the accepted set equals the
// reference's up to the choice among points tied at the k-th distance.
#include "ckd_internal.h"

namespace ckd {

namespace {

constexpr int QT = 256;  // threads per block
constexpr int64_t ORDER_MIN_QUERIES = 4096;  // below this, queries run in input order

template <typename T, int D>
__device__ __forceinline__ void load_pt(const T *pts, int64_t p, T (&x)[D]) {
    constexpr int PS = pstride(D);
    const T *src = pts + p * PS;
    if constexpr (sizeof(T) == 4 && PS == 4) {
        float4 v = *(const float4 *)src;
        x[0] = v.x;
        if (D > 1) x[1] = v.y;
        if (D > 2) x[2] = v.z;
        if (D > 3) x[3] = v.w;
    } else {
#pragma unroll
        for (int d = 0; d < D; d++) x[d] = src[d];
    }
}

template <typename T, int D>
__device__ __forceinline__ void load_box(const T *bbox, int32_t node, T (&b)[2 * D]) {
    const T *src = bbox + (int64_t)node * 2 * D;
#pragma unroll
    for (int d = 0; d < 2 * D; d++) b[d] = src[d];
}

template <typename T, int D>
__device__ __forceinline__ T node_mindist(const TreeView<T> &tv, int32_t node, const T (&q)[D], bool per) {
    T b[2 * D];
    load_box<T, D>(tv.bbox, node, b);
    return per ? mindist2<T, D, true>(b, q, tv.L) : mindist2<T, D, false>(b, q, tv.L);
}

template <typename T, int KMAX>
__device__ __forceinline__ void list_insert(T (&bd)[KMAX], int32_t (&bs)[KMAX], T d, int32_t s) {
#pragma unroll
    for (int j = KMAX - 1; j > 0; --j) {
        const bool sh = d < bd[j - 1];
        const bool here = !sh && d < bd[j];
        bd[j] = sh ? bd[j - 1] : (here ? d : bd[j]);
        bs[j] = sh ? bs[j - 1] : (here ? s : bs[j]);
    }
    const bool here = d < bd[0];
    bd[0] = here ? d : bd[0];
    bs[0] = here ? s : bs[0];
}

// The leaf a point falls in: the nearer child at every level (the first
// descent of knn_search_recursive_hyperp, cosmokdtree.f90:820-836).
template <typename T, int D>
__device__ __forceinline__ int32_t find_leaf(const TreeView<T> &tv, const T (&q)[D]) {
    int32_t node = 0;
    for (;;) {
        const NodeRec<T> nr = tv.nodes[node];
        if (nr.axis < 0) return node;
        node = nr.child + (q[nr.axis] < nr.split ? 0 : 1);
    }
}

// Query ordering: each query's leaf, and its rank among the queries of
// that leaf; a counting sort then lists the queries leaf by leaf so a wave
// works on neighbouring queries.
template <typename T, int D>
__global__ __launch_bounds__(QT) void k_locate(TreeView<T> tv, const T *queries, int64_t nq, int32_t *leaf_of,
                                               int32_t *rank, int64_t *leaf_cnt) {
    const int64_t qi = (int64_t)blockIdx.x * QT + threadIdx.x;
    if (qi >= nq) return;
    T q[D];
#pragma unroll
    for (int d = 0; d < D; d++) q[d] = queries[qi * D + d];
    const int32_t leaf = find_leaf<T, D>(tv, q);
    leaf_of[qi] = leaf;
    rank[qi] = (int32_t)atomicAdd((unsigned long long *)&leaf_cnt[leaf], 1ull);
}

__global__ void k_order(const int32_t *leaf_of, const int32_t *rank, const int64_t *leaf_off, int64_t nq,
                        int32_t *order) {
    const int64_t qi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    order[leaf_off[leaf_of[qi]] + rank[qi]] = (int32_t)qi;
}

__global__ void k_fill_i64(int64_t *p, int64_t n, int64_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename T, int D, bool PER, int KMAX>
__device__ __forceinline__ void knn_scan_leaf(const TreeView<T> &tv, int32_t leaf, const T (&q)[D], T (&bd)[KMAX],
                                              int32_t (&bs)[KMAX], int64_t &tested) {
    const Range r = tv.range[leaf];
    for (int64_t p = r.start; p < r.start + r.count; p++) {
        T x[D];
        load_pt<T, D>(tv.pts, p, x);
        const T d = dist2<T, D, PER>(x, q, tv.L);
        if (d < bd[KMAX - 1]) list_insert<T, KMAX>(bd, bs, d, (int32_t)p);
    }
    tested += r.count;
}

template <typename T, int D, bool PER, int KMAX>
__global__ __launch_bounds__(QT) void k_knn(TreeView<T> tv, const T *queries, int64_t nq, int k, int sorted,
                                            const float *seed, const int32_t *order, const int32_t *leaf_of,
                                            T *dout, int64_t *iout, int64_t *stats) {
    const int64_t t = (int64_t)blockIdx.x * QT + threadIdx.x;
    int64_t visited = 0, tested = 0, reruns = 0;
    if (t < nq) {
    const int64_t qi = order ? (int64_t)order[t] : t;
    T q[D];
#pragma unroll
    for (int d = 0; d < D; d++) q[d] = queries[qi * D + d];
    T bd[KMAX];
    int32_t bs[KMAX];
    // the leaf the query falls in: nearer child at every level
    int32_t leaf = leaf_of ? leaf_of[qi] : find_leaf<T, D>(tv, q);
    T init = T(FLT_MAX);
    if (seed) {
        const T sr = (T)seed[qi];
        const T s2 = sr * sr;
        if (s2 < init) init = s2;
    }
    for (int pass = 0; pass < 2; pass++) {
#pragma unroll
        for (int j = 0; j < KMAX; j++) {
            bd[j] = j < KMAX - k ? T(-INFINITY) : init;
            bs[j] = -1;
        }
        knn_scan_leaf<T, D, PER, KMAX>(tv, leaf, q, bd, bs, tested);
        visited++;
        int32_t from = leaf, cur = tv.nodes[leaf].parent;
        while (cur >= 0) {
            const NodeRec<T> nr = tv.nodes[cur];
            const int32_t nearc = nr.child + (q[nr.axis] < nr.split ? 0 : 1);
            bool moved = false;
            if (from == nearc) {
                const int32_t farc = nearc == nr.child ? nr.child + 1 : nr.child;
                if (node_mindist<T, D>(tv, farc, q, PER) < bd[KMAX - 1]) {
                    int32_t n2 = farc;
                    for (;;) {
                        visited++;
                        const NodeRec<T> r2 = tv.nodes[n2];
                        if (r2.axis < 0) {
                            knn_scan_leaf<T, D, PER, KMAX>(tv, n2, q, bd, bs, tested);
                            from = n2;
                            cur = r2.parent;
                            break;
                        }
                        const int32_t nc2 = r2.child + (q[r2.axis] < r2.split ? 0 : 1);
                        if (node_mindist<T, D>(tv, nc2, q, PER) < bd[KMAX - 1]) {
                            n2 = nc2;
                            continue;
                        }
                        from = nc2;  // nearer child pruned: continue as if back from it
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
        if (!seed || bs[KMAX - 1] >= 0 || init == T(FLT_MAX)) break;
        init = T(FLT_MAX);  // fewer than k points inside the seed radius
        reruns++;
    }
    // write (distance, original index) in ascending order
    T *drow = dout + qi * k;
    int64_t *irow = iout + qi * k;
    bool fix = false;
    T pd = T(-1);
    int64_t pi = 0;
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
        if (j >= KMAX - k) {
            const int32_t s = bs[j];
            const int64_t oi = s >= 0 ? load_idx(tv.idx, tv.idx64, s) : (int64_t)CKD_EMPTY_SLOT;
            const T dv = sqrt_rn<T>(bd[j]);
            fix |= (dv == pd && oi < pi);
            pd = dv;
            pi = oi;
            drow[j - (KMAX - k)] = dv;
            irow[j - (KMAX - k)] = oi;
        }
    }
    if (sorted && fix) {  // order equal distances by index (insertion sort, rare)
        for (int a = 1; a < k; a++) {
            const T dv = drow[a];
            const int64_t iv = irow[a];
            int b = a - 1;
            while (b >= 0 && drow[b] == dv && irow[b] > iv) {
                drow[b + 1] = drow[b];
                irow[b + 1] = irow[b];
                b--;
            }
            drow[b + 1] = dv;
            irow[b + 1] = iv;
        }
    }
    }
    if (stats) {  // one atomic per wave and counter
        int64_t nqw = t < nq ? 1 : 0;
#pragma unroll
        for (int m = WAVE / 2; m >= 1; m >>= 1) {
            visited += __shfl_xor(visited, m, WAVE);
            tested += __shfl_xor(tested, m, WAVE);
            reruns += __shfl_xor(reruns, m, WAVE);
            nqw += __shfl_xor(nqw, m, WAVE);
        }
        if ((threadIdx.x & (WAVE - 1)) == 0 && nqw > 0) {
            atomicAdd((unsigned long long *)&stats[0], (unsigned long long)visited);
            atomicAdd((unsigned long long *)&stats[1], (unsigned long long)tested);
            atomicAdd((unsigned long long *)&stats[2], (unsigned long long)reruns);
            atomicAdd((unsigned long long *)&stats[3], (unsigned long long)nqw);
        }
    }
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
    int32_t leaf = 0;
    for (;;) {
        const NodeRec<T> nr = tv.nodes[leaf];
        if (nr.axis < 0) break;
        leaf = nr.child + (q[nr.axis] < nr.split ? 0 : 1);
    }
    scan(leaf);
    int32_t from = leaf, cur = tv.nodes[leaf].parent;
    while (cur >= 0) {
        const NodeRec<T> nr = tv.nodes[cur];
        const int32_t nearc = nr.child + (q[nr.axis] < nr.split ? 0 : 1);
        bool moved = false;
        if (from == nearc) {
            const int32_t farc = nearc == nr.child ? nr.child + 1 : nr.child;
            if (node_mindist<T, D>(tv, farc, q, PER) < hd[0]) {
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
                    if (node_mindist<T, D>(tv, nc2, q, PER) < hd[0]) {
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

// ---------------------------------------------------------------- ball
// Depth-first, left child first (ball_search_recursive_bbox,
// cosmokdtree.f90:1076-1139): a cell entirely farther than r is skipped, a
// cell entirely within r contributes its whole point range
// (ball_reduction_recursive, :1332-1366), leaves test d^2 <= r^2 inclusively
// (:1126).  The whole-cell bound uses the same rounded per-axis differences
// as the point distance, so it includes exactly the points the per-point
// test would.
template <typename T, int D, bool PER, bool FILL>
__global__ __launch_bounds__(QT) void k_ball(TreeView<T> tv, const T *queries, int64_t nq, T r2, int64_t *counts,
                                             const int64_t *offsets, T *dout, int64_t *iout) {
    const int64_t qi = (int64_t)blockIdx.x * QT + threadIdx.x;
    if (qi >= nq) return;
    T q[D];
#pragma unroll
    for (int d = 0; d < D; d++) q[d] = queries[qi * D + d];
    int64_t cnt = 0;
    int64_t w = FILL ? offsets[qi] : 0;
    auto emit = [&](int64_t p, T d) {
        if (FILL) {
            dout[w] = sqrt_rn<T>(d);
            iout[w] = load_idx(tv.idx, tv.idx64, p);
            w++;
        }
    };
    int32_t cur = 0, from = -1;  // from == -1: arriving from above
    while (cur >= 0) {
        const NodeRec<T> nr = tv.nodes[cur];
        if (from == -1) {  // first visit of cur
            T b[2 * D];
            load_box<T, D>(tv.bbox, cur, b);
            const T mn = mindist2<T, D, PER>(b, q, tv.L);
            bool descend = false;
            if (!(mn > r2)) {
                if (maxdist2<T, D>(b, q) <= r2) {
                    const Range r = tv.range[cur];
                    if (FILL)
                        for (int64_t p = r.start; p < r.start + r.count; p++) {
                            T x[D];
                            load_pt<T, D>(tv.pts, p, x);
                            emit(p, dist2<T, D, PER>(x, q, tv.L));
                        }
                    cnt += r.count;
                } else if (nr.axis < 0) {
                    const Range r = tv.range[cur];
                    for (int64_t p = r.start; p < r.start + r.count; p++) {
                        T x[D];
                        load_pt<T, D>(tv.pts, p, x);
                        const T d = dist2<T, D, PER>(x, q, tv.L);
                        if (d <= r2) {
                            emit(p, d);
                            cnt++;
                        }
                    }
                } else {
                    descend = true;
                }
            }
            if (descend) {
                cur = nr.child;  // left first
                continue;
            }
            from = cur;
            cur = nr.parent;
            continue;
        }
        // back from child `from`
        if (from == nr.child) {
            cur = nr.child + 1;
            from = -1;
            continue;
        }
        from = cur;
        cur = nr.parent;
    }
    if (!FILL) counts[qi] = cnt;
}

// ---------------------------------------------------------------- box
// box_search_recursive (cosmokdtree.f90:1258-1328): visit the left child iff
// box_lo[axis] < split and the right child iff box_hi[axis] > split; leaves
// test box_lo <= x <= box_hi on every axis.  A cell strictly inside the box
// on every axis is taken whole: every split below it passes both visit tests
// and every point passes the leaf test.
template <typename T, int D, bool FILL>
__global__ __launch_bounds__(QT) void k_box(TreeView<T> tv, const T *boxes, int64_t nb, int64_t *counts,
                                            const int64_t *offsets, int64_t *iout) {
    const int64_t bi = (int64_t)blockIdx.x * QT + threadIdx.x;
    if (bi >= nb) return;
    T bx[2 * D];
#pragma unroll
    for (int d = 0; d < 2 * D; d++) bx[d] = boxes[bi * 2 * D + d];
    int64_t cnt = 0;
    int64_t w = FILL ? offsets[bi] : 0;
    int32_t cur = 0, from = -1;
    while (cur >= 0) {
        const NodeRec<T> nr = tv.nodes[cur];
        if (from == -1) {
            T b[2 * D];
            load_box<T, D>(tv.bbox, cur, b);
            bool inside = true;
#pragma unroll
            for (int d = 0; d < D; d++) inside = inside && bx[2 * d] < b[d] && b[D + d] < bx[2 * d + 1];
            if (inside) {
                const Range r = tv.range[cur];
                if (FILL)
                    for (int64_t p = r.start; p < r.start + r.count; p++) iout[w++] = load_idx(tv.idx, tv.idx64, p);
                cnt += r.count;
            } else if (nr.axis < 0) {
                const Range r = tv.range[cur];
                for (int64_t p = r.start; p < r.start + r.count; p++) {
                    T x[D];
                    load_pt<T, D>(tv.pts, p, x);
                    bool in = true;
#pragma unroll
                    for (int d = 0; d < D; d++) in = in && !(x[d] < bx[2 * d] || x[d] > bx[2 * d + 1]);
                    if (in) {
                        if (FILL) iout[w++] = load_idx(tv.idx, tv.idx64, p);
                        cnt++;
                    }
                }
            } else {
                if (bx[2 * nr.axis] < nr.split) {
                    cur = nr.child;
                    continue;
                }
                if (bx[2 * nr.axis + 1] > nr.split) {
                    cur = nr.child + 1;
                    continue;
                }
            }
            from = cur;
            cur = nr.parent;
            continue;
        }
        if (from == nr.child && bx[2 * nr.axis + 1] > nr.split) {
            cur = nr.child + 1;
            from = -1;
            continue;
        }
        from = cur;
        cur = nr.parent;
    }
    if (!FILL) counts[bi] = cnt;
}

// ------------------------------------------------------------ dispatch

template <typename T, int D, bool PER>
int knn_launch(const ckd_tree *t, const T *q, int64_t nq, int k, int sorted, const float *seed, T *dist,
               int64_t *idx, int64_t *stats, hipStream_t s) {
    const TreeView<T> tv = view<T>(t);
    const unsigned g = (unsigned)ceil_div(nq, QT);
    if (k <= 64 && t->n < INT32_MAX) {
        const int32_t *order = nullptr, *leaf_of = nullptr;
        if (nq >= ORDER_MIN_QUERIES && nq < INT32_MAX && t->nnodes > 1) {
            // workspace: leaf_of, rank, order (int32 x nq), leaf counts and
            // offsets (int64 x nnodes+1), scan scratch
            const size_t a = 3 * sizeof(int32_t) * (size_t)nq, b = sizeof(int64_t) * (size_t)(t->nnodes + 1);
            void *ws = nullptr;
            int rc = workspace(const_cast<ckd_tree *>(t), a + 2 * b + scan_tmp_bytes(t->nnodes) + 256, &ws, s);
            if (rc) return rc;
            int32_t *lo = (int32_t *)ws, *rk = lo + nq, *ord = rk + nq;
            int64_t *cnt = (int64_t *)((char *)ws + ((a + 255) & ~(size_t)255));
            int64_t *off = cnt + t->nnodes + 1;
            void *scr = off + t->nnodes + 1;
            k_fill_i64<<<(unsigned)ceil_div(t->nnodes, 256), 256, 0, s>>>(cnt, t->nnodes, 0);
            k_locate<T, D><<<g, QT, 0, s>>>(tv, q, nq, lo, rk, cnt);
            if ((rc = scan_i64(cnt, off, t->nnodes, false, scr, s))) return rc;
            k_order<<<(unsigned)ceil_div(nq, 256), 256, 0, s>>>(lo, rk, off, nq, ord);
            CKD_HIP(hipGetLastError());
            order = ord;
            leaf_of = lo;
        }
        if (k <= 8) k_knn<T, D, PER, 8><<<g, QT, 0, s>>>(tv, q, nq, k, sorted, seed, order, leaf_of, dist, idx, stats);
        else if (k <= 16) k_knn<T, D, PER, 16><<<g, QT, 0, s>>>(tv, q, nq, k, sorted, seed, order, leaf_of, dist, idx, stats);
        else if (k <= 32) k_knn<T, D, PER, 32><<<g, QT, 0, s>>>(tv, q, nq, k, sorted, seed, order, leaf_of, dist, idx, stats);
        else k_knn<T, D, PER, 64><<<g, QT, 0, s>>>(tv, q, nq, k, sorted, seed, order, leaf_of, dist, idx, stats);
        CKD_HIP(hipGetLastError());
        return CKD_OK;
    }
    k_knn_heap<T, D, PER><<<g, QT, 0, s>>>(tv, q, nq, k, dist, idx);
    CKD_HIP(hipGetLastError());
    if (sorted) return segmented_sort<T>(dist, idx, nullptr, nq, k, s);
    return CKD_OK;
}

template <typename T, int D>
int ball_launch(const ckd_tree *t, const T *q, int64_t nq, T r2, int64_t *counts, const int64_t *off, T *dist,
                int64_t *idx, hipStream_t s) {
    const TreeView<T> tv = view<T>(t);
    const unsigned g = (unsigned)ceil_div(nq, QT);
    const bool fill = off != nullptr;
    if (t->periodic) {
        if (fill) k_ball<T, D, true, true><<<g, QT, 0, s>>>(tv, q, nq, r2, counts, off, dist, idx);
        else k_ball<T, D, true, false><<<g, QT, 0, s>>>(tv, q, nq, r2, counts, off, dist, idx);
    } else {
        if (fill) k_ball<T, D, false, true><<<g, QT, 0, s>>>(tv, q, nq, r2, counts, off, dist, idx);
        else k_ball<T, D, false, false><<<g, QT, 0, s>>>(tv, q, nq, r2, counts, off, dist, idx);
    }
    CKD_HIP(hipGetLastError());
    return CKD_OK;
}

template <typename T, int D>
int box_launch(const ckd_tree *t, const T *b, int64_t nb, int64_t *counts, const int64_t *off, int64_t *idx,
               hipStream_t s) {
    const TreeView<T> tv = view<T>(t);
    const unsigned g = (unsigned)ceil_div(nb, QT);
    if (off) k_box<T, D, true><<<g, QT, 0, s>>>(tv, b, nb, counts, off, idx);
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

template <typename T>
int knn_t(const ckd_tree *t, const void *q, int64_t nq, int k, int sorted, const float *seed, void *dist,
          int64_t *idx, int64_t *stats, hipStream_t s) {
    if (t->periodic) {
        CKD_DIM_SWITCH(t->ndim, (knn_launch<T, D_, true>(t, (const T *)q, nq, k, sorted, seed, (T *)dist, idx, stats, s)))
    } else {
        CKD_DIM_SWITCH(t->ndim, (knn_launch<T, D_, false>(t, (const T *)q, nq, k, sorted, seed, (T *)dist, idx, stats, s)))
    }
    return CKD_ERR_ARG;
}

template <typename T>
int ball_t(const ckd_tree *t, const void *q, int64_t nq, double r, int64_t *counts, const int64_t *off, void *dist,
           int64_t *idx, hipStream_t s) {
    const T rr = (T)r;   // radius is a real(prec) dummy (cosmokdtree.f90:1019)
    const T r2 = rr * rr;
    CKD_DIM_SWITCH(t->ndim, (ball_launch<T, D_>(t, (const T *)q, nq, r2, counts, off, (T *)dist, idx, s)))
    return CKD_ERR_ARG;
}

template <typename T>
int box_t(const ckd_tree *t, const void *b, int64_t nb, int64_t *counts, const int64_t *off, int64_t *idx,
          hipStream_t s) {
    CKD_DIM_SWITCH(t->ndim, (box_launch<T, D_>(t, (const T *)b, nb, counts, off, idx, s)))
    return CKD_ERR_ARG;
}

}  // namespace

int query_knn(const ckd_tree *t, const void *q, int64_t nq, int k, int sorted, const float *seed, void *dist,
              int64_t *idx, int64_t *stats, hipStream_t s) {
    return t->prec == 4 ? knn_t<float>(t, q, nq, k, sorted, seed, dist, idx, stats, s)
                        : knn_t<double>(t, q, nq, k, sorted, seed, dist, idx, stats, s);
}

int query_ball(const ckd_tree *t, const void *q, int64_t nq, double r, int64_t *counts, const int64_t *off,
               void *dist, int64_t *idx, hipStream_t s) {
    return t->prec == 4 ? ball_t<float>(t, q, nq, r, counts, off, dist, idx, s)
                        : ball_t<double>(t, q, nq, r, counts, off, dist, idx, s);
}

int query_box(const ckd_tree *t, const void *b, int64_t nb, int64_t *counts, const int64_t *off, int64_t *idx,
              hipStream_t s) {
    return t->prec == 4 ? box_t<float>(t, b, nb, counts, off, idx, s) : box_t<double>(t, b, nb, counts, off, idx, s);
}

}  // namespace ckd
