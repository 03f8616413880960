// ckd_batch_impl.h -- batched queries over a forest of independent point
// sets in persistent kernels that drain a device work queue (the kernels and
// their launch; the plan and the C entry points are in ckd_batch.hip, each
// precision and periodicity is instantiated in its own ckd_batch_<prec>_
// <per|open>.hip so they compile in parallel).
//
// A plan lists work items (set, kind, first query, query count <= NT):
// kNN items run the stack walk from the set's root (whole waves, queries in
// leaf order within their set, 64 / G queries per wave; ckd_batch_config)
// or one walk per query;
// the kNN items (largest k first) run in k_batch, the ball-count items and
// then the ball-fill items in k_batch_ball (one walk per query, few
// registers).  Each launch starts one grid of NT-thread blocks that pull
// items from an atomic head counter until its queue is empty, so sets of
// very different sizes, k and radii share a launch.  Between the ball
// counts and fills a device scan turns the counts into CSR offsets.
// ckd_batch_run only launches kernels (no allocation, copy or host wait), so
// a caller can capture it in a hipGraph and replay it.
//
// Per set r: queries are rows [qoff[r], qoff[r+1]) of the query array, kNN
// rows of k[r] entries start at element knn_off[r] of the kNN outputs, and
// indices are relative to the set (0 = the set's first point).  Results are
// those of the single-tree calls on the set's own tree.
#pragma once
#include <algorithm>
#include <vector>

#include "ckd_traverse.h"

struct BatchItem {
    int32_t set;
    int32_t kind;  // 0 kNN, 1 ball count, 2 ball fill
    int64_t q0;
    int32_t nq;
    int32_t pad;
};

struct ckd_batch {
    const ckd_tree *tree = nullptr;
    int64_t nsets = 0, nq = 0, knn_total = 0, capacity = 0;
    int sorted = 1, nt = 256, grid = 1024, has_ball = 0, has_knn = 0;
    int knn_kernel = CKD_KNN_STACK, qpw = 64, stack = 0;  // set by ckd_batch_plan / ckd_batch_config
    int ball_kernel = CKD_KNN_LANE;  // ball items: one walk per query, or CKD_KNN_STACK (ckd_batch_config_ball)
    // items_a: kNN (k_batch); items_c: ball counts, items_b: ball fills (k_batch_ball)
    BatchItem *items_a = nullptr, *items_b = nullptr, *items_c = nullptr;
    int64_t n_a = 0, n_b = 0, n_c = 0;
    int32_t *kset = nullptr;
    double *rset = nullptr;
    int64_t *qoff = nullptr, *knn_off = nullptr, *counters = nullptr, *counts = nullptr;
    void *scan_tmp = nullptr;
    // query ordering of the stack kNN path (leaf order within each set)
    int32_t *leaf_of = nullptr, *rank = nullptr, *order = nullptr;
    int64_t *leaf_cnt = nullptr, *leaf_off = nullptr, *setpos = nullptr;
    void *order_tmp = nullptr;
};

namespace ckd {
namespace {

__global__ void k_batch_reset(int64_t *counters) {
    counters[0] = 0;  // phase A head
    counters[1] = 0;  // phase B head
    counters[2] = 0;  // ball overflow flag
    counters[3] = 0;  // ball-count head
}

// ---- query ordering for the stack kNN path: each query's set (binary
// search of qoff), its leaf in that set's tree and its rank there; then a
// counting sort by leaf.  Every set's leaves are contiguous in id order, so
// position qoff[r] + (leaf offset + rank - first offset of set r) lists set
// r's queries in leaf order inside its own range.
__device__ __forceinline__ int32_t set_of(const int64_t *qoff, int64_t nsets, int64_t qi) {
    int64_t lo = 0, hi = nsets;  // qoff[lo] <= qi < qoff[hi]
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) / 2;
        if (qoff[mid] <= qi) lo = mid;
        else hi = mid;
    }
    return (int32_t)lo;
}

template <typename T, int D>
__global__ void k_blocate(TreeView<T> tv, const T *queries, int64_t nq, const int64_t *qoff, int64_t nsets,
                          int32_t *leaf_of, int32_t *rank, int64_t *leaf_cnt) {
    const int64_t qi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    T q[D];
#pragma unroll
    for (int d = 0; d < D; d++) q[d] = queries[qi * D + d];
    const int32_t leaf = find_leaf<T, D>(tv, q, set_of(qoff, nsets, qi));
    leaf_of[qi] = leaf;
    rank[qi] = (int32_t)atomicAdd((unsigned long long *)&leaf_cnt[leaf], 1ull);
}

__global__ void k_bsetpos(const int32_t *leaf_of, const int64_t *leaf_off, const int64_t *qoff, int64_t nsets,
                          int64_t nq, int64_t *setpos) {
    const int64_t qi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    atomicMin((long long *)&setpos[set_of(qoff, nsets, qi)], (long long)leaf_off[leaf_of[qi]]);
}

__global__ void k_border(const int32_t *leaf_of, const int32_t *rank, const int64_t *leaf_off, const int64_t *qoff,
                         int64_t nsets, int64_t nq, const int64_t *setpos, int32_t *order) {
    const int64_t qi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    const int32_t r = set_of(qoff, nsets, qi);
    order[qoff[r] + leaf_off[leaf_of[qi]] + rank[qi] - setpos[r]] = (int32_t)qi;
}

__global__ void k_bfill(int64_t *p, int64_t n, int64_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// kNN of one work item by the stack walk from the set's root: the block's
// waves take turns on slots of 64 / G queries (in leaf order), each query
// held by G lanes.
template <typename T, int D, bool PER, int KMAX, int SD, int NT>
__device__ __forceinline__ void batch_knn_stack(const TreeView<T> &tv, const BatchItem &w, const T *queries,
                                                const int32_t *order, const int32_t *leaf_of, int k, int64_t base,
                                                int64_t kbase, int64_t q0set, int sorted, T *kdist, int64_t *kidx,
                                                int G, T *buf_d, int32_t *buf_s, StackEntry<T, D> *stk) {
    constexpr int CB = KMAX < 16 ? 0 : (sizeof(T) == 4 ? 16 : 8);
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE, sub = lane & (G - 1),
              gbase = lane & ~(G - 1);
    const int qw = WAVE / G;
    const int nslots = (w.nq + qw - 1) / qw;
    const int2 root = tv.wroot[w.set];
    for (int sl = wv; sl < nslots; sl += NT / WAVE) {
        const int j = sl * qw + lane / G;
        const bool active = j < w.nq;
        const int64_t qi = active ? (int64_t)order[w.q0 + j] : 0;
        T q[D];
#pragma unroll
        for (int d = 0; d < D; d++) q[d] = active ? queries[qi * D + d] : T(0);
        const int64_t row = kbase + (qi - q0set) * k;
        Counters c;
        int found = 0;
        const int32_t leaf = G == 1 && active && root.y < 0 ? leaf_of[qi] : -1;  // list prefill
        knn_stack<T, D, PER, KMAX, SD, CB>(tv, root, q, active, k, sorted, T(FLT_MAX), kdist + row, kidx + row, base,
                                          c, buf_d, buf_s, threadIdx.x, NT, stk, found, G, sub, gbase, leaf);
    }
}

// GC = 1 compiles the one-lane-per-query stack walk with G folded away (as
// k_knn_s does); GC = 0 takes G from the argument.
template <typename T, int D, bool PER, int NT, int SD, int GC>
__global__ __launch_bounds__(NT, 3) void k_batch(TreeView<T> tv, const int64_t *set_off, const T *queries,
                                              const BatchItem *items, int64_t nitems, int64_t *head,
                                              const int32_t *kset, const int64_t *qoff,
                                              const int64_t *knn_off, int sorted, T *kdist, int64_t *kidx,
                                              const int32_t *order, const int32_t *leaf_of, int use_stack,
                                              int Grt) {
    const int G = GC ? GC : Grt;
    __shared__ int64_t s_item;
    constexpr int CBMAX = sizeof(T) == 4 ? 16 : 8;
    __shared__ T buf_d[CBMAX * NT];
    __shared__ int32_t buf_s[CBMAX * NT];
    __shared__ StackEntry<T, D> stk[NT / WAVE][SD];
    for (;;) {
        if (threadIdx.x == 0) s_item = (int64_t)atomicAdd((unsigned long long *)head, 1ull);
        __syncthreads();
        const int64_t it = s_item;
        __syncthreads();
        if (it >= nitems) break;
        const BatchItem w = items[it];
        if (w.kind == 0 && use_stack) {  // stack walk, whole waves
            const int k = kset[w.set];
            StackEntry<T, D> *ws = stk[threadIdx.x / WAVE];
            const int64_t base = set_off[w.set], kb = knn_off[w.set], q0s = qoff[w.set];
            if (k <= 8)
                batch_knn_stack<T, D, PER, 8, SD, NT>(tv, w, queries, order, leaf_of, k, base, kb, q0s, sorted, kdist, kidx, G,
                                                      buf_d, buf_s, ws);
            else if (k <= 16)
                batch_knn_stack<T, D, PER, 16, SD, NT>(tv, w, queries, order, leaf_of, k, base, kb, q0s, sorted, kdist, kidx, G,
                                                       buf_d, buf_s, ws);
            else if (k <= 32)
                batch_knn_stack<T, D, PER, 32, SD, NT>(tv, w, queries, order, leaf_of, k, base, kb, q0s, sorted, kdist, kidx, G,
                                                       buf_d, buf_s, ws);
            else
                batch_knn_stack<T, D, PER, 64, SD, NT>(tv, w, queries, order, leaf_of, k, base, kb, q0s, sorted, kdist, kidx, G,
                                                       buf_d, buf_s, ws);
            continue;
        }
        if ((int)threadIdx.x >= w.nq) continue;
        // one walk per query, the item's queries taken in leaf order so
        // that neighbouring lanes walk the same nodes
        const int64_t qg = order ? (int64_t)order[w.q0 + threadIdx.x] : w.q0 + threadIdx.x;
        T q[D];
#pragma unroll
        for (int d = 0; d < D; d++) q[d] = queries[qg * D + d];
        const int32_t root = w.set;
        const int64_t base = set_off[w.set];
        {  // one kNN walk per query
            const int k = kset[w.set];
            const int64_t local = qg - qoff[w.set];
            T *drow = kdist + knn_off[w.set] + local * k;
            int64_t *irow = kidx + knn_off[w.set] + local * k;
            const int32_t leaf = leaf_of ? leaf_of[qg] : find_leaf<T, D>(tv, q, root);
            Counters c;
            if (k <= 8) knn_one<T, D, PER, 8>(tv, q, leaf, k, sorted, T(FLT_MAX), drow, irow, base, c);
            else if (k <= 16) knn_one<T, D, PER, 16>(tv, q, leaf, k, sorted, T(FLT_MAX), drow, irow, base, c);
            else if (k <= 32) knn_one<T, D, PER, 32>(tv, q, leaf, k, sorted, T(FLT_MAX), drow, irow, base, c);
            else knn_one<T, D, PER, 64>(tv, q, leaf, k, sorted, T(FLT_MAX), drow, irow, base, c);
        }
    }
}

// The ball items (kind 1: count, kind 2: fill), one walk per query, in a
// persistent kernel of their own: without the kNN lists it needs far fewer
// registers than k_batch, so more waves hide the walks' load latency.
template <typename T, int D, bool PER, int NT>
__global__ __launch_bounds__(NT) void k_batch_ball(TreeView<T> tv, const int64_t *set_off, const T *queries,
                                                   const BatchItem *items, int64_t nitems, int64_t *head,
                                                   const double *rset, const int32_t *order, int64_t *ball_counts,
                                                   const int64_t *ball_off, T *bdist, int64_t *bidx,
                                                   int64_t capacity, int64_t *overflow, int64_t step_limit) {
    __shared__ int64_t s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int64_t)atomicAdd((unsigned long long *)head, 1ull);
        __syncthreads();
        const int64_t it = s_item;
        __syncthreads();
        if (it >= nitems) break;
        const BatchItem w = items[it];
        if ((int)threadIdx.x >= w.nq) continue;
        // the item's queries in the set's leaf order, so neighbouring lanes
        // walk the same nodes
        const int64_t qg = order ? (int64_t)order[w.q0 + threadIdx.x] : w.q0 + threadIdx.x;
        T q[D];
#pragma unroll
        for (int d = 0; d < D; d++) q[d] = queries[qg * D + d];
        const int64_t base = set_off[w.set];
        const T r = (T)rset[w.set];
        const T r2 = r * r;
        if (tv.wparent) {  // over the wide records (present when rows fit int32)
            const int2 rp = tv.wroot[w.set];
            const T *rb = tv.dbox + (int64_t)w.set * 2 * D;
            bool bad = false;
            if (w.kind == 1) {
                ball_counts[qg] = ball_wide_one<T, D, PER, false>(tv, q, rp, rb, r2, nullptr, nullptr, base,
                                                                  step_limit, bad);
            } else {
                const int64_t o = ball_off[qg];
                if (ball_off[qg + 1] <= capacity)
                    ball_wide_one<T, D, PER, true>(tv, q, rp, rb, r2, bdist + o, bidx + o, base, step_limit, bad);
                else atomicOr((unsigned long long *)overflow, 1ull);
            }
            if (bad) atomicOr((unsigned long long *)overflow, 2ull);
        } else if (w.kind == 1) {
            ball_counts[qg] = ball_one<T, D, PER, false>(tv, q, w.set, r2, nullptr, nullptr, base);
        } else {
            const int64_t o = ball_off[qg];
            if (ball_off[qg + 1] <= capacity) ball_one<T, D, PER, true>(tv, q, w.set, r2, bdist + o, bidx + o, base);
            else atomicOr((unsigned long long *)overflow, 1ull);
        }
    }
}

// The ball items with one walk per wave (ball_stack) over each set's wide
// records: the item's queries are in the set's leaf order and its radius is
// the set's, so a wave's 64 lanes share their walk.  Every lane of a wave
// takes part (a lane past the item's queries is inactive), because the walk
// is wave-cooperative: its votes, and the drain of whole-subtree rows,
// which spreads each queued range over all 64 lanes.  Each wave has its own
// LDS stack and emit queue, shared with nothing else.  A lane whose rows
// would pass the ball capacity writes nothing and raises the overflow flag.
template <typename T, int D, bool PER, int NT, int SD>
__global__ __launch_bounds__(NT) void k_batch_ball_s(TreeView<T> tv, const int64_t *set_off, const T *queries,
                                                     const BatchItem *items, int64_t nitems, int64_t *head,
                                                     const double *rset, const int32_t *order, int64_t *ball_counts,
                                                     const int64_t *ball_off, T *bdist, int64_t *bidx,
                                                     int64_t capacity, int64_t *overflow) {
    __shared__ int64_t s_item;
    __shared__ BallEntry stk[NT / WAVE][SD];
    __shared__ EmitEntry eq[NT / WAVE][EMIT_CAP];
    const int wv = threadIdx.x / WAVE;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int64_t)atomicAdd((unsigned long long *)head, 1ull);
        __syncthreads();
        const int64_t it = s_item;
        __syncthreads();
        if (it >= nitems) break;
        const BatchItem w = items[it];
        if ((int)(threadIdx.x & ~(WAVE - 1)) >= w.nq) continue;  // whole wave without queries
        const bool has = (int)threadIdx.x < w.nq;
        const int64_t qg = has ? (order ? (int64_t)order[w.q0 + threadIdx.x] : w.q0 + threadIdx.x) : 0;
        T q[D];
#pragma unroll
        for (int d = 0; d < D; d++) q[d] = has ? queries[qg * D + d] : T(0);
        const int64_t base = set_off[w.set];
        const T r = (T)rset[w.set];
        const T r2 = r * r;
        const int2 rp = tv.wroot[w.set];
        const int2 rows = rp.y >= 0 ? rp : tv.wrange[rp.x];
        const T *rb = tv.dbox + (int64_t)w.set * 2 * D;
        if (w.kind == 1) {
            const int64_t c = ball_stack<T, D, PER, false, SD>(tv, rp, rb, rows, q, has, r2, nullptr, nullptr, 0, base,
                                                               stk[wv], eq[wv]);
            if (has) ball_counts[qg] = c;
        } else {
            bool act = has;
            int64_t o = 0;
            if (has) {
                o = ball_off[qg];
                if (ball_off[qg + 1] > capacity) {
                    act = false;
                    atomicOr((unsigned long long *)overflow, 1ull);
                }
            }
            ball_stack<T, D, PER, true, SD>(tv, rp, rb, rows, q, act, r2, bdist, bidx, o, base, stk[wv], eq[wv]);
        }
    }
}

template <typename T, int D, bool PER, int NT, int SD>
int run_t(ckd_batch *b, const T *queries, T *kdist, int64_t *kidx, int64_t *ball_off, T *bdist, int64_t *bidx,
          hipStream_t s) {
    const ckd_tree *t = b->tree;
    const TreeView<T> tv = view<T>(t);
    k_batch_reset<<<1, 1, 0, s>>>(b->counters);
    // leaf order of each set's queries: the stack kNN walk needs it, the
    // per-query walks take their queries in it
    const int32_t *order = nullptr;
    const int use_stack = b->has_knn && b->knn_kernel == CKD_KNN_STACK ? 1 : 0;
    if (b->nq > 0) {
        const unsigned gq = (unsigned)ceil_div(b->nq, 256), gn = (unsigned)ceil_div(t->nnodes, 256);
        k_bfill<<<gn, 256, 0, s>>>(b->leaf_cnt, t->nnodes, 0);
        k_bfill<<<(unsigned)ceil_div(b->nsets, 256), 256, 0, s>>>(b->setpos, b->nsets, INT64_MAX);
        k_blocate<T, D><<<gq, 256, 0, s>>>(tv, queries, b->nq, b->qoff, b->nsets, b->leaf_of, b->rank, b->leaf_cnt);
        int rc = scan_i64(b->leaf_cnt, b->leaf_off, t->nnodes, false, b->order_tmp, s);
        if (rc) return rc;
        k_bsetpos<<<gq, 256, 0, s>>>(b->leaf_of, b->leaf_off, b->qoff, b->nsets, b->nq, b->setpos);
        k_border<<<gq, 256, 0, s>>>(b->leaf_of, b->rank, b->leaf_off, b->qoff, b->nsets, b->nq, b->setpos, b->order);
        CKD_HIP(hipGetLastError());
        order = b->order;
    }
    const int G = WAVE / b->qpw;
    const int64_t step_limit = 3 * t->nwide + 3;  // bound of the ball walk over the wide records
    if (b->n_a > 0) {  // kNN items
        if (G == 1 || !use_stack)
            k_batch<T, D, PER, NT, SD, 1><<<b->grid, NT, 0, s>>>(tv, t->set_off, queries, b->items_a, b->n_a,
                                                                 b->counters, b->kset, b->qoff, b->knn_off, b->sorted,
                                                                 kdist, kidx, order, b->leaf_of, use_stack, 1);
        else
            k_batch<T, D, PER, NT, SD, 0><<<b->grid, NT, 0, s>>>(tv, t->set_off, queries, b->items_a, b->n_a,
                                                                 b->counters, b->kset, b->qoff, b->knn_off, b->sorted,
                                                                 kdist, kidx, order, b->leaf_of, use_stack, G);
    }
    CKD_HIP(hipGetLastError());
    if (b->has_ball) {
        // the shared walk needs the wide records, queries in leaf order and a
        // stack that holds the forest's depth
        const bool ball_stack_walk = b->ball_kernel == CKD_KNN_STACK && tv.wparent && order && t->depth <= SD;
        if (b->n_c > 0) {  // ball counts
            if (ball_stack_walk)
                k_batch_ball_s<T, D, PER, NT, SD><<<b->grid, NT, 0, s>>>(tv, t->set_off, queries, b->items_c, b->n_c,
                                                                         b->counters + 3, b->rset, order, b->counts,
                                                                         ball_off, bdist, bidx, b->capacity,
                                                                         b->counters + 2);
            else
                k_batch_ball<T, D, PER, NT><<<b->grid, NT, 0, s>>>(tv, t->set_off, queries, b->items_c, b->n_c,
                                                                   b->counters + 3, b->rset, order, b->counts, ball_off,
                                                                   bdist, bidx, b->capacity, b->counters + 2,
                                                                   step_limit);
        }
        int rc = scan_i64(b->counts, ball_off, b->nq, true, b->scan_tmp, s);
        if (rc) return rc;
        if (b->n_b > 0) {  // ball fills
            if (ball_stack_walk)
                k_batch_ball_s<T, D, PER, NT, SD><<<b->grid, NT, 0, s>>>(tv, t->set_off, queries, b->items_b, b->n_b,
                                                                         b->counters + 1, b->rset, order, b->counts,
                                                                         ball_off, bdist, bidx, b->capacity,
                                                                         b->counters + 2);
            else
                k_batch_ball<T, D, PER, NT><<<b->grid, NT, 0, s>>>(tv, t->set_off, queries, b->items_b, b->n_b,
                                                                   b->counters + 1, b->rset, order, b->counts, ball_off,
                                                                   bdist, bidx, b->capacity, b->counters + 2,
                                                                   step_limit);
            // the fill walks wrote d^2; rows past the capacity were not written
            k_sqrt_rows<T><<<SQRT_BLOCKS, 256, 0, s>>>(bdist, ball_off + b->nq, b->capacity);
        }
        CKD_HIP(hipGetLastError());
    }
    return CKD_OK;
}

template <typename T, bool PER>
int run_nt(ckd_batch *b, const void *q, void *kd, int64_t *ki, int64_t *bo, void *bd, int64_t *bi, hipStream_t s) {
    if (b->tree->ndim != 3) {
        set_error("the batched engine supports 3-D forests");
        return CKD_ERR_ARG;
    }
    // a stack shorter than the tree is never used: 64 entries if 32 do not
    // fit, one walk per query if 64 do not
    if ((b->tree->depth > 64 || !b->tree->wide) && b->knn_kernel == CKD_KNN_STACK) b->knn_kernel = CKD_KNN_LANE;
    const bool sd32 = b->tree->depth <= 32 && b->stack != 64;
    switch (b->nt) {
        case 64:
            return sd32 ? run_t<T, 3, PER, 64, 32>(b, (const T *)q, (T *)kd, ki, bo, (T *)bd, bi, s)
                        : run_t<T, 3, PER, 64, 64>(b, (const T *)q, (T *)kd, ki, bo, (T *)bd, bi, s);
        case 128:
            return sd32 ? run_t<T, 3, PER, 128, 32>(b, (const T *)q, (T *)kd, ki, bo, (T *)bd, bi, s)
                        : run_t<T, 3, PER, 128, 64>(b, (const T *)q, (T *)kd, ki, bo, (T *)bd, bi, s);
        default:
            return sd32 ? run_t<T, 3, PER, 256, 32>(b, (const T *)q, (T *)kd, ki, bo, (T *)bd, bi, s)
                        : run_t<T, 3, PER, 256, 64>(b, (const T *)q, (T *)kd, ki, bo, (T *)bd, bi, s);
    }
}

}  // namespace
}  // namespace ckd

