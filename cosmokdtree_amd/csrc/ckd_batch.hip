// ckd_batch.hip -- batched queries over a forest of independent point sets
// in persistent kernels that drain a device work queue.
//
// A plan lists work items (set, kind, first query, query count <= NT):
// phase A holds the kNN items (largest k first) and the ball-count items,
// phase B the ball-fill items.  Each launch starts one grid of NT-thread
// blocks that pull items from an atomic head counter until the queue is
// empty, so sets of very different sizes, k and radii share one launch.
// Between the phases a device scan turns the ball counts into CSR offsets.
// ckd_batch_run only launches kernels (no allocation, copy or host wait), so
// a caller can capture it in a hipGraph and replay it.
//
// Per set r: queries are rows [qoff[r], qoff[r+1]) of the query array, kNN
// rows of k[r] entries start at element knn_off[r] of the kNN outputs, and
// indices are relative to the set (0 = the set's first point).  Results are
// those of the single-tree calls on the set's own tree.
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
    int sorted = 1, nt = 256, grid = 1024, has_ball = 0;
    BatchItem *items_a = nullptr, *items_b = nullptr;
    int64_t n_a = 0, n_b = 0;
    int32_t *kset = nullptr;
    double *rset = nullptr;
    int64_t *qoff = nullptr, *knn_off = nullptr, *counters = nullptr, *counts = nullptr;
    void *scan_tmp = nullptr;
};

namespace ckd {
namespace {

__global__ void k_batch_reset(int64_t *counters) {
    counters[0] = 0;  // phase A head
    counters[1] = 0;  // phase B head
    counters[2] = 0;  // ball overflow flag
}

template <typename T, int D, bool PER, int NT>
__global__ __launch_bounds__(NT) void k_batch(TreeView<T> tv, const int64_t *set_off, const T *queries,
                                              const BatchItem *items, int64_t nitems, int64_t *head,
                                              const int32_t *kset, const double *rset, const int64_t *qoff,
                                              const int64_t *knn_off, int sorted, T *kdist, int64_t *kidx,
                                              int64_t *ball_counts, const int64_t *ball_off, T *bdist,
                                              int64_t *bidx, int64_t capacity, int64_t *overflow) {
    __shared__ int64_t s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int64_t)atomicAdd((unsigned long long *)head, 1ull);
        __syncthreads();
        const int64_t it = s_item;
        __syncthreads();
        if (it >= nitems) break;
        const BatchItem w = items[it];
        if ((int)threadIdx.x >= w.nq) continue;
        const int64_t qg = w.q0 + threadIdx.x;
        T q[D];
#pragma unroll
        for (int d = 0; d < D; d++) q[d] = queries[qg * D + d];
        const int32_t root = w.set;
        const int64_t base = set_off[w.set];
        if (w.kind == 0) {
            const int k = kset[w.set];
            const int64_t local = qg - qoff[w.set];
            T *drow = kdist + knn_off[w.set] + local * k;
            int64_t *irow = kidx + knn_off[w.set] + local * k;
            const int32_t leaf = find_leaf<T, D>(tv, q, root);
            Counters c;
            if (k <= 8) knn_one<T, D, PER, 8>(tv, q, leaf, k, sorted, T(FLT_MAX), drow, irow, base, c);
            else if (k <= 16) knn_one<T, D, PER, 16>(tv, q, leaf, k, sorted, T(FLT_MAX), drow, irow, base, c);
            else if (k <= 32) knn_one<T, D, PER, 32>(tv, q, leaf, k, sorted, T(FLT_MAX), drow, irow, base, c);
            else knn_one<T, D, PER, 64>(tv, q, leaf, k, sorted, T(FLT_MAX), drow, irow, base, c);
        } else {
            const T r = (T)rset[w.set];
            const T r2 = r * r;
            if (w.kind == 1) {
                ball_counts[qg] = ball_one<T, D, PER, false>(tv, q, root, r2, nullptr, nullptr, base);
            } else {
                const int64_t o = ball_off[qg];
                if (ball_off[qg + 1] <= capacity) ball_one<T, D, PER, true>(tv, q, root, r2, bdist + o, bidx + o, base);
                else atomicOr((unsigned long long *)overflow, 1ull);
            }
        }
    }
}

template <typename T, int D, bool PER, int NT>
int run_t(ckd_batch *b, const T *queries, T *kdist, int64_t *kidx, int64_t *ball_off, T *bdist, int64_t *bidx,
          hipStream_t s) {
    const ckd_tree *t = b->tree;
    const TreeView<T> tv = view<T>(t);
    k_batch_reset<<<1, 1, 0, s>>>(b->counters);
    if (b->n_a > 0)
        k_batch<T, D, PER, NT><<<b->grid, NT, 0, s>>>(tv, t->set_off, queries, b->items_a, b->n_a, b->counters,
                                                      b->kset, b->rset, b->qoff, b->knn_off, b->sorted, kdist, kidx,
                                                      b->counts, ball_off, bdist, bidx, b->capacity, b->counters + 2);
    CKD_HIP(hipGetLastError());
    if (b->has_ball) {
        int rc = scan_i64(b->counts, ball_off, b->nq, true, b->scan_tmp, s);
        if (rc) return rc;
        if (b->n_b > 0)
            k_batch<T, D, PER, NT><<<b->grid, NT, 0, s>>>(tv, t->set_off, queries, b->items_b, b->n_b,
                                                          b->counters + 1, b->kset, b->rset, b->qoff, b->knn_off,
                                                          b->sorted, kdist, kidx, b->counts, ball_off, bdist, bidx,
                                                          b->capacity, b->counters + 2);
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
    switch (b->nt) {
        case 64: return run_t<T, 3, PER, 64>(b, (const T *)q, (T *)kd, ki, bo, (T *)bd, bi, s);
        case 128: return run_t<T, 3, PER, 128>(b, (const T *)q, (T *)kd, ki, bo, (T *)bd, bi, s);
        default: return run_t<T, 3, PER, 256>(b, (const T *)q, (T *)kd, ki, bo, (T *)bd, bi, s);
    }
}

}  // namespace
}  // namespace ckd

using namespace ckd;

extern "C" {

int ckd_batch_free(ckd_batch *b) {
    if (!b) return CKD_OK;
    void *ps[] = {b->items_a, b->items_b, b->kset, b->rset, b->qoff, b->knn_off, b->counters, b->counts, b->scan_tmp};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    delete b;
    return CKD_OK;
}

int ckd_batch_plan(const ckd_tree *t, const int64_t *qoff, const int32_t *k, const double *radius, int32_t sorted,
                   int32_t threads, int32_t blocks, int64_t ball_capacity, ckd_batch **out, int64_t *knn_total) {
    if (!t || !qoff || !k || !out) {
        set_error("invalid ckd_batch_plan arguments");
        return CKD_ERR_ARG;
    }
    *out = nullptr;
    if (threads != 64 && threads != 128 && threads != 256) {
        set_error("threads per block must be 64, 128 or 256");
        return CKD_ERR_ARG;
    }
    if (t->ndim != 3) {
        set_error("the batched engine supports 3-D forests");
        return CKD_ERR_ARG;
    }
    const int64_t ns = t->nsets;
    if (qoff[0] != 0) {
        set_error("query offsets must start at 0");
        return CKD_ERR_ARG;
    }
    std::vector<int64_t> knn_off(ns + 1, 0);
    std::vector<BatchItem> a, bb;
    bool any_ball = false;
    for (int64_t r = 0; r < ns; r++) {
        const int64_t nq = qoff[r + 1] - qoff[r];
        if (nq < 0 || k[r] < 0 || k[r] > 64) {
            set_error("set %lld: negative query count or k outside [0, 64]", (long long)r);
            return CKD_ERR_ARG;
        }
        knn_off[r + 1] = knn_off[r] + nq * (int64_t)k[r];
        const bool ball = radius && radius[r] > 0;
        any_ball |= ball;
        for (int64_t q0 = 0; q0 < nq; q0 += threads) {
            BatchItem w;
            w.set = (int32_t)r;
            w.q0 = qoff[r] + q0;
            w.nq = (int32_t)std::min<int64_t>(threads, nq - q0);
            w.pad = 0;
            if (k[r] > 0) {
                w.kind = 0;
                a.push_back(w);
            }
            if (ball) {
                w.kind = 1;
                a.push_back(w);
                w.kind = 2;
                bb.push_back(w);
            }
        }
    }
    // heaviest items first: kNN by descending k, then ball counts
    std::stable_sort(a.begin(), a.end(), [&](const BatchItem &x, const BatchItem &y) {
        const int kx = x.kind == 0 ? k[x.set] : -1, ky = y.kind == 0 ? k[y.set] : -1;
        return kx > ky;
    });
    ckd_batch *b = new ckd_batch();
    b->tree = t;
    b->nsets = ns;
    b->nq = qoff[ns];
    b->knn_total = knn_off[ns];
    b->capacity = ball_capacity;
    b->sorted = sorted ? 1 : 0;
    b->nt = threads;
    b->has_ball = any_ball ? 1 : 0;
    int dev_cu = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, t->device) == hipSuccess) dev_cu = prop.multiProcessorCount;
    b->grid = blocks > 0 ? blocks : dev_cu * (2048 / threads);
    b->n_a = (int64_t)a.size();
    b->n_b = (int64_t)bb.size();
    std::vector<double> rs(ns, 0.0);
    if (radius)
        for (int64_t r = 0; r < ns; r++) rs[r] = radius[r];
    auto up = [&](void **p, const void *src, size_t bytes) -> hipError_t {
        hipError_t e = hipMalloc(p, bytes ? bytes : 16);
        if (e == hipSuccess && bytes && src) e = hipMemcpy(*p, src, bytes, hipMemcpyHostToDevice);
        return e;
    };
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = up((void **)&b->items_a, a.data(), sizeof(BatchItem) * a.size());
    if (e == hipSuccess) e = up((void **)&b->items_b, bb.data(), sizeof(BatchItem) * bb.size());
    if (e == hipSuccess) e = up((void **)&b->kset, k, sizeof(int32_t) * ns);
    if (e == hipSuccess) e = up((void **)&b->rset, rs.data(), sizeof(double) * ns);
    if (e == hipSuccess) e = up((void **)&b->qoff, qoff, sizeof(int64_t) * (ns + 1));
    if (e == hipSuccess) e = up((void **)&b->knn_off, knn_off.data(), sizeof(int64_t) * (ns + 1));
    if (e == hipSuccess) e = up((void **)&b->counters, nullptr, sizeof(int64_t) * 4);
    if (e == hipSuccess) e = up((void **)&b->counts, nullptr, sizeof(int64_t) * (b->nq > 0 ? b->nq : 1));
    if (e == hipSuccess) e = up(&b->scan_tmp, nullptr, scan_tmp_bytes(b->nq));
    if (e != hipSuccess) {
        ckd_batch_free(b);
        return hip_fail(e, "ckd_batch_plan", __FILE__, __LINE__);
    }
    if (knn_total) *knn_total = b->knn_total;
    *out = b;
    return CKD_OK;
}

int ckd_batch_run(ckd_batch *b, const void *queries, void *knn_dist, int64_t *knn_idx, int64_t *ball_off,
                  void *ball_dist, int64_t *ball_idx, void *stream) {
    if (!b || (b->nq > 0 && !queries) || (b->knn_total > 0 && (!knn_dist || !knn_idx)) ||
        (b->has_ball && (!ball_off || (b->capacity > 0 && (!ball_dist || !ball_idx))))) {
        set_error("invalid ckd_batch_run arguments");
        return CKD_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    const ckd_tree *t = b->tree;
    if (t->prec == 4)
        return t->periodic ? run_nt<float, true>(b, queries, knn_dist, knn_idx, ball_off, ball_dist, ball_idx, s)
                           : run_nt<float, false>(b, queries, knn_dist, knn_idx, ball_off, ball_dist, ball_idx, s);
    return t->periodic ? run_nt<double, true>(b, queries, knn_dist, knn_idx, ball_off, ball_dist, ball_idx, s)
                       : run_nt<double, false>(b, queries, knn_dist, knn_idx, ball_off, ball_dist, ball_idx, s);
}

int ckd_batch_status(const ckd_batch *b, const int64_t *ball_off, int64_t *ball_total, int32_t *overflow,
                     void *stream) {
    if (!b) {
        set_error("plan must not be NULL");
        return CKD_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    int64_t flag = 0, tot = 0;
    CKD_HIP(hipMemcpyAsync(&flag, b->counters + 2, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    if (b->has_ball && ball_off)
        CKD_HIP(hipMemcpyAsync(&tot, ball_off + b->nq, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    CKD_HIP(hipStreamSynchronize(s));
    if (overflow) *overflow = flag ? 1 : 0;
    if (ball_total) *ball_total = tot;
    return CKD_OK;
}

}  // extern "C"
