// ckd_batch.hip -- the plan and the C entry points of the batched engine
// (kernels in ckd_batch_impl.h).
#include "ckd_batch_impl.h"

namespace ckd {
int batch_run_f32_per(ckd_batch *, const void *, void *, int64_t *, int64_t *, void *, int64_t *, hipStream_t);
int batch_run_f32_open(ckd_batch *, const void *, void *, int64_t *, int64_t *, void *, int64_t *, hipStream_t);
int batch_run_f64_per(ckd_batch *, const void *, void *, int64_t *, int64_t *, void *, int64_t *, hipStream_t);
int batch_run_f64_open(ckd_batch *, const void *, void *, int64_t *, int64_t *, void *, int64_t *, hipStream_t);
}  // namespace ckd

using namespace ckd;

extern "C" {

int ckd_batch_free(ckd_batch *b) {
    if (!b) return CKD_OK;
    void *ps[] = {b->items_a, b->items_b, b->items_c, b->kset, b->rset, b->qoff, b->knn_off, b->counters,
                  b->counts,  b->scan_tmp, b->leaf_of, b->rank, b->order, b->leaf_cnt, b->leaf_off,
                  b->setpos,  b->order_tmp};
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
    // the batch kernels hold rows, query positions and leaf ranks in 32 bits
    if (t->n >= INT32_MAX || qoff[ns] >= INT32_MAX || t->nnodes >= INT32_MAX) {
        set_error("the batched engine takes forests and query batches below 2^31 - 1 rows "
                  "(use ckd_knn / ckd_ball_* per tree beyond)");
        return CKD_ERR_LIMIT;
    }
    std::vector<int64_t> knn_off(ns + 1, 0);
    std::vector<BatchItem> a, bb, cc;
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
                cc.push_back(w);
                w.kind = 2;
                bb.push_back(w);
            }
        }
    }
    // heaviest kNN items first (descending k)
    std::stable_sort(a.begin(), a.end(), [&](const BatchItem &x, const BatchItem &y) { return k[x.set] > k[y.set]; });
    ckd_batch *b = new ckd_batch();
    b->tree = t;
    b->nsets = ns;
    b->nq = qoff[ns];
    b->knn_total = knn_off[ns];
    b->capacity = ball_capacity;
    b->sorted = sorted ? 1 : 0;
    b->nt = threads;
    b->has_ball = any_ball ? 1 : 0;
    b->has_knn = knn_off[ns] > 0 ? 1 : 0;
    // default kNN and ball kernels (measured on MI355X): shared walks pay
    // when the queries are at least 1/20 of the points; sparser batches run
    // one walk per query.  On 2e8 points in 4096 sets with queries 1/10 of
    // the points, the shared ball walk took 190 ms against 461 ms.
    b->knn_kernel = qoff[ns] * 20 >= t->n ? CKD_KNN_STACK : CKD_KNN_LANE;
    b->ball_kernel = b->knn_kernel;
    int dev_cu = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, t->device) == hipSuccess) dev_cu = prop.multiProcessorCount;
    b->grid = blocks > 0 ? blocks : dev_cu * (2048 / threads);
    b->n_a = (int64_t)a.size();
    b->n_b = (int64_t)bb.size();
    b->n_c = (int64_t)cc.size();
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
    if (e == hipSuccess) e = up((void **)&b->items_c, cc.data(), sizeof(BatchItem) * cc.size());
    if (e == hipSuccess) e = up((void **)&b->kset, k, sizeof(int32_t) * ns);
    if (e == hipSuccess) e = up((void **)&b->rset, rs.data(), sizeof(double) * ns);
    if (e == hipSuccess) e = up((void **)&b->qoff, qoff, sizeof(int64_t) * (ns + 1));
    if (e == hipSuccess) e = up((void **)&b->knn_off, knn_off.data(), sizeof(int64_t) * (ns + 1));
    if (e == hipSuccess) e = up((void **)&b->counters, nullptr, sizeof(int64_t) * 4);
    if (e == hipSuccess) e = up((void **)&b->counts, nullptr, sizeof(int64_t) * (b->nq > 0 ? b->nq : 1));
    // queries of sets without a ball search keep a zero count: only the
    // ball-count items write counts
    if (e == hipSuccess) e = hipMemset(b->counts, 0, sizeof(int64_t) * (b->nq > 0 ? b->nq : 1));
    if (e == hipSuccess) e = up(&b->scan_tmp, nullptr, scan_tmp_bytes(b->nq));
    if (b->nq > 0) {  // leaf ordering of the queries (stack kNN walk, query order of the other walks)
        const size_t nqb = (size_t)(b->nq > 0 ? b->nq : 1);
        if (e == hipSuccess) e = up((void **)&b->leaf_of, nullptr, sizeof(int32_t) * nqb);
        if (e == hipSuccess) e = up((void **)&b->rank, nullptr, sizeof(int32_t) * nqb);
        if (e == hipSuccess) e = up((void **)&b->order, nullptr, sizeof(int32_t) * nqb);
        if (e == hipSuccess) e = up((void **)&b->leaf_cnt, nullptr, sizeof(int64_t) * t->nnodes);
        if (e == hipSuccess) e = up((void **)&b->leaf_off, nullptr, sizeof(int64_t) * (t->nnodes + 1));
        if (e == hipSuccess) e = up((void **)&b->setpos, nullptr, sizeof(int64_t) * ns);
        if (e == hipSuccess) e = up(&b->order_tmp, nullptr, scan_tmp_bytes(t->nnodes));
    }
    if (e != hipSuccess) {
        ckd_batch_free(b);
        return hip_fail(e, "ckd_batch_plan", __FILE__, __LINE__);
    }
    if (knn_total) *knn_total = b->knn_total;
    *out = b;
    return CKD_OK;
}

int ckd_batch_config(ckd_batch *b, int32_t knn_kernel, int32_t queries_per_wave, int32_t stack_depth) {
    if (!b || (knn_kernel != CKD_KNN_LANE && knn_kernel != CKD_KNN_STACK) ||
        (queries_per_wave < 1 || queries_per_wave > 64 || (queries_per_wave & (queries_per_wave - 1)) != 0) ||
        (stack_depth != 0 && stack_depth != 32 && stack_depth != 64)) {
        set_error("ckd_batch_config: kernel CKD_KNN_LANE or CKD_KNN_STACK, queries per wave a power of two up to 64, "
                  "stack depth 0, 32 or 64");
        return CKD_ERR_ARG;
    }
    b->knn_kernel = knn_kernel;
    b->qpw = queries_per_wave;
    b->stack = stack_depth;
    return CKD_OK;
}

int ckd_batch_config_ball(ckd_batch *b, int32_t ball_kernel) {
    if (!b || (ball_kernel != CKD_KNN_LANE && ball_kernel != CKD_KNN_STACK)) {
        set_error("ckd_batch_config_ball: kernel CKD_KNN_LANE or CKD_KNN_STACK");
        return CKD_ERR_ARG;
    }
    b->ball_kernel = ball_kernel;
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
        return t->periodic ? batch_run_f32_per(b, queries, knn_dist, knn_idx, ball_off, ball_dist, ball_idx, s)
                           : batch_run_f32_open(b, queries, knn_dist, knn_idx, ball_off, ball_dist, ball_idx, s);
    return t->periodic ? batch_run_f64_per(b, queries, knn_dist, knn_idx, ball_off, ball_dist, ball_idx, s)
                       : batch_run_f64_open(b, queries, knn_dist, knn_idx, ball_off, ball_dist, ball_idx, s);
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
    if (overflow) *overflow = (flag & 1) ? 1 : 0;
    if (ball_total) *ball_total = tot;
    if (flag & 2) {  // a ball walk ran past its step bound: the tree's parent links are inconsistent
        set_error("ball walk exceeded its step bound (corrupt wide records)");
        return CKD_ERR_HIP;
    }
    return CKD_OK;
}

}  // extern "C"
