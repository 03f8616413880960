// ckd_api.hip -- the extern "C" entry points declared in
// include/cosmokdtree_amd.h: argument checks, tree lifetime, error text.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "ckd_internal.h"

namespace ckd {
int query_knn(const ckd_tree *t, const void *q, int64_t nq, int k, int sorted, const float *seed, void *dist,
              int64_t *idx, int64_t *stats, Scratch &sc, hipStream_t s);
int query_knn_learned(const ckd_tree *t, const void *q, int64_t nq, int k, int sorted, const float *params,
                      float margin, int G, const uint16_t *table, void *dist, int64_t *idx, int64_t *stats,
                      Scratch &sc, hipStream_t s);
int query_sph(const ckd_tree *t, const void *q, int64_t nq, int k, const void *mass, void *h, void *rho,
              Scratch &sc, hipStream_t s);
int query_ball(const ckd_tree *t, const void *q, int64_t nq, double r, int64_t *counts, const int64_t *off,
               void *dist, int64_t *idx, Scratch &sc, hipStream_t s);
int query_box(const ckd_tree *t, const void *b, int64_t nb, int64_t *counts, const int64_t *off, int64_t *idx,
              int wrap, hipStream_t s);
int query_ball_search(const ckd_tree *t, const void *q, int64_t nq, double r, int64_t *off, int64_t cap, void *dist,
                      int64_t *idx, int64_t *total, int32_t *overflow, Scratch &sc, hipStream_t s);

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    set_error("HIP error %s (%d) in %s at %s:%d", hipGetErrorString(e), (int)e, what, file, line);
    return e == hipErrorOutOfMemory ? CKD_ERR_MEMORY : CKD_ERR_HIP;
}

int Scratch::take(size_t bytes, void **p) {
    *p = nullptr;
    if (arena) {
        const size_t at = (used + 255) & ~(size_t)255;
        if (at + bytes > cap) {
            set_error("the call needs %zu bytes of scratch; the static scratch holds %zu", at + bytes, cap);
            return CKD_ERR_LIMIT;
        }
        used = at + bytes;
        *p = arena + at;
        return CKD_OK;
    }
    const int rc = dev_alloc(p, bytes, s);
    if (rc == CKD_OK) owned.push_back(*p);
    return rc;
}

// Frees every device array of the tree.  Like hipFree, it first waits for
// all work on the tree's device, so no stream still reads the arrays.
static void release(ckd_tree *t) {
    for (auto &ev : t->knn_events) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    t->knn_events.clear();
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (cur != t->device) (void)hipSetDevice(t->device);
    (void)hipDeviceSynchronize();
    void *arrays[] = {t->set_off, t->nodes, t->bbox, t->dbox,  t->range,  t->pts,
                      t->idx,     t->wide,  t->wroot, t->wrange, t->wparent};
    for (void *a : arrays) dev_free_idle(a);
    if (cur != t->device) (void)hipSetDevice(cur);
    t->set_off = nullptr;
    t->nodes = t->bbox = t->dbox = t->pts = t->idx = t->wide = nullptr;
    t->wroot = t->wrange = nullptr;
    t->wparent = nullptr;
    t->nwide = 0;
    t->range = nullptr;
}

// CSR offsets from per-row counts held in offsets[1..n]: exclusive scan in
// place through a temporary buffer, total copied to the host.
static int counts_to_offsets(int64_t *offsets, int64_t n, int64_t *total, hipStream_t s) {
    int64_t *tmp = nullptr;
    void *scr = nullptr;
    int rc = dev_alloc((void **)&tmp, sizeof(int64_t) * (n > 0 ? n : 1), s);
    if (!rc) rc = dev_alloc(&scr, scan_tmp_bytes(n), s);
    if (rc) {
        dev_free(tmp, s);
        return rc;
    }
    if (n > 0) CKD_HIP(hipMemcpyAsync(tmp, offsets + 1, sizeof(int64_t) * n, hipMemcpyDeviceToDevice, s));
    rc = scan_i64(tmp, offsets, n, true, scr, s);
    int64_t tot = 0;
    if (!rc) {
        hipError_t e = hipMemcpyAsync(&tot, offsets + n, sizeof(int64_t), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = hip_fail(e, "offsets total", __FILE__, __LINE__);
    }
    dev_free(tmp, s);
    dev_free(scr, s);
    if (total) *total = tot;
    return rc;
}

}  // namespace ckd

using namespace ckd;

extern "C" {

const char *ckd_last_error(void) { return g_err; }

const char *ckd_version(void) { return "cosmokdtree_amd 0.1 (gfx950)"; }

static int build_common(const void *points, int64_t n, const int64_t *set_offsets, int64_t nsets, int32_t ndim,
                        int32_t prec, int32_t leaf, const double *boxsize, int32_t idx_bits, void *stream,
                        ckd_tree **out) {
    g_err[0] = 0;
    if (!out) {
        set_error("out must not be NULL");
        return CKD_ERR_ARG;
    }
    *out = nullptr;
    if (ndim < 1 || ndim > CKD_MAX_DIM) {
        set_error("Input points must have between 1 and %d dimensions, got %d", CKD_MAX_DIM, ndim);
        return CKD_ERR_ARG;
    }
    if (prec != 4 && prec != 8) {
        set_error("prec must be 4 or 8, got %d", prec);
        return CKD_ERR_ARG;
    }
    if (leaf < 1) {
        set_error("Leaf size must be a positive integer");
        return CKD_ERR_ARG;
    }
    if (n < 1 || !points) {
        set_error("at least one point is required");
        return CKD_ERR_ARG;
    }
    if (nsets < 1 || nsets > INT32_MAX / 4) {
        set_error("invalid number of point sets %lld", (long long)nsets);
        return CKD_ERR_ARG;
    }
    if (set_offsets) {
        if (set_offsets[0] != 0 || set_offsets[nsets] != n) {
            set_error("set offsets must start at 0 and end at n");
            return CKD_ERR_ARG;
        }
        for (int64_t r = 0; r < nsets; r++)
            if (set_offsets[r + 1] <= set_offsets[r]) {
                set_error("point set %lld is empty or offsets decrease", (long long)r);
                return CKD_ERR_ARG;
            }
    }
    if (idx_bits == 0) idx_bits = n < INT32_MAX ? 32 : 64;
    if (idx_bits != 32 && idx_bits != 64) {
        set_error("idx_bits must be 0, 32 or 64");
        return CKD_ERR_ARG;
    }
    if (idx_bits == 32 && n >= INT32_MAX) {
        set_error("idx_bits=32 cannot index %lld points", (long long)n);
        return CKD_ERR_LIMIT;
    }
    ckd_tree *t = new ckd_tree();
    t->n = n;
    t->ndim = ndim;
    t->prec = prec;
    t->leafsize = leaf;
    t->idx_bits = idx_bits;
    t->nsets = nsets;
    if (boxsize) {
        t->periodic = 1;
        for (int d = 0; d < ndim; d++) {
            if (!(boxsize[d] > 0)) {
                delete t;
                set_error("boxsize entries must be positive, got %g on axis %d", boxsize[d], d);
                return CKD_ERR_ARG;
            }
            t->L[d] = boxsize[d];
        }
    }
    (void)hipGetDevice(&t->device);
    hipStream_t s = (hipStream_t)stream;
    int rc = dev_alloc((void **)&t->set_off, sizeof(int64_t) * (nsets + 1), s);
    if (!rc) {
        const int64_t two[2] = {0, n};
        hipError_t e = hipMemcpyAsync(t->set_off, set_offsets ? set_offsets : two, sizeof(int64_t) * (nsets + 1),
                                      hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);  // the host array may go away on return
        rc = e == hipSuccess ? build_tree(t, points, s) : hip_fail(e, "set offsets", __FILE__, __LINE__);
    }
    if (rc) {
        release(t);
        delete t;
        return rc;
    }
    *out = t;
    return CKD_OK;
}

int ckd_build(const void *points, int64_t n, int32_t ndim, int32_t prec, int32_t leaf, const double *boxsize,
              int32_t idx_bits, void *stream, ckd_tree **out) {
    return build_common(points, n, nullptr, 1, ndim, prec, leaf, boxsize, idx_bits, stream, out);
}

int ckd_build_forest(const void *points, const int64_t *set_offsets, int64_t nsets, int32_t ndim, int32_t prec,
                     int32_t leaf, const double *boxsize, int32_t idx_bits, void *stream, ckd_tree **out) {
    if (!set_offsets || nsets < 1) {
        set_error("set_offsets must not be NULL and nsets must be positive");
        return CKD_ERR_ARG;
    }
    return build_common(points, set_offsets[nsets], set_offsets, nsets, ndim, prec, leaf, boxsize, idx_bits, stream,
                        out);
}

int ckd_free(ckd_tree *t) {
    if (!t) return CKD_OK;
    release(t);
    delete t;
    return CKD_OK;
}

int ckd_get_info(const ckd_tree *t, ckd_tree_info *info) {
    if (!t || !info) {
        set_error("tree and info must not be NULL");
        return CKD_ERR_ARG;
    }
    memset(info, 0, sizeof(*info));
    info->n = t->n;
    info->ndim = t->ndim;
    info->prec = t->prec;
    info->leafsize = t->leafsize;
    info->periodic = t->periodic;
    info->idx_bits = t->idx_bits;
    info->depth = t->depth;
    info->nnodes = t->nnodes;
    info->nleaves = t->nleaves;
    for (int d = 0; d < CKD_MAX_DIM; d++) info->boxsize[d] = t->L[d];
    info->build_ms = t->build_ms;
    info->nsets = t->nsets;
    return CKD_OK;
}

int ckd_knn(const ckd_tree *t, const void *queries, int64_t nq, int32_t k, int32_t sorted, void *dist, int64_t *idx,
            const float *seed_r, int64_t *stats, void *stream) {
    g_err[0] = 0;
    if (!t) {
        set_error("tree must not be NULL");
        return CKD_ERR_ARG;
    }
    if (k < 1) {
        set_error("k must be a positive integer, got %d", k);
        return CKD_ERR_ARG;
    }
    if (nq < 0 || (nq > 0 && (!queries || !dist || !idx))) {
        set_error("invalid query or output arrays");
        return CKD_ERR_ARG;
    }
    if (nq == 0) return CKD_OK;
    Scratch sc((hipStream_t)stream);
    return query_knn(t, queries, nq, k, sorted, seed_r, dist, idx, stats, sc, (hipStream_t)stream);
}

int ckd_knn_learned(const ckd_tree *t, const void *queries, int64_t nq, int32_t k, int32_t sorted,
                    const float *params, float margin, int32_t G, const uint16_t *table, void *dist, int64_t *idx,
                    int64_t *stats, void *stream) {
    g_err[0] = 0;
    if (!t || !params || !table) {
        set_error("tree, params and table must not be NULL");
        return CKD_ERR_ARG;
    }
    if (t->ndim != 3 || !t->periodic) {
        set_error("learned seeds need a periodic 3-D tree");
        return CKD_ERR_ARG;
    }
    if (k < 1 || k > 64) {
        set_error("learned seeds need 1 <= k <= 64, got %d", k);
        return CKD_ERR_ARG;
    }
    if (!(margin > 0.0f) || G < 1 || G > 1024) {
        set_error("invalid margin or grid size");
        return CKD_ERR_ARG;
    }
    if (nq < 0 || (nq > 0 && (!queries || !dist || !idx))) {
        set_error("invalid query or output arrays");
        return CKD_ERR_ARG;
    }
    if (nq == 0) return CKD_OK;
    Scratch sc((hipStream_t)stream);
    return query_knn_learned(t, queries, nq, k, sorted, params, margin, G, table, dist, idx, stats, sc,
                             (hipStream_t)stream);
}

int ckd_sph_density(const ckd_tree *t, const void *queries, int64_t nq, int32_t k, const void *mass, void *h,
                    void *rho, void *stream) {
    g_err[0] = 0;
    if (!t) {
        set_error("tree must not be NULL");
        return CKD_ERR_ARG;
    }
    if (t->ndim != 3) {
        set_error("the SPH density step needs a 3-D tree, got ndim = %d", t->ndim);
        return CKD_ERR_ARG;
    }
    if (k < 2 || k > 64 || k > t->n) {
        set_error("the SPH density step needs 2 <= k <= min(64, n), got k = %d", k);
        return CKD_ERR_ARG;
    }
    if (nq < 0 || (nq > 0 && (!queries || !h || !rho))) {
        set_error("invalid query or output arrays");
        return CKD_ERR_ARG;
    }
    if (nq == 0) return CKD_OK;
    Scratch sc((hipStream_t)stream);
    return query_sph(t, queries, nq, k, mass, h, rho, sc, (hipStream_t)stream);
}

int ckd_knn_config(ckd_tree *t, int32_t kernel, int32_t stack_depth, int32_t queries_per_wave) {
    g_err[0] = 0;
    if (!t) {
        set_error("tree must not be NULL");
        return CKD_ERR_ARG;
    }
    if (kernel < CKD_KNN_AUTO || kernel > CKD_KNN_STACK) {
        set_error("unknown kNN kernel %d", kernel);
        return CKD_ERR_ARG;
    }
    if (stack_depth != 0 && stack_depth != 32 && stack_depth != 64) {
        set_error("stack depth must be 0, 32 or 64, got %d", stack_depth);
        return CKD_ERR_ARG;
    }
    if (queries_per_wave < 0 || queries_per_wave > 64 || (queries_per_wave & (queries_per_wave - 1)) != 0) {
        set_error("queries per wave must be 0 or a power of two up to 64, got %d", queries_per_wave);
        return CKD_ERR_ARG;
    }
    t->knn_kernel = kernel;
    t->knn_stack = stack_depth;
    t->knn_qpw = queries_per_wave;
    return CKD_OK;
}

int ckd_knn_timing(ckd_tree *t, int32_t on, double *total_ms, int64_t *launches) {
    g_err[0] = 0;
    if (!t) {
        set_error("tree must not be NULL");
        return CKD_ERR_ARG;
    }
    double sum = 0;
    int64_t cnt = 0;
    int rc = CKD_OK;
    std::lock_guard<std::mutex> lock(t->timing_mu);
    for (auto &ev : t->knn_events) {
        float ms = 0;
        hipError_t e = hipEventSynchronize(ev.second);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev.first, ev.second);
        if (e != hipSuccess && rc == CKD_OK) rc = hip_fail(e, "hipEventElapsedTime", __FILE__, __LINE__);
        sum += ms;
        cnt++;
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    t->knn_events.clear();
    if (total_ms) *total_ms = sum;
    if (launches) *launches = cnt;
    t->timing = on != 0;
    return rc;
}

int ckd_ball_count(const ckd_tree *t, const void *queries, int64_t nq, double radius, int64_t *offsets,
                   int64_t *total, void *stream) {
    g_err[0] = 0;
    if (!t || !offsets) {
        set_error("tree and offsets must not be NULL");
        return CKD_ERR_ARG;
    }
    if (!(radius > 0)) {
        set_error("Radius must be a positive number");
        return CKD_ERR_ARG;
    }
    if (nq < 0 || (nq > 0 && !queries)) {
        set_error("invalid query array");
        return CKD_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    if (nq > 0) {
        Scratch sc(s);
        int rc = query_ball(t, queries, nq, radius, offsets + 1, nullptr, nullptr, nullptr, sc, s);
        if (rc) return rc;
    }
    return counts_to_offsets(offsets, nq, total, s);
}

int ckd_ball_fill(const ckd_tree *t, const void *queries, int64_t nq, double radius, int32_t sorted,
                  const int64_t *offsets, void *dist, int64_t *idx, void *stream) {
    g_err[0] = 0;
    if (!t || !offsets || !(radius > 0)) {
        set_error("invalid ball_fill arguments");
        return CKD_ERR_ARG;
    }
    if (nq <= 0) return CKD_OK;
    hipStream_t s = (hipStream_t)stream;
    int rc;
    {
        Scratch sc(s);
        rc = query_ball(t, queries, nq, radius, nullptr, offsets, dist, idx, sc, s);
    }
    if (rc || !sorted) return rc;
    return t->prec == 4 ? segmented_sort<float>((float *)dist, idx, offsets, nq, 0, s)
                        : segmented_sort<double>((double *)dist, idx, offsets, nq, 0, s);
}

int ckd_ball_search(const ckd_tree *t, const void *queries, int64_t nq, double radius, int32_t sorted,
                    int64_t *offsets, int64_t capacity, void *dist, int64_t *idx, int64_t *total, void *stream) {
    g_err[0] = 0;
    if (!t || !offsets || !total) {
        set_error("tree, offsets and total must not be NULL");
        return CKD_ERR_ARG;
    }
    if (!(radius > 0)) {
        set_error("Radius must be a positive number");
        return CKD_ERR_ARG;
    }
    if (nq < 0 || (nq > 0 && !queries) || capacity < 0 || (capacity > 0 && (!dist || !idx))) {
        set_error("invalid ball_search arguments");
        return CKD_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    if (nq == 0) return counts_to_offsets(offsets, 0, total, s);
    int64_t tot = 0;
    int32_t unwritten = 0;  // rows not (all) written by the one walk
    int rc;
    // the one walk needs the stack walk's conditions (wide records, depth
    // <= 64, 32-bit rows, ndim <= 3); otherwise, or with CKD_BALL_KERNEL=lane
    // (one walk per query), count, then fill
    const char *kenv = getenv("CKD_BALL_KERNEL");
    const bool lane = kenv && kenv[0] == 'l';
    if (!lane && t->ndim <= 3 && t->wide && t->depth <= 64 && t->n < INT32_MAX && nq < INT32_MAX) {
        Scratch sc(s);
        rc = query_ball_search(t, queries, nq, radius, offsets, capacity, dist, idx, &tot, &unwritten, sc, s);
        if (rc) return rc;
    } else {
        {
            Scratch sc(s);
            rc = query_ball(t, queries, nq, radius, offsets + 1, nullptr, nullptr, nullptr, sc, s);
        }
        if (rc) return rc;
        if ((rc = counts_to_offsets(offsets, nq, &tot, s))) return rc;
        unwritten = 1;
    }
    *total = tot;
    if (tot > capacity) return CKD_OK;  // the caller fills buffers of tot rows (ckd_ball_fill)
    if (unwritten) {
        Scratch sc(s);
        if ((rc = query_ball(t, queries, nq, radius, nullptr, offsets, dist, idx, sc, s))) return rc;
    }
    if (!sorted || tot == 0) return CKD_OK;
    return t->prec == 4 ? segmented_sort<float>((float *)dist, idx, offsets, nq, 0, s)
                        : segmented_sort<double>((double *)dist, idx, offsets, nq, 0, s);
}

static int box_count(const ckd_tree *t, const void *boxes, int64_t nb, int64_t *offsets, int64_t *total, int wrap,
                     void *stream) {
    g_err[0] = 0;
    if (!t || !offsets || nb < 0 || (nb > 0 && !boxes)) {
        set_error("invalid box_count arguments");
        return CKD_ERR_ARG;
    }
    if (wrap && !t->periodic) {
        set_error("boxes taken modulo the box need a periodic tree");
        return CKD_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    if (nb > 0) {
        int rc = query_box(t, boxes, nb, offsets + 1, nullptr, nullptr, wrap, s);
        if (rc) return rc;
    }
    return counts_to_offsets(offsets, nb, total, s);
}

static int box_fill(const ckd_tree *t, const void *boxes, int64_t nb, const int64_t *offsets, int64_t *idx, int wrap,
                    void *stream) {
    g_err[0] = 0;
    if (!t || !offsets) {
        set_error("invalid box_fill arguments");
        return CKD_ERR_ARG;
    }
    if (wrap && !t->periodic) {
        set_error("boxes taken modulo the box need a periodic tree");
        return CKD_ERR_ARG;
    }
    if (nb <= 0) return CKD_OK;
    return query_box(t, boxes, nb, nullptr, offsets, idx, wrap, (hipStream_t)stream);
}

int ckd_box_count(const ckd_tree *t, const void *boxes, int64_t nb, int64_t *offsets, int64_t *total,
                  void *stream) {
    return box_count(t, boxes, nb, offsets, total, 0, stream);
}

int ckd_box_fill(const ckd_tree *t, const void *boxes, int64_t nb, const int64_t *offsets, int64_t *idx,
                 void *stream) {
    return box_fill(t, boxes, nb, offsets, idx, 0, stream);
}

int ckd_box_count_periodic(const ckd_tree *t, const void *boxes, int64_t nb, int64_t *offsets, int64_t *total,
                           void *stream) {
    return box_count(t, boxes, nb, offsets, total, 1, stream);
}

int ckd_box_fill_periodic(const ckd_tree *t, const void *boxes, int64_t nb, const int64_t *offsets, int64_t *idx,
                          void *stream) {
    return box_fill(t, boxes, nb, offsets, idx, 1, stream);
}

int ckd_tree_dump(const ckd_tree *t, int32_t *axis, void *split, void *lo, void *hi, int32_t *child, int32_t *parent,
                  int64_t *start, int64_t *count, void *points, int64_t *idx) {
    g_err[0] = 0;
    if (!t) {
        set_error("tree must not be NULL");
        return CKD_ERR_ARG;
    }
    const int64_t nn = t->nnodes, n = t->n;
    const int D = t->ndim, es = t->prec;
    const size_t nrec = t->prec == 4 ? sizeof(NodeRec<float>) : sizeof(NodeRec<double>);
    char *hn = (char *)malloc(nrec * nn);
    char *hb = (char *)malloc((size_t)es * 2 * D * nn);
    Range *hr = (Range *)malloc(sizeof(Range) * nn);
    CKD_HIP(hipMemcpy(hn, t->nodes, nrec * nn, hipMemcpyDeviceToHost));
    CKD_HIP(hipMemcpy(hb, t->bbox, (size_t)es * 2 * D * nn, hipMemcpyDeviceToHost));
    CKD_HIP(hipMemcpy(hr, t->range, sizeof(Range) * nn, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < nn; i++) {
        const char *rec = hn + nrec * i;
        const int32_t *ints = (const int32_t *)(rec + es);
        if (axis) axis[i] = ints[0];
        if (child) child[i] = ints[1];
        if (parent) parent[i] = ints[2];
        if (split) memcpy((char *)split + (size_t)es * i, rec, es);
        if (lo) memcpy((char *)lo + (size_t)es * D * i, hb + (size_t)es * 2 * D * i, (size_t)es * D);
        if (hi) memcpy((char *)hi + (size_t)es * D * i, hb + (size_t)es * (2 * D * i + D), (size_t)es * D);
        if (start) start[i] = hr[i].start;
        if (count) count[i] = hr[i].count;
    }
    free(hn);
    free(hb);
    free(hr);
    if (points) {
        const int PS = pstride(D);
        char *hp = (char *)malloc((size_t)es * PS * n);
        CKD_HIP(hipMemcpy(hp, t->pts, (size_t)es * PS * n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; i++) memcpy((char *)points + (size_t)es * D * i, hp + (size_t)es * PS * i, (size_t)es * D);
        free(hp);
    }
    if (idx) {
        if (t->idx_bits == 64) {
            CKD_HIP(hipMemcpy(idx, t->idx, sizeof(int64_t) * n, hipMemcpyDeviceToHost));
        } else {
            int32_t *h = (int32_t *)malloc(sizeof(int32_t) * n);
            CKD_HIP(hipMemcpy(h, t->idx, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; i++) idx[i] = h[i];
            free(h);
        }
    }
    return CKD_OK;
}

}  // extern "C"
