// ckd_query_f64.hip -- the double instantiation of the ball, box and SPH kernels, and the
// kNN dispatch to ckd_knn_f64_per.hip / ckd_knn_f64_open.hip.
#include "ckd_query_impl.h"

namespace ckd {

int query_knn_f64_per(const ckd_tree *, const void *, int64_t, int, int, const float *, const float *, float,
                      int, const uint16_t *, void *, int64_t *, int64_t *, Scratch &, hipStream_t);
int query_knn_f64_open(const ckd_tree *, const void *, int64_t, int, int, const float *, const float *, float,
                       int, const uint16_t *, void *, int64_t *, int64_t *, Scratch &, hipStream_t);

int query_knn_f64(const ckd_tree *t, const void *q, int64_t nq, int k, int sorted, const float *seed, void *dist,
                  int64_t *idx, int64_t *stats, Scratch &sc, hipStream_t s) {
    return t->periodic
               ? query_knn_f64_per(t, q, nq, k, sorted, seed, nullptr, 0.0f, 0, nullptr, dist, idx, stats, sc, s)
               : query_knn_f64_open(t, q, nq, k, sorted, seed, nullptr, 0.0f, 0, nullptr, dist, idx, stats, sc, s);
}

int query_knn_learned_f64(const ckd_tree *t, const void *q, int64_t nq, int k, int sorted, const float *params,
                          float margin, int G, const uint16_t *table, void *dist, int64_t *idx, int64_t *stats,
                          Scratch &sc, hipStream_t s) {
    return t->periodic
               ? query_knn_f64_per(t, q, nq, k, sorted, nullptr, params, margin, G, table, dist, idx, stats, sc, s)
               : query_knn_f64_open(t, q, nq, k, sorted, nullptr, params, margin, G, table, dist, idx, stats, sc, s);
}

int query_ball_f64(const ckd_tree *t, const void *q, int64_t nq, double r, int64_t *counts, const int64_t *off,
                   void *dist, int64_t *idx, Scratch &sc, hipStream_t s) {
    return ball_t<double>(t, q, nq, r, counts, off, dist, idx, sc, s);
}

int query_ball_search_f64(const ckd_tree *t, const void *q, int64_t nq, double r, int64_t *off, int64_t cap,
                          void *dist, int64_t *idx, int64_t *total, int32_t *overflow, Scratch &sc, hipStream_t s) {
    return ball_search_t<double>(t, q, nq, r, off, cap, dist, idx, total, overflow, sc, s);
}

int query_sph_f64(const ckd_tree *t, const void *q, int64_t nq, int k, const void *mass, void *h, void *rho,
                   Scratch &sc, hipStream_t s) {
    return t->periodic ? sph_launch<double, true>(t, (const double *)q, nq, k, (const double *)mass, (double *)h, (double *)rho, sc, s)
                       : sph_launch<double, false>(t, (const double *)q, nq, k, (const double *)mass, (double *)h, (double *)rho, sc, s);
}

int query_box_f64(const ckd_tree *t, const void *b, int64_t nb, int64_t *counts, const int64_t *off, int64_t *idx,
                  int wrap, hipStream_t s) {
    return box_t<double>(t, b, nb, counts, off, idx, wrap, s);
}

}  // namespace ckd
