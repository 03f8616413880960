// ckd_query.hip -- precision dispatch of the single-tree queries.
#include "ckd_internal.h"

namespace ckd {

int query_knn_f32(const ckd_tree *, const void *, int64_t, int, int, const float *, void *, int64_t *, int64_t *,
                  Scratch &, hipStream_t);
int query_knn_f64(const ckd_tree *, const void *, int64_t, int, int, const float *, void *, int64_t *, int64_t *,
                  Scratch &, hipStream_t);
int query_knn_learned_f32(const ckd_tree *, const void *, int64_t, int, int, const float *, float, int,
                          const uint16_t *, void *, int64_t *, int64_t *, Scratch &, hipStream_t);
int query_knn_learned_f64(const ckd_tree *, const void *, int64_t, int, int, const float *, float, int,
                          const uint16_t *, void *, int64_t *, int64_t *, Scratch &, hipStream_t);
int query_ball_f32(const ckd_tree *, const void *, int64_t, double, int64_t *, const int64_t *, void *, int64_t *,
                   Scratch &, hipStream_t);
int query_ball_f64(const ckd_tree *, const void *, int64_t, double, int64_t *, const int64_t *, void *, int64_t *,
                   Scratch &, hipStream_t);
int query_ball_search_f32(const ckd_tree *, const void *, int64_t, double, int64_t *, int64_t, void *, int64_t *,
                          int64_t *, int32_t *, Scratch &, hipStream_t);
int query_ball_search_f64(const ckd_tree *, const void *, int64_t, double, int64_t *, int64_t, void *, int64_t *,
                          int64_t *, int32_t *, Scratch &, hipStream_t);
int query_box_f32(const ckd_tree *, const void *, int64_t, int64_t *, const int64_t *, int64_t *, int, hipStream_t);
int query_sph_f32(const ckd_tree *, const void *, int64_t, int, const void *, void *, void *, Scratch &, hipStream_t);
int query_sph_f64(const ckd_tree *, const void *, int64_t, int, const void *, void *, void *, Scratch &, hipStream_t);
int query_box_f64(const ckd_tree *, const void *, int64_t, int64_t *, const int64_t *, int64_t *, int, hipStream_t);

int query_knn(const ckd_tree *t, const void *q, int64_t nq, int k, int sorted, const float *seed, void *dist,
              int64_t *idx, int64_t *stats, Scratch &sc, hipStream_t s) {
    return t->prec == 4 ? query_knn_f32(t, q, nq, k, sorted, seed, dist, idx, stats, sc, s)
                        : query_knn_f64(t, q, nq, k, sorted, seed, dist, idx, stats, sc, s);
}

int query_knn_learned(const ckd_tree *t, const void *q, int64_t nq, int k, int sorted, const float *params,
                      float margin, int G, const uint16_t *table, void *dist, int64_t *idx, int64_t *stats,
                      Scratch &sc, hipStream_t s) {
    return t->prec == 4 ? query_knn_learned_f32(t, q, nq, k, sorted, params, margin, G, table, dist, idx, stats, sc, s)
                        : query_knn_learned_f64(t, q, nq, k, sorted, params, margin, G, table, dist, idx, stats, sc, s);
}

int query_ball(const ckd_tree *t, const void *q, int64_t nq, double r, int64_t *counts, const int64_t *off,
               void *dist, int64_t *idx, Scratch &sc, hipStream_t s) {
    return t->prec == 4 ? query_ball_f32(t, q, nq, r, counts, off, dist, idx, sc, s)
                        : query_ball_f64(t, q, nq, r, counts, off, dist, idx, sc, s);
}

int query_ball_search(const ckd_tree *t, const void *q, int64_t nq, double r, int64_t *off, int64_t cap, void *dist,
                      int64_t *idx, int64_t *total, int32_t *overflow, Scratch &sc, hipStream_t s) {
    return t->prec == 4 ? query_ball_search_f32(t, q, nq, r, off, cap, dist, idx, total, overflow, sc, s)
                        : query_ball_search_f64(t, q, nq, r, off, cap, dist, idx, total, overflow, sc, s);
}

int query_sph(const ckd_tree *t, const void *q, int64_t nq, int k, const void *mass, void *h, void *rho,
              Scratch &sc, hipStream_t s) {
    return t->prec == 4 ? query_sph_f32(t, q, nq, k, mass, h, rho, sc, s)
                        : query_sph_f64(t, q, nq, k, mass, h, rho, sc, s);
}

int query_box(const ckd_tree *t, const void *b, int64_t nb, int64_t *counts, const int64_t *off, int64_t *idx,
              int wrap, hipStream_t s) {
    return t->prec == 4 ? query_box_f32(t, b, nb, counts, off, idx, wrap, s)
                        : query_box_f64(t, b, nb, counts, off, idx, wrap, s);
}

}  // namespace ckd
