// ckd_knn_f64_open.hip -- double kNN kernels of non-periodic trees.
#include "ckd_query_impl.h"

namespace ckd {

int query_knn_f64_open(const ckd_tree *t, const void *q, int64_t nq, int k, int sorted, const float *seed,
                      const float *params, float margin, int G, const uint16_t *table, void *dist, int64_t *idx,
                      int64_t *stats, Scratch &sc, hipStream_t s) {
    return knn_per<double, false>(t, q, nq, k, sorted, seed, params, margin, G, table, dist, idx, stats, sc, s);
}

}  // namespace ckd
