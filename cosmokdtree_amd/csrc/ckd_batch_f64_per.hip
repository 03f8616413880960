// ckd_batch_f64_per.hip -- the batched engine's kernels for double periodic forests.
#include "ckd_batch_impl.h"

namespace ckd {

int batch_run_f64_per(ckd_batch *b, const void *q, void *kd, int64_t *ki, int64_t *bo, void *bd, int64_t *bi,
                      hipStream_t s) {
    return run_nt<double, true>(b, q, kd, ki, bo, bd, bi, s);
}

}  // namespace ckd
