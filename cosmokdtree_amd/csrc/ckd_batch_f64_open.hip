// ckd_batch_f64_open.hip -- the batched engine's kernels for double non-periodic forests.
#include "ckd_batch_impl.h"

namespace ckd {

int batch_run_f64_open(ckd_batch *b, const void *q, void *kd, int64_t *ki, int64_t *bo, void *bd, int64_t *bi,
                      hipStream_t s) {
    return run_nt<double, false>(b, q, kd, ki, bo, bd, bi, s);
}

}  // namespace ckd
