// ckd_batch_f32_per.hip -- the batched engine's kernels for float periodic forests.
#include "ckd_batch_impl.h"

namespace ckd {

int batch_run_f32_per(ckd_batch *b, const void *q, void *kd, int64_t *ki, int64_t *bo, void *bd, int64_t *bi,
                      hipStream_t s) {
    return run_nt<float, true>(b, q, kd, ki, bo, bd, bi, s);
}

}  // namespace ckd
