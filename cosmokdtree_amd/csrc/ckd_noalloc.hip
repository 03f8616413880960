// ckd_noalloc.hip -- ckd_knn_noalloc: the kNN entry point for callers that
// cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  Its scratch (the leaf
// ordering of the queries, the re-search lists of seeded calls) is carved
// from one static device array per GPU instead of the caching allocator.
#include <mutex>

#include "ckd_query_impl.h"

namespace ckd {
int query_knn(const ckd_tree *t, const void *q, int64_t nq, int k, int sorted, const float *seed, void *dist,
              int64_t *idx, int64_t *stats, Scratch &sc, hipStream_t s);

// The static scratch, one copy per GPU, held until the process ends.  Host
// code never passes `scratch` itself to a kernel (on the host it names a
// stand-in, not device memory): it passes the device address that
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) char scratch[CKD_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;

std::mutex g_mu;
void *g_addr[MAX_DEVICES] = {};

// Device address of the current device's copy of `scratch`, looked up once
// per device (with that device current) on its first call.
int scratch_address(void **out) {
    int dev = 0;
    CKD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error("device %d is beyond the %d devices the library supports", dev, MAX_DEVICES);
        return CKD_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_addr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)", __FILE__,
                            __LINE__);
        }
        g_addr[dev] = p;
    }
    *out = g_addr[dev];
    return CKD_OK;
}

}  // namespace
}  // namespace ckd

using namespace ckd;

extern "C" {

size_t ckd_knn_noalloc_scratch_bytes(const ckd_tree *t, int64_t nq, int32_t seeded) {
    if (!t || nq <= 0) return 0;
    size_t need = 0;
    if (orders_queries(t, nq)) need += order_scratch_bytes(t, nq) + 256;
    if (seeded) need += retry_scratch_bytes(nq) + 256;
    return need;
}

int ckd_knn_noalloc(const ckd_tree *t, const void *queries, int64_t nq, int32_t k, int32_t sorted, void *dist,
                    int64_t *idx, const float *seed_r, int64_t *stats, void *stream) {
    if (!t || nq < 0 || k < 0 || (nq > 0 && (!queries || (k > 0 && (!dist || !idx))))) {
        set_error("invalid ckd_knn_noalloc arguments");
        return CKD_ERR_ARG;
    }
    if (nq == 0 || k == 0) return CKD_OK;
    if (k > 64 || t->n >= INT32_MAX) {
        set_error("ckd_knn_noalloc takes k <= 64 and trees below 2^31 - 1 points (use ckd_knn)");
        return CKD_ERR_LIMIT;
    }
    const size_t need = ckd_knn_noalloc_scratch_bytes(t, nq, seed_r != nullptr);
    if (need > (size_t)CKD_NOALLOC_SCRATCH_BYTES) {
        set_error("ckd_knn_noalloc: %lld queries on a tree of %lld nodes need %zu bytes of scratch, more than the "
                  "%d of the static scratch (use ckd_knn, or smaller batches)",
                  (long long)nq, (long long)t->nnodes, need, (int)CKD_NOALLOC_SCRATCH_BYTES);
        return CKD_ERR_LIMIT;
    }
    void *arena = nullptr;
    int rc = scratch_address(&arena);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    Scratch sc(s, arena, (size_t)CKD_NOALLOC_SCRATCH_BYTES);
    return query_knn(t, queries, nq, k, sorted, seed_r, dist, idx, stats, sc, s);
}

}  // extern "C"
