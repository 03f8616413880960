// ckd_scan.hip -- device exclusive scan of int64 values (CSR offsets,
// chunk bases, node compaction).  Reduce-then-scan over tiles of 2048
// values (256 threads x 8), recursive on the tile sums.
#include "ckd_internal.h"

namespace ckd {

namespace {
constexpr int SCAN_NT = 256;
constexpr int SCAN_IPT = 8;
constexpr int64_t SCAN_TILE = SCAN_NT * SCAN_IPT;

__global__ __launch_bounds__(SCAN_NT) void k_tile_sums(const int64_t *in, int64_t n, int64_t *sums) {
    __shared__ int64_t lds[SCAN_NT / WAVE + 1];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_IPT;
    int64_t s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_IPT; j++)
        if (base + j < n) s += in[base + j];
    int64_t tot;
    block_exclusive_scan<int64_t, SCAN_NT>(s, lds, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_NT) void k_tile_scan(const int64_t *in, int64_t *out, int64_t n,
                                                        const int64_t *tile_off, int with_total) {
    __shared__ int64_t lds[SCAN_NT / WAVE + 1];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_IPT;
    int64_t v[SCAN_IPT];
    int64_t s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_IPT; j++) {
        v[j] = base + j < n ? in[base + j] : 0;
        s += v[j];
    }
    int64_t tot;
    int64_t run = block_exclusive_scan<int64_t, SCAN_NT>(s, lds, &tot);
    const int64_t off = tile_off ? tile_off[blockIdx.x] : 0;
    run += off;
#pragma unroll
    for (int j = 0; j < SCAN_IPT; j++) {
        if (base + j < n) out[base + j] = run;
        run += v[j];
    }
    if (with_total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = off + tot;
}
}  // namespace

size_t scan_tmp_bytes(int64_t n) {
    size_t b = 0;
    int64_t nb = ceil_div(n, SCAN_TILE);
    while (nb > 1) {
        b += (size_t)(nb + 1) * 2 * sizeof(int64_t);
        nb = ceil_div(nb, SCAN_TILE);
    }
    return b + 256;
}

int scan_i64(const int64_t *in, int64_t *out, int64_t n, bool with_total, void *tmp, hipStream_t s) {
    if (n <= 0) {
        if (with_total) {
            k_tile_scan<<<1, SCAN_NT, 0, s>>>(in, out, 0, nullptr, 1);
            CKD_HIP(hipGetLastError());
        }
        return CKD_OK;
    }
    const int64_t nb = ceil_div(n, SCAN_TILE);
    if (nb == 1) {
        k_tile_scan<<<1, SCAN_NT, 0, s>>>(in, out, n, nullptr, with_total ? 1 : 0);
        CKD_HIP(hipGetLastError());
        return CKD_OK;
    }
    int64_t *sums = (int64_t *)tmp;
    int64_t *sums_off = sums + nb + 1;
    void *rest = sums_off + nb + 1;
    k_tile_sums<<<(unsigned)nb, SCAN_NT, 0, s>>>(in, n, sums);
    CKD_HIP(hipGetLastError());
    int rc = scan_i64(sums, sums_off, nb, false, rest, s);
    if (rc) return rc;
    k_tile_scan<<<(unsigned)nb, SCAN_NT, 0, s>>>(in, out, n, sums_off, with_total ? 1 : 0);
    CKD_HIP(hipGetLastError());
    return CKD_OK;
}

}  // namespace ckd
