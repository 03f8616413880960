// ckd_sort.hip -- segmented sort of result rows by (distance, index).
//
// Replaces the per-query quicksort of the reference (quicksort /
// partition2, cosmokdtree.f90:1553-1643), which orders by distance only and
// is unstable; ordering equal distances by index makes the output
// deterministic while listing the same (distance, index) pairs.
//
// One workgroup per row.  Rows of up to LDS_MAX elements are sorted in LDS
// by a bitonic network.  Longer rows are cut into LDS-sorted runs which are
// then merged pairwise in global memory: every element finds its rank in
// the partner run by binary search (pairs are unique within a row because
// the indices are), so each merge pass is one scatter.
#include "ckd_internal.h"

namespace ckd {

namespace {

constexpr int ST = 256;
constexpr int LDS_MAX = 2048;

template <typename T>
__device__ __forceinline__ bool pair_less(T ka, int64_t va, T kb, int64_t vb) {
    return ka < kb || (ka == kb && va < vb);
}

template <typename T>
__device__ void lds_bitonic(T *k, int64_t *v, int n2) {
    for (int size = 2; size <= n2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            for (int i = threadIdx.x; i < n2 / 2; i += ST) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool up = (lo & size) == 0;
                const bool gt = pair_less(k[hi], v[hi], k[lo], v[lo]);
                if (gt == up) {
                    T tk = k[lo];
                    k[lo] = k[hi];
                    k[hi] = tk;
                    int64_t tv = v[lo];
                    v[lo] = v[hi];
                    v[hi] = tv;
                }
            }
        }
    }
    __syncthreads();
}

// Sort runs of LDS_MAX elements of every row in place.
template <typename T>
__global__ __launch_bounds__(ST) void k_sort_runs(T *keys, int64_t *vals, const int64_t *off, int64_t width) {
    __shared__ T lk[LDS_MAX];
    __shared__ int64_t lv[LDS_MAX];
    const int64_t row = blockIdx.x;
    const int64_t b = off ? off[row] : row * width;
    const int64_t e = off ? off[row + 1] : (row + 1) * width;
    const int64_t n = e - b;
    if (n <= 1) return;
    for (int64_t r0 = 0; r0 < n; r0 += LDS_MAX) {
        const int m = (int)(n - r0 < LDS_MAX ? n - r0 : LDS_MAX);
        int n2 = 1;
        while (n2 < m) n2 <<= 1;
        __syncthreads();
        for (int i = threadIdx.x; i < n2; i += ST) {
            if (i < m) {
                lk[i] = keys[b + r0 + i];
                lv[i] = vals[b + r0 + i];
            } else {
                lk[i] = T(INFINITY);
                lv[i] = INT64_MAX;
            }
        }
        lds_bitonic<T>(lk, lv, n2);
        for (int i = threadIdx.x; i < m; i += ST) {
            keys[b + r0 + i] = lk[i];
            vals[b + r0 + i] = lv[i];
        }
    }
}

// One merge pass: runs of `run` elements merged pairwise from src to dst.
template <typename T>
__global__ __launch_bounds__(ST) void k_merge_pass(const T *sk, const int64_t *sv, T *dk, int64_t *dv,
                                                   const int64_t *off, int64_t width, int64_t run) {
    const int64_t row = blockIdx.x;
    const int64_t b = off ? off[row] : row * width;
    const int64_t e = off ? off[row + 1] : (row + 1) * width;
    const int64_t n = e - b;
    if (n <= run) {  // nothing to merge in this row: keep both buffers in step
        for (int64_t i = threadIdx.x; i < n; i += ST) {
            dk[b + i] = sk[b + i];
            dv[b + i] = sv[b + i];
        }
        return;
    }
    for (int64_t i = threadIdx.x; i < n; i += ST) {
        const int64_t pair0 = (i / (2 * run)) * (2 * run);
        const int64_t a0 = pair0, a1 = min(pair0 + run, n);
        const int64_t b0 = a1, b1 = min(pair0 + 2 * run, n);
        const T key = sk[b + i];
        const int64_t val = sv[b + i];
        int64_t pos;
        if (i < a1) {  // rank among the second run: elements less than (key,val)
            int64_t lo = b0, hi = b1;
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (pair_less(sk[b + mid], sv[b + mid], key, val)) lo = mid + 1;
                else hi = mid;
            }
            pos = pair0 + (i - a0) + (lo - b0);
        } else {
            int64_t lo = a0, hi = a1;
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (pair_less(sk[b + mid], sv[b + mid], key, val)) lo = mid + 1;
                else hi = mid;
            }
            pos = pair0 + (i - b0) + (lo - a0);
        }
        dk[b + pos] = key;
        dv[b + pos] = val;
    }
}

template <typename T>
__global__ void k_max_row(const int64_t *off, int64_t nrows, unsigned long long *mx) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    atomicMax(mx, (unsigned long long)(off[r + 1] - off[r]));
}

}  // namespace

template <typename T>
int segmented_sort(T *keys, int64_t *vals, const int64_t *off, int64_t nrows, int64_t width, hipStream_t s) {
    if (nrows <= 0) return CKD_OK;
    k_sort_runs<T><<<(unsigned)nrows, ST, 0, s>>>(keys, vals, off, width);
    CKD_HIP(hipGetLastError());
    int64_t maxlen = width;
    if (off) {
        unsigned long long *dmx = nullptr, hmx = 0;
        int rc = dev_alloc((void **)&dmx, sizeof(unsigned long long), s);
        if (rc) return rc;
        hipError_t e = hipMemsetAsync(dmx, 0, sizeof(unsigned long long), s);
        if (e == hipSuccess) {
            k_max_row<T><<<(unsigned)ceil_div(nrows, 256), 256, 0, s>>>(off, nrows, dmx);
            e = hipMemcpyAsync(&hmx, dmx, sizeof(hmx), hipMemcpyDeviceToHost, s);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        dev_free(dmx, s);
        if (e != hipSuccess) return hip_fail(e, "longest row", __FILE__, __LINE__);
        maxlen = (int64_t)hmx;
    }
    if (maxlen <= LDS_MAX) return CKD_OK;
    // merge passes need a second buffer the size of the whole output
    int64_t total = width * nrows;
    if (off) CKD_HIP(hipMemcpy(&total, off + nrows, sizeof(int64_t), hipMemcpyDeviceToHost));
    T *tk = nullptr;
    int64_t *tv = nullptr;
    int rc = dev_alloc((void **)&tk, sizeof(T) * total, s);
    if (!rc) rc = dev_alloc((void **)&tv, sizeof(int64_t) * total, s);
    if (rc) {
        dev_free(tk, s);
        return rc;
    }
    T *sk = keys, *dk = tk;
    int64_t *sv = vals, *dv = tv;
    hipError_t e = hipSuccess;
    for (int64_t run = LDS_MAX; run < maxlen && e == hipSuccess; run *= 2) {
        k_merge_pass<T><<<(unsigned)nrows, ST, 0, s>>>(sk, sv, dk, dv, off, width, run);
        e = hipGetLastError();
        std::swap(sk, dk);
        std::swap(sv, dv);
    }
    if (e == hipSuccess && sk != keys) {  // rows longer than LDS_MAX ended in the temporary buffer
        e = hipMemcpyAsync(keys, sk, sizeof(T) * total, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(vals, sv, sizeof(int64_t) * total, hipMemcpyDeviceToDevice, s);
    }
    dev_free(tk, s);  // after the merge passes and copies enqueued above
    dev_free(tv, s);
    if (e != hipSuccess) return hip_fail(e, "merge passes", __FILE__, __LINE__);
    return CKD_OK;
}

template int segmented_sort<float>(float *, int64_t *, const int64_t *, int64_t, int64_t, hipStream_t);
template int segmented_sort<double>(double *, int64_t *, const int64_t *, int64_t, int64_t, hipStream_t);

}  // namespace ckd
