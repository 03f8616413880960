// ckd_internal.h -- shared definitions of the gfx950 k-d tree library.
//
// Layout of a built tree in HBM (all arrays owned by ckd_tree):
//   nodes[nnodes]   NodeRec<T> {split, axis, child, parent}: axis < 0 marks a
//                   leaf; the children of node i are child[i] and child[i]+1
//                   (allocated as a pair), parent[0] = -1.
//   bbox[nnodes]    2*D values of T: the node's cell, lo[0..D) then hi[0..D)
//                   (KDTreeNode%minbounds/maxbounds, cosmokdtree.f90:99-100).
//   dbox[nnodes]    2*D values of T: the tight bounding box of the node's
//                   points (inside the cell; what the traversals prune with).
//   range[nnodes]   {start, count} int64: the node's points are rows
//                   [start, start+count) of the leaf-ordered arrays below,
//                   for inner nodes too (a subtree is one contiguous range).
//   pts[n]          leaf-ordered points, AoS with row stride pstride(D)
//                   elements (D=3 -> 4, so one 16/32-byte load per point),
//                   followed by PTS_PAD zero rows.
//   idx[n]          original 0-based index of each row, int32 or int64.
//   wide[nwide]     one record per inner node, in depth-first order, for the
//                   stack traversals: the tight boxes of both children
//                   (lo0, hi0, lo1, hi1; 4*D values of T) followed by two
//                   int32 pairs (a, b) saying where each child lives:
//                   b < 0 -> inner node, a = its wide id; b >= 0 -> leaf
//                   holding rows [a, a + b).  One load decides both children
//                   and hands over a leaf's rows without touching nodes[].
//   wroot[nsets]    the (a, b) pair of each set's root, same convention.
//   wrange[nwide]   (first row, count) of each inner node, by wide id (the
//                   ball walk takes a whole subtree without descending).
//   wparent[nwide]  wide id of each inner node's parent, -1 for set roots
//                   (the per-query ball walk of the batched engine climbs
//                   by it); derived from the wide records at build and load.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cfloat>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/cosmokdtree_amd.h"

#pragma clang fp contract(off)

namespace ckd {

template <typename T>
struct alignas(sizeof(T) == 4 ? 16 : 8) NodeRec {
    T split;
    int32_t axis;    // split axis, -1 for a leaf
    int32_t child;   // first child id (the second is child+1), -1 for a leaf
    int32_t parent;  // -1 for the root
};

struct Range {
    int64_t start;
    int64_t count;
};

__host__ __device__ constexpr int pstride(int D) { return D == 3 ? 4 : (D <= 4 ? D : 8); }

// Spare rows after the n leaf-ordered points (zero-filled), so that batched
// row loads may run past the last leaf.
constexpr int PTS_PAD = 8;

// Bytes of one wide record (a multiple of 16 for 4- and 8-byte T).
template <typename T, int D>
__host__ __device__ constexpr size_t wide_bytes() { return 4 * D * sizeof(T) + 4 * sizeof(int32_t); }
__host__ __device__ constexpr size_t wide_bytes_rt(int D, int tsize) { return 4 * (size_t)D * tsize + 16; }

// Device view of a tree, passed by value to kernels.
template <typename T>
struct TreeView {
    const NodeRec<T> *nodes;
    const T *bbox;
    const T *dbox;
    const Range *range;
    const T *pts;
    const void *idx;
    const char *wide;     // wide records (nullptr when not built)
    const int2 *wroot;    // root pair of each set
    const int2 *wrange;   // rows of each inner node, by wide id
    const int32_t *wparent;  // parent of each inner node, by wide id (-1 for roots)
    int idx64;
    T L[CKD_MAX_DIM];  // box size per axis (periodic only)
    int64_t nrows;     // points (rows of pts and idx; pts has PTS_PAD more)
    int64_t nnodes, nwide;
};

// Bounds checks of the debug build (python -m cosmokdtree_amd.build --debug
// compiles with -DCKD_DEBUG): a failed check prints where and stops the
// kernel.  They cost nothing in the normal build.
#ifdef CKD_DEBUG
#define CKD_DASSERT(c)                                                               \
    do {                                                                             \
        if (!(c)) {                                                                  \
            printf("CKD_DEBUG check failed: %s at %s:%d\n", #c, __FILE__, __LINE__);  \
            __builtin_trap();                                                        \
        }                                                                            \
    } while (0)
#else
#define CKD_DASSERT(c) ((void)0)
#endif

// ---------------------------------------------------------------- errors
void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);
#define CKD_HIP(call)                                                  \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (e_ != hipSuccess) return hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

// ------------------------------------------------------ device memory
// Device memory from the library's caching allocator on the current device
// (freed blocks stay mapped for the next allocation; ckd_release_memory
// returns them to the driver).  dev_alloc's block is ready for work
// enqueued on `s` after the call.  dev_free hands a block back once the
// work enqueued on `s` before the free has run; dev_free_idle hands back a
// block that no pending work uses (e.g. after a device synchronisation).
int dev_alloc(void **p, size_t bytes, hipStream_t s);
void dev_free(void *p, hipStream_t s);
void dev_free_idle(void *p);

// Device scratch of one library call, released when the call returns.  By
// default each take() is a block of the caching allocator, handed back
// stream-ordered at the end of the call, so calls that overlap on one tree
// (other streams, other host threads) never share scratch.  Given an arena
// (the entry points that allocate nothing), take() carves 256-byte aligned
// pieces out of it instead and fails with CKD_ERR_LIMIT past its end.
struct Scratch {
    hipStream_t s;
    char *arena = nullptr;
    size_t cap = 0, used = 0;
    std::vector<void *> owned;
    explicit Scratch(hipStream_t s_) : s(s_) {}
    Scratch(hipStream_t s_, void *a, size_t c) : s(s_), arena((char *)a), cap(c) {}
    Scratch(const Scratch &) = delete;
    Scratch &operator=(const Scratch &) = delete;
    ~Scratch() {
        for (void *b : owned) dev_free(b, s);
    }
    int take(size_t bytes, void **p);
};

// ---------------------------------------------------- device arithmetic
// |x|, min, max as single instructions (source modifiers, v_min/v_max);
// the operands are never NaN, and each returns one of its inputs exactly.
__device__ __forceinline__ float absv(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double absv(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float minv(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double minv(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ float maxv(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double maxv(double a, double b) { return __builtin_fmax(a, b); }
// median of three (lo <= hi): max(lo, min(x, hi)), one v_med3 in float
__device__ __forceinline__ float med3v(float lo, float x, float hi) { return __builtin_amdgcn_fmed3f(lo, x, hi); }
__device__ __forceinline__ double med3v(double lo, double x, double hi) { return maxv(lo, minv(x, hi)); }

// Squared distance exactly as distance() of the reference
// (cosmokdtree.f90:1441-1465): dx = p - q per axis, minimum image
// min(|dx|, L - |dx|) when periodic, then ((0 + dx1*dx1) + dx2*dx2) + ...
// in axis order with one rounding per operation (no FMA contraction).
template <typename T, int D, bool PER>
__device__ __forceinline__ T dist2(const T *p, const T *q, const T *L) {
    T s = T(0);
#pragma unroll
    for (int i = 0; i < D; i++) {
        T dx = p[i] - q[i];
        if (PER) {
            const T a = absv(dx);
            dx = minv(L[i] - a, a);
        }
        s = s + dx * dx;
    }
    return s;
}

// Lower bound of dist2 over every point of the box [lo, hi].  Per axis the
// gap is max(lo - q, q - hi, 0), built from the same rounded differences
// the point distance uses, so the bound never exceeds the distance computed
// for any point of the box (rounding is monotone).  Periodic: the gap is
// also capped by the wrapped gap L - (far face distance), where the far
// face distance is -min(lo - q, q - hi) (a rounded difference negates
// exactly).
template <typename T, int D, bool PER>
__device__ __forceinline__ T mindist2(const T *box, const T *q, const T *L) {
    T s = T(0);
#pragma unroll
    for (int i = 0; i < D; i++) {
        const T dl = box[i] - q[i], dh = q[i] - box[D + i];
        T d;
        if (PER) d = maxv(minv(maxv(dl, dh), L[i] + minv(dl, dh)), T(0));
        else d = maxv(maxv(dl, dh), T(0));
        s = s + d * d;
    }
    return s;
}

// Upper bound of dist2 over the cell: per axis max(|lo - q|, |hi - q|) with
// the same rounded differences, summed like dist2.  The minimum image never
// exceeds the direct difference, so the bound also holds when periodic.
template <typename T, int D>
__device__ __forceinline__ T maxdist2(const T *box, const T *q) {
    T s = T(0);
#pragma unroll
    for (int i = 0; i < D; i++) {
        T a = box[i] - q[i];
        T b = box[D + i] - q[i];
        a = a < T(0) ? -a : a;
        b = b < T(0) ? -b : b;
        T m = a > b ? a : b;
        s = s + m * m;
    }
    return s;
}

template <typename T>
__device__ __forceinline__ T sqrt_rn(T x);
template <>
__device__ __forceinline__ float sqrt_rn<float>(float x) { return __builtin_sqrtf(x); }
template <>
__device__ __forceinline__ double sqrt_rn<double>(double x) { return __builtin_sqrt(x); }

__device__ __forceinline__ int64_t load_idx(const void *idx, int idx64, int64_t i) {
    return idx64 ? ((const int64_t *)idx)[i] : (int64_t)((const int32_t *)idx)[i];
}

// ----------------------------------------------------- wave / block scans
constexpr int WAVE = 64;

// Wave votes on a bool: the mask of the active lanes where p holds, and
// whether there is one.  HIP's __ballot / __any take an int; the bool ->
// int -> bool round trip cost two vector instructions per vote.
__device__ __forceinline__ uint64_t wave_vote(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

// The same straight from one compare (its lane mask; inactive lanes 0), for
// votes combined with another lane mask by a scalar AND: a vote on
// `run && a < b` was materialised as a 0/1 value and compared again.
// Predicates as LLVM numbers them: 4 = ordered less than, 10 = unordered or
// greater than, 38 = signed greater than.
__device__ __forceinline__ uint64_t lanes_lt(float a, float b) { return __builtin_amdgcn_fcmpf(a, b, 4); }
__device__ __forceinline__ uint64_t lanes_lt(double a, double b) { return __builtin_amdgcn_fcmp(a, b, 4); }
__device__ __forceinline__ uint64_t lanes_ugt(float a, float b) { return __builtin_amdgcn_fcmpf(a, b, 10); }
__device__ __forceinline__ uint64_t lanes_ugt(double a, double b) { return __builtin_amdgcn_fcmp(a, b, 10); }
__device__ __forceinline__ uint64_t lanes_gt(int a, int b) { return __builtin_amdgcn_sicmp(a, b, 38); }

// Set bits of the lane mask m below this lane (v_mbcnt).
__device__ __forceinline__ int lane_rank(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// XCD-aware block order.  Workgroups are dealt to the 8 XCDs (each with its
// own L2) round-robin by block id.  This maps the blocks an XCD receives
// within each group of 8 * XCD_RUN consecutive ids to a run of XCD_RUN
// consecutive logical blocks, so neighbouring work (queries in leaf order)
// shares an L2, while runs still rotate over the XCDs (balanced finish).
// A bijection of [0, gridDim.x); a trailing partial group keeps its order.
constexpr unsigned NXCD = 8, XCD_RUN = 32;
__device__ __forceinline__ unsigned xcd_block() {
    const unsigned b = blockIdx.x, grp = NXCD * XCD_RUN;
    if (b >= gridDim.x / grp * grp) return b;
    const unsigned w = b % grp;
    return b - w + (w % NXCD) * XCD_RUN + w / NXCD;
}

template <typename V>
__device__ __forceinline__ V wave_inclusive_scan(V v) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        V t = __shfl_up(v, off, WAVE);
        if (lane >= off) v += t;
    }
    return v;
}

// Exclusive scan over a block of NT threads; returns the block total in
// *total.  `lds` must hold NT / WAVE values.
template <typename V, int NT>
__device__ __forceinline__ V block_exclusive_scan(V v, V *lds, V *total) {
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    V inc = wave_inclusive_scan(v);
    if (lane == WAVE - 1) lds[w] = inc;
    __syncthreads();
    if (w == 0) {
        V s = lane < NT / WAVE ? lds[lane] : V(0);
        V si = wave_inclusive_scan(s);
        if (lane < NT / WAVE) lds[lane] = si - s;
        if (lane == NT / WAVE - 1) lds[NT / WAVE] = si;
    }
    __syncthreads();
    V r = inc - v + lds[w];
    *total = lds[NT / WAVE];
    __syncthreads();
    return r;
}

// -------------------------------------------------------- host helpers
// Exclusive scan of n int64 values into out[0..n); out[n] receives the
// total when with_total.  Uses `tmp` scratch of scan_tmp_bytes(n).
size_t scan_tmp_bytes(int64_t n);
int scan_i64(const int64_t *in, int64_t *out, int64_t n, bool with_total, void *tmp, hipStream_t s);

// Segmented sort of rows by (key, idx) ascending; row r is
// [off[r], off[r+1]) of keys/vals.  off == NULL means fixed rows of
// `width` elements.
template <typename T>
int segmented_sort(T *keys, int64_t *vals, const int64_t *off, int64_t nrows, int64_t width, hipStream_t s);

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace ckd

// Host-side tree object behind the opaque ckd_tree handle.
struct ckd_tree {
    int64_t n = 0;
    int ndim = 0, prec = 8, leafsize = 16, periodic = 0, idx_bits = 32, depth = 0;
    int device = 0;
    int64_t nnodes = 0, nleaves = 0;
    double L[CKD_MAX_DIM] = {0};
    double build_ms = 0;
    void *nodes = nullptr;   // NodeRec<T>[nnodes]
    void *bbox = nullptr;    // T[nnodes * 2 * ndim] cells
    void *dbox = nullptr;    // T[nnodes * 2 * ndim] tight point boxes
    ckd::Range *range = nullptr;
    void *pts = nullptr;     // T[n * pstride(ndim)]
    void *idx = nullptr;     // int32 / int64 [n]
    int64_t nsets = 1;        // independent point sets (a forest); root of set r is node r
    int64_t *set_off = nullptr;  // device [nsets + 1]: set r holds input rows [set_off[r], set_off[r+1])
    void *wide = nullptr;    // wide records of the inner nodes
    int64_t nwide = 0;
    int2 *wroot = nullptr;   // device [nsets] root pairs
    int2 *wrange = nullptr;  // device [nwide] inner-node rows
    int32_t *wparent = nullptr;  // device [nwide] parent wide ids, -1 for roots
    int knn_kernel = -1;     // kNN kernel choice, -1 = automatic (ckd_knn_config)
    int knn_stack = 0;       // stack depth of the stack kernel, 0 = automatic
    int knn_qpw = 0;         // queries per wave of the stack kernel (1..64), 0 = 64
    int timing = 0;          // record events around kNN traversal launches (ckd_knn_timing)
    std::mutex timing_mu;    // guards knn_events: queries may run from several host threads
    std::vector<std::pair<hipEvent_t, hipEvent_t>> knn_events;
};

namespace ckd {
template <typename T>
TreeView<T> view(const ckd_tree *t) {
    TreeView<T> v;
    v.nodes = (const NodeRec<T> *)t->nodes;
    v.bbox = (const T *)t->bbox;
    v.dbox = (const T *)t->dbox;
    v.range = t->range;
    v.pts = (const T *)t->pts;
    v.idx = t->idx;
    v.wide = (const char *)t->wide;
    v.wroot = t->wroot;
    v.wrange = t->wrange;
    v.wparent = t->wparent;
    v.idx64 = t->idx_bits == 64;
    for (int i = 0; i < CKD_MAX_DIM; i++) v.L[i] = (T)t->L[i];
    v.nrows = t->n;
    v.nnodes = t->nnodes;
    v.nwide = t->wide ? t->nwide : 0;
    return v;
}
int build_tree(ckd_tree *t, const void *points, hipStream_t s);
// wparent from the wide records (after a build or a load of a tree with them).
int wide_parents(ckd_tree *t, hipStream_t s);
}  // namespace ckd
