// ckd_traverse.h -- per-query traversals shared by the single-tree kernels
// (ckd_query_impl.h) and the batched persistent kernel (ckd_batch.hip).
//
// One thread per query.  Traversals are stack-free: a node's parent pointer
// replaces the recursion stack of the reference, so any tree depth
// (sliding-midpoint trees can be very deep on duplicates) is handled
// without a stack bound.
//
// kNN keeps the k best candidates in registers as an ascending list
// (distance, leaf row) of static length KMAX >= k: the k live entries are
// the top k slots, the slots below are -inf and inert, so the pruning bound
// is always slot KMAX-1.  A candidate is accepted only when strictly closer
// than that bound, as in cosmokdtree.f90:658,808; the list starts filled
// with FLT_MAX (HUGE(0.0), cosmokdtree.f90:517) or with the squared seed
// radius.  The visit order is depth-first, nearer child first (the order
// of knn_search_recursive_hyperp, cosmokdtree.f90:820-846), with every
// subtree pruned by the minimum distance to the tight box of its points
// (the reference prunes with the cell, bbox_mindist_kd,
// cosmokdtree.f90:1525-1546; the points lie in both, so the tighter box
// prunes more and drops nothing; minimum image when periodic).  Pruning never
// drops a candidate that could be accepted, so the accepted set equals the
// reference's up to the choice among points tied at the k-th distance.
#pragma once
#include "ckd_internal.h"

namespace ckd {

template <typename T, int D>
__device__ __forceinline__ void load_pt(const T *pts, int64_t p, T (&x)[D]) {
    constexpr int PS = pstride(D);
    const T *src = pts + p * PS;
    if constexpr (sizeof(T) == 4 && PS == 4) {
        float4 v = *(const float4 *)src;
        x[0] = v.x;
        if (D > 1) x[1] = v.y;
        if (D > 2) x[2] = v.z;
        if (D > 3) x[3] = v.w;
    } else {
#pragma unroll
        for (int d = 0; d < D; d++) x[d] = src[d];
    }
}

template <typename T, int D>
__device__ __forceinline__ void load_box(const T *bbox, int32_t node, T (&b)[2 * D]) {
    const T *src = bbox + (int64_t)node * 2 * D;
#pragma unroll
    for (int d = 0; d < 2 * D; d++) b[d] = src[d];
}

template <typename T, int D, bool PER>
__device__ __forceinline__ T node_mindist(const TreeView<T> &tv, int32_t node, const T (&q)[D]) {
    T b[2 * D];
    load_box<T, D>(tv.dbox, node, b);
    return mindist2<T, D, PER>(b, q, tv.L);
}

// Insert (d, s), d < bd[KMAX-1], into the ascending list; the last entry
// drops out and d goes after the entries equal to it.  Slot j becomes the
// median of (old j-1, d, old j); the rows follow the same choice.  The row
// of slot j is chosen in two steps (s or the old row, then the previous
// slot's row): written as one nested conditional, the compiler turned it
// into two divergent branches per slot inside the callers' `if`s, about
// three times the instructions of the two selects.
//
// The slots are updated in chunks of 8 from the top; a chunk whose entries
// are all at or below d for every lane taking part is left as it is, and so
// is everything under it (the update leaves slot j unchanged when
// d >= old j).  Candidates found late in a walk mostly enter near the end
// of the list, so many insertions stop after one or two chunks.
template <typename T, int KMAX>
__device__ __forceinline__ void list_insert(T (&bd)[KMAX], int32_t (&bs)[KMAX], T d, int32_t s) {
    constexpr int C = KMAX >= 16 ? 8 : KMAX;
#pragma unroll
    for (int c = KMAX / C - 1; c >= 0; --c) {
        // slots below c*C + C change only for lanes with d < old bd[c*C + C - 1]
        if (c < KMAX / C - 1 && !lanes_lt(d, bd[c * C + C - 1])) break;
#pragma unroll
        for (int j = c * C + C - 1; j >= (c == 0 ? 1 : c * C); --j) {
            const bool below = d < bd[j - 1];
            const bool here = d < bd[j];
            const int32_t keep = here ? s : bs[j];
            bs[j] = below ? bs[j - 1] : keep;
            bd[j] = med3v(bd[j - 1], d, bd[j]);
        }
        if (c == 0) {
            bs[0] = d < bd[0] ? s : bs[0];
            bd[0] = minv(bd[0], d);
        }
    }
}

// list_insert without the rows: slot j becomes med3(old j-1, d, old j), one
// instruction per slot (callers that need only the distances).
template <typename T, int KMAX>
__device__ __forceinline__ void list_insert_d(T (&bd)[KMAX], T d) {
#pragma unroll
    for (int j = KMAX - 1; j > 0; --j) bd[j] = med3v(bd[j - 1], d, bd[j]);
    bd[0] = minv(bd[0], d);
}

// Insert into the list, with its rows unless NOIDX.
template <bool NOIDX, typename T, int KMAX>
__device__ __forceinline__ void list_put(T (&bd)[KMAX], int32_t (&bs)[KMAX], T d, int32_t s) {
    if constexpr (NOIDX) list_insert_d<T, KMAX>(bd, d);
    else list_insert<T, KMAX>(bd, bs, d, s);
}

// The leaf a point falls in below `root`: the nearer child at every level
// (the first descent of knn_search_recursive_hyperp, cosmokdtree.f90:820-836).
template <typename T, int D>
__device__ __forceinline__ int32_t find_leaf(const TreeView<T> &tv, const T (&q)[D], int32_t root) {
    int32_t node = root;
    for (;;) {
        const NodeRec<T> nr = tv.nodes[node];
        if (nr.axis < 0) return node;
        node = nr.child + (q[nr.axis] < nr.split ? 0 : 1);
    }
}

constexpr int CBUF = 16;  // leaf candidates buffered per lane between list updates

struct Counters {  // per-query work counts (kernel statistics)
    int64_t visited = 0, tested = 0, reruns = 0, inserted = 0, climbed = 0;
};

template <typename T, int D, bool PER, int KMAX>
__device__ __forceinline__ void knn_scan_leaf(const TreeView<T> &tv, int32_t leaf, const T (&q)[D], T (&bd)[KMAX],
                                              int32_t (&bs)[KMAX], Counters &c) {
    const Range r = tv.range[leaf];
    for (int64_t p = r.start; p < r.start + r.count; p++) {
        T x[D];
        load_pt<T, D>(tv.pts, p, x);
        const T d = dist2<T, D, PER>(x, q, tv.L);
        if (d < bd[KMAX - 1]) {
            list_insert<T, KMAX>(bd, bs, d, (int32_t)p);
            c.inserted++;
        }
    }
    c.tested += r.count;
}

// k nearest neighbours of q, starting at its leaf.  Writes k (distance,
// index - idx_base) pairs ascending to drow/irow; with `sorted`, equal
// distances are put in index order.  seed2 < FLT_MAX seeds the bound; a
// search that finds fewer than k points inside it runs again unseeded.
template <typename T, int D, bool PER, int KMAX>
__device__ __forceinline__ void knn_one(const TreeView<T> &tv, const T (&q)[D], int32_t leaf, int k, int sorted,
                                        T seed2, T *drow, int64_t *irow, int64_t idx_base, Counters &c) {
    T bd[KMAX];
    int32_t bs[KMAX];
    T init = seed2 < T(FLT_MAX) ? seed2 : T(FLT_MAX);
    for (int pass = 0; pass < 2; pass++) {
#pragma unroll
        for (int j = 0; j < KMAX; j++) {
            bd[j] = j < KMAX - k ? T(-INFINITY) : init;
            bs[j] = -1;
        }
        knn_scan_leaf<T, D, PER, KMAX>(tv, leaf, q, bd, bs, c);
        c.visited++;
        int32_t from = leaf, cur = tv.nodes[leaf].parent;
        while (cur >= 0) {
            c.climbed++;
            const NodeRec<T> nr = tv.nodes[cur];
            const int32_t nearc = nr.child + (q[nr.axis] < nr.split ? 0 : 1);
            bool moved = false;
            if (from == nearc) {
                const int32_t farc = nearc == nr.child ? nr.child + 1 : nr.child;
                if (node_mindist<T, D, PER>(tv, farc, q) < bd[KMAX - 1]) {
                    int32_t n2 = farc;
                    for (;;) {
                        c.visited++;
                        const NodeRec<T> r2 = tv.nodes[n2];
                        if (r2.axis < 0) {
                            knn_scan_leaf<T, D, PER, KMAX>(tv, n2, q, bd, bs, c);
                            from = n2;
                            cur = r2.parent;
                            break;
                        }
                        const int32_t nc2 = r2.child + (q[r2.axis] < r2.split ? 0 : 1);
                        if (node_mindist<T, D, PER>(tv, nc2, q) < bd[KMAX - 1]) {
                            n2 = nc2;
                            continue;
                        }
                        from = nc2;  // nearer child pruned: continue as if back from it
                        cur = n2;
                        break;
                    }
                    moved = true;
                }
            }
            if (!moved) {
                from = cur;
                cur = nr.parent;
            }
        }
        if (bs[KMAX - 1] >= 0 || init == T(FLT_MAX)) break;
        init = T(FLT_MAX);  // fewer than k points inside the seed radius
        c.reruns++;
    }
    bool fix = false;
    T pd = T(-1);
    int64_t pi = 0;
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
        if (j >= KMAX - k) {
            const int32_t s = bs[j];
            CKD_DASSERT(s >= -1 && s < tv.nrows);
            const int64_t oi = s >= 0 ? load_idx(tv.idx, tv.idx64, s) - idx_base : (int64_t)CKD_EMPTY_SLOT;
            const T dv = sqrt_rn<T>(bd[j]);
            fix |= (dv == pd && oi < pi);
            pd = dv;
            pi = oi;
            drow[j - (KMAX - k)] = dv;
            irow[j - (KMAX - k)] = oi;
        }
    }
    if (sorted && fix) {  // order equal distances by index (insertion sort, rare)
        for (int a = 1; a < k; a++) {
            const T dv = drow[a];
            const int64_t iv = irow[a];
            int b = a - 1;
            while (b >= 0 && drow[b] == dv && irow[b] > iv) {
                drow[b + 1] = drow[b];
                irow[b + 1] = irow[b];
                b--;
            }
            drow[b + 1] = dv;
            irow[b + 1] = iv;
        }
    }
}

// Wave-uniform kNN: the 64 queries of a wave share one depth-first
// traversal of the tree below `root`.  Node records, boxes and leaf points
// are read once per wave (uniform addresses); a child is visited when any
// lane's own bound still reaches its point box, children in near-first
// order for the mean of the wave's queries; at a leaf every lane tests every
// point against its own query with its own strict bound.  Every lane visits
// a superset of the nodes its own search would need, in a fixed order, so
// each lane's accepted set is exactly the k nearest (ties at the k-th
// distance aside).  All lanes of the wave must call it; lanes without a
// query pass active = false.  One pass: a seeded lane that finds fewer than
// k points inside its seed writes nothing and returns false, so the caller
// can search it again unseeded (seeded waves never pay for a second pass).
template <typename T>
__device__ __forceinline__ T wave_uniform(T v) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int32_t, v)));
    } else {
        const int64_t b = __builtin_bit_cast(int64_t, v);
        const int32_t lo = __builtin_amdgcn_readfirstlane((int32_t)(b & 0xffffffff));
        const int32_t hi = __builtin_amdgcn_readfirstlane((int32_t)(b >> 32));
        return __builtin_bit_cast(T, (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
    }
}

template <typename T, int D, bool PER, int KMAX>
__device__ __forceinline__ bool knn_wave(const TreeView<T> &tv, const T (&q)[D], bool active, int32_t root, int k,
                                         int sorted, T seed2, T *drow, int64_t *irow, int64_t idx_base,
                                         Counters &c, T *buf_d, int32_t *buf_s, int tid, int nthreads, int &found) {
    T bd[KMAX];
    int32_t bs[KMAX];
    // near-first reference: mean of the active queries, made wave-uniform
    T ref[D];
    {
        T n = active ? T(1) : T(0);
#pragma unroll
        for (int m = WAVE / 2; m >= 1; m >>= 1) n += __shfl_xor(n, m, WAVE);
#pragma unroll
        for (int d = 0; d < D; d++) {
            T v = active ? q[d] : T(0);
#pragma unroll
            for (int m = WAVE / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, WAVE);
            ref[d] = wave_uniform<T>(n > T(0) ? v / n : T(0));
        }
    }
    const T init = seed2 < T(FLT_MAX) ? seed2 : T(FLT_MAX);
    const bool run = active;
    {
#pragma unroll
        for (int j = 0; j < KMAX; j++) {
            bd[j] = j < KMAX - k ? T(-INFINITY) : init;
            bs[j] = -1;
        }
        int32_t cur = root, from = -1;
        for (;;) {
            const NodeRec<T> nr = tv.nodes[cur];
            if (from == -1) {
                c.visited++;
                if (nr.axis < 0) {
                    // every lane tests the leaf's points (uniform loads); the
                    // candidates are buffered in LDS and inserted in rounds,
                    // so the list update runs as often as the busiest lane
                    // needs, not once per point some lane accepts
                    // Leaves longer than CBUF points go in chunks of CBUF,
                    // so a lane never buffers more than CBUF candidates;
                    // candidates are inserted in point order and re-tested
                    // against the current k-th distance, as in a plain scan.
                    const Range r = tv.range[cur];
                    const int32_t pb = (int32_t)r.start, pe = (int32_t)(r.start + r.count);
#pragma unroll 1
                    for (int32_t cb = pb; cb < pe; cb += CBUF) {
                        const int32_t ce = cb + CBUF < pe ? cb + CBUF : pe;
                        const T bound = bd[KMAX - 1];
                        int cnt = 0;
                        for (int32_t p = cb; p < ce; p++) {
                            T x[D];
                            load_pt<T, D>(tv.pts, p, x);
                            const T d = dist2<T, D, PER>(x, q, tv.L);
                            if (run && d < bound) {
                                buf_d[cnt * nthreads + tid] = d;
                                buf_s[cnt * nthreads + tid] = p;
                                cnt++;
                            }
                        }
                        const int jmax = __reduce_max_sync(~0ull, cnt);
#pragma unroll 1
                        for (int j = 0; j < jmax; j++) {
                            if (j < cnt) {
                                const T d = buf_d[j * nthreads + tid];
                                if (d < bd[KMAX - 1]) {
                                    list_insert<T, KMAX>(bd, bs, d, buf_s[j * nthreads + tid]);
                                    c.inserted++;
                                }
                            }
                        }
                    }
                    c.tested += pe - pb;
                } else {
                    const int32_t nearc = nr.child + (ref[nr.axis] < nr.split ? 0 : 1);
                    const int32_t farc = nearc == nr.child ? nr.child + 1 : nr.child;
                    if (wave_any(run && node_mindist<T, D, PER>(tv, nearc, q) < bd[KMAX - 1])) {
                        cur = nearc;
                        continue;
                    }
                    if (wave_any(run && node_mindist<T, D, PER>(tv, farc, q) < bd[KMAX - 1])) {
                        cur = farc;
                        continue;
                    }
                }
            } else {
                c.climbed++;
                const int32_t nearc = nr.child + (ref[nr.axis] < nr.split ? 0 : 1);
                const int32_t farc = nearc == nr.child ? nr.child + 1 : nr.child;
                if (from == nearc && wave_any(run && node_mindist<T, D, PER>(tv, farc, q) < bd[KMAX - 1])) {
                    cur = farc;
                    from = -1;
                    continue;
                }
            }
            if (cur == root) break;
            from = cur;
            cur = nr.parent;
        }
    }
    // a seeded search that found fewer than k points is incomplete: the
    // caller runs it again unseeded
    if (!active) return true;
    if (bs[KMAX - 1] < 0 && init != T(FLT_MAX)) {
        c.reruns++;
        found = 0;
#pragma unroll
        for (int j = 0; j < KMAX; j++) found += bs[j] >= 0 ? 1 : 0;
        return false;
    }
    bool fix = false;
    T pd = T(-1);
    int64_t pi = 0;
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
        if (j >= KMAX - k) {
            const int32_t s = bs[j];
            CKD_DASSERT(s >= -1 && s < tv.nrows);
            const int64_t oi = s >= 0 ? load_idx(tv.idx, tv.idx64, s) - idx_base : (int64_t)CKD_EMPTY_SLOT;
            const T dv = sqrt_rn<T>(bd[j]);
            fix |= (dv == pd && oi < pi);
            pd = dv;
            pi = oi;
            drow[j - (KMAX - k)] = dv;
            irow[j - (KMAX - k)] = oi;
        }
    }
    if (sorted && fix) {  // order equal distances by index (insertion sort, rare)
        for (int a = 1; a < k; a++) {
            const T dv = drow[a];
            const int64_t iv = irow[a];
            int b = a - 1;
            while (b >= 0 && drow[b] == dv && irow[b] > iv) {
                drow[b + 1] = drow[b];
                irow[b + 1] = irow[b];
                b--;
            }
            drow[b + 1] = dv;
            irow[b + 1] = iv;
        }
    }
    return true;
}

// ------------------------------------------------------------------
// Stack traversal over the wide records (layout in ckd_internal.h).
//
// Same contract and result as knn_wave: the 64 queries of a wave share one
// depth-first traversal, a subtree is entered when any lane's own bound
// reaches its tight point box, each lane accepts with its own strict bound.
// It differs in how the tree is walked: one wide record per inner node
// gives both children's boxes and locations, so deciding a node costs one
// (wave-uniform) load and a leaf's rows come with its parent's record; the
// far child of a node whose two children are both needed is pushed on a
// wave-uniform stack in LDS together with its box, and re-tested against
// the bounds of that moment when popped.  The stack holds at most one entry
// per level of the current path, so SD >= tree depth makes it overflow-free
// (the launcher checks this).
template <typename T, int D>
__device__ __forceinline__ void load_wide(const char *wide, int32_t w, T (&b0)[2 * D], T (&b1)[2 * D], int4 &ref) {
    const char *base = wide + (size_t)w * wide_bytes<T, D>();
    const T *bx = (const T *)base;
#pragma unroll
    for (int d = 0; d < 2 * D; d++) {
        b0[d] = bx[d];
        b1[d] = bx[2 * D + d];
    }
    ref = *(const int4 *)(base + 4 * D * sizeof(T));
}

template <typename T, int D>
struct StackEntry {
    T box[2 * D];
    int32_t a, b;
};

// PB consecutive rows from a wave-uniform row p on, as one contiguous
// block, so all PB loads are issued (and merged) before the first wait
// instead of one load latency per row.  Rows past the leaf may be read
// (the caller masks them): the point array carries PTS_PAD spare rows.
template <typename T, int D, int PB>
__device__ __forceinline__ void load_rows(const T *pts, int32_t p, T (&x)[PB][D]) {
    static_assert(PB <= PTS_PAD, "row batch longer than the padding");
    constexpr int PS = pstride(D);
    const T *src = pts + (int64_t)p * PS;
#pragma unroll
    for (int j = 0; j < PB; j++)
#pragma unroll
        for (int d = 0; d < D; d++) x[j][d] = src[j * PS + d];
}

// Candidates of the stack walk wait in a per-lane LDS buffer (CB > 0 slots
// per lane, slot j of lane tid at j * nthreads + tid; `pend` of them filled)
// and enter the list in rounds: round j inserts every lane's j-th pending
// candidate, re-tested against the list's current k-th distance, so a flush
// costs as many list updates as the fullest lane holds.
template <bool NOIDX, typename T, int KMAX, int CB>
__device__ __forceinline__ void flush_candidates(T (&bd)[KMAX], int32_t (&bs)[KMAX], Counters &c, const T *buf_d,
                                                 const int32_t *buf_s, int tid, int nthreads, int &pend) {
    const int jmax = __reduce_max_sync(~0ull, pend);
#pragma unroll 1
    for (int j = 0; j < jmax; j++) {
        if (j < pend) {
            const T d = buf_d[j * nthreads + tid];
            if (d < bd[KMAX - 1]) {
                list_put<NOIDX, T, KMAX>(bd, bs, d, NOIDX ? 0 : buf_s[j * nthreads + tid]);
                c.inserted++;
            }
        }
    }
    pend = 0;
}

// Leaf scan shared by the stack kernels: rows [pb, pe).  With CB > 0 a row
// closer than the lane's k-th distance joins the lane's pending candidates;
// they stay pending across leaves and are flushed only when some lane's
// buffer could overflow on the next rows (and once when the walk
// ends).  Flushing after every leaf instead made each leaf cost the rounds
// of its busiest lane: most rounds found one or two lanes with a candidate.
// The bound used meanwhile is the list's, which only shrinks at flushes, so
// pending candidates can only make the walk visit more, never less.  With
// CB == 0 rows are inserted directly.  A lane skips the leaf when its rows
// lie in the lane's prefilled range [xs, xs + xm) (knn_stack): a leaf's rows
// are either all inside an ancestor's row range or all outside it.
template <typename T, int D, bool PER, int KMAX, int CB, bool NOIDX = false>
__device__ __forceinline__ void wave_scan_rows(const TreeView<T> &tv, int32_t pb, int32_t pe, const T (&q)[D],
                                               bool run_q, T (&bd)[KMAX], int32_t (&bs)[KMAX], Counters &c,
                                               T *buf_d, int32_t *buf_s, int tid, int nthreads, int32_t xs,
                                               uint32_t xm, int &pend) {
    constexpr int PB = sizeof(T) == 4 ? 8 : 4;  // rows loaded per batch (CB is a multiple)
    const bool run = run_q && (uint32_t)(pb - xs) >= xm;
    if constexpr (CB > 0 && NOIDX) {
        // distances only (the flushes are cheap): blocks of PB rows, a lane's
        // accepted rows written to its next slots, one flush test per block
#pragma unroll 1
        for (int32_t p = pb; p < pe; p += PB) {
            const T bound = bd[KMAX - 1];
            int cnt = pend;
            T *wd = buf_d + cnt * nthreads + tid;  // this lane's next free slot
            T x[PB][D];
            CKD_DASSERT(p >= 0 && p + PB <= tv.nrows + PTS_PAD);
            load_rows<T, D, PB>(tv.pts, p, x);
#pragma unroll
            for (int j = 0; j < PB; j++) {
                const T d = dist2<T, D, PER>(x[j], q, tv.L);
                if (run && p + j < pe && d < bound) {
                    CKD_DASSERT(cnt < CB);
                    *wd = d;
                    wd += nthreads;
                    cnt++;
                }
            }
            pend = cnt;
            if (!lanes_gt(cnt, CB - PB)) continue;
            pend = 0;
            const int jmax = __reduce_max_sync(~0ull, cnt);
#pragma unroll 1
            for (int j = 0; j < jmax; j++) {
                if (j < cnt) {
                    const T dj = buf_d[j * nthreads + tid];
                    if (dj < bd[KMAX - 1]) {
                        list_insert_d<T, KMAX>(bd, dj);
                        c.inserted++;
                    }
                }
            }
        }
    } else if constexpr (CB > 0) {
        // groups of RS rows, loaded together; before each group the buffers
        // are flushed only if some lane could not take all of its rows (a
        // later flush than one test per 8 rows: each flush costs the rounds
        // of its fullest lane, and lanes fill in bursts, at different
        // times).  The rows are loaded after the test, so they are not held
        // across a flush.
        constexpr int RS = 4;
        static_assert(CB > RS, "row groups");
#pragma unroll 1
        for (int32_t p = pb; p < pe; p += RS) {
            if (lanes_gt(pend, CB - RS)) {
                const int jmax = __reduce_max_sync(~0ull, pend);
#pragma unroll 1
                for (int j = 0; j < jmax; j++) {
                    if (j < pend) {
                        const T dj = buf_d[j * nthreads + tid];
                        if (dj < bd[KMAX - 1]) {
                            list_put<NOIDX, T, KMAX>(bd, bs, dj, NOIDX ? 0 : buf_s[j * nthreads + tid]);
                            c.inserted++;
                        }
                    }
                }
                pend = 0;
            }
            const T bound = bd[KMAX - 1];
            T x[RS][D];
            CKD_DASSERT(p >= 0 && p + RS <= tv.nrows + PTS_PAD);
            load_rows<T, D, RS>(tv.pts, p, x);
#pragma unroll
            for (int j = 0; j < RS; j++) {
                const T d = dist2<T, D, PER>(x[j], q, tv.L);
                // every lane writes its next free slot, taken only when the
                // row is accepted: no branch per row
                CKD_DASSERT(pend < CB);
                const int slot = pend * nthreads + tid;
                buf_d[slot] = d;
                if constexpr (!NOIDX) buf_s[slot] = p + j;
                pend += (run && p + j < pe && d < bound) ? 1 : 0;
            }
        }
    } else {
#pragma unroll 1
        for (int32_t p = pb; p < pe; p += PB) {
            T x[PB][D];
            load_rows<T, D, PB>(tv.pts, p, x);
#pragma unroll
            for (int j = 0; j < PB; j++) {
                const T d = dist2<T, D, PER>(x[j], q, tv.L);
                if (run && p + j < pe && d < bd[KMAX - 1]) {
                    list_put<NOIDX, T, KMAX>(bd, bs, d, p + j);
                    c.inserted++;
                }
            }
        }
    }
    c.tested += pe - pb;
}

// Leaf scan of a query held by a group of G lanes (G = 2..64, the lanes
// gbase .. gbase+G-1 of the wave, this one being gbase+sub): each lane
// computes the distance of one row of every G-row slice, then the group
// inserts the slice's rows in row order, so every lane of the group keeps
// the same list as a one-lane scan would.
template <typename T, int D, bool PER, int KMAX>
__device__ __forceinline__ void group_scan_rows(const TreeView<T> &tv, int32_t pb, int32_t pe, const T (&q)[D],
                                                bool run, T (&bd)[KMAX], int32_t (&bs)[KMAX], Counters &c, int G,
                                                int sub, int gbase) {
#pragma unroll 1
    for (int32_t base = pb; base < pe; base += G) {
        const int32_t p = base + sub;
        T d = T(INFINITY);  // rows past the leaf are never accepted
        if (p < pe) {
            T x[D];
            load_pt<T, D>(tv.pts, p, x);
            d = dist2<T, D, PER>(x, q, tv.L);
        }
        const int ns = pe - base < G ? pe - base : G;
#pragma unroll 1
        for (int j = 0; j < ns; j++) {
            const T dj = __shfl(d, gbase + j, WAVE);
            if (run && dj < bd[KMAX - 1]) {
                list_insert<T, KMAX>(bd, bs, dj, base + j);
                c.inserted++;
            }
        }
    }
    c.tested += pe - pb;
}

// Final rows of a kNN list: k (sqrt distance, index - idx_base) pairs,
// equal distances in index order when `sorted`.
template <typename T, int KMAX>
__device__ __forceinline__ void knn_write_row(const TreeView<T> &tv, const T (&bd)[KMAX], const int32_t (&bs)[KMAX],
                                              int k, int sorted, T *drow, int64_t *irow, int64_t idx_base) {
    bool fix = false;
    T pd = T(-1);
    int64_t pi = 0;
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
        if (j >= KMAX - k) {
            const int32_t s = bs[j];
            CKD_DASSERT(s >= -1 && s < tv.nrows);
            const int64_t oi = s >= 0 ? load_idx(tv.idx, tv.idx64, s) - idx_base : (int64_t)CKD_EMPTY_SLOT;
            const T dv = sqrt_rn<T>(bd[j]);
            fix |= (dv == pd && oi < pi);
            pd = dv;
            pi = oi;
            drow[j - (KMAX - k)] = dv;
            irow[j - (KMAX - k)] = oi;
        }
    }
    if (sorted && fix) {  // order equal distances by index (insertion sort, rare)
        for (int a = 1; a < k; a++) {
            const T dv = drow[a];
            const int64_t iv = irow[a];
            int b = a - 1;
            while (b >= 0 && drow[b] == dv && irow[b] > iv) {
                drow[b + 1] = drow[b];
                irow[b + 1] = irow[b];
                b--;
            }
            drow[b + 1] = dv;
            irow[b + 1] = iv;
        }
    }
}

// Periodic shortcut of the stack walk.  A point whose minimum image wraps
// on some axis (L - |dx| < |dx|) is at least w_i = L - |dx_i| >= L/2 - |q_i|
// away on that axis, so its distance can never be accepted once the
// lane's bound is at most m^2, m = min_i (L/2 - |q_i|) minus a margin that
// covers the rounding of |dx| and L - |dx| (1e-3 L, far above it).  When
// every lane of the wave is there, the walk continues with the plain
// (non-periodic) formulas: for the points that do not wrap they give the
// same bits, the ones that wrap are rejected either way, and the plain box
// bound is a lower bound of every acceptable point's distance.  Bounds
// only shrink, so the switch is made once.
template <typename T, int D>
__device__ __forceinline__ T wrap_free_bound(const T (&q)[D], const T *L) {
    T m = T(FLT_MAX);
#pragma unroll
    for (int d = 0; d < D; d++) m = minv(m, L[d] * T(0.5) - absv(q[d]) - L[d] * T(1e-3));
    m = maxv(m, T(0));
    return m * m;
}

// The walk of knn_stack from the item (a, b) with sp stacked entries.
// Returns true when the traversal is complete; with CHECK it instead
// returns false, (a, b) not yet processed, as soon as every running lane's
// bound is within its wrap-free bound `safe2`.
template <typename T, int D, bool PER, int KMAX, int SD, int CB, bool CHECK, bool NOIDX = false>
__device__ __forceinline__ bool stack_walk(const TreeView<T> &tv, const T (&q)[D], bool run, T safe2, T (&bd)[KMAX],
                                           int32_t (&bs)[KMAX], int32_t &a, int32_t &b, int &sp, Counters &c,
                                           T *buf_d, int32_t *buf_s, int tid, int nthreads, StackEntry<T, D> *stk,
                                           int G, int sub, int gbase, int32_t xs, uint32_t xm, int &pend) {
    const uint64_t runm = wave_vote(run);  // votes below: lanes of a compare & runm
    for (;;) {
        if (CHECK && !(lanes_ugt(bd[KMAX - 1], safe2) & runm)) return false;  // !(safe2 >= bound) nowhere
        c.visited++;
        bool descend = false;
        if (b >= 0) {
            if (G == 1)
                wave_scan_rows<T, D, PER, KMAX, CB, NOIDX>(tv, a, a + b, q, run, bd, bs, c, buf_d, buf_s, tid, nthreads,
                                                           xs, xm, pend);
            else group_scan_rows<T, D, PER, KMAX>(tv, a, a + b, q, run, bd, bs, c, G, sub, gbase);
        } else {
            T b0[2 * D], b1[2 * D];
            int4 r;
            CKD_DASSERT(a >= 0 && a < tv.nwide);
            load_wide<T, D>(tv.wide, a, b0, b1, r);
            const T bound = bd[KMAX - 1];
            const T m0 = mindist2<T, D, PER>(b0, q, tv.L), m1 = mindist2<T, D, PER>(b1, q, tv.L);
            const bool n0 = (lanes_lt(m0, bound) & runm) != 0;
            const bool n1 = (lanes_lt(m1, bound) & runm) != 0;
            if (n0 && n1) {
                // both: enter first the child more lanes are nearer to, push the other
                const bool first1 =
                    (int)__popcll(lanes_lt(m1, m0) & runm) > (int)__popcll(lanes_lt(m0, m1) & runm);
                CKD_DASSERT(sp < SD);
                StackEntry<T, D> &e = stk[sp];
#pragma unroll
                for (int d = 0; d < 2 * D; d++) e.box[d] = first1 ? b0[d] : b1[d];
                e.a = first1 ? r.x : r.z;
                e.b = first1 ? r.y : r.w;
                sp++;
                a = first1 ? r.z : r.x;
                b = first1 ? r.w : r.y;
                descend = true;
            } else if (n0 || n1) {
                a = n0 ? r.x : r.z;
                b = n0 ? r.y : r.w;
                descend = true;
            }
        }
        if (descend) {  // the values are uniform; say so, for scalar branches and loads
            a = __builtin_amdgcn_readfirstlane(a);
            b = __builtin_amdgcn_readfirstlane(b);
            continue;
        }
        // pop until a deferred subtree is still needed by some lane
        bool found = false;
        while (sp > 0) {
            sp--;
            const StackEntry<T, D> &e = stk[sp];
            T bx[2 * D];
#pragma unroll
            for (int d = 0; d < 2 * D; d++) bx[d] = e.box[d];
            const int32_t ea = __builtin_amdgcn_readfirstlane(e.a), eb = __builtin_amdgcn_readfirstlane(e.b);
            c.climbed++;
            if (lanes_lt(mindist2<T, D, PER>(bx, q, tv.L), bd[KMAX - 1]) & runm) {
                a = ea;
                b = eb;
                found = true;
                break;
            }
        }
        if (!found) return true;
    }
}

// SPH density of a query from its k nearest neighbours (the step that
// follows kNN in SPH codes): smoothing length h = r_k / 2, so the cubic
// spline kernel's support 2h reaches the k-th neighbour, and
// rho = sum_j m_j W(r_j, h), W(r, h) = 1/(pi h^3) w(r/h),
// w(q) = 1 - 1.5 q^2 + 0.75 q^3 (q < 1), 0.25 (2 - q)^3 (q < 2), else 0
// (Monaghan & Lattanzio 1985).  The query itself counts when it is a tree
// point (r = 0).  mass: per original point, nullptr = unit masses.  With
// NOIDX the list carries no rows (unit masses only); every live slot holds
// a point since k <= n.
template <typename T, int KMAX, bool NOIDX = false>
__device__ __forceinline__ void sph_row(const TreeView<T> &tv, const T (&bd)[KMAX], const int32_t (&bs)[KMAX], int k,
                                        const T *mass, int64_t idx_base, T *h_out, T *rho_out) {
    const T h = sqrt_rn<T>(bd[KMAX - 1]) * T(0.5);
    const T inv_h = T(1) / h;
    T rho = T(0);
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
        if (j >= KMAX - k && (NOIDX || bs[j] >= 0)) {
            const T q = sqrt_rn<T>(bd[j]) * inv_h;
            const T tq = T(2) - q;
            const T w = q < T(1) ? T(1) - T(1.5) * q * q + T(0.75) * q * q * q
                                 : (q < T(2) ? T(0.25) * tq * tq * tq : T(0));
            const T m = !NOIDX && mass ? mass[load_idx(tv.idx, tv.idx64, bs[j]) - idx_base] : T(1);
            rho = rho + m * w;
        }
    }
    *h_out = h;
    *rho_out = rho * inv_h * inv_h * inv_h * T(0.31830988618379067154);  // 1/pi
}

// Returns false for a seeded lane whose seed held fewer than k points (it
// writes nothing then and sets `found` to the number it held; the caller
// searches it again with a larger or no seed).  `stk` is this wave's stack
// of SD entries in LDS.  With G > 1 each query is held by G lanes (lane
// gbase + sub of its group; 64 / G queries per wave): they split the leaf
// rows, keep identical lists, and lane sub == 0 writes the row.  EPI = 1
// writes the SPH smoothing length to *drow and density to *rho_out instead
// of the row (sph_row).
//
// With G == 1 and the query's leaf given (leaf >= 0), the lane first fills
// its list from the rows of the leaf's smallest ancestor holding at least k
// rows (or the root): every lane of the wave inserts about as many rows, so
// the rounds are balanced, and the list's k-th distance is already close
// when the shared walk starts, which then skips those rows.  Far points
// that the walk's first leaves would have inserted, only to be displaced,
// are never inserted.
template <typename T, int D, bool PER, int KMAX, int SD, int CB, int EPI = 0, bool NOIDX = false>
__device__ __forceinline__ bool knn_stack(const TreeView<T> &tv, int2 root, const T (&q)[D], bool active, int k,
                                          int sorted, T seed2, T *drow, int64_t *irow, int64_t idx_base, Counters &c,
                                          T *buf_d, int32_t *buf_s, int tid, int nthreads, StackEntry<T, D> *stk,
                                          int &found, int G = 1, int sub = 0, int gbase = 0, int32_t leaf = -1,
                                          const T *mass = nullptr, T *rho_out = nullptr) {
    T bd[KMAX];
    int32_t bs[KMAX];
    const T init = seed2 < T(FLT_MAX) ? seed2 : T(FLT_MAX);
    const bool run = active;
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
        bd[j] = j < KMAX - k ? T(-INFINITY) : init;
        bs[j] = -1;
    }
    int sp = 0;
    int32_t a = root.x, b = root.y;  // wave-uniform: the subtree to enter
    int32_t xs = 0;                  // prefilled rows [xs, xs + xm) of this lane
    uint32_t xm = 0;
    if (G == 1) {
        if (run && leaf >= 0) {
            int32_t node = leaf;
            Range rg = tv.range[node];
            while (rg.count < k) {
                const int32_t par = tv.nodes[node].parent;
                if (par < 0) break;
                node = par;
                rg = tv.range[node];
            }
            xs = (int32_t)rg.start;
            xm = (uint32_t)rg.count;
        }
        // Rows xs + r .. xs + r + PB - 1 are read as one block, masked past
        // the lane's range.  The wave runs to the largest range, so a lane
        // past its own (r >= xm) reads its block from row 0 instead; within
        // it, the block ends at row xs + xm - 1 + PB - 1 < n + PTS_PAD.  Two
        // rows per block: more would hold row registers beside the list
        // past the 128 VGPRs of four waves per SIMD.
        constexpr int PB = 2;
        static_assert(PB <= PTS_PAD, "prefill block longer than the padding");
        const int mmax = __reduce_max_sync(~0ull, (int)xm);
#pragma unroll 1
        for (int r = 0; r < mmax; r += PB) {
            T x[PB][D];
            CKD_DASSERT((r < (int)xm ? xs + r : 0) + PB <= tv.nrows + PTS_PAD);
            const T *src = tv.pts + (int64_t)(r < (int)xm ? xs + r : 0) * pstride(D);
#pragma unroll
            for (int j = 0; j < PB; j++)
#pragma unroll
                for (int d = 0; d < D; d++) x[j][d] = src[j * pstride(D) + d];
#pragma unroll
            for (int j = 0; j < PB; j++) {
                // unconditional: list_insert leaves the list as it is for a
                // distance at or above the k-th (a masked row inserts +inf),
                // which keeps the list in place instead of a selected copy
                const T d = r + j < (int)xm ? dist2<T, D, PER>(x[j], q, tv.L) : T(INFINITY);
                c.inserted += d < bd[KMAX - 1] ? 1 : 0;
                list_put<NOIDX, T, KMAX>(bd, bs, d, xs + r + j);
            }
        }
        c.tested += xm;
    }
    int pend = 0;  // candidates waiting in this lane's buffer
    if constexpr (PER) {
        const T safe2 = wrap_free_bound<T, D>(q, tv.L);
        if (!stack_walk<T, D, true, KMAX, SD, CB, true, NOIDX>(tv, q, run, safe2, bd, bs, a, b, sp, c, buf_d, buf_s, tid,
                                                        nthreads, stk, G, sub, gbase, xs, xm, pend))
            stack_walk<T, D, false, KMAX, SD, CB, false, NOIDX>(tv, q, run, safe2, bd, bs, a, b, sp, c, buf_d, buf_s, tid,
                                                         nthreads, stk, G, sub, gbase, xs, xm, pend);
    } else {
        stack_walk<T, D, false, KMAX, SD, CB, false, NOIDX>(tv, q, run, T(0), bd, bs, a, b, sp, c, buf_d, buf_s, tid,
                                                     nthreads, stk, G, sub, gbase, xs, xm, pend);
    }
    if constexpr (CB > 0) {
        if (G == 1) flush_candidates<NOIDX, T, KMAX, CB>(bd, bs, c, buf_d, buf_s, tid, nthreads, pend);
    }
    if (!active) return true;
    if (bs[KMAX - 1] < 0 && init != T(FLT_MAX)) {
        c.reruns++;
        found = 0;
#pragma unroll
        for (int j = 0; j < KMAX; j++) found += bs[j] >= 0 ? 1 : 0;
        return false;
    }
    if (sub == 0) {
        if constexpr (EPI == 0) knn_write_row<T, KMAX>(tv, bd, bs, k, sorted, drow, irow, idx_base);
        else sph_row<T, KMAX, NOIDX>(tv, bd, bs, k, mass, idx_base, drow, rho_out);  // EPI 1: h to *drow
    }
    return true;
}

// Ball search below `root`, depth-first, left child first
// (ball_search_recursive_bbox, cosmokdtree.f90:1076-1139): a subtree whose
// point box is entirely farther than r is skipped, one entirely within r
// contributes its whole point range (ball_reduction_recursive,
// :1332-1366), leaves test d^2 <= r^2 inclusively (:1126).  The box bounds
// use the same rounded per-axis differences as the point distance, so the
// whole-range rule includes exactly the points the per-point test would.  Returns the count; with
// FILL writes (d^2, index - idx_base) from dout/iout on: the square roots
// are taken afterwards over the whole output by k_sqrt_rows, so the walk
// does not run the correctly rounded square root for every row a lane of
// the wave accepts.
template <typename T, int D, bool PER, bool FILL>
__device__ __forceinline__ int64_t ball_one(const TreeView<T> &tv, const T (&q)[D], int32_t root, T r2, T *dout,
                                            int64_t *iout, int64_t idx_base) {
    int64_t cnt = 0, w = 0;
    auto emit = [&](int64_t p, T d) {
        if (FILL) {
            dout[w] = d;
            iout[w] = load_idx(tv.idx, tv.idx64, p) - idx_base;
            w++;
        }
    };
    int32_t cur = root, from = -1;  // from == -1: arriving from above
    while (true) {
        const NodeRec<T> nr = tv.nodes[cur];
        if (from == -1) {  // first visit of cur
            T b[2 * D];
            load_box<T, D>(tv.dbox, cur, b);
            const T mn = mindist2<T, D, PER>(b, q, tv.L);
            bool descend = false;
            if (!(mn > r2)) {
                if (maxdist2<T, D>(b, q) <= r2) {
                    const Range r = tv.range[cur];
                    if (FILL)
                        for (int64_t p = r.start; p < r.start + r.count; p++) {
                            T x[D];
                            load_pt<T, D>(tv.pts, p, x);
                            emit(p, dist2<T, D, PER>(x, q, tv.L));
                        }
                    cnt += r.count;
                } else if (nr.axis < 0) {
                    const Range r = tv.range[cur];
                    const int64_t pe = r.start + r.count;
                    constexpr int RB = 8;  // rows whose loads are issued together
                    for (int64_t p0 = r.start; p0 < pe; p0 += RB) {
                        T x[RB][D];
#pragma unroll
                        for (int u = 0; u < RB; u++) load_pt<T, D>(tv.pts, p0 + u < pe ? p0 + u : r.start, x[u]);
#pragma unroll
                        for (int u = 0; u < RB; u++) {
                            const int64_t p = p0 + u;
                            if (p >= pe) break;
                            const T d = dist2<T, D, PER>(x[u], q, tv.L);
                            if (d <= r2) {
                                emit(p, d);
                                cnt++;
                            }
                        }
                    }
                } else {
                    descend = true;
                }
            }
            if (descend) {
                cur = nr.child;  // left first
                continue;
            }
        } else if (from == nr.child) {  // back from the left child
            cur = nr.child + 1;
            from = -1;
            continue;
        }
        if (cur == root) break;
        from = cur;
        cur = nr.parent;
    }
    return cnt;
}

// The same ball search (same decisions, same output order) over the wide
// records, for one query per lane: one record load decides both children
// of a node, where ball_one loads a child's node, box and range apart.
// Stack-free like ball_one: the walk climbs by wparent and reads the
// parent's child pairs to learn which side it came from.  root is the
// subtree's (a, b) pair (wroot convention), rootbox its tight box.  Every
// inner node is loaded at most three times (arrival, return from each
// child), so the walk ends within 3 * nwide + 3 steps; past step_limit it
// stops and sets `bad` (a corrupt tree) instead of looping.
template <typename T, int D, bool PER, bool FILL>
__device__ __forceinline__ int64_t ball_wide_one(const TreeView<T> &tv, const T (&q)[D], int2 root,
                                                 const T *rootbox, T r2, T *dout, int64_t *iout, int64_t idx_base,
                                                 int64_t step_limit, bool &bad) {
    int64_t cnt = 0, w = 0;
    // rows [s, s + n): every one (a subtree within r) or those within r
    auto take_rows = [&](int32_t s, int32_t n, bool whole) {
        if (whole && !FILL) {
            cnt += n;
            return;
        }
        constexpr int RB = 8;  // rows whose loads are issued together
        for (int32_t p0 = s; p0 < s + n; p0 += RB) {
            T x[RB][D];
#pragma unroll
            for (int u = 0; u < RB; u++) load_pt<T, D>(tv.pts, p0 + u < s + n ? p0 + u : s, x[u]);
#pragma unroll
            for (int u = 0; u < RB; u++) {
                const int32_t p = p0 + u;
                if (p >= s + n) break;
                const T d = dist2<T, D, PER>(x[u], q, tv.L);
                if (whole || d <= r2) {
                    if (FILL) {
                        dout[w] = d;  // squared, see ball_one
                        iout[w] = load_idx(tv.idx, tv.idx64, p) - idx_base;
                        w++;
                    }
                    cnt++;
                }
            }
        }
    };
    // one subtree (tight box b, pair (a, n)): the wide id to descend into, or -1
    auto visit = [&](const T *b, int32_t a, int32_t n) -> int32_t {
        if (mindist2<T, D, PER>(b, q, tv.L) > r2) return -1;
        if (maxdist2<T, D>(b, q) <= r2) {
            const int2 rr = n >= 0 ? make_int2(a, n) : tv.wrange[a];
            take_rows(rr.x, rr.y, true);
            return -1;
        }
        if (n >= 0) {
            take_rows(a, n, false);
            return -1;
        }
        return a;
    };
    const int32_t top = visit(rootbox, root.x, root.y);
    if (top < 0) return cnt;
    int32_t cur = top, from = -1, prev = -1;  // from: -1 arriving from above, 0/1 back from that child
    for (int64_t step = 0;; step++) {
        if (step > step_limit) {
            bad = true;
            break;
        }
        T b0[2 * D], b1[2 * D];
        int4 r;
        load_wide<T, D>(tv.wide, cur, b0, b1, r);
        if (from == -2) from = (r.y < 0 && r.x == prev) ? 0 : 1;  // back from a child: which one
        int32_t next = -1;
        if (from < 0) next = visit(b0, r.x, r.y);
        if (next < 0 && from < 1) next = visit(b1, r.z, r.w);
        if (next >= 0) {
            cur = next;
            from = -1;
            continue;
        }
        if (cur == top) break;
        prev = cur;
        cur = tv.wparent[cur];
        from = -2;
    }
    return cnt;
}

// ------------------------------------------------------------------
// Ball search with one walk per wave (queries in leaf order), over the wide
// records: the same decisions as ball_one for every lane -- a subtree whose
// point box is farther than r is skipped, one entirely within r is taken
// whole, a leaf otherwise tests d^2 <= r^2 -- but the subtrees are visited
// once for all the lanes that touch them (the lanes that still need a
// deferred subtree travel with it on the stack as a mask), and a subtree
// taken whole is not copied by its lane: the lane queues (its rows, its
// output offset) and the whole wave writes the queued rows together, 64
// consecutive rows per step (coalesced loads and stores).
struct BallEntry {  // a deferred subtree and the lanes that must visit it
    int32_t a, b;
    uint64_t mask;
};
struct EmitEntry {  // rows [start, start + count) for lane `lane`, out from `out`
    int32_t start, count;
    int64_t out;
    int32_t lane, pad;
};
constexpr int EMIT_CAP = 2 * WAVE;  // two children per lane between drains

template <typename T, int D, bool PER>
__device__ __forceinline__ void drain_emits(const TreeView<T> &tv, const T (&q)[D], EmitEntry *eq, int &qn,
                                            T *dout, int64_t *iout, int64_t idx_base) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll 1
    for (int e = 0; e < qn; e++) {
        const EmitEntry E = eq[e];
        const int owner = __builtin_amdgcn_readfirstlane(E.lane);
        T qo[D];
#pragma unroll
        for (int d = 0; d < D; d++) qo[d] = __shfl(q[d], owner, WAVE);
#pragma unroll 1
        for (int32_t j = lane; j < E.count; j += WAVE) {
            const int32_t p = E.start + j;
            T x[D];
            load_pt<T, D>(tv.pts, p, x);
            dout[E.out + j] = dist2<T, D, PER>(x, qo, tv.L);  // squared, see ball_one
            iout[E.out + j] = load_idx(tv.idx, tv.idx64, p) - idx_base;
        }
    }
    qn = 0;
}

// Returns this lane's count; with FILL writes its rows (d^2, index) from
// dout/iout + w0.
// rootbox/rootrows: the tight box and rows of the subtree at `root`.
template <typename T, int D, bool PER, bool FILL, int SD>
__device__ __forceinline__ int64_t ball_stack(const TreeView<T> &tv, int2 root, const T *rootbox, int2 rootrows,
                                              const T (&q)[D], bool active, T r2, T *dout, int64_t *iout, int64_t w0,
                                              int64_t idx_base, BallEntry *stk, EmitEntry *eq) {
    const int lane = threadIdx.x & (WAVE - 1);
    const uint64_t below = lane ? (~0ull >> (WAVE - lane)) : 0ull;  // lanes under this one
    int64_t cnt = 0, w = w0;
    int qn = 0;  // queued emits (wave-uniform)
    // take rows [s, s + n) whole for the lanes in `in`
    auto take = [&](bool in, int32_t s, int32_t n) {
        const uint64_t m = wave_vote(in);
        if (in) {
            cnt += n;
            if (FILL) {
                CKD_DASSERT(qn + __popcll(m) <= EMIT_CAP);
                EmitEntry &E = eq[qn + __popcll(m & below)];
                E.start = s;
                E.count = n;
                E.out = w;
                E.lane = lane;
                w += n;
            }
        }
        if (FILL) qn += __popcll(m);
    };
    bool run = active;
    {  // the root itself
        T b[2 * D];
#pragma unroll
        for (int d = 0; d < 2 * D; d++) b[d] = rootbox[d];
        const bool reach = run && !(mindist2<T, D, PER>(b, q, tv.L) > r2);
        const bool whole = reach && maxdist2<T, D>(b, q) <= r2;
        take(whole, rootrows.x, rootrows.y);
        run = reach && !whole;
        if (FILL && qn > 0) drain_emits<T, D, PER>(tv, q, eq, qn, dout, iout, idx_base);
    }
    int32_t a = root.x, b = root.y;
    int sp = 0;
    bool go = wave_any(run);
    while (go) {
        if (b >= 0) {  // a leaf some lanes only partly cover: test its rows
            constexpr int PB = sizeof(T) == 4 ? 8 : 4;  // rows per load batch (within the row padding)
#pragma unroll 1
            for (int32_t p0 = a; p0 < a + b; p0 += PB) {
                T x[PB][D];
                CKD_DASSERT(p0 >= 0 && p0 + PB <= tv.nrows + PTS_PAD);
                load_rows<T, D, PB>(tv.pts, p0, x);
#pragma unroll
                for (int j = 0; j < PB; j++) {
                    const int32_t p = p0 + j;
                    if (p >= a + b) break;
                    const T d = dist2<T, D, PER>(x[j], q, tv.L);
                    if (run && d <= r2) {
                        if (FILL) {
                            dout[w] = d;  // squared, see ball_one
                            iout[w] = load_idx(tv.idx, tv.idx64, p) - idx_base;
                            w++;
                        }
                        cnt++;
                    }
                }
            }
        } else {
            T b0[2 * D], b1[2 * D];
            int4 r;
            CKD_DASSERT(a >= 0 && a < tv.nwide);
            load_wide<T, D>(tv.wide, a, b0, b1, r);
            const bool reach0 = run && !(mindist2<T, D, PER>(b0, q, tv.L) > r2);
            const bool reach1 = run && !(mindist2<T, D, PER>(b1, q, tv.L) > r2);
            const bool whole0 = reach0 && maxdist2<T, D>(b0, q) <= r2;
            const bool whole1 = reach1 && maxdist2<T, D>(b1, q) <= r2;
            if (wave_any(whole0)) {
                const int2 rr = r.y >= 0 ? make_int2(r.x, r.y) : tv.wrange[r.x];
                take(whole0, rr.x, rr.y);
            }
            if (wave_any(whole1)) {
                const int2 rr = r.w >= 0 ? make_int2(r.z, r.w) : tv.wrange[r.z];
                take(whole1, rr.x, rr.y);
            }
            if (FILL && qn > 0) drain_emits<T, D, PER>(tv, q, eq, qn, dout, iout, idx_base);
            const bool t0 = reach0 && !whole0, t1 = reach1 && !whole1;
            const uint64_t m0 = wave_vote(t0), m1 = wave_vote(t1);
            if (m0 && m1) {  // left first, the right one deferred with its lanes
                CKD_DASSERT(sp < SD);
                stk[sp].a = r.z;
                stk[sp].b = r.w;
                stk[sp].mask = m1;
                sp++;
            }
            if (m0 || m1) {
                a = __builtin_amdgcn_readfirstlane(m0 ? r.x : r.z);
                b = __builtin_amdgcn_readfirstlane(m0 ? r.y : r.w);
                run = m0 ? t0 : t1;
                continue;
            }
        }
        if (sp == 0) break;  // nothing deferred: done
        sp--;
        a = __builtin_amdgcn_readfirstlane(stk[sp].a);
        b = __builtin_amdgcn_readfirstlane(stk[sp].b);
        run = (stk[sp].mask >> lane) & 1ull;
    }
    if (FILL && qn > 0) drain_emits<T, D, PER>(tv, q, eq, qn, dout, iout, idx_base);
    return cnt;
}

// ----------------------------------------------- one-walk ball search
// The ball walk of ball_stack without a second (fill) walk: rows are
// recorded, as the walk finds them, in a staging area of 16-byte records
// (k_ball_scatter moves them to their CSR places once the counts are
// scanned).  A record is (query, rank of its first row in the query's list,
// first row of a batch of 8, mask of the lane's accepted rows of the batch),
// or, with STAGE_RANGE in the query word, (query, rank, first row, count)
// for a subtree taken whole; the scatter computes the distances again, to
// the same bits.  One record per lane and batch instead of one per row (with
// its distance) made the search 4.5 -> 4.0 ms at 1M queries (the rows
// accepted in a batch average several per lane); reading the original
// indices in the walk (scalar loads of each batch) instead of in the scatter
// made it slower.  Each wave takes the staging area in chunks of STAGE_CHUNK
// records from a global counter (one atomic per chunk); records that do not
// fit the rest of the chunk continue in the next one, so only each wave's
// last chunk is left partly filled.  A full staging area (more chunks than
// it holds) raises the overflow flag and stops the records; the counts stay
// exact.  Record slots are 32-bit (fewer than 2^31 records).
constexpr int STAGE_CHUNK = 512;  // >= the records of one batch (64 lanes)
constexpr uint32_t STAGE_RANGE = 0x80000000u;

struct StageOut {
    int4 *rec;            // nchunks * STAGE_CHUNK records
    int32_t *chunk_fill;  // records written in each taken chunk
    uint32_t *next_chunk; // chunks taken so far
    int32_t nchunks;      // nchunks * STAGE_CHUNK < 2^31
    int32_t *overflow;
};

// The wave's place in the staging area (wave-uniform).
struct StageCursor {
    int32_t cur = -1;  // chunk being filled, -1: none yet
    int fill = STAGE_CHUNK;
    bool dead = false;  // the staging area is full: no more records
};

__device__ __forceinline__ int32_t stage_take_chunk(const StageOut &so) {
    uint32_t c = 0;
    if ((threadIdx.x & (WAVE - 1)) == 0) c = atomicAdd(so.next_chunk, 1u);
    c = __builtin_amdgcn_readfirstlane(c);
    return c < (uint32_t)so.nchunks ? (int32_t)c : -1;
}

// Room for n <= STAGE_CHUNK records (wave-uniform n): the first record
// slot in the current chunk and the chunk the overflow part goes to (-1
// none).  Returns false once the staging area is full.
__device__ __forceinline__ bool stage_reserve(const StageOut &so, StageCursor &sc, int n, int32_t &base,
                                              int32_t &spill) {
    spill = -1;
    if (sc.dead) return false;
    if (sc.fill + n > STAGE_CHUNK || sc.cur < 0) {
        const int32_t c = stage_take_chunk(so);
        if (c < 0) {
            sc.dead = true;
            if ((threadIdx.x & (WAVE - 1)) == 0) {
                *so.overflow = 1;
                if (sc.cur >= 0) so.chunk_fill[sc.cur] = sc.fill;
            }
            return false;
        }
        if (sc.cur < 0) {
            sc.cur = c;
            sc.fill = 0;
        } else {
            spill = c;
        }
    }
    base = sc.cur * STAGE_CHUNK + sc.fill;
    return true;
}

// Slot of the v-th record of a reservation (v < n).
__device__ __forceinline__ int32_t stage_slot(const StageCursor &sc, int32_t base, int32_t spill, int v) {
    const int room = STAGE_CHUNK - sc.fill;
    return v < room ? base + v : spill * STAGE_CHUNK + (v - room);
}

// After writing n records of a reservation.
__device__ __forceinline__ void stage_commit(const StageOut &so, StageCursor &sc, int n, int32_t spill) {
    sc.fill += n;
    if (spill >= 0) {
        if ((threadIdx.x & (WAVE - 1)) == 0) so.chunk_fill[sc.cur] = STAGE_CHUNK;
        sc.cur = spill;
        sc.fill -= STAGE_CHUNK;
    }
}

__device__ __forceinline__ void stage_close(const StageOut &so, const StageCursor &sc) {
    if (!sc.dead && sc.cur >= 0 && (threadIdx.x & (WAVE - 1)) == 0) so.chunk_fill[sc.cur] = sc.fill;
}

// The ball walk of ball_stack (count pass) writing staging records for the
// rows it accepts; returns this lane's count.  qi: the lane's query (< 2^31).
template <typename T, int D, bool PER, int SD>
__device__ __forceinline__ int64_t ball_stage(const TreeView<T> &tv, int2 root, const T *rootbox, int2 rootrows,
                                              const T (&q)[D], bool active, T r2, int32_t qi, const StageOut &so,
                                              BallEntry *stk) {
    const int lane = threadIdx.x & (WAVE - 1);
    int64_t cnt = 0;
    StageCursor sc;
    // rows [s, s + n) whole for the lanes in `in`: one range record each
    auto take = [&](bool in, int32_t s, int32_t n) {
        const uint64_t m = wave_vote(in);
        const int nm = __popcll(m);
        int32_t base, spill;
        if (stage_reserve(so, sc, nm, base, spill)) {
            if (in) {
                const int32_t slot = stage_slot(sc, base, spill, lane_rank(m));
                CKD_DASSERT(slot >= 0 && slot < so.nchunks * STAGE_CHUNK);
                so.rec[slot] = make_int4((int32_t)((uint32_t)qi | STAGE_RANGE), (int32_t)cnt, s, n);
            }
            stage_commit(so, sc, nm, spill);
        }
        if (in) cnt += n;
    };
    bool run = active;
    {  // the root itself
        T b[2 * D];
#pragma unroll
        for (int d = 0; d < 2 * D; d++) b[d] = rootbox[d];
        const bool reach = run && !(mindist2<T, D, PER>(b, q, tv.L) > r2);
        const bool whole = reach && maxdist2<T, D>(b, q) <= r2;
        if (wave_any(whole)) take(whole, rootrows.x, rootrows.y);
        run = reach && !whole;
    }
    int32_t a = root.x, b = root.y;
    int sp = 0;
    bool go = wave_any(run);
    while (go) {
        if (b >= 0) {  // a leaf some lanes only partly cover: test its rows
            constexpr int PB = 8;  // rows per load batch (within the row padding)
            const int32_t pe = a + b;
#pragma unroll 1
            for (int32_t p0 = a; p0 < pe; p0 += PB) {
                T x[PB][D];
                CKD_DASSERT(p0 >= 0 && p0 + PB <= tv.nrows + PTS_PAD);
                load_rows<T, D, PB>(tv.pts, p0, x);
                uint32_t mk = 0;  // this lane's accepted rows of the batch
#pragma unroll
                for (int j = 0; j < PB; j++) {
                    const T dd = dist2<T, D, PER>(x[j], q, tv.L);
                    mk |= (run && p0 + j < pe && dd <= r2) ? (1u << j) : 0u;
                }
                const uint64_t mw = wave_vote(mk != 0);
                if (!mw) continue;
                const int nm = __popcll(mw);
                int32_t base, spill;
                if (stage_reserve(so, sc, nm, base, spill)) {
                    if (mk) {
                        const int32_t slot = stage_slot(sc, base, spill, lane_rank(mw));
                        CKD_DASSERT(slot >= 0 && slot < so.nchunks * STAGE_CHUNK);
                        so.rec[slot] = make_int4((int32_t)qi, (int32_t)cnt, p0, (int32_t)mk);
                    }
                    stage_commit(so, sc, nm, spill);
                }
                cnt += __builtin_popcount(mk);
            }
        } else {
            T b0[2 * D], b1[2 * D];
            int4 r;
            CKD_DASSERT(a >= 0 && a < tv.nwide);
            load_wide<T, D>(tv.wide, a, b0, b1, r);
            const bool reach0 = run && !(mindist2<T, D, PER>(b0, q, tv.L) > r2);
            const bool reach1 = run && !(mindist2<T, D, PER>(b1, q, tv.L) > r2);
            const bool whole0 = reach0 && maxdist2<T, D>(b0, q) <= r2;
            const bool whole1 = reach1 && maxdist2<T, D>(b1, q) <= r2;
            if (wave_any(whole0)) {
                const int2 rr = r.y >= 0 ? make_int2(r.x, r.y) : tv.wrange[r.x];
                take(whole0, rr.x, rr.y);
            }
            if (wave_any(whole1)) {
                const int2 rr = r.w >= 0 ? make_int2(r.z, r.w) : tv.wrange[r.z];
                take(whole1, rr.x, rr.y);
            }
            const bool t0 = reach0 && !whole0, t1 = reach1 && !whole1;
            const uint64_t m0 = wave_vote(t0), m1 = wave_vote(t1);
            if (m0 && m1) {  // left first, the right one deferred with its lanes
                CKD_DASSERT(sp < SD);
                stk[sp].a = r.z;
                stk[sp].b = r.w;
                stk[sp].mask = m1;
                sp++;
            }
            if (m0 || m1) {
                a = __builtin_amdgcn_readfirstlane(m0 ? r.x : r.z);
                b = __builtin_amdgcn_readfirstlane(m0 ? r.y : r.w);
                run = m0 ? t0 : t1;
                continue;
            }
        }
        if (sp == 0) break;  // nothing deferred: done
        sp--;
        a = __builtin_amdgcn_readfirstlane(stk[sp].a);
        b = __builtin_amdgcn_readfirstlane(stk[sp].b);
        run = (stk[sp].mask >> lane) & 1ull;
    }
    stage_close(so, sc);
    return cnt;
}

// Pieces of the interval [lo, hi] of one axis of a box taken modulo the
// periodic box: [lo - kL, hi - kL] clipped to [-L/2, L/2] for k = m+1, m,
// m-1 (ascending positions), m = floor((lo + L/2) / L) in double; pieces
// that touch are merged, so they are disjoint.  The whole axis when
// hi - lo >= L, no piece when lo > hi.  No reference counterpart:
// box_search does not wrap (cosmokdtree.f90:1260); this is the library's
// rule for boxes that straddle the faces of a periodic tree.
template <typename T>
__device__ __forceinline__ int wrap_pieces(T lo, T hi, T L, T (&pa)[3], T (&pb)[3]) {
    const T H = L * T(0.5);
    if (!(lo <= hi)) return 0;
    if (hi - lo >= L) {
        pa[0] = -H;
        pb[0] = H;
        return 1;
    }
    const double m = floor(((double)lo + (double)H) / (double)L);
    int n = 0;
#pragma unroll
    for (int j = 1; j >= -1; j--) {
        const T sh = (T)(m + j) * L;
        T a = lo - sh, b = hi - sh;
        a = a < -H ? -H : a;
        b = b > H ? H : b;
        if (a <= b) {
            if (n > 0 && a <= pb[n - 1]) {
                pb[n - 1] = b > pb[n - 1] ? b : pb[n - 1];
            } else {
                pa[n] = a;
                pb[n] = b;
                n++;
            }
        }
    }
    return n;
}


// Box search below `root` (box_search_recursive, cosmokdtree.f90:1258-1328):
// visit the left child iff box_lo[axis] < split and the right child iff
// box_hi[axis] > split; leaves test box_lo <= x <= box_hi on every axis.  A
// subtree whose point box lies strictly inside the query box is taken
// whole: each split lies within its node's point range, so every split
// below passes both visit tests and every point passes the leaf test.  A
// subtree whose point box misses the query box is skipped: none of its
// points could pass the leaf test.
template <typename T, int D, bool FILL>
__device__ __forceinline__ int64_t box_one(const TreeView<T> &tv, const T (&bx)[2 * D], int32_t root, int64_t *iout,
                                           int64_t idx_base) {
    int64_t cnt = 0, w = 0;
    int32_t cur = root, from = -1;
    while (true) {
        CKD_DASSERT(cur >= 0 && cur < tv.nnodes);
        const NodeRec<T> nr = tv.nodes[cur];
        if (from == -1) {
            T b[2 * D];
            load_box<T, D>(tv.dbox, cur, b);
            bool inside = true, apart = false;
#pragma unroll
            for (int d = 0; d < D; d++) {
                inside = inside && bx[2 * d] < b[d] && b[D + d] < bx[2 * d + 1];
                apart = apart || b[D + d] < bx[2 * d] || b[d] > bx[2 * d + 1];
            }
            if (apart) {
                // no point of the subtree can pass the inclusive leaf test
            } else if (inside) {
                const Range r = tv.range[cur];
                if (FILL)
                    for (int64_t p = r.start; p < r.start + r.count; p++)
                        iout[w++] = load_idx(tv.idx, tv.idx64, p) - idx_base;
                cnt += r.count;
            } else if (nr.axis < 0) {
                const Range r = tv.range[cur];
                const int64_t pe = r.start + r.count;
                constexpr int RB = 8;  // rows whose loads are issued together
                for (int64_t p0 = r.start; p0 < pe; p0 += RB) {
                    T x[RB][D];
#pragma unroll
                    for (int u = 0; u < RB; u++) load_pt<T, D>(tv.pts, p0 + u < pe ? p0 + u : r.start, x[u]);
#pragma unroll
                    for (int u = 0; u < RB; u++) {
                        const int64_t p = p0 + u;
                        if (p >= pe) break;
                        bool in = true;
#pragma unroll
                        for (int d = 0; d < D; d++) in = in && !(x[u][d] < bx[2 * d] || x[u][d] > bx[2 * d + 1]);
                        if (in) {
                            if (FILL) iout[w++] = load_idx(tv.idx, tv.idx64, p) - idx_base;
                            cnt++;
                        }
                    }
                }
            } else {
                if (bx[2 * nr.axis] < nr.split) {
                    cur = nr.child;
                    continue;
                }
                if (bx[2 * nr.axis + 1] > nr.split) {
                    cur = nr.child + 1;
                    continue;
                }
            }
        } else if (from == nr.child && bx[2 * nr.axis + 1] > nr.split) {
            cur = nr.child + 1;
            from = -1;
            continue;
        }
        if (cur == root) break;
        from = cur;
        cur = nr.parent;
    }
    return cnt;
}

// Square roots of the ball rows, in place: rows [0, min(*total, cap)) of
// dout hold d^2 from the fill walks (ball_one, ball_stack).  A grid-stride
// loop, so the count stays on the device.
template <typename T>
__global__ void k_sqrt_rows(T *dout, const int64_t *total, int64_t cap) {
    const int64_t n = *total < cap ? *total : cap;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dout[i] = sqrt_rn<T>(dout[i]);
}
constexpr int SQRT_BLOCKS = 2048;  // k_sqrt_rows grid (256 threads each)

}  // namespace ckd
