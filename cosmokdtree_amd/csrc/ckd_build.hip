// ckd_build.hip -- level-synchronous sliding-midpoint build on the GPU.
//
// Restates build_kdtree / build_kdtree_recursive (cosmokdtree.f90:124-451)
// one tree level at a time.  Every node of a level that holds more than
// `leafsize` points ("active") is split in the same pass:
//   1. its point range is cut into chunks of CH points; one wave per chunk
//      reduces per-axis min/max (with first positions) -- the
//      minval/maxval/minloc/maxloc of cosmokdtree.f90:349-350,372-373;
//   2. one wave per node combines its chunks and applies the split rule:
//      longest cell side, exact ties to the larger spread (:355-365),
//      midpoint 0.5*(lo+hi) (:367), slide to the min (:376-388) or max
//      (:391-403) when every point lies on one side;
//   3. a stable partition (points with x < mid first, :458-479; for a slide
//      only the extreme point moves), one wave per chunk taking 64
//      consecutive points per round and placing them by ballot ranks,
//      scatters coordinates and indices into the other ping-pong buffer;
//   4. both children are created as an adjacent pair with their cells
//      (:418-422); children with <= leafsize points become leaves and their
//      points are copied to the final AoS leaf buckets.
// The left part keeps its input order here, while the reference's Lomuto
// swap permutes the right part; the split sets and therefore the tree are
// the same whenever the slide's extreme point is unique (the GPU takes the
// first in its current order).
#include <vector>

#include "ckd_internal.h"

namespace ckd {

namespace {

constexpr int CH = 2048;          // points per chunk
constexpr int BT = 256;           // threads per workgroup
constexpr int CPB = BT / WAVE;    // chunks (or nodes) per workgroup, one per wave
constexpr int64_t NODE_THREAD_MIN_CH = 4;  // decide with one thread per node when chunks <= 4 x nodes
constexpr int RU = 4;             // rounds of 64 points whose loads a chunk wave issues together
constexpr int STATS_G = 16;       // lanes per chunk of the extremes pass on levels of small nodes
constexpr int64_t DECIDE_BLOCK_MAX = 16;  // levels of at most this many nodes decide with a workgroup per node
constexpr int64_t DECIDE_THREAD_MIN = 8192;  // a thread per node from this many nodes (fewer leave most CUs idle)

struct Chunk {
    int64_t begin, end;  // positions in the working buffers
    int32_t act;         // index into the level's active-node list
    int32_t pad;
};

template <typename T>
struct Ext {  // extreme value + first position
    T mn, mx;
    int64_t mnpos, mxpos;
};

template <typename T>
__device__ __forceinline__ void ext_merge(Ext<T> &a, const Ext<T> &b) {
    if (b.mn < a.mn || (b.mn == a.mn && b.mnpos < a.mnpos)) {
        a.mn = b.mn;
        a.mnpos = b.mnpos;
    }
    if (b.mx > a.mx || (b.mx == a.mx && b.mxpos < a.mxpos)) {
        a.mx = b.mx;
        a.mxpos = b.mxpos;
    }
}

template <typename T>
__device__ __forceinline__ Ext<T> ext_empty() {
    Ext<T> e;
    e.mn = T(INFINITY);
    e.mx = T(-INFINITY);
    e.mnpos = INT64_MAX;
    e.mxpos = INT64_MAX;
    return e;
}

template <typename T>
__device__ __forceinline__ Ext<T> ext_shfl_xor(const Ext<T> &e, int m) {
    Ext<T> o;
    o.mn = __shfl_xor(e.mn, m, WAVE);
    o.mx = __shfl_xor(e.mx, m, WAVE);
    o.mnpos = __shfl_xor(e.mnpos, m, WAVE);
    o.mxpos = __shfl_xor(e.mxpos, m, WAVE);
    return o;
}

// Reduce one Ext per axis across the block; result valid in thread 0.
template <typename T, int D>
__device__ __forceinline__ void block_reduce_ext(Ext<T> (&e)[D], Ext<T> (*lds)[D]) {
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
        for (int m = WAVE / 2; m >= 1; m >>= 1) ext_merge(e[d], ext_shfl_xor(e[d], m));
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    if (lane == 0)
#pragma unroll
        for (int d = 0; d < D; d++) lds[w][d] = e[d];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 1; i < BT / WAVE; i++)
#pragma unroll
            for (int d = 0; d < D; d++) ext_merge(e[d], lds[i][d]);
}

// Reduce one Ext per axis across the wave; result valid in every lane.
template <typename T, int D>
__device__ __forceinline__ void wave_reduce_ext(Ext<T> (&e)[D]) {
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
        for (int m = WAVE / 2; m >= 1; m >>= 1) ext_merge(e[d], ext_shfl_xor(e[d], m));
}

// Decision for one active node.
template <typename T>
struct Decision {
    T mid;
    int64_t selpos;  // slide: position of the point that moves
    int32_t axis;
    int32_t mode;    // 0 partition on x < mid, 1 slide to min, 2 slide to max
};

template <typename T>
__device__ __forceinline__ bool goes_left(const Decision<T> &dc, T x, int64_t pos) {
    return dc.mode == 0 ? (x < dc.mid) : (dc.mode == 1 ? pos == dc.selpos : pos != dc.selpos);
}

// ---------------------------------------------------------------- kernels

template <typename T, int D>
__global__ __launch_bounds__(BT) void k_transpose_in(const T *in, int64_t n, T *P, void *I, int idx64) {
    const int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int d = 0; d < D; d++) P[(int64_t)d * n + i] = in[i * D + d];
    if (idx64) ((int64_t *)I)[i] = i;
    else ((int32_t *)I)[i] = (int32_t)i;
}

__global__ void k_chunk_count(const int32_t *act, int64_t nact, const Range *range, int64_t *nch) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nact) return;
    nch[i] = ceil_div(range[act[i]].count, CH);
}

__global__ void k_chunk_fill(const int32_t *act, const Range *range, const int64_t *chunk_base, Chunk *chunks) {
    const int64_t a = blockIdx.x;
    const Range r = range[act[a]];
    const int64_t nc = ceil_div(r.count, CH), base = chunk_base[a];
    for (int64_t c = threadIdx.x; c < nc; c += blockDim.x) {
        Chunk ch;
        ch.begin = r.start + c * CH;
        ch.end = r.start + (c + 1) * CH < r.start + r.count ? r.start + (c + 1) * CH : r.start + r.count;
        ch.act = (int32_t)a;
        ch.pad = 0;
        chunks[base + c] = ch;
    }
}

// One thread per active node (levels whose nodes have one or a few chunks).
__global__ void k_chunk_fill_t(const int32_t *act, int64_t nact, const Range *range, const int64_t *chunk_base,
                               Chunk *chunks) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nact) return;
    const Range r = range[act[a]];
    const int64_t nc = ceil_div(r.count, CH), base = chunk_base[a];
    for (int64_t c = 0; c < nc; c++) {
        Chunk ch;
        ch.begin = r.start + c * CH;
        ch.end = r.start + (c + 1) * CH < r.start + r.count ? r.start + (c + 1) * CH : r.start + r.count;
        ch.act = (int32_t)a;
        ch.pad = 0;
        chunks[base + c] = ch;
    }
}

// One wave per chunk: lanes take points p = begin + lane, + 64, ...
// (coalesced column reads); strict comparisons keep each lane's first
// position, ext_merge the smallest position on ties.
template <typename T, int D>
__global__ __launch_bounds__(BT) void k_chunk_stats(const Chunk *chunks, const int64_t *nchunks, const T *P, int64_t n,
                                                    Ext<T> *cstat) {
    const int64_t ci = (int64_t)blockIdx.x * CPB + threadIdx.x / WAVE;
    if (ci >= *nchunks) return;  // whole wave
    const int lane = threadIdx.x & (WAVE - 1);
    const Chunk ch = chunks[ci];
    Ext<T> e[D];
#pragma unroll
    for (int d = 0; d < D; d++) e[d] = ext_empty<T>();
    for (int64_t base = ch.begin; base < ch.end; base += RU * WAVE) {
        T x[RU][D];  // RU rounds' loads in flight together
#pragma unroll
        for (int u = 0; u < RU; u++) {
            const int64_t p = base + u * WAVE + lane;
            const int64_t q = p < ch.end ? p : base;
#pragma unroll
            for (int d = 0; d < D; d++) x[u][d] = P[(int64_t)d * n + q];
        }
#pragma unroll
        for (int u = 0; u < RU; u++) {  // rounds in order: first positions stay first
            const int64_t p = base + u * WAVE + lane;
            if (p < ch.end)
#pragma unroll
                for (int d = 0; d < D; d++) {
                    if (x[u][d] < e[d].mn) { e[d].mn = x[u][d]; e[d].mnpos = p; }
                    if (x[u][d] > e[d].mx) { e[d].mx = x[u][d]; e[d].mxpos = p; }
                }
        }
    }
    wave_reduce_ext<T, D>(e);
    if (lane == 0)
#pragma unroll
        for (int d = 0; d < D; d++) cstat[ci * D + d] = e[d];
}

// Same with G lanes per chunk (WAVE / G chunks per wave), for levels of
// small nodes where most chunks hold a few dozen points: shorter
// reductions and fewer waves.
template <typename T, int D, int G>
__global__ __launch_bounds__(BT) void k_chunk_stats_g(const Chunk *chunks, const int64_t *nchunks, const T *P,
                                                      int64_t n, Ext<T> *cstat) {
    const int64_t ci = ((int64_t)blockIdx.x * BT + threadIdx.x) / G;
    if (ci >= *nchunks) return;  // whole group
    const int gl = threadIdx.x & (G - 1);
    const Chunk ch = chunks[ci];
    Ext<T> e[D];
#pragma unroll
    for (int d = 0; d < D; d++) e[d] = ext_empty<T>();
    for (int64_t base = ch.begin; base < ch.end; base += RU * G) {
        T x[RU][D];
#pragma unroll
        for (int u = 0; u < RU; u++) {
            const int64_t p = base + u * G + gl;
            const int64_t q = p < ch.end ? p : base;
#pragma unroll
            for (int d = 0; d < D; d++) x[u][d] = P[(int64_t)d * n + q];
        }
#pragma unroll
        for (int u = 0; u < RU; u++) {
            const int64_t p = base + u * G + gl;
            if (p < ch.end)
#pragma unroll
                for (int d = 0; d < D; d++) {
                    if (x[u][d] < e[d].mn) { e[d].mn = x[u][d]; e[d].mnpos = p; }
                    if (x[u][d] > e[d].mx) { e[d].mx = x[u][d]; e[d].mxpos = p; }
                }
        }
    }
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
        for (int m = G / 2; m >= 1; m >>= 1) {
            Ext<T> o;
            o.mn = __shfl_xor(e[d].mn, m, G);
            o.mx = __shfl_xor(e[d].mx, m, G);
            o.mnpos = __shfl_xor(e[d].mnpos, m, G);
            o.mxpos = __shfl_xor(e[d].mxpos, m, G);
            ext_merge(e[d], o);
        }
    if (gl == 0)
#pragma unroll
        for (int d = 0; d < D; d++) cstat[ci * D + d] = e[d];
}

// The split rule for active node a (id) from its merged extremes e:
// longest cell side with the spread tie-break, midpoint, slide.
template <typename T, int D>
__device__ __forceinline__ void decide_node(int64_t a, int32_t id, const Ext<T> (&e)[D], NodeRec<T> *nodes, T *bbox,
                                            T *dbox, Decision<T> *dec, int root_from_data) {
    T *bb = bbox + (int64_t)id * 2 * D;
    T *db = dbox + (int64_t)id * 2 * D;
#pragma unroll
    for (int d = 0; d < D; d++) {
        db[d] = e[d].mn;
        db[D + d] = e[d].mx;
    }
    if (root_from_data) {  // non-periodic root cell = data bounds (cosmokdtree.f90:226-229)
#pragma unroll
        for (int d = 0; d < D; d++) {
            bb[d] = e[d].mn;
            bb[D + d] = e[d].mx;
        }
    }
    T side[D], spread[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        side[d] = bb[D + d] - bb[d];
        spread[d] = e[d].mx - e[d].mn;
    }
    T maxside = side[0];
    int axis = 0;
#pragma unroll
    for (int d = 1; d < D; d++) {
        if (side[d] > maxside) {
            maxside = side[d];
            axis = d;
        }
        if (side[d] == maxside && spread[d] > spread[axis]) axis = d;
    }
    Decision<T> dc;
    dc.axis = axis;
    dc.mid = T(0.5) * (bb[axis] + bb[D + axis]);
    dc.mode = 0;
    dc.selpos = -1;
    if (e[axis].mn >= dc.mid) {
        dc.mode = 1;
        dc.mid = e[axis].mn;
        dc.selpos = e[axis].mnpos;
    } else if (e[axis].mx < dc.mid) {
        dc.mode = 2;
        dc.mid = e[axis].mx;
        dc.selpos = e[axis].mxpos;
    }
    dec[a] = dc;
    nodes[id].split = dc.mid;
    nodes[id].axis = axis;}

// One wave per active node: its lanes merge the node's chunk statistics,
// lane 0 applies the split rule (levels of few nodes with many chunks).
// With slot_of, node a is the child slot slot_of[a] = 2 x parent + side of
// the previous level, and its statistics are that level's chunk records of
// the parent (chunk_base, nch of that level; cstat left, cstat_r right).
template <typename T, int D>
__device__ __forceinline__ void decide_chunks(int64_t a, const int64_t *chunk_base, const int64_t *nch,
                                              const Ext<T> *cstat, const int64_t *slot_of, const Ext<T> *cstat_r,
                                              int64_t &c0, int64_t &c1, const Ext<T> *&cs) {
    int64_t pa = a;
    cs = cstat;
    if (slot_of) {
        const int64_t m = slot_of[a];
        pa = m >> 1;
        cs = (m & 1) ? cstat_r : cstat;
    }
    c0 = chunk_base[pa];
    c1 = c0 + nch[pa];
}

template <typename T, int D>
__global__ __launch_bounds__(BT) void k_node_decide(const int32_t *act, int64_t nact, const int64_t *chunk_base,
                                                    const int64_t *nch, const Ext<T> *cstat, NodeRec<T> *nodes,
                                                    T *bbox, T *dbox, Decision<T> *dec, int root_from_data,
                                                    const int64_t *slot_of = nullptr, const Ext<T> *cstat_r = nullptr) {
    const int64_t a = (int64_t)blockIdx.x * CPB + threadIdx.x / WAVE;
    if (a >= nact) return;  // whole wave
    const int lane = threadIdx.x & (WAVE - 1);
    Ext<T> e[D];
#pragma unroll
    for (int d = 0; d < D; d++) e[d] = ext_empty<T>();
    int64_t c0, c1;
    decide_chunks<T, D>(a, chunk_base, nch, cstat, slot_of, cstat_r, c0, c1, cstat);
    for (int64_t base = c0; base < c1; base += RU * WAVE) {
        Ext<T> x[RU][D];  // RU rounds' loads in flight together
#pragma unroll
        for (int u = 0; u < RU; u++) {
            const int64_t c = base + u * WAVE + lane;
            const int64_t q = c < c1 ? c : base;
#pragma unroll
            for (int d = 0; d < D; d++) x[u][d] = cstat[q * D + d];
        }
#pragma unroll
        for (int u = 0; u < RU; u++)
            if (base + u * WAVE + lane < c1)
#pragma unroll
                for (int d = 0; d < D; d++) ext_merge(e[d], x[u][d]);
    }
    wave_reduce_ext<T, D>(e);
    if (lane == 0) decide_node<T, D>(a, act[a], e, nodes, bbox, dbox, dec, root_from_data);
}

// One workgroup per active node (the top levels: few nodes with thousands
// of chunks each, where one wave per node waited on a chain of 32 rounds of
// loads at the root: 67 -> 27 us there; from 32 nodes on, a wave per node
// is faster).  The merge keeps the smallest (value, position), so its order
// does not change the result.
template <typename T, int D>
__global__ __launch_bounds__(BT) void k_node_decide_b(const int32_t *act, int64_t nact, const int64_t *chunk_base,
                                                      const int64_t *nch, const Ext<T> *cstat, NodeRec<T> *nodes,
                                                      T *bbox, T *dbox, Decision<T> *dec, int root_from_data,
                                                      const int64_t *slot_of = nullptr,
                                                      const Ext<T> *cstat_r = nullptr) {
    __shared__ Ext<T> lds[BT / WAVE][D];
    const int64_t a = blockIdx.x;
    Ext<T> e[D];
#pragma unroll
    for (int d = 0; d < D; d++) e[d] = ext_empty<T>();
    int64_t c0, c1;
    decide_chunks<T, D>(a, chunk_base, nch, cstat, slot_of, cstat_r, c0, c1, cstat);
    for (int64_t base = c0; base < c1; base += RU * BT) {
        Ext<T> x[RU][D];  // RU rounds' loads in flight together
#pragma unroll
        for (int u = 0; u < RU; u++) {
            const int64_t c = base + u * BT + threadIdx.x;
            const int64_t q = c < c1 ? c : base;
#pragma unroll
            for (int d = 0; d < D; d++) x[u][d] = cstat[q * D + d];
        }
#pragma unroll
        for (int u = 0; u < RU; u++)
            if (base + u * BT + threadIdx.x < c1)
#pragma unroll
                for (int d = 0; d < D; d++) ext_merge(e[d], x[u][d]);
    }
    block_reduce_ext<T, D>(e, lds);
    if (threadIdx.x == 0) decide_node<T, D>(a, act[a], e, nodes, bbox, dbox, dec, root_from_data);
}

// One thread per active node, merging its chunks in order (levels of many
// nodes with one or a few chunks each).
template <typename T, int D>
__global__ __launch_bounds__(BT) void k_node_decide_t(const int32_t *act, int64_t nact, const int64_t *chunk_base,
                                                      const int64_t *nch, const Ext<T> *cstat, NodeRec<T> *nodes,
                                                      T *bbox, T *dbox, Decision<T> *dec, int root_from_data) {
    const int64_t a = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (a >= nact) return;
    Ext<T> e[D];
#pragma unroll
    for (int d = 0; d < D; d++) e[d] = ext_empty<T>();
    const int64_t c0 = chunk_base[a], c1 = c0 + nch[a];
    for (int64_t c = c0; c < c1; c++)
#pragma unroll
        for (int d = 0; d < D; d++) ext_merge(e[d], cstat[c * D + d]);
    decide_node<T, D>(a, act[a], e, nodes, bbox, dbox, dec, root_from_data);
}

// Left count of each chunk (one wave per chunk, 64 points per ballot);
// chunk slots past the level's count get 0 for the scan over nub slots.
template <typename T>
__global__ __launch_bounds__(BT) void k_chunk_countleft(const Chunk *chunks, const int64_t *nchunks, int64_t nub,
                                                        const T *P, int64_t n, const Decision<T> *dec,
                                                        int64_t *cleft) {
    const int64_t ci = (int64_t)blockIdx.x * CPB + threadIdx.x / WAVE;
    if (ci >= nub) return;  // whole wave
    const int lane = threadIdx.x & (WAVE - 1);
    if (ci >= *nchunks) {
        if (lane == 0) cleft[ci] = 0;
        return;
    }
    const Chunk ch = chunks[ci];
    const Decision<T> dc = dec[ch.act];
    const T *col = P + (int64_t)dc.axis * n;
    int64_t c = 0;
    for (int64_t base = ch.begin; base < ch.end; base += RU * WAVE) {
        T x[RU];
#pragma unroll
        for (int u = 0; u < RU; u++) {
            const int64_t p = base + u * WAVE + lane;
            x[u] = col[p < ch.end ? p : base];
        }
#pragma unroll
        for (int u = 0; u < RU; u++) {
            const int64_t p = base + u * WAVE + lane;
            c += __popcll(wave_vote(p < ch.end && goes_left(dc, x[u], p)));
        }
    }
    if (lane == 0) cleft[ci] = c;
}

// Stable partition of one chunk by one wave: rounds of 64 consecutive
// points; a point's place among the round's left (right) points is its
// ballot rank, added to the left (right) points of earlier rounds, earlier
// chunks of the node (cleft_off) and, for the right side, the node's left
// count.  Reads and writes of a round are contiguous runs.
template <typename T, int D, typename IdxT, bool FUSE = false>
__global__ __launch_bounds__(BT) void k_chunk_scatter(const Chunk *chunks, const int64_t *nchunks,
                                                      const int64_t *chunk_base, const int64_t *nch,
                                                      const int32_t *act, const Range *range, const T *P,
                                                      const IdxT *I, T *P2, IdxT *I2, int64_t n,
                                                      const Decision<T> *dec, const int64_t *cleft_off,
                                                      Ext<T> *cs_l = nullptr, Ext<T> *cs_r = nullptr) {
    const int64_t ci = (int64_t)blockIdx.x * CPB + threadIdx.x / WAVE;
    if (ci >= *nchunks) return;  // whole wave
    const int lane = threadIdx.x & (WAVE - 1);
    const uint64_t below = (1ull << lane) - 1;  // lanes before this one
    const Chunk ch = chunks[ci];
    const Decision<T> dc = dec[ch.act];
    const int64_t nstart = range[act[ch.act]].start;
    const int64_t first = chunk_base[ch.act];
    const int64_t left_base = nstart + (cleft_off[ci] - cleft_off[first]);
    const int64_t n_left = cleft_off[first + nch[ch.act]] - cleft_off[first];
    const int64_t right_base = nstart + n_left + (ch.begin - nstart) - (cleft_off[ci] - cleft_off[first]);
    int64_t lacc = 0, racc = 0;
    Ext<T> el[D], er[D];  // FUSE: this lane's extremes of each child, positions in P2
    if constexpr (FUSE) {
#pragma unroll
        for (int d = 0; d < D; d++) el[d] = er[d] = ext_empty<T>();
    }
    for (int64_t base = ch.begin; base < ch.end; base += RU * WAVE) {
        T x[RU][D];  // RU rounds' loads in flight together
        IdxT iv[RU];
#pragma unroll
        for (int u = 0; u < RU; u++) {
            const int64_t p = base + u * WAVE + lane;
            const int64_t q = p < ch.end ? p : base;  // a valid row for the idle lanes' loads
#pragma unroll
            for (int d = 0; d < D; d++) x[u][d] = P[(int64_t)d * n + q];
            iv[u] = I[q];
        }
#pragma unroll
        for (int u = 0; u < RU; u++) {
            const int64_t rb = base + u * WAVE, p = rb + lane;
            const bool in = p < ch.end;
            T xa = x[u][0];
#pragma unroll
            for (int d = 1; d < D; d++)
                if (d == dc.axis) xa = x[u][d];
            const bool f = in && goes_left(dc, xa, p);
            const uint64_t m = wave_vote(f);
            const int rank = __popcll(m & below);
            const int nl = __popcll(m);
            const int cnt = ch.end - rb >= WAVE ? WAVE : (ch.end > rb ? (int)(ch.end - rb) : 0);
            if (in) {
                const int64_t dst = f ? left_base + lacc + rank : right_base + racc + (lane - rank);
#pragma unroll
                for (int d = 0; d < D; d++) P2[(int64_t)d * n + dst] = x[u][d];
                I2[dst] = iv[u];
                if constexpr (FUSE) {
#pragma unroll
                    for (int d = 0; d < D; d++) {
                        Ext<T> pt;
                        pt.mn = pt.mx = x[u][d];
                        pt.mnpos = pt.mxpos = dst;
                        ext_merge(f ? el[d] : er[d], pt);
                    }
                }
            }
            lacc += nl;
            racc += cnt - nl;
        }
    }
    if constexpr (FUSE) {  // the chunk's record of each child's extremes
        wave_reduce_ext<T, D>(el);
        wave_reduce_ext<T, D>(er);
        if (lane == 0)
#pragma unroll
            for (int d = 0; d < D; d++) {
                cs_l[ci * D + d] = el[d];
                cs_r[ci * D + d] = er[d];
            }
    }
}

template <typename T, int D>
__global__ void k_make_children(const int32_t *act, int64_t nact, const Decision<T> *dec, const int64_t *chunk_base,
                                const int64_t *nch, const int64_t *cleft_off, NodeRec<T> *nodes, T *bbox,
                                Range *range, int32_t next_id, int leafsize, int64_t *child_active) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nact) return;
    const int32_t id = act[a];
    const Decision<T> dc = dec[a];
    const int64_t first = chunk_base[a];
    const int64_t n_left = cleft_off[first + nch[a]] - cleft_off[first];
    const Range r = range[id];
    const int32_t c0 = next_id + 2 * (int32_t)a;
    nodes[id].child = c0;
    const T *pb = bbox + (int64_t)id * 2 * D;
    for (int s = 0; s < 2; s++) {
        const int32_t c = c0 + s;
        T *cb = bbox + (int64_t)c * 2 * D;
#pragma unroll
        for (int d = 0; d < 2 * D; d++) cb[d] = pb[d];
        if (s == 0) cb[D + dc.axis] = dc.mid;  // left upper bound (cosmokdtree.f90:421)
        else cb[dc.axis] = dc.mid;             // right lower bound (:422)
        Range cr;
        cr.start = s == 0 ? r.start : r.start + n_left;
        cr.count = s == 0 ? n_left : r.count - n_left;
        range[c] = cr;
        NodeRec<T> nr;
        nr.split = T(0);
        nr.axis = -1;
        nr.child = -1;
        nr.parent = id;
        nodes[c] = nr;
        child_active[2 * a + s] = cr.count > leafsize ? 1 : 0;
    }
}

__global__ void k_compact_active(const int64_t *child_active, const int64_t *pos, int64_t m, int32_t next_id,
                                 int32_t *next_act, int64_t *slot_of = nullptr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    if (child_active[i]) {
        next_act[pos[i]] = next_id + (int32_t)i;
        if (slot_of) slot_of[pos[i]] = i;  // the child slot (2 x parent + side) of each next active node
    }
}

// Copy the points of the leaves created this level (children that are not
// active) into the final AoS leaf buckets.
template <typename T, int D, typename IdxT>
__global__ void k_copy_leaves(const int64_t *child_active, int64_t m, int32_t next_id, const Range *range,
                              int leafsize, const T *P, const IdxT *I, int64_t n, T *F, IdxT *FI) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ci = t / leafsize, j = t % leafsize;
    if (ci >= m || child_active[ci]) return;
    const Range r = range[next_id + ci];
    if (j >= r.count) return;
    const int64_t p = r.start + j;
    constexpr int PS = pstride(D);
    T row[PS];
#pragma unroll
    for (int d = 0; d < PS; d++) row[d] = d < D ? P[(int64_t)d * n + p] : T(0);
    if constexpr (sizeof(T) == 4 && PS == 4) {  // the row as one 16-byte store (rows are 16-byte aligned)
        *(float4 *)(F + p * PS) = make_float4(row[0], row[1], row[2], row[3]);
    } else if constexpr (sizeof(T) == 8 && PS == 4) {
        *(double2 *)(F + p * PS) = make_double2(row[0], row[1]);
        *(double2 *)(F + p * PS + 2) = make_double2(row[2], row[3]);
    } else {
#pragma unroll
        for (int d = 0; d < PS; d++) F[p * PS + d] = row[d];
    }
    FI[p] = I[p];
}

// Roots of all point sets (node r = set r, points [set_off[r], set_off[r+1])).
// A root is active when it holds more than leafsize points.
template <typename T, int D>
__global__ void k_init_roots(NodeRec<T> *nodes, T *bbox, Range *range, const int64_t *set_off, int64_t nsets,
                             TreeView<T> tv, int periodic, int leafsize, int64_t *root_active) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nsets) return;
    NodeRec<T> nr;
    nr.split = T(0);
    nr.axis = -1;
    nr.child = -1;
    nr.parent = -1;
    nodes[r] = nr;
    Range rg;
    rg.start = set_off[r];
    rg.count = set_off[r + 1] - set_off[r];
    range[r] = rg;
    root_active[r] = rg.count > leafsize ? 1 : 0;
    if (periodic) {  // cosmokdtree.f90:218-224
        for (int d = 0; d < D; d++) {
            bbox[r * 2 * D + d] = -tv.L[d] / T(2);
            bbox[r * 2 * D + D + d] = tv.L[d] / T(2);
        }
    }
}

// Roots that are leaves already (at most leafsize points): cell bounds
// from the data when non-periodic (cosmokdtree.f90:226-229), and the copy
// into the leaf buckets.  One block per root.
template <typename T, int D, typename IdxT>
__global__ __launch_bounds__(BT) void k_root_leaves(const T *P, const IdxT *I, int64_t n, const int64_t *set_off,
                                                    const int64_t *root_active, T *bbox, int root_from_data, T *F,
                                                    IdxT *FI) {
    __shared__ Ext<T> lds[BT / WAVE][D];
    const int64_t r = blockIdx.x;
    if (root_active[r]) return;
    const int64_t b = set_off[r], e_ = set_off[r + 1];
    Ext<T> e[D];
#pragma unroll
    for (int d = 0; d < D; d++) e[d] = ext_empty<T>();
    for (int64_t p = b + threadIdx.x; p < e_; p += BT)
#pragma unroll
        for (int d = 0; d < D; d++) {
            T x = P[(int64_t)d * n + p];
            if (x < e[d].mn) { e[d].mn = x; e[d].mnpos = p; }
            if (x > e[d].mx) { e[d].mx = x; e[d].mxpos = p; }
        }
    block_reduce_ext<T, D>(e, lds);
    constexpr int PS = pstride(D);
    for (int64_t p = b + threadIdx.x; p < e_; p += BT) {
#pragma unroll
        for (int d = 0; d < D; d++) F[p * PS + d] = P[(int64_t)d * n + p];
#pragma unroll
        for (int d = D; d < PS; d++) F[p * PS + d] = T(0);
        FI[p] = I[p];
    }
    if (threadIdx.x == 0 && root_from_data)
#pragma unroll
        for (int d = 0; d < D; d++) {
            bbox[r * 2 * D + d] = e[d].mn;
            bbox[r * 2 * D + D + d] = e[d].mx;
        }
}

// Tight point boxes of the leaves, from the final leaf buckets.
template <typename T, int D>
__global__ void k_leaf_boxes(const NodeRec<T> *nodes, const Range *range, int64_t nnodes, const T *F, T *dbox) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nnodes || nodes[v].axis >= 0) return;
    const Range r = range[v];
    constexpr int PS = pstride(D);
    T mn[D], mx[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        mn[d] = T(INFINITY);
        mx[d] = T(-INFINITY);
    }
    // four rows' loads in flight together; past the leaf's last row the
    // last row is read again, which leaves the extremes as they are
    const int64_t e = r.start + r.count;
    for (int64_t p = r.start; p < e; p += 4) {
        T x[4][D];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t q = p + u < e ? p + u : e - 1;
            if constexpr (sizeof(T) == 4 && PS == 4) {  // one 16-byte load per row (rows are 16-byte aligned)
                const float4 v = *(const float4 *)(F + q * PS);
                x[u][0] = v.x;
                if constexpr (D > 1) x[u][1] = v.y;
                if constexpr (D > 2) x[u][2] = v.z;
                if constexpr (D > 3) x[u][3] = v.w;
            } else {
#pragma unroll
                for (int d = 0; d < D; d++) x[u][d] = F[q * PS + d];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int d = 0; d < D; d++) {
                mn[d] = x[u][d] < mn[d] ? x[u][d] : mn[d];
                mx[d] = x[u][d] > mx[d] ? x[u][d] : mx[d];
            }
    }
#pragma unroll
    for (int d = 0; d < D; d++) {
        dbox[v * 2 * D + d] = mn[d];
        dbox[v * 2 * D + D + d] = mx[d];
    }
}

// ---- depth-first renumbering of the sibling pairs
// The level loop numbers nodes level by level.  Afterwards the pairs of
// children are renumbered in depth-first order: the children of v keep
// adjacent ids (pair base b(v)), the descendants of the first child follow
// at b(v)+2, those of the second child after them.  A subtree's nodes are
// then contiguous, and leaves sorted by id are in depth-first (spatial)
// order.
template <typename T>
__global__ void k_subtree_size(const NodeRec<T> *nodes, int64_t lo, int64_t hi, int64_t *desc) {
    const int64_t v = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= hi) return;
    const NodeRec<T> nr = nodes[v];
    desc[v] = nr.axis < 0 ? 0 : 2 + desc[nr.child] + desc[nr.child + 1];
}

__global__ void k_root_ids(const int64_t *desc, const int64_t *root_off, int64_t nsets, int64_t *newid,
                           int64_t *pairbase) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nsets) return;
    newid[r] = r;
    pairbase[r] = nsets + root_off[r];
}

template <typename T>
__global__ void k_child_ids(const NodeRec<T> *nodes, int64_t lo, int64_t hi, const int64_t *desc, int64_t *newid,
                            int64_t *pairbase) {
    const int64_t v = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= hi) return;
    const NodeRec<T> nr = nodes[v];
    if (nr.axis < 0) return;
    const int64_t b = pairbase[v], c0 = nr.child, c1 = nr.child + 1;
    newid[c0] = b;
    newid[c1] = b + 1;
    pairbase[c0] = b + 2;
    pairbase[c1] = b + 2 + desc[c0];
}

// oldid[newid[v]] = v: the node each depth-first position takes.
__global__ void k_invert_ids(const int64_t *newid, int64_t nn, int64_t *oldid) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    CKD_DASSERT(newid[v] >= 0 && newid[v] < nn);
    oldid[newid[v]] = v;
}

// The level-order arrays in depth-first order, gathered: thread w writes
// position w of each array (contiguous stores) from node oldid[w], its
// child and parent renamed.  Scattering each node's 80 bytes to its new
// place instead took 0.52 ms of a 2^24-point build.
template <typename T, int D>
__global__ void k_permute_nodes(const NodeRec<T> *nodes, const T *bbox, const T *dbox, const Range *range,
                                int64_t nn, const int64_t *newid, const int64_t *oldid, NodeRec<T> *nodes2,
                                T *bbox2, T *dbox2, Range *range2) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nn) return;
    const int64_t v = oldid[w];
    NodeRec<T> nr = nodes[v];
    if (nr.axis >= 0) nr.child = (int32_t)newid[nr.child];
    if (nr.parent >= 0) nr.parent = (int32_t)newid[nr.parent];
    nodes2[w] = nr;
#pragma unroll
    for (int d = 0; d < 2 * D; d++) {
        bbox2[w * 2 * D + d] = bbox[v * 2 * D + d];
        dbox2[w * 2 * D + d] = dbox[v * 2 * D + d];
    }
    range2[w] = range[v];
}

__global__ void k_root_desc(const int64_t *desc, int64_t nsets, int64_t *out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nsets) out[r] = desc[r];
}

// ---- wide records (layout in ckd_internal.h)
template <typename T>
__global__ void k_inner_flags(const NodeRec<T> *nodes, int64_t nn, int64_t *flag) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nn) flag[v] = nodes[v].axis >= 0 ? 1 : 0;
}

template <typename T>
__device__ __forceinline__ int2 child_pair(const NodeRec<T> *nodes, const Range *range, const int64_t *widx,
                                           int64_t c) {
    if (nodes[c].axis >= 0) return make_int2((int32_t)widx[c], -1);
    const Range r = range[c];
    return make_int2((int32_t)r.start, (int32_t)r.count);
}

template <typename T, int D>
__global__ void k_fill_wide(const NodeRec<T> *nodes, const T *dbox, const Range *range, const int64_t *widx,
                            int64_t nn, char *wide, int2 *wrange) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    const NodeRec<T> nr = nodes[v];
    if (nr.axis < 0) return;
    const Range rv = range[v];
    wrange[widx[v]] = make_int2((int32_t)rv.start, (int32_t)rv.count);
    char *base = wide + widx[v] * wide_bytes<T, D>();
    const int2 p0 = child_pair(nodes, range, widx, nr.child), p1 = child_pair(nodes, range, widx, nr.child + 1);
    if constexpr (sizeof(T) == 4 && D == 3) {
        // the two children's boxes are 12 consecutive floats (8-byte aligned);
        // the 64-byte record takes them as three 16-byte stores
        const float2 *src = (const float2 *)(dbox + (int64_t)nr.child * 2 * D);
        float2 h[6];
#pragma unroll
        for (int i = 0; i < 6; i++) h[i] = src[i];
        float4 *dst = (float4 *)base;
#pragma unroll
        for (int i = 0; i < 3; i++) dst[i] = make_float4(h[2 * i].x, h[2 * i].y, h[2 * i + 1].x, h[2 * i + 1].y);
    } else {
        T *bx = (T *)base;
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int d = 0; d < 2 * D; d++) bx[c * 2 * D + d] = dbox[(int64_t)(nr.child + c) * 2 * D + d];
    }
    *(int4 *)(base + 4 * D * sizeof(T)) = make_int4(p0.x, p0.y, p1.x, p1.y);
}

__global__ void k_set_i32(int32_t *p, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Parent of each inner child, from the child pairs of the wide records.
__global__ void k_wide_parents(const char *wide, size_t rec_bytes, size_t pair_off, int64_t nwide,
                               int32_t *wparent) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nwide) return;
    const int4 r = *(const int4 *)(wide + (size_t)v * rec_bytes + pair_off);
    if (r.y < 0) wparent[r.x] = (int32_t)v;
    if (r.w < 0) wparent[r.z] = (int32_t)v;
}

template <typename T>
__global__ void k_root_pairs(const NodeRec<T> *nodes, const Range *range, const int64_t *widx, int64_t nsets,
                             int2 *wroot) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nsets) wroot[r] = child_pair(nodes, range, widx, r);
}

// Periodic range check (cosmokdtree.f90:176-190): counts points outside.
template <typename T, int D>
__global__ void k_check_range(const T *in, int64_t n, TreeView<T> tv, int *bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int d = 0; d < D; d++) {
        T x = in[i * D + d];
        if (x < -tv.L[d] / T(2) || x > tv.L[d] / T(2)) atomicOr(bad, 1);
    }
}

// ------------------------------------------------------------ host driver

// Scratch device buffer from the library's cache, handed back on its stream
// (after the work enqueued there) when it goes out of scope.
struct DevBuf {
    hipStream_t s;
    void *p = nullptr;
    explicit DevBuf(hipStream_t stream) : s(stream) {}
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { dev_free(p, s); }
    int alloc(size_t bytes) {
        dev_free(p, s);
        p = nullptr;
        return dev_alloc(&p, bytes, s);
    }
    template <typename X>
    X *as() const { return (X *)p; }
};

// One pinned int64 per host thread for the per-level count read-back,
// allocated on first use and kept until the process ends.
int pinned_slot(int64_t **p) {
    static thread_local int64_t *slot = nullptr;
    if (!slot) CKD_HIP(hipHostMalloc((void **)&slot, sizeof(int64_t), hipHostMallocDefault));
    *p = slot;
    return CKD_OK;
}

// Waits until the copy of a count into the pinned slot *p (set to -1 before
// the copy was queued) has landed and returns the count in *out.  A value
// that a half-written slot could show (negative, or 2^32 - 1) keeps it
// polling; after 2^20 polls it synchronises the stream instead, which also
// reports an error of the queued work.
int wait_pinned(const int64_t *p, hipStream_t s, int64_t *out) {
    const volatile int64_t *v = p;
    for (int i = 0; i < (1 << 20); i++) {
        const int64_t a = *v;
        if (a >= 0 && a != 0xFFFFFFFFLL) {
            *out = a;
            return CKD_OK;
        }
    }
    CKD_HIP(hipStreamSynchronize(s));
    *out = *v;
    return CKD_OK;
}

template <typename T, int D>
int renumber_dfs(ckd_tree *t, const std::vector<int64_t> &lvl_start, hipStream_t s) {
    const int64_t nn = t->nnodes, nsets = t->nsets;
    const int nlev = (int)lvl_start.size() - 1;
    DevBuf desc(s), newid(s), pairbase(s), rdesc(s), roff(s), scr(s);
    int rc;
    if ((rc = desc.alloc(sizeof(int64_t) * nn)) || (rc = newid.alloc(sizeof(int64_t) * nn)) ||
        (rc = pairbase.alloc(sizeof(int64_t) * nn)) || (rc = rdesc.alloc(sizeof(int64_t) * nsets)) ||
        (rc = roff.alloc(sizeof(int64_t) * (nsets + 1))) || (rc = scr.alloc(scan_tmp_bytes(nsets))))
        return rc;
    const NodeRec<T> *nodes = (const NodeRec<T> *)t->nodes;
    for (int L = nlev - 1; L >= 0; L--) {  // subtree sizes, deepest level first
        const int64_t lo = lvl_start[L], hi = lvl_start[L + 1];
        if (hi > lo)
            k_subtree_size<T><<<(unsigned)ceil_div(hi - lo, 256), 256, 0, s>>>(nodes, lo, hi, desc.as<int64_t>());
    }
    k_root_desc<<<(unsigned)ceil_div(nsets, 256), 256, 0, s>>>(desc.as<int64_t>(), nsets, rdesc.as<int64_t>());
    if ((rc = scan_i64(rdesc.as<int64_t>(), roff.as<int64_t>(), nsets, true, scr.p, s))) return rc;
    k_root_ids<<<(unsigned)ceil_div(nsets, 256), 256, 0, s>>>(desc.as<int64_t>(), roff.as<int64_t>(), nsets,
                                                             newid.as<int64_t>(), pairbase.as<int64_t>());
    for (int L = 0; L < nlev; L++) {  // ids of the children, top level first
        const int64_t lo = lvl_start[L], hi = lvl_start[L + 1];
        if (hi > lo)
            k_child_ids<T><<<(unsigned)ceil_div(hi - lo, 256), 256, 0, s>>>(nodes, lo, hi, desc.as<int64_t>(),
                                                                            newid.as<int64_t>(),
                                                                            pairbase.as<int64_t>());
    }
    void *n2 = nullptr, *b2 = nullptr, *d2 = nullptr, *r2 = nullptr;
    if ((rc = dev_alloc(&n2, sizeof(NodeRec<T>) * nn, s)) || (rc = dev_alloc(&b2, sizeof(T) * 2 * D * nn, s)) ||
        (rc = dev_alloc(&d2, sizeof(T) * 2 * D * nn, s)) || (rc = dev_alloc(&r2, sizeof(Range) * nn, s))) {
        for (void *x : {n2, b2, d2, r2}) dev_free(x, s);
        return rc;
    }
    // pairbase is done with: it takes the inverse permutation
    k_invert_ids<<<(unsigned)ceil_div(nn, 256), 256, 0, s>>>(newid.as<int64_t>(), nn, pairbase.as<int64_t>());
    k_permute_nodes<T, D><<<(unsigned)ceil_div(nn, 256), 256, 0, s>>>(
        nodes, (const T *)t->bbox, (const T *)t->dbox, t->range, nn, newid.as<int64_t>(), pairbase.as<int64_t>(),
        (NodeRec<T> *)n2, (T *)b2, (T *)d2, (Range *)r2);
    CKD_HIP(hipGetLastError());
    // the level-order arrays go once the permutation has read them
    for (void *x : {t->nodes, t->bbox, t->dbox, (void *)t->range}) dev_free(x, s);
    t->nodes = n2;
    t->bbox = b2;
    t->dbox = d2;
    t->range = (Range *)r2;
    return CKD_OK;
}

// Wide records of the inner nodes, numbered by an exclusive scan of the
// inner-node flags in (depth-first) node order.  Needs row starts below 2^31.
template <typename T, int D>
int build_wide(ckd_tree *t, hipStream_t s) {
    const int64_t nn = t->nnodes, nsets = t->nsets;
    if (t->n >= INT32_MAX) return CKD_OK;  // stack traversals not offered
    DevBuf flag(s), widx(s), scr(s);
    int rc;
    if ((rc = flag.alloc(sizeof(int64_t) * nn)) || (rc = widx.alloc(sizeof(int64_t) * (nn + 1))) ||
        (rc = scr.alloc(scan_tmp_bytes(nn))))
        return rc;
    const NodeRec<T> *nodes = (const NodeRec<T> *)t->nodes;
    const unsigned g = (unsigned)ceil_div(nn, 256);
    k_inner_flags<T><<<g, 256, 0, s>>>(nodes, nn, flag.as<int64_t>());
    if ((rc = scan_i64(flag.as<int64_t>(), widx.as<int64_t>(), nn, true, scr.p, s))) return rc;
    int64_t ninner = 0;
    CKD_HIP(hipMemcpyAsync(&ninner, widx.as<int64_t>() + nn, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    CKD_HIP(hipStreamSynchronize(s));
    if ((rc = dev_alloc(&t->wide, wide_bytes<T, D>() * (size_t)(ninner > 0 ? ninner : 1), s)) ||
        (rc = dev_alloc((void **)&t->wroot, sizeof(int2) * (size_t)nsets, s)) ||
        (rc = dev_alloc((void **)&t->wrange, sizeof(int2) * (size_t)(ninner > 0 ? ninner : 1), s)))
        return rc;
    t->nwide = ninner;
    k_fill_wide<T, D><<<g, 256, 0, s>>>(nodes, (const T *)t->dbox, t->range, widx.as<int64_t>(), nn,
                                        (char *)t->wide, t->wrange);
    k_root_pairs<T><<<(unsigned)ceil_div(nsets, 256), 256, 0, s>>>(nodes, t->range, widx.as<int64_t>(), nsets,
                                                                    t->wroot);
    CKD_HIP(hipGetLastError());
    return wide_parents(t, s);
}

template <typename T, int D, typename IdxT>
int build_impl(ckd_tree *t, const T *in, hipStream_t s) {
    const int64_t n = t->n;
    const int leaf = t->leafsize;
    const TreeView<T> tvL = view<T>(t);
    struct Events {  // destroyed on every return
        hipEvent_t a = nullptr, b = nullptr;
        ~Events() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } ev;
    CKD_HIP(hipEventCreate(&ev.a));
    CKD_HIP(hipEventCreate(&ev.b));
    hipEvent_t ev0 = ev.a, ev1 = ev.b;
    CKD_HIP(hipEventRecord(ev0, s));

    if (t->periodic) {
        DevBuf bad(s);
        int rc = bad.alloc(sizeof(int));
        if (rc) return rc;
        CKD_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
        k_check_range<T, D><<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(in, n, tvL, bad.as<int>());
        int64_t *hbad = nullptr;
        if ((rc = pinned_slot(&hbad))) return rc;
        *hbad = 0;
        CKD_HIP(hipMemcpyAsync(hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
        CKD_HIP(hipStreamSynchronize(s));
        if (*hbad) {
            set_error("Points outside (-L/2, L/2) range !!");
            return CKD_ERR_RANGE;
        }
    }

    constexpr int PS = pstride(D);
    DevBuf P(s), P2(s), I(s), I2(s);
    int rc;
    if ((rc = P.alloc(sizeof(T) * D * n)) || (rc = P2.alloc(sizeof(T) * D * n)) ||
        (rc = I.alloc(sizeof(IdxT) * n)) || (rc = I2.alloc(sizeof(IdxT) * n)))
        return rc;
    // final arrays, owned by the tree
    if ((rc = dev_alloc(&t->pts, sizeof(T) * PS * (n + PTS_PAD), s))) return rc;
    CKD_HIP(hipMemsetAsync((T *)t->pts + (size_t)PS * n, 0, sizeof(T) * PS * PTS_PAD, s));
    if ((rc = dev_alloc(&t->idx, sizeof(IdxT) * n, s))) return rc;

    const int64_t nsets_ = t->nsets;
    int64_t cap = 4 * n / (leaf > 1 ? leaf : 1) + 64;
    if (cap > 2 * n + 1) cap = 2 * n + 1;
    if (cap < 2 * nsets_ + 64) cap = 2 * nsets_ + 64;
    auto alloc_nodes = [&](int64_t c) -> int {
        void *nn = nullptr, *nb = nullptr, *nd = nullptr, *nr = nullptr;
        int e;
        if ((e = dev_alloc(&nn, sizeof(NodeRec<T>) * c, s)) || (e = dev_alloc(&nb, sizeof(T) * 2 * D * c, s)) ||
            (e = dev_alloc(&nd, sizeof(T) * 2 * D * c, s)) || (e = dev_alloc(&nr, sizeof(Range) * c, s))) {
            for (void *x : {nn, nb, nd, nr}) dev_free(x, s);
            return e;
        }
        if (t->nodes) {
            CKD_HIP(hipMemcpyAsync(nn, t->nodes, sizeof(NodeRec<T>) * t->nnodes, hipMemcpyDeviceToDevice, s));
            CKD_HIP(hipMemcpyAsync(nb, t->bbox, sizeof(T) * 2 * D * t->nnodes, hipMemcpyDeviceToDevice, s));
            CKD_HIP(hipMemcpyAsync(nd, t->dbox, sizeof(T) * 2 * D * t->nnodes, hipMemcpyDeviceToDevice, s));
            CKD_HIP(hipMemcpyAsync(nr, t->range, sizeof(Range) * t->nnodes, hipMemcpyDeviceToDevice, s));
            for (void *x : {t->nodes, t->bbox, t->dbox, (void *)t->range}) dev_free(x, s);
        }
        t->nodes = nn;
        t->bbox = nb;
        t->dbox = nd;
        t->range = (Range *)nr;
        cap = c;
        return CKD_OK;
    };
    t->nnodes = 0;
    if ((rc = alloc_nodes(cap))) return rc;

    k_transpose_in<T, D><<<(unsigned)ceil_div(n, BT), BT, 0, s>>>(in, n, P.as<T>(), I.p, sizeof(IdxT) == 8);
    CKD_HIP(hipGetLastError());
    const int64_t nsets = t->nsets;
    t->depth = 0;
    {
        // per-level work arrays, sized for the largest possible level
        const int64_t nact_max = (n / (leaf + 1) + 1 > nsets ? n / (leaf + 1) + 1 : nsets);
        const int64_t nchunk_max = n / CH + nact_max + 1;
        DevBuf act(s), act2(s), nch(s), chunk_base(s), chunks(s), cstat(s), dec(s), cleft(s), cleft_off(s),
            child_active(s), child_pos(s), scratch(s);
        if ((rc = act.alloc(sizeof(int32_t) * nact_max)) || (rc = act2.alloc(sizeof(int32_t) * nact_max)) ||
            (rc = nch.alloc(sizeof(int64_t) * nact_max)) ||
            (rc = chunk_base.alloc(sizeof(int64_t) * (nact_max + 1))) ||
            (rc = chunks.alloc(sizeof(Chunk) * nchunk_max)) || (rc = cstat.alloc(sizeof(Ext<T>) * D * nchunk_max)) ||
            (rc = dec.alloc(sizeof(Decision<T>) * nact_max)) || (rc = cleft.alloc(sizeof(int64_t) * nchunk_max)) ||
            (rc = cleft_off.alloc(sizeof(int64_t) * (nchunk_max + 1))) ||
            (rc = child_active.alloc(sizeof(int64_t) * 2 * nact_max)) ||
            (rc = child_pos.alloc(sizeof(int64_t) * (2 * nact_max + 1))) ||
            (rc = scratch.alloc(scan_tmp_bytes(2 * nact_max + nchunk_max + 2))))
            return rc;
        // the children's extremes from the partition of the levels of large
        // nodes (fewer than n / (3 CH) + 2 nodes: 4 nodes < n / CH + nodes +
        // 1 chunk slots): chunk records of the right children beside cstat
        // (left), the level's chunk bases kept for the next level's
        // decisions, and each next active node's child slot
        const int64_t fuse_nodes = n / (3 * CH) + 2;
        DevBuf cstat_r(s), chunk_base2(s), nch2(s), slot_of(s);
        if ((rc = cstat_r.alloc(sizeof(Ext<T>) * D * (n / CH + fuse_nodes + 1))) ||
            (rc = chunk_base2.alloc(sizeof(int64_t) * (nact_max + 1))) || (rc = nch2.alloc(sizeof(int64_t) * nact_max)) ||
            (rc = slot_of.alloc(sizeof(int64_t) * 2 * fuse_nodes)))
            return rc;
        bool fused_in = false;  // this level's extremes are the previous level's chunk records
        T *Pc = P.as<T>(), *Pn = P2.as<T>();
        IdxT *Ic = I.as<IdxT>(), *In = I2.as<IdxT>();
        int32_t *actc = act.as<int32_t>(), *actn = act2.as<int32_t>();
        int64_t *h_nact = nullptr;
        if ((rc = pinned_slot(&h_nact))) return rc;
        // roots: one per point set; the active ones start the level loop
        k_init_roots<T, D><<<(unsigned)ceil_div(nsets, 256), 256, 0, s>>>(
            (NodeRec<T> *)t->nodes, (T *)t->bbox, t->range, t->set_off, nsets, tvL, t->periodic, leaf,
            child_active.as<int64_t>());
        k_root_leaves<T, D, IdxT><<<(unsigned)nsets, BT, 0, s>>>(Pc, Ic, n, t->set_off, child_active.as<int64_t>(),
                                                                 (T *)t->bbox, !t->periodic, (T *)t->pts,
                                                                 (IdxT *)t->idx);
        if ((rc = scan_i64(child_active.as<int64_t>(), child_pos.as<int64_t>(), nsets, true, scratch.p, s)))
            return rc;
        k_compact_active<<<(unsigned)ceil_div(nsets, 256), 256, 0, s>>>(
            child_active.as<int64_t>(), child_pos.as<int64_t>(), nsets, 0, actc);
        CKD_HIP(hipGetLastError());
        CKD_HIP(hipMemcpyAsync(h_nact, child_pos.as<int64_t>() + nsets, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        CKD_HIP(hipStreamSynchronize(s));
        int64_t nact = *h_nact;
        int32_t next_id = (int32_t)nsets;
        t->nnodes = nsets;
        int64_t nleaves = nsets - nact;
        int level = 0;
        std::vector<int64_t> lvl_start(1, 0);  // level L holds ids [lvl_start[L], lvl_start[L+1])
        lvl_start.push_back(nsets);
        while (nact > 0) {
            if ((int64_t)next_id + 2 * nact > cap) {
                int64_t nc = cap * 2;
                while (nc < (int64_t)next_id + 2 * nact) nc *= 2;
                if (nc > INT32_MAX) {
                    set_error("tree needs more than 2^31 nodes");
                    return CKD_ERR_LIMIT;
                }
                t->nnodes = next_id;
                if ((rc = alloc_nodes(nc))) return rc;
            }
            const unsigned ga = (unsigned)ceil_div(nact, 256);
            const int64_t nub = n / CH + nact + 1;  // upper bound of chunks this level
            // chunk bases alternate between two buffers: a fused level's
            // decisions read the previous level's
            int64_t *d_nch = (level & 1) ? nch2.as<int64_t>() : nch.as<int64_t>();
            int64_t *d_cb = (level & 1) ? chunk_base2.as<int64_t>() : chunk_base.as<int64_t>();
            const int64_t *p_nch = (level & 1) ? nch.as<int64_t>() : nch2.as<int64_t>();
            const int64_t *p_cb = (level & 1) ? chunk_base.as<int64_t>() : chunk_base2.as<int64_t>();
            k_chunk_count<<<ga, 256, 0, s>>>(actc, nact, t->range, d_nch);
            if ((rc = scan_i64(d_nch, d_cb, nact, true, scratch.p, s))) break;
            // levels of small nodes: at most a few chunks per node
            const bool small_nodes = nact * NODE_THREAD_MIN_CH >= nub;
            if (small_nodes)
                k_chunk_fill_t<<<ga, 256, 0, s>>>(actc, nact, t->range, d_cb, chunks.as<Chunk>());
            else
                k_chunk_fill<<<(unsigned)nact, 64, 0, s>>>(actc, t->range, d_cb, chunks.as<Chunk>());
            const int64_t *d_ntot = d_cb + nact;  // total chunks (device)
            const unsigned gch = (unsigned)ceil_div(nub, CPB);  // one wave per chunk slot
            if (fused_in) {  // the previous level's chunk records of each node's parent
                const int64_t *so = slot_of.as<int64_t>();
                if (nact <= DECIDE_BLOCK_MAX)
                    k_node_decide_b<T, D><<<(unsigned)nact, BT, 0, s>>>(
                        actc, nact, p_cb, p_nch, cstat.as<Ext<T>>(), (NodeRec<T> *)t->nodes, (T *)t->bbox,
                        (T *)t->dbox, dec.as<Decision<T>>(), level == 0 && !t->periodic, so, cstat_r.as<Ext<T>>());
                else
                    k_node_decide<T, D><<<(unsigned)ceil_div(nact, CPB), BT, 0, s>>>(
                        actc, nact, p_cb, p_nch, cstat.as<Ext<T>>(), (NodeRec<T> *)t->nodes, (T *)t->bbox,
                        (T *)t->dbox, dec.as<Decision<T>>(), level == 0 && !t->periodic, so, cstat_r.as<Ext<T>>());
            } else {
                if (small_nodes)  // 16 lanes per chunk
                    k_chunk_stats_g<T, D, STATS_G><<<(unsigned)ceil_div(nub * STATS_G, BT), BT, 0, s>>>(
                        chunks.as<Chunk>(), d_ntot, Pc, n, cstat.as<Ext<T>>());
                else
                    k_chunk_stats<T, D><<<gch, BT, 0, s>>>(chunks.as<Chunk>(), d_ntot, Pc, n, cstat.as<Ext<T>>());
                if (small_nodes && nact >= DECIDE_THREAD_MIN)  // one thread per node
                    k_node_decide_t<T, D><<<(unsigned)ceil_div(nact, BT), BT, 0, s>>>(
                        actc, nact, d_cb, d_nch, cstat.as<Ext<T>>(), (NodeRec<T> *)t->nodes, (T *)t->bbox, (T *)t->dbox,
                        dec.as<Decision<T>>(), level == 0 && !t->periodic);
                else if (nact <= DECIDE_BLOCK_MAX)  // a workgroup per node
                    k_node_decide_b<T, D><<<(unsigned)nact, BT, 0, s>>>(
                        actc, nact, d_cb, d_nch, cstat.as<Ext<T>>(), (NodeRec<T> *)t->nodes, (T *)t->bbox, (T *)t->dbox,
                        dec.as<Decision<T>>(), level == 0 && !t->periodic);
                else
                    k_node_decide<T, D><<<(unsigned)ceil_div(nact, CPB), BT, 0, s>>>(
                        actc, nact, d_cb, d_nch, cstat.as<Ext<T>>(), (NodeRec<T> *)t->nodes, (T *)t->bbox, (T *)t->dbox,
                        dec.as<Decision<T>>(), level == 0 && !t->periodic);
            }
            k_chunk_countleft<T><<<gch, BT, 0, s>>>(chunks.as<Chunk>(), d_ntot, nub, Pc, n,
                                                               dec.as<Decision<T>>(), cleft.as<int64_t>());
            if ((rc = scan_i64(cleft.as<int64_t>(), cleft_off.as<int64_t>(), nub, true, scratch.p, s))) break;
            // levels of large nodes: the partition also records each chunk's
            // extremes of both children (the next level's extremes pass)
            const bool fuse_out = !small_nodes && nact <= fuse_nodes;
            if (fuse_out)
                k_chunk_scatter<T, D, IdxT, true><<<gch, BT, 0, s>>>(
                    chunks.as<Chunk>(), d_ntot, d_cb, d_nch, actc, t->range, Pc, Ic, Pn, In, n,
                    dec.as<Decision<T>>(), cleft_off.as<int64_t>(), cstat.as<Ext<T>>(), cstat_r.as<Ext<T>>());
            else
                k_chunk_scatter<T, D, IdxT><<<gch, BT, 0, s>>>(
                    chunks.as<Chunk>(), d_ntot, d_cb, d_nch, actc, t->range, Pc, Ic, Pn, In, n,
                    dec.as<Decision<T>>(), cleft_off.as<int64_t>());
            k_make_children<T, D><<<ga, 256, 0, s>>>(actc, nact, dec.as<Decision<T>>(), d_cb, d_nch,
                                                     cleft_off.as<int64_t>(), (NodeRec<T> *)t->nodes, (T *)t->bbox,
                                                     t->range, next_id, leaf, child_active.as<int64_t>());
            const int64_t m = 2 * nact;
            if ((rc = scan_i64(child_active.as<int64_t>(), child_pos.as<int64_t>(), m, true, scratch.p, s))) break;
            k_compact_active<<<(unsigned)ceil_div(m, 256), 256, 0, s>>>(
                child_active.as<int64_t>(), child_pos.as<int64_t>(), m, next_id, actn,
                fuse_out ? slot_of.as<int64_t>() : nullptr);
            fused_in = fuse_out;
            k_copy_leaves<T, D, IdxT><<<(unsigned)ceil_div(m * leaf, 256), 256, 0, s>>>(
                child_active.as<int64_t>(), m, next_id, t->range, leaf, Pn, In, n, (T *)t->pts, (IdxT *)t->idx);
            CKD_HIP(hipGetLastError());
            // the next level's node count: the copy lands in pinned memory,
            // which the host polls (a stream synchronisation per level cost
            // more); past a bound on the polls it synchronises after all
            *(volatile int64_t *)h_nact = -1;
            CKD_HIP(hipMemcpyAsync(h_nact, child_pos.as<int64_t>() + m, sizeof(int64_t), hipMemcpyDeviceToHost, s));
            int64_t nact_next = 0;
            if ((rc = wait_pinned(h_nact, s, &nact_next))) return rc;
            nleaves += m - nact_next;
            next_id += (int32_t)m;
            lvl_start.push_back(next_id);
            nact = nact_next;
            std::swap(Pc, Pn);
            std::swap(Ic, In);
            std::swap(actc, actn);
            level++;
        }
        if (rc) return rc;
        t->nnodes = next_id;
        t->nleaves = nleaves;
        t->depth = level;
        k_leaf_boxes<T, D><<<(unsigned)ceil_div(t->nnodes, 256), 256, 0, s>>>(
            (const NodeRec<T> *)t->nodes, t->range, t->nnodes, (const T *)t->pts, (T *)t->dbox);
        CKD_HIP(hipGetLastError());
        if ((rc = renumber_dfs<T, D>(t, lvl_start, s))) return rc;
        if ((rc = build_wide<T, D>(t, s))) return rc;
    }
    CKD_HIP(hipEventRecord(ev1, s));
    CKD_HIP(hipEventSynchronize(ev1));
    float ms = 0;
    CKD_HIP(hipEventElapsedTime(&ms, ev0, ev1));
    t->build_ms = ms;
    return CKD_OK;
}

template <typename T, typename IdxT>
int build_dispatch(ckd_tree *t, const void *in, hipStream_t s) {
    switch (t->ndim) {
        case 1: return build_impl<T, 1, IdxT>(t, (const T *)in, s);
        case 2: return build_impl<T, 2, IdxT>(t, (const T *)in, s);
        case 3: return build_impl<T, 3, IdxT>(t, (const T *)in, s);
        case 4: return build_impl<T, 4, IdxT>(t, (const T *)in, s);
        case 5: return build_impl<T, 5, IdxT>(t, (const T *)in, s);
        case 6: return build_impl<T, 6, IdxT>(t, (const T *)in, s);
        case 7: return build_impl<T, 7, IdxT>(t, (const T *)in, s);
        case 8: return build_impl<T, 8, IdxT>(t, (const T *)in, s);
    }
    set_error("ndim must be in [1, %d], got %d", CKD_MAX_DIM, t->ndim);
    return CKD_ERR_ARG;
}

}  // namespace

int wide_parents(ckd_tree *t, hipStream_t s) {
    if (!t->wide || t->nwide <= 0) return CKD_OK;
    dev_free(t->wparent, s);
    t->wparent = nullptr;
    int rc = dev_alloc((void **)&t->wparent, sizeof(int32_t) * (size_t)t->nwide, s);
    if (rc) return rc;
    const unsigned g = (unsigned)ceil_div(t->nwide, 256);
    const size_t rec = wide_bytes_rt(t->ndim, t->prec), pair_off = rec - 16;
    k_set_i32<<<g, 256, 0, s>>>(t->wparent, t->nwide, -1);
    k_wide_parents<<<g, 256, 0, s>>>((const char *)t->wide, rec, pair_off, t->nwide, t->wparent);
    CKD_HIP(hipGetLastError());
    return CKD_OK;
}

int build_tree(ckd_tree *t, const void *points, hipStream_t s) {
    if (t->prec == 4)
        return t->idx_bits == 64 ? build_dispatch<float, int64_t>(t, points, s)
                                 : build_dispatch<float, int32_t>(t, points, s);
    return t->idx_bits == 64 ? build_dispatch<double, int64_t>(t, points, s)
                             : build_dispatch<double, int32_t>(t, points, s);
}

}  // namespace ckd
