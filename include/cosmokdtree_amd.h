/*
 * cosmokdtree_amd.h -- C ABI of the MI355X (gfx950) k-d tree library.
 *
 * The library replaces the f2py extension module `pycosmokdtree` of the
 * reference (src/cosmokdtree2py.f90) underneath the Python API of
 * src/cosmokdtree/kdtree.py.  Every entry point below names the reference
 * routine it replaces.
 *
 * Conventions
 *  - All array arguments marked [dev] are device pointers (HBM) on the
 *    current HIP device; [host] arguments are host pointers.  `stream` is a
 *    hipStream_t (NULL = the default stream); work is enqueued on it and
 *    the call returns when the work is enqueued, except where stated.
 *  - Coordinates are row-major (n, ndim) in the tree precision: `prec` = 4
 *    (float) or 8 (double).  The reference Python path is always 8
 *    (src/Makefile:22-25); 4 reproduces a DOUBLEPRECISION=0 build.
 *  - Indices returned to the caller are int64 and 0-based, i.e. what
 *    src/cosmokdtree/kdtree.py returns after its `idx -= 1`.  An empty kNN
 *    slot (k > n) has index CKD_EMPTY_SLOT (-2, the value kdtree.py:52 turns
 *    the Fortran sentinel -1 into) and distance sqrt(FLT_MAX) in `prec`
 *    (cosmokdtree.f90:517-518,538).
 *  - Periodic trees take `boxsize` (ndim entries, each > 0); every point
 *    must lie in [-L/2, L/2] per axis (cosmokdtree.f90:176-190).  Distances
 *    then use the minimum image per axis (cosmokdtree.f90:1453-1457).
 *  - Every function returns CKD_OK (0) or a negative error code; the text
 *    of the last error on the calling thread is given by ckd_last_error().
 *    The reference STOPs the process instead (cosmokdtree.f90:161,170,190;
 *    cosmokdtree2py.f90:58,276).
 */
#ifndef COSMOKDTREE_AMD_H
#define COSMOKDTREE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CKD_OK 0
#define CKD_ERR_ARG -1     /* invalid argument (ndim, leaf, k, radius, sizes) */
#define CKD_ERR_RANGE -2   /* periodic point outside [-L/2, L/2] */
#define CKD_ERR_HIP -3     /* HIP runtime error */
#define CKD_ERR_MEMORY -4  /* device allocation failed */
#define CKD_ERR_LIMIT -5   /* a documented size limit was exceeded */

#define CKD_EMPTY_SLOT (-2)
#define CKD_MAX_DIM 8

typedef struct ckd_tree ckd_tree;

typedef struct {
    int64_t n;          /* number of points */
    int32_t ndim;       /* dimensions */
    int32_t prec;       /* 4 or 8 */
    int32_t leafsize;   /* maximum points per leaf */
    int32_t periodic;   /* 1 if built with a boxsize */
    int32_t idx_bits;   /* stored original-index width, 32 or 64 */
    int32_t depth;      /* levels below the root (a single leaf has 0) */
    int64_t nnodes;     /* nodes in the flat node array */
    int64_t nleaves;
    double boxsize[CKD_MAX_DIM];
    double build_ms;    /* device time of the last build, milliseconds */
    int64_t nsets;      /* independent point sets (1 unless built by ckd_build_forest) */
} ckd_tree_info;

/*
 * Build a tree over `points` [dev, (n, ndim) row-major, prec].
 * Replaces pybuild_kdtree (cosmokdtree2py.f90:29-74) -> build_kdtree
 * (cosmokdtree.f90:124-238): sliding-midpoint rule, leaves of at most
 * `leaf` points (default 16 in kdtree.py:19-20), out of place (the input
 * is not modified).  boxsize [host] = NULL builds a non-periodic tree.
 * idx_bits = 32 or 64 selects the stored index width (0 = 32 when n < 2^31).
 * The call synchronises `stream` (the level loop is host driven).
 */
int ckd_build(const void *points, int64_t n, int32_t ndim, int32_t prec, int32_t leaf, const double *boxsize,
              int32_t idx_bits, void *stream, ckd_tree **out);

/*
 * Build one tree per point set, all sets in one level-synchronous pass (a
 * forest).  Set r holds rows [set_offsets[r], set_offsets[r+1]) of `points`
 * (set_offsets [host, nsets+1], starting at 0, strictly increasing); its
 * root is node r and its trees' results use indices relative to the set.
 * All sets share `boxsize` (NULL = non-periodic).  Each set's tree is the
 * tree ckd_build would build from that set alone.  No reference counterpart
 * (the reference builds one tree per call, cosmokdtree.f90:124-238); it
 * serves the batched query engine (ckd_batch_*).
 */
int ckd_build_forest(const void *points, const int64_t *set_offsets, int64_t nsets, int32_t ndim, int32_t prec,
                     int32_t leaf, const double *boxsize, int32_t idx_bits, void *stream, ckd_tree **out);

/* Release a tree.  Replaces pydeallocate_tree (cosmokdtree2py.f90:80-97). */
int ckd_free(ckd_tree *tree);

int ckd_get_info(const ckd_tree *tree, ckd_tree_info *info);

/*
 * k nearest neighbours of nq query rows `queries` [dev, (nq, ndim), prec].
 * Replaces pyknn_search (cosmokdtree2py.f90:103-143) -> knn_search
 * (cosmokdtree.f90:504-540).  Outputs [dev]: dist (nq, k) in prec, idx
 * (nq, k) int64.  sorted != 0 orders each row by (distance, index); with
 * sorted == 0 the row order is unspecified (the reference returns its heap
 * layout).  The distance of each neighbour is computed bit-for-bit as the
 * reference's distance() (cosmokdtree.f90:1441-1465).
 * seed_r [dev, nq floats] or NULL: a per-query search radius proposal; a
 * query whose seeded search finds fewer than k points is searched again
 * (with a larger seed, then without), so results never depend on the seed.
 * stats [dev, 6 int64] or NULL: accumulates {nodes visited, leaf points
 * tested, queries re-run after a short seed, queries, candidates inserted
 * into a result list, parent steps climbed or deferred subtrees popped}.
 * The wave kernels count per wave: nodes visited and pops are wave totals
 * summed over the lanes.
 * A large batch is walked in the order of the leaves its queries fall in.
 * When nq equals the tree's point count (one point set), the library first
 * checks on the device whether query idx[r] is the tree's row r, for every
 * 16th row r (the self-query of the tree's own points, in any array); if
 * so the tree's row order is that ordering and the locate pass is skipped
 * (for the walk shared per wave; the walk per query keeps the locate pass).
 * Results do not depend on which ordering ran.
 */
int ckd_knn(const ckd_tree *tree, const void *queries, int64_t nq, int32_t k, int32_t sorted, void *dist,
            int64_t *idx, const float *seed_r, int64_t *stats, void *stream);

/*
 * SPH smoothing length and density of nq query rows [dev, (nq, 3), prec],
 * fused into the kNN traversal: no kNN rows are written.  No counterpart in
 * the reference library; it is the step its callers take next with the
 * k = 32 neighbours (the cosmology use named in README.md:18), in one
 * kernel instead of knn_search (cosmokdtree.f90:504-540) followed by a pass
 * over the (nq, k) rows.  For each query, with r_1 <= ... <= r_k the
 * distances ckd_knn returns:
 *   h   = r_k / 2  (the cubic spline's support 2h reaches the k-th point)
 *   rho = sum_j m_j W(r_j, h),  W(r, h) = w(r / h) / (pi h^3),
 *   w(u) = 1 - 1.5 u^2 + 0.75 u^3 (u < 1), 0.25 (2 - u)^3 (u < 2), else 0
 * (Monaghan & Lattanzio 1985); a query that is a tree point counts itself.
 * mass [dev, n, prec] indexed by original point index, or NULL for unit
 * masses.  Outputs h, rho [dev, nq, prec].  Needs a 3-D tree of depth <= 64
 * and 2 <= k <= min(64, n); k coincident points give h = 0 and a
 * non-finite rho.  h equals half of ckd_knn's k-th distance bit for bit;
 * rho is summed in order of increasing distance.
 */
int ckd_sph_density(const ckd_tree *tree, const void *queries, int64_t nq, int32_t k, const void *mass, void *h,
                    void *rho, void *stream);

/*
 * kNN that allocates nothing: the same arguments, kernels and results as
 * ckd_knn, with the call's scratch (the leaf ordering of the queries, the
 * re-search lists of seeded calls) taken from one static device array of
 * CKD_NOALLOC_SCRATCH_BYTES per GPU instead of the library's allocator, and
 * no events created (ckd_knn_timing does not see its launches).  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Limits: k <= 64, trees below 2^31 - 1
 * points, and a batch whose scratch (ckd_knn_noalloc_scratch_bytes: about
 * 12 B per query plus 16 B per tree node, when there are at least 4096
 * queries; seeds add about 36 B per query) fits the array; anything beyond
 * returns CKD_ERR_LIMIT before any work is enqueued.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams or
 * host threads) race and must be serialised by the caller.
 */
#define CKD_NOALLOC_SCRATCH_BYTES (64 << 20)
int ckd_knn_noalloc(const ckd_tree *tree, const void *queries, int64_t nq, int32_t k, int32_t sorted, void *dist,
                    int64_t *idx, const float *seed_r, int64_t *stats, void *stream);
/* Scratch bytes ckd_knn_noalloc needs for nq queries (seeded != 0: with seeds). */
size_t ckd_knn_noalloc_scratch_bytes(const ckd_tree *tree, int64_t nq, int32_t seeded);

#define CKD_KNN_AUTO (-1)
#define CKD_KNN_LANE 0   /* one traversal per query */
#define CKD_KNN_WAVE 1   /* the queries of a wave share one traversal (parent pointers, any depth) */
#define CKD_KNN_STACK 2  /* shared traversal over wide records with an LDS stack (ndim <= 3) */

/*
 * Choose the kNN kernel of one tree (k <= 64).  No reference counterpart:
 * the reference has a single CPU traversal.  kernel is one of CKD_KNN_*;
 * stack_depth is 0 (automatic), 32 or 64, the LDS stack entries per wave of
 * CKD_KNN_STACK, which is used only when ndim <= 3 and the tree depth fits
 * the stack (otherwise CKD_KNN_WAVE runs); queries_per_wave is 0 (= 64), 1,
 * 4, 16 or 64 (any power of two up to 64): how many queries share one wave
 * of CKD_KNN_STACK (each then held by 64 / queries_per_wave lanes that split
 * the leaf rows).  With CKD_KNN_AUTO, a call of at most 65536 queries with
 * k <= 32 and fewer than n/4 queries runs one query per wave; otherwise a
 * call with at least n/20 queries, or k > 32, shares walks 64 per wave; a
 * larger sparse call with k <= 32 runs one walk per query.  Every choice
 * returns identical results; only the speed differs.  Used by the
 * autotuner.
 */
int ckd_knn_config(ckd_tree *tree, int32_t kernel, int32_t stack_depth, int32_t queries_per_wave);

/*
 * Launch timing of the kNN traversal kernel of one tree (no reference
 * counterpart; for benchmarks and roofline figures).  on != 0 records a
 * pair of HIP events, on the call's stream, around every traversal kernel
 * that ckd_knn launches for this tree (the main kernel only: not the query
 * ordering nor the retry of short seeds).  Each call first reports and
 * forgets the launches recorded so far -- *total_ms [host] their summed
 * duration, *launches [host] their count (it synchronises those events;
 * either pointer may be NULL) -- then applies `on`.
 */
int ckd_knn_timing(ckd_tree *tree, int32_t on, double *total_ms, int64_t *launches);

/*
 * Ball search, first pass: offsets [dev, nq+1 int64] receives the CSR row
 * offsets of the points with squared distance <= radius^2 (radius is cast
 * to prec first, cosmokdtree.f90:1019,1126).  *total [host] receives
 * offsets[nq] (the call synchronises `stream` to read it).
 * Replaces pyball_search_call_1 (cosmokdtree2py.f90:153-201).
 */
int ckd_ball_count(const ckd_tree *tree, const void *queries, int64_t nq, double radius, int64_t *offsets,
                   int64_t *total, void *stream);

/*
 * Ball search, second pass: fills dist [dev, total, prec] and idx
 * [dev, total int64] at the offsets of ckd_ball_count.  sorted != 0 orders
 * each row by (distance, index).  Replaces pyball_search_call_2
 * (cosmokdtree2py.f90:203-246) with CSR output instead of a zero-padded
 * (ntar, nballmax) matrix.
 */
int ckd_ball_fill(const ckd_tree *tree, const void *queries, int64_t nq, double radius, int32_t sorted,
                  const int64_t *offsets, void *dist, int64_t *idx, void *stream);

/*
 * Ball search in one walk: the two calls above (pyball_search_call_1 and
 * _2, cosmokdtree2py.f90:153-246) as one, where the rows fit.  Writes
 * offsets [dev, nq+1] and *total [host] as ckd_ball_count does and, when
 * *total <= capacity, the rows to dist [dev, capacity, prec] and idx
 * [dev, capacity int64] as ckd_ball_fill does (sorted != 0: each row by
 * (distance, index)).  When *total > capacity the rows are not written and
 * the call still returns CKD_OK: ckd_ball_fill with buffers of *total rows
 * and these offsets completes it (a second walk).  The walk stages its rows
 * in library scratch of 16 bytes per row of capacity plus 8 KiB per 64
 * queries; the call synchronises `stream`.
 */
int ckd_ball_search(const ckd_tree *tree, const void *queries, int64_t nq, double radius, int32_t sorted,
                    int64_t *offsets, int64_t capacity, void *dist, int64_t *idx, int64_t *total, void *stream);

/*
 * Box search over nb boxes [dev, (nb, 2*ndim) = (xmin, xmax, ymin, ...)],
 * first pass (CSR offsets + host total).  Replaces pybox_search_call_1
 * (cosmokdtree2py.f90:252-287) -> box_search (cosmokdtree.f90:1220-1253),
 * with the reference's inclusive point test and strict child-visit tests
 * (cosmokdtree.f90:1296-1302,1309-1325).  No periodic wrap, as in the
 * reference (cosmokdtree.f90:1260); ckd_box_count_periodic wraps.
 */
int ckd_box_count(const ckd_tree *tree, const void *boxes, int64_t nb, int64_t *offsets, int64_t *total,
                  void *stream);

/* Box search second pass.  Replaces pybox_search_call_2
   (cosmokdtree2py.f90:289-315). */
int ckd_box_fill(const ckd_tree *tree, const void *boxes, int64_t nb, const int64_t *offsets, int64_t *idx,
                 void *stream);

/*
 * Box search in a periodic tree with boxes taken modulo the box: a box may
 * straddle the faces or lie outside [-L/2, L/2].  No reference counterpart
 * (box_search does not wrap, cosmokdtree.f90:1260); the same two passes and
 * arguments as ckd_box_count / ckd_box_fill.  The rule: on each axis the
 * interval [lo, hi] (in prec) becomes the pieces [lo - kL, hi - kL] clipped
 * to [-L/2, L/2] for k = m+1, m, m-1, m = floor((lo + L/2) / L) evaluated
 * in double, with pieces that touch merged; the whole axis when
 * hi - lo >= L; nothing when lo > hi.  Every sub-box of the product of the
 * pieces (axis 0 varying fastest) is searched with the reference's rules
 * (inclusive point test, strict child visits, cosmokdtree.f90:1296-1325)
 * and its rows follow the previous sub-box's.  The pieces are disjoint, so
 * a row holds each point at most once.  CKD_ERR_ARG for a non-periodic tree.
 */
int ckd_box_count_periodic(const ckd_tree *tree, const void *boxes, int64_t nb, int64_t *offsets, int64_t *total,
                           void *stream);
int ckd_box_fill_periodic(const ckd_tree *tree, const void *boxes, int64_t nb, const int64_t *offsets, int64_t *idx,
                          void *stream);

/*
 * Copy the tree structure to host arrays (any pointer may be NULL), in the
 * library's node numbering (root = 0, children of node i at child[i] and
 * child[i] + 1).  Sizes: nnodes for node arrays, nnodes*ndim for lo/hi,
 * n for points (n*ndim) and idx.  Leaf points of node i occupy rows
 * [start[i], start[i] + count[i]) of points/idx.  No reference counterpart
 * beyond the node fields of KDTreeNode (cosmokdtree.f90:93-105); used by
 * tests and by persistence.
 */
int ckd_tree_dump(const ckd_tree *tree, int32_t *axis, void *split, void *lo, void *hi, int32_t *child,
                  int32_t *parent, int64_t *start, int64_t *count, void *points, int64_t *idx);

/*
 * Save a built tree (or forest) to a file [host path] and load it back on
 * the current device without rebuilding: a fixed header, then every device
 * array as it lies in HBM.  A loaded tree answers every query exactly as
 * the saved one.  No reference counterpart: the reference's tree handle is
 * an opaque pointer (cosmokdtree2py.f90:70-73) that cannot be persisted.
 */
int ckd_tree_save(const ckd_tree *tree, const char *path);
int ckd_tree_load(const char *path, void *stream, ckd_tree **out);

/*
 * Batched queries over a forest (ckd_build_forest), drained by persistent
 * kernels from a device work queue.  No reference counterpart (the
 * reference answers one tree per call); each set's results equal the
 * single-tree calls on that set's tree.
 *
 * ckd_batch_plan: set r has queries qoff[r]..qoff[r+1] [host, nsets+1] of the
 * query array, k[r] neighbours (0..64, 0 = no kNN) and ball radius
 * radius[r] (radius NULL or radius[r] <= 0 = no ball search).  kNN rows of
 * set r start at element sum_{s<r} nq_s*k_s of the kNN outputs
 * (*knn_total = total elements).  threads = queries per work item and
 * workgroup (64, 128 or 256); blocks = persistent grid size (0 = 2048
 * threads per CU).  Ball output holds at most ball_capacity entries.
 *
 * ckd_batch_run only enqueues kernels (no allocation, copy or wait), so it
 * can be captured in a hipGraph.  Outputs [dev]: knn_dist/knn_idx
 * (knn_total), ball_off (nq+1, CSR over all queries), ball_dist/ball_idx
 * (ball_capacity).  Indices are relative to each set.  Ball rows are
 * unsorted; kNN rows follow `sorted`.
 *
 * ckd_batch_status: waits for `stream`, then reports the ball total and
 * whether it exceeded ball_capacity (rows past the capacity are not
 * written).  It returns CKD_ERR_HIP if a ball walk stopped at its step
 * bound, which only an inconsistent tree can cause.
 */
typedef struct ckd_batch ckd_batch;
int ckd_batch_plan(const ckd_tree *forest, const int64_t *qoff, const int32_t *k, const double *radius,
                   int32_t sorted, int32_t threads, int32_t blocks, int64_t ball_capacity, ckd_batch **out,
                   int64_t *knn_total);
/* kNN kernel of a plan: CKD_KNN_STACK (the default when the queries are at
   least 1/20 as many as the points; queries sorted by leaf
   within each set, queries_per_wave in {1, 4, 16, 64} share a wave's walk,
   LDS stack of stack_depth = 32 or 64 entries, 0 = the smallest that holds
   the forest's depth; a stack shorter than the depth is never used) or
   CKD_KNN_LANE (one walk per query; the default for sparser batches).
   Results never depend on it. */
int ckd_batch_config(ckd_batch *plan, int32_t knn_kernel, int32_t queries_per_wave, int32_t stack_depth);
/* Ball kernel of a plan: CKD_KNN_STACK (the 64 queries of a wave share one
   walk over the set's wide records, as ckd_ball_count/fill do; used when
   the forest has wide records and its depth fits the stack; the default
   when the queries are at least 1/20 as many as the points) or
   CKD_KNN_LANE (one walk per query; the default for sparser batches).
   Results never depend on it; row order within a query may differ. */
int ckd_batch_config_ball(ckd_batch *plan, int32_t ball_kernel);
int ckd_batch_run(ckd_batch *plan, const void *queries, void *knn_dist, int64_t *knn_idx, int64_t *ball_off,
                  void *ball_dist, int64_t *ball_idx, void *stream);
int ckd_batch_status(const ckd_batch *plan, const int64_t *ball_off, int64_t *ball_total, int32_t *overflow,
                     void *stream);
int ckd_batch_free(ckd_batch *plan);

/*
 * Learned search-radius proposal (no reference counterpart: it only seeds
 * the kNN pruning bound through ckd_knn's seed_r, and a short seed falls
 * back to the plain search).  Network 32 -> 64 -> 64 -> 1 with GELU, bf16
 * operands on MFMA, fp32 accumulation; parameters are one fp32 vector of
 * ckd_mlp_nparam() values: W1[32][64], b1[64], W2[64][64], b2[64], w3[64], b3.
 */
int ckd_mlp_nparam(void);

/* Point counts of a periodic 3-D tree in a G^3 grid over [-L/2, L/2)^3:
   grid [dev, G^3 int32], cell (x, y, z) at (z*G + y)*G + x. */
int ckd_occupancy(const ckd_tree *tree, int32_t G, int32_t *grid, void *stream);

/* bf16 log1p of every count of a G^3 grid: table [dev, G^3 uint16]. */
int ckd_mlp_table(const int32_t *grid, int32_t G, uint16_t *table, void *stream);

/* Features of nq queries [dev, (nq, 3), tree precision] against a periodic
   3-D tree into X [dev, (nq, 32) bf16]: (q + L/2)/L per axis; the 27
   log1p counts of the query's 3x3x3 periodic grid neighbourhood (dz, dy, dx
   in -1..1, dz-major), looked up in `table` (ckd_mlp_table); then the local
   density of the tree around the query: (1/3) ln(count / volume) of the
   tight point box of the smallest ancestor of the query's leaf holding at
   least 16 points, then at least 64 points (each box extent at least
   1e-6 L).  leaf_of [dev, nq int32] or NULL: the query's leaf, when known
   (otherwise found by descent). */
int ckd_mlp_features(const ckd_tree *tree, const void *queries, int64_t nq, int32_t G, const uint16_t *table,
                     const int32_t *leaf_of, uint16_t *X, void *stream);

/* Forward pass: mode 0 writes y = predicted log radius, mode 1 writes
   scale * exp(y) (a seed radius for ckd_knn).  out [dev, n float]. */
int ckd_mlp_forward(const float *params, const uint16_t *X, int64_t n, float scale, int32_t mode, float *out,
                    void *stream);

/*
 * kNN with learned seeds, in one call: the queries' leaves (found once, for
 * both the query ordering and the tree features), the features
 * (ckd_mlp_features with `table`, G), the network's seed radius
 * margin * exp(y) per query (ckd_mlp_forward), then the seeded traversal.
 * Same outputs and guarantee as ckd_knn: a query whose seed holds fewer than
 * k points is searched again (first with a larger seed, then without), so
 * the results never depend on the network.  Periodic 3-D trees, k <= 64.
 * No reference counterpart (the reference has no learned component).
 */
int ckd_knn_learned(const ckd_tree *tree, const void *queries, int64_t nq, int32_t k, int32_t sorted,
                    const float *params, float margin, int32_t G, const uint16_t *table, void *dist, int64_t *idx,
                    int64_t *stats, void *stream);

/* Mean-squared-error gradient over B rows against target [dev, B float]:
   grad [dev, nparam float] and *loss [dev] (sum of squared errors) are
   overwritten. */
int ckd_mlp_grad(const float *params, const uint16_t *X, const float *target, int64_t B, float *grad, float *loss,
                 void *stream);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected for
   `step` >= 1) on params with moment buffers m, v [dev, nparam float]. */
int ckd_mlp_adam(float *params, const float *grad, float *m, float *v, int64_t step, float lr, void *stream);

/*
 * Device memory: every device array of a tree, and the scratch of builds
 * and queries, comes from a caching allocator (one free list per device)
 * that keeps freed blocks mapped, so that repeated builds and queries take
 * their memory without hipMalloc/hipFree (page mapping, device-wide
 * synchronisation).  This call waits for the work on every device the
 * library has used and returns the cached blocks to the driver (memory of
 * live trees stays).  An allocation that finds the device full does the
 * same for its device before it fails.  No reference counterpart (the
 * reference allocates host arrays with Fortran ALLOCATE,
 * cosmokdtree.f90:151-153,206-211).
 */
int ckd_release_memory(void);

/* Text of the last error on this thread ("" if none). */
const char *ckd_last_error(void);

/* Library version string. */
const char *ckd_version(void);

#ifdef __cplusplus
}
#endif
#endif /* COSMOKDTREE_AMD_H */
