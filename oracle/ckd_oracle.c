/*
 * ckd_oracle.c -- CPU oracle for the cosmokdtree build and query path.
 *
 * TEST INFRASTRUCTURE ONLY: this library is the checker the GPU library is
 * compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it; the product (cosmokdtree_amd) never
 * does.
 *
 * It restates, routine by routine, the Fortran reference
 * (src/cosmokdtree.f90 and src/cosmokdtree2py.f90, cited per function in
 * ckd_oracle_impl.h): the out-of-place sliding-midpoint build with the
 * Lomuto partition and first-occurrence minloc/maxloc, best-first kNN
 * (non-periodic) and hyperplane kNN (periodic), bbox/hyperplane ball search
 * with whole-subtree reduction, strict-visit box search and the
 * median-of-three quicksort.  Outputs follow src/cosmokdtree/kdtree.py:
 * 0-based indices (an empty kNN slot is -2), distances after sqrt.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp -fPIC -shared (oracle/Makefile).
 * -ffp-contract=off is required: it makes every a + b*c round twice, as the
 * Fortran source is written.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define CKDO_MAXD 8

#define REAL float
#define SFX f32
#define SQRT sqrtf
#include "ckd_oracle_impl.h"
#undef REAL
#undef SFX
#undef SQRT

#define REAL double
#define SFX f64
#define SQRT sqrt
#include "ckd_oracle_impl.h"
#undef REAL
#undef SFX
#undef SQRT

/* ---- exported C API (ctypes); prec = 4 or 8 selects the instantiation ---- */

typedef struct {
    int prec;
    void *t;
} ckdo_handle;

void *ckdo_build(const void *pts, int64_t n, int ndim, int prec, int leaf, const void *L, int *err) {
    ckdo_handle *h;
    void *t;
    if (prec == 4) t = build_f32((const float *)pts, n, ndim, leaf, (const float *)L, err);
    else t = build_f64((const double *)pts, n, ndim, leaf, (const double *)L, err);
    if (!t) return NULL;
    h = (ckdo_handle *)malloc(sizeof(ckdo_handle));
    h->prec = prec;
    h->t = t;
    return h;
}

void ckdo_free(void *hv) {
    ckdo_handle *h = (ckdo_handle *)hv;
    if (!h) return;
    if (h->prec == 4) tree_free_f32((otree_f32 *)h->t);
    else tree_free_f64((otree_f64 *)h->t);
    free(h);
}

int64_t ckdo_nnodes(void *hv) {
    ckdo_handle *h = (ckdo_handle *)hv;
    return h->prec == 4 ? ((otree_f32 *)h->t)->nnodes : ((otree_f64 *)h->t)->nnodes;
}

void ckdo_dump(void *hv, int8_t *is_leaf, int32_t *axis, void *split, void *minb, void *maxb, int64_t *left,
               int64_t *right, int64_t *start, int64_t *count, void *lpts, int64_t *lidx) {
    ckdo_handle *h = (ckdo_handle *)hv;
    if (h->prec == 4)
        dump_f32((otree_f32 *)h->t, is_leaf, axis, (float *)split, (float *)minb, (float *)maxb, left, right, start,
                 count, (float *)lpts, lidx);
    else
        dump_f64((otree_f64 *)h->t, is_leaf, axis, (double *)split, (double *)minb, (double *)maxb, left, right,
                 start, count, (double *)lpts, lidx);
}

void ckdo_knn(void *hv, const void *q, int64_t nq, int k, int sorted, void *dist, int64_t *idx) {
    ckdo_handle *h = (ckdo_handle *)hv;
    if (h->prec == 4) knn_f32((otree_f32 *)h->t, (const float *)q, nq, k, sorted, (float *)dist, idx);
    else knn_f64((otree_f64 *)h->t, (const double *)q, nq, k, sorted, (double *)dist, idx);
}

typedef struct {
    int prec;
    void *r;
} ckdo_rhandle;

void *ckdo_ball_run(void *hv, const void *q, int64_t nq, double r, int sorted, int64_t *counts) {
    ckdo_handle *h = (ckdo_handle *)hv;
    ckdo_rhandle *rh = (ckdo_rhandle *)malloc(sizeof(ckdo_rhandle));
    rh->prec = h->prec;
    if (h->prec == 4) rh->r = ball_run_f32((otree_f32 *)h->t, (const float *)q, nq, (float)r, sorted, counts);
    else rh->r = ball_run_f64((otree_f64 *)h->t, (const double *)q, nq, r, sorted, counts);
    return rh;
}

void *ckdo_box_run(void *hv, const void *boxes, int64_t nb, int64_t *counts) {
    ckdo_handle *h = (ckdo_handle *)hv;
    ckdo_rhandle *rh = (ckdo_rhandle *)malloc(sizeof(ckdo_rhandle));
    rh->prec = h->prec;
    if (h->prec == 4) rh->r = box_run_f32((otree_f32 *)h->t, (const float *)boxes, nb, counts);
    else rh->r = box_run_f64((otree_f64 *)h->t, (const double *)boxes, nb, counts);
    return rh;
}

void *ckdo_box_wrap_run(void *hv, const void *boxes, int64_t nb, int64_t *counts) {
    ckdo_handle *h = (ckdo_handle *)hv;
    ckdo_rhandle *rh = (ckdo_rhandle *)malloc(sizeof(ckdo_rhandle));
    rh->prec = h->prec;
    if (h->prec == 4) rh->r = box_wrap_run_f32((otree_f32 *)h->t, (const float *)boxes, nb, counts);
    else rh->r = box_wrap_run_f64((otree_f64 *)h->t, (const double *)boxes, nb, counts);
    return rh;
}

void ckdo_res_fetch(void *rv, const int64_t *offsets, void *dist, int64_t *idx) {
    ckdo_rhandle *rh = (ckdo_rhandle *)rv;
    if (rh->prec == 4) res_fetch_f32((ores_f32 *)rh->r, offsets, (float *)dist, idx);
    else res_fetch_f64((ores_f64 *)rh->r, offsets, (double *)dist, idx);
}

void ckdo_res_free(void *rv) {
    ckdo_rhandle *rh = (ckdo_rhandle *)rv;
    if (!rh) return;
    if (rh->prec == 4) res_free_f32((ores_f32 *)rh->r);
    else res_free_f64((ores_f64 *)rh->r);
    free(rh);
}
