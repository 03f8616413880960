/*
 * ckd_oracle_impl.h -- precision-generic body of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the GPU library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  It is never on the product path.
 *
 * Included twice by ckd_oracle.c with REAL = float / double and
 * SFX = f32 / f64.  Every function restates one routine of the reference
 * (src/cosmokdtree.f90, src/cosmokdtree2py.f90) and cites it.  Arithmetic is
 * written so that gcc -ffp-contract=off evaluates exactly the Fortran source
 * expression (one rounding per operation, left-to-right sums).
 */

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define T(name) CAT(name, SFX)

typedef struct {
    REAL point;        /* split value                   (cosmokdtree.f90:95) */
    int axis;          /* 0-based split axis            (cosmokdtree.f90:96) */
    int is_leaf;       /*                               (cosmokdtree.f90:102) */
    int64_t left;      /* child node ids, -1 = null     (cosmokdtree.f90:97-98) */
    int64_t right;
    int64_t leaf_off;  /* leaf: offset into the leaf pools */
    int64_t leaf_n;    /* leaf: number of points */
} T(onode);

typedef struct {
    int ndim;
    int periodic;
    int leafsize;
    int64_t n;
    REAL L[CKDO_MAXD];
    int64_t nnodes, cap;
    T(onode) *nodes;
    REAL *minb, *maxb;   /* nnodes x ndim cell bounds (cosmokdtree.f90:99-100) */
    REAL *lpts;          /* leaf point pool, AoS (ndim, n) (cosmokdtree.f90:330-338) */
    int64_t *lidx;       /* leaf index pool, 1-based original indices */
    /* build scratch: column-major points(n, ndim) and indices(n)
       (cosmokdtree.f90:194-197) */
    REAL *bp;
    int64_t *bi;
} T(otree);

static int64_t T(new_node)(T(otree) *t) {
    if (t->nnodes == t->cap) {
        int64_t nc = t->cap ? 2 * t->cap : 1024;
        t->nodes = (T(onode) *)realloc(t->nodes, (size_t)nc * sizeof(T(onode)));
        t->minb = (REAL *)realloc(t->minb, (size_t)nc * t->ndim * sizeof(REAL));
        t->maxb = (REAL *)realloc(t->maxb, (size_t)nc * t->ndim * sizeof(REAL));
        t->cap = nc;
    }
    return t->nnodes++;
}

/* swap rows i, j of points(:, :) and indices(:)  (cosmokdtree.f90:481-497) */
static void T(swap_rows)(T(otree) *t, int64_t base, int64_t i, int64_t j) {
    int d;
    int64_t ti;
    for (d = 0; d < t->ndim; d++) {
        REAL *col = t->bp + (size_t)d * t->n + base;
        REAL tmp = col[i];
        col[i] = col[j];
        col[j] = tmp;
    }
    ti = t->bi[base + i];
    t->bi[base + i] = t->bi[base + j];
    t->bi[base + j] = ti;
}

/*
 * build_kdtree_recursive (cosmokdtree.f90:286-451) on the slice
 * [base, base+n) of the column-major build arrays.  Returns the node id.
 * Nodes are numbered in preorder (node, left subtree, right subtree).
 */
static int64_t T(build_rec)(T(otree) *t, int64_t base, int64_t n, const REAL *bounds) {
    int ndim = t->ndim, j, axis;
    int64_t id, i, n_left;
    REAL minv[CKDO_MAXD], maxv[CKDO_MAXD], spread[CKDO_MAXD], side[CKDO_MAXD];
    REAL maxside, midpoint, minx, maxx;
    REAL bl[2 * CKDO_MAXD], br[2 * CKDO_MAXD];
    int64_t minind, maxind;

    if (n == 0) return -1; /* cosmokdtree.f90:315-319 (never reached) */
    id = T(new_node)(t);
    for (j = 0; j < ndim; j++) { /* :322-325 */
        t->minb[(size_t)id * ndim + j] = bounds[2 * j];
        t->maxb[(size_t)id * ndim + j] = bounds[2 * j + 1];
    }
    t->nodes[id].left = t->nodes[id].right = -1;
    t->nodes[id].point = 0;
    t->nodes[id].axis = -1;

    if (n <= t->leafsize) { /* :327-341 leaf: copy (ndim, n) AoS + indices */
        t->nodes[id].is_leaf = 1;
        t->nodes[id].leaf_off = base;
        t->nodes[id].leaf_n = n;
        for (i = 0; i < n; i++) {
            for (j = 0; j < ndim; j++)
                t->lpts[(size_t)(base + i) * ndim + j] = t->bp[(size_t)j * t->n + base + i];
            t->lidx[base + i] = t->bi[base + i];
        }
        return id;
    }
    t->nodes[id].is_leaf = 0;
    t->nodes[id].leaf_off = base;
    t->nodes[id].leaf_n = n;

    /* :347-353 per-axis statistics */
    for (j = 0; j < ndim; j++) {
        const REAL *col = t->bp + (size_t)j * t->n + base;
        REAL mn = col[0], mx = col[0];
        for (i = 1; i < n; i++) {
            if (col[i] < mn) mn = col[i];
            if (col[i] > mx) mx = col[i];
        }
        minv[j] = mn;
        maxv[j] = mx;
        side[j] = bounds[2 * j + 1] - bounds[2 * j];
        spread[j] = maxv[j] - minv[j];
    }
    /* :355-365 longest cell side; exact ties go to the larger spread */
    maxside = side[0];
    axis = 0;
    for (j = 1; j < ndim; j++) {
        if (side[j] > maxside) {
            maxside = side[j];
            axis = j;
        }
        if (side[j] == maxside && spread[j] > spread[axis]) axis = j;
    }
    /* :367 */
    midpoint = (REAL)0.5 * (bounds[2 * axis] + bounds[2 * axis + 1]);

    /* :370-373 minloc / maxloc return the first occurrence */
    minx = minv[axis];
    maxx = maxv[axis];
    {
        const REAL *col = t->bp + (size_t)axis * t->n + base;
        minind = 0;
        maxind = 0;
        for (i = 1; i < n; i++) {
            if (col[i] < col[minind]) minind = i;
            if (col[i] > col[maxind]) maxind = i;
        }
    }

    if (minx >= midpoint) { /* :376-388 slide to the minimum */
        midpoint = minx;
        T(swap_rows)(t, base, 0, minind);
        n_left = 1;
    } else if (maxx < midpoint) { /* :391-403 slide to the maximum */
        midpoint = maxx;
        T(swap_rows)(t, base, n - 1, maxind);
        n_left = n - 1;
    } else { /* :406-409 -> partition (:458-479), Lomuto order */
        const REAL *col = t->bp + (size_t)axis * t->n + base;
        n_left = 0;
        for (i = 0; i < n; i++) {
            if (col[i] < midpoint) {
                if (i != n_left) T(swap_rows)(t, base, i, n_left);
                n_left++;
            }
        }
    }
    t->nodes[id].point = midpoint; /* :415-416 */
    t->nodes[id].axis = axis;

    for (j = 0; j < 2 * ndim; j++) bl[j] = br[j] = bounds[j]; /* :419-422 */
    bl[2 * axis + 1] = midpoint;
    br[2 * axis] = midpoint;

    /* :425-448 (task parallelism does not change the result) */
    {
        int64_t l = T(build_rec)(t, base, n_left, bl);
        int64_t r = T(build_rec)(t, base + n_left, n - n_left, br);
        t->nodes[id].left = l;
        t->nodes[id].right = r;
    }
    return id;
}

/* build_kdtree (cosmokdtree.f90:124-238).  pts is C-order (n, ndim).
   L == NULL selects the non-periodic build. */
static T(otree) *T(build)(const REAL *pts, int64_t n, int ndim, int leaf, const REAL *L, int *err) {
    T(otree) *t;
    REAL bounds[2 * CKDO_MAXD];
    int64_t i;
    int j;
    *err = 0;
    if (ndim < 1 || ndim > CKDO_MAXD) { *err = 1; return NULL; } /* :160-162 */
    if (leaf < 1) { *err = 2; return NULL; }                       /* kdtree.py:21 */
    if (n < 1) { *err = 4; return NULL; }
    if (L) { /* :176-190 points must lie in [-L/2, L/2] */
        for (j = 0; j < ndim; j++) {
            if (!(L[j] > 0)) { *err = 5; return NULL; }
            for (i = 0; i < n; i++) {
                REAL x = pts[(size_t)i * ndim + j];
                if (x < -L[j] / 2 || x > L[j] / 2) { *err = 3; return NULL; }
            }
        }
    }
    t = (T(otree) *)calloc(1, sizeof(T(otree)));
    t->ndim = ndim;
    t->n = n;
    t->leafsize = leaf;
    t->periodic = L != NULL;
    if (L) for (j = 0; j < ndim; j++) t->L[j] = L[j];
    t->bp = (REAL *)malloc((size_t)n * ndim * sizeof(REAL));
    t->bi = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    t->lpts = (REAL *)malloc((size_t)n * ndim * sizeof(REAL));
    t->lidx = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    for (i = 0; i < n; i++) { /* :194-197 out-of-place copies, indices = 1..n */
        for (j = 0; j < ndim; j++) t->bp[(size_t)j * n + i] = pts[(size_t)i * ndim + j];
        t->bi[i] = i + 1;
    }
    if (L) { /* :218-224 (the reference hard-codes axes 1..3; all axes here) */
        for (j = 0; j < ndim; j++) {
            bounds[2 * j] = -L[j] / 2;
            bounds[2 * j + 1] = L[j] / 2;
        }
    } else { /* :226-229 */
        for (j = 0; j < ndim; j++) {
            const REAL *col = t->bp + (size_t)j * n;
            REAL mn = col[0], mx = col[0];
            for (i = 1; i < n; i++) {
                if (col[i] < mn) mn = col[i];
                if (col[i] > mx) mx = col[i];
            }
            bounds[2 * j] = mn;
            bounds[2 * j + 1] = mx;
        }
    }
    T(build_rec)(t, 0, n, bounds);
    free(t->bp); /* :235-236 */
    free(t->bi);
    t->bp = NULL;
    t->bi = NULL;
    return t;
}

static void T(tree_free)(T(otree) *t) {
    if (!t) return;
    free(t->nodes);
    free(t->minb);
    free(t->maxb);
    free(t->lpts);
    free(t->lidx);
    free(t);
}

/* distance (cosmokdtree.f90:1441-1465): squared, axis order, p1 - p2. */
static inline REAL T(distance)(const T(otree) *t, const REAL *p1, const REAL *p2) {
    REAL dx[CKDO_MAXD], d;
    int i;
    for (i = 0; i < t->ndim; i++) dx[i] = p1[i] - p2[i];
    if (t->periodic) {
        for (i = 0; i < t->ndim; i++) {
            REAL a = dx[i] < 0 ? -dx[i] : dx[i];
            REAL b = t->L[i] - a;
            dx[i] = b < a ? b : a;
        }
    }
    d = 0;
    for (i = 0; i < t->ndim; i++) d = d + dx[i] * dx[i];
    return d;
}

/* distance_1D (cosmokdtree.f90:1472-1485) */
static inline REAL T(distance_1d)(const T(otree) *t, REAL x1, REAL x2, int dim) {
    REAL dx = x1 - x2;
    if (t->periodic) {
        REAL a = dx < 0 ? -dx : dx;
        REAL b = t->L[dim] - a;
        dx = b < a ? b : a;
    }
    return dx < 0 ? -dx : dx;
}

/* bbox_mindist_kd (cosmokdtree.f90:1525-1546) */
static inline REAL T(bbox_mindist)(const T(otree) *t, int64_t node, const REAL *p) {
    const REAL *mn = t->minb + (size_t)node * t->ndim, *mx = t->maxb + (size_t)node * t->ndim;
    REAL s = 0, dmin;
    int d;
    for (d = 0; d < t->ndim; d++) {
        if (p[d] < mn[d]) dmin = T(distance_1d)(t, p[d], mn[d], d);
        else if (p[d] > mx[d]) dmin = T(distance_1d)(t, p[d], mx[d], d);
        else dmin = 0;
        s = s + dmin * dmin;
    }
    return s;
}

/* max_heap_insert (cosmokdtree.f90:859-905): sift the new root down. */
static inline void T(max_heap_insert)(REAL *dist, int64_t *idx, int k) {
    int i = 0;
    for (;;) {
        int l = 2 * i + 1, r = 2 * i + 2, largest = i;
        if (l < k && dist[l] > dist[largest]) largest = l;
        if (r < k && dist[r] > dist[largest]) largest = r;
        if (largest == i) break;
        {
            REAL td = dist[i];
            int64_t ti = idx[i];
            dist[i] = dist[largest];
            dist[largest] = td;
            idx[i] = idx[largest];
            idx[largest] = ti;
        }
        i = largest;
    }
}

/* leaf scan shared by K1 and K2 (cosmokdtree.f90:654-664, :804-814) */
static inline void T(knn_leaf)(const T(otree) *t, int64_t node, const REAL *q, REAL *dist, int64_t *idx, int k) {
    const T(onode) *nd = &t->nodes[node];
    REAL far = dist[0];
    int64_t i;
    for (i = 0; i < nd->leaf_n; i++) {
        REAL dc = T(distance)(t, t->lpts + (size_t)(nd->leaf_off + i) * t->ndim, q);
        if (dc < far) {
            dist[0] = dc;
            idx[0] = t->lidx[nd->leaf_off + i];
            T(max_heap_insert)(dist, idx, k);
            far = dist[0];
        }
    }
}

/* min_heap_queue_down (cosmokdtree.f90:912-958) on queue[0..nq) */
static inline void T(pq_down)(int64_t *qn, REAL *qd, int64_t nq) {
    int64_t i = 0;
    for (;;) {
        int64_t l = 2 * i + 1, r = 2 * i + 2, s = i;
        if (l < nq && qd[l] < qd[s]) s = l;
        if (r < nq && qd[r] < qd[s]) s = r;
        if (s == i) break;
        {
            REAL td = qd[i];
            int64_t tn = qn[i];
            qd[i] = qd[s];
            qd[s] = td;
            qn[i] = qn[s];
            qn[s] = tn;
        }
        i = s;
    }
}

/* min_heap_queue_up (cosmokdtree.f90:965-1008) for the element at nq-1 */
static inline void T(pq_up)(int64_t *qn, REAL *qd, int64_t nq) {
    int64_t i = nq - 1;
    while (i > 0) {
        int64_t f = (i - 1) / 2; /* 1-based father = i/2 */
        if (qd[f] > qd[i]) {
            REAL td = qd[i];
            int64_t tn = qn[i];
            qd[i] = qd[f];
            qd[f] = td;
            qn[i] = qn[f];
            qn[f] = tn;
            i = f;
        } else {
            break;
        }
    }
}

/* knn_search_queue (cosmokdtree.f90:549-675): best-first, non-periodic. */
static void T(knn_queue)(const T(otree) *t, const REAL *q, REAL *dist, int64_t *idx, int k) {
    int64_t cap = 128, nq = 1;
    int64_t *qn = (int64_t *)malloc((size_t)cap * sizeof(int64_t));
    REAL *qd = (REAL *)malloc((size_t)cap * sizeof(REAL));
    qn[0] = 0;
    qd[0] = T(bbox_mindist)(t, 0, q);
    while (nq > 0) {
        int64_t node;
        REAL md, far;
        if (nq + 2 > cap) {
            cap *= 4;
            qn = (int64_t *)realloc(qn, (size_t)cap * sizeof(int64_t));
            qd = (REAL *)realloc(qd, (size_t)cap * sizeof(REAL));
        }
        /* :603-611 pop the root, move the last element up, sift down */
        node = qn[0];
        md = qd[0];
        qn[0] = qn[nq - 1];
        qd[0] = qd[nq - 1];
        nq--;
        if (nq > 1) T(pq_down)(qn, qd, nq);
        if (node < 0) continue;
        far = dist[0]; /* :618 */
        if (md > far) continue; /* :621 */
        if (!t->nodes[node].is_leaf) {
            int64_t l = t->nodes[node].left, r = t->nodes[node].right;
            REAL ml = T(bbox_mindist)(t, l, q);
            REAL mr = T(bbox_mindist)(t, r, q);
            if (ml <= far) { /* :634-639 */
                qn[nq] = l;
                qd[nq] = ml;
                nq++;
                T(pq_up)(qn, qd, nq);
            }
            if (mr <= far) { /* :642-647 */
                qn[nq] = r;
                qd[nq] = mr;
                nq++;
                T(pq_up)(qn, qd, nq);
            }
        } else {
            T(knn_leaf)(t, node, q, dist, idx, k);
        }
    }
    free(qn);
    free(qd);
}

/* knn_search_recursive_hyperp (cosmokdtree.f90:782-853): periodic. */
static void T(knn_hyperp)(const T(otree) *t, int64_t node, const REAL *q, REAL *dist, int64_t *idx, int k) {
    const T(onode) *nd;
    if (node < 0) return;
    nd = &t->nodes[node];
    if (nd->is_leaf) {
        T(knn_leaf)(t, node, q, dist, idx, k);
    } else {
        int axis = nd->axis;
        REAL d1d = q[axis] - nd->point, far;
        int look;
        if (d1d < 0) {
            T(knn_hyperp)(t, nd->left, q, dist, idx, k);
            far = dist[0];
            look = d1d * d1d < far;
            if (t->periodic && q[axis] - SQRT(far) <= -t->L[axis] / 2) look = 1;
            if (look) T(knn_hyperp)(t, nd->right, q, dist, idx, k);
        } else {
            T(knn_hyperp)(t, nd->right, q, dist, idx, k);
            far = dist[0];
            look = d1d * d1d < far;
            if (t->periodic && q[axis] + SQRT(far) >= t->L[axis] / 2) look = 1;
            if (look) T(knn_hyperp)(t, nd->left, q, dist, idx, k);
        }
    }
}

/* quicksort / partition2 / swap2 (cosmokdtree.f90:1553-1643), 0-based. */
static void T(qsort_rec)(REAL *d, int64_t *x, int64_t low, int64_t high) {
    while (low < high) {
        int64_t mid = (low + high) / 2, i, j, p;
        REAL pv, td;
        int64_t ti;
#define SWP(a, b) do { td = d[a]; d[a] = d[b]; d[b] = td; ti = x[a]; x[a] = x[b]; x[b] = ti; } while (0)
        if (d[low] > d[mid]) SWP(low, mid);
        if (d[low] > d[high]) SWP(low, high);
        if (d[mid] > d[high]) SWP(mid, high);
        pv = d[mid];
        SWP(mid, high);
        i = low - 1;
        for (j = low; j < high; j++) {
            if (d[j] <= pv) {
                i++;
                SWP(i, j);
            }
        }
        i++;
        SWP(i, high);
#undef SWP
        p = i;
        T(qsort_rec)(d, x, low, p - 1);
        low = p + 1; /* tail call of quicksort_recursive(pivot+1, high) */
    }
}

/* knn_search (cosmokdtree.f90:504-540) + pyknn_search loop
   (cosmokdtree2py.f90:132-141) + kdtree.py:52 (idx -= 1). */
static void T(knn)(const T(otree) *t, const REAL *q, int64_t nq, int k, int sorted, REAL *dout, int64_t *iout) {
    int64_t qi;
#pragma omp parallel for schedule(static)
    for (qi = 0; qi < nq; qi++) {
        REAL *dist = dout + (size_t)qi * k;
        int64_t *idx = iout + (size_t)qi * k;
        const REAL *qq = q + (size_t)qi * t->ndim;
        int j;
        for (j = 0; j < k; j++) { /* :517-518 HUGE(0.0) is the float max */
            dist[j] = (REAL)FLT_MAX;
            idx[j] = -1;
        }
        if (!t->periodic) T(knn_queue)(t, qq, dist, idx, k);
        else T(knn_hyperp)(t, 0, qq, dist, idx, k);
        if (sorted) T(qsort_rec)(dist, idx, 0, k - 1);
        for (j = 0; j < k; j++) {
            dist[j] = SQRT(dist[j]);
            idx[j] -= 1;
        }
    }
}

/* growable per-query result (cosmokdtree.f90:1394-1420) */
typedef struct {
    int64_t n, cap;
    REAL *d;
    int64_t *i;
} T(olist);

static inline void T(olist_add)(T(olist) *l, int64_t id, REAL d) {
    if (l->n == l->cap) {
        l->cap = l->cap ? l->cap * 4 : 64;
        l->d = (REAL *)realloc(l->d, (size_t)l->cap * sizeof(REAL));
        l->i = (int64_t *)realloc(l->i, (size_t)l->cap * sizeof(int64_t));
    }
    l->d[l->n] = d;
    l->i[l->n] = id;
    l->n++;
}

/* ball_reduction_recursive (cosmokdtree.f90:1332-1366) */
static void T(ball_reduce)(const T(otree) *t, int64_t node, const REAL *q, T(olist) *l) {
    const T(onode) *nd;
    if (node < 0) return;
    nd = &t->nodes[node];
    if (nd->is_leaf) {
        int64_t i;
        for (i = 0; i < nd->leaf_n; i++)
            T(olist_add)(l, t->lidx[nd->leaf_off + i],
                         T(distance)(t, t->lpts + (size_t)(nd->leaf_off + i) * t->ndim, q));
    } else {
        T(ball_reduce)(t, nd->left, q, l);
        T(ball_reduce)(t, nd->right, q, l);
    }
}

/* ball_search_recursive_bbox (cosmokdtree.f90:1076-1139) + bbox_distance_1D
   (:1492-1516): non-periodic. */
static void T(ball_bbox)(const T(otree) *t, int64_t node, const REAL *q, REAL r2, T(olist) *l) {
    const T(onode) *nd;
    const REAL *mn, *mx;
    REAL s_min = 0, s_max = 0;
    int d;
    if (node < 0) return;
    nd = &t->nodes[node];
    mn = t->minb + (size_t)node * t->ndim;
    mx = t->maxb + (size_t)node * t->ndim;
    for (d = 0; d < t->ndim; d++) {
        REAL dmin, dmax, a, b;
        if (q[d] < mn[d]) dmin = T(distance_1d)(t, q[d], mn[d], d);
        else if (q[d] > mx[d]) dmin = T(distance_1d)(t, q[d], mx[d], d);
        else dmin = 0;
        a = T(distance_1d)(t, q[d], mn[d], d);
        b = T(distance_1d)(t, q[d], mx[d], d);
        dmax = a > b ? a : b; /* Fortran max(a, b) */
        s_min = s_min + dmin * dmin; /* sum(mindist**2), axis order */
        s_max = s_max + dmax * dmax;
    }
    if (s_min > r2) return; /* :1112 */
    if (s_max <= r2) { /* :1116-1118 whole subtree */
        T(ball_reduce)(t, node, q, l);
        return;
    }
    if (nd->is_leaf) { /* :1122-1129 inclusive test */
        int64_t i;
        for (i = 0; i < nd->leaf_n; i++) {
            REAL dc = T(distance)(t, t->lpts + (size_t)(nd->leaf_off + i) * t->ndim, q);
            if (dc <= r2) T(olist_add)(l, t->lidx[nd->leaf_off + i], dc);
        }
    } else {
        T(ball_bbox)(t, nd->left, q, r2, l);
        T(ball_bbox)(t, nd->right, q, r2, l);
    }
}

/* ball_search_recursive_hyperp (cosmokdtree.f90:1143-1212): periodic. */
static void T(ball_hyperp)(const T(otree) *t, int64_t node, const REAL *q, REAL r, REAL r2, T(olist) *l) {
    const T(onode) *nd;
    if (node < 0) return;
    nd = &t->nodes[node];
    if (nd->is_leaf) {
        int64_t i;
        for (i = 0; i < nd->leaf_n; i++) {
            REAL dc = T(distance)(t, t->lpts + (size_t)(nd->leaf_off + i) * t->ndim, q);
            if (dc <= r2) T(olist_add)(l, t->lidx[nd->leaf_off + i], dc);
        }
    } else {
        int axis = nd->axis;
        REAL d1d = q[axis] - nd->point;
        REAL ad = d1d < 0 ? -d1d : d1d;
        int look;
        if (d1d < 0) {
            T(ball_hyperp)(t, nd->left, q, r, r2, l);
            look = ad <= r;
            if (t->periodic && q[axis] - r <= -t->L[axis] / 2) look = 1;
            if (look) T(ball_hyperp)(t, nd->right, q, r, r2, l);
        } else {
            T(ball_hyperp)(t, nd->right, q, r, r2, l);
            look = ad <= r;
            if (t->periodic && q[axis] + r >= t->L[axis] / 2) look = 1;
            if (look) T(ball_hyperp)(t, nd->left, q, r, r2, l);
        }
    }
}

typedef struct {
    int64_t nq;
    T(olist) *lists;
} T(ores);

/* ball_search (cosmokdtree.f90:1015-1072) + pyball_search_call_1
   (cosmokdtree2py.f90:153-201).  counts[q] receives the result size. */
static T(ores) *T(ball_run)(const T(otree) *t, const REAL *q, int64_t nq, REAL r, int sorted, int64_t *counts) {
    T(ores) *res = (T(ores) *)calloc(1, sizeof(T(ores)));
    int64_t qi;
    REAL r2 = r * r; /* radius**2 in tree precision */
    res->nq = nq;
    res->lists = (T(olist) *)calloc((size_t)(nq > 0 ? nq : 1), sizeof(T(olist)));
#pragma omp parallel for schedule(dynamic, 64)
    for (qi = 0; qi < nq; qi++) {
        const REAL *qq = q + (size_t)qi * t->ndim;
        T(olist) *l = &res->lists[qi];
        int64_t j;
        if (!t->periodic) T(ball_bbox)(t, 0, qq, r2, l);
        else T(ball_hyperp)(t, 0, qq, r, r2, l);
        if (l->n > 0 && sorted) T(qsort_rec)(l->d, l->i, 0, l->n - 1); /* :1062-1066 */
        for (j = 0; j < l->n; j++) { /* :1069-1070 sqrt; kdtree.py:68 idx -= 1 */
            l->d[j] = SQRT(l->d[j]);
            l->i[j] -= 1;
        }
        counts[qi] = l->n;
    }
    return res;
}

/* pyball_search_call_2 (cosmokdtree2py.f90:203-246) into CSR rows */
static void T(res_fetch)(const T(ores) *res, const int64_t *offsets, REAL *dout, int64_t *iout) {
    int64_t qi;
    for (qi = 0; qi < res->nq; qi++) {
        const T(olist) *l = &res->lists[qi];
        if (l->n == 0) continue; /* an empty row has no buffers (memcpy from NULL is undefined) */
        if (dout) memcpy(dout + offsets[qi], l->d, (size_t)l->n * sizeof(REAL));
        memcpy(iout + offsets[qi], l->i, (size_t)l->n * sizeof(int64_t));
    }
}

static void T(res_free)(T(ores) *res) {
    int64_t qi;
    if (!res) return;
    for (qi = 0; qi < res->nq; qi++) {
        free(res->lists[qi].d);
        free(res->lists[qi].i);
    }
    free(res->lists);
    free(res);
}

/* box_search_recursive (cosmokdtree.f90:1258-1328): strict visit tests,
   inclusive leaf test, left before right. */
static void T(box_rec)(const T(otree) *t, int64_t node, const REAL *box, T(olist) *l) {
    const T(onode) *nd;
    if (node < 0) return;
    nd = &t->nodes[node];
    if (nd->is_leaf) {
        int64_t i;
        for (i = 0; i < nd->leaf_n; i++) {
            const REAL *p = t->lpts + (size_t)(nd->leaf_off + i) * t->ndim;
            int in = 1, j;
            for (j = 0; j < t->ndim; j++) { /* is_in_box :1309-1325 */
                if (p[j] < box[2 * j] || p[j] > box[2 * j + 1]) {
                    in = 0;
                    break;
                }
            }
            if (in) T(olist_add)(l, t->lidx[nd->leaf_off + i], 0);
        }
    } else {
        int axis = nd->axis;
        if (box[2 * axis] < nd->point) T(box_rec)(t, nd->left, box, l);
        if (box[2 * axis + 1] > nd->point) T(box_rec)(t, nd->right, box, l);
    }
}

/* box_search (cosmokdtree.f90:1220-1253) + pybox_search_call_1/2
   (cosmokdtree2py.f90:252-315), one call per box, boxes[b] =
   (xmin, xmax, ymin, ymax, ...). */
static T(ores) *T(box_run)(const T(otree) *t, const REAL *boxes, int64_t nb, int64_t *counts) {
    T(ores) *res = (T(ores) *)calloc(1, sizeof(T(ores)));
    int64_t b;
    res->nq = nb;
    res->lists = (T(olist) *)calloc((size_t)(nb > 0 ? nb : 1), sizeof(T(olist)));
#pragma omp parallel for schedule(dynamic, 16)
    for (b = 0; b < nb; b++) {
        T(olist) *l = &res->lists[b];
        int64_t j;
        T(box_rec)(t, 0, boxes + (size_t)b * 2 * t->ndim, l);
        for (j = 0; j < l->n; j++) l->i[j] -= 1; /* kdtree.py:86 */
        counts[b] = l->n;
    }
    return res;
}

/* Box search in a periodic tree with boxes that may straddle the faces (no
   reference counterpart: box_search does not wrap, cosmokdtree.f90:1260).
   The rule this library defines: on each axis the interval [lo, hi] is
   taken modulo L as the pieces [lo - kL, hi - kL] clipped to [-L/2, L/2]
   for k = m+1, m, m-1, m = floor((lo + L/2) / L) in double, pieces that
   touch merged, the whole axis when hi - lo >= L, nothing when lo > hi.
   Every sub-box of the product of the pieces (axis 0 fastest) is searched
   with the reference's rules above, and the rows are concatenated: the
   pieces are disjoint, so no point comes twice. */
static int T(wrap_pieces)(REAL lo, REAL hi, REAL L, REAL *pa, REAL *pb) {
    const REAL H = L * (REAL)0.5;
    double m;
    int j, n = 0;
    if (!(lo <= hi)) return 0;
    if (hi - lo >= L) {
        pa[0] = -H;
        pb[0] = H;
        return 1;
    }
    m = floor(((double)lo + (double)H) / (double)L);
    for (j = 1; j >= -1; j--) {
        const REAL sh = (REAL)(m + j) * L;
        REAL a = lo - sh, b = hi - sh;
        if (a < -H) a = -H;
        if (b > H) b = H;
        if (a <= b) {
            if (n > 0 && a <= pb[n - 1]) {
                if (b > pb[n - 1]) pb[n - 1] = b;
            } else {
                pa[n] = a;
                pb[n] = b;
                n++;
            }
        }
    }
    return n;
}

static T(ores) *T(box_wrap_run)(const T(otree) *t, const REAL *boxes, int64_t nb, int64_t *counts) {
    T(ores) *res = (T(ores) *)calloc(1, sizeof(T(ores)));
    int64_t b;
    res->nq = nb;
    res->lists = (T(olist) *)calloc((size_t)(nb > 0 ? nb : 1), sizeof(T(olist)));
#pragma omp parallel for schedule(dynamic, 16)
    for (b = 0; b < nb; b++) {
        T(olist) *l = &res->lists[b];
        const REAL *bx = boxes + (size_t)b * 2 * t->ndim;
        REAL pa[CKDO_MAXD][3], pb[CKDO_MAXD][3], sub[2 * CKDO_MAXD];
        int np[CKDO_MAXD], d;
        int64_t total = 1, c, j;
        for (d = 0; d < t->ndim; d++) {
            np[d] = T(wrap_pieces)(bx[2 * d], bx[2 * d + 1], t->L[d], pa[d], pb[d]);
            total *= np[d];
        }
        for (c = 0; c < total; c++) {
            int64_t r = c;
            for (d = 0; d < t->ndim; d++) {
                const int64_t jd = r % np[d];
                r /= np[d];
                sub[2 * d] = pa[d][jd];
                sub[2 * d + 1] = pb[d][jd];
            }
            T(box_rec)(t, 0, sub, l);
        }
        for (j = 0; j < l->n; j++) l->i[j] -= 1;
        counts[b] = l->n;
    }
    return res;
}

/* Preorder dump of the tree structure. */
static void T(dump)(const T(otree) *t, int8_t *is_leaf, int32_t *axis, REAL *split, REAL *minb, REAL *maxb,
                    int64_t *left, int64_t *right, int64_t *start, int64_t *count, REAL *lpts, int64_t *lidx) {
    int64_t i;
    for (i = 0; i < t->nnodes; i++) {
        const T(onode) *nd = &t->nodes[i];
        if (is_leaf) is_leaf[i] = (int8_t)nd->is_leaf;
        if (axis) axis[i] = nd->axis;
        if (split) split[i] = nd->point;
        if (left) left[i] = nd->left;
        if (right) right[i] = nd->right;
        if (start) start[i] = nd->leaf_off;
        if (count) count[i] = nd->leaf_n;
    }
    if (minb) memcpy(minb, t->minb, (size_t)t->nnodes * t->ndim * sizeof(REAL));
    if (maxb) memcpy(maxb, t->maxb, (size_t)t->nnodes * t->ndim * sizeof(REAL));
    if (lpts) memcpy(lpts, t->lpts, (size_t)t->n * t->ndim * sizeof(REAL));
    if (lidx) for (i = 0; i < t->n; i++) lidx[i] = t->lidx[i] - 1;
}

#undef T
#undef CAT
#undef CAT2
