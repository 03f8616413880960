// ckd_persist.hip -- a built tree saved to and loaded from a file, without
// rebuilding: a fixed header, then every device array of the tree as it
// lies in HBM.  No reference counterpart: the reference's tree is an opaque
// pointer that cannot be serialised (src/cosmokdtree2py.f90:70-73).
//
// File layout (little-endian): PersistHeader, then the arrays in the order
// of the table in save_arrays(), each as raw bytes.
#include <cstdio>
#include <cstring>
#include <vector>

#include "ckd_internal.h"

namespace ckd {
namespace {

constexpr char MAGIC[8] = {'C', 'K', 'D', 'T', 'R', 'E', 'E', '1'};

struct PersistHeader {
    char magic[8];
    int64_t n, nnodes, nleaves, nsets, nwide;
    int32_t ndim, prec, leafsize, periodic, idx_bits, depth;
    double L[CKD_MAX_DIM];
};

struct Array {
    void **ptr;
    size_t bytes;
};

// The device arrays of a tree and their sizes.
std::vector<Array> arrays(ckd_tree *t) {
    const size_t T = (size_t)t->prec, nn = (size_t)t->nnodes, D = (size_t)t->ndim;
    const size_t node_bytes = T == 4 ? 16 : 24;  // NodeRec<float> / NodeRec<double>
    const size_t nw = (size_t)(t->nwide > 0 ? t->nwide : 1);
    return {
        {&t->nodes, node_bytes * nn},
        {&t->bbox, T * 2 * D * nn},
        {&t->dbox, T * 2 * D * nn},
        {(void **)&t->range, sizeof(Range) * nn},
        {&t->pts, T * (size_t)pstride((int)D) * (size_t)(t->n + PTS_PAD)},
        {&t->idx, (size_t)(t->idx_bits / 8) * (size_t)t->n},
        {&t->wide, t->wide ? wide_bytes_rt((int)D, (int)T) * nw : 0},
        {(void **)&t->wroot, t->wroot ? sizeof(int2) * (size_t)t->nsets : 0},
        {(void **)&t->wrange, t->wrange ? sizeof(int2) * nw : 0},
        {(void **)&t->set_off, t->set_off ? sizeof(int64_t) * (size_t)(t->nsets + 1) : 0},
    };
}

// Checks of one loaded array before it goes to the device: every index a
// kernel follows must stay inside its array, whatever the file holds.
// Returns an empty string when the array is consistent.  `depth` receives
// the depth recomputed from the node links (array 0).
template <typename T>
std::string check_nodes(const char *raw, const PersistHeader &h, int32_t &depth) {
    const NodeRec<T> *nd = (const NodeRec<T> *)raw;
    const int64_t nn = h.nnodes;
    int64_t leaves = 0;
    for (int64_t i = 0; i < nn; i++) {
        const NodeRec<T> &r = nd[i];
        if (r.axis < 0) {
            if (r.child != -1) return "a leaf with children";
            leaves++;
        } else {
            if (r.axis >= h.ndim) return "a split axis beyond ndim";
            if (r.child < h.nsets || (int64_t)r.child + 1 >= nn) return "a child index out of range";
        }
        if (r.parent < -1 || r.parent >= nn) return "a parent index out of range";
    }
    if (leaves != h.nleaves) return "the leaf count differs from the header";
    // walk every set's tree from its root: each node reached once, depth counted
    std::vector<int32_t> lvl(nn, -1);
    std::vector<int64_t> stk;
    int64_t seen = 0;
    depth = 0;
    for (int64_t r = 0; r < h.nsets; r++) {
        if (nd[r].parent != -1) return "a set root with a parent";
        lvl[r] = 0;
        stk.push_back(r);
        while (!stk.empty()) {
            const int64_t v = stk.back();
            stk.pop_back();
            seen++;
            depth = lvl[v] > depth ? lvl[v] : depth;
            if (nd[v].axis >= 0)
                for (int64_t c = nd[v].child; c <= (int64_t)nd[v].child + 1; c++) {
                    if (lvl[c] >= 0 || nd[c].parent != v) return "node links do not form a tree";
                    lvl[c] = lvl[v] + 1;
                    stk.push_back(c);
                }
        }
    }
    if (seen != nn) return "nodes unreachable from the roots";
    return "";
}

std::string check_pair(int2 p, int64_t nwide, int64_t n) {
    if (p.y < 0) return p.x >= 0 && p.x < nwide ? "" : "a wide id out of range";
    return p.x >= 0 && (int64_t)p.x + p.y <= n ? "" : "leaf rows out of range";
}

// Child pairs of the wide records (array 6), kept for the check of the roots
// (array 7): every wide record is reached once from the set roots, and the
// deepest chain of inner nodes is no deeper than the node links say (the
// stack walks size their stacks by that depth).
struct WideLinks {
    std::vector<int4> refs;
};

std::string check_wide_tree(const WideLinks &wl, const int2 *roots, int64_t nsets, int32_t depth) {
    const int64_t nw = (int64_t)wl.refs.size();
    std::vector<uint8_t> seen(nw, 0);
    std::vector<std::pair<int32_t, int32_t>> stk;  // (wide id, inner nodes above it)
    int64_t reached = 0;
    for (int64_t r = 0; r < nsets; r++) {
        if (roots[r].y >= 0) continue;  // a root that is a leaf
        stk.push_back({roots[r].x, 0});
        while (!stk.empty()) {
            const auto [w, lv] = stk.back();
            stk.pop_back();
            if (seen[w]) return "a wide record reached twice";
            seen[w] = 1;
            reached++;
            if (lv + 1 > depth) return "wide records deeper than the node links";
            const int4 f = wl.refs[w];
            if (f.y < 0) stk.push_back({f.x, lv + 1});
            if (f.w < 0) stk.push_back({f.z, lv + 1});
        }
    }
    return reached == nw ? "" : "wide records unreachable from the roots";
}

std::string check_array(int which, const char *raw, size_t bytes, const PersistHeader &h, int32_t &depth,
                        WideLinks &wl) {
    const int64_t n = h.n, nn = h.nnodes;
    switch (which) {
        case 0:
            return h.prec == 4 ? check_nodes<float>(raw, h, depth) : check_nodes<double>(raw, h, depth);
        case 3: {
            const Range *rg = (const Range *)raw;
            for (int64_t i = 0; i < nn; i++)
                if (rg[i].start < 0 || rg[i].count < 0 || rg[i].start + rg[i].count > n) return "a row range out of range";
            return "";
        }
        case 5: {
            for (int64_t i = 0; i < n; i++) {
                const int64_t v = h.idx_bits == 64 ? ((const int64_t *)raw)[i] : ((const int32_t *)raw)[i];
                if (v < 0 || v >= n) return "an original index out of range";
            }
            return "";
        }
        case 6: {
            const size_t wb = wide_bytes_rt(h.ndim, h.prec);
            wl.refs.resize((size_t)h.nwide);
            for (int64_t w = 0; w < h.nwide; w++) {
                const int4 ref = *(const int4 *)(raw + w * wb + 4 * (size_t)h.ndim * h.prec);
                std::string e = check_pair(make_int2(ref.x, ref.y), h.nwide, n);
                if (e.empty()) e = check_pair(make_int2(ref.z, ref.w), h.nwide, n);
                if (!e.empty()) return e;
                // depth-first numbering: an inner child comes after its parent,
                // so no walk can cycle
                if ((ref.y < 0 && ref.x <= w) || (ref.w < 0 && ref.z <= w)) return "a wide child before its parent";
                wl.refs[(size_t)w] = ref;
            }
            return "";
        }
        case 7: {
            for (int64_t r = 0; r < h.nsets; r++) {
                std::string e = check_pair(((const int2 *)raw)[r], h.nwide, n);
                if (!e.empty()) return e;
            }
            return check_wide_tree(wl, (const int2 *)raw, h.nsets, depth);
        }
        case 8: {
            for (int64_t w = 0; w < h.nwide; w++) {
                const int2 p = ((const int2 *)raw)[w];
                if (p.x < 0 || p.y < 0 || (int64_t)p.x + p.y > n) return "inner-node rows out of range";
            }
            return "";
        }
        case 9: {
            const int64_t *so = (const int64_t *)raw;
            if (so[0] != 0 || so[h.nsets] != n) return "set offsets do not span the points";
            for (int64_t r = 0; r < h.nsets; r++)
                if (so[r + 1] <= so[r]) return "set offsets not increasing";
            return "";
        }
        default:
            (void)bytes;
            return "";
    }
}

}  // namespace
}  // namespace ckd

using namespace ckd;

extern "C" {

int ckd_tree_save(const ckd_tree *tc, const char *path) {
    if (!tc || !path) {
        set_error("tree and path must not be NULL");
        return CKD_ERR_ARG;
    }
    ckd_tree *t = const_cast<ckd_tree *>(tc);
    CKD_HIP(hipSetDevice(t->device));
    CKD_HIP(hipDeviceSynchronize());
    FILE *f = fopen(path, "wb");
    if (!f) {
        set_error("cannot open %s for writing", path);
        return CKD_ERR_ARG;
    }
    PersistHeader h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, MAGIC, 8);
    h.n = t->n;
    h.nnodes = t->nnodes;
    h.nleaves = t->nleaves;
    h.nsets = t->nsets;
    h.nwide = t->wide ? t->nwide : -1;
    h.ndim = t->ndim;
    h.prec = t->prec;
    h.leafsize = t->leafsize;
    h.periodic = t->periodic;
    h.idx_bits = t->idx_bits;
    h.depth = t->depth;
    for (int d = 0; d < CKD_MAX_DIM; d++) h.L[d] = t->L[d];
    bool ok = fwrite(&h, sizeof(h), 1, f) == 1;
    std::vector<char> host;
    for (const Array &a : arrays(t)) {
        if (!ok || a.bytes == 0) continue;
        host.resize(a.bytes);
        hipError_t e = hipMemcpy(host.data(), *a.ptr, a.bytes, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            fclose(f);
            return hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
        }
        ok = fwrite(host.data(), 1, a.bytes, f) == a.bytes;
    }
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        set_error("writing %s failed", path);
        return CKD_ERR_ARG;
    }
    return CKD_OK;
}

int ckd_tree_load(const char *path, void *stream, ckd_tree **out) {
    if (!path || !out) {
        set_error("path and out must not be NULL");
        return CKD_ERR_ARG;
    }
    *out = nullptr;
    FILE *f = fopen(path, "rb");
    if (!f) {
        set_error("cannot open %s", path);
        return CKD_ERR_ARG;
    }
    PersistHeader h;
    if (fread(&h, sizeof(h), 1, f) != 1 || memcmp(h.magic, MAGIC, 8) != 0 || h.n < 1 || h.ndim < 1 ||
        h.ndim > CKD_MAX_DIM || (h.prec != 4 && h.prec != 8) || (h.idx_bits != 32 && h.idx_bits != 64) ||
        h.nnodes < 1 || h.nsets < 1) {
        fclose(f);
        set_error("%s is not a saved cosmokdtree_amd tree", path);
        return CKD_ERR_ARG;
    }
    bool bad = h.nnodes >= INT32_MAX || h.nsets > h.nnodes || h.nleaves < 1 || h.nleaves > h.nnodes ||
               h.nwide < -1 || h.nwide >= h.nnodes || h.leafsize < 1 || h.depth < 0 ||
               (h.periodic != 0 && h.periodic != 1) || (h.idx_bits == 32 && h.n >= INT32_MAX) ||
               (h.nwide >= 0 && h.n >= INT32_MAX);
    for (int d = 0; d < h.ndim && h.periodic; d++) bad |= !(h.L[d] > 0);
    if (bad) {
        fclose(f);
        set_error("%s: inconsistent tree header", path);
        return CKD_ERR_ARG;
    }
    ckd_tree *t = new ckd_tree();
    t->n = h.n;
    t->nnodes = h.nnodes;
    t->nleaves = h.nleaves;
    t->nsets = h.nsets;
    t->ndim = h.ndim;
    t->prec = h.prec;
    t->leafsize = h.leafsize;
    t->periodic = h.periodic;
    t->idx_bits = h.idx_bits;
    t->depth = h.depth;
    for (int d = 0; d < CKD_MAX_DIM; d++) t->L[d] = h.L[d];
    t->nwide = h.nwide >= 0 ? h.nwide : 0;
    (void)hipGetDevice(&t->device);
    // allocate what the header says the file holds (the wide records only if saved)
    const bool has_wide = h.nwide >= 0;
    const size_t nw = (size_t)(t->nwide > 0 ? t->nwide : 1);
    std::vector<Array> as = arrays(t);
    // sizes of the optional arrays follow presence, as at save time
    as[6].bytes = has_wide ? wide_bytes_rt(t->ndim, t->prec) * nw : 0;
    as[7].bytes = has_wide ? sizeof(int2) * (size_t)t->nsets : 0;
    as[8].bytes = has_wide ? sizeof(int2) * nw : 0;
    as[9].bytes = sizeof(int64_t) * (size_t)(t->nsets + 1);  // always saved
    std::vector<char> host;
    int rc = CKD_OK;
    int32_t depth = 0;
    WideLinks wl;
    for (size_t ai = 0; ai < as.size(); ai++) {
        Array &a = as[ai];
        if (a.bytes == 0) continue;
        host.resize(a.bytes);
        if (fread(host.data(), 1, a.bytes, f) != a.bytes) {
            set_error("%s is truncated", path);
            rc = CKD_ERR_ARG;
            break;
        }
        const std::string why = check_array((int)ai, host.data(), a.bytes, h, depth, wl);
        if (!why.empty()) {
            set_error("%s: %s", path, why.c_str());
            rc = CKD_ERR_ARG;
            break;
        }
        if ((rc = dev_alloc(a.ptr, a.bytes, (hipStream_t)stream))) break;
        hipError_t e = hipMemcpyAsync(*a.ptr, host.data(), a.bytes, hipMemcpyHostToDevice, (hipStream_t)stream);
        if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
        if (e != hipSuccess) {
            rc = hip_fail(e, "loading a tree", __FILE__, __LINE__);
            break;
        }
    }
    if (rc == CKD_OK && fgetc(f) != EOF) {
        set_error("%s: data after the last array", path);
        rc = CKD_ERR_ARG;
    }
    fclose(f);
    // the traversals size their stacks by the depth: take the one the links give
    t->depth = depth;
    if (rc == CKD_OK && has_wide) {  // parent pointers are derived, not stored
        rc = wide_parents(t, (hipStream_t)stream);
        if (rc == CKD_OK) {
            hipError_t e = hipStreamSynchronize((hipStream_t)stream);
            if (e != hipSuccess) rc = hip_fail(e, "loading a tree", __FILE__, __LINE__);
        }
    }
    if (rc != CKD_OK) {
        ckd_free(t);
        return rc;
    }
    *out = t;
    return CKD_OK;
}

}  // extern "C"
