/*
 * Binned-SAH BVH builder (host, C).
 *
 * Replaces the reference build (src/ray_tracing/bvh/bvh_build.c:17-124,
 * qshort_axis.c:88-111): that one splits at the object median of the longest
 * axis down to one primitive per leaf and recomputes bounds per node.  Here:
 * 32-bin SAH on centroids per node, leaves of at most max_leaf primitives,
 * and an object-median fallback that bounds the depth so the device stack
 * never overflows.  The output is the node / primitive layout in mrt_host.h.
 */
#include "mrt_host_internal.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NBINS 32

typedef struct {
    float lo[3], hi[3];
} box3;

/* A primitive's bounds, centroid and source index.  The build partitions
 * these records in place, so every pass over a range reads memory in order. */
typedef struct {
    box3     b;
    float    c[3];
    uint32_t id;
} prec;

typedef struct {
    const mrt_scene *sc;
    uint32_t  max_leaf;
    uint32_t  n;
    prec     *rec;       /* leaf order once built */
    mrt_bvh_node *nodes;
    uint32_t  n_nodes, cap_nodes;
    uint32_t  max_depth, n_leaves;
    double    sah;
    double    root_area;
    int       err;
} builder;

/* min / max that compile to single instructions (bounds are never NaN:
 * mrt_bvh_build rejects non-finite input) */
static inline float mnf(float a, float b) { return b < a ? b : a; }
static inline float mxf(float a, float b) { return b > a ? b : a; }

static inline void box_empty(box3 *b)
{
    for (int k = 0; k < 3; k++) {
        b->lo[k] = FLT_MAX;
        b->hi[k] = -FLT_MAX;
    }
}

static inline void box_grow(box3 *b, const box3 *o)
{
    for (int k = 0; k < 3; k++) {
        b->lo[k] = mnf(b->lo[k], o->lo[k]);
        b->hi[k] = mxf(b->hi[k], o->hi[k]);
    }
}

static inline double box_area(const box3 *b)
{
    double d[3];
    for (int k = 0; k < 3; k++) {
        d[k] = (double)b->hi[k] - (double)b->lo[k];
        if (d[k] < 0.0)
            return 0.0;
    }
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
}

/* Box of a cylinder, cone or hyperboloid: the finite solid that bounds its
 * surface (the cylinder of radius r between its caps, the cone from apex
 * to base disk, the hyperboloid's slab |z| <= height / 2 inside the
 * circular cylinder of radius max(a, b) sqrt(shape + (height/2)^2 / c^2)),
 * computed in double and widened by a relative 1e-5 so the float box holds
 * every float hit point. */
static void quadric_box(const mrt_quadric *q, box3 *b)
{
    double A[3] = {q->axis[0], q->axis[1], q->axis[2]};
    double P[3] = {q->pos[0], q->pos[1], q->pos[2]};
    double h = q->height, lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        double side = sqrt(fmax(0.0, 1.0 - A[k] * A[k]));
        if (q->kind == MRT_QUADRIC_CONE) {
            double base = P[k] + A[k] * h, rr = (double)q->radius * side;
            lo[k] = fmin(P[k], base - rr);
            hi[k] = fmax(P[k], base + rr);
        } else {
            double rad = q->radius;
            if (q->kind == MRT_QUADRIC_HYPERBOLOID) {
                double hc = 0.5 * h / q->c;
                rad = fmax(q->a, q->b) * sqrt(fmax(0.0, q->shape + hc * hc));
            }
            double e = fabs(A[k]) * 0.5 * h + rad * side;
            lo[k] = P[k] - e;
            hi[k] = P[k] + e;
        }
    }
    double scale = 0.0;
    for (int k = 0; k < 3; k++)
        scale = fmax(scale, fmax(fabs(lo[k]), fabs(hi[k])));
    double pad = 1e-5 * (scale + h + q->radius + q->a + q->b + q->c);
    for (int k = 0; k < 3; k++) {
        b->lo[k] = nextafterf((float)(lo[k] - pad), -INFINITY);
        b->hi[k] = nextafterf((float)(hi[k] + pad), INFINITY);
    }
}

/* Fills the primitive records; MRT_E_INVALID for a non-finite vertex,
 * centre or radius. */
static int prim_bounds(builder *B)
{
    const mrt_scene *s = B->sc;
    int finite = 1;
    for (uint32_t i = 0; i < s->n_triangles; i++) {
        const float *v = s->tri_pos + (size_t)i * 9;
        prec *r = &B->rec[i];
        for (int k = 0; k < 3; k++) {
            r->b.lo[k] = mnf(v[k], mnf(v[3 + k], v[6 + k]));
            r->b.hi[k] = mxf(v[k], mxf(v[3 + k], v[6 + k]));
            r->c[k] = 0.5f * (r->b.lo[k] + r->b.hi[k]);
            finite &= isfinite(v[k]) && isfinite(v[3 + k]) && isfinite(v[6 + k]);
        }
        r->id = i;
    }
    for (uint32_t j = 0; j < s->n_spheres; j++) {
        const float *q = s->sph + (size_t)j * 4;
        uint32_t i = s->n_triangles + j;
        prec *r = &B->rec[i];
        for (int k = 0; k < 3; k++) {
            /* one ulp outward so the float box holds the exact sphere */
            r->b.lo[k] = nextafterf(q[k] - q[3], -INFINITY);
            r->b.hi[k] = nextafterf(q[k] + q[3], INFINITY);
            r->c[k] = q[k];
            finite &= isfinite(r->b.lo[k]) && isfinite(r->b.hi[k]);
        }
        finite &= q[3] >= 0.0f;
        r->id = i;
    }
    for (uint32_t j = 0; j < s->n_quadrics; j++) {
        uint32_t i = s->n_triangles + s->n_spheres + j;
        prec *r = &B->rec[i];
        quadric_box(&s->quadrics[j], &r->b);
        for (int k = 0; k < 3; k++) {
            r->c[k] = 0.5f * (r->b.lo[k] + r->b.hi[k]);
            finite &= isfinite(r->b.lo[k]) && isfinite(r->b.hi[k]);
        }
        r->id = i;
    }
    return finite ? MRT_OK : MRT_E_INVALID;
}

static int32_t leaf_code(uint32_t first, uint32_t count)
{
    return ~(int32_t)((first << 3) | (count - 1));
}

static uint32_t new_node(builder *B)
{
    if (B->n_nodes == B->cap_nodes) {
        uint32_t nc = B->cap_nodes ? B->cap_nodes * 2 : 1024;
        mrt_bvh_node *nn = realloc(B->nodes, (size_t)nc * sizeof(*nn));
        if (!nn) {
            B->err = MRT_E_NOMEM;
            return UINT32_MAX;
        }
        B->nodes = nn;
        B->cap_nodes = nc;
    }
    memset(&B->nodes[B->n_nodes], 0, sizeof(mrt_bvh_node));
    return B->n_nodes++;
}

static void set_child(builder *B, uint32_t node, int side, const box3 *b,
                      int32_t code)
{
    mrt_bvh_node *n = &B->nodes[node];
    float *lo = side ? n->lo1 : n->lo0;
    float *hi = side ? n->hi1 : n->hi0;
    memcpy(lo, b->lo, sizeof(b->lo));
    memcpy(hi, b->hi, sizeof(b->hi));
    n->child[side] = code;
}

static uint32_t ceil_log2(uint32_t x)
{
    uint32_t l = 0;
    while ((1u << l) < x)
        l++;
    return l;
}

/* comparator state of the median fallback: per thread, since scenes of a
 * batch are built concurrently */
static _Thread_local int cmp_axis;
static int cmp_prim(const void *a, const void *b)
{
    const prec *ra = a, *rb = b;
    float ca = ra->c[cmp_axis], cb = rb->c[cmp_axis];
    if (ca < cb)
        return -1;
    if (ca > cb)
        return 1;
    return (ra->id > rb->id) - (ra->id < rb->id);
}

static inline void cbox_grow(box3 *b, const float *c)
{
    for (int k = 0; k < 3; k++) {
        b->lo[k] = mnf(b->lo[k], c[k]);
        b->hi[k] = mxf(b->hi[k], c[k]);
    }
}

static inline int bin_of(float c, float lo, float scale)
{
    int bin = (int)((c - lo) * scale);
    if (bin >= NBINS)
        bin = NBINS - 1;
    if (bin < 0)
        bin = 0;
    return bin;
}

typedef struct {
    box3     bb, cb;          /* primitive and centroid bounds of the bin */
    uint32_t cnt;
} bin_t;

/* Splits rec[begin,end) (centroid bounds cb): 32-bin SAH over the three
 * axes binned in one pass, or the object median on the longest centroid
 * axis near the depth limit or when no SAH split separates the range.
 * Returns mid (begin < mid < end) and the primitive and centroid bounds of
 * both halves. */
static uint32_t split_range(builder *B, uint32_t begin, uint32_t end,
                            uint32_t depth, const box3 *cb, box3 *lb,
                            box3 *rb, box3 *lcb, box3 *rcb)
{
    uint32_t count = end - begin;
    int axis_long = 0;
    float ext_long = cb->hi[0] - cb->lo[0];
    for (int k = 1; k < 3; k++)
        if (cb->hi[k] - cb->lo[k] > ext_long) {
            ext_long = cb->hi[k] - cb->lo[k];
            axis_long = k;
        }
    uint32_t leaves_needed = (count + B->max_leaf - 1) / B->max_leaf;
    int median = (depth + 1 + ceil_log2(leaves_needed) >= MRT_BVH_MAX_DEPTH)
                 || !(ext_long > 0.0f);
    if (!median) {
        bin_t bins[3][NBINS];
        float scale[3];
        int live[3];
        for (int axis = 0; axis < 3; axis++) {
            float ext = cb->hi[axis] - cb->lo[axis];
            live[axis] = ext > 0.0f;
            scale[axis] = live[axis] ? (float)NBINS / ext : 0.0f;
            for (int k = 0; k < NBINS; k++) {
                box_empty(&bins[axis][k].bb);
                box_empty(&bins[axis][k].cb);
                bins[axis][k].cnt = 0;
            }
        }
        for (uint32_t i = begin; i < end; i++) {
            const float *c = B->rec[i].c;
            const box3 *pb = &B->rec[i].b;
            for (int axis = 0; axis < 3; axis++) {
                if (!live[axis])
                    continue;
                bin_t *bn = &bins[axis][bin_of(c[axis], cb->lo[axis], scale[axis])];
                bn->cnt++;
                box_grow(&bn->bb, pb);
                cbox_grow(&bn->cb, c);
            }
        }
        double best = INFINITY;
        int best_axis = -1, best_bin = -1;
        for (int axis = 0; axis < 3; axis++) {
            if (!live[axis])
                continue;
            double right_area[NBINS];
            uint32_t right_cnt[NBINS];
            box3 acc;
            box_empty(&acc);
            uint32_t cnt = 0;
            for (int k = NBINS - 1; k > 0; k--) {
                box_grow(&acc, &bins[axis][k].bb);
                cnt += bins[axis][k].cnt;
                right_area[k] = box_area(&acc);
                right_cnt[k] = cnt;
            }
            box_empty(&acc);
            cnt = 0;
            for (int k = 0; k < NBINS - 1; k++) {
                box_grow(&acc, &bins[axis][k].bb);
                cnt += bins[axis][k].cnt;
                if (cnt == 0 || right_cnt[k + 1] == 0)
                    continue;
                double cost = box_area(&acc) * cnt
                              + right_area[k + 1] * right_cnt[k + 1];
                if (cost < best) {
                    best = cost;
                    best_axis = axis;
                    best_bin = k;
                }
            }
        }
        if (best_axis >= 0) {
            const float lo = cb->lo[best_axis], sc = scale[best_axis];
            uint32_t i = begin, j = end;
            while (i < j) {
                if (bin_of(B->rec[i].c[best_axis], lo, sc) <= best_bin) {
                    i++;
                } else {
                    j--;
                    prec t = B->rec[i];
                    B->rec[i] = B->rec[j];
                    B->rec[j] = t;
                }
            }
            if (i > begin && i < end) {
                box_empty(lb);
                box_empty(rb);
                box_empty(lcb);
                box_empty(rcb);
                for (int k = 0; k < NBINS; k++) {
                    const bin_t *bn = &bins[best_axis][k];
                    box_grow(k <= best_bin ? lb : rb, &bn->bb);
                    box_grow(k <= best_bin ? lcb : rcb, &bn->cb);
                }
                return i;
            }
        }
    }
    /* object median on the longest centroid axis (ties by index) */
    cmp_axis = axis_long;
    qsort(B->rec + begin, count, sizeof(prec), cmp_prim);
    uint32_t mid = begin + count / 2;
    box_empty(lb);
    box_empty(rb);
    box_empty(lcb);
    box_empty(rcb);
    for (uint32_t i = begin; i < end; i++) {
        box_grow(i < mid ? lb : rb, &B->rec[i].b);
        cbox_grow(i < mid ? lcb : rcb, B->rec[i].c);
    }
    return mid;
}

static void range_box(builder *B, uint32_t begin, uint32_t end, box3 *out)
{
    box_empty(out);
    for (uint32_t i = begin; i < end; i++)
        box_grow(out, &B->rec[i].b);
}

static void range_cbox(builder *B, uint32_t begin, uint32_t end, box3 *out)
{
    box_empty(out);
    for (uint32_t i = begin; i < end; i++)
        cbox_grow(out, B->rec[i].c);
}

/* Builds the subtree of rec[begin,end) whose box is bx and centroid box cb;
 * returns its child code (leaf or inner node index). depth = inner levels
 * above it. */
static int32_t build_rec(builder *B, uint32_t begin, uint32_t end,
                         const box3 *bx, const box3 *cb, uint32_t depth)
{
    uint32_t count = end - begin;
    if (B->err)
        return 0;
    if (count <= B->max_leaf) {
        B->n_leaves++;
        B->sah += box_area(bx) / B->root_area * (double)count;
        return leaf_code(begin, count);
    }
    uint32_t node = new_node(B);
    if (node == UINT32_MAX)
        return 0;
    if (depth + 1 > B->max_depth)
        B->max_depth = depth + 1;
    B->sah += box_area(bx) / B->root_area;
    box3 lb, rb, lcb, rcb;
    uint32_t mid = split_range(B, begin, end, depth, cb, &lb, &rb, &lcb, &rcb);
    int32_t lc = build_rec(B, begin, mid, &lb, &lcb, depth + 1);
    set_child(B, node, 0, &lb, lc);
    int32_t rc = build_rec(B, mid, end, &rb, &rcb, depth + 1);
    set_child(B, node, 1, &rb, rc);
    return (int32_t)node;
}

static inline float bits_f(uint32_t u)
{
    float f;
    memcpy(&f, &u, sizeof(f));
    return f;
}

int mrt_bvh_build(const mrt_scene *s, uint32_t max_leaf, mrt_bvh *out)
{
    if (!s || !out || max_leaf < 1 || max_leaf > MRT_BVH_MAX_LEAF)
        return MRT_E_INVALID;
    memset(out, 0, sizeof(*out));
    if ((s->n_triangles && (!s->tri_pos || !s->tri_mat))
        || (s->n_spheres && (!s->sph || !s->sph_mat)))
        return MRT_E_INVALID;
    if (s->n_quadrics && !s->quadrics)
        return MRT_E_INVALID;
    uint64_t n64 = (uint64_t)s->n_triangles + s->n_spheres + s->n_quadrics;
    if (n64 >= (1u << 28))
        return MRT_E_LIMIT;
    uint32_t n = (uint32_t)n64;
    if (n == 0)
        return MRT_OK;

    builder B;
    memset(&B, 0, sizeof(B));
    B.sc = s;
    B.max_leaf = max_leaf;
    B.n = n;
    B.rec = malloc((size_t)n * sizeof(prec));
    if (!B.rec)
        return MRT_E_NOMEM;
    if (prim_bounds(&B) != MRT_OK) {
        free(B.rec);
        return MRT_E_INVALID;
    }
    box3 root;
    range_box(&B, 0, n, &root);
    B.root_area = box_area(&root);
    if (!(B.root_area > 0.0))
        B.root_area = 1.0;

    int32_t rc;
    if (n <= max_leaf) {
        /* a single leaf: one inner node with an absent second child */
        uint32_t node = new_node(&B);
        B.max_depth = 1;
        B.n_leaves = 1;
        B.sah = 1.0 + (double)n;
        set_child(&B, node, 0, &root, leaf_code(0, n));
        box3 far;
        for (int k = 0; k < 3; k++)
            far.lo[k] = far.hi[k] = 1e30f;
        set_child(&B, node, 1, &far, leaf_code(0, 1));
        rc = 0;
    } else {
        box3 croot;
        range_cbox(&B, 0, n, &croot);
        rc = build_rec(&B, 0, n, &root, &croot, 0);
    }
    (void)rc;
    if (B.err) {
        free(B.nodes);
        free(B.rec);
        return B.err;
    }

    out->prims = malloc((size_t)n * sizeof(mrt_prim));
    out->shade_nrm = malloc((size_t)n * 9 * sizeof(float));
    if (!out->prims || !out->shade_nrm) {
        free(out->prims);
        free(out->shade_nrm);
        free(B.nodes);
        free(B.rec);
        memset(out, 0, sizeof(*out));
        return MRT_E_NOMEM;
    }
    for (uint32_t i = 0; i < n; i++) {
        uint32_t p = B.rec[i].id;
        mrt_prim *r = &out->prims[i];
        float *nrm = out->shade_nrm + (size_t)i * 9;
        if (p < s->n_triangles) {
            const float *v = s->tri_pos + (size_t)p * 9;
            for (int k = 0; k < 3; k++) {
                r->a[k] = v[k];
                r->b[k] = v[3 + k] - v[k];
                r->c[k] = v[6 + k] - v[k];
            }
            r->a[3] = bits_f(MRT_PRIM_TRIANGLE | p);
            r->b[3] = bits_f(MRT_PRIMKIND_TRIANGLE);
            r->c[3] = bits_f(s->tri_mat[p]);
            if (s->tri_nrm) {
                memcpy(nrm, s->tri_nrm + (size_t)p * 9, 9 * sizeof(float));
            } else {
                double e1[3], e2[3];
                for (int k = 0; k < 3; k++) {
                    e1[k] = r->b[k];
                    e2[k] = r->c[k];
                }
                double c[3] = {e1[1] * e2[2] - e1[2] * e2[1],
                               e1[2] * e2[0] - e1[0] * e2[2],
                               e1[0] * e2[1] - e1[1] * e2[0]};
                double l = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
                l = l > 0.0 ? 1.0 / l : 0.0;
                for (int k = 0; k < 3; k++)
                    for (int j = 0; j < 3; j++)
                        nrm[k * 3 + j] = (float)(c[j] * l);
            }
        } else if (p >= s->n_triangles + s->n_spheres) {
            uint32_t j = p - s->n_triangles - s->n_spheres;
            const mrt_quadric *q = &s->quadrics[j];
            for (int k = 0; k < 3; k++) {
                r->a[k] = q->pos[k];
                r->b[k] = q->axis[k];
            }
            r->a[3] = bits_f(MRT_PRIM_QUADRIC | j);
            r->b[3] = bits_f(MRT_PRIMKIND_QUADRIC + q->kind);
            int hyp = q->kind == MRT_QUADRIC_HYPERBOLOID;
            r->c[0] = hyp ? q->a : q->radius;
            r->c[1] = hyp ? q->b : 0.0f;
            r->c[2] = hyp ? q->c : 0.0f;
            r->c[3] = q->height;
            memset(nrm, 0, 9 * sizeof(float));
            nrm[0] = hyp ? q->shape : 0.0f;
            nrm[1] = bits_f(q->mat);
            nrm[2] = bits_f(q->mat_top);
            nrm[3] = bits_f(q->mat_bot);
        } else {
            uint32_t j = p - s->n_triangles;
            const float *q = s->sph + (size_t)j * 4;
            r->a[0] = q[0];
            r->a[1] = q[1];
            r->a[2] = q[2];
            r->a[3] = bits_f(MRT_PRIM_SPHERE | j);
            r->b[0] = q[3];
            r->b[1] = q[3] * q[3];
            r->b[2] = 0.0f;
            r->b[3] = bits_f(MRT_PRIMKIND_SPHERE);
            r->c[0] = r->c[1] = r->c[2] = 0.0f;
            r->c[3] = bits_f(s->sph_mat[j]);
            memset(nrm, 0, 9 * sizeof(float));
        }
    }
    free(B.rec);
    out->nodes = B.nodes;
    out->n_nodes = B.n_nodes;
    out->n_prims = n;
    out->depth = B.max_depth;
    out->n_leaves = B.n_leaves;
    out->sah_cost = B.sah;
    return MRT_OK;
}

void mrt_bvh_free(mrt_bvh *b)
{
    if (!b)
        return;
    free(b->nodes);
    free(b->prims);
    free(b->shade_nrm);
    memset(b, 0, sizeof(*b));
}

/* ---- 4-wide collapse ----------------------------------------------------- */

typedef struct {
    int32_t code;
    float   lo[3], hi[3];
} child4;

typedef struct {
    const mrt_bvh *b;
    mrt_bvh4_node *nodes;
    uint32_t       n, cap, depth;
    int            err;
} col4;

static double child_area(const child4 *c)
{
    double dx = (double)c->hi[0] - c->lo[0], dy = (double)c->hi[1] - c->lo[1],
           dz = (double)c->hi[2] - c->lo[2];
    if (dx < 0 || dy < 0 || dz < 0)
        return 0.0;
    return dx * dy + dy * dz + dz * dx;
}

static void bin_children(const mrt_bvh_node *n, child4 *a, child4 *b)
{
    a->code = n->child[0];
    b->code = n->child[1];
    memcpy(a->lo, n->lo0, sizeof(a->lo));
    memcpy(a->hi, n->hi0, sizeof(a->hi));
    memcpy(b->lo, n->lo1, sizeof(b->lo));
    memcpy(b->hi, n->hi1, sizeof(b->hi));
}

/* Builds the 4-wide node for binary inner node `code`; returns its index. */
static int32_t collapse4(col4 *C, int32_t code, uint32_t level)
{
    if (C->err)
        return 0;
    if (level > C->depth)
        C->depth = level;
    child4 ch[4];
    int nc = 2;
    bin_children(&C->b->nodes[code], &ch[0], &ch[1]);
    while (nc < 4) {
        int best = -1;
        double ba = -1.0;
        for (int i = 0; i < nc; i++)
            if (ch[i].code >= 0 && child_area(&ch[i]) > ba) {
                ba = child_area(&ch[i]);
                best = i;
            }
        if (best < 0)
            break;
        child4 a, b;
        bin_children(&C->b->nodes[ch[best].code], &a, &b);
        for (int i = nc; i > best + 1; i--)
            ch[i] = ch[i - 1];
        ch[best] = a;
        ch[best + 1] = b;
        nc++;
    }
    if (C->n == C->cap) {
        uint32_t nc2 = C->cap ? C->cap * 2 : 1024;
        mrt_bvh4_node *p = realloc(C->nodes, (size_t)nc2 * sizeof(*p));
        if (!p) {
            C->err = MRT_E_NOMEM;
            return 0;
        }
        C->nodes = p;
        C->cap = nc2;
    }
    const uint32_t k = C->n++;
    mrt_bvh4_node node;
    memset(&node, 0, sizeof(node));
    for (int i = 0; i < 4; i++) {
        if (i >= nc) {
            node.lo_x[i] = node.hi_x[i] = 1e30f;
            node.lo_y[i] = node.hi_y[i] = 1e30f;
            node.lo_z[i] = node.hi_z[i] = 1e30f;
            node.child[i] = ~(int32_t)0;       /* leaf (0, 1), never reached */
            continue;
        }
        node.lo_x[i] = ch[i].lo[0];
        node.hi_x[i] = ch[i].hi[0];
        node.lo_y[i] = ch[i].lo[1];
        node.hi_y[i] = ch[i].hi[1];
        node.lo_z[i] = ch[i].lo[2];
        node.hi_z[i] = ch[i].hi[2];
        node.child[i] = ch[i].code;
    }
    C->nodes[k] = node;
    for (int i = 0; i < nc; i++)
        if (ch[i].code >= 0) {
            int32_t sub = collapse4(C, ch[i].code, level + 1);
            C->nodes[k].child[i] = sub;
        }
    return (int32_t)k;
}

/* Renumbers the nodes so that the first n_top of a breadth-first walk from
 * the root come first, in that order, and the others follow in their
 * pre-order: the top levels then form one contiguous block at the front,
 * which a kernel can copy into LDS with one linear loop. */
static int top_first(mrt_bvh4_node **nodes, uint32_t n, uint32_t n_top)
{
    if (n <= 1)
        return MRT_OK;
    if (n_top > n)
        n_top = n;
    uint32_t *order = malloc((size_t)n_top * sizeof(uint32_t));
    uint32_t *nid = malloc((size_t)n * sizeof(uint32_t));
    mrt_bvh4_node *out = malloc((size_t)n * sizeof(mrt_bvh4_node));
    if (!order || !nid || !out) {
        free(order);
        free(nid);
        free(out);
        return MRT_E_NOMEM;
    }
    for (uint32_t k = 0; k < n; k++)
        nid[k] = UINT32_MAX;
    const mrt_bvh4_node *in = *nodes;
    uint32_t head = 0, tail = 0;
    order[tail++] = 0;
    nid[0] = 0;
    while (head < tail && tail < n_top) {
        const mrt_bvh4_node *nd = &in[order[head++]];
        for (int i = 0; i < 4 && tail < n_top; i++)
            if (nd->child[i] >= 0) {
                nid[nd->child[i]] = tail;
                order[tail++] = (uint32_t)nd->child[i];
            }
    }
    uint32_t next = tail;
    for (uint32_t k = 0; k < n; k++)
        if (nid[k] == UINT32_MAX)
            nid[k] = next++;
    for (uint32_t k = 0; k < n; k++) {
        mrt_bvh4_node nd = in[k];
        for (int i = 0; i < 4; i++)
            if (nd.child[i] >= 0)
                nd.child[i] = (int32_t)nid[nd.child[i]];
        out[nid[k]] = nd;
    }
    free(order);
    free(nid);
    free(*nodes);
    *nodes = out;
    return MRT_OK;
}

int mrt_bvh4_build(const mrt_bvh *b, mrt_bvh4 *out)
{
    if (!b || !out)
        return MRT_E_INVALID;
    memset(out, 0, sizeof(*out));
    if (b->n_nodes == 0)
        return MRT_OK;
    col4 C;
    memset(&C, 0, sizeof(C));
    C.b = b;
    (void)collapse4(&C, 0, 1);
    if (!C.err)
        C.err = top_first(&C.nodes, C.n, MRT_BVH4_TOP);
    if (C.err) {
        free(C.nodes);
        return C.err;
    }
    out->nodes = C.nodes;
    out->n_nodes = C.n;
    out->depth = C.depth;
    return MRT_OK;
}

void mrt_bvh4_free(mrt_bvh4 *b)
{
    if (!b)
        return;
    free(b->nodes);
    memset(b, 0, sizeof(*b));
}
