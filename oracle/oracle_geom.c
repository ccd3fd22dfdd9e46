/*
 * oracle_geom.c -- CPU restatement of Rumi-Renderer's intersection and BVH
 * code.  TEST INFRASTRUCTURE ONLY: the checker for the GPU library, never
 * linked into it.  See oracle.h for what pins each function.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, so the FP32 operation
 * order written below is the order executed).
 */
#include <pthread.h>
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------ */
/* Vector helpers with the reference's operation order
 * (src/Kernel/vector.h:34-68).                                        */

static inline float dot3(const float* u, const float* v)
{
    return u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
}

static inline void cross3(const float* u, const float* v, float* c)
{
    float cx = u[1] * v[2] - v[1] * u[2];
    float cy = u[2] * v[0] - v[2] * u[0];
    float cz = u[0] * v[1] - v[0] * u[1];
    c[0] = cx;
    c[1] = cy;
    c[2] = cz;
}

static inline void sub3(const float* a, const float* b, float* c)
{
    c[0] = a[0] - b[0];
    c[1] = a[1] - b[1];
    c[2] = a[2] - b[2];
}

/* ------------------------------------------------------------------ */
/* Triangle::Intersect (src/Shape/triangle.cpp:41-118).  Moller-Trumbore with
 * the reference's rejections: |det| < 1e-8, u outside [0,1], v < 0,
 * u + v > 1, t < eps, (tmin > eps && t > tmin), tmax - t < eps.          */
int or_triangle_intersect(const float v0[3], const float v1[3], const float v2[3],
                          or_ray* ray, float* u_out, float* v_out, float* det_out)
{
    float e1[3], e2[3], p[3], tv[3], q[3];
    sub3(v1, v0, e1);
    sub3(v2, v0, e2);
    cross3(ray->d, e2, p);
    float det = dot3(p, e1);
    if (-1e-8f < det && det < 1e-8f)
        return 0;
    float inv_det = 1.0f / det;
    sub3(ray->o, v0, tv);
    float u = dot3(p, tv) * inv_det;
    if (u < 0.0f || u > 1.0f)
        return 0;
    cross3(tv, e1, q);
    float v = dot3(q, ray->d) * inv_det;
    if (v < 0.0f || u + v > 1.0f)
        return 0;
    float t = dot3(q, e2) * inv_det;
    if (t < OR_EPS)
        return 0;
    if (ray->tmin > OR_EPS && t > ray->tmin)
        return 0;
    if (ray->tmax - t < OR_EPS)
        return 0;
    ray->tmin = t;
    if (u_out) *u_out = u;
    if (v_out) *v_out = v;
    if (det_out) *det_out = det;
    return 1;
}

/* SolveQuadratic (src/Kernel/clibrary.h:62-87). */
static int solve_quadratic(float A, float B, float C, float* t0, float* t1)
{
    float disc = B * B - 4 * A * C;
    if (disc < 0.0f)
        return 0;
    float inv_2a = 1.0f / (2.0f * A);
    if (disc == 0.0f || disc == -0.0f) {
        *t0 = *t1 = -B * inv_2a;
        return 1;
    }
    float s = sqrtf(disc);
    *t0 = (-B - s) * inv_2a;
    *t1 = (-B + s) * inv_2a;
    return 1;
}

/* Length (src/Kernel/vector.h:39-42): the C++ float overload of sqrt. */
static inline float length3(const float* v)
{
    return sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

/* Sphere::Intersect (src/Shape/sphere.cpp:51-132).  A = 1 (unit
 * direction assumed); far root first when the ray is inside the object or
 * for a textured emitter seen from inside; reject t0 < eps, tmax - t0 <
 * eps, (t0 > tmin && tmin > 0).                                          */
int or_sphere_intersect(const or_prim* s, or_ray* ray, int in_object)
{
    float oc[3];
    sub3(ray->o, s->a, oc);
    float A = 1.0f;
    float B = 2.0f * dot3(ray->d, oc);
    float C = dot3(oc, oc) - s->r * s->r;
    float t0 = 0.0f, t1 = 0.0f;
    if (!solve_quadratic(A, B, C, &t0, &t1))
        return 0;
    if (t0 < 0.0f && t1 < 0.0f)
        return 0;
    if (in_object) {
        float tmp = t0;
        t0 = t1;
        t1 = tmp;
    }
    if (s->le_textured && length3(oc) < s->r) {
        float tmp = t0;
        t0 = t1;
        t1 = tmp;
    }
    if (t0 < OR_EPS)
        return 0;
    if (ray->tmax - t0 < OR_EPS)
        return 0;
    if (t0 > ray->tmin && ray->tmin > 0.0f)
        return 0;
    ray->tmin = t0;
    return 1;
}

/* Rect::Intersect (src/Shape/rect.cpp:51-122).  Back faces are culled when
 * the material emits nothing (0 < d.n < 1).                              */
int or_rect_intersect(const or_prim* q, or_ray* ray)
{
    float cost = dot3(ray->d, q->n);
    if (cost == 0.0f || cost == -0.0f)
        return 0;
    float op[3];
    sub3(ray->o, q->a, op);
    float t = -dot3(op, q->n) / cost;
    if (t < 0.0f)
        return 0;
    if (t < OR_EPS)
        return 0;
    if (ray->tmax - t < OR_EPS)
        return 0;
    if (!q->emissive && (0.0f < cost && cost < 1.0f))
        return 0;
    if (t > ray->tmin && ray->tmin > 0.0f)
        return 0;
    float ip[3], onp[3];
    ip[0] = ray->o[0] + ray->d[0] * t;
    ip[1] = ray->o[1] + ray->d[1] * t;
    ip[2] = ray->o[2] + ray->d[2] * t;
    sub3(ip, q->a, onp);
    float il = dot3(onp, q->b) * (q->inv_area * q->width);
    if (il < 0.0f || il > q->length)
        return 0;
    float iw = dot3(onp, q->c) * (q->inv_area * q->length);
    if (iw < 0.0f || iw > q->width)
        return 0;
    ray->tmin = t;
    return 1;
}

int or_prim_intersect(const or_prim* p, or_ray* ray, int in_object)
{
    switch (p->kind) {
    case OR_TRI:
        return or_triangle_intersect(p->a, p->b, p->c, ray, 0, 0, 0);
    case OR_SPHERE:
        return or_sphere_intersect(p, ray, in_object);
    case OR_RECT:
        return or_rect_intersect(p, ray);
    }
    return 0;
}

/* BBox::Intersect (src/Kernel/bbox.cpp:30-92): x slab initialises the
 * interval, y and z reject when disjoint and then clamp.               */
int or_bbox_intersect(const float lo[3], const float hi[3], const or_ray* ray,
                      float* tnear, float* tfar)
{
    float ix = 1.0f / ray->d[0];
    float iy = 1.0f / ray->d[1];
    float iz = 1.0f / ray->d[2];
    float t0 = (lo[0] - ray->o[0]) * ix;
    float t1 = (hi[0] - ray->o[0]) * ix;
    float tmp;
    if (t0 > t1) { tmp = t0; t0 = t1; t1 = tmp; }
    *tnear = t0;
    *tfar = t1;
    t0 = (lo[1] - ray->o[1]) * iy;
    t1 = (hi[1] - ray->o[1]) * iy;
    if (t0 > t1) { tmp = t0; t0 = t1; t1 = tmp; }
    if (*tnear > t1 || t0 > *tfar)
        return 0;
    if (*tnear < t0) *tnear = t0;
    if (*tfar > t1) *tfar = t1;
    t0 = (lo[2] - ray->o[2]) * iz;
    t1 = (hi[2] - ray->o[2]) * iz;
    if (t0 > t1) { tmp = t0; t0 = t1; t1 = tmp; }
    if (*tnear > t1 || t0 > *tfar)
        return 0;
    if (*tnear < t0) *tnear = t0;
    if (*tfar > t1) *tfar = t1;
    return 1;
}

/* Triangle::Bound (triangle.cpp:120-148), Sphere::Bound (sphere.cpp:139-145),
 * Rect::Bound (rect.cpp:141-157: two opposite corners only).            */
void or_prim_bound(const or_prim* p, float lo[3], float hi[3])
{
    int k;
    if (p->kind == OR_TRI) {
        for (k = 0; k < 3; ++k) {
            lo[k] = p->a[k] < p->b[k] ? p->a[k] : p->b[k];
            if (lo[k] > p->c[k]) lo[k] = p->c[k];
            hi[k] = p->a[k] > p->b[k] ? p->a[k] : p->b[k];
            if (hi[k] < p->c[k]) hi[k] = p->c[k];
        }
    } else if (p->kind == OR_SPHERE) {
        for (k = 0; k < 3; ++k) {
            lo[k] = p->a[k] - p->r;
            hi[k] = p->a[k] + p->r;
        }
    } else {
        /* another_p = _p + _vect_l + _vect_w, evaluated left to right */
        float ap[3];
        for (k = 0; k < 3; ++k) ap[k] = (p->a[k] + p->b[k]) + p->c[k];
        for (k = 0; k < 3; ++k) {
            lo[k] = p->a[k] < ap[k] ? p->a[k] : ap[k];
            hi[k] = p->a[k] > ap[k] ? p->a[k] : ap[k];
        }
    }
}

/* ------------------------------------------------------------------ */
/* Median-6 BVH (src/Acceleration/bvh.cpp:48-148). */

typedef struct {
    float lo[3], hi[3];
} box_t;

typedef struct {
    or_bvh* out;
    int cap;
    int* objs;     /* permuted primitive list (the reference's _object_list) */
    box_t* bounds; /* permuted bounds, moved together with objs */
} build_ctx;

static box_t box_empty(void)
{
    box_t b;
    int k;
    for (k = 0; k < 3; ++k) {
        b.lo[k] = FLT_MAX;
        b.hi[k] = -FLT_MAX;
    }
    return b;
}

/* Combine (src/Kernel/bbox.cpp:127-140) with MSVC __min/__max semantics. */
static box_t box_combine(const box_t* a, const box_t* b)
{
    box_t r;
    int k;
    for (k = 0; k < 3; ++k) {
        r.lo[k] = a->lo[k] < b->lo[k] ? a->lo[k] : b->lo[k];
        r.hi[k] = a->hi[k] > b->hi[k] ? a->hi[k] : b->hi[k];
    }
    return r;
}

static int new_node(build_ctx* c)
{
    or_bvh* o = c->out;
    if (o->n_nodes == c->cap) {
        int nc = c->cap ? 2 * c->cap : 64;
        o->lo = (float*)realloc(o->lo, sizeof(float) * 3 * nc);
        o->hi = (float*)realloc(o->hi, sizeof(float) * 3 * nc);
        o->left = (int*)realloc(o->left, sizeof(int) * nc);
        o->right = (int*)realloc(o->right, sizeof(int) * nc);
        o->first = (int*)realloc(o->first, sizeof(int) * nc);
        o->count = (int*)realloc(o->count, sizeof(int) * nc);
        c->cap = nc;
    }
    int id = o->n_nodes++;
    o->left[id] = o->right[id] = -1;
    o->first[id] = o->count[id] = 0;
    return id;
}

static void set_box(or_bvh* o, int id, const box_t* b)
{
    memcpy(o->lo + 3 * id, b->lo, sizeof(float) * 3);
    memcpy(o->hi + 3 * id, b->hi, sizeof(float) * 3);
}

/* BBox::Center()[dim] (bbox.cpp:94-97). */
static inline float centre(const box_t* b, int dim)
{
    return (b->hi[dim] + b->lo[dim]) * 0.5f;
}

static int create_node(build_ctx* c, int start, int end, int pre_dim, box_t* out_box)
{
    box_t ext = box_empty();
    int i;
    for (i = start; i <= end; ++i)
        ext = box_combine(&ext, &c->bounds[i]);
    int count = end - start + 1;
    float dx = ext.hi[0] - ext.lo[0];
    float dy = ext.hi[1] - ext.lo[1];
    float dz = ext.hi[2] - ext.lo[2];
    int dim = 0;
    if (dx < dy && dz < dy) dim = 1;
    if (dx < dz && dy < dz) dim = 2;
    if (dim != pre_dim) {
        /* Shell sort with halving gaps; element moves step by step. */
        int gap, ti, ji;
        for (gap = count >> 1; gap >= 1; gap >>= 1) {
            for (ti = start + gap; ti <= end; ++ti) {
                int ts = c->objs[ti];
                box_t tb = c->bounds[ti];
                float tdim = centre(&tb, dim);
                for (ji = ti - gap; ji >= start; ji -= gap) {
                    if (centre(&c->bounds[ji], dim) > tdim) {
                        c->objs[ji + gap] = c->objs[ji];
                        c->bounds[ji + gap] = c->bounds[ji];
                        c->objs[ji] = ts;
                        c->bounds[ji] = tb;
                    } else {
                        break;
                    }
                }
            }
        }
    }
    int node = new_node(c);
    if (count > 6) {
        int cut = (start + end) >> 1;
        box_t lb, rb;
        int l = create_node(c, start, cut, dim, &lb);
        int r = create_node(c, cut + 1, end, dim, &rb);
        box_t nb = box_combine(&lb, &rb);
        c->out->left[node] = l;
        c->out->right[node] = r;
        set_box(c->out, node, &nb);
        *out_box = nb;
    } else {
        c->out->first[node] = start;
        c->out->count[node] = count;
        set_box(c->out, node, &ext);
        *out_box = ext;
    }
    return node;
}

int or_bvh_build_median6(const float* prim_lo, const float* prim_hi, int n, or_bvh* out)
{
    memset(out, 0, sizeof(*out));
    out->n_prims = n;
    out->root = -1;
    if (n <= 0)
        return 0;
    build_ctx c;
    c.out = out;
    c.cap = 0;
    c.objs = (int*)malloc(sizeof(int) * n);
    c.bounds = (box_t*)malloc(sizeof(box_t) * n);
    int i;
    for (i = 0; i < n; ++i) {
        c.objs[i] = i;
        memcpy(c.bounds[i].lo, prim_lo + 3 * i, sizeof(float) * 3);
        memcpy(c.bounds[i].hi, prim_hi + 3 * i, sizeof(float) * 3);
    }
    box_t rb;
    out->root = create_node(&c, 0, n - 1, 3, &rb);
    out->order = c.objs;
    free(c.bounds);
    return 0;
}

void or_bvh_free(or_bvh* b)
{
    free(b->lo);
    free(b->hi);
    free(b->left);
    free(b->right);
    free(b->first);
    free(b->count);
    free(b->order);
    memset(b, 0, sizeof(*b));
}

int or_bvh_nodes(const or_bvh* b) { return b->n_nodes; }

void or_bvh_copy(const or_bvh* b, float* lo, float* hi, int* left, int* right,
                 int* first, int* count, int* order, int* root)
{
    if (lo) memcpy(lo, b->lo, sizeof(float) * 3 * b->n_nodes);
    if (hi) memcpy(hi, b->hi, sizeof(float) * 3 * b->n_nodes);
    if (left) memcpy(left, b->left, sizeof(int) * b->n_nodes);
    if (right) memcpy(right, b->right, sizeof(int) * b->n_nodes);
    if (first) memcpy(first, b->first, sizeof(int) * b->n_nodes);
    if (count) memcpy(count, b->count, sizeof(int) * b->n_nodes);
    if (order) memcpy(order, b->order, sizeof(int) * b->n_prims);
    if (root) *root = b->root;
}

/* Child-box layout -> per-node boxes. */
typedef struct {
    const float* cb;
    const int* cr;
    or_bvh* out;
    int cap;
} cvt_ctx;

static int cvt_ref(cvt_ctx* c, int ref, const float* lo, const float* hi)
{
    build_ctx bc;
    bc.out = c->out;
    bc.cap = c->cap;
    int node = new_node(&bc);
    c->cap = bc.cap;
    memcpy(c->out->lo + 3 * node, lo, sizeof(float) * 3);
    memcpy(c->out->hi + 3 * node, hi, sizeof(float) * 3);
    if (ref < 0) {
        int enc = ~ref;
        c->out->first[node] = enc >> 4;
        c->out->count[node] = enc & 15;
    } else {
        const float* b = c->cb + 12 * ref;
        int l = cvt_ref(c, c->cr[2 * ref + 0], b + 0, b + 3);
        int r = cvt_ref(c, c->cr[2 * ref + 1], b + 6, b + 9);
        c->out->left[node] = l;
        c->out->right[node] = r;
    }
    return node;
}

int or_bvh_from_childbox(int n_inner, const float* child_boxes, const int* child_refs,
                         const float root_lo[3], const float root_hi[3], int root_ref,
                         const int* leaf_prims, int n_prims, or_bvh* out)
{
    (void)n_inner;
    memset(out, 0, sizeof(*out));
    out->n_prims = n_prims;
    out->root = -1;
    if (n_prims <= 0)
        return 0;
    cvt_ctx c;
    c.cb = child_boxes;
    c.cr = child_refs;
    c.out = out;
    c.cap = 0;
    out->root = cvt_ref(&c, root_ref, root_lo, root_hi);
    out->order = (int*)malloc(sizeof(int) * n_prims);
    memcpy(out->order, leaf_prims, sizeof(int) * n_prims);
    return 0;
}

/* ------------------------------------------------------------------ */
/* BVHAccel::Travel (src/Acceleration/bvh.cpp:150-212).  The node accept
 * test (bvh.cpp:167-169) is split in a tmin-independent part evaluated
 * when the children of an entered node are examined and the tmin part
 * evaluated when the child is actually visited (immediately for the child
 * descended into, at pop time for the stacked one).  That evaluates the
 * same predicates at the same moments as the reference.               */

typedef struct {
    int node;
    float tnear;
} stk_t;

static inline int static_accept(const or_bvh* b, int node, const or_ray* r, float* tn)
{
    float tf;
    if (!or_bbox_intersect(b->lo + 3 * node, b->hi + 3 * node, r, tn, &tf))
        return 0;
    return tf >= 0.0f && r->tmax >= *tn;
}

static inline int tmin_accept(const or_ray* r, float tn)
{
    return r->tmin == 0.0f || r->tmin > tn;
}

/* Near-first order with postponed leaves (the GPU's order for SAH trees).
 * One node to visit (cur) and one leaf under test advance together per
 * step: the leaf's next primitive is tested first, then cur's children are
 * tested against the updated tmin (nearer accepted child next, farther one
 * stacked); a leaf reached while the leaf slot is free takes the slot and
 * the next stack entry is popped, otherwise it waits in cur.  Restates the
 * product's traversal order so hits and work counts compare exactly; the
 * closest hit distance equals the reference order's (parity pinned against
 * brute force). */
static int travel_near(const or_bvh* b, const or_prim* prims, or_ray* ray, int any,
                       int in_object, or_stats* st)
{
    stk_t stack[256];
    int sp = 0, hit = -1;
    int cur = b->root, leaf = -1, lpos = 0, lend = 0;
    for (;;) {
        int at_node = cur >= 0 && b->left[cur] >= 0;
        if (leaf >= 0) {
            int p = b->order[lpos];
            if (st) st->prims++;
            if (or_prim_intersect(&prims[p], ray, in_object))
                hit = p;
            if (++lpos == lend) {
                leaf = -1;
                if (any && ray->tmin > 0.0f)
                    return hit;
            }
        }
        if (at_node) {
            if (st) st->nodes++;
            int l = b->left[cur], r = b->right[cur];
            float tl, tr;
            int al = static_accept(b, l, ray, &tl) && tmin_accept(ray, tl);
            int ar = static_accept(b, r, ray, &tr) && tmin_accept(ray, tr);
            if (al && ar) {
                int swap = tr < tl;
                stack[sp].node = swap ? l : r;
                stack[sp].tnear = swap ? tl : tr;
                ++sp;
                cur = swap ? r : l;
            } else {
                cur = al ? l : (ar ? r : -1);
            }
        }
        for (;;) {
            if (cur >= 0 && b->left[cur] < 0) {
                if (b->count[cur] == 0) {
                    cur = -1;
                } else if (leaf >= 0) {
                    break;
                } else {
                    leaf = cur;
                    lpos = b->first[cur];
                    lend = lpos + b->count[cur];
                    cur = -1;
                }
            }
            if (cur >= 0 || sp == 0)
                break;
            --sp;
            if (tmin_accept(ray, stack[sp].tnear))
                cur = stack[sp].node;
        }
        if (cur < 0 && leaf < 0)
            return hit;
    }
}

/* The 4-wide form of an interior node, as the GPU builds it (widen,
 * rumi-renderer_amd/csrc/bvh_build.cpp): its two children, the interior child
 * of largest surface area (first on ties) repeatedly replaced in place by its
 * own two children, until four. */
static int wide_children(const or_bvh* b, int node, int out[4])
{
    int n = 2, k;
    out[0] = b->left[node];
    out[1] = b->right[node];
    while (n < 4) {
        int best = -1;
        double best_a = -1.0;
        for (k = 0; k < n; ++k) {
            if (b->left[out[k]] < 0) continue;
            const float* lo = b->lo + 3 * out[k];
            const float* hi = b->hi + 3 * out[k];
            double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
            double a = 2.0 * (dx * dy + dy * dz + dz * dx);
            if (a > best_a) {
                best_a = a;
                best = k;
            }
        }
        if (best < 0) break;
        int c = out[best];
        for (k = n; k > best + 1; --k) out[k] = out[k - 1];
        out[best] = b->left[c];
        out[best + 1] = b->right[c];
        ++n;
    }
    return n;
}

/* Order key of an accepted child of a 4-wide node: the entry distance's
 * bits mapped to an unsigned integer of the same order (either zero as +0),
 * the two low bits replaced by the child's slot.  key_distance reads back
 * the distance with those two bits cleared, never larger than the entry
 * distance; it is what the stack keeps. */
#define NO_KEY 0xffffffffu

static uint32_t order_key(float tn, int slot)
{
    float z = tn + 0.0f;
    int32_t bits;
    uint32_t u;
    memcpy(&bits, &z, 4);
    u = bits < 0 ? ~(uint32_t)bits : (uint32_t)bits | 0x80000000u;
    return (u & ~3u) | (uint32_t)slot;
}

static float key_distance(uint32_t key)
{
    uint32_t u = key & ~3u;
    uint32_t bits = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

/* travel_near over the 4-wide form: the accepted children are ordered by
 * order_key, the first is visited next and the others stacked last first
 * with their key_distance (trav_step_near4 in csrc/traverse.h). */
static int travel_near4(const or_bvh* b, const or_prim* prims, or_ray* ray, int any,
                        int in_object, or_stats* st)
{
    stk_t stack[256];
    int sp = 0, hit = -1;
    int cur = b->root, leaf = -1, lpos = 0, lend = 0;
    for (;;) {
        int at_node = cur >= 0 && b->left[cur] >= 0;
        if (leaf >= 0) {
            /* the leaf's next four primitives (as many as remain), one
             * after the other: the GPU's mesh kernels test up to four per
             * step */
            int k, use = lend - lpos < 4 ? lend - lpos : 4;
            for (k = 0; k < use; ++k) {
                int p = b->order[lpos++];
                if (st) st->prims++;
                if (or_prim_intersect(&prims[p], ray, in_object))
                    hit = p;
            }
            if (lpos == lend) {
                leaf = -1;
                if (any && ray->tmin > 0.0f)
                    return hit;
            }
        }
        if (at_node) {
            int ch[4], i, j, n;
            uint32_t key[4];
            if (st) st->nodes++;
            n = wide_children(b, cur, ch);
            for (i = 0; i < 4; ++i) {
                float tn = 0.0f;
                key[i] = NO_KEY;
                if (i < n && static_accept(b, ch[i], ray, &tn) && tmin_accept(ray, tn))
                    key[i] = order_key(tn, i);
            }
            /* ascending keys (distinct when accepted), rejected last */
            for (i = 1; i < 4; ++i)
                for (j = i; j > 0 && key[j] < key[j - 1]; --j) {
                    uint32_t tk = key[j]; int tc = ch[j];
                    key[j] = key[j - 1]; ch[j] = ch[j - 1];
                    key[j - 1] = tk; ch[j - 1] = tc;
                }
            cur = key[0] != NO_KEY ? ch[0] : -1;
            for (i = 3; i >= 1; --i)
                if (key[i] != NO_KEY) {
                    stack[sp].node = ch[i];
                    stack[sp].tnear = key_distance(key[i]);
                    ++sp;
                }
        }
        for (;;) {
            if (cur >= 0 && b->left[cur] < 0) {
                if (b->count[cur] == 0) {
                    cur = -1;
                } else if (leaf >= 0) {
                    break;
                } else {
                    leaf = cur;
                    lpos = b->first[cur];
                    lend = lpos + b->count[cur];
                    cur = -1;
                }
            }
            if (cur >= 0 || sp == 0)
                break;
            --sp;
            if (tmin_accept(ray, stack[sp].tnear))
                cur = stack[sp].node;
        }
        if (cur < 0 && leaf < 0)
            return hit;
    }
}

int or_travel(const or_bvh* b, const or_prim* prims, or_ray* ray, int any,
              int ordered, int in_object, or_stats* st)
{
    int hit = -1;
    if (b->root < 0)
        return -1;
    if (ray->d[0] == 0.0f && ray->d[1] == 0.0f && ray->d[2] == 0.0f)
        return -1;
    stk_t stack[256];
    int sp = 0;
    float tn;
    if (!(static_accept(b, b->root, ray, &tn) && tmin_accept(ray, tn)))
        return -1;
    if (ordered == 2)
        return travel_near4(b, prims, ray, any, in_object, st);
    if (ordered)
        return travel_near(b, prims, ray, any, in_object, st);
    int cur = b->root;
    for (;;) {
        if (b->left[cur] < 0) {
            int i, e = b->first[cur] + b->count[cur];
            for (i = b->first[cur]; i < e; ++i) {
                int p = b->order[i];
                if (st) st->prims++;
                if (or_prim_intersect(&prims[p], ray, in_object))
                    hit = p;
            }
            if (any && ray->tmin > 0.0f)
                return hit;
        } else {
            if (st) st->nodes++;
            int l = b->left[cur], r = b->right[cur];
            float tl, tr;
            int hl = static_accept(b, l, ray, &tl);
            int hr = static_accept(b, r, ray, &tr);
            if (!ordered) {
                if (hr) {
                    stack[sp].node = r;
                    stack[sp].tnear = tr;
                    ++sp;
                }
                if (hl && tmin_accept(ray, tl)) {
                    cur = l;
                    continue;
                }
            } else {
                int al = hl && tmin_accept(ray, tl);
                int ar = hr && tmin_accept(ray, tr);
                if (al && ar) {
                    int near = l, far = r;
                    float tfar_ = tr;
                    if (tr < tl) {
                        near = r;
                        far = l;
                        tfar_ = tl;
                    }
                    stack[sp].node = far;
                    stack[sp].tnear = tfar_;
                    ++sp;
                    cur = near;
                    continue;
                }
                if (al) { cur = l; continue; }
                if (ar) { cur = r; continue; }
            }
        }
        int found = 0;
        while (sp > 0) {
            --sp;
            if (tmin_accept(ray, stack[sp].tnear)) {
                cur = stack[sp].node;
                found = 1;
                break;
            }
        }
        if (!found)
            break;
    }
    return hit;
}

/* ------------------------------------------------------------------ */
/* Mesh helpers and batch drivers. */

int or_mesh_prims(const float* xyz, const uint32_t* idx, int nt, or_prim* out)
{
    int i, k;
    memset(out, 0, sizeof(or_prim) * (size_t)nt);
    for (i = 0; i < nt; ++i) {
        out[i].kind = OR_TRI;
        for (k = 0; k < 3; ++k) {
            out[i].a[k] = xyz[3 * idx[3 * i + 0] + k];
            out[i].b[k] = xyz[3 * idx[3 * i + 1] + k];
            out[i].c[k] = xyz[3 * idx[3 * i + 2] + k];
        }
    }
    return 0;
}

int or_mesh_bounds(const float* xyz, const uint32_t* idx, int nt, float* lo, float* hi)
{
    int i;
    for (i = 0; i < nt; ++i) {
        or_prim p;
        int k;
        p.kind = OR_TRI;
        for (k = 0; k < 3; ++k) {
            p.a[k] = xyz[3 * idx[3 * i + 0] + k];
            p.b[k] = xyz[3 * idx[3 * i + 1] + k];
            p.c[k] = xyz[3 * idx[3 * i + 2] + k];
        }
        or_prim_bound(&p, lo + 3 * i, hi + 3 * i);
    }
    return 0;
}

static void load_ray(const float* r, or_ray* ray)
{
    ray->o[0] = r[0];
    ray->o[1] = r[1];
    ray->o[2] = r[2];
    ray->tmin = r[3];
    ray->d[0] = r[4];
    ray->d[1] = r[5];
    ray->d[2] = r[6];
    ray->tmax = r[7];
}

int or_trace_mesh(const or_bvh* bvh, const float* xyz, const uint32_t* idx, int nt,
                  const float* rays, int n, int any, int ordered,
                  int32_t* prim_out, float* t_out, int32_t* node_cnt, int32_t* prim_cnt)
{
    or_prim* prims = (or_prim*)malloc(sizeof(or_prim) * (size_t)(nt > 0 ? nt : 1));
    or_mesh_prims(xyz, idx, nt, prims);
    int i;
    for (i = 0; i < n; ++i) {
        or_ray ray;
        load_ray(rays + 8 * (size_t)i, &ray);
        float t0 = ray.tmin;
        or_stats st = {0, 0};
        int h = or_travel(bvh, prims, &ray, any, ordered, 0, &st);
        prim_out[i] = h;
        if (t_out) t_out[i] = h >= 0 ? ray.tmin : t0;
        if (node_cnt) node_cnt[i] = (int32_t)st.nodes;
        if (prim_cnt) prim_cnt[i] = (int32_t)st.prims;
    }
    free(prims);
    return 0;
}

int or_brute_mesh(const float* xyz, const uint32_t* idx, int nt, const float* rays, int n,
                  int32_t* prim_out, float* t_out)
{
    int i, j, k;
    for (i = 0; i < n; ++i) {
        or_ray ray;
        load_ray(rays + 8 * (size_t)i, &ray);
        float t0 = ray.tmin;
        int h = -1;
        for (j = 0; j < nt; ++j) {
            float a[3], b[3], c[3];
            for (k = 0; k < 3; ++k) {
                a[k] = xyz[3 * idx[3 * j + 0] + k];
                b[k] = xyz[3 * idx[3 * j + 1] + k];
                c[k] = xyz[3 * idx[3 * j + 2] + k];
            }
            if (or_triangle_intersect(a, b, c, &ray, 0, 0, 0))
                h = j;
        }
        prim_out[i] = h;
        if (t_out) t_out[i] = h >= 0 ? ray.tmin : t0;
    }
    return 0;
}

static double now_seconds(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* One thread's share of or_time_median6_mesh's ray loop: rays [i0, i1). */
typedef struct {
    const or_bvh* b;
    const or_prim* prims;
    const float* rays;
    int i0, i1, any;
    int32_t* prim_out;
    float* t_out;
} travel_job;

static void* travel_rays(void* arg)
{
    const travel_job* j = (const travel_job*)arg;
    int i;
    for (i = j->i0; i < j->i1; ++i) {
        or_ray ray;
        load_ray(j->rays + 8 * (size_t)i, &ray);
        float tin = ray.tmin;
        int h = or_travel(j->b, j->prims, &ray, j->any, 0, 0, 0);
        if (j->prim_out) j->prim_out[i] = h;
        if (j->t_out) j->t_out[i] = h >= 0 ? ray.tmin : tin;
    }
    return 0;
}

double or_time_median6_mesh(const float* xyz, const uint32_t* idx, int nt,
                            const float* rays, int n, int any,
                            int32_t* prim_out, float* t_out, double* build_seconds)
{
    return or_time_median6_mesh_mt(xyz, idx, nt, rays, n, any, prim_out, t_out, build_seconds, 1);
}

double or_time_median6_mesh_mt(const float* xyz, const uint32_t* idx, int nt,
                               const float* rays, int n, int any,
                               int32_t* prim_out, float* t_out, double* build_seconds,
                               int nthreads)
{
    float* lo = (float*)malloc(sizeof(float) * 3 * (size_t)nt);
    float* hi = (float*)malloc(sizeof(float) * 3 * (size_t)nt);
    or_prim* prims = (or_prim*)malloc(sizeof(or_prim) * (size_t)nt);
    if (!lo || !hi || !prims)
        return -1.0;
    or_mesh_bounds(xyz, idx, nt, lo, hi);
    or_mesh_prims(xyz, idx, nt, prims);
    double t0 = now_seconds();
    or_bvh b;
    or_bvh_build_median6(lo, hi, nt, &b);
    double t1 = now_seconds();
    if (build_seconds) *build_seconds = t1 - t0;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    travel_job jobs[256];
    pthread_t th[256];
    int t, started = 0;
    double s0 = now_seconds();
    for (t = 0; t < nthreads; ++t) {
        jobs[t].b = &b;
        jobs[t].prims = prims;
        jobs[t].rays = rays;
        jobs[t].i0 = (int)((long long)n * t / nthreads);
        jobs[t].i1 = (int)((long long)n * (t + 1) / nthreads);
        jobs[t].any = any;
        jobs[t].prim_out = prim_out;
        jobs[t].t_out = t_out;
    }
    if (nthreads == 1) {
        travel_rays(&jobs[0]);
    } else {
        for (t = 0; t < nthreads; ++t)
            if (pthread_create(&th[t], 0, travel_rays, &jobs[t]) == 0) started = t + 1;
            else break;
        for (t = 0; t < started; ++t) pthread_join(th[t], 0);
        for (t = started; t < nthreads; ++t) travel_rays(&jobs[t]);  /* threads that did not start */
    }
    double s1 = now_seconds();
    or_bvh_free(&b);
    free(lo);
    free(hi);
    free(prims);
    return s1 - s0;
}
