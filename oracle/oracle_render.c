/*
 * oracle_render.c -- FP64 CPU restatement of Rumi-Renderer's path tracer,
 * materials, area lights, homogeneous medium and subsurface walk.
 * TEST INFRASTRUCTURE ONLY (the checker; never linked into the product).
 *
 * The per-sample algorithm follows PathTracer::Trace(w, h, ...)
 * (src/Integrator/pathtracer.cpp:117-354) literally: a forward loop that
 * records one Path per vertex, then a backward loop that evaluates the
 * direct lighting and accumulates L = direct + vol * fr * L * cos / pdf.
 * All arithmetic is double precision.  Randomness: one PCG32 stream per
 * sample (path tracer) or per pixel (bidirectional tracer), uniform doubles
 * in [0, 1) (the reference draws rand()/RAND_MAX in [0, 1]); pixel jitter
 * and light points are drawn continuously instead of from the reference's
 * fixed stratified point sets.
 *
 * Extensions beyond the reference (flags in the medium record): a
 * sphere-bounded medium and a Henyey-Greenstein phase function; with the
 * box bound and g = 0 the reference's code paths are taken.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define PI_D 3.14159265358979323846
#define INV_PI_D 0.3183098865475128
#define EPS_D 1e-4

typedef struct {
    double x, y, z;
} v3;

static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 u, v3 v)
{
    return mk(u.y * v.z - v.y * u.z, u.z * v.x - v.z * u.x, u.x * v.y - v.x * u.y);
}
static inline double len(v3 a) { return sqrt(dot(a, a)); }
static inline v3 norm(v3 a) { double il = 1.0 / sqrt(dot(a, a)); return mul(a, il); }
static inline double vget(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

typedef struct {
    double r, g, b;
} col;
static inline col cmk(double r, double g, double b) { col c = {r, g, b}; return c; }
static inline col cs(double s) { return cmk(s, s, s); }
static inline col cadd(col a, col b) { return cmk(a.r + b.r, a.g + b.g, a.b + b.b); }
static inline col cmul(col a, col b) { return cmk(a.r * b.r, a.g * b.g, a.b * b.b); }
static inline col cscale(col a, double s) { return cmk(a.r * s, a.g * s, a.b * s); }
static inline col cdiv(col a, col b) { return cmk(a.r / b.r, a.g / b.g, a.b / b.b); }
static inline double cavg(col a) { return (a.r + a.g + a.b) / 3.0; }
static inline int cnan(col a) { return a.r != a.r || a.g != a.g || a.b != a.b; }
static inline int cis0(col a) { return a.r == 0.0 && a.g == 0.0 && a.b == 0.0; }

/* ------------------------------------------------------------------ RNG */
typedef struct {
    uint64_t state, inc;
} pcg_t;

static uint32_t pcg_next(pcg_t* g)
{
    uint64_t old = g->state;
    g->state = old * 6364136223846793005ULL + g->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

static void pcg_seed(pcg_t* g, uint64_t initstate, uint64_t initseq)
{
    g->state = 0;
    g->inc = (initseq << 1u) | 1u;
    pcg_next(g);
    g->state += initstate;
    pcg_next(g);
}

static inline double rnd(pcg_t* g) { return (double)pcg_next(g) * (1.0 / 4294967296.0); }

/* Planted mistakes (or_plant, oracle.h): 0 in every ordinary use.  A test
 * sets one to show that a pin fails when the restated step is wrong. */
static int g_plant = 0;
void or_plant(int id) { g_plant = id; }

/* ------------------------------------------------------------------ scene */
enum { MAT_MATTE = 0, MAT_GLOSSY = 1, MAT_SPECULAR = 2, MAT_EMISSIVE = 3,
       MAT_DIELECTRIC = 4, MAT_SUBSURFACE = 5, MAT_GGX = 6 };

typedef struct {
    int kind;
    double kd, ks, e, kd_pdf, phong_norm;
    col color;
    double eta_in, eta_out;
    col absorption;
    col sigma_s, sigma_a, sigma_t;
    double avg_sigmat, absorb;
    double alpha; /* GGX roughness */
    int opaque;
} mat_t;

typedef struct {
    int kind; /* OR_TRI / OR_SPHERE / OR_RECT */
    int mat;
    int light;
    v3 a, b, c;       /* tri v0 v1 v2; sphere centre a; rect p l w */
    double r;
    v3 n0, n1, n2;    /* tri vertex normals */
    v3 n;             /* rect normal */
    double length, width, inv_area;
    double lo[3], hi[3];
    int tex;          /* texture index or -1 */
    double uv[6];     /* tri vertex uv: u0 v0 u1 v1 u2 v2 */
} shape_t;

typedef struct {
    int kind; /* 0 none, 1 box (reference), 2 sphere */
    double lo[3], hi[3];
    v3 c;
    double r;
    col sigma_s, sigma_a, sigma_t;
    double avg_sigma_t, absorb, g;
} medium_t;

typedef struct {
    v3 p;
    double R[9];
    double tanfov, y_ratio;
    double focal, lens; /* Pinhole::DepthOfField; focal 0 = pinhole */
} camera_t;

typedef struct {
    int first, count, left, right;
    double lo[3], hi[3];
} bnode_t;

typedef struct {
    shape_t* shapes;
    int n_shapes;
    mat_t* mats;
    int n_mats;
    int* lights;
    int n_lights;
    medium_t med;
    camera_t cam;
    bnode_t* nodes;
    int n_nodes;
    int* order;
    const float* texels;    /* RGB texel pool */
    const int* tex_table;   /* per texture: width, height, first float of its texels, kind
                               (1: checker, its w, h, c1, c2 as 8 floats at that offset) */
    int n_tex;
    int whitted;            /* RayTracer: only specular and dielectric surfaces continue */
} scene_t;

/* Ray + hit record, the reference's Ray/Intersection pair. */
typedef struct {
    v3 o, d;
    double tmin, tmax;
} ray_t;

/* --------------------------------------------------------- intersectors */
static int hit_tri(const shape_t* s, ray_t* ray, double* u_out, double* v_out, double* det_out)
{
    v3 e1 = sub(s->b, s->a), e2 = sub(s->c, s->a);
    v3 p = cross(ray->d, e2);
    double det = dot(p, e1);
    if (-1e-8 < det && det < 1e-8) return 0;
    double inv = 1.0 / det;
    v3 tv = sub(ray->o, s->a);
    double u = dot(p, tv) * inv;
    if (u < 0.0 || u > 1.0) return 0;
    v3 q = cross(tv, e1);
    double v = dot(q, ray->d) * inv;
    if (v < 0.0 || u + v > 1.0) return 0;
    double t = dot(q, e2) * inv;
    if (t < EPS_D) return 0;
    if (ray->tmin > EPS_D && t > ray->tmin) return 0;
    if (ray->tmax - t < EPS_D) return 0;
    ray->tmin = t;
    *u_out = u;
    *v_out = v;
    *det_out = det;
    return 1;
}

static int hit_sphere(const shape_t* s, ray_t* ray, int in_object, int ibl)
{
    v3 oc = sub(ray->o, s->a);
    double B = 2.0 * dot(ray->d, oc);
    double C = dot(oc, oc) - s->r * s->r;
    double disc = B * B - 4.0 * C;
    if (disc < 0.0) return 0;
    double t0, t1;
    if (disc == 0.0) {
        t0 = t1 = -B * 0.5;
    } else {
        double sq = sqrt(disc);
        t0 = (-B - sq) * 0.5;
        t1 = (-B + sq) * 0.5;
    }
    if (t0 < 0.0 && t1 < 0.0) return 0;
    if (in_object) {
        double tmp = t0;
        t0 = t1;
        t1 = tmp;
    }
    /* image-based light: an emissive textured sphere seen from inside
       (Sphere::Intersect, sphere.cpp:85-92) swaps the roots again */
    if (ibl && sqrt(dot(oc, oc)) < s->r) {
        double tmp = t0;
        t0 = t1;
        t1 = tmp;
    }
    if (t0 < EPS_D) return 0;
    if (ray->tmax - t0 < EPS_D) return 0;
    if (t0 > ray->tmin && ray->tmin > 0.0) return 0;
    ray->tmin = t0;
    return 1;
}

static int hit_rect(const shape_t* q, const mat_t* m, ray_t* ray)
{
    double cost = dot(ray->d, q->n);
    if (cost == 0.0) return 0;
    double t = -dot(sub(ray->o, q->a), q->n) / cost;
    if (t < 0.0 || t < EPS_D) return 0;
    if (ray->tmax - t < EPS_D) return 0;
    if (m->kind != MAT_EMISSIVE && (0.0 < cost && cost < 1.0)) return 0;
    if (t > ray->tmin && ray->tmin > 0.0) return 0;
    v3 ip = add(ray->o, mul(ray->d, t));
    v3 onp = sub(ip, q->a);
    double il = dot(onp, q->b) * (q->inv_area * q->width);
    if (il < 0.0 || il > q->length) return 0;
    double iw = dot(onp, q->c) * (q->inv_area * q->length);
    if (iw < 0.0 || iw > q->width) return 0;
    ray->tmin = t;
    return 1;
}

typedef struct {
    int obj;
    double u, v, det;
} hitrec_t;

static int hit_prim(const scene_t* sc, int i, ray_t* ray, int in_object, hitrec_t* h)
{
    const shape_t* s = &sc->shapes[i];
    double u = 0, v = 0, det = 0;
    int ok = 0;
    if (s->kind == OR_TRI)
        ok = hit_tri(s, ray, &u, &v, &det);
    else if (s->kind == OR_SPHERE)
        ok = hit_sphere(s, ray, in_object, sc->mats[s->mat].kind == MAT_EMISSIVE && s->tex >= 0);
    else if (s->kind == OR_RECT)
        ok = hit_rect(s, &sc->mats[s->mat], ray);
    /* OR_POINT: a PointLight is never hit */
    if (ok && h) {
        h->obj = i;
        h->u = u;
        h->v = v;
        h->det = det;
    }
    return ok;
}

static int slab(const double* lo, const double* hi, const ray_t* r, double* tn, double* tf)
{
    double t0, t1, tmp;
    int k;
    *tn = -INFINITY;
    *tf = INFINITY;
    for (k = 0; k < 3; ++k) {
        double inv = 1.0 / vget(r->d, k);
        double o = vget(r->o, k);
        t0 = (lo[k] - o) * inv;
        t1 = (hi[k] - o) * inv;
        if (t0 > t1) { tmp = t0; t0 = t1; t1 = tmp; }
        if (k == 0) {
            *tn = t0;
            *tf = t1;
            continue;
        }
        if (*tn > t1 || t0 > *tf) return 0;
        if (*tn < t0) *tn = t0;
        if (*tf > t1) *tf = t1;
    }
    return 1;
}

/* Closest hit (any = 0) or shadow query (any = 1): BVHAccel::Travel. */
static int travel(const scene_t* sc, ray_t* ray, int any, int in_object, hitrec_t* h)
{
    int stack[128], sp = 0, found = -1;
    if (sc->n_nodes == 0) return -1;
    if (ray->d.x == 0.0 && ray->d.y == 0.0 && ray->d.z == 0.0) return -1;
    stack[sp++] = 0;
    while (sp > 0) {
        const bnode_t* nd = &sc->nodes[stack[--sp]];
        double tn, tf;
        if (!(slab(nd->lo, nd->hi, ray, &tn, &tf) && tf >= 0.0 && ray->tmax >= tn &&
              (ray->tmin == 0.0 || ray->tmin > tn)))
            continue;
        if (nd->left < 0) {
            int i;
            for (i = nd->first; i < nd->first + nd->count; ++i)
                if (hit_prim(sc, sc->order[i], ray, in_object, any ? 0 : h)) found = sc->order[i];
            if (any && ray->tmin > 0.0) return found;
        } else {
            stack[sp++] = nd->right;
            stack[sp++] = nd->left;
        }
    }
    return found;
}

static void build_bvh(scene_t* sc)
{
    int n = sc->n_shapes, i;
    sc->order = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    for (i = 0; i < n; ++i) sc->order[i] = i;
    sc->nodes = (bnode_t*)malloc(sizeof(bnode_t) * (2 * (size_t)n + 1));
    sc->n_nodes = 0;
    if (n == 0) return;
    /* explicit work stack: (node, first, count) */
    int* work = (int*)malloc(sizeof(int) * 3 * (2 * (size_t)n + 1));
    int wp = 0;
    sc->nodes[0].first = 0;
    sc->nodes[0].count = n;
    sc->n_nodes = 1;
    work[wp++] = 0;
    while (wp > 0) {
        int ni = work[--wp];
        bnode_t* nd = &sc->nodes[ni];
        int f = nd->first, c = nd->count, k;
        for (k = 0; k < 3; ++k) {
            nd->lo[k] = INFINITY;
            nd->hi[k] = -INFINITY;
        }
        double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (i = f; i < f + c; ++i) {
            const shape_t* s = &sc->shapes[sc->order[i]];
            for (k = 0; k < 3; ++k) {
                double m = 0.5 * (s->lo[k] + s->hi[k]);
                if (s->lo[k] < nd->lo[k]) nd->lo[k] = s->lo[k];
                if (s->hi[k] > nd->hi[k]) nd->hi[k] = s->hi[k];
                if (m < clo[k]) clo[k] = m;
                if (m > chi[k]) chi[k] = m;
            }
        }
        nd->left = nd->right = -1;
        if (c <= 4) continue;
        int ax = 0;
        for (k = 1; k < 3; ++k)
            if (chi[k] - clo[k] > chi[ax] - clo[ax]) ax = k;
        double mid = 0.5 * (clo[ax] + chi[ax]);
        int l = f, r = f + c - 1;
        while (l <= r) {
            const shape_t* s = &sc->shapes[sc->order[l]];
            if (0.5 * (s->lo[ax] + s->hi[ax]) < mid) {
                ++l;
            } else {
                int t = sc->order[l];
                sc->order[l] = sc->order[r];
                sc->order[r] = t;
                --r;
            }
        }
        int nl = l - f;
        if (nl == 0 || nl == c) nl = c / 2;
        int li = sc->n_nodes++, ri = sc->n_nodes++;
        nd = &sc->nodes[ni];
        nd->left = li;
        nd->right = ri;
        sc->nodes[li].first = f;
        sc->nodes[li].count = nl;
        sc->nodes[ri].first = f + nl;
        sc->nodes[ri].count = c - nl;
        work[wp++] = ri;
        work[wp++] = li;
    }
    free(work);
}

/* ------------------------------------------------------- shading helpers */
/* OrthonormalBasis (src/Kernel/vector.h:70-95), its +-Y special case (tangents X and
 * -Z) taken within 1e-6 as the GPU renderer does: interpolated normals of
 * axis-aligned surfaces land on or beside +-Y by rounding, which would
 * otherwise decide the frame's orientation (a choice no distribution sampled
 * in the frame depends on). */
static void basis(v3 v0, v3* v1, v3* v2)
{
    if (fabs(v0.x) <= 1e-6 && fabs(v0.z) <= 1e-6 && fabs(fabs(v0.y) - 1.0) <= 1e-6) {
        *v1 = mk(1, 0, 0);
        *v2 = mk(0, 0, -1);
    } else {
        v3 c = cross(v0, mk(0, 1, 0));
        *v1 = norm(c);
        if (v1->x != v1->x || v1->y != v1->y || v1->z != v1->z) *v1 = mk(1, 0, 0);
        *v2 = cross(v0, *v1);
    }
}

static v3 frame(v3 u, v3 v, v3 w, double su, double sv, double sw)
{
    return add(add(mul(u, su), mul(v, sv)), mul(w, sw));
}

/* Lambertian::BSDFSampling (src/Kernel/bsdf.cpp:43-69). */
static v3 lambert_sample(v3 n, pcg_t* g, double* cost, double* pdf)
{
    v3 u, w;
    basis(n, &u, &w);
    double phi = rnd(g) * 2.0 * PI_D;
    double theta = acos(1.0 - rnd(g) * 2.0) * 0.5;
    double st = sin(theta);
    v3 wi = frame(u, n, w, st * sin(phi), cos(theta), st * cos(phi));
    *cost = dot(wi, n);
    if (*cost <= 0.0) *cost = 1e-4;
    *pdf = INV_PI_D * *cost;
    if (g_plant == OR_PLANT_LAMBERT_PDF) *pdf *= 0.5;
    if (*pdf <= 0.0) *pdf = 1e-4;
    return wi;
}

/* Phong::F (bsdf.cpp:91-104), white lobe. */
static double phong_f(const mat_t* m, v3 n, v3 wi, v3 wo)
{
    double c = dot(wi, n);
    if (c <= 0.0) c = 1e-4;
    v3 r = add(neg(wi), mul(n, c * 2.0));
    double ca = pow(dot(wo, r), m->e);
    return ca >= 1e-4 ? m->phong_norm * ca : m->phong_norm * 1e-4;
}

/* Phong::BSDFSampling (bsdf.cpp:127-170): the returned cosine is 1. */
static v3 phong_sample(const mat_t* m, v3 wo, v3 n, pcg_t* g, double* cost, double* pdf)
{
    double won = dot(wo, n);
    if (won <= 0.0) won = 1e-4;
    v3 v = add(neg(wo), mul(n, won * 2.0));
    v3 u, w;
    basis(v, &u, &w);
    double phi = rnd(g) * 2.0 * PI_D;
    double theta = acos(pow(rnd(g), 1.0 / (1.0 + m->e)));
    double su = sin(theta) * sin(phi), sv = cos(theta), sw = sin(theta) * cos(phi);
    v3 wi = frame(u, v, w, su, sv, sw);
    *cost = dot(wi, n);
    if (*cost < 0.0) {
        *cost = -*cost;
        wi = frame(u, v, w, -su, sv, -sw);
    }
    if (*cost == 0.0) *cost = 1e-4;
    double ca = pow(dot(wi, v), m->e);
    *pdf = ca >= 1e-4 ? m->phong_norm * ca : m->phong_norm * 1e-4;
    if (*pdf <= 0.0) *pdf = 1e-4;
    *cost = 1.0;
    return wi;
}

/* FresnelReflection::Fresnel (bsdf.cpp:255-283). */
static double fresnel(v3 dir, v3 n, double eta_in, double eta_out)
{
    double eta_i = eta_out, eta_t = eta_in, ci = -dot(dir, n);
    if (g_plant == OR_PLANT_ETA_SWAP) eta_i = eta_in, eta_t = eta_out;
    if (ci < 0.0) {
        eta_i = eta_in;
        eta_t = eta_out;
        ci = -ci;
    }
    double inv_eta = 1.0 / (eta_t / eta_i);
    double ct = 1.0 - (1.0 - ci * ci) * (inv_eta * inv_eta);
    if (ct <= 0.0) return 1.0;
    ct = sqrt(ct);
    double rs = (eta_i * ci - eta_t * ct) / (eta_i * ci + eta_t * ct);
    double rp = (eta_i * ct - eta_t * ci) / (eta_i * ct + eta_t * ci);
    return (rs * rs + rp * rp) * 0.5;
}

/* FresnelReflection::BSDFSampling (bsdf.cpp:232-253). */
static v3 fresnel_reflect(v3 wo, v3 n, double* cost)
{
    *cost = dot(wo, n);
    double won = *cost;
    if (won < 1e-4) {
        if (won > -1e-4)
            won = 1e-4;
        else {
            n = neg(n);
            won = -won;
        }
    }
    return add(neg(wo), mul(n, 2.0 * won));
}

/* FresnelTransparency::BSDFSampling (bsdf.cpp:322-342). */
static v3 fresnel_transmit(const mat_t* m, v3 wo, v3 n, double* cost)
{
    double eta = m->eta_in / m->eta_out;
    double ci = *cost = dot(wo, n);
    if (ci < 0.0) {
        n = neg(n);
        eta = m->eta_out / m->eta_in;
        ci = -ci;
    }
    double inv = g_plant == OR_PLANT_SNELL ? eta : 1.0 / eta;
    double ct = sqrt(1.0 - (1.0 - ci * ci) * (inv * inv));
    return add(mul(neg(wo), inv), mul(n, ci * inv - ct));
}

/* GGX microfacet reflection (Walter et al. 2007), Schlick Fresnel with
 * F0 = colour; an extension with no reference counterpart. */
static double ggx_d(double a, double ch)
{
    double a2 = a * a, d = ch * ch * (a2 - 1.0) + 1.0;
    return a2 / (PI_D * d * d);
}

static double ggx_lambda(double a, double c)
{
    double t2 = (1.0 - c * c) / (c * c);
    return 0.5 * (-1.0 + sqrt(1.0 + a * a * t2));
}

static col ggx_f(const mat_t* m, v3 n, v3 wi, v3 wo)
{
    double ci = dot(n, wi), co = dot(n, wo);
    if (ci <= 0.0 || co <= 0.0) return cs(0.0);
    v3 h = norm(add(wi, wo));
    double D = ggx_d(m->alpha, dot(n, h));
    double G = 1.0 / (1.0 + ggx_lambda(m->alpha, ci) + ggx_lambda(m->alpha, co));
    double x = 1.0 - fmax(dot(wi, h), 0.0), x5 = x * x * x * x * x;
    col F = cmk(m->color.r + (1 - m->color.r) * x5, m->color.g + (1 - m->color.g) * x5,
                m->color.b + (1 - m->color.b) * x5);
    return cscale(F, D * G / (4.0 * ci * co));
}

static double ggx_pdf(const mat_t* m, v3 n, v3 wi, v3 wo)
{
    double ci = dot(n, wi), co = dot(n, wo);
    if (ci <= 0.0 || co <= 0.0) return 0.0;
    v3 h = norm(add(wi, wo));
    double ch = dot(n, h);
    if (g_plant == OR_PLANT_GGX_JACOBIAN) return ggx_d(m->alpha, ch) * ch;
    return ggx_d(m->alpha, ch) * ch / (4.0 * dot(wo, h));
}

static v3 ggx_sample(const mat_t* m, v3 n, v3 wo, pcg_t* g)
{
    double u1 = rnd(g), u2 = rnd(g);
    double t2 = m->alpha * m->alpha * u1 / (1.0 - u1);
    double ch = 1.0 / sqrt(1.0 + t2), sh = sqrt(fmax(0.0, 1.0 - ch * ch));
    double phi = 2.0 * PI_D * u2;
    v3 u, w;
    basis(n, &u, &w);
    v3 h = frame(u, n, w, sh * sin(phi), ch, sh * cos(phi));
    return sub(mul(h, 2.0 * dot(wo, h)), wo);
}

/* ------------------------------------------------------------ medium */
static int med_intersect(const medium_t* md, const ray_t* r, double* tn, double* tf)
{
    *tn = 0.0;
    *tf = 0.0;
    if (md->kind == 1) {
        if (!slab(md->lo, md->hi, r, tn, tf)) return 0;
    } else {
        v3 oc = sub(r->o, md->c);
        double B = dot(r->d, oc), C = dot(oc, oc) - md->r * md->r;
        double A = dot(r->d, r->d);
        double disc = B * B - A * C;
        if (disc < 0.0) return 0;
        double sq = sqrt(disc);
        *tn = (-B - sq) / A;
        *tf = (-B + sq) / A;
    }
    if (*tn != *tn || *tf != *tf) return 0;
    if (0.0 < r->tmin && r->tmin < *tn) return 0;
    if (*tf < 0.0) return 0;
    return 1;
}

static double med_maxdist(double thit, double tn, double tf)
{
    if (tn < 0.0) {
        if (thit == 0.0) return tf;
        return thit < tf ? thit : tf;
    }
    if (thit == 0.0) return tf - tn;
    return thit < tf ? thit - tn : tf - tn;
}

static col exp_tau(col st, double d)
{
    return cmk(exp(st.r * -d), exp(st.g * -d), exp(st.b * -d));
}

/* HomogeneousVolume::DistanceSampling (homogeneousvolume.cpp:168-215) and
 * Subsurface::DistanceSampling (subsurface.cpp:49-99): returns 1 when a
 * scattering distance s was sampled; *absorbed when the roulette killed it. */
static int distance_sample(col sigma_t, double avg, double absorb, double smax, pcg_t* g,
                           double* s, int* absorbed, col* tau, col* pdf)
{
    double xi = rnd(g);
    if (xi >= 1.0 - exp(avg * -smax)) {
        *tau = exp_tau(sigma_t, smax);
        if (cis0(*tau)) *tau = cs(1e-4);
        *pdf = *tau;
        return 0;
    }
    double ds = -log(1.0 - xi) / avg;
    if (smax - ds < 1e-4) {
        *tau = exp_tau(sigma_t, smax);
        if (cis0(*tau)) *tau = cs(1e-4);
        *pdf = *tau;
        return 0;
    }
    if ((g_plant == OR_PLANT_NO_ROULETTE ? rnd(g) < 0.0 : rnd(g) < absorb) || ds < 1e-4) {
        *s = -1.0;
        *absorbed = 1;
        return 0;
    }
    *s = ds;
    *tau = exp_tau(sigma_t, ds);
    *pdf = g_plant == OR_PLANT_GREY_WEIGHT ? cscale(*tau, avg) : cmul(*tau, sigma_t);
    if (cis0(*pdf)) *pdf = cs(1e-4);
    return 1;
}

/* Henyey-Greenstein value for the cosine between propagation d and wi. */
static double hg_eval(double g, double cos_dw)
{
    double den = 1.0 + g * g - 2.0 * g * cos_dw;
    return (1.0 - g * g) / (4.0 * PI_D * den * sqrt(den));
}

/* Isotropic::IsotropicSampling (phase.cpp:40-56) for g = 0; HG otherwise. */
static v3 phase_sample(double g, v3 d, pcg_t* gen, double* pdf)
{
    v3 v = d, u, w;
    basis(v, &u, &w);
    double phi = rnd(gen) * 2.0 * PI_D;
    double ct;
    if (g == 0.0) {
        ct = cos(acos(rnd(gen) * 2.0 - 1.0));
        *pdf = 1.0 / (4.0 * PI_D);
    } else {
        double x = rnd(gen);
        double sq = (1.0 - g * g) / (1.0 - g + 2.0 * g * x);
        ct = (1.0 + g * g - sq * sq) / (2.0 * g);
        if (g_plant == OR_PLANT_HG_SIGN) ct = -ct;
        if (ct > 1.0) ct = 1.0;
        if (ct < -1.0) ct = -1.0;
        *pdf = hg_eval(g, ct);
    }
    double st = sqrt(1.0 - ct * ct);
    return frame(u, v, w, st * sin(phi), ct, st * cos(phi));
}

/* ------------------------------------------- the reference's point sets */
/* Q9: the reference draws pixel jitter and area-light points from fixed
 * point sets made once per frame, each draw taken by Sampler::Sampling
 * (sampler.cpp:17-38): an incremental shuffle of the x and of the y values,
 * independently, restarting after every pass over the set.  ref_set_t is
 * one such set and its cursor. */
typedef struct {
    int n, cursor;
    double *x, *y;
} ref_set_t;

static void ref_set_alloc(ref_set_t* s, int n)
{
    s->n = n;
    s->cursor = 0;
    s->x = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    s->y = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
}

static void ref_set_free(ref_set_t* s)
{
    free(s->x);
    free(s->y);
    s->x = s->y = 0;
    s->n = 0;
}

static void ref_set_copy(ref_set_t* d, const ref_set_t* s)
{
    memcpy(d->x, s->x, sizeof(double) * (size_t)s->n);
    memcpy(d->y, s->y, sizeof(double) * (size_t)s->n);
    d->n = s->n;
    d->cursor = 0;
}

/* Sampler::Sampling (sampler.cpp:17-38); rand() % m becomes a uniform
 * integer below m. */
static void ref_set_draw(ref_set_t* s, pcg_t* g, double* x, double* y)
{
    if (s->cursor >= s->n) s->cursor = 0;
    int m = s->n - s->cursor;
    int i1 = s->cursor + (int)(rnd(g) * m), i2 = s->cursor + (int)(rnd(g) * m);
    if (i1 >= s->n) i1 = s->n - 1;
    if (i2 >= s->n) i2 = s->n - 1;
    double sx = s->x[i1], sy = s->y[i2];
    s->x[i1] = s->x[s->cursor];
    s->y[i2] = s->y[s->cursor];
    s->x[s->cursor] = sx;
    s->y[s->cursor] = sy;
    ++s->cursor;
    *x = sx;
    *y = sy;
}

/* Sampler::ShuffleSamples (sampler.cpp:40-66). */
static void ref_set_shuffle(ref_set_t* s, pcg_t* g)
{
    int i;
    for (i = 0; i < s->n; ++i) {
        int j = (int)(rnd(g) * s->n);
        if (j >= s->n) j = s->n - 1;
        double t = s->x[i];
        s->x[i] = s->x[j];
        s->x[j] = t;
    }
    for (i = 0; i < s->n; ++i) {
        int j = (int)(rnd(g) * s->n);
        if (j >= s->n) j = s->n - 1;
        double t = s->y[i];
        s->y[i] = s->y[j];
        s->y[j] = t;
    }
}

/* UniformSample(n, jitter = n > 1) + GenerateSamples (uniformsample.cpp:
 * 18-58): d = (int)sqrt(n), d^2 points, one per cell of a d x d grid,
 * jittered in the cell or at its centre. */
static void ref_uniform_set(ref_set_t* s, int n, pcg_t* g)
{
    int d = (int)sqrt((float)n), ix, iy;
    ref_set_alloc(s, d * d);
    double inv = 1.0 / d;
    for (iy = 0; iy < d; ++iy)
        for (ix = 0; ix < d; ++ix) {
            double jx = n > 1 ? rnd(g) : 0.5, jy = n > 1 ? rnd(g) : 0.5;
            s->x[iy * d + ix] = (ix + jx) * inv;
            s->y[iy * d + ix] = (iy + jy) * inv;
        }
}

/* RectSample(n, length, width) + GenerateSamples (rectsample.cpp:18-102):
 * a jittered (int)sqrt(n) grid on the unit square, the rows' points with
 * y < ratio (short side / long side) kept until a row keeps none, mapped to
 * (x, y / ratio) -- or (y / ratio, x) when width > length. */
static void ref_rect_set(ref_set_t* s, int n, double length, double width, pcg_t* g)
{
    int d = (int)sqrt((float)n), ix, iy, kept = 0, k;
    double ratio = length >= width ? width / length : length / width;
    double* tx = (double*)malloc(sizeof(double) * (size_t)(d * d > 0 ? d * d : 1));
    double* ty = (double*)malloc(sizeof(double) * (size_t)(d * d > 0 ? d * d : 1));
    int rows = 0;
    for (iy = 0; iy < d; ++iy) {
        int any = 0;
        for (ix = 0; ix < d; ++ix) {
            tx[iy * d + ix] = (ix + rnd(g)) / d;
            ty[iy * d + ix] = (iy + rnd(g)) / d;
            if (ty[iy * d + ix] < ratio) {
                any = 1;
                ++kept;
            }
        }
        ++rows;
        if (!any) break;
    }
    ref_set_alloc(s, kept);
    for (k = 0, ix = 0; k < kept && ix < rows * d; ++ix)
        if (ty[ix] < ratio) {
            if (length >= width) {
                s->x[k] = tx[ix];
                s->y[k] = ty[ix] / ratio;
            } else {
                s->x[k] = ty[ix] / ratio;
                s->y[k] = tx[ix];
            }
            ++k;
        }
    free(tx);
    free(ty);
}

/* SphereSample(n) + GenerateSamples (spheresample.cpp:20-48): phi, theta
 * of a golden-ratio spiral (one random direction when n = 1). */
static void ref_sphere_set(ref_set_t* s, int n, pcg_t* g)
{
    int i;
    ref_set_alloc(s, n);
    if (n == 1) {
        s->x[0] = rnd(g) * 2.0 * PI_D;
        s->y[0] = acos(1.0 - rnd(g) * 2.0);
        return;
    }
    double golden = (1.0 + sqrt(5.0)) * 0.5;
    for (i = 0; i < n; ++i) {
        s->x[i] = PI_D * golden * (i + 0.5);
        s->y[i] = acos(1.0 - 2.0 * (i + 0.5) / n);
    }
}

/* The light sets' state of the pixel a thread is rendering (reference
 * sampling only): one set per scene shape, NULL where the shape has none. */
static __thread ref_set_t* tl_light_sets = 0;

/* ------------------------------------------------------------ lights */
typedef struct {
    v3 p, n;
    double pdf; /* (1 / n_lights) * (1 / area); image-based light: (1 / n_lights) cos / pi */
    col le;
    int ibl;    /* image-based light: the direction is wi, G = 1 */
    v3 wi;
    int tex;    /* image-based light: the texel le came from (pool index), else -1 */
} lsample_t;

static col pattern(const scene_t* sc, int obj, v3 p, double bu, double bv, int* index);

/* The stream of a vertex's light sample: a second PCG32 stream seeded from
 * the path stream's state on arrival at the vertex (recorded in the forward
 * loop), so the numbers the path itself draws do not depend on how many its
 * light samples take -- the same numbering as the GPU renderer, which takes
 * each light sample at its vertex rather than after the whole path. */
#define NEE_MIX 0xDA3E39CB94B95BDBull
static void nee_stream(pcg_t* n, const pcg_t* path, uint64_t key)
{
    n->state = 0;
    n->inc = path->inc ^ 2u;
    pcg_next(n);
    n->state += key ^ NEE_MIX;
    pcg_next(n);
}

static void light_sample(const scene_t* sc, v3 x, v3 nrm, pcg_t* g, lsample_t* ls)
{
    int li = (int)(rnd(g) * sc->n_lights);
    if (li >= sc->n_lights) li = sc->n_lights - 1;
    const shape_t* s = &sc->shapes[sc->lights[li]];
    const mat_t* m = &sc->mats[s->mat];
    double area;
    ls->ibl = 0;
    ls->tex = -1;
    if (s->kind == OR_POINT) {
        /* PointLight (pointlight.cpp:44-72): N points from the light toward
         * x, so the area-light G, max(-wi.n, 0) / d^2, is its 1 / d^2 */
        ls->p = s->a;
        ls->n = norm(sub(x, s->a));
        ls->pdf = 1.0 / sc->n_lights; /* Pdf() = 1 */
        ls->le = cmk(s->n0.x, s->n0.y, s->n0.z);
        return;
    }
    if (s->light == 2) {
        /* ImageBasedLight::UniformSampling (imagebasedlight.cpp:35-64) */
        v3 u, w;
        basis(nrm, &u, &w);
        double phi = rnd(g) * 2.0 * PI_D;
        double theta = acos(sqrt(1.0 - rnd(g)));
        double st = sin(theta), c;
        v3 wi = frame(u, nrm, w, st * sin(phi), cos(theta), st * cos(phi));
        ray_t r = {x, wi, 0.0, 1e300};
        int idx;
        hit_sphere(s, &r, 0, 1);
        ls->p = add(x, mul(wi, r.tmin));
        ls->n = neg(wi);
        idx = -1;
        ls->le = pattern(sc, sc->lights[li], ls->p, 0.0, 0.0, &idx);
        ls->tex = idx;
        c = dot(wi, nrm);
        if (c <= 0.0) c = 1e-4;
        ls->pdf = (1.0 / sc->n_lights) * (INV_PI_D * c > 0.0 ? INV_PI_D * c : 1e-4);
        ls->ibl = 1;
        ls->wi = wi;
        return;
    }
    ref_set_t* set = tl_light_sets ? &tl_light_sets[sc->lights[li]] : 0;
    if (set && set->n == 0) set = 0;
    if (s->kind == OR_RECT) {
        double a, b;
        if (set)
            ref_set_draw(set, g, &a, &b); /* Rect::ShapeSampling (rect.cpp:176-181) */
        else
            a = rnd(g), b = rnd(g);
        ls->p = add(add(s->a, mul(s->b, a)), mul(s->c, b));
        ls->n = s->n;
        area = s->length * s->width;
    } else if (set) {
        /* Sphere::ShapeSampling (sphere.cpp:152-170), identity transform */
        double phi, theta;
        ref_set_draw(set, g, &phi, &theta);
        v3 dir = mk(sin(theta) * sin(phi), cos(theta), sin(theta) * cos(phi));
        ls->p = add(s->a, mul(dir, s->r));
        ls->n = norm(sub(ls->p, s->a));
        area = 4.0 * PI_D * s->r * s->r;
    } else {
        double z = 1.0 - 2.0 * rnd(g), phi = 2.0 * PI_D * rnd(g);
        double r = sqrt(1.0 - z * z > 0 ? 1.0 - z * z : 0.0);
        v3 dir = mk(r * cos(phi), r * sin(phi), z);
        ls->p = add(s->a, mul(dir, s->r));
        ls->n = norm(sub(ls->p, s->a));
        area = 4.0 * PI_D * s->r * s->r;
    }
    ls->pdf = (1.0 / sc->n_lights) * (1.0 / area);
    ls->le = cscale(m->color, m->kd);
}

/* HomogeneousVolume::ShadowRayTr (homogeneousvolume.cpp:56-93): a grey
 * Bernoulli transmittance, 0 or 1. */
static double shadow_tr(const scene_t* sc, const ray_t* shadow, pcg_t* g, double* smax_out)
{
    const medium_t* md = &sc->med;
    double tn, tf;
    *smax_out = 0.0;
    if (md->kind == 0) return 1.0;
    if (!med_intersect(md, shadow, &tn, &tf)) return 1.0;
    double smax = med_maxdist(shadow->tmax, tn, tf);
    double xi = rnd(g);
    if (g_plant != OR_PLANT_SHADOW_TR && xi < 1.0 - exp(md->avg_sigma_t * -smax)) return 0.0;
    *smax_out = smax; /* the segment survived a Bernoulli trial of this length */
    return 1.0;
}

/* Scene::VisibilityTest (src/Scene/scene.cpp:99-123). */
static int visibility(const scene_t* sc, v3 x, v3 nrm, pcg_t* g, lsample_t* ls, v3* wi, ray_t* shadow)
{
    light_sample(sc, x, nrm, g, ls);
    *wi = ls->ibl ? ls->wi : norm(sub(ls->p, x));
    shadow->o = x;
    shadow->d = *wi;
    shadow->tmin = 0.0;
    shadow->tmax = len(sub(ls->p, x));
    if (dot(shadow->d, ls->n) >= 0.0) return 0;
    ray_t r = *shadow;
    travel(sc, &r, 1, 0, 0);
    if (r.tmin > 0.0) return 0;
    return 1;
}

/* AreaLight::G (arealight.cpp:106-112). */
static double geom(v3 x, const lsample_t* ls, v3 wi)
{
    if (ls->ibl) return 1.0; /* ImageBasedLight::G */
    double d = len(sub(x, ls->p));
    double c = -dot(wi, ls->n);
    if (c < 0.0) c = 0.0;
    return c / (d * d);
}

/* Matte::DirectLighting (matte.cpp:54-82).  *smax_s: medium length of the
 * shadow segment when its Bernoulli transmittance survived (else 0); *ltex:
 * the image-based light's texel the estimate is linear in (else -1). */
static col matte_direct(const scene_t* sc, const mat_t* m, v3 p, v3 n, pcg_t* g, double* smax_s,
                        int* ltex)
{
    lsample_t ls;
    v3 wi;
    ray_t sh;
    *smax_s = 0.0;
    *ltex = -1;
    if (!visibility(sc, p, n, g, &ls, &wi, &sh)) return cs(0.0);
    *ltex = ls.tex;
    col f = cscale(m->color, INV_PI_D * m->kd);
    double c = dot(wi, n);
    if (c < 0.0) c = 0.0;
    col lo = cscale(cmul(f, ls.le), c * geom(p, &ls, wi) / ls.pdf);
    if (sc->med.kind) lo = cscale(lo, shadow_tr(sc, &sh, g, smax_s));
    return lo;
}

/* HomogeneousVolume::DirectLighting (homogeneousvolume.cpp:231-250), with
 * the HG value in place of 1/(4 pi) when g != 0.  Returns the estimate
 * without its sigma_s factor; *cos_l is the cosine between d and the light
 * direction, *smax_s as for matte_direct. */
static col medium_direct0(const scene_t* sc, v3 p, v3 d, pcg_t* g, double* smax_s, double* cos_l)
{
    lsample_t ls;
    v3 wi;
    ray_t sh;
    *smax_s = 0.0;
    *cos_l = 0.0;
    if (!visibility(sc, p, mk(0, 1, 0), g, &ls, &wi, &sh)) return cs(0.0); /* no image-based light with media */
    double tau = shadow_tr(sc, &sh, g, smax_s);
    *cos_l = dot(d, wi);
    double fp = sc->med.g == 0.0 ? 1.0 / (4.0 * PI_D) : hg_eval(sc->med.g, *cos_l);
    return cscale(ls.le, tau * fp * geom(p, &ls, wi) / ls.pdf);
}

/* d log HG(c; g) / dg; at g = 0 (the isotropic phase) it is 3c. */
static double dlog_hg(double g, double c)
{
    return -2.0 * g / (1.0 - g * g) - 3.0 * (g - c) / (1.0 + g * g - 2.0 * g * c);
}

/* ------------------------------------------------------------ textures */
/* ImageTexture::Pixel (imagetexture.cpp:32-61), nearest texel; *index is
 * the texel's position in the pool. */
static col texel(const scene_t* sc, int t, double u, double v, int* index)
{
    int W = sc->tex_table[4 * t], H = sc->tex_table[4 * t + 1], off = sc->tex_table[4 * t + 2];
    if (sc->tex_table[4 * t + 3] == 2) {
        /* HDRTexture::Pixel (hdrtexture.cpp:29-41): rows as loaded, u mirrored */
        int y = (int)((1.0 - v) * (H - 1)), x = (int)((1.0 - u) * (W - 1));
        if (y < 0) y = 0;
        if (y > H - 1) y = H - 1;
        if (x < 0) x = 0;
        if (x > W - 1) x = W - 1;
        *index = off / 3 + y * W + x;
        const float* q = sc->texels + off + 3 * ((size_t)y * W + x);
        return cmk(q[0], q[1], q[2]);
    }
    if (sc->tex_table[4 * t + 3] == 1) {
        /* Checker::Pixel (checker.cpp:24-32), FP32 products as the reference's */
        const float* c = sc->texels + off;
        int a = (int)((float)u * c[0]) % 2, b = (int)((float)v * c[1]) % 2;
        const float* pick = ((a == 0 && b == 0) || (a != 0 && b != 0)) ? c + 5 : c + 2;
        *index = -1;
        return cmk(pick[0], pick[1], pick[2]);
    }
    int row = (int)((1.0 - v) * (H - 1));
    int cl = (int)(u * (W - 1));
    int y = H - 1 - row, x = cl;
    if (y < 0) y = 0;
    if (y > H - 1) y = H - 1;
    if (x < 0) x = 0;
    if (x > W - 1) x = W - 1;
    *index = off / 3 + y * W + x;
    const float* q = sc->texels + off + 3 * ((size_t)y * W + x);
    return cmk(q[0], q[1], q[2]);
}

/* Texture::Pattern (texture.cpp:13-24) with the shapes' TextureMapping:
 * Sphere (sphere.cpp:172-195, identity transform), Rect (rect.cpp:183-197);
 * triangles interpolate their vertex uv (the reference's Shape default maps
 * every point to (0, 0), which zero uv reproduces). */
static col pattern(const scene_t* sc, int obj, v3 p, double bu, double bv, int* index)
{
    const shape_t* s = &sc->shapes[obj];
    double u, v;
    if (s->kind == OR_TRI) {
        double w0 = 1.0 - bu - bv;
        u = w0 * s->uv[0] + bu * s->uv[2] + bv * s->uv[4];
        v = w0 * s->uv[1] + bu * s->uv[3] + bv * s->uv[5];
    } else if (s->kind == OR_SPHERE) {
        /* angles against the rows of the shape's transform (n0, n1, n2;
           all zero: identity) */
        v3 uu = s->n0, vv = s->n1, ww = s->n2;
        if (dot(uu, uu) == 0.0 && dot(vv, vv) == 0.0 && dot(ww, ww) == 0.0) {
            uu = mk(1, 0, 0);
            vv = mk(0, 1, 0);
            ww = mk(0, 0, 1);
        }
        v3 pc = norm(sub(p, s->a));
        double theta = acos(dot(pc, vv));
        v3 pxz = mul(pc, sin(theta));
        double phi = atan2(dot(pxz, uu), dot(pxz, ww));
        u = phi >= 0.0 ? phi / (2.0 * PI_D) : (phi + 2.0 * PI_D) / (2.0 * PI_D);
        v = 1.0 - theta / PI_D;
    } else {
        v3 on = sub(p, s->a);
        u = dot(on, s->b) / (s->length * s->length);
        v = dot(on, s->c) / (s->width * s->width);
        if (u >= 1.0) u = 1.0;
        if (u <= 0.0) u = 0.0;
        if (v >= 1.0) v = 1.0;
        if (v <= 0.0) v = 0.0;
    }
    return texel(sc, s->tex, u, v, index);
}

/* ------------------------------------------------------------ path tracer */
typedef struct {
    /* vertex */
    v3 p, n;
    double costheta_i, pdf_path, pdf_wi;
    int onsurface;
    /* ray */
    v3 o, d;
    double tmin;
    /* info */
    int obj;
    int le, in_object;
    double rand;
    double vs, vsmax;
    col vtau, vpdf;
    double bu, bv;   /* triangle barycentrics of the hit */
    uint64_t nee;    /* path-stream state on arrival: seeds the light sample */
} pv_t;

static void pv_clean(pv_t* v)
{
    memset(v, 0, sizeof(*v));
    v->costheta_i = 1.0;
    v->pdf_path = 1.0;
    v->pdf_wi = 1.0;
    v->onsurface = 1;
    v->obj = -1;
    v->rand = 1.0;
    v->vtau = cs(1.0);
    v->vpdf = cs(1.0);
}

/* Surface point and normal after a closest hit: Triangle/Sphere/Rect
 * Intersect's vertex writes (triangle.cpp:104-114, sphere.cpp:110-129,
 * rect.cpp:112-118). */
static void surface(const scene_t* sc, const hitrec_t* h, const ray_t* r, v3* p, v3* n)
{
    const shape_t* s = &sc->shapes[h->obj];
    const mat_t* m = &sc->mats[s->mat];
    *p = add(r->o, mul(r->d, r->tmin));
    if (s->kind == OR_TRI) {
        *n = add(add(mul(s->n0, 1.0 - h->u - h->v), mul(s->n1, h->u)), mul(s->n2, h->v));
        if (h->det < 0.0 && m->opaque) *n = neg(*n);
    } else if (s->kind == OR_SPHERE) {
        if (m->kind == MAT_EMISSIVE && !(len(sub(r->o, s->a)) > s->r))
            *n = norm(sub(s->a, *p));
        else
            *n = norm(sub(*p, s->a));
    } else {
        *n = s->n;
    }
}

/* Forward-mode derivative of one sample's radiance, carried through the
 * backward loop: dL[k] = dL/d theta_k per channel for the medium's
 * (sigma_a rgb, sigma_s rgb, g), and one entry per textured vertex with
 * dL/d texel (per channel).  Sampling decisions that depend on theta add
 * their score d log p / d theta times the radiance they lead to, so the
 * result is the derivative of the estimator's expectation. */
typedef struct {
    col dL[7];
    int n_tex;
    int* tix;   /* 2 max_depth entries */
    col* tval;
} grad_t;

/* PathTracer::Trace(w, h, camera, film) (pathtracer.cpp:117-354). */
static col trace_pt(const scene_t* sc, ray_t cam_ray, int max_depth, pcg_t* g, pv_t* P, grad_t* G)
{
    const medium_t* md = &sc->med;
    int i, di;
    for (i = 0; i < max_depth; ++i) pv_clean(&P[i]);
    pv_t* cur = &P[0];
    cur->o = cam_ray.o;
    cur->d = cam_ray.d;
    int le = 1, in_object = 0, absorb = 0;
    cur->le = le;
    cur->in_object = in_object;
    for (di = 0; di < max_depth; ++di) {
        cur = &P[di];
        cur->nee = g->state;
        ray_t r = {cur->o, cur->d, 0.0, 1e300};
        hitrec_t h = {-1, 0, 0, 0};
        int hit = travel(sc, &r, 0, cur->in_object, &h);
        cur->tmin = r.tmin;
        if (hit >= 0) {
            cur->obj = hit;
            cur->bu = h.u;
            cur->bv = h.v;
            surface(sc, &h, &r, &cur->p, &cur->n);
        }
        if (md->kind && !cur->in_object && !sc->whitted) {
            double tn, tf;
            if (med_intersect(md, &r, &tn, &tf)) {
                cur->vsmax = med_maxdist(r.tmin, tn, tf);
                if (distance_sample(md->sigma_t, md->avg_sigma_t, md->absorb, cur->vsmax, g, &cur->vs,
                                    &absorb, &cur->vtau, &cur->vpdf)) {
                    cur->p = add(r.o, mul(r.d, tn < 0.0 ? cur->vs : tn + cur->vs));
                    cur->onsurface = 0;
                }
            }
            if (absorb) break;
        }
        double xi = 1.0, cost = 1.0, pdf_wi = 1.0, pdf_path = 1.0;
        v3 w = mk(0, 0, 0);
        if (md->kind && cur->vs > 0.0 && !cur->in_object) {
            le = 0;
            in_object = 0;
            w = phase_sample(md->g, r.d, g, &pdf_wi);
        } else {
            if (cur->obj < 0) break;
            const mat_t* m = &sc->mats[sc->shapes[cur->obj].mat];
            if (m->kind == MAT_EMISSIVE) break;
            /* RayTracer::Trace (raytracer.cpp:37-78): every material but
               Specular and Dielectric gives its DirectLighting and ends the
               path; the lobe variable stays 1 (intersection.cpp:18) */
            if (sc->whitted && m->kind != MAT_SPECULAR && m->kind != MAT_DIELECTRIC) break;
            v3 wo = neg(r.d), n = cur->n;
            int sample_bsdf = 1;
            if (m->kind == MAT_SUBSURFACE && in_object) {
                /* Subsurface::PhaseSampling (subsurface.cpp:135-155) */
                int ab = 0;
                if (distance_sample(m->sigma_t, m->avg_sigmat, m->absorb, r.tmin, g, &cur->vs, &ab,
                                    &cur->vtau, &cur->vpdf)) {
                    le = 1;
                    in_object = 1;
                    cur->p = add(r.o, mul(r.d, cur->vs));
                    w = phase_sample(0.0, neg(wo), g, &pdf_wi);
                    sample_bsdf = 0;
                } else {
                    if (ab) cur->vs = -1.0;
                    if (cur->vs < 0.0) break;
                }
            }
            if (sample_bsdf) {
                switch (m->kind) {
                case MAT_MATTE:
                    le = 0;
                    in_object = 0;
                    xi = 1.0;
                    w = lambert_sample(n, g, &cost, &pdf_wi);
                    break;
                case MAT_GLOSSY:
                    xi = rnd(g);
                    if (m->kd != 0.0 && xi <= m->kd_pdf) {
                        w = lambert_sample(n, g, &cost, &pdf_wi);
                        le = 0;
                        pdf_path = m->kd_pdf;
                    } else {
                        w = phong_sample(m, wo, n, g, &cost, &pdf_wi);
                        le = 1;
                        pdf_path = 1.0 - m->kd_pdf;
                    }
                    in_object = 0;
                    break;
                case MAT_SPECULAR: {
                    le = 1;
                    in_object = 0;
                    xi = 1.0;
                    double won = dot(wo, n);
                    if (won <= 0.0) won = 1e-4;
                    w = add(neg(wo), mul(n, 2.0 * won));
                    cost = 1.0;
                    pdf_wi = 1.0;
                    break;
                }
                case MAT_GGX:
                    le = 1;
                    in_object = 0;
                    w = ggx_sample(m, n, wo, g);
                    cost = dot(n, w);
                    pdf_wi = ggx_pdf(m, n, w, wo);
                    if (!(cost > 0.0 && pdf_wi > 0.0)) cost = 0.0, pdf_wi = 1.0;
                    break;
                case MAT_DIELECTRIC:
                case MAT_SUBSURFACE: {
                    double kr = fresnel(neg(wo), n, m->eta_in, m->eta_out);
                    double kt = 1.0 - kr;
                    double c;
                    if (kr != 0.0 && rnd(g) <= kr) {
                        w = fresnel_reflect(wo, n, &c);
                        in_object = c < 0.0;
                        xi = 1.0;
                        pdf_path = kr;
                    } else {
                        c = 0.0;
                        w = fresnel_transmit(m, wo, n, &c);
                        in_object = c > 0.0;
                        xi = 0.0;
                        pdf_path = kt;
                    }
                    le = 1;
                    cost = 1.0;
                    pdf_wi = 1.0;
                    break;
                }
                }
            }
        }
        cur->rand = xi;
        cur->costheta_i = cost;
        cur->pdf_wi = pdf_wi;
        cur->pdf_path = pdf_path;
        if (di + 1 > max_depth - 1) break;
        pv_t* nx = &P[di + 1];
        nx->o = cur->p;
        nx->d = w;
        nx->le = le;
        nx->in_object = in_object;
    }
    if (di >= max_depth) di = max_depth - 1;

    col L = cs(0.0);
    int k, e;
    if (G) {
        for (k = 0; k < 7; ++k) G->dL[k] = cs(0.0);
        G->n_tex = 0;
    }
    for (; di >= 0; --di) {
        cur = &P[di];
        v3 n = cur->n, wo = neg(cur->d);
        v3 w = di == max_depth - 1 ? mk(0, 0, 0) : P[di + 1].d;
        col direct = cs(0.0), indirect, fr = cs(1.0), vol = cs(1.0);
        double cost = cur->costheta_i;
        double pdf = cur->pdf_wi * cur->pdf_path;
        if (md->kind && cur->vs != 0.0 && !cur->in_object) {
            if (cur->vs == -1.0) continue;
            if (cur->vs > 0.0) {
                double smax_s, cos_l;
                pcg_t gn;
                nee_stream(&gn, g, cur->nee);
                col dir0 = medium_direct0(sc, cur->p, cur->d, &gn, &smax_s, &cos_l);
                direct = cmul(dir0, md->sigma_s);
                double fp = md->g == 0.0 ? 1.0 / (4.0 * PI_D) : hg_eval(md->g, dot(cur->d, w));
                col L_in = L;
                indirect = cscale(cmul(L, md->sigma_s), fp / pdf);
                vol = cdiv(cur->vtau, cur->vpdf);
                L = cmul(cadd(direct, indirect), vol);
                if (G) {
                    /* scattering at distance ds after surviving the absorption
                       roulette: density avg e^(-avg ds) (1 - A), A = mean(sa/st);
                       weight vol = 1/sigma_t per channel (tau / pdf), times
                       sigma_s and the phase value (whose sampling pdf it equals,
                       so the direction's score is d log HG) */
                    const double avg = md->avg_sigma_t, A = md->absorb;
                    const double hg_w = dlog_hg(md->g, dot(cur->d, w)), hg_l = dlog_hg(md->g, cos_l);
                    for (k = 0; k < 7; ++k) {
                        col dd = cs(0.0), dn = cscale(cmul(G->dL[k], md->sigma_s), fp / pdf), dv = cs(0.0);
                        double score = 0.0;
                        if (k < 6) {
                            const int c = k % 3;
                            const double st = c == 0 ? md->sigma_t.r : (c == 1 ? md->sigma_t.g : md->sigma_t.b);
                            const double sa = c == 0 ? md->sigma_a.r : (c == 1 ? md->sigma_a.g : md->sigma_a.b);
                            const double ss = st - sa;
                            double* ddc = c == 0 ? &dd.r : (c == 1 ? &dd.g : &dd.b);
                            double* dnc = c == 0 ? &dn.r : (c == 1 ? &dn.g : &dn.b);
                            double* dvc = c == 0 ? &dv.r : (c == 1 ? &dv.g : &dv.b);
                            if (k >= 3) { /* sigma_s_c multiplies both terms of channel c */
                                *ddc += c == 0 ? dir0.r : (c == 1 ? dir0.g : dir0.b);
                                *dnc += (c == 0 ? L_in.r : (c == 1 ? L_in.g : L_in.b)) * fp / pdf;
                            }
                            *dvc = -(c == 0 ? vol.r : (c == 1 ? vol.g : vol.b)) / st;
                            const double dA = k < 3 ? ss / (3.0 * st * st) : -sa / (3.0 * st * st);
                            score = (1.0 / avg - cur->vs) / 3.0 - dA / (1.0 - A);
                            if (smax_s > 0.0) dd = cadd(dd, cscale(direct, -smax_s / 3.0));
                        } else {
                            dd = cscale(direct, hg_l);
                            dn = cadd(dn, cscale(indirect, hg_w));
                        }
                        G->dL[k] = cadd(cadd(cmul(cadd(dd, dn), vol), cmul(cadd(direct, indirect), dv)),
                                        cscale(L, score));
                    }
                    for (e = 0; e < G->n_tex; ++e)
                        G->tval[e] = cmul(cscale(cmul(G->tval[e], md->sigma_s), fp / pdf), vol);
                }
            }
        } else {
            if (cur->obj < 0) continue;
            const shape_t* s = &sc->shapes[cur->obj];
            const mat_t* m = &sc->mats[s->mat];
            int rinfo_in = cur->in_object;
            double smax_s = 0.0;
            int ltex = -1;
            pcg_t gn;
            nee_stream(&gn, g, cur->nee);
            switch (m->kind) {
            case MAT_MATTE:
                fr = cscale(m->color, INV_PI_D * m->kd);
                direct = matte_direct(sc, m, cur->p, n, &gn, &smax_s, &ltex);
                break;
            case MAT_GLOSSY:
                if (m->kd != 0.0 && cur->rand <= m->kd_pdf) {
                    fr = cscale(m->color, INV_PI_D * m->kd);
                    direct = cscale(matte_direct(sc, m, cur->p, n, &gn, &smax_s, &ltex), 1.0 / m->kd_pdf);
                } else {
                    fr = cs(phong_f(m, n, w, wo) * m->ks);
                }
                break;
            case MAT_SPECULAR:
                fr = cscale(m->color, m->kd);
                break;
            case MAT_GGX:
                fr = ggx_f(m, n, w, wo);
                break;
            case MAT_EMISSIVE:
                if (rinfo_in == 0 && dot(n, wo) > 0.0) fr = cscale(m->color, m->kd);
                else fr = cs(0.0);
                if (rinfo_in == 0 && cur->le && -dot(cur->d, n) > 0.0)
                    direct = cscale(m->color, m->kd);
                break;
            case MAT_DIELECTRIC:
            case MAT_SUBSURFACE:
                if (m->kind == MAT_SUBSURFACE && cur->vs != 0.0) {
                    vol = cdiv(cur->vtau, cur->vpdf);
                    vol = cscale(cmul(vol, m->sigma_s), 1.0 / (4.0 * PI_D));
                    break;
                }
                if (cur->rand == 1.0) {
                    fr = cs(cur->pdf_path);
                } else {
                    double eta = rinfo_in ? m->eta_in / m->eta_out : m->eta_out / m->eta_in;
                    if (g_plant == OR_PLANT_NO_ETA2) eta = 1.0;
                    fr = cs(1.0 / (eta * eta) * cur->pdf_path);
                }
                if (rinfo_in) {
                    if (m->kind == MAT_DIELECTRIC && g_plant != OR_PLANT_NO_BEER)
                        fr = cmul(fr, exp_tau(m->absorption, cur->tmin));
                    else
                        fr = cmul(fr, cdiv(cur->vtau, cur->vpdf));
                }
                break;
            }
            col k_ind = cscale(cmul(vol, fr), cost / pdf);
            indirect = cmul(k_ind, L);
            L = cadd(direct, indirect);
            col pass = cs(1.0);
            if (md->kind && !cur->in_object) pass = cdiv(cur->vtau, cur->vpdf);
            L = cmul(L, pass);
            /* the segment into this vertex crossed the medium without
               scattering: probability e^(-avg (smax - 1e-4)) */
            const double pass_score =
                (md->kind && !cur->in_object && cur->vsmax > 1e-4) ? -(cur->vsmax - 1e-4) / 3.0 : 0.0;
            col tx = cs(1.0);
            int tix = -1;
            if (s->tex >= 0) tx = pattern(sc, cur->obj, cur->p, cur->bu, cur->bv, &tix);
            if (G) {
                for (k = 0; k < 7; ++k) {
                    col d = cmul(G->dL[k], k_ind);
                    if (k < 6 && smax_s > 0.0) d = cadd(d, cscale(direct, -smax_s / 3.0));
                    d = cmul(d, pass);
                    if (k < 6) d = cadd(d, cscale(L, pass_score));
                    G->dL[k] = cmul(d, tx);
                }
                for (e = 0; e < G->n_tex; ++e) G->tval[e] = cmul(cmul(cmul(G->tval[e], k_ind), pass), tx);
                if (tix >= 0) {
                    G->tix[G->n_tex] = tix;
                    G->tval[G->n_tex] = L; /* d (L tx) / d tx per channel */
                    ++G->n_tex;
                }
                if (ltex >= 0) {
                    /* the light sample is linear in the image-based light's
                       texel: d direct / d texel = direct / texel per channel */
                    const float* tv = sc->texels + 3 * (size_t)ltex;
                    col d = cmk(tv[0] > 0 ? direct.r / tv[0] : 0.0, tv[1] > 0 ? direct.g / tv[1] : 0.0,
                                tv[2] > 0 ? direct.b / tv[2] : 0.0);
                    G->tix[G->n_tex] = ltex;
                    G->tval[G->n_tex] = cmul(cmul(d, pass), tx);
                    ++G->n_tex;
                }
            }
            L = cmul(L, tx);
        }
    }
    return L;
}

/* Pinhole::CastViewRay (src/Camera/pinhole.cpp:97-136); with g and a focal
 * distance > 0 the thin lens draws two more numbers after the pixel jitter. */
static ray_t camera_ray(const camera_t* c, int W, int H, int x, int y, double sx, double sy, pcg_t* g)
{
    double hw = W * 0.5, hh = H * 0.5;
    double px = -(((double)x - hw) + sx) / hw * c->tanfov;
    double py = -(((double)y - hh) + sy) / hh * c->tanfov * c->y_ratio;
    v3 l = norm(mk(px, py, 1.0));
    ray_t r;
    r.o = c->p;
    r.d = mk(c->R[0] * l.x + c->R[1] * l.y + c->R[2] * l.z,
             c->R[3] * l.x + c->R[4] * l.y + c->R[5] * l.z,
             c->R[6] * l.x + c->R[7] * l.y + c->R[8] * l.z);
    if (c->focal > 0.0 && g) {
        /* thin lens (Pinhole::CastViewRay, pinhole.cpp:111-122) */
        v3 fp = add(r.o, mul(r.d, c->focal / l.z));
        double rr = sqrt(rnd(g)) * c->lens;
        double th = rnd(g) * 2.0 * PI_D;
        v3 up = mk(c->R[1], c->R[4], c->R[7]), right = mk(c->R[0], c->R[3], c->R[6]);
        r.o = add(add(r.o, mul(mul(up, rr), sin(th))), mul(mul(right, rr), cos(th)));
        r.d = norm(sub(fp, r.o));
    }
    r.tmin = 0.0;
    r.tmax = 1e300;
    return r;
}

/* ------------------------------------------------------------ scene load */
static void load_scene(scene_t* sc, const float* shapes, int n_shapes, const float* mats,
                       int n_mats, const float* cam, const float* med)
{
    /* textures are attached by the caller after loading */
    int i, k;
    memset(sc, 0, sizeof(*sc));
    sc->n_shapes = n_shapes;
    sc->shapes = (shape_t*)calloc((size_t)(n_shapes > 0 ? n_shapes : 1), sizeof(shape_t));
    sc->lights = (int*)malloc(sizeof(int) * (n_shapes > 0 ? n_shapes : 1));
    sc->n_mats = n_mats;
    sc->mats = (mat_t*)calloc((size_t)(n_mats > 0 ? n_mats : 1), sizeof(mat_t));
    for (i = 0; i < n_mats; ++i) {
        const float* m = mats + 32 * (size_t)i;
        mat_t* o = &sc->mats[i];
        o->kind = (int)m[0];
        o->kd = m[1];
        o->ks = m[2];
        o->e = m[3];
        o->color = cmk(m[4], m[5], m[6]);
        o->eta_in = m[7];
        o->eta_out = m[8];
        o->absorption = cmk(m[9], m[10], m[11]);
        o->sigma_s = cmk(m[12], m[13], m[14]);
        o->sigma_a = cmk(m[15], m[16], m[17]);
        o->sigma_t = cadd(o->sigma_s, o->sigma_a);
        double density = m[18] > 0 && g_plant != OR_PLANT_DENSITY ? m[18] : 1.0;
        o->avg_sigmat = (o->sigma_t.r + o->sigma_t.g + o->sigma_t.b) / 3.0 * density;
        if (o->kind == MAT_SUBSURFACE)
            o->absorb = (o->sigma_a.r / o->sigma_t.r + o->sigma_a.g / o->sigma_t.g +
                         o->sigma_a.b / o->sigma_t.b) / 3.0;
        o->kd_pdf = (o->kd + o->ks) > 0 ? o->kd / (o->kd + o->ks) : 0.0;
        o->phong_norm = (o->e + (g_plant == OR_PLANT_PHONG_NORM ? 2.0 : 1.0)) / (2.0 * PI_D);
        o->alpha = m[19];
        o->opaque = !(o->kind == MAT_DIELECTRIC || o->kind == MAT_SUBSURFACE);
    }
    for (i = 0; i < n_shapes; ++i) {
        const float* s = shapes + 32 * (size_t)i;
        shape_t* o = &sc->shapes[i];
        o->kind = (int)s[0];
        o->mat = (int)s[1];
        o->light = (int)s[2]; /* 1 area light, 2 image-based light */
        o->a = mk(s[3], s[4], s[5]);
        o->b = mk(s[6], s[7], s[8]);
        o->c = mk(s[9], s[10], s[11]);
        o->r = s[12];
        o->n0 = mk(s[13], s[14], s[15]);
        o->n1 = mk(s[16], s[17], s[18]);
        o->n2 = mk(s[19], s[20], s[21]);
        o->tex = (int)s[22];
        for (k = 0; k < 6; ++k) o->uv[k] = s[23 + k];
        if (o->kind == OR_TRI) {
            for (k = 0; k < 3; ++k) {
                double a = vget(o->a, k), b = vget(o->b, k), c = vget(o->c, k);
                o->lo[k] = fmin(a, fmin(b, c));
                o->hi[k] = fmax(a, fmax(b, c));
            }
        } else if (o->kind == OR_SPHERE) {
            for (k = 0; k < 3; ++k) {
                o->lo[k] = vget(o->a, k) - o->r;
                o->hi[k] = vget(o->a, k) + o->r;
            }
        } else if (o->kind == OR_POINT) {
            for (k = 0; k < 3; ++k) o->lo[k] = o->hi[k] = vget(o->a, k);
        } else {
            o->n = norm(cross(o->b, o->c));
            o->length = len(o->b);
            o->width = len(o->c);
            o->inv_area = 1.0 / (o->length * o->width);
            /* full parallelogram bounds (a box slightly larger than the
               reference's two-corner box for skewed rects: conservative) */
            for (k = 0; k < 3; ++k) {
                double p0 = vget(o->a, k), l = vget(o->b, k), w = vget(o->c, k);
                o->lo[k] = p0 + fmin(0.0, l) + fmin(0.0, w);
                o->hi[k] = p0 + fmax(0.0, l) + fmax(0.0, w);
            }
        }
        if (o->light) sc->lights[sc->n_lights++] = i;
    }
    sc->cam.p = mk(cam[0], cam[1], cam[2]);
    for (k = 0; k < 9; ++k) sc->cam.R[k] = cam[3 + k];
    sc->cam.tanfov = cam[12];
    sc->cam.y_ratio = cam[13];
    sc->cam.focal = cam[14];
    sc->cam.lens = cam[15];
    medium_t* md = &sc->med;
    md->kind = med ? (int)med[0] : 0;
    if (md->kind) {
        for (k = 0; k < 3; ++k) {
            md->lo[k] = med[1 + k];
            md->hi[k] = med[4 + k];
        }
        md->c = mk(med[1], med[2], med[3]);
        md->r = med[7];
        md->sigma_s = cmk(med[8], med[9], med[10]);
        md->sigma_a = cmk(med[11], med[12], med[13]);
        md->sigma_t = cadd(md->sigma_s, md->sigma_a);
        md->avg_sigma_t = cavg(md->sigma_t);
        md->absorb = cavg(cdiv(md->sigma_a, md->sigma_t));
        md->g = med[14];
    }
    build_bvh(sc);
}

static void free_scene(scene_t* sc)
{
    free(sc->shapes);
    free(sc->mats);
    free(sc->lights);
    free(sc->nodes);
    free(sc->order);
}

typedef struct {
    const scene_t* sc;
    int W, H, spp, max_depth, row0, row1;
    uint64_t seed;
    double* mean;
    double* m2;
    /* adjoint: dLoss/dImage in, per-thread sums out */
    const double* dimg;
    double grad[7], grad_var[7];
    double* gtex; /* 3 per texel, or NULL */
    /* reference sampling: the frame's pixel set and one light set per shape
       (NULL: continuous sampling) */
    const ref_set_t* pixel_set;
    const ref_set_t* light_sets;
    int n_shapes;
} job_t;

static void* render_rows(void* arg)
{
    job_t* j = (job_t*)arg;
    const int md = j->max_depth > 0 ? j->max_depth : 1;
    pv_t* P = (pv_t*)malloc(sizeof(pv_t) * (size_t)md);
    grad_t G, *Gp = 0;
    int x, y, s, k, e;
    if (j->dimg) {
        /* per vertex: its own texel and an image-based light's */
        G.tix = (int*)malloc(sizeof(int) * 2 * (size_t)md);
        G.tval = (col*)malloc(sizeof(col) * 2 * (size_t)md);
        Gp = &G;
        for (k = 0; k < 7; ++k) j->grad[k] = j->grad_var[k] = 0.0;
    }
    /* reference sampling: this thread's copies of the frame's sets */
    ref_set_t pix = {0, 0, 0, 0}, *lsets = 0;
    if (j->pixel_set) {
        ref_set_alloc(&pix, j->pixel_set->n);
        lsets = (ref_set_t*)calloc((size_t)(j->n_shapes > 0 ? j->n_shapes : 1), sizeof(ref_set_t));
        for (k = 0; k < j->n_shapes; ++k)
            if (j->light_sets[k].n) ref_set_alloc(&lsets[k], j->light_sets[k].n);
        tl_light_sets = lsets;
    }
    for (y = j->row0; y < j->row1; ++y)
        for (x = 0; x < j->W; ++x) {
            size_t o = 3 * ((size_t)y * j->W + x);
            col sum = cs(0.0), sq = cs(0.0);
            double psum[7] = {0}, psq[7] = {0};
            pcg_t gp;
            if (j->pixel_set) {
                /* Scene::Rendering takes the pixel's samples in one pass over
                   the pixel set; the light sets are met wherever the frame's
                   earlier draws left them: a random arrangement and cursor */
                pcg_seed(&gp, j->seed ^ 0xA0761D6478BD642Full, (uint64_t)y * (uint64_t)j->W + (uint64_t)x);
                ref_set_copy(&pix, j->pixel_set);
                for (k = 0; k < j->n_shapes; ++k)
                    if (lsets[k].n) {
                        ref_set_copy(&lsets[k], &j->light_sets[k]);
                        ref_set_shuffle(&lsets[k], &gp);
                        lsets[k].cursor = (int)(rnd(&gp) * lsets[k].n);
                    }
            }
            for (s = 0; s < j->spp; ++s) {
                /* one stream per sample, so that renders at nearby parameters
                   with one seed stay correlated sample by sample */
                pcg_t g;
                pcg_seed(&g, j->seed + (uint64_t)s * 0x9E3779B97F4A7C15ull,
                         (uint64_t)y * (uint64_t)j->W + (uint64_t)x);
                double sx = rnd(&g), sy = rnd(&g);
                if (j->pixel_set) ref_set_draw(&pix, &gp, &sx, &sy); /* Film::PixelSampler()->Sampling() */
                ray_t r = camera_ray(&j->sc->cam, j->W, j->H, x, y, sx, sy, &g);
                col L = trace_pt(j->sc, r, j->max_depth, &g, P, Gp);
                if (cnan(L)) continue; /* Scene::Rendering drops NaN samples */
                sum = cadd(sum, L);
                sq = cadd(sq, cmul(L, L));
                if (Gp) {
                    const col w = cscale(cmk(j->dimg[o], j->dimg[o + 1], j->dimg[o + 2]), 1.0 / j->spp);
                    for (k = 0; k < 7; ++k) {
                        const double v = w.r * G.dL[k].r + w.g * G.dL[k].g + w.b * G.dL[k].b;
                        if (v != v) continue;
                        psum[k] += v;
                        psq[k] += v * v;
                    }
                    if (j->gtex)
                        for (e = 0; e < G.n_tex; ++e) {
                            double* q = j->gtex + 3 * (size_t)G.tix[e];
                            q[0] += w.r * G.tval[e].r;
                            q[1] += w.g * G.tval[e].g;
                            q[2] += w.b * G.tval[e].b;
                        }
                }
            }
            j->mean[o + 0] = sum.r / j->spp;
            j->mean[o + 1] = sum.g / j->spp;
            j->mean[o + 2] = sum.b / j->spp;
            if (j->m2) {
                j->m2[o + 0] = sq.r / j->spp;
                j->m2[o + 1] = sq.g / j->spp;
                j->m2[o + 2] = sq.b / j->spp;
            }
            if (Gp)
                for (k = 0; k < 7; ++k) {
                    j->grad[k] += psum[k];
                    /* variance of this pixel's sum of spp independent terms */
                    if (j->spp > 1)
                        j->grad_var[k] += (psq[k] - psum[k] * psum[k] / j->spp) * j->spp / (j->spp - 1);
                }
        }
    if (Gp) {
        free(G.tix);
        free(G.tval);
    }
    if (j->pixel_set) {
        ref_set_free(&pix);
        for (k = 0; k < j->n_shapes; ++k) ref_set_free(&lsets[k]);
        free(lsets);
        tl_light_sets = 0;
    }
    free(P);
    return 0;
}

static int run_pt(const float* shapes, int n_shapes, const float* mats, int n_mats,
                  const float* camera, const float* medium, const float* texels,
                  const int* tex_table, int n_tex, int width, int height, int spp, int max_depth,
                  uint64_t seed, int nthreads, double* mean_rgb, double* m2_rgb,
                  const double* dimg, double* grad7, double* grad7_var, double* grad_tex,
                  size_t n_texels, int whitted, const ref_set_t* pixel_set, const ref_set_t* light_sets)
{
    if (width <= 0 || height <= 0 || spp <= 0 || max_depth <= 0 || !mean_rgb) return 1;
    if (n_tex > 0 && (!texels || !tex_table)) return 1;
    scene_t sc;
    load_scene(&sc, shapes, n_shapes, mats, n_mats, camera, medium);
    sc.texels = texels;
    sc.tex_table = tex_table;
    sc.n_tex = n_tex;
    sc.whitted = whitted;
    int i, t, k;
    for (i = 0; i < sc.n_shapes; ++i)
        if (sc.shapes[i].tex >= n_tex || (sc.shapes[i].light == 2 && sc.med.kind)) {
            /* a texture index out of range; an image-based light with a medium */
            free_scene(&sc);
            return 1;
        }
    if (nthreads < 1) nthreads = 1;
    if (nthreads > height) nthreads = height;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    job_t* jobs = (job_t*)calloc((size_t)nthreads, sizeof(job_t));
    for (t = 0; t < nthreads; ++t) {
        jobs[t].sc = &sc;
        jobs[t].W = width;
        jobs[t].H = height;
        jobs[t].spp = spp;
        jobs[t].max_depth = max_depth;
        jobs[t].row0 = height * t / nthreads;
        jobs[t].row1 = height * (t + 1) / nthreads;
        jobs[t].seed = seed;
        jobs[t].mean = mean_rgb;
        jobs[t].m2 = m2_rgb;
        jobs[t].dimg = dimg;
        jobs[t].gtex = (dimg && grad_tex) ? (double*)calloc(3 * n_texels + 1, sizeof(double)) : 0;
        jobs[t].pixel_set = pixel_set;
        jobs[t].light_sets = light_sets;
        jobs[t].n_shapes = sc.n_shapes;
        if (nthreads == 1)
            render_rows(&jobs[t]);
        else
            pthread_create(&th[t], 0, render_rows, &jobs[t]);
    }
    if (nthreads > 1)
        for (t = 0; t < nthreads; ++t) pthread_join(th[t], 0);
    if (dimg) {
        for (k = 0; k < 7; ++k) {
            grad7[k] = 0.0;
            if (grad7_var) grad7_var[k] = 0.0;
        }
        if (grad_tex) memset(grad_tex, 0, sizeof(double) * 3 * n_texels);
        for (t = 0; t < nthreads; ++t) {
            for (k = 0; k < 7; ++k) {
                grad7[k] += jobs[t].grad[k];
                if (grad7_var) grad7_var[k] += jobs[t].grad_var[k];
            }
            if (jobs[t].gtex) {
                size_t q;
                for (q = 0; q < 3 * n_texels; ++q) grad_tex[q] += jobs[t].gtex[q];
                free(jobs[t].gtex);
            }
        }
    }
    free(th);
    free(jobs);
    free_scene(&sc);
    return 0;
}

int or_render_pt(const float* shapes, int n_shapes, const float* mats, int n_mats,
                 const float* camera, const float* medium, const float* texels,
                 const int* tex_table, int n_tex, int width, int height, int spp, int max_depth,
                 uint64_t seed, int nthreads, double* mean_rgb, double* m2_rgb)
{
    return run_pt(shapes, n_shapes, mats, n_mats, camera, medium, texels, tex_table, n_tex, width,
                  height, spp, max_depth, seed, nthreads, mean_rgb, m2_rgb, 0, 0, 0, 0, 0, 0, 0, 0);
}

int or_render_pt_refsampling(const float* shapes, int n_shapes, const float* mats, int n_mats,
                             const float* camera, const float* medium, const float* texels,
                             const int* tex_table, int n_tex, int width, int height, int spp,
                             int max_depth, int light_points, uint64_t seed, int nthreads,
                             double* mean_rgb, double* m2_rgb)
{
    if (spp <= 0 || light_points <= 0 || n_shapes < 0) return 1;
    int i, rc, d = (int)sqrt((float)spp);
    pcg_t gf;
    /* the frame's sets, made once as Scene::Build makes them (scene.cpp:
       145, 193): the film's UniformSample, each area light's set, shuffled */
    pcg_seed(&gf, seed ^ 0x2545F4914F6CDD1Dull, 0xF00Du);
    ref_set_t pix;
    ref_uniform_set(&pix, spp, &gf);
    ref_set_t* lights = (ref_set_t*)calloc((size_t)(n_shapes > 0 ? n_shapes : 1), sizeof(ref_set_t));
    for (i = 0; i < n_shapes; ++i) {
        const float* sh = shapes + 32 * (size_t)i;
        if ((int)sh[2] != 1) continue;
        if ((int)sh[0] == OR_RECT) {
            double l = sqrt((double)sh[6] * sh[6] + (double)sh[7] * sh[7] + (double)sh[8] * sh[8]);
            double w = sqrt((double)sh[9] * sh[9] + (double)sh[10] * sh[10] + (double)sh[11] * sh[11]);
            ref_rect_set(&lights[i], light_points, l, w, &gf);
        } else if ((int)sh[0] == OR_SPHERE) {
            ref_sphere_set(&lights[i], light_points, &gf);
        } else {
            continue;
        }
        ref_set_shuffle(&lights[i], &gf);
    }
    rc = run_pt(shapes, n_shapes, mats, n_mats, camera, medium, texels, tex_table, n_tex, width,
                height, d * d, max_depth, seed, nthreads, mean_rgb, m2_rgb, 0, 0, 0, 0, 0, 0, &pix,
                lights);
    for (i = 0; i < n_shapes; ++i) ref_set_free(&lights[i]);
    free(lights);
    ref_set_free(&pix);
    return rc;
}

int or_probe_refsets(int kind, int n, double length, double width, uint64_t seed, double* xy,
                     int* count)
{
    pcg_t g;
    ref_set_t s;
    pcg_seed(&g, seed, 7u);
    if (kind == 0)
        ref_uniform_set(&s, n, &g);
    else if (kind == 1)
        ref_rect_set(&s, n, length, width, &g);
    else if (kind == 2)
        ref_sphere_set(&s, n, &g);
    else
        return 1;
    *count = s.n;
    if (xy) {
        int i;
        for (i = 0; i < s.n; ++i) {
            xy[2 * i] = s.x[i];
            xy[2 * i + 1] = s.y[i];
        }
    }
    ref_set_free(&s);
    return 0;
}

/* RayTracer::Trace (src/Integrator/raytracer.cpp:37-86) with
 * RayTracer(scene, max_depth): the path tracer's walk with every surface
 * but Specular and Dielectric ending the path after its DirectLighting, no
 * medium sampling along rays, and trace depths 0 .. max_depth. */
int or_render_whitted(const float* shapes, int n_shapes, const float* mats, int n_mats,
                      const float* camera, const float* medium, const float* texels,
                      const int* tex_table, int n_tex, int width, int height, int spp,
                      int max_depth, uint64_t seed, int nthreads, double* mean_rgb, double* m2_rgb)
{
    return run_pt(shapes, n_shapes, mats, n_mats, camera, medium, texels, tex_table, n_tex, width,
                  height, spp, max_depth + 1, seed, nthreads, mean_rgb, m2_rgb, 0, 0, 0, 0, 0, 1, 0, 0);
}

int or_adjoint_pt(const float* shapes, int n_shapes, const float* mats, int n_mats,
                  const float* camera, const float* medium, const float* texels,
                  const int* tex_table, int n_tex, int width, int height, int spp, int max_depth,
                  uint64_t seed, int nthreads, const double* dloss_dimage, double* mean_rgb,
                  double* grad7, double* grad7_var, double* grad_tex, size_t n_texels)
{
    if (!dloss_dimage || !grad7) return 1;
    return run_pt(shapes, n_shapes, mats, n_mats, camera, medium, texels, tex_table, n_tex, width,
                  height, spp, max_depth, seed, nthreads, mean_rgb, 0, dloss_dimage, grad7,
                  grad7_var, grad_tex, n_texels, 0, 0, 0);
}

/* ====================================================================== */
/* FP64 bidirectional path tracer.  Follows BidirectionalPathTracer::Trace
 * (src/Integrator/bidirectionalpathtracer.cpp:121-394): an eye subpath
 * (ConstructViewPath :396-711) and a light subpath (ConstructLightPath
 * :713-1016) with area-measure forward/reverse pdfs, then every (s, t)
 * connection -- t = 1 splats through the pinhole onto the film
 * (Pinhole::FilmPlaneIntersection, pinhole.cpp:138-167) -- weighted by MIS
 * over all strategies with exponent beta (1 = the reference's balance
 * heuristic, MISWeight :1096-1586).  Mirror, dielectric and subsurface
 * boundary vertices and subsurface interior events are never connected.
 * Materials: matte, glossy, GGX, specular, dielectric, subsurface, emissive;
 * the homogeneous medium (medium vertices on both subpaths, a grey Bernoulli
 * survival on every connection, as VolumetricTest :1037-1077).
 *
 * Glossy picks one lobe per vertex as Glossy::BSDFSampling does
 * (src/Material/glossy.cpp:56-78): the Lambertian lobe with probability
 * kd_pdf (then the vertex is a matte vertex whose f and throughput carry
 * 1/kd_pdf, Glossy::Fr :80-90 over _pdf_path), else the Phong lobe, which
 * Glossy::IsDirac (:35-41) marks as never connected.  The Phong lobe is
 * sampled with Phong::BSDFSampling's rules (bsdf.cpp:127-170: a mirrored
 * below-horizon sample, a returned cosine of 1), so its integrand is the path
 * tracer's and is only reproduced from the eye side: light subpaths end at a
 * Phong-lobe vertex, and the MIS sums stop at one (BV_PHONG), since every
 * strategy beyond it would need a light subpath through it. */

#define BV_DELTA 1
#define BV_EMIT 2
#define BV_CAM 4
#define BV_LIGHT 8
#define BV_PHONG 16
#define BV_MEDIUM 32

typedef struct {
    v3 p, n;
    double pdf_fwd, pdf_rev;
    col beta;
    int flags, mat, prim;
} bvert_t;

/* Medium length of the segment a -> b (0 without a medium). */
static double seg_medium(const scene_t* sc, v3 a, v3 b)
{
    const medium_t* md = &sc->med;
    if (!md->kind) return 0.0;
    v3 dv = sub(b, a);
    double dist = len(dv);
    if (!(dist > 0.0)) return 0.0;
    ray_t r = {a, mul(dv, 1.0 / dist), dist, 1e300};
    double tn, tf;
    if (!med_intersect(md, &r, &tn, &tf)) return 0.0;
    double L = med_maxdist(dist, tn, tf);
    return L > 0.0 ? L : 0.0;
}

/* Density of reaching a vertex across a segment of medium length L, as the
 * path tracer samples segments (distance_sample): a surface vertex passes the
 * medium (probability exp(-sbar L)), a medium vertex is a sampled distance
 * that scattered (sbar (1 - A) exp(-sbar L)). */
static double seg_density(const scene_t* sc, double L, int medium_vertex)
{
    const medium_t* md = &sc->med;
    if (!md->kind) return 1.0;
    double tr = exp(-md->avg_sigma_t * L);
    return medium_vertex ? md->avg_sigma_t * (1.0 - md->absorb) * tr : tr;
}

/* Area (surface vertex) or volume (medium vertex) density at v of a direction
 * sampled at `from` with solid-angle pdf pdf_w, the segment holding L of
 * medium. */
static double to_meas(const scene_t* sc, double pdf_w, v3 from, const bvert_t* v, double L)
{
    v3 d = sub(v->p, from);
    double d2 = dot(d, d);
    int med = (v->flags & BV_MEDIUM) != 0;
    double c = med ? 1.0 : fabs(dot(v->n, d)) / sqrt(d2);
    return pdf_w * c / d2 * seg_density(sc, L, med);
}

/* Phase value (HomogeneousVolume::Fp without sigma_s) between the two
 * directions leaving a medium vertex toward its neighbours: light arrives
 * along -wi and leaves along wo. */
static double phase_v(const scene_t* sc, v3 wi, v3 wo)
{
    double g = sc->med.g;
    return g == 0.0 ? 1.0 / (4.0 * PI_D) : hg_eval(g, -dot(wi, wo));
}

static col bsdf_f_v(const scene_t* sc, const bvert_t* v, v3 wi, v3 wo)
{
    if (v->flags & BV_MEDIUM) return cscale(sc->med.sigma_s, phase_v(sc, wi, wo));
    const mat_t* m = &sc->mats[v->mat];
    if (m->kind == MAT_MATTE) return cscale(m->color, m->kd * INV_PI_D);
    /* only Lambertian-lobe glossy vertices are connected */
    if (m->kind == MAT_GLOSSY) return cscale(m->color, m->kd * INV_PI_D / m->kd_pdf);
    if (m->kind == MAT_GGX) return ggx_f(m, v->n, wi, wo);
    return cs(0.0);
}

static double bsdf_pdf_v(const scene_t* sc, const bvert_t* v, v3 wi, v3 wo)
{
    if (v->flags & BV_MEDIUM) return phase_v(sc, wi, wo);
    const mat_t* m = &sc->mats[v->mat];
    if (m->kind == MAT_MATTE || m->kind == MAT_GLOSSY) {
        double p = INV_PI_D * dot(v->n, wi); /* Lambertian::Pdfw (bsdf.cpp:34-41) */
        return p <= 0.0 ? 1e-4 : p;
    }
    if (m->kind == MAT_GGX) return ggx_pdf(m, v->n, wi, wo);
    return 0.0;
}

static int bd_walk(const scene_t* sc, bvert_t* arr, int start, int maxv, v3 o, v3 d, col beta,
                   double pdf_w, pcg_t* g, int eye)
{
    const medium_t* md = &sc->med;
    int nv = start, in_obj = 0;
    v3 prev_p = o;
    while (nv < maxv) {
        ray_t r = {o, d, 0.0, 1e300};
        hitrec_t h = {-1, 0, 0, 0};
        int hit = travel(sc, &r, 0, in_obj, &h);
        /* the medium along the segment, sampled as the path tracer does
         * (pathtracer.cpp:156-178; bidirectionalpathtracer.cpp:452-473) */
        double seg_L = 0.0;
        int scatter = 0;
        v3 p, n;
        if (md->kind && !in_obj) {
            double tn, tf;
            if (med_intersect(md, &r, &tn, &tf)) {
                double smax = med_maxdist(r.tmin, tn, tf), sd = 0.0;
                int absorbed = 0;
                col tau, pdf;
                scatter = distance_sample(md->sigma_t, md->avg_sigma_t, md->absorb, smax, g, &sd,
                                          &absorbed, &tau, &pdf);
                if (absorbed) break;
                beta = cmul(beta, cdiv(tau, pdf));
                if (scatter) {
                    seg_L = sd;
                    p = add(r.o, mul(r.d, tn < 0.0 ? sd : tn + sd));
                } else {
                    seg_L = smax;
                }
            }
        }
        if (!scatter && hit < 0) break;
        if (scatter) {
            /* a medium vertex: HomogeneousVolume::PhaseSampling, Fp / pdf = sigma_s */
            bvert_t* v = &arr[nv];
            v->p = p;
            v->n = d; /* no surface: the arrival direction */
            v->pdf_rev = 0.0;
            v->beta = beta;
            v->flags = BV_MEDIUM;
            v->mat = -1;
            v->prim = -1;
            v->pdf_fwd = to_meas(sc, pdf_w, prev_p, v, seg_L);
            ++nv;
            double pdf_ph;
            v3 w = phase_sample(md->g, d, g, &pdf_ph);
            if (nv >= 2) {
                bvert_t* pv = &arr[nv - 2];
                pv->pdf_rev = to_meas(sc, pdf_ph, p, pv, seg_L);
            }
            beta = cmul(beta, md->sigma_s);
            pdf_w = pdf_ph;
            prev_p = p;
            o = p;
            d = w;
            in_obj = 0;
            continue;
        }
        surface(sc, &h, &r, &p, &n);
        const mat_t* m = &sc->mats[sc->shapes[hit].mat];
        col vol = cs(1.0); /* subsurface pass-through tau / pdf at the exit boundary */
        if (m->kind == MAT_SUBSURFACE && in_obj) {
            /* Subsurface::PhaseSampling (subsurface.cpp:135-155): the interior
             * walk over the segment to the boundary.  Interior vertices are
             * never connected (a connection from inside ends on the
             * boundary), so they enter the MIS walk like Dirac vertices. */
            double sd = 0.0;
            int absorbed = 0;
            col tau, pdf;
            int sc_ev = distance_sample(m->sigma_t, m->avg_sigmat, m->absorb, r.tmin, g, &sd, &absorbed,
                                        &tau, &pdf);
            if (absorbed) break;
            if (sc_ev) {
                v3 ps = add(r.o, mul(r.d, sd));
                bvert_t* v = &arr[nv];
                v->p = ps;
                v->n = d;
                v->pdf_fwd = 0.0;
                v->pdf_rev = 0.0;
                v->beta = beta;
                v->flags = BV_DELTA;
                v->mat = sc->shapes[hit].mat;
                v->prim = hit;
                ++nv;
                if (nv >= 2) arr[nv - 2].pdf_rev = 0.0;
                beta = cmul(cmul(beta, cdiv(tau, pdf)), m->sigma_s);
                double pdf_ph;
                v3 w = phase_sample(0.0, d, g, &pdf_ph);
                pdf_w = 0.0;
                prev_p = ps;
                o = ps;
                d = w;
                continue;
            }
            vol = cdiv(tau, pdf);
        }
        bvert_t* v = &arr[nv];
        v->p = p;
        v->n = n;
        v->pdf_rev = 0.0;
        v->beta = beta;
        v->flags = 0;
        v->mat = sc->shapes[hit].mat;
        v->prim = hit;
        v->pdf_fwd = to_meas(sc, pdf_w, prev_p, v, seg_L);
        ++nv;
        if (m->kind == MAT_EMISSIVE) {
            v->flags = BV_EMIT;
            break;
        }
        int was_in = in_obj;
        v3 wo = neg(d), w;
        double pdf_next = 0.0, pdf_rev = 0.0;
        int delta = 0;
        col weight;
        if (m->kind == MAT_MATTE) {
            double cost, pdf;
            w = lambert_sample(n, g, &cost, &pdf);
            weight = cscale(m->color, m->kd * INV_PI_D * cost / pdf);
            pdf_next = pdf;
            double pr = INV_PI_D * dot(n, wo);
            pdf_rev = pr <= 0.0 ? 1e-4 : pr;
        } else if (m->kind == MAT_GLOSSY) {
            if (m->kd != 0.0 && rnd(g) <= m->kd_pdf) {
                double cost, pdf;
                w = lambert_sample(n, g, &cost, &pdf);
                weight = cscale(m->color, m->kd * INV_PI_D * cost / (pdf * m->kd_pdf));
                pdf_next = pdf;
                double pr = INV_PI_D * dot(n, wo);
                pdf_rev = pr <= 0.0 ? 1e-4 : pr;
            } else {
                v->flags = BV_DELTA | BV_PHONG;
                if (!eye) break;
                double cost, pdf;
                w = phong_sample(m, wo, n, g, &cost, &pdf);
                weight = cs(phong_f(m, n, w, wo) * m->ks * cost / (pdf * (1.0 - m->kd_pdf)));
                delta = 1;
            }
        } else if (m->kind == MAT_GGX) {
            w = ggx_sample(m, n, wo, g);
            double cost = dot(n, w);
            pdf_next = ggx_pdf(m, n, w, wo);
            if (!(cost > 0.0 && pdf_next > 0.0)) break;
            weight = cscale(ggx_f(m, n, w, wo), cost / pdf_next);
            pdf_rev = ggx_pdf(m, n, wo, w);
        } else if (m->kind == MAT_SPECULAR) {
            double won = dot(wo, n);
            if (won <= 0.0) won = 1e-4;
            w = add(neg(wo), mul(n, 2.0 * won));
            weight = cscale(m->color, m->kd);
            delta = 1;
        } else if (m->kind == MAT_DIELECTRIC || m->kind == MAT_SUBSURFACE) {
            double kr = fresnel(neg(wo), n, m->eta_in, m->eta_out), c;
            delta = 1;
            if (kr != 0.0 && rnd(g) <= kr) {
                w = fresnel_reflect(wo, n, &c);
                in_obj = c < 0.0;
                weight = cs(1.0);
            } else {
                c = 0.0;
                w = fresnel_transmit(m, wo, n, &c);
                in_obj = c > 0.0;
                double er = was_in ? m->eta_in / m->eta_out : m->eta_out / m->eta_in;
                weight = cs(eye ? 1.0 / (er * er) : 1.0);
            }
            if (was_in && m->kind == MAT_DIELECTRIC) weight = cmul(weight, exp_tau(m->absorption, r.tmin));
            if (was_in && m->kind == MAT_SUBSURFACE) weight = cmul(weight, vol);
        } else {
            break;
        }
        if (delta) v->flags |= BV_DELTA;
        if (nv >= 2) {
            bvert_t* pv = &arr[nv - 2];
            pv->pdf_rev = delta ? 0.0 : to_meas(sc, pdf_rev, p, pv, seg_L);
        }
        beta = cmul(beta, weight);
        pdf_w = delta ? 0.0 : pdf_next;
        prev_p = p;
        o = p;
        d = w;
        if (m->kind != MAT_DIELECTRIC && m->kind != MAT_SUBSURFACE) in_obj = 0;
    }
    return nv;
}

/* Inverse of Pinhole::CastViewRay's raster mapping for a world direction. */
static int camera_project(const camera_t* c, int W, int H, v3 dc, int* pixel, double* cos_c)
{
    double lx = c->R[0] * dc.x + c->R[3] * dc.y + c->R[6] * dc.z;
    double ly = c->R[1] * dc.x + c->R[4] * dc.y + c->R[7] * dc.z;
    double lz = c->R[2] * dc.x + c->R[5] * dc.y + c->R[8] * dc.z;
    if (!(lz > 0.0)) return 0;
    double px = lx / lz, py = ly / lz, hw = W * 0.5, hh = H * 0.5;
    double rx = hw - px * hw / c->tanfov, ry = hh - py * hh / (c->tanfov * c->y_ratio);
    if (!(rx >= 0.0 && rx < W && ry >= 0.0 && ry < H)) return 0;
    int ix = (int)rx, iy = (int)ry;
    if (ix > W - 1) ix = W - 1;
    if (iy > H - 1) iy = H - 1;
    *pixel = iy * W + ix;
    *cos_c = lz;
    return 1;
}

static int visible(const scene_t* sc, v3 a, v3 b)
{
    v3 dv = sub(b, a);
    double dist = len(dv);
    ray_t r = {a, mul(dv, 1.0 / dist), 0.0, dist};
    travel(sc, &r, 1, 0, 0);
    return !(r.tmin > 0.0);
}

static double rm0(double f) { return f != 0.0 ? f : 1.0; }

typedef struct {
    const scene_t* sc;
    int W, H, spp, nE, nL, max_len, row0, row1;
    double beta;
    uint64_t seed;
    int* light_of; /* shape -> light index or -1 */
    double* film;  /* per-thread accumulation, W*H*3 */
} bjob_t;

#define BDPT_LIGHT_MIX 0x2545F4914F6CDD1Dull
#define BDPT_CONN_MIX 0x9FB21C651E98DF25ull

static void* bdpt_rows(void* arg)
{
    bjob_t* j = (bjob_t*)arg;
    const scene_t* sc = j->sc;
    const camera_t* cam = &sc->cam;
    double film_area = (2.0 * cam->tanfov) * (2.0 * cam->tanfov * cam->y_ratio);
    bvert_t* E = (bvert_t*)calloc((size_t)j->nE, sizeof(bvert_t));
    bvert_t* Lv = (bvert_t*)calloc((size_t)j->nL, sizeof(bvert_t));
    int x, y, smp, s, t, i;
    for (y = j->row0; y < j->row1; ++y)
        for (x = 0; x < j->W; ++x) {
            uint64_t pixel = (uint64_t)y * j->W + x;
            pcg_t g, gl;
            for (smp = 0; smp < j->spp; ++smp) {
                /* every sample's eye and light subpaths draw from their own
                 * PCG32 streams, numbered as the GPU tracer numbers them */
                pcg_seed(&g, j->seed + (uint64_t)smp * 0x9E3779B97F4A7C15ull, pixel);
                pcg_seed(&gl, (j->seed ^ BDPT_LIGHT_MIX) + (uint64_t)smp * 0x9E3779B97F4A7C15ull, pixel);
                /* eye subpath */
                double sx = rnd(&g), sy = rnd(&g);
                ray_t cr = camera_ray(cam, j->W, j->H, x, y, sx, sy, 0);
                double hw = j->W * 0.5, hh = j->H * 0.5;
                double lpx = -(((double)x - hw) + sx) / hw * cam->tanfov;
                double lpy = -(((double)y - hh) + sy) / hh * cam->tanfov * cam->y_ratio;
                double lz = 1.0 / sqrt(lpx * lpx + lpy * lpy + 1.0);
                E[0].p = cam->p;
                E[0].n = mk(cam->R[2], cam->R[5], cam->R[8]);
                E[0].pdf_fwd = 1.0;
                E[0].pdf_rev = 0.0;
                E[0].beta = cs(1.0);
                E[0].flags = BV_CAM;
                double pdf_dir = 1.0 / (lz * lz * lz * film_area);
                int nE = bd_walk(sc, E, 1, j->nE, cr.o, cr.d, cs(1.0), pdf_dir, &g, 1);
                /* light subpath */
                int nL = 0;
                if (sc->n_lights > 0) {
                    lsample_t ls;
                    light_sample(sc, mk(0, 0, 0), mk(0, 1, 0), &gl, &ls);
                    Lv[0].p = ls.p;
                    Lv[0].n = ls.n;
                    Lv[0].pdf_fwd = ls.pdf;
                    Lv[0].pdf_rev = 0.0;
                    Lv[0].beta = cscale(ls.le, 1.0 / ls.pdf);
                    Lv[0].flags = BV_LIGHT;
                    Lv[0].mat = -1;
                    double cost, pdfd;
                    v3 w = lambert_sample(ls.n, &gl, &cost, &pdfd);
                    col b = cscale(ls.le, cost / (ls.pdf * pdfd));
                    nL = bd_walk(sc, Lv, 1, j->nL, ls.p, w, b, pdfd, &gl, 0);
                }
                col own = cs(0.0);
                for (t = 1; t <= nE; ++t)
                    for (s = 0; s <= nL; ++s) {
                        if (s + t < 2 || (s == 1 && t == 1) || s + t - 1 > j->max_len) continue;
                        const bvert_t* pt = &E[t - 1];
                        double pt_rev = 0, ptm_rev = 0, qs_rev = 0, qsm_rev = 0;
                        int trivial = 0, splat = -1;
                        col c;
                        v3 va, vb;
                        int vis = 0;
                        double medL = 0.0; /* medium length of the connecting segment */
                        if (s == 0) {
                            if (!(pt->flags & BV_EMIT) || t < 2) continue;
                            const bvert_t* pm = &E[t - 2];
                            v3 to_prev = norm(sub(pm->p, pt->p));
                            if (!(dot(pt->n, to_prev) > 0.0)) continue;
                            const mat_t* m = &sc->mats[pt->mat];
                            c = cmul(pt->beta, cscale(m->color, m->kd));
                            int li = j->light_of[pt->prim];
                            if (li < 0) {
                                trivial = 1;
                            } else {
                                const shape_t* sh = &sc->shapes[sc->lights[li]];
                                double area = sh->kind == OR_RECT ? sh->length * sh->width
                                                                  : 4.0 * PI_D * sh->r * sh->r;
                                pt_rev = (1.0 / sc->n_lights) * (1.0 / area);
                                if (!(pm->flags & BV_CAM))
                                    ptm_rev = to_meas(sc, INV_PI_D * dot(pt->n, to_prev), pt->p, pm,
                                                      seg_medium(sc, pt->p, pm->p));
                            }
                        } else {
                            const bvert_t* qs = &Lv[s - 1];
                            if ((qs->flags & BV_DELTA) || (qs->flags & BV_EMIT)) continue;
                            if (t == 1) {
                                v3 dc = norm(sub(qs->p, cam->p));
                                double cos_c;
                                if (!camera_project(cam, j->W, j->H, dc, &splat, &cos_c)) continue;
                                v3 to_cam = neg(dc);
                                v3 dd = sub(qs->p, cam->p);
                                double d2 = dot(dd, dd);
                                double cos_q = (qs->flags & BV_MEDIUM) ? 1.0 : dot(qs->n, to_cam);
                                if (!(cos_q > 0.0)) continue;
                                col fL = cs(1.0);
                                if (s >= 2) {
                                    const bvert_t* qm = &Lv[s - 2];
                                    v3 to_qm = norm(sub(qm->p, qs->p));
                                    fL = bsdf_f_v(sc, qs, to_cam, to_qm);
                                    qsm_rev = to_meas(sc, bsdf_pdf_v(sc, qs, to_qm, to_cam), qs->p, qm,
                                                      seg_medium(sc, qs->p, qm->p));
                                }
                                double We = 1.0 / (film_area * cos_c * cos_c * cos_c * cos_c);
                                double G = cos_q * cos_c / d2;
                                c = cscale(cmul(qs->beta, fL), We * G);
                                medL = seg_medium(sc, qs->p, cam->p);
                                qs_rev = to_meas(sc, 1.0 / (cos_c * cos_c * cos_c * film_area), cam->p, qs, medL);
                                va = qs->p;
                                vb = cam->p;
                                vis = 1;
                            } else {
                                if ((pt->flags & BV_DELTA) || (pt->flags & BV_EMIT)) continue;
                                const bvert_t* pm = &E[t - 2];
                                v3 dpq = sub(qs->p, pt->p);
                                double d2 = dot(dpq, dpq);
                                v3 wpq = mul(dpq, 1.0 / sqrt(d2));
                                double cos_p = (pt->flags & BV_MEDIUM) ? 1.0 : dot(pt->n, wpq);
                                double cos_q = (qs->flags & BV_MEDIUM) ? 1.0 : -dot(qs->n, wpq);
                                if (!(cos_p > 0.0 && cos_q > 0.0)) continue;
                                v3 to_pm = norm(sub(pm->p, pt->p));
                                col fE = bsdf_f_v(sc, pt, wpq, to_pm);
                                col fL = cs(1.0);
                                medL = seg_medium(sc, pt->p, qs->p);
                                if (s == 1) {
                                    pt_rev = to_meas(sc, INV_PI_D * cos_q, qs->p, pt, medL);
                                } else {
                                    const bvert_t* qm = &Lv[s - 2];
                                    v3 to_qm = norm(sub(qm->p, qs->p));
                                    fL = bsdf_f_v(sc, qs, neg(wpq), to_qm);
                                    pt_rev = to_meas(sc, bsdf_pdf_v(sc, qs, neg(wpq), to_qm), qs->p, pt, medL);
                                    qsm_rev = to_meas(sc, bsdf_pdf_v(sc, qs, to_qm, neg(wpq)), qs->p, qm,
                                                      seg_medium(sc, qs->p, qm->p));
                                }
                                c = cscale(cmul(cmul(cmul(pt->beta, fE), fL), qs->beta), cos_p * cos_q / d2);
                                qs_rev = to_meas(sc, bsdf_pdf_v(sc, pt, wpq, to_pm), pt->p, qs, medL);
                                if (!(pm->flags & BV_CAM))
                                    ptm_rev = to_meas(sc, bsdf_pdf_v(sc, pt, to_pm, wpq), pt->p, pm,
                                                      seg_medium(sc, pt->p, pm->p));
                                va = qs->p;
                                vb = pt->p;
                                vis = 1;
                            }
                        }
                        if (cis0(c)) continue;
                        double w = 1.0;
                        if (s + t > 2 && !trivial) {
                            double sum = 0.0, ri = 1.0;
                            for (i = t - 1; i > 0; --i) {
                                double pr = i == t - 1 ? pt_rev : (i == t - 2 ? ptm_rev : E[i].pdf_rev);
                                ri *= rm0(pr) / rm0(E[i].pdf_fwd);
                                int di = i == t - 1 ? 0 : (E[i].flags & BV_DELTA) != 0;
                                int dp = (E[i - 1].flags & BV_DELTA) != 0;
                                if (!di && !dp) sum += pow(ri, j->beta);
                                if (E[i].flags & BV_PHONG) break;
                            }
                            ri = 1.0;
                            for (i = s - 1; i >= 0; --i) {
                                double pr = i == s - 1 ? qs_rev : (i == s - 2 ? qsm_rev : Lv[i].pdf_rev);
                                ri *= rm0(pr) / rm0(Lv[i].pdf_fwd);
                                int di = i == s - 1 ? 0 : (Lv[i].flags & BV_DELTA) != 0;
                                int dp = i > 0 ? (Lv[i - 1].flags & BV_DELTA) != 0 : 0;
                                if (!di && !dp) sum += pow(ri, j->beta);
                            }
                            w = 1.0 / (1.0 + sum);
                        }
                        /* the connection's medium: a grey Bernoulli survival, as
                         * VolumetricTest (bidirectionalpathtracer.cpp:1050-1062)
                         * and the path tracer's shadow rays; one number from the
                         * connection's own stream */
                        if (medL > 0.0) {
                            pcg_t gc;
                            uint64_t k = (uint64_t)smp * (uint64_t)j->W * (uint64_t)j->H + pixel;
                            pcg_seed(&gc, (j->seed ^ BDPT_CONN_MIX) + k * 0x9E3779B97F4A7C15ull,
                                     ((uint64_t)s << 8) | (uint64_t)t);
                            if (rnd(&gc) < 1.0 - exp(-sc->med.avg_sigma_t * medL)) continue;
                        }
                        if (vis && !visible(sc, va, vb)) continue;
                        c = cscale(c, w);
                        if (cnan(c)) continue;
                        if (splat >= 0) {
                            double* f = j->film + 3 * (size_t)splat;
                            f[0] += c.r;
                            f[1] += c.g;
                            f[2] += c.b;
                        } else {
                            own = cadd(own, c);
                        }
                    }
                double* f = j->film + 3 * pixel;
                f[0] += own.r;
                f[1] += own.g;
                f[2] += own.b;
            }
        }
    free(E);
    free(Lv);
    return 0;
}

int or_render_bdpt(const float* shapes, int n_shapes, const float* mats, int n_mats,
                   const float* camera, const float* medium, int width, int height, int spp,
                   int n_eye, int n_light, int max_len, double mis_beta, uint64_t seed,
                   int nthreads, double* mean_rgb)
{
    if (width <= 0 || height <= 0 || spp <= 0 || n_eye < 2 || n_light < 1 || !mean_rgb) return 1;
    scene_t sc;
    load_scene(&sc, shapes, n_shapes, mats, n_mats, camera, medium);
    int i, t;
    for (i = 0; i < sc.n_shapes; ++i)
        if (sc.shapes[i].light == 2) { /* no image-based light in the bidirectional tracer */
            free_scene(&sc);
            return 1;
        }
    int* light_of = (int*)malloc(sizeof(int) * (size_t)(n_shapes > 0 ? n_shapes : 1));
    for (i = 0; i < n_shapes; ++i) light_of[i] = -1;
    for (i = 0; i < sc.n_lights; ++i) light_of[sc.lights[i]] = i;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > height) nthreads = height;
    size_t fn = 3 * (size_t)width * height;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    bjob_t* jobs = (bjob_t*)malloc(sizeof(bjob_t) * nthreads);
    for (t = 0; t < nthreads; ++t) {
        bjob_t* b = &jobs[t];
        b->sc = &sc;
        b->W = width;
        b->H = height;
        b->spp = spp;
        b->nE = n_eye;
        b->nL = n_light;
        b->max_len = max_len;
        b->row0 = height * t / nthreads;
        b->row1 = height * (t + 1) / nthreads;
        b->beta = mis_beta;
        b->seed = seed;
        b->light_of = light_of;
        b->film = (double*)calloc(fn, sizeof(double));
        if (nthreads == 1)
            bdpt_rows(b);
        else
            pthread_create(&th[t], 0, bdpt_rows, b);
    }
    if (nthreads > 1)
        for (t = 0; t < nthreads; ++t) pthread_join(th[t], 0);
    memset(mean_rgb, 0, sizeof(double) * fn);
    for (t = 0; t < nthreads; ++t) {
        size_t k;
        for (k = 0; k < fn; ++k) mean_rgb[k] += jobs[t].film[k] / spp;
        free(jobs[t].film);
    }
    free(th);
    free(jobs);
    free(light_of);
    free_scene(&sc);
    return 0;
}

/* ------------------------------------------------------------ probes */
/* or_probe (oracle.h): the samplers and evaluators above, one call per
 * sample, for the tests that pin them against closed-form densities. */
static mat_t probe_mat(const double* par)
{
    mat_t m;
    memset(&m, 0, sizeof(m));
    m.kind = MAT_GLOSSY;
    m.e = par[0];
    m.phong_norm = (m.e + (g_plant == OR_PLANT_PHONG_NORM ? 2.0 : 1.0)) / (2.0 * PI_D);
    m.alpha = par[0];
    m.color = cs(1.0);
    m.eta_in = par[0];
    m.eta_out = par[1];
    return m;
}

int or_probe(int kind, const double* par, const double* in, int n, uint64_t seed, double* out)
{
    pcg_t g;
    int i;
    if (n < 0 || !par || !out) return 1;
    pcg_seed(&g, seed, 1u);
    mat_t m = probe_mat(par);
    for (i = 0; i < n; ++i) {
        double* o = out + 8 * (size_t)i;
        const double* x = in ? in + 6 * (size_t)i : 0;
        memset(o, 0, 8 * sizeof(double));
        switch (kind) {
        case OR_PROBE_LAMBERT: {
            double c, p;
            v3 w = lambert_sample(mk(par[0], par[1], par[2]), &g, &c, &p);
            o[0] = w.x, o[1] = w.y, o[2] = w.z, o[3] = c, o[4] = p;
            break;
        }
        case OR_PROBE_PHONG: {
            double c, p;
            v3 wo = mk(par[1], par[2], par[3]), nn = mk(par[4], par[5], par[6]);
            v3 w = phong_sample(&m, wo, nn, &g, &c, &p);
            o[0] = w.x, o[1] = w.y, o[2] = w.z, o[3] = c, o[4] = p;
            o[5] = phong_f(&m, nn, w, wo);
            break;
        }
        case OR_PROBE_PHASE: {
            double p;
            v3 d = mk(par[1], par[2], par[3]);
            v3 w = phase_sample(par[0], d, &g, &p);
            o[0] = w.x, o[1] = w.y, o[2] = w.z, o[3] = dot(d, w), o[4] = p;
            break;
        }
        case OR_PROBE_GGX: {
            v3 wo = mk(par[1], par[2], par[3]), nn = mk(par[4], par[5], par[6]);
            v3 w = x ? mk(x[0], x[1], x[2]) : ggx_sample(&m, nn, wo, &g);
            o[0] = w.x, o[1] = w.y, o[2] = w.z, o[3] = dot(nn, w);
            o[4] = ggx_pdf(&m, nn, w, wo);
            o[5] = ggx_f(&m, nn, w, wo).r;
            break;
        }
        case OR_PROBE_FRESNEL:
            if (!x) return 1;
            o[0] = fresnel(mk(x[0], x[1], x[2]), mk(x[3], x[4], x[5]), par[0], par[1]);
            break;
        case OR_PROBE_TRANSMIT:
        case OR_PROBE_REFLECT: {
            double c = 0.0;
            if (!x) return 1;
            v3 wo = mk(x[0], x[1], x[2]), nn = mk(x[3], x[4], x[5]);
            v3 w = kind == OR_PROBE_TRANSMIT ? fresnel_transmit(&m, wo, nn, &c) : fresnel_reflect(wo, nn, &c);
            o[0] = w.x, o[1] = w.y, o[2] = w.z, o[3] = c;
            break;
        }
        case OR_PROBE_DISTANCE: {
            col st = cmk(par[0], par[1], par[2]), tau, pdf;
            double s = 0.0;
            int ab = 0;
            o[0] = distance_sample(st, par[3], par[4], par[5], &g, &s, &ab, &tau, &pdf);
            o[1] = s, o[2] = ab;
            if (!ab) o[3] = tau.r / pdf.r, o[4] = tau.g / pdf.g, o[5] = tau.b / pdf.b;
            break;
        }
        default:
            return 1;
        }
    }
    return 0;
}
