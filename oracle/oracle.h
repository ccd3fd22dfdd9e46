/*
 * oracle.h -- CPU restatement of Rumi-Renderer's hot path, used ONLY as the
 * checker in tests/, in __graft_entry__.smoke() and in bench.py's CPU-baseline
 * leg.  Nothing in the product (rumi-renderer_amd/) links, loads or calls it.
 *
 * Every function cites the reference source it restates (paths relative to
 * the Rumi-Renderer source tree).  Arithmetic keeps the reference's FP32
 * operation order; the library is compiled with -ffp-contract=off so no
 * multiply-add is fused.
 *
 * Parity pins (what checks this file, besides the product it checks):
 *   or_triangle_intersect / or_sphere_intersect / or_rect_intersect /
 *   or_bbox_intersect : closed-form cases in tests/test_oracle_geometry.py
 *   or_bvh_build_median6 : structural invariants (leaf size, median cut,
 *                          sortedness on the split axis, bounds nesting)
 *   or_travel_*          : brute force over all primitives (BVH-independent
 *                          closest hit; ties within 1e-6 relative t)
 */
#ifndef RUMI_ORACLE_H
#define RUMI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Ray::_epsilon (src/Kernel/ray.cpp:4). */
#define OR_EPS 1e-4f

/* Ray (src/Kernel/ray.h:9-31): tmin is "closest hit so far", 0 = none. */
typedef struct {
    float o[3];
    float d[3];
    float tmin;
    float tmax;
} or_ray;

/* Primitive kinds of the shape table. */
enum { OR_TRI = 0, OR_SPHERE = 1, OR_RECT = 2, OR_POINT = 3 };

/* One scene primitive, restating the data members the reference's shapes
 * keep (src/Shape/triangle.h:11-17, sphere.h, rect.h). */
typedef struct {
    int kind;
    float a[3], b[3], c[3];  /* tri: v0,v1,v2; sphere: a=centre; rect: a=p, b=l, c=w */
    float n[3];              /* rect: unit normal */
    float r;                 /* sphere radius */
    float length, width, inv_area; /* rect */
    int emissive;            /* material Le() != 0 (rect back-face rule, sphere normal) */
    int le_textured;         /* sphere: IsLe() && texture (image-based light) */
} or_prim;

/* ------------------------------------------------------------------ */
/* Intersectors.  Each returns 1 when the primitive is accepted and then
 * has written ray->tmin = t, exactly like the reference.                */
int or_triangle_intersect(const float v0[3], const float v1[3], const float v2[3],
                          or_ray* ray, float* u, float* v, float* det);
int or_sphere_intersect(const or_prim* s, or_ray* ray, int in_object);
int or_rect_intersect(const or_prim* q, or_ray* ray);
int or_prim_intersect(const or_prim* p, or_ray* ray, int in_object);

/* BBox::Intersect (src/Kernel/bbox.cpp:30-92). */
int or_bbox_intersect(const float lo[3], const float hi[3], const or_ray* ray,
                      float* tnear, float* tfar);

/* Shape::Bound overrides: triangle.cpp:120-148, sphere.cpp:139-145,
 * rect.cpp:141-157. */
void or_prim_bound(const or_prim* p, float lo[3], float hi[3]);

/* ------------------------------------------------------------------ */
/* Flat BVH: every node carries its own box, as BVHNode does
 * (src/Acceleration/bvh.h:9-36).  left/right < 0 marks a leaf whose
 * primitives are order[first .. first+count). */
typedef struct {
    int n_nodes;
    int root;
    int n_prims;
    float* lo;     /* 3 * n_nodes */
    float* hi;     /* 3 * n_nodes */
    int* left;
    int* right;
    int* first;
    int* count;
    int* order;    /* n_prims: primitive index per leaf slot */
} or_bvh;

/* BVHAccel::BuildBVH + CreateBVHNode (src/Acceleration/bvh.cpp:48-148):
 * median split on the longest axis after a shell sort of box centres,
 * sort skipped when the axis repeats, leaves of at most 6.  Input is the
 * per-primitive bounds (prim_lo/prim_hi, 3 floats each) in scene order. */
int or_bvh_build_median6(const float* prim_lo, const float* prim_hi, int n, or_bvh* out);

/* Rebuild the flat layout from a child-box BVH (two child boxes per
 * interior node, child ref >= 0 = interior node, < 0 = leaf), the form the
 * GPU library exports.  Leaf ref encoding: ~((first << 4) | count). */
int or_bvh_from_childbox(int n_inner, const float* child_boxes /* 12 per node */,
                         const int* child_refs /* 2 per node */,
                         const float root_lo[3], const float root_hi[3], int root_ref,
                         const int* leaf_prims, int n_prims, or_bvh* out);

void or_bvh_free(or_bvh* b);

/* Traversal statistics: interior nodes entered, primitive tests. */
typedef struct {
    int64_t nodes;
    int64_t prims;
} or_stats;

/* BVHAccel::Travel (src/Acceleration/bvh.cpp:150-212).
 *   any      = 0: closest hit (isect != NULL); 1: shadow ray (isect == NULL),
 *              return after the first leaf that produced a hit.
 *   ordered  = 0: reference order (push right, descend left).
 *            = 1: near child first by box entry distance with postponed
 *              leaves (the GPU's order for SAH trees): the same closest
 *              distance; ids differ only between primitives at exactly
 *              equal distance.
 *            = 2: near-first over the tree's 4-wide form, children
 *              ordered by order_key, the stack keeping key_distance, up
 *              to four of a leaf's primitives per step (the GPU's
 *              width-4 mesh traversal).
 * Returns the primitive index (scene order) or -1; ray->tmin holds t. */
int or_travel(const or_bvh* bvh, const or_prim* prims, or_ray* ray, int any,
              int ordered, int in_object, or_stats* st);

/* Batched form over a triangle mesh given as indexed vertices (the
 * product's input format).  rays: n x 8 floats (ox,oy,oz,tmin,dx,dy,dz,tmax).
 * Outputs prim id (-1 miss) and t (0 on miss) per ray; per-ray node and
 * primitive counts when node_cnt / prim_cnt are non-NULL. */
int or_trace_mesh(const or_bvh* bvh, const float* xyz, const uint32_t* idx, int nt,
                  const float* rays, int n, int any, int ordered,
                  int32_t* prim_out, float* t_out, int32_t* node_cnt, int32_t* prim_cnt);

/* Brute force closest hit over all triangles in index order with the
 * Triangle::Intersect acceptance rules (BVH-independent pin). */
int or_brute_mesh(const float* xyz, const uint32_t* idx, int nt, const float* rays, int n,
                  int32_t* prim_out, float* t_out);

/* Convenience: build the scene-order triangle primitive table. */
int or_mesh_prims(const float* xyz, const uint32_t* idx, int nt, or_prim* out);
int or_mesh_bounds(const float* xyz, const uint32_t* idx, int nt, float* lo, float* hi);

/* Query the layout of a median6 BVH built by or_bvh_build_median6 (for
 * structural comparison with the product's reference-mode builder). */
int or_bvh_nodes(const or_bvh* b);
void or_bvh_copy(const or_bvh* b, float* lo, float* hi, int* left, int* right,
                 int* first, int* count, int* order, int* root);

/* Median6 build + traversal driver over a triangle mesh, for timing the
 * reference's CPU path (bench CPU baseline).  Returns the seconds spent in
 * the traversal loop only (std-like monotonic clock), -1 on error. */
double or_time_median6_mesh(const float* xyz, const uint32_t* idx, int nt,
                            const float* rays, int n, int any,
                            int32_t* prim_out, float* t_out, double* build_seconds);
/* The same over nthreads host threads (the ray batch split in equal
 * contiguous ranges, one shared tree); the wall time of the ray loop. */
double or_time_median6_mesh_mt(const float* xyz, const uint32_t* idx, int nt,
                               const float* rays, int n, int any,
                               int32_t* prim_out, float* t_out, double* build_seconds,
                               int nthreads);

/* ------------------------------------------------------------------ */
/* FP64 path tracer (oracle_render.c): PathTracer::Trace(w, h, ...)
 * (src/Integrator/pathtracer.cpp:117-354) over a flat scene description.
 *   shapes    n_shapes x 32 floats: [0] kind (OR_TRI/OR_SPHERE/OR_RECT),
 *             [1] material, [2] area-light flag, [3..11] a,b,c (tri v0 v1 v2;
 *             sphere centre in a; rect p, l, w), [12] radius, [13..21] tri
 *             vertex normals n0 n1 n2, [22] texture index (-1 none),
 *             [23..28] tri vertex uv u0 v0 u1 v1 u2 v2.  OR_POINT: a
 *             PointLight, never hit: position in a, Le = rgb * ls in
 *             [13..15], flag [2] = 1.
 *   mats      n_mats x 32 floats: [0] kind (0 matte, 1 glossy, 2 specular,
 *             3 emissive, 4 dielectric, 5 subsurface, 6 GGX), [1] kd
 *             (specular kr, emissive ls), [2] ks, [3] Phong exponent,
 *             [4..6] rgb, [7] eta_in, [8] eta_out, [9..11] absorption,
 *             [12..14] sigma_s, [15..17] sigma_a, [18] density, [19] GGX
 *             roughness.
 *   camera    16 floats: position, row-major 3x3 basis, tan(fov/2), y_ratio,
 *             focal distance and lens radius (thin lens; focal 0 = pinhole).
 *   medium    16 floats or NULL: [0] 0 none / 1 box / 2 sphere, [1..3] lo or
 *             centre, [4..6] hi, [7] radius, [8..10] sigma_s, [11..13]
 *             sigma_a, [14] Henyey-Greenstein g (0 = the reference's isotropic).
 *   texels    RGB floats of every texture, row-major, row 0 = the image's
 *             first row; tex_table: per texture width, height and the index
 *             of its first float in texels.
 * Writes the per-pixel mean (and mean of squares when m2_rgb != NULL), RGB,
 * row-major, y = 0 at the top row as the reference's film indexes it. */
int or_render_pt(const float* shapes, int n_shapes, const float* mats, int n_mats,
                 const float* camera, const float* medium, const float* texels,
                 const int* tex_table, int n_tex, int width, int height, int spp, int max_depth,
                 uint64_t seed, int nthreads, double* mean_rgb, double* m2_rgb);

/* or_render_pt with the reference's sampling (Q9): (int)sqrt(spp)^2
 * samples per pixel (Scene::Rendering takes SampleCount(), scene.cpp:214,
 * and the film divides by it), their jitter drawn by Sampler::Sampling
 * (sampler.cpp:17-38) from UniformSample's fixed jittered set
 * (uniformsample.cpp:18-58), and area-light points from RectSample's
 * (rectsample.cpp:18-102, light_points requested) or SphereSample's
 * (spheresample.cpp:20-48) fixed sets, made once per frame and shuffled
 * (Sampler::ShuffleSamples).  The pixel's pass over its set restarts at
 * every pixel as in the reference; a light set is met in a random
 * arrangement and cursor at each pixel (where the frame's earlier draws
 * would have left it). */
int or_render_pt_refsampling(const float* shapes, int n_shapes, const float* mats, int n_mats,
                             const float* camera, const float* medium, const float* texels,
                             const int* tex_table, int n_tex, int width, int height, int spp,
                             int max_depth, int light_points, uint64_t seed, int nthreads,
                             double* mean_rgb, double* m2_rgb);
/* The point sets themselves: kind 0 UniformSample(n, n > 1), 1
 * RectSample(n, length, width), 2 SphereSample(n) (phi, theta); xy: 2 per
 * point (may be NULL), count: the set's size. */
int or_probe_refsets(int kind, int n, double length, double width, uint64_t seed, double* xy,
                     int* count);

/* RayTracer (src/Integrator/raytracer.cpp:30-86) over the same scene
 * records: or_render_pt's walk where every surface but specular and
 * dielectric ends the path after its light sample (Glossy's only when its
 * lobe variable, left at 1, is within kd_pdf), no medium sampling along
 * rays, trace depths 0 .. max_depth. */
int or_render_whitted(const float* shapes, int n_shapes, const float* mats, int n_mats,
                      const float* camera, const float* medium, const float* texels,
                      const int* tex_table, int n_tex, int width, int height, int spp,
                      int max_depth, uint64_t seed, int nthreads, double* mean_rgb, double* m2_rgb);

/* Derivative of the expected image of or_render_pt, contracted with
 * dloss_dimage (height x width x 3): grad7 = dLoss/d(sigma_a rgb, sigma_s
 * rgb, g) of the medium, grad7_var the variance of that estimate,
 * grad_tex (3 per texel, n_texels in the pool, may be NULL) dLoss/dtexel.
 * Forward-mode derivatives carried through the backward loop, with the
 * score of every parameter-dependent sampling decision. */
int or_adjoint_pt(const float* shapes, int n_shapes, const float* mats, int n_mats,
                  const float* camera, const float* medium, const float* texels,
                  const int* tex_table, int n_tex, int width, int height, int spp, int max_depth,
                  uint64_t seed, int nthreads, const double* dloss_dimage, double* mean_rgb,
                  double* grad7, double* grad7_var, double* grad_tex, size_t n_texels);

/* FP64 bidirectional path tracer over the same scene records (no medium):
 * eye subpaths of at most n_eye vertices (camera included), light subpaths
 * of at most n_light vertices, paths of at most max_len edges, MIS exponent
 * mis_beta.  Writes the per-pixel mean radiance. */
int or_render_bdpt(const float* shapes, int n_shapes, const float* mats, int n_mats,
                   const float* camera, const float* medium, int width, int height, int spp,
                   int n_eye, int n_light, int max_len, double mis_beta, uint64_t seed,
                   int nthreads, double* mean_rgb);

/* OBJ model loading as Model::Model(file, transform) does it
 * (src/Shape/model.cpp:72-338): 9 vertex and 9 normal floats per triangle
 * (malloc'ed; free with or_obj_free), the triangle count and the file's
 * v / vt / vn / f record counts.  transform: row-major 4x4 or NULL.
 * Returns -1 on a face whose reference behaviour is undefined. */
int or_obj_load(const char* path, const float* transform, float** xyz, float** normals,
                int* n_triangles, int counts[4]);
void or_obj_free(float* p);
/* Film::Save's pixel conversion (src/Camera/film.cpp:152-203): film is
 * width * height * 3 floats, B, G, R sums (bgr_sum = 1, divided by
 * (int)sqrt(n_samples)^2) or R, G, B means (bgr_sum = 0); out:
 * width * height * bpp bytes B, G, R (, 0). */
void or_film_encode(const float* film, int width, int height, int bgr_sum, int n_samples,
                    float gamma, int bpp, unsigned char* out);
/* ImageTexture's texel values from BMP pixel bytes (imagetexture.cpp:32-61). */
void or_bmp_texels(const unsigned char* pixels, int width, int height, int bpp, float gamma,
                   float* rgb);
/* LoadHDR (src/Kernel/hdr.cpp:4-177) over a Radiance RGBE file image:
 * rows top-down, (r, g, b) * 2^(e - 136); rgb NULL: only the size. */
int or_hdr_read(const unsigned char* data, size_t n, int* width, int* height, float* rgb);
/* SaveBmp's 54-byte header (src/Kernel/bmp.cpp:59-97). */
void or_bmp_header(int width, int height, int bpp, unsigned char* header);
/* FirefliesFiltering (src/Scene/scene.cpp:261-431) over pixel bytes. */
void or_fireflies(const unsigned char* src, int width, int height, int bpp, unsigned char* dst);
/* Inverse (src/Kernel/matrix.cpp:67-133), row-major 4x4. */
void or_mat4_inverse(const float* a, float* out);

/* ------------------------------------------------------------------ */
/* Probes of oracle_render.c's samplers and evaluators, for the tests that
 * pin them against closed forms.  n samples (or n inputs of 6 doubles in
 * `in`), 8 doubles out per sample; PCG32 stream (seed, 1).
 *   LAMBERT   par n[3]: Lambertian::BSDFSampling (bsdf.cpp:43-69) ->
 *             wi[3], cos, pdf
 *   PHONG     par e, wo[3], n[3]: Phong::BSDFSampling (bsdf.cpp:127-170)
 *             -> wi[3], cos (1), pdf, Phong::F(n, wi, wo) (bsdf.cpp:91-104)
 *   PHASE     par g, d[3]: Isotropic::PhaseSampling (phase.cpp:40-63) at
 *             g = 0, Henyey-Greenstein otherwise -> wi[3], d.wi, pdf
 *   GGX       par alpha, wo[3], n[3]; in: wi[3] or NULL (sampled) -> wi[3],
 *             n.wi, pdf, f (F0 = 1)
 *   FRESNEL   par eta_in, eta_out; in: direction[3], normal[3]:
 *             FresnelReflection::Fresnel (bsdf.cpp:255-283) -> kr
 *   TRANSMIT  par eta_in, eta_out; in: wo[3], n[3]:
 *             FresnelTransparency::BSDFSampling (bsdf.cpp:322-342) -> w[3], cos
 *   REFLECT   in: wo[3], n[3]: FresnelReflection::BSDFSampling
 *             (bsdf.cpp:232-253) -> w[3], cos
 *   DISTANCE  par sigma_t[3], avg sigma_t, absorb, smax:
 *             HomogeneousVolume::DistanceSampling (homogeneousvolume.cpp:
 *             168-215) -> scattered, s, absorbed, tau/pdf[3] */
enum { OR_PROBE_LAMBERT = 0, OR_PROBE_PHONG = 1, OR_PROBE_PHASE = 2, OR_PROBE_GGX = 3,
       OR_PROBE_FRESNEL = 4, OR_PROBE_TRANSMIT = 5, OR_PROBE_REFLECT = 6, OR_PROBE_DISTANCE = 7 };
int or_probe(int kind, const double* par, const double* in, int n, uint64_t seed, double* out);

/* Planted mistakes, for the tests that show each pin fails when the
 * restated step is wrong; 0 (none) in every ordinary use.  Global to the
 * library: set it only around single-threaded test code. */
enum {
    OR_PLANT_NONE = 0,
    OR_PLANT_ETA_SWAP = 1,      /* Fresnel with eta_in and eta_out exchanged */
    OR_PLANT_NO_ETA2 = 2,       /* transmission without FresnelTransparency::F's 1/eta^2 */
    OR_PLANT_PHONG_NORM = 3,    /* Phong normalised by (e + 2) / 2 pi */
    OR_PLANT_HG_SIGN = 4,       /* Henyey-Greenstein cosine with the wrong sign */
    OR_PLANT_GREY_WEIGHT = 5,   /* scattering weight tau / (tau avg sigma_t), not tau / (tau sigma_t) */
    OR_PLANT_NO_ROULETTE = 6,   /* no absorption roulette at a scattering event */
    OR_PLANT_NO_BEER = 7,       /* no Beer-Lambert factor inside a dielectric */
    OR_PLANT_LAMBERT_PDF = 8,   /* Lambertian pdf cos / 2 pi */
    OR_PLANT_GGX_JACOBIAN = 9,  /* GGX pdf without the half-vector Jacobian 1 / (4 wo.h) */
    OR_PLANT_SHADOW_TR = 10,    /* shadow rays through the medium always survive */
    OR_PLANT_SNELL = 11,        /* refraction with eta in place of 1 / eta */
    OR_PLANT_DENSITY = 12       /* Subsurface density ignored by the sampling sigma */
};
void or_plant(int id);

#ifdef __cplusplus
}
#endif
#endif
