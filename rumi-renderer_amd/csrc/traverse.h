// Device-side BVH traversal shared by the ray-query kernels (trace.hip) and
// the wavefront renderer's extension and shadow stages (render.hip).
//
// Stack: LDS, one column per lane (entry k of lane t at k * BLOCK + t);
// trees deeper than the LDS stack spill the excess to a private array.  A
// stack entry is (child ref, entry distance of its box): the box test runs
// when the parent is entered and the "closer than the current hit" test when
// the entry is popped, the moment BVHAccel::Travel tests a popped node
// (src/Acceleration/bvh.cpp:164-169).
//
// All FP32 expressions keep the reference's operation order; the library is
// compiled with -ffp-contract=off and IEEE division, so the box, triangle,
// sphere and rect tests produce the bits of BBox::Intersect
// (src/Kernel/bbox.cpp:30-92), Triangle::Intersect (src/Shape/triangle.cpp:
// 41-118), Sphere::Intersect (src/Shape/sphere.cpp:51-132) and
// Rect::Intersect (src/Shape/rect.cpp:51-122).
#pragma once

#include <hip/hip_runtime.h>

#include "internal.h"

namespace rumi {
namespace dev {

constexpr float kEps = 1e-4f;  // Ray::_epsilon (src/Kernel/ray.cpp:4)

// Primitive tags stored in word c.z of a leaf record (mixed scenes only).
constexpr int32_t kTagTri = 0;
constexpr int32_t kTagSphere = 1;
constexpr int32_t kTagRect = 2;

struct RootInfo {
    float lo[3];
    float hi[3];
    int32_t ref;
};

// BBox::Intersect with a precomputed reciprocal direction (1.0f / d is the
// same value whether computed per node or once per ray).
__device__ __forceinline__ bool slab(float lx, float ly, float lz, float hx, float hy, float hz,
                                     float ox, float oy, float oz, float ix, float iy, float iz,
                                     float& tn, float& tf) {
    float t0 = (lx - ox) * ix;
    float t1 = (hx - ox) * ix;
    float tmp;
    if (t0 > t1) { tmp = t0; t0 = t1; t1 = tmp; }
    tn = t0;
    tf = t1;
    t0 = (ly - oy) * iy;
    t1 = (hy - oy) * iy;
    if (t0 > t1) { tmp = t0; t0 = t1; t1 = tmp; }
    if (tn > t1 || t0 > tf) return false;
    if (tn < t0) tn = t0;
    if (tf > t1) tf = t1;
    t0 = (lz - oz) * iz;
    t1 = (hz - oz) * iz;
    if (t0 > t1) { tmp = t0; t0 = t1; t1 = tmp; }
    if (tn > t1 || t0 > tf) return false;
    if (tn < t0) tn = t0;
    if (tf > t1) tf = t1;
    return true;
}

// The tmin-independent part of the node accept test (bvh.cpp:167-168): the
// decision of slab() followed by tf >= 0 and tmax >= tn, and, for an accepted
// box, slab()'s tn up to the sign of a zero (tn only meets comparisons), but
// without branches: each axis's pair is ordered with min/max and the three
// merged with max/min.  The pair can hold a NaN (0 * inf: origin on a face's
// plane, direction component zero), where min/max and slab()'s swap-and-
// compare part ways; slab() never replaces a NaN on x (tn or tf stays NaN and
// the box is rejected), and never takes one on y or z (no comparison with it
// is true), which here is an entry of -inf or an exit of +inf.
//
// CHECK_NAN = false is the same test for a ray whose origin is finite and
// whose reciprocal direction is finite and non-zero: with finite box
// corners a difference can overflow to an infinity but never meets a zero
// factor, so no product is a NaN and the NaN handling drops out (five
// instructions per box fewer).
template <bool CHECK_NAN = true>
__device__ __forceinline__ bool box_static(float lx, float ly, float lz, float hx, float hy,
                                           float hz, float ox, float oy, float oz, float ix,
                                           float iy, float iz, float tmax, float& tn) {
    const float x0 = (lx - ox) * ix, x1 = (hx - ox) * ix;
    const float y0 = (ly - oy) * iy, y1 = (hy - oy) * iy;
    const float z0 = (lz - oz) * iz, z1 = (hz - oz) * iz;
    // maxnum/minnum return the other operand for a NaN
    const float y0n = CHECK_NAN ? fmaxf(y0, -INFINITY) : y0, y1n = CHECK_NAN ? fminf(y1, INFINITY) : y1;
    const float z0n = CHECK_NAN ? fmaxf(z0, -INFINITY) : z0, z1n = CHECK_NAN ? fminf(z1, INFINITY) : z1;
    tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0n, y1n)), fminf(z0n, z1n));
    const float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0n, y1n)), fmaxf(z0n, z1n));
    const bool x_ok = !CHECK_NAN || !__builtin_isunordered(x0, x1);
    return x_ok & (tn <= tf) & (tf >= 0.0f) & (tmax >= tn);
}

// True when no slab product of the ray can be a NaN (box_static<false>
// applies): origin finite, reciprocal direction finite and non-zero.
__device__ __forceinline__ bool finite_ray(float ox, float oy, float oz, float ix, float iy, float iz) {
    return __builtin_isfinite(ox) & __builtin_isfinite(oy) & __builtin_isfinite(oz) &
           __builtin_isfinite(ix) & __builtin_isfinite(iy) & __builtin_isfinite(iz) & (ix != 0.0f) &
           (iy != 0.0f) & (iz != 0.0f);
}

// The tmin part (bvh.cpp:169): no hit yet, or the box starts before it.
__device__ __forceinline__ bool closer(float tmin, float tn) { return tmin == 0.0f || tmin > tn; }

struct Ray {
    float ox, oy, oz, dx, dy, dz, tmax;
    float tmin;  // closest hit so far (0 = none), updated by the tests
};

// Triangle::Intersect on a leaf record (v0, e1 = v1 - v0, e2 = v2 - v0).
__device__ __forceinline__ bool hit_triangle(const float4& ta, const float4& tb, const float4& tc,
                                             Ray& r, float& u_out, float& v_out) {
    const float e1x = ta.w, e1y = tb.x, e1z = tb.y;
    const float e2x = tb.z, e2y = tb.w, e2z = tc.x;
    const float px = r.dy * e2z - e2y * r.dz;
    const float py = r.dz * e2x - e2z * r.dx;
    const float pz = r.dx * e2y - e2x * r.dy;
    const float det = px * e1x + py * e1y + pz * e1z;
    const float inv_det = 1.0f / det;
    const float tx = r.ox - ta.x, ty = r.oy - ta.y, tz = r.oz - ta.z;
    const float u = (px * tx + py * ty + pz * tz) * inv_det;
    const float qx = ty * e1z - e1y * tz;
    const float qy = tz * e1x - e1z * tx;
    const float qz = tx * e1y - e1x * ty;
    const float v = (qx * r.dx + qy * r.dy + qz * r.dz) * inv_det;
    const float t = (qx * e2x + qy * e2y + qz * e2z) * inv_det;
    // the early returns of the sequential test, as one condition: in a wave
    // some lane nearly always passes each of them, so branching saves nothing
    const bool ok = !(-1e-8f < det && det < 1e-8f) & !(u < 0.0f || u > 1.0f) &
                    !(v < 0.0f || u + v > 1.0f) & !(t < kEps) &
                    !(r.tmin > kEps && t > r.tmin) & !(r.tmax - t < kEps);
    if (ok) {
        r.tmin = t;
        u_out = u;
        v_out = v;
    }
    return ok;
}

// hit_triangle's tests except the one against the current hit (t > tmin),
// which depends on the order of the tests: the triangle's distance and
// barycentrics are returned and r is not changed.  The same arithmetic, so
// the same bits.
__device__ __forceinline__ bool tri_base(const float4& ta, const float4& tb, const float4& tc, const Ray& r,
                                         float& t_out, float& u_out, float& v_out) {
    const float e1x = ta.w, e1y = tb.x, e1z = tb.y;
    const float e2x = tb.z, e2y = tb.w, e2z = tc.x;
    const float px = r.dy * e2z - e2y * r.dz;
    const float py = r.dz * e2x - e2z * r.dx;
    const float pz = r.dx * e2y - e2x * r.dy;
    const float det = px * e1x + py * e1y + pz * e1z;
    const float inv_det = 1.0f / det;
    const float tx = r.ox - ta.x, ty = r.oy - ta.y, tz = r.oz - ta.z;
    const float u = (px * tx + py * ty + pz * tz) * inv_det;
    const float qx = ty * e1z - e1y * tz;
    const float qy = tz * e1x - e1z * tx;
    const float qz = tx * e1y - e1x * ty;
    const float v = (qx * r.dx + qy * r.dy + qz * r.dz) * inv_det;
    const float t = (qx * e2x + qy * e2y + qz * e2z) * inv_det;
    t_out = t;
    u_out = u;
    v_out = v;
    return !(-1e-8f < det && det < 1e-8f) & !(u < 0.0f || u > 1.0f) & !(v < 0.0f || u + v > 1.0f) &
           !(t < kEps) & !(r.tmax - t < kEps);
}

// Sphere::Intersect: a = centre.xyz, radius.  A = 1 (unit direction assumed),
// far root first inside an object, and swapped (again) for an emissive
// textured sphere seen from inside (ibl: the image-based light's sphere).
__device__ __forceinline__ bool hit_sphere(const float4& a, Ray& r, bool in_object, bool ibl = false) {
    const float ocx = r.ox - a.x, ocy = r.oy - a.y, ocz = r.oz - a.z;
    const float A = 1.0f;
    const float B = 2.0f * (r.dx * ocx + r.dy * ocy + r.dz * ocz);
    const float C = (ocx * ocx + ocy * ocy + ocz * ocz) - a.w * a.w;
    const float disc = B * B - 4 * A * C;
    if (disc < 0.0f) return false;
    const float inv_2a = 1.0f / (2.0f * A);
    float t0, t1;
    if (disc == 0.0f || disc == -0.0f) {
        t0 = t1 = -B * inv_2a;
    } else {
        const float s = sqrtf(disc);
        t0 = (-B - s) * inv_2a;
        t1 = (-B + s) * inv_2a;
    }
    if (t0 < 0.0f && t1 < 0.0f) return false;
    if (in_object) {
        const float tmp = t0;
        t0 = t1;
        t1 = tmp;
    }
    if (ibl && sqrtf(ocx * ocx + ocy * ocy + ocz * ocz) < a.w) {
        const float tmp = t0;
        t0 = t1;
        t1 = tmp;
    }
    if (t0 < kEps) return false;
    if (r.tmax - t0 < kEps) return false;
    if (t0 > r.tmin && r.tmin > 0.0f) return false;
    r.tmin = t0;
    return true;
}

// Rect record in the side table: p.xyz n.x | n.yz l.xy | l.z w.xyz |
// (inv_area * width) (inv_area * length) length width.
__device__ __forceinline__ bool hit_rect(const float4* rect, bool emissive, Ray& r) {
    const float4 q0 = rect[0], q1 = rect[1], q2 = rect[2], q3 = rect[3];
    const float nx = q0.w, ny = q1.x, nz = q1.y;
    const float cost = r.dx * nx + r.dy * ny + r.dz * nz;
    if (cost == 0.0f || cost == -0.0f) return false;
    const float opx = r.ox - q0.x, opy = r.oy - q0.y, opz = r.oz - q0.z;
    const float t = -(opx * nx + opy * ny + opz * nz) / cost;
    if (t < 0.0f) return false;
    if (t < kEps) return false;
    if (r.tmax - t < kEps) return false;
    if (!emissive && (0.0f < cost && cost < 1.0f)) return false;
    if (t > r.tmin && r.tmin > 0.0f) return false;
    const float ipx = r.ox + r.dx * t, ipy = r.oy + r.dy * t, ipz = r.oz + r.dz * t;
    const float ux = ipx - q0.x, uy = ipy - q0.y, uz = ipz - q0.z;
    const float il = (ux * q1.z + uy * q1.w + uz * q2.x) * q3.x;
    if (il < 0.0f || il > q3.z) return false;
    const float iw = (ux * q2.y + uy * q2.z + uz * q2.w) * q3.y;
    if (iw < 0.0f || iw > q3.w) return false;
    r.tmin = t;
    return true;
}

// Traversal state of one ray between steps: the ray (r.tmin = closest hit
// so far), its reciprocal direction, the node or leaf to visit next, the
// stack depth and the hit.
struct TravState {
    Ray r;
    float ix, iy, iz;
    int32_t cur;
    int32_t sp;
    int32_t hit;
    float hu, hv;
    int32_t leaf;  // near-first order: the leaf whose primitives are being tested (0 = none)
};

// "No entry" in TravState::cur (near-first order).  A leaf reference equals
// it only for a leaf at primitive 2^27 - 16 holding 15 primitives, beyond
// the 2^27 primitives a leaf reference can address.
constexpr int32_t kNoRef = -2147483647 - 1;

// Starts a traversal at the root: false when the ray has a zero direction or
// misses the root box (BVHAccel::Travel's first node test, bvh.cpp:164-169).
__device__ __forceinline__ bool trav_begin(const RootInfo& root, TravState& s) {
    s.hit = -1;
    s.sp = 0;
    s.cur = root.ref;
    s.leaf = 0;
    s.hu = 0.0f;
    s.hv = 0.0f;
    if (s.r.dx == 0.0f && s.r.dy == 0.0f && s.r.dz == 0.0f) return false;
    s.ix = 1.0f / s.r.dx;
    s.iy = 1.0f / s.r.dy;
    s.iz = 1.0f / s.r.dz;
    float tn;
    return box_static(root.lo[0], root.lo[1], root.lo[2], root.hi[0], root.hi[1], root.hi[2],
                      s.r.ox, s.r.oy, s.r.oz, s.ix, s.iy, s.iz, s.r.tmax, tn) &&
           closer(s.r.tmin, tn);
}

// Stack push/pop of (reference, entry distance): LDS for the first STACK
// entries, private memory beyond (SPILL only).
template <int BLOCK, int STACK, bool SPILL>
__device__ __forceinline__ void stack_push(TravState& s, int32_t ref, float tn, int2* s_stk,
                                           uint32_t tid, int2* p_stk) {
    const int2 e = make_int2(ref, __float_as_int(tn));
    if (!SPILL || s.sp < STACK) {
        s_stk[s.sp * BLOCK + tid] = e;
    } else {
        p_stk[s.sp - STACK] = e;
        // keeps this private-memory access apart from the LDS one: merged,
        // they become generic (flat) accesses on every push
        __asm__ volatile("" ::: "memory");
    }
    ++s.sp;
}

template <int BLOCK, int STACK, bool SPILL>
__device__ __forceinline__ void stack_pop(TravState& s, int32_t& ref, float& tn, const int2* s_stk,
                                          uint32_t tid, const int2* p_stk) {
    --s.sp;
    int2 e;
    if (!SPILL || s.sp < STACK) {
        e = s_stk[s.sp * BLOCK + tid];
        // one 8-byte read: without this the reference's read is sunk below
        // the distance test, a second LDS round trip
        __asm__ volatile("" : "+v"(e.x), "+v"(e.y));
    } else {
        e = p_stk[s.sp - STACK];
        __asm__ volatile("" : "+v"(e.x), "+v"(e.y));  // as in stack_push
    }
    ref = e.x;
    tn = __int_as_float(e.y);
}

// Picks the next node to visit: a leaf in s.cur moves to the free leaf slot
// (an empty leaf is dropped), and while s.cur is empty stack entries are
// popped, those farther than the closest hit so far dropped.  A leaf found
// while the leaf slot is taken waits in s.cur.
template <int BLOCK, int STACK, bool SPILL>
__device__ __forceinline__ void next_node(TravState& s, const int2* s_stk, uint32_t tid,
                                          const int2* p_stk) {
    while (true) {
        const int32_t c = s.cur;
        const bool is_leaf = c < 0 && c != kNoRef;
        const bool empty = (~c & 15) == 0;
        const bool take = is_leaf && !empty && s.leaf == 0;
        s.leaf = take ? c : s.leaf;
        s.cur = is_leaf && (empty || take) ? kNoRef : c;
        if (s.cur != kNoRef || s.sp == 0) break;
        int32_t rr;
        float t;
        stack_pop<BLOCK, STACK, SPILL>(s, rr, t, s_stk, tid, p_stk);
        s.cur = closer(s.r.tmin, t) ? rr : kNoRef;
    }
}

// Near-first traversal with postponed leaves.  A lane holds one node to
// visit (s.cur) and one leaf being tested (s.leaf), and each step advances
// both: the next primitive of s.leaf is tested first (it may lower
// r.tmin), then node s.cur's children are tested against r.tmin, the nearer
// accepted child becoming s.cur and the farther pushed.  A leaf reached
// while s.leaf is free moves there at once and traversal goes on with the
// next stack entry; otherwise it waits in s.cur.  Node and primitive fetches
// of a step are independent, so a wave waits on one fetch round trip per
// step while its lanes do node and primitive work together.  The order
// depends only on the ray (not on other lanes), so hits and work counts are
// reproducible; the closest hit distance is that of every other order, and
// only which of two primitives at exactly equal distance is reported can
// differ from them.  Any-hit queries end after the leaf that found a hit.
template <int BLOCK, int STACK, bool ANY, bool MIXED, bool COUNT, bool SPILL>
__device__ __forceinline__ bool trav_step_near(const float4* __restrict__ nodes,
                                               const float4* __restrict__ prims,
                                               const float4* __restrict__ rects, TravState& s,
                                               bool in_object, int32_t& nn, int32_t& np,
                                               int2* s_stk, uint32_t tid, int2* p_stk) {
    Ray& r = s.r;
    const bool at_node = s.cur >= 0;
    const bool at_leaf = s.leaf != 0;
    float4 a, b, c, d, ta, tb, tc;
    if (at_node) {
        const float4* np4 = nodes + 4 * (size_t)s.cur;
        a = np4[0];
        b = np4[1];
        c = np4[2];
        d = np4[3];
    }
    const int32_t enc = ~s.leaf;
    if (at_leaf) {
        const float4* t4 = prims + 3 * (size_t)(enc >> 4);
        ta = t4[0];
        tb = t4[1];
        tc = t4[2];
    }
    if (at_leaf) {
        if (COUNT) ++np;
        float u = 0.0f, v = 0.0f;
        bool ok;
        if (!MIXED) {
            ok = hit_triangle(ta, tb, tc, r, u, v);
        } else {
            const int32_t tag = __float_as_int(tc.z);
            if (tag == kTagTri)
                ok = hit_triangle(ta, tb, tc, r, u, v);
            else if (tag == kTagSphere)
                ok = hit_sphere(ta, r, in_object, __float_as_int(tc.w) != 0);
            else
                ok = hit_rect(rects + 4 * (size_t)__float_as_int(tc.x), __float_as_int(tc.w) != 0, r);
        }
        if (ok) {
            s.hit = __float_as_int(tc.y);
            s.hu = u;
            s.hv = v;
        }
        const int32_t count = enc & 15;
        if (count > 1) {
            s.leaf = ~((((enc >> 4) + 1) << 4) | (count - 1));
        } else {
            s.leaf = 0;
            if (ANY && r.tmin > 0.0f) return false;
        }
    }
    if (at_node) {
        if (COUNT) ++nn;
        const int32_t ref0 = __float_as_int(d.x);
        const int32_t ref1 = __float_as_int(d.y);
        float tn0, tn1;
        const bool a0 = box_static(a.x, a.y, a.z, a.w, b.x, b.y, r.ox, r.oy, r.oz, s.ix, s.iy, s.iz,
                                   r.tmax, tn0) && closer(r.tmin, tn0);
        const bool a1 = box_static(b.z, b.w, c.x, c.y, c.z, c.w, r.ox, r.oy, r.oz, s.ix, s.iy, s.iz,
                                   r.tmax, tn1) && closer(r.tmin, tn1);
        if (a0 && a1) {
            const bool swap = tn1 < tn0;
            s.cur = swap ? ref1 : ref0;
            stack_push<BLOCK, STACK, SPILL>(s, swap ? ref0 : ref1, swap ? tn0 : tn1, s_stk, tid, p_stk);
        } else {
            s.cur = a0 ? ref0 : (a1 ? ref1 : kNoRef);
        }
    }
    next_node<BLOCK, STACK, SPILL>(s, s_stk, tid, p_stk);
    return s.cur != kNoRef || s.leaf != 0;
}

// Order keys of a 4-wide node's accepted children: the entry distance's bits
// mapped to an unsigned integer of the same order (a zero of either sign as
// +0), the two low bits replaced by the child's slot, so keys are distinct
// and sort by distance, ties and near-ties (within 4 units in the last place)
// by slot.  key_distance() reads back a distance no larger than the entry
// distance (the two bits cleared), which the stack keeps for the test on pop.
constexpr uint32_t kNoKey = 0xffffffffu;

__device__ __forceinline__ uint32_t order_key(float tn, int slot) {
    const int32_t b = __float_as_int(tn + 0.0f);
    const uint32_t u = b < 0 ? ~(uint32_t)b : (uint32_t)b | 0x80000000u;
    return (u & ~3u) | (uint32_t)slot;
}

// Set bits of mask in the lanes below this one.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ float key_distance(uint32_t key) {
    const uint32_t u = key & ~3u;
    return __int_as_float((int32_t)((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u));
}

// Near-first order over 4-wide nodes (DevNode4, 8 float4 per node): the
// accepted children are ordered by order_key, the first becomes s.cur and the
// others are pushed last first, each with its key_distance.  Half the node
// steps of the binary tree, four boxes per fetch; up to four triangles of
// the postponed leaf per step (meshes; two in mixed scenes).
//
// RD (meshes only): every lane of the wave calls the step, `live` telling
// whether it holds a ray, and the step's triangle tests -- up to four per
// lane, tested in turn on that lane otherwise -- are spread over the wave's
// lanes, 64 per pass (slots: K * 64 entries (triangle, owner lane) of LDS
// per wave); each lane then takes its tests' outcomes back in their order
// and applies the test against its current hit, so hits, distances and
// counts are those of the per-lane steps.
template <int BLOCK, int STACK, bool ANY, bool MIXED, bool COUNT, bool SPILL, bool RD = false>
__device__ __forceinline__ bool trav_step_near4(const float4* __restrict__ nodes,
                                                const float4* __restrict__ prims,
                                                const float4* __restrict__ rects, TravState& s,
                                                bool in_object, int32_t& nn, int32_t& np,
                                                int2* s_stk, uint32_t tid, int2* p_stk, bool live = true,
                                                uint8_t* slots = nullptr) {
    static_assert(!(RD && MIXED), "spread triangle tests are for meshes");
    Ray& r = s.r;
    const bool at_node = (!RD || live) && s.cur >= 0;
    const bool at_leaf = (!RD || live) && s.leaf != 0;
    float4 lx, ly, lz, hx, hy, hz, rf, ta, tb, tc;
    if (at_node) {
        const float4* np4 = nodes + 8 * (size_t)s.cur;
        lx = np4[0];
        ly = np4[1];
        lz = np4[2];
        hx = np4[3];
        hy = np4[4];
        hz = np4[5];
        rf = np4[6];
    }
    const int32_t enc = ~s.leaf;
    // A leaf's next K primitive records in one step (the wave waits on one
    // fetch round trip for all of them, and tmin drops sooner for the node
    // tests of the same step): meshes up to four triangles (spread over the
    // wave: 1.92 -> 1.87 ms closest-hit on the 1.3M-triangle mesh, against
    // three), mixed scenes two when both are triangles, other primitives
    // one per step.
    constexpr int K = MIXED ? 2 : 4;
    const int32_t rem = at_leaf ? (enc & 15) : 0;
    int use = rem < K ? rem : K;
    if constexpr (RD) {
        const uint32_t lane = tid & 63u;
        const uint64_t b0 = __ballot((use & 1) != 0), b1 = __ballot((use & 2) != 0),
                       b2 = K > 3 ? __ballot((use & 4) != 0) : 0ull;
        const uint32_t total =
            (uint32_t)__popcll(b0) + 2u * (uint32_t)__popcll(b1) + 4u * (uint32_t)__popcll(b2);
        if (total != 0) {
            // test k of lane L is the wave's test pre(L) + k
            const uint32_t pre = lanes_below(b0) + 2u * lanes_below(b1) + 4u * lanes_below(b2);
            int2* sl = reinterpret_cast<int2*>(slots);
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (k < use) sl[pre + k] = make_int2((enc >> 4) + k, (int32_t)lane);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (uint32_t r0 = 0; r0 < total; r0 += 64u) {
                const uint32_t p = r0 + lane;
                const bool tester = p < total;
                const int2 e = tester ? sl[p] : make_int2(0, (int32_t)lane);
                // the owner's ray (every lane takes part)
                const int oa = e.y << 2;
                Ray q;
                q.ox = __int_as_float(__builtin_amdgcn_ds_bpermute(oa, __float_as_int(r.ox)));
                q.oy = __int_as_float(__builtin_amdgcn_ds_bpermute(oa, __float_as_int(r.oy)));
                q.oz = __int_as_float(__builtin_amdgcn_ds_bpermute(oa, __float_as_int(r.oz)));
                q.dx = __int_as_float(__builtin_amdgcn_ds_bpermute(oa, __float_as_int(r.dx)));
                q.dy = __int_as_float(__builtin_amdgcn_ds_bpermute(oa, __float_as_int(r.dy)));
                q.dz = __int_as_float(__builtin_amdgcn_ds_bpermute(oa, __float_as_int(r.dz)));
                q.tmax = __int_as_float(__builtin_amdgcn_ds_bpermute(oa, __float_as_int(r.tmax)));
                q.tmin = 0.0f;  // not read: tri_base leaves the test against the hit to the owner
                int32_t rid = -1;
                float rt = 0.0f, ru = 0.0f, rv = 0.0f;
                if (tester) {
                    const float4* t4 = prims + 3 * (size_t)e.x;
                    const float4 a = t4[0], b = t4[1], c = t4[2];
                    float t, u, v;
                    if (tri_base(a, b, c, q, t, u, v)) {
                        rid = __float_as_int(c.y);
                        rt = t;
                        ru = u;
                        rv = v;
                    }
                }
                // each lane takes back its tests of this pass, in order
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const uint32_t pos = pre + (uint32_t)k;
                    const bool mine = k < use && pos >= r0 && pos < r0 + 64u;
                    const int sa = (int)((mine ? pos - r0 : lane) << 2);
                    const int32_t id = __builtin_amdgcn_ds_bpermute(sa, rid);
                    const float t = __int_as_float(__builtin_amdgcn_ds_bpermute(sa, __float_as_int(rt)));
                    const float u = __int_as_float(__builtin_amdgcn_ds_bpermute(sa, __float_as_int(ru)));
                    const float v = __int_as_float(__builtin_amdgcn_ds_bpermute(sa, __float_as_int(rv)));
                    if (mine && id >= 0 && !(r.tmin > kEps && t > r.tmin)) {
                        r.tmin = t;
                        s.hit = id;
                        s.hu = u;
                        s.hv = v;
                    }
                }
            }
            // the next step's slot writes stay after this step's reads
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
        if (at_leaf) {
            if (COUNT) np += use;
            const int32_t count = enc & 15;
            if (count > use) {
                s.leaf = ~((((enc >> 4) + use) << 4) | (count - use));
            } else {
                s.leaf = 0;
                if (ANY && r.tmin > 0.0f) return false;
            }
        }
    }
    float4 tt[K][3];
    if (!RD && at_leaf) {
        const float4* t4 = prims + 3 * (size_t)(enc >> 4);
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (k < use) {
                tt[k][0] = t4[3 * k + 0];
                tt[k][1] = t4[3 * k + 1];
                tt[k][2] = t4[3 * k + 2];
            }
    }
    if (MIXED && use > 1 && !(__float_as_int(tt[0][2].z) == kTagTri && __float_as_int(tt[1][2].z) == kTagTri))
        use = 1;
    if (!RD && at_leaf) {
        ta = tt[0][0];
        tb = tt[0][1];
        tc = tt[0][2];
        if (COUNT) ++np;
        float u = 0.0f, v = 0.0f;
        bool ok;
        if (!MIXED) {
            ok = hit_triangle(ta, tb, tc, r, u, v);
        } else {
            const int32_t tag = __float_as_int(tc.z);
            if (tag == kTagTri)
                ok = hit_triangle(ta, tb, tc, r, u, v);
            else if (tag == kTagSphere)
                ok = hit_sphere(ta, r, in_object, __float_as_int(tc.w) != 0);
            else
                ok = hit_rect(rects + 4 * (size_t)__float_as_int(tc.x), __float_as_int(tc.w) != 0, r);
        }
        if (ok) {
            s.hit = __float_as_int(tc.y);
            s.hu = u;
            s.hv = v;
        }
#pragma unroll
        for (int k = 1; k < K; ++k)
            if (k < use) {
                if (COUNT) ++np;
                float u2 = 0.0f, v2 = 0.0f;
                if (hit_triangle(tt[k][0], tt[k][1], tt[k][2], r, u2, v2)) {
                    s.hit = __float_as_int(tt[k][2].y);
                    s.hu = u2;
                    s.hv = v2;
                }
            }
        const int32_t count = enc & 15;
        if (count > use) {
            s.leaf = ~((((enc >> 4) + use) << 4) | (count - use));
        } else {
            s.leaf = 0;
            if (ANY && r.tmin > 0.0f) return false;
        }
    }
    if (at_node) {
        if (COUNT) ++nn;
        int32_t ref[4] = {__float_as_int(rf.x), __float_as_int(rf.y), __float_as_int(rf.z),
                          __float_as_int(rf.w)};
        const float blx[4] = {lx.x, lx.y, lx.z, lx.w}, bly[4] = {ly.x, ly.y, ly.z, ly.w},
                    blz[4] = {lz.x, lz.y, lz.z, lz.w}, bhx[4] = {hx.x, hx.y, hx.z, hx.w},
                    bhy[4] = {hy.x, hy.y, hy.z, hy.w}, bhz[4] = {hz.x, hz.y, hz.z, hz.w};
        uint32_t key[4];
        // the NaN-free test unless some lane of the wave needs the other
        // (a wave-uniform choice)
        if (__ballot(!finite_ray(r.ox, r.oy, r.oz, s.ix, s.iy, s.iz)) == 0) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                float tn;
                const int acc = (int)(ref[c] != kNoRef) &
                                (int)box_static<false>(blx[c], bly[c], blz[c], bhx[c], bhy[c], bhz[c], r.ox,
                                                       r.oy, r.oz, s.ix, s.iy, s.iz, r.tmax, tn) &
                                (int)closer(r.tmin, tn);
                key[c] = acc ? order_key(tn, c) : kNoKey;
            }
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                float tn;
                const int acc = (int)(ref[c] != kNoRef) &
                                (int)box_static(blx[c], bly[c], blz[c], bhx[c], bhy[c], bhz[c], r.ox, r.oy,
                                                r.oz, s.ix, s.iy, s.iz, r.tmax, tn) &
                                (int)closer(r.tmin, tn);
                key[c] = acc ? order_key(tn, c) : kNoKey;
            }
        }
        // sorting network over (key, ref): ascending keys, rejected last
        auto cmpx = [&](int a, int b) {
            const bool sw = key[b] < key[a];
            const uint32_t ka = key[a], kb = key[b];
            key[a] = sw ? kb : ka;
            key[b] = sw ? ka : kb;
            const int32_t ra = ref[a], rb = ref[b];
            ref[a] = sw ? rb : ra;
            ref[b] = sw ? ra : rb;
        };
        cmpx(0, 1);
        cmpx(2, 3);
        cmpx(0, 2);
        cmpx(1, 3);
        cmpx(1, 2);
        s.cur = key[0] != kNoKey ? ref[0] : kNoRef;
        if (!SPILL || __all(s.sp + 3 <= STACK)) {
            // no lane of the wave can reach private memory: LDS pushes only
#pragma unroll
            for (int c = 3; c >= 1; --c)
                if (key[c] != kNoKey)
                    stack_push<BLOCK, STACK, false>(s, ref[c], key_distance(key[c]), s_stk, tid, p_stk);
        } else {
#pragma unroll
            for (int c = 3; c >= 1; --c)
                if (key[c] != kNoKey)
                    stack_push<BLOCK, STACK, SPILL>(s, ref[c], key_distance(key[c]), s_stk, tid, p_stk);
        }
    }
    if (RD && !live) return false;
    next_node<BLOCK, STACK, SPILL>(s, s_stk, tid, p_stk);
    return s.cur != kNoRef || s.leaf != 0;
}

// One traversal step: visits node s.cur or the next primitive of leaf s.cur;
// after a leaf's last primitive, or a node with no child to enter, pops
// entries until one is closer than the current hit.
// Returns false when the traversal is over (stack empty, or an any-hit query
// found its hit).  MIXED enables sphere/rect leaf records (tag in c.z, rect
// side table index in c.x, emissive flag in c.w).  Stack entries beyond
// STACK live in p_stk (SPILL only).
template <int BLOCK, int STACK, bool ANY, bool ORDERED, bool MIXED, bool COUNT, bool SPILL,
          bool WIDE = false>
__device__ __forceinline__ bool trav_step(const float4* __restrict__ nodes,
                                          const float4* __restrict__ prims,
                                          const float4* __restrict__ rects, TravState& s,
                                          bool in_object, int32_t& nn, int32_t& np,
                                          int2* s_stk, uint32_t tid, int2* p_stk) {
    if (ORDERED && WIDE)
        return trav_step_near4<BLOCK, STACK, ANY, MIXED, COUNT, SPILL>(
            nodes, prims, rects, s, in_object, nn, np, s_stk, tid, p_stk);
    if (ORDERED)
        return trav_step_near<BLOCK, STACK, ANY, MIXED, COUNT, SPILL>(
            nodes, prims, rects, s, in_object, nn, np, s_stk, tid, p_stk);
    Ray& r = s.r;
    if (s.cur >= 0) {
        const float4* np4 = nodes + 4 * (size_t)s.cur;
        const float4 a = np4[0];
        const float4 b = np4[1];
        const float4 c = np4[2];
        const float4 d = np4[3];
        if (COUNT) ++nn;
        const int32_t ref0 = __float_as_int(d.x);
        const int32_t ref1 = __float_as_int(d.y);
        float tn0, tn1;
        const bool h0 = box_static(a.x, a.y, a.z, a.w, b.x, b.y, r.ox, r.oy, r.oz, s.ix, s.iy, s.iz,
                                   r.tmax, tn0);
        const bool h1 = box_static(b.z, b.w, c.x, c.y, c.z, c.w, r.ox, r.oy, r.oz, s.ix, s.iy, s.iz,
                                   r.tmax, tn1);
        const bool a0 = h0 && closer(r.tmin, tn0);
        const bool a1 = h1 && closer(r.tmin, tn1);
        int32_t push_ref = 0;
        float push_tn = 0.0f;
        bool do_push = false;
        bool descend = true;
        if (ORDERED) {
            if (a0 && a1) {
                const bool swap = tn1 < tn0;
                s.cur = swap ? ref1 : ref0;
                push_ref = swap ? ref0 : ref1;
                push_tn = swap ? tn0 : tn1;
                do_push = true;
            } else if (a0) {
                s.cur = ref0;
            } else if (a1) {
                s.cur = ref1;
            } else {
                descend = false;
            }
        } else {
            // Reference order: the right child waits on the stack, the left
            // child is visited next.
            do_push = a1;
            push_ref = ref1;
            push_tn = tn1;
            if (a0)
                s.cur = ref0;
            else
                descend = false;
        }
        if (do_push) {
            stack_push<BLOCK, STACK, SPILL>(s, push_ref, push_tn, s_stk, tid, p_stk);
        }
        if (descend) return true;
    } else {
        // One primitive per step: a wave then waits on one fetch per step
        // whether its lanes are at nodes or in leaves (a whole leaf per step
        // would cost its count of fetch round trips for the entire wave).
        // The leaf's remaining primitives stay encoded in s.cur, so the
        // primitives are tested in the same order as a loop over the leaf.
        const int32_t enc = ~s.cur;
        const int32_t first = enc >> 4;
        const int32_t count = enc & 15;
        if (count > 0) {
            const float4* t4 = prims + 3 * (size_t)first;
            const float4 ta = t4[0];
            const float4 tb = t4[1];
            const float4 tc = t4[2];
            if (COUNT) ++np;
            float u = 0.0f, v = 0.0f;
            bool ok;
            if (!MIXED) {
                ok = hit_triangle(ta, tb, tc, r, u, v);
            } else {
                const int32_t tag = __float_as_int(tc.z);
                if (tag == kTagTri)
                    ok = hit_triangle(ta, tb, tc, r, u, v);
                else if (tag == kTagSphere)
                    ok = hit_sphere(ta, r, in_object, __float_as_int(tc.w) != 0);
                else
                    ok = hit_rect(rects + 4 * (size_t)__float_as_int(tc.x), __float_as_int(tc.w) != 0,
                                  r);
            }
            if (ok) {
                s.hit = __float_as_int(tc.y);
                s.hu = u;
                s.hv = v;
            }
            if (count > 1) {
                s.cur = ~(((first + 1) << 4) | (count - 1));
                return true;
            }
        }
        // an any-hit query ends after the leaf that found a hit (bvh.cpp:181-182)
        if (ANY && r.tmin > 0.0f) return false;
    }
    while (s.sp > 0) {
        int32_t rr;
        float t;
        stack_pop<BLOCK, STACK, SPILL>(s, rr, t, s_stk, tid, p_stk);
        if (closer(r.tmin, t)) {
            s.cur = rr;
            return true;
        }
    }
    return false;
}

// One whole traversal.  Returns the hit primitive's input index or -1;
// r.tmin holds the hit distance.
template <int BLOCK, int STACK, bool ANY, bool ORDERED, bool MIXED, bool COUNT, bool SPILL,
          bool WIDE = false>
__device__ __forceinline__ int32_t traverse(const float4* __restrict__ nodes,
                                            const float4* __restrict__ prims,
                                            const float4* __restrict__ rects, const RootInfo& root,
                                            Ray& r, bool in_object, float& hu, float& hv,
                                            int32_t& nn, int32_t& np, int2* s_stk,
                                            uint32_t tid) {
    constexpr int kSpill = SPILL ? kMaxDepth - STACK : 1;
    int2 p_stk[kSpill];
    TravState s;
    s.r = r;
    if (trav_begin(root, s))
        while (trav_step<BLOCK, STACK, ANY, ORDERED, MIXED, COUNT, SPILL, WIDE>(
            nodes, prims, rects, s, in_object, nn, np, s_stk, tid, p_stk)) {
        }
    r = s.r;
    hu = s.hu;
    hv = s.hv;
    return s.hit;
}

// Rays a wave takes from the work counter per atomic (before the tail).
#ifndef RUMI_RUN_CHUNK
#define RUMI_RUN_CHUNK 256
#endif
constexpr uint32_t kRunChunk = RUMI_RUN_CHUNK;
// Rays per atomic while more than 16 rays per lane of the grid remain:
// queues many times the grid's width (the renderer's connection queues hold
// up to 2^25 rays) would otherwise wait on the counter's same-address
// atomics, which serialise.
constexpr uint32_t kRunChunkBig = 1024;
// Lanes with a fetched ray that start together (see run_lanes).
#ifndef RUMI_START_BATCH
#define RUMI_START_BATCH 16
#endif
constexpr uint32_t kStartBatch = RUMI_START_BATCH;

// Persistent-lane driver over a queue of n rays.  Rays differ in cost by two
// orders of magnitude, so one ray per lane leaves a wave waiting on its
// slowest lane.  Here every lane steps its own ray one node or primitive at
// a time; a lane whose ray ends swaps in the ray it has already prefetched
// and, in the same iteration, issues the loads of its next one, which
// complete while the wave waits on its node fetches anyway.  Ray indices come
// from *counter (zero before the launch) in wave chunks of kRunChunk (64 in
// the tail), the next requested while the current one still holds up to
// kRunChunk.  Each ray's arithmetic is the one-ray-per-lane arithmetic, so
// results are bit-identical.
//
// Item (per lane, all inlined):
//   void prefetch(uint32_t k)      issue the loads of ray k
//   bool start(TravState& s)       take the prefetched ray: fill s.r, select
//                                  the BVH, return trav_begin(...)
//   bool step(TravState& s)        one trav_step on the selected BVH
//   void finish(const TravState&)  consume the result of the current ray
template <class Item>
__device__ __forceinline__ void run_lanes(uint32_t n, uint32_t* counter, Item& item) {
    const uint32_t lane = threadIdx.x & 63u;
    // Every wave draws from one counter, whose atomics serialise; large
    // chunks keep them rare, and chunks shrink to a wave's width once fewer
    // than two rays per lane of the grid remain, to balance the tail.
    const uint32_t tail = gridDim.x * blockDim.x * 2u;
    const uint64_t big_tail = (uint64_t)gridDim.x * blockDim.x * 16u;
    TravState s = {};
    bool active = false, has_next = false, drained = false, unfinished = false;
    uint32_t cbase = 0, cleft = 0, pend_val = 0, pend_sz = 0;  // wave-uniform chunk state
    bool pend = false;
    while (true) {
        // Ending a ray and starting the next one run for few lanes at a time;
        // they wait until kStartBatch lanes are ready (or nothing else is
        // left to do), so the wave pays for those instructions less often.
        const uint64_t ready = __ballot(!active && has_next);
        if (ready != 0 && ((uint32_t)__popcll(ready) >= kStartBatch || __ballot(active) == 0 ||
                           __ballot(drained) != 0)) {
            if (!active) {
                if (unfinished) {
                    item.finish(s);
                    unfinished = false;
                }
                if (has_next) {
                    has_next = false;
                    active = item.start(s);
                    unfinished = !active;
                }
            }
        }
        if (!pend && cleft < kRunChunk) {
            pend_sz = (uint64_t)cbase + big_tail < n ? kRunChunkBig : (cbase + tail < n ? kRunChunk : 64u);
            if (lane == 0) pend_val = atomicAdd(counter, pend_sz);
            pend = true;
        }
        const uint64_t want = __ballot(!has_next && !drained);
        if (want != 0) {
            if (cleft == 0) {
                cbase = __shfl(pend_val, 0);
                cleft = pend_sz;
                pend = false;
            }
            const uint32_t need = (uint32_t)__popcll(want);
            const uint32_t take = need < cleft ? need : cleft;
            const uint32_t rank = lanes_below(want);
            if (!has_next && !drained && rank < take) {
                const uint32_t k = cbase + rank;
                if (k >= n) {
                    drained = true;
                } else {
                    item.prefetch(k);
                    has_next = true;
                }
            }
            cbase += take;
            cleft -= take;
        }
        if constexpr (Item::kWaveStep) {
            // the whole wave steps; lanes without a ray only lend their ALUs
            const bool a = item.step(s, active);
            if (active) {
                active = a;
                unfinished = !a;
            }
        } else if (active) {
            active = item.step(s);
            unfinished = !active;
        }
        if (__ballot(active || has_next || !drained) == 0) break;
    }
    if (unfinished) item.finish(s);
}

}  // namespace dev
}  // namespace rumi
