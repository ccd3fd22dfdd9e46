// Internal declarations shared by the host code and the HIP kernels of
// librumi.  Nothing here is part of the C ABI (include/rumi.h).
#pragma once

#include <atomic>

#include <cstdint>
#include <string>
#include <vector>

#include "rumi.h"

namespace rumi {

// ---------------------------------------------------------------- errors
// Records the message returned by rumi_last_error() on this thread and
// returns the code, so call sites read `return fail(RUMI_ERR_INVALID, ...)`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// ------------------------------------------------------------ PCG32
// O'Neill's PCG32 (XSH-RR 64/32): state advance by the 64-bit LCG with
// multiplier 6364136223846793005, odd increment derived from the stream.
struct Pcg32 {
    uint64_t state = 0;
    uint64_t inc = 1;
    Pcg32(uint64_t initstate, uint64_t initseq) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
    }
    // Uniform in [0, 1) with 24 random bits.
    float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

// ------------------------------------------------------------ device layout
// Interior node of the GPU BVH, 64 B: both child boxes plus child refs.
//   a = c0.lo.x c0.lo.y c0.lo.z c0.hi.x
//   b = c0.hi.y c0.hi.z c1.lo.x c1.lo.y
//   c = c1.lo.z c1.hi.x c1.hi.y c1.hi.z
//   d = ref0 ref1 (unused) (unused)
// A ref >= 0 is an interior node index; a ref < 0 is a leaf encoded as
// ~((first << 4) | count) over the leaf-ordered triangle array.
struct alignas(16) DevNode {
    float a[4], b[4], c[4];
    int32_t d[4];
};
static_assert(sizeof(DevNode) == 64, "DevNode must be 64 bytes");

// Interior node of the 4-wide GPU BVH, 128 B: up to four child boxes as one
// float4 per bound and axis (lo.x of the four children, then lo.y, ...), the
// four child refs (as in DevNode; kNoChild marks an unused slot) and a pad.
// Built by collapsing the binary tree: a node takes its two children and
// keeps replacing the interior child of largest surface area by that child's
// two children until it holds four.
struct alignas(16) DevNode4 {
    float lo[3][4];
    float hi[3][4];
    int32_t ref[4];
    int32_t pad[4];
};
static_assert(sizeof(DevNode4) == 128, "DevNode4 must be 128 bytes");
constexpr int32_t kNoChild = -2147483647 - 1;

// Leaf-ordered triangle, 48 B: v0, e1 = v1 - v0, e2 = v2 - v0 (the same FP32
// subtractions Triangle::Intersect performs per test), input index.
//   a = v0.x v0.y v0.z e1.x ; b = e1.y e1.z e2.x e2.y ; c = e2.z id 0 0
struct alignas(16) DevTri {
    float a[4], b[4], c[4];
};
static_assert(sizeof(DevTri) == 48, "DevTri must be 48 bytes");

constexpr int kMaxLeafCount = 15;        // 4-bit count in a leaf ref
constexpr int kMaxLeafFirst = (1 << 27) - 1;
constexpr int kMaxDepth = 96;            // traversal stack bound (LDS + private spill)

inline int32_t leaf_ref(int first, int count) { return ~((first << 4) | count); }

// ------------------------------------------------------------ host BVH
// Build-time node: own box, children or leaf range over `order`.
struct BuildNode {
    float lo[3], hi[3];
    int32_t left = -1, right = -1;  // < 0 => leaf
    int32_t first = 0, count = 0;
};

struct BuildResult {
    std::vector<BuildNode> nodes;   // nodes[root] is the root
    std::vector<int32_t> order;     // leaf slot -> primitive index
    int32_t root = -1;
    double sah_cost = 0.0;
};

// Binned SAH builder with FP64 cost sums (prim boxes as lo/hi float3).
void build_sah(const float* lo, const float* hi, int n, int max_leaf, int n_bins,
               BuildResult* out);
// The reference's median builder (src/Acceleration/bvh.cpp:48-148).
void build_median6(const float* lo, const float* hi, int n, BuildResult* out);

// Flattened GPU layout shared by host export and device upload.
struct FlatBvh {
    std::vector<DevNode> nodes;
    std::vector<int32_t> leaf_prims;  // leaf slot -> primitive index
    float root_lo[3] = {0, 0, 0}, root_hi[3] = {0, 0, 0};
    int32_t root_ref = 0;
    int32_t n_leaves = 0;
    int32_t max_depth = 0;
};
void flatten(const BuildResult& br, FlatBvh* out);

// The 4-wide form of a flattened tree (leaf refs and triangles unchanged).
struct FlatBvh4 {
    std::vector<DevNode4> nodes;
    int32_t root_ref = 0;
    int32_t max_depth = 0;  // wide levels to the deepest leaf
};
void widen(const FlatBvh& b2, FlatBvh4* out);

double sah_cost_of(const BuildResult& br);

}  // namespace rumi

// Opaque handles of the C ABI.
struct rumi_ctx {
    int device = 0;
    int cu_count = 0;
    // Work counters of the persistent kernels: one slot per launch, taken
    // round robin, so launches on one stream never share a live slot and
    // launches on different streams share one only after kCounterSlots
    // further launches.
    static constexpr uint32_t kCounterSlots = 1024;
    uint32_t* dev_counters = nullptr;
    std::atomic<uint32_t> next_counter{0};
    uint32_t* take_counter() { return dev_counters + (next_counter.fetch_add(1) % kCounterSlots); }
    rumi_trace_tuning tuning{256, 16, 16, 2, 4};
};

struct rumi_bvh {
    rumi::FlatBvh flat;
    int32_t mode = RUMI_BVH_SAH_BINNED;
    int32_t n_prims = 0;
    double sah_cost = 0.0;
    double build_seconds = 0.0;
    int device = -1;
    void* dev_nodes = nullptr;  // DevNode[n_inner]
    void* dev_tris = nullptr;   // DevTri[n_prims]
    std::vector<rumi::DevTri> host_tris;
    // SAH trees: the 4-wide form, usable when its stack bound (three entries
    // per wide level) fits the traversal stack
    rumi::FlatBvh4 wide;
    bool wide_ok = false;
    void* dev_nodes4 = nullptr;  // DevNode4[wide.nodes.size()]
};
