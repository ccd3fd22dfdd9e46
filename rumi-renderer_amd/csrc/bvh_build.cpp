// Host BVH builders: binned SAH (FP64 cost sums) and the reference's
// median-6 builder, plus flattening into the 64-byte GPU node layout.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>

#include "internal.h"

namespace rumi {

namespace {

struct Box {
    float lo[3], hi[3];
};

inline Box empty_box() {
    Box b;
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = FLT_MAX;
        b.hi[k] = -FLT_MAX;
    }
    return b;
}

// Union with the reference's __min/__max selection (src/Kernel/bbox.cpp:127-140).
inline void grow(Box& b, const Box& o) {
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = b.lo[k] < o.lo[k] ? b.lo[k] : o.lo[k];
        b.hi[k] = b.hi[k] > o.hi[k] ? b.hi[k] : o.hi[k];
    }
}

inline double area(const Box& b) {
    double dx = (double)b.hi[0] - (double)b.lo[0];
    double dy = (double)b.hi[1] - (double)b.lo[1];
    double dz = (double)b.hi[2] - (double)b.lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0;
    return 2.0 * (dx * dy + dx * dz + dy * dz);
}

inline Box prim_box(const float* lo, const float* hi, int i) {
    Box b;
    std::memcpy(b.lo, lo + 3 * (size_t)i, sizeof(float) * 3);
    std::memcpy(b.hi, hi + 3 * (size_t)i, sizeof(float) * 3);
    return b;
}

// ---------------------------------------------------------------- SAH
// Depth at which the SAH builder stops looking for SAH splits and halves
// the range instead, keeping the tree within kMaxDepth levels for any input.
constexpr int kSahDepthLimit = 64;

struct SahBuilder {
    const float* lo;
    const float* hi;
    std::vector<float> cen;  // 3 per prim: (hi + lo) * 0.5f
    int max_leaf;
    int nb;
    std::vector<int32_t> ids;
    std::vector<BuildNode> nodes;

    int make_leaf(int node, int begin, int end) {
        nodes[node].left = nodes[node].right = -1;
        nodes[node].first = begin;
        nodes[node].count = end - begin;
        return node;
    }

    int build(int begin, int end, int depth) {
        const int count = end - begin;
        Box nb_box = empty_box(), cb = empty_box();
        for (int i = begin; i < end; ++i) {
            const int p = ids[i];
            grow(nb_box, prim_box(lo, hi, p));
            for (int k = 0; k < 3; ++k) {
                float c = cen[3 * (size_t)p + k];
                cb.lo[k] = cb.lo[k] < c ? cb.lo[k] : c;
                cb.hi[k] = cb.hi[k] > c ? cb.hi[k] : c;
            }
        }
        const int node = (int)nodes.size();
        nodes.emplace_back();
        std::memcpy(nodes[node].lo, nb_box.lo, sizeof(float) * 3);
        std::memcpy(nodes[node].hi, nb_box.hi, sizeof(float) * 3);
        if (count <= 1) return make_leaf(node, begin, end);

        int best_axis = -1, best_bin = -1;
        double best_cost = std::numeric_limits<double>::infinity();
        if (depth < kSahDepthLimit) {
            std::vector<int> bcount(nb);
            std::vector<Box> bbox(nb);
            std::vector<double> right_area(nb), right_cnt(nb);
            for (int axis = 0; axis < 3; ++axis) {
                const double cmin = cb.lo[axis], cmax = cb.hi[axis];
                if (!(cmax > cmin)) continue;
                const double scale = (double)nb / (cmax - cmin);
                std::fill(bcount.begin(), bcount.end(), 0);
                std::fill(bbox.begin(), bbox.end(), empty_box());
                for (int i = begin; i < end; ++i) {
                    const int p = ids[i];
                    int b = (int)(((double)cen[3 * (size_t)p + axis] - cmin) * scale);
                    b = b < 0 ? 0 : (b >= nb ? nb - 1 : b);
                    bcount[b]++;
                    grow(bbox[b], prim_box(lo, hi, p));
                }
                Box acc = empty_box();
                double acc_n = 0;
                for (int b = nb - 1; b > 0; --b) {
                    grow(acc, bbox[b]);
                    acc_n += bcount[b];
                    right_area[b] = area(acc);
                    right_cnt[b] = acc_n;
                }
                acc = empty_box();
                acc_n = 0;
                for (int b = 0; b + 1 < nb; ++b) {
                    grow(acc, bbox[b]);
                    acc_n += bcount[b];
                    const double rn = right_cnt[b + 1];
                    if (acc_n == 0 || rn == 0) continue;
                    const double cost = area(acc) * acc_n + right_area[b + 1] * rn;
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = axis;
                        best_bin = b;
                    }
                }
            }
        }
        const double a_node = area(nb_box);
        if (count <= max_leaf) {
            const double split = best_axis >= 0 && a_node > 0 ? 1.0 + best_cost / a_node
                                                              : std::numeric_limits<double>::infinity();
            if ((double)count <= split) return make_leaf(node, begin, end);
        }
        int mid;
        if (best_axis >= 0) {
            const double cmin = cb.lo[best_axis], cmax = cb.hi[best_axis];
            const double scale = (double)nb / (cmax - cmin);
            const int axis = best_axis, split_bin = best_bin;
            auto it = std::partition(ids.begin() + begin, ids.begin() + end, [&](int32_t p) {
                int b = (int)(((double)cen[3 * (size_t)p + axis] - cmin) * scale);
                b = b < 0 ? 0 : (b >= nb ? nb - 1 : b);
                return b <= split_bin;
            });
            mid = (int)(it - ids.begin());
            if (mid == begin || mid == end) mid = begin + count / 2;
        } else {
            // No SAH split (identical centroids or depth limit): halve the
            // range on the longest centroid axis, ties by primitive index.
            int axis = 0;
            float ext = cb.hi[0] - cb.lo[0];
            for (int k = 1; k < 3; ++k)
                if (cb.hi[k] - cb.lo[k] > ext) {
                    ext = cb.hi[k] - cb.lo[k];
                    axis = k;
                }
            mid = begin + count / 2;
            std::nth_element(ids.begin() + begin, ids.begin() + mid, ids.begin() + end,
                             [&](int32_t a, int32_t b) {
                                 float ca = cen[3 * (size_t)a + axis], cb2 = cen[3 * (size_t)b + axis];
                                 return ca < cb2 || (ca == cb2 && a < b);
                             });
        }
        const int l = build(begin, mid, depth + 1);
        const int r = build(mid, end, depth + 1);
        nodes[node].left = l;
        nodes[node].right = r;
        return node;
    }
};

// ---------------------------------------------------------------- median-6
// BVHAccel::CreateBVHNode (src/Acceleration/bvh.cpp:48-127): the extents of
// the range pick the longest axis (ties resolved as the reference does), a
// shell sort of the box centres along it runs unless the parent used the
// same axis, ranges of more than 6 are cut at (start + end) >> 1.
struct MedianBuilder {
    std::vector<int32_t> objs;
    std::vector<Box> bounds;
    std::vector<BuildNode> nodes;

    static float centre(const Box& b, int dim) { return (b.hi[dim] + b.lo[dim]) * 0.5f; }

    int create(int start, int end, int pre_dim, Box* out) {
        Box ext = empty_box();
        for (int i = start; i <= end; ++i) grow(ext, bounds[i]);
        const int count = end - start + 1;
        const float dx = ext.hi[0] - ext.lo[0];
        const float dy = ext.hi[1] - ext.lo[1];
        const float dz = ext.hi[2] - ext.lo[2];
        int dim = 0;
        if (dx < dy && dz < dy) dim = 1;
        if (dx < dz && dy < dz) dim = 2;
        if (dim != pre_dim) {
            for (int gap = count >> 1; gap >= 1; gap >>= 1) {
                for (int ti = start + gap; ti <= end; ++ti) {
                    const int32_t ts = objs[ti];
                    const Box tb = bounds[ti];
                    const float key = centre(tb, dim);
                    for (int ji = ti - gap; ji >= start; ji -= gap) {
                        if (centre(bounds[ji], dim) > key) {
                            objs[ji + gap] = objs[ji];
                            bounds[ji + gap] = bounds[ji];
                            objs[ji] = ts;
                            bounds[ji] = tb;
                        } else {
                            break;
                        }
                    }
                }
            }
        }
        const int node = (int)nodes.size();
        nodes.emplace_back();
        if (count > 6) {
            const int cut = (start + end) >> 1;
            Box lb, rb;
            const int l = create(start, cut, dim, &lb);
            const int r = create(cut + 1, end, dim, &rb);
            Box nbx = lb;
            grow(nbx, rb);
            nodes[node].left = l;
            nodes[node].right = r;
            std::memcpy(nodes[node].lo, nbx.lo, sizeof(float) * 3);
            std::memcpy(nodes[node].hi, nbx.hi, sizeof(float) * 3);
            *out = nbx;
        } else {
            nodes[node].first = start;
            nodes[node].count = count;
            std::memcpy(nodes[node].lo, ext.lo, sizeof(float) * 3);
            std::memcpy(nodes[node].hi, ext.hi, sizeof(float) * 3);
            *out = ext;
        }
        return node;
    }
};

}  // namespace

void build_sah(const float* lo, const float* hi, int n, int max_leaf, int n_bins,
               BuildResult* out) {
    out->nodes.clear();
    out->order.clear();
    out->root = -1;
    if (n <= 0) return;
    SahBuilder b;
    b.lo = lo;
    b.hi = hi;
    b.max_leaf = max_leaf;
    b.nb = n_bins;
    b.cen.resize(3 * (size_t)n);
    for (size_t i = 0; i < 3 * (size_t)n; ++i) b.cen[i] = (hi[i] + lo[i]) * 0.5f;
    b.ids.resize(n);
    for (int i = 0; i < n; ++i) b.ids[i] = i;
    b.nodes.reserve(2 * (size_t)n / (max_leaf > 1 ? max_leaf - 1 : 1) + 1);
    out->root = b.build(0, n, 0);
    out->nodes.swap(b.nodes);
    out->order.swap(b.ids);
    out->sah_cost = sah_cost_of(*out);
}

void build_median6(const float* lo, const float* hi, int n, BuildResult* out) {
    out->nodes.clear();
    out->order.clear();
    out->root = -1;
    if (n <= 0) return;
    MedianBuilder b;
    b.objs.resize(n);
    b.bounds.resize(n);
    for (int i = 0; i < n; ++i) {
        b.objs[i] = i;
        b.bounds[i] = prim_box(lo, hi, i);
    }
    Box rb;
    out->root = b.create(0, n - 1, 3, &rb);
    out->nodes.swap(b.nodes);
    out->order.swap(b.objs);
    out->sah_cost = sah_cost_of(*out);
}

double sah_cost_of(const BuildResult& br) {
    if (br.root < 0) return 0.0;
    Box root;
    std::memcpy(root.lo, br.nodes[br.root].lo, sizeof(float) * 3);
    std::memcpy(root.hi, br.nodes[br.root].hi, sizeof(float) * 3);
    const double ar = area(root);
    if (!(ar > 0)) return 0.0;
    double cost = 0.0;
    for (const BuildNode& n : br.nodes) {
        Box b;
        std::memcpy(b.lo, n.lo, sizeof(float) * 3);
        std::memcpy(b.hi, n.hi, sizeof(float) * 3);
        const double w = area(b) / ar;
        cost += n.left < 0 ? w * n.count : w;
    }
    return cost;
}

namespace {

struct Flattener {
    const BuildResult& br;
    FlatBvh* out;

    int32_t emit(int node, int depth) {
        const BuildNode& n = br.nodes[node];
        if (n.left < 0) {
            out->n_leaves++;
            if (depth > out->max_depth) out->max_depth = depth;
            return leaf_ref(n.first, n.count);
        }
        const int32_t idx = (int32_t)out->nodes.size();
        out->nodes.emplace_back();
        const int32_t r0 = emit(n.left, depth + 1);
        const int32_t r1 = emit(n.right, depth + 1);
        const BuildNode& c0 = br.nodes[n.left];
        const BuildNode& c1 = br.nodes[n.right];
        DevNode& d = out->nodes[idx];
        d.a[0] = c0.lo[0]; d.a[1] = c0.lo[1]; d.a[2] = c0.lo[2]; d.a[3] = c0.hi[0];
        d.b[0] = c0.hi[1]; d.b[1] = c0.hi[2]; d.b[2] = c1.lo[0]; d.b[3] = c1.lo[1];
        d.c[0] = c1.lo[2]; d.c[1] = c1.hi[0]; d.c[2] = c1.hi[1]; d.c[3] = c1.hi[2];
        d.d[0] = r0;
        d.d[1] = r1;
        d.d[2] = 0;
        d.d[3] = 0;
        return idx;
    }
};

}  // namespace

void flatten(const BuildResult& br, FlatBvh* out) {
    out->nodes.clear();
    out->leaf_prims = br.order;
    out->n_leaves = 0;
    out->max_depth = 0;
    if (br.root < 0) {
        for (int k = 0; k < 3; ++k) {
            out->root_lo[k] = FLT_MAX;
            out->root_hi[k] = -FLT_MAX;
        }
        out->root_ref = leaf_ref(0, 0);
        return;
    }
    std::memcpy(out->root_lo, br.nodes[br.root].lo, sizeof(float) * 3);
    std::memcpy(out->root_hi, br.nodes[br.root].hi, sizeof(float) * 3);
    out->nodes.reserve(br.nodes.size() / 2 + 1);
    Flattener f{br, out};
    out->root_ref = f.emit(br.root, 0);
}

namespace {

struct WideChild {
    float lo[3], hi[3];
    int32_t ref;
};

WideChild child_of(const DevNode& d, int c) {
    WideChild w;
    if (c == 0) {
        w.lo[0] = d.a[0]; w.lo[1] = d.a[1]; w.lo[2] = d.a[2];
        w.hi[0] = d.a[3]; w.hi[1] = d.b[0]; w.hi[2] = d.b[1];
    } else {
        w.lo[0] = d.b[2]; w.lo[1] = d.b[3]; w.lo[2] = d.c[0];
        w.hi[0] = d.c[1]; w.hi[1] = d.c[2]; w.hi[2] = d.c[3];
    }
    w.ref = d.d[c];
    return w;
}

double wide_area(const WideChild& c) {
    const double dx = (double)c.hi[0] - c.lo[0], dy = (double)c.hi[1] - c.lo[1],
                 dz = (double)c.hi[2] - c.lo[2];
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

struct Widener {
    const FlatBvh& b2;
    FlatBvh4* out;

    // Converts binary interior node `node`; returns its wide index.
    int32_t emit(int32_t node, int depth) {
        WideChild ch[4];
        int n = 2;
        ch[0] = child_of(b2.nodes[node], 0);
        ch[1] = child_of(b2.nodes[node], 1);
        while (n < 4) {
            int best = -1;
            double best_area = -1.0;
            for (int k = 0; k < n; ++k)
                if (ch[k].ref >= 0) {
                    const double a = wide_area(ch[k]);
                    if (a > best_area) {
                        best_area = a;
                        best = k;
                    }
                }
            if (best < 0) break;
            const DevNode& e = b2.nodes[ch[best].ref];
            for (int k = n; k > best + 1; --k) ch[k] = ch[k - 1];
            ch[best] = child_of(e, 0);
            ch[best + 1] = child_of(e, 1);
            ++n;
        }
        const int32_t idx = (int32_t)out->nodes.size();
        out->nodes.emplace_back();
        int32_t refs[4] = {kNoChild, kNoChild, kNoChild, kNoChild};
        for (int k = 0; k < n; ++k) {
            if (ch[k].ref >= 0) {
                refs[k] = emit(ch[k].ref, depth + 1);
            } else {
                refs[k] = ch[k].ref;
                if (depth + 1 > out->max_depth) out->max_depth = depth + 1;
            }
        }
        DevNode4& d = out->nodes[idx];
        for (int k = 0; k < 4; ++k) {
            for (int a = 0; a < 3; ++a) {
                d.lo[a][k] = k < n ? ch[k].lo[a] : 0.0f;
                d.hi[a][k] = k < n ? ch[k].hi[a] : 0.0f;
            }
            d.ref[k] = refs[k];
            d.pad[k] = 0;
        }
        return idx;
    }
};

}  // namespace

void widen(const FlatBvh& b2, FlatBvh4* out) {
    out->nodes.clear();
    out->max_depth = 0;
    if (b2.root_ref < 0) {
        out->root_ref = b2.root_ref;
        return;
    }
    out->nodes.reserve(b2.nodes.size() / 3 + 1);
    Widener w{b2, out};
    out->root_ref = w.emit(b2.root_ref, 0);
}

}  // namespace rumi
