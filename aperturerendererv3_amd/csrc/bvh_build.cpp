// Host BVH2 builder: binned SAH over all three axes, parallel over subtrees.
//
// Semantics follow the reference builder (core/bvh/bvh.cuh:142-422): a
// binary tree of primitive ranges in pre-order with left child = parent+1,
// leaves of at most leaf_size primitives, SAH split with node and primitive
// costs of 1, and an object-median fallback.  Differences, chosen for the GPU
// traversal: all three axes are binned (the reference bins the longest
// centroid axis only), leaves never exceed leaf_size (identical centroids are
// split by index instead of forming one large leaf), and the depth is bounded
// by max_depth (switching to median splits when the remaining budget runs
// short) so the traversal stack has a fixed size.
#include "ar_internal.hpp"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <thread>

namespace ar {
namespace {

constexpr int kMaxBins = 64;

struct Box {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const float* l, const float* h) {
        for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], l[a]); hi[a] = std::max(hi[a], h[a]); }
    }
    void grow(const Box& b) { grow(b.lo, b.hi); }
    float area() const {
        const float dx = std::max(0.0f, hi[0] - lo[0]);
        const float dy = std::max(0.0f, hi[1] - lo[1]);
        const float dz = std::max(0.0f, hi[2] - lo[2]);
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct Builder {
    const float *lo, *hi, *cen;
    uint32_t* order;
    uint32_t leaf_size;
    int bins;
    int max_depth;
    std::atomic<int> threads_free{0};

    static int ceil_log2(uint32_t x) {
        int r = 0;
        while ((1u << r) < x) ++r;
        return r;
    }

    // Appends the subtree over order[start, end) to out in pre-order.
    // Returns the subtree depth.
    int build(uint32_t start, uint32_t end, int depth, std::vector<BuildNode>& out) {
        const uint32_t count = end - start;
        const size_t self = out.size();
        out.emplace_back();
        Box bounds, cbounds;
        for (uint32_t i = start; i < end; ++i) {
            const uint32_t p = order[i];
            bounds.grow(&lo[3 * p], &hi[3 * p]);
            cbounds.grow(&cen[3 * p], &cen[3 * p]);
        }
        {
            BuildNode& n = out[self];
            for (int a = 0; a < 3; ++a) { n.lo[a] = bounds.lo[a]; n.hi[a] = bounds.hi[a]; }
        }
        if (count <= leaf_size) {
            BuildNode& n = out[self];
            n.left = n.right = -1;
            n.first = start;
            n.count = count;
            return 0;
        }

        uint32_t mid = start;
        bool split = false;
        const int budget = max_depth - depth;
        const bool need_median = ceil_log2((count + leaf_size - 1) / leaf_size) >= budget;
        float ext[3];
        for (int a = 0; a < 3; ++a) ext[a] = cbounds.hi[a] - cbounds.lo[a];
        int long_axis = 0;
        if (ext[1] > ext[long_axis]) long_axis = 1;
        if (ext[2] > ext[long_axis]) long_axis = 2;

        if (!need_median && ext[long_axis] > 0.0f) {
            // bin on all three axes
            Box bbox[3][kMaxBins];
            uint32_t bcnt[3][kMaxBins] = {};
            float scale[3];
            for (int a = 0; a < 3; ++a)
                scale[a] = ext[a] > 0.0f ? (float)bins / ext[a] : 0.0f;
            for (uint32_t i = start; i < end; ++i) {
                const uint32_t p = order[i];
                for (int a = 0; a < 3; ++a) {
                    int b = (int)((cen[3 * p + a] - cbounds.lo[a]) * scale[a]);
                    b = std::min(std::max(b, 0), bins - 1);
                    bbox[a][b].grow(&lo[3 * p], &hi[3 * p]);
                    bcnt[a][b]++;
                }
            }
            float best = FLT_MAX;
            int best_axis = -1, best_bin = -1;
            for (int a = 0; a < 3; ++a) {
                if (scale[a] == 0.0f) continue;
                float right_cost[kMaxBins];
                Box acc;
                uint32_t n_acc = 0;
                for (int b = bins - 1; b > 0; --b) {
                    if (bcnt[a][b]) acc.grow(bbox[a][b]);
                    n_acc += bcnt[a][b];
                    right_cost[b] = n_acc ? acc.area() * (float)n_acc : 0.0f;
                }
                Box lacc;
                uint32_t nl = 0;
                for (int b = 0; b < bins - 1; ++b) {
                    if (bcnt[a][b]) lacc.grow(bbox[a][b]);
                    nl += bcnt[a][b];
                    if (nl == 0 || nl == count) continue;
                    const float c = lacc.area() * (float)nl + right_cost[b + 1];
                    if (c < best) { best = c; best_axis = a; best_bin = b; }
                }
            }
            if (best_axis >= 0) {
                const int a = best_axis;
                const float cl = cbounds.lo[a], sc = scale[a];
                const int bb = best_bin, nb = bins;
                uint32_t* m = std::partition(order + start, order + end, [&](uint32_t p) {
                    int b = (int)((cen[3 * p + a] - cl) * sc);
                    b = std::min(std::max(b, 0), nb - 1);
                    return b <= bb;
                });
                mid = (uint32_t)(m - order);
                split = mid > start && mid < end;
            }
        }
        if (!split) {
            // object median on the longest centroid axis (index order when all
            // centroids coincide)
            mid = start + count / 2;
            const int a = long_axis;
            std::nth_element(order + start, order + mid, order + end, [&](uint32_t x, uint32_t y) {
                const float kx = cen[3 * x + a], ky = cen[3 * y + a];
                return kx < ky || (kx == ky && x < y);
            });
        }

        int dl, dr;
        if (count > 65536 && threads_free.fetch_sub(1) > 0) {
            std::vector<BuildNode> left_nodes;
            left_nodes.reserve(2 * (mid - start));
            std::thread t([&] { dl = build(start, mid, depth + 1, left_nodes); });
            std::vector<BuildNode> right_nodes;
            right_nodes.reserve(2 * (end - mid));
            dr = build(mid, end, depth + 1, right_nodes);
            t.join();
            threads_free.fetch_add(1);
            const int32_t lbase = (int32_t)out.size();
            for (BuildNode n : left_nodes) {
                if (n.left >= 0) { n.left += lbase; n.right += lbase; }
                out.push_back(n);
            }
            const int32_t rbase = (int32_t)out.size();
            for (BuildNode n : right_nodes) {
                if (n.left >= 0) { n.left += rbase; n.right += rbase; }
                out.push_back(n);
            }
            out[self].left = lbase;
            out[self].right = rbase;
        } else {
            if (count > 65536) threads_free.fetch_add(1);
            const int32_t l = (int32_t)out.size();
            dl = build(start, mid, depth + 1, out);
            const int32_t r = (int32_t)out.size();
            dr = build(mid, end, depth + 1, out);
            out[self].left = l;
            out[self].right = r;
        }
        out[self].first = 0;
        out[self].count = 0;
        return 1 + std::max(dl, dr);
    }
};

} // namespace

BuildResult build_bvh2(const float* lo, const float* hi, const float* cen, uint32_t n,
                       uint32_t leaf_size, uint32_t bins, int max_depth) {
    BuildResult r;
    if (n == 0) return r;
    r.order.resize(n);
    for (uint32_t i = 0; i < n; ++i) r.order[i] = i;
    Builder b;
    b.lo = lo; b.hi = hi; b.cen = cen;
    b.order = r.order.data();
    b.leaf_size = std::max(1u, std::min(leaf_size, kMaxLeafTris));
    b.bins = (int)std::max(4u, std::min(bins, (uint32_t)kMaxBins));
    b.max_depth = max_depth;
    const unsigned hc = std::max(1u, std::thread::hardware_concurrency());
    b.threads_free.store((int)std::min(hc, 64u) - 1);
    r.nodes.reserve(2 * (size_t)n);
    r.depth = b.build(0, n, 0, r.nodes);

    // SAH cost of the result, normalised by the root area
    Box root;
    root.grow(r.nodes[0].lo, r.nodes[0].hi);
    const double ra = std::max(root.area(), 1e-30f);
    double cost = 0.0;
    for (const BuildNode& nd : r.nodes) {
        Box bx;
        bx.grow(nd.lo, nd.hi);
        cost += (double)bx.area() / ra * (nd.left >= 0 ? 1.0 : (double)nd.count);
    }
    r.sah_cost = cost;
    return r;
}

} // namespace ar
