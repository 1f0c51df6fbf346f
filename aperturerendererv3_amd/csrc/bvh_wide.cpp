// Collapse of the binary SAH tree into the quantised 4- or 8-wide device
// layouts.
//
// Each wide node takes a binary node's two children and repeatedly opens the
// interior child of largest surface area (its two children replace it) until
// four / eight children or only leaves remain, as in the wide-BVH construction of
// Wald et al. (RT'08) and Ylitie et al. (HPG'17).  Child boxes are padded
// by the scene pad, then quantised outward to 8 bits on a per-node grid
// p + q * 2^e; every decoded bound is checked in float against the padded
// box and nudged outward if rounding put it inside.
#include "ar_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ar {
namespace {

float area_of(const BuildNode& n) {
    const float dx = std::max(0.0f, n.hi[0] - n.lo[0]);
    const float dy = std::max(0.0f, n.hi[1] - n.lo[1]);
    const float dz = std::max(0.0f, n.hi[2] - n.lo[2]);
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// quantised child bounds: lo.x..lo.z then hi.x..hi.z, K bytes each
template <int K> struct Quant;
template <> struct Quant<8> {
    using Node = GpuNode8;
    static void set(Node& g, int a, int k, uint8_t lo, uint8_t hi) { g.q[8 * a + k] = lo; g.q[24 + 8 * a + k] = hi; }
};
template <> struct Quant<4> {
    using Node = GpuNode4;
    static void set(Node& g, int a, int k, uint8_t lo, uint8_t hi) { g.q[4 * a + k] = lo; g.q[12 + 4 * a + k] = hi; }
};

template <int K>
struct Collapser {
    using Node = typename Quant<K>::Node;
    const BuildResult& br;
    LeafWriter& lw;
    float pad;
    std::vector<Node> nodes;
    std::vector<float> tris;
    uint64_t leaves = 0;
    int max_depth = 0;

    // A leaf's slots (48 B each) are placed so that they touch no more
    // 128-byte lines than their size needs (one for 1-2 slots, two for 3):
    // up to two unused slots are inserted before it.  The slot array starts
    // on a 256-byte boundary (hipMalloc).
    static uint32_t lines(size_t slot, uint32_t count) {
        const size_t a = 48 * slot, b = a + 48 * (size_t)count - 1;
        return (uint32_t)(b / 128 - a / 128 + 1);
    }

    uint32_t leaf_ref(const BuildNode& n) {
        const uint32_t need = (48 * n.count + 127) / 128;
        for (int pad_slots = 0; pad_slots < 2 && lines(tris.size() / 12, n.count) > need; ++pad_slots)
            tris.resize(tris.size() + 12, 0.0f);
        const uint32_t first = (uint32_t)(tris.size() / 12);
        for (uint32_t k = 0; k < n.count; ++k) lw.write(tris, br.order[n.first + k]);
        ++leaves;
        return make_leaf_ref(first, n.count);
    }

    // biased exponent e (1..254) with 255 * 2^(e-127) >= ext
    static int exponent_for(float ext) {
        if (!(ext > 0.0f)) return 1;
        int e = (int)std::ceil(std::log2((double)ext / 255.0));
        while (std::ldexp(255.0, e) < (double)ext) ++e;
        return std::min(254, std::max(1, e + 127));
    }

    static float scale_of(int biased) {
        uint32_t bits = (uint32_t)biased << 23;
        float s;
        std::memcpy(&s, &bits, 4);
        return s;
    }

    // children: binary-node indices of the wide node's children
    uint32_t build(int bi, int depth) {
        max_depth = std::max(max_depth, depth);
        std::vector<int> ch;
        const BuildNode& n = br.nodes[bi];
        if (n.left < 0) {
            ch.push_back(bi);
        } else {
            ch.push_back(n.left);
            ch.push_back(n.right);
            while (ch.size() < (size_t)K) {
                int best = -1;
                float best_a = -1.0f;
                for (size_t k = 0; k < ch.size(); ++k) {
                    const BuildNode& c = br.nodes[ch[k]];
                    if (c.left >= 0 && area_of(c) > best_a) { best_a = area_of(c); best = (int)k; }
                }
                if (best < 0) break;
                const BuildNode& c = br.nodes[ch[best]];
                ch[best] = c.left;
                ch.push_back(c.right);
            }
        }
        // children in ascending order of their centre along the axis of
        // largest extent of their union (bits 28-29 of meta): traversal pops
        // the hit children after the nearest in that order or its reverse,
        // by the sign of the ray's direction on the axis
        int axis = 0;
        {
            float ulo[3] = {INFINITY, INFINITY, INFINITY}, uhi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int c : ch)
                for (int a = 0; a < 3; ++a) {
                    ulo[a] = std::min(ulo[a], br.nodes[c].lo[a]);
                    uhi[a] = std::max(uhi[a], br.nodes[c].hi[a]);
                }
            for (int a = 1; a < 3; ++a)
                if (uhi[a] - ulo[a] > uhi[axis] - ulo[axis]) axis = a;
            std::stable_sort(ch.begin(), ch.end(), [&](int x, int y) {
                return br.nodes[x].lo[axis] + br.nodes[x].hi[axis] < br.nodes[y].lo[axis] + br.nodes[y].hi[axis];
            });
        }
        const uint32_t gi = (uint32_t)nodes.size();
        nodes.emplace_back();
        uint32_t refs[K];
        for (size_t k = 0; k < (size_t)K; ++k) refs[k] = kRefEmpty;
        for (size_t k = 0; k < ch.size(); ++k) {
            const BuildNode& c = br.nodes[ch[k]];
            refs[k] = c.left < 0 ? leaf_ref(c) : build(ch[k], depth + 1);
        }
        // quantisation grid over the union of the padded child boxes
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int c : ch)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], br.nodes[c].lo[a] - pad);
                hi[a] = std::max(hi[a], br.nodes[c].hi[a] + pad);
            }
        Node g{};
        int eb[3];
        float s[3];
        for (int a = 0; a < 3; ++a) {
            g.p[a] = lo[a];
            eb[a] = exponent_for(hi[a] - lo[a]);
            s[a] = scale_of(eb[a]);
        }
        g.meta = (uint32_t)eb[0] | ((uint32_t)eb[1] << 8) | ((uint32_t)eb[2] << 16) | ((uint32_t)ch.size() << 24) |
                 ((uint32_t)axis << 28);
        for (int k = 0; k < K; ++k) g.ref[k] = refs[k];
        for (size_t k = 0; k < ch.size(); ++k) {
            const BuildNode& c = br.nodes[ch[k]];
            for (int a = 0; a < 3; ++a) {
                const float clo = c.lo[a] - pad, chi = c.hi[a] + pad;
                long ql = (long)std::floor(((double)clo - (double)g.p[a]) / (double)s[a]);
                long qh = (long)std::ceil(((double)chi - (double)g.p[a]) / (double)s[a]);
                ql = std::min(255L, std::max(0L, ql));
                qh = std::min(255L, std::max(0L, qh));
                while (ql > 0 && std::fma((float)ql, s[a], g.p[a]) > clo) --ql;
                while (qh < 255 && std::fma((float)qh, s[a], g.p[a]) < chi) ++qh;
                Quant<K>::set(g, a, (int)k, (uint8_t)ql, (uint8_t)qh);
            }
        }
        nodes[gi] = g;
        return gi;
    }
};

} // namespace

void LeafWriter::write(std::vector<float>& slots, uint32_t prim) {
    const size_t at = slots.size();
    slots.resize(at + 12, 0.0f);
    float* s = &slots[at];
    if (prim < ntri) {
        for (int v = 0; v < 3; ++v) {
            const float* p = &verts[3 * (size_t)idx[3 * (size_t)prim + v]];
            s[4 * v + 0] = p[0];
            s[4 * v + 1] = p[1];
            s[4 * v + 2] = p[2];
        }
        std::memcpy(&s[3], &prim, 4);
    } else {
        const uint32_t k = prim - ntri;
        const uint32_t w = kObjSlot | k;
        std::memcpy(&s[3], &w, 4);
        if (obj_slot.size() <= k) obj_slot.resize(k + 1, kRefEmpty);
        obj_slot[k] = (uint32_t)(at / 12);
    }
}

template <int K, class Node>
static void convert(const BuildResult& br, LeafWriter& lw, float pad, std::vector<Node>& nodes,
                    std::vector<float>& tris, uint64_t& leaves, int& depth) {
    Collapser<K> c{br, lw, pad, {}, {}, 0, 0};
    c.nodes.reserve(br.nodes.size() / (K / 2) + 1);
    c.tris.reserve(12 * br.order.size());
    c.build(0, 0);
    nodes.swap(c.nodes);
    tris.swap(c.tris);
    leaves = c.leaves;
    depth = c.max_depth;
}

bool convert_bvh8(const BuildResult& br, LeafWriter& lw, float pad, std::vector<GpuNode8>& nodes,
                  std::vector<float>& tris, uint64_t& leaves, int& depth) {
    convert<8>(br, lw, pad, nodes, tris, leaves, depth);
    return true;
}

bool convert_bvh4(const BuildResult& br, LeafWriter& lw, float pad, std::vector<GpuNode4>& nodes,
                  std::vector<float>& tris, uint64_t& leaves, int& depth) {
    convert<4>(br, lw, pad, nodes, tris, leaves, depth);
    return true;
}

} // namespace ar
