// Host-side check of the BVH builder and the wide-node collapse, built with
// AddressSanitizer and UndefinedBehaviorSanitizer by test_sanitizers.py (the
// host code of the HIP library compiled by g++, no device code).
//
// For seeded meshes (a random soup, coincident copies with zero-area
// triangles, a regular grid, one triangle, a soup spanning 1e-3..1e4) and
// several leaf sizes, bin counts and depth bounds it checks:
//   binary tree:  pre-order with left = parent + 1, leaf ranges partition the
//                 primitive order, the order is a permutation, leaves hold at
//                 most leaf_size primitives, children lie inside the parent
//                 box, the reported depth is the deepest leaf;
//   4- and 8-wide nodes: every primitive is reached exactly once through the
//                 child references, every decoded child box (p + q * 2^(e-127)
//                 in float) contains the vertices of every triangle below it,
//                 the reported depth is that of the deepest wide node
//                 (root 0), from which the traversal stack is sized.
// Exit status 0 and "ok" on success; the first failure is printed.
#include "../../aperturerendererv3_amd/csrc/ar_internal.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

using namespace ar;

namespace {

struct Mesh {
    std::string name;
    std::vector<float> v;
    std::vector<uint32_t> f;
    uint32_t ntri() const { return (uint32_t)(f.size() / 3); }
};

int failures = 0;

void fail(const std::string& what) {
    if (failures++ < 10) std::printf("FAIL %s\n", what.c_str());
}

void add_tri(Mesh& m, const float* a, const float* b, const float* c) {
    const uint32_t base = (uint32_t)(m.v.size() / 3);
    m.v.insert(m.v.end(), a, a + 3);
    m.v.insert(m.v.end(), b, b + 3);
    m.v.insert(m.v.end(), c, c + 3);
    m.f.insert(m.f.end(), {base, base + 1, base + 2});
}

Mesh soup(uint32_t n, float lo_scale, float hi_scale, uint32_t seed, const char* name) {
    Mesh m{name, {}, {}};
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    std::uniform_real_distribution<float> lg(std::log(lo_scale), std::log(hi_scale));
    for (uint32_t t = 0; t < n; ++t) {
        const float s = std::exp(lg(rng));
        const float c[3] = {u(rng) * s, u(rng) * s, u(rng) * s};
        float p[3][3];
        for (int k = 0; k < 3; ++k)
            for (int a = 0; a < 3; ++a) p[k][a] = c[a] + 0.05f * s * u(rng);
        add_tri(m, p[0], p[1], p[2]);
    }
    return m;
}

Mesh coincident() {
    Mesh m{"coincident", {}, {}};
    const float a[3] = {-0.5f, -0.5f, 0.0f}, b[3] = {0.5f, -0.5f, 0.0f}, c[3] = {0.0f, 0.5f, 0.0f};
    const float pt[3] = {0.1f, 0.1f, 0.0f};
    for (int k = 0; k < 1000; ++k) add_tri(m, a, b, c);
    for (int k = 0; k < 64; ++k) add_tri(m, pt, pt, pt);
    Mesh s = soup(100, 0.5f, 1.0f, 9, "");
    for (uint32_t t = 0; t < s.ntri(); ++t) add_tri(m, &s.v[9 * t], &s.v[9 * t + 3], &s.v[9 * t + 6]);
    return m;
}

Mesh grid(int n) {
    Mesh m{"grid", {}, {}};
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const float p00[3] = {(float)x, (float)y, 0.0f}, p10[3] = {(float)x + 1, (float)y, 0.0f};
            const float p01[3] = {(float)x, (float)y + 1, 0.0f}, p11[3] = {(float)x + 1, (float)y + 1, 0.0f};
            add_tri(m, p00, p10, p11);
            add_tri(m, p00, p11, p01);
        }
    return m;
}

void bounds(const Mesh& m, std::vector<float>& lo, std::vector<float>& hi, std::vector<float>& cen) {
    const uint32_t n = m.ntri();
    lo.assign(3 * (size_t)n, 0.0f);
    hi.assign(3 * (size_t)n, 0.0f);
    cen.assign(3 * (size_t)n, 0.0f);
    for (uint32_t t = 0; t < n; ++t)
        for (int a = 0; a < 3; ++a) {
            float l = INFINITY, h = -INFINITY;
            for (int k = 0; k < 3; ++k) {
                const float x = m.v[3 * (size_t)m.f[3 * (size_t)t + k] + a];
                l = std::fmin(l, x);
                h = std::fmax(h, x);
            }
            lo[3 * (size_t)t + a] = l;
            hi[3 * (size_t)t + a] = h;
            cen[3 * (size_t)t + a] = 0.5f * (l + h);
        }
}

int check_binary(const BuildResult& br, uint32_t n, uint32_t leaf_size, const std::string& tag) {
    std::vector<int> seen(n, 0);
    for (uint32_t p : br.order) {
        if (p >= n) { fail(tag + ": order holds an out-of-range primitive"); return 0; }
        seen[p]++;
    }
    for (uint32_t p = 0; p < n; ++p)
        if (seen[p] != 1) { fail(tag + ": order is not a permutation"); break; }
    std::vector<int> covered(n, 0);
    int deepest = 0;
    struct Item { int32_t node; int depth; };
    std::vector<Item> st{{0, 0}};
    size_t visited = 0;
    while (!st.empty()) {
        const Item it = st.back();
        st.pop_back();
        if (it.node < 0 || (size_t)it.node >= br.nodes.size()) { fail(tag + ": child index out of range"); return 0; }
        ++visited;
        const BuildNode& nd = br.nodes[it.node];
        deepest = std::max(deepest, it.depth);
        if (nd.left < 0) {
            if (nd.count == 0 || nd.count > leaf_size) fail(tag + ": leaf size " + std::to_string(nd.count));
            if ((size_t)nd.first + nd.count > n) { fail(tag + ": leaf range out of range"); continue; }
            for (uint32_t i = nd.first; i < nd.first + nd.count; ++i) covered[i]++;
            continue;
        }
        if (nd.left != it.node + 1) fail(tag + ": left child is not parent + 1");
        for (int32_t c : {nd.left, nd.right}) {
            if (c < 0 || (size_t)c >= br.nodes.size()) { fail(tag + ": bad child"); continue; }
            const BuildNode& ch = br.nodes[c];
            for (int a = 0; a < 3; ++a)
                if (ch.lo[a] < nd.lo[a] || ch.hi[a] > nd.hi[a]) { fail(tag + ": child box outside its parent"); break; }
            st.push_back({c, it.depth + 1});
        }
    }
    if (visited != br.nodes.size()) fail(tag + ": unreachable nodes");
    for (uint32_t i = 0; i < n; ++i)
        if (covered[i] != 1) { fail(tag + ": leaf ranges do not partition the order"); break; }
    if (deepest != br.depth) fail(tag + ": depth " + std::to_string(br.depth) + " vs deepest leaf " + std::to_string(deepest));
    return deepest;
}

template <int K, class Node>
void check_wide(const Mesh& m, const std::vector<Node>& nodes, const std::vector<float>& slots, int depth,
                const std::string& tag) {
    const uint32_t n = m.ntri();
    std::vector<int> seen(n, 0);
    int deepest = 0;
    struct Item { uint32_t node; int depth; };
    std::vector<Item> st{{0, 0}};
    while (!st.empty()) {
        const Item it = st.back();
        st.pop_back();
        if (it.node >= nodes.size()) { fail(tag + ": node index out of range"); return; }
        const Node& g = nodes[it.node];
        deepest = std::max(deepest, it.depth);
        const uint32_t meta = g.meta;
        const uint32_t nchild = (meta >> 24) & 15u;
        if (nchild < 1 || nchild > (uint32_t)K) { fail(tag + ": child count " + std::to_string(nchild)); continue; }
        float scale[3];
        for (int a = 0; a < 3; ++a) {
            const uint32_t e = (meta >> (8 * a)) & 255u;
            uint32_t bits = e << 23;
            std::memcpy(&scale[a], &bits, 4);
        }
        for (uint32_t k = 0; k < (uint32_t)K; ++k) {
            const uint32_t ref = g.ref[k];
            if (k >= nchild) {
                if (ref != kRefEmpty) fail(tag + ": a child beyond the count");
                continue;
            }
            if (ref == kRefEmpty) { fail(tag + ": empty child inside the count"); continue; }
            float blo[3], bhi[3];
            for (int a = 0; a < 3; ++a) {
                const uint8_t ql = g.q[K * a + k], qh = g.q[3 * K + K * a + k];
                blo[a] = g.p[a] + (float)ql * scale[a];
                bhi[a] = g.p[a] + (float)qh * scale[a];
            }
            std::vector<uint32_t> prims;
            if (ref & kLeafBit) {
                const uint32_t first = ref & 0x07FFFFFFu, count = ((ref >> 27) & 15u) + 1u;
                if ((size_t)(first + count) * 12 > slots.size()) { fail(tag + ": leaf slots out of range"); continue; }
                for (uint32_t s = first; s < first + count; ++s) {
                    uint32_t p;
                    std::memcpy(&p, &slots[12 * (size_t)s + 3], 4);
                    if (p >= n) { fail(tag + ": slot holds an out-of-range primitive"); continue; }
                    seen[p]++;
                    prims.push_back(p);
                }
            } else {
                // every triangle below the interior child: walk its subtree
                std::vector<uint32_t> sub{ref};
                while (!sub.empty()) {
                    const uint32_t x = sub.back();
                    sub.pop_back();
                    if (x >= nodes.size()) break;
                    for (uint32_t j = 0; j < (uint32_t)K; ++j) {
                        const uint32_t r = nodes[x].ref[j];
                        if (r == kRefEmpty) continue;
                        if (r & kLeafBit) {
                            const uint32_t first = r & 0x07FFFFFFu, count = ((r >> 27) & 15u) + 1u;
                            for (uint32_t s = first; s < first + count && (size_t)(s + 1) * 12 <= slots.size(); ++s) {
                                uint32_t p;
                                std::memcpy(&p, &slots[12 * (size_t)s + 3], 4);
                                if (p < n) prims.push_back(p);
                            }
                        } else {
                            sub.push_back(r);
                        }
                    }
                }
                st.push_back({ref, it.depth + 1});
            }
            for (uint32_t p : prims)
                for (int v = 0; v < 3; ++v)
                    for (int a = 0; a < 3; ++a) {
                        const float x = m.v[3 * (size_t)m.f[3 * (size_t)p + v] + a];
                        if (x < blo[a] || x > bhi[a]) {
                            fail(tag + ": decoded child box misses a vertex of primitive " + std::to_string(p));
                            a = 3; v = 3;
                        }
                    }
        }
    }
    for (uint32_t p = 0; p < n; ++p)
        if (seen[p] != 1) { fail(tag + ": primitive " + std::to_string(p) + " reached " + std::to_string(seen[p]) + " times"); break; }
    if (deepest != depth) fail(tag + ": wide depth " + std::to_string(depth) + " vs deepest " + std::to_string(deepest));
}

} // namespace

int main() {
    std::vector<Mesh> meshes;
    meshes.push_back(soup(20000, 1.0f, 1.0f, 1, "soup"));
    meshes.push_back(coincident());
    meshes.push_back(grid(48));
    meshes.push_back(soup(1, 1.0f, 1.0f, 2, "one"));
    meshes.push_back(soup(5000, 1e-3f, 1e4f, 3, "scales"));
    const uint32_t leaf_sizes[] = {1, 3, 15};
    const uint32_t bin_counts[] = {4, 16, 64};
    const int depth_bounds[] = {4, 60};
    size_t builds = 0;
    for (const Mesh& m : meshes) {
        std::vector<float> lo, hi, cen;
        bounds(m, lo, hi, cen);
        float mag = 0.0f;
        for (float x : m.v) mag = std::fmax(mag, std::fabs(x));
        const float pad = kPadRel * mag;
        for (uint32_t ls : leaf_sizes)
            for (uint32_t bins : bin_counts)
                for (int db : depth_bounds) {
                    const std::string tag = m.name + " leaf " + std::to_string(ls) + " bins " + std::to_string(bins) +
                                            " depth " + std::to_string(db);
                    BuildResult br = build_bvh2(lo.data(), hi.data(), cen.data(), m.ntri(), ls, bins, db);
                    check_binary(br, m.ntri(), ls, tag);
                    {
                        LeafWriter lw{m.v, m.f, m.ntri(), {}};
                        std::vector<GpuNode8> n8;
                        std::vector<float> slots;
                        uint64_t leaves = 0;
                        int d = 0;
                        convert_bvh8(br, lw, pad, n8, slots, leaves, d);
                        check_wide<8>(m, n8, slots, d, tag + " (8-wide)");
                    }
                    {
                        LeafWriter lw{m.v, m.f, m.ntri(), {}};
                        std::vector<GpuNode4> n4;
                        std::vector<float> slots;
                        uint64_t leaves = 0;
                        int d = 0;
                        convert_bvh4(br, lw, pad, n4, slots, leaves, d);
                        check_wide<4>(m, n4, slots, d, tag + " (4-wide)");
                    }
                    ++builds;
                }
    }
    std::printf("%zu builds, %d failures\n", builds, failures);
    if (failures) return 1;
    std::printf("ok\n");
    return 0;
}
