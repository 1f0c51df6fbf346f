// Closest-hit and any-hit BVH traversal for gfx950.
//
// Replaces TraverseBVH<bAnyHit> / AnyHitSegmentBVH of the reference
// (core/render/intersection.cuh:148-310) for triangle meshes and the
// analytic box / sphere / SDF objects (objects.hpp).  Design:
//  * persistent grid with per-lane refill: each wave owns a private chunk
//    of 128 ray indices taken with one atomic, and hands the next index of
//    its chunk to every lane whose ray finished, so lanes stay busy despite
//    divergent ray lengths; a bounded step budget per refill round;
//  * default layout: the quantised 8-wide BVH, one 128-byte node (one L2
//    line) per interior node, children decoded with one FMA per plane and
//    stored along the node's largest axis so pushes need no sort; the
//    quantised 4-wide layout (64-byte nodes) is width 4 and the binary
//    layout (64-byte nodes holding both child boxes) width 2;
//  * per-lane traversal stack: the top kLdsStack entries in LDS (indexed
//    [entry][thread], conflict-free), deeper entries spill to an HBM buffer
//    laid out [entry][global thread];
//  * leaf deferral: lanes at a leaf wait until a third of the live lanes
//    are at one, so triangle tests run with more lanes active;
//  * watertight triangle test (Woop, Benthin, Wald 2013) compiled without
//    FMA contraction and with IEEE division, so hit ids, t and barycentrics
//    are bit-identical to a CPU evaluation of the same formula.
#include "ar_internal.hpp"
#include "objects.hpp"

#include <algorithm>
#include <cfloat>
#include <mutex>
#include <string>

namespace ar {
namespace {

// Closest hit pushes the hit children after the nearest in the node's slot
// order (children are stored along the node's largest axis) rather than
// sorting them far to near (PUSH_SLOT_ORDER 0): on the 1M-triangle workload
// the sort costs more VALU than its better culling saves.
#ifndef AR_TRACE_PUSH_SLOT_ORDER
#define AR_TRACE_PUSH_SLOT_ORDER 1
#endif
// Leaf deferral: a wave runs its leaf branch only when at least 1 / N of its
// live lanes sit at a leaf (0: every step).
// Hit children pushed by rank without a branch per child (Stack::push_mask)
// rather than one conditional push each.
#ifndef AR_TRACE_PUSH_MASK
#define AR_TRACE_PUSH_MASK 1
#endif
#ifndef AR_TRACE_LEAF_DEFER_CLOSEST
#define AR_TRACE_LEAF_DEFER_CLOSEST 3
#endif
#ifndef AR_TRACE_LEAF_DEFER_ANY
#define AR_TRACE_LEAF_DEFER_ANY 3
#endif

struct RayPre {
    float ox, oy, oz;
    float dx, dy, dz;      // raw direction (analytic-object tests only)
    float idx, idy, idz;   // 1/d, +-FLT_MAX for zero components
    float nox, noy, noz;   // -o/d
    int kx, ky, kz;
    float sx, sy, sz;
    float tpad;            // extra t padding for far-away origins
    bool valid;
};

// 1/x for the box tests, magnitude clamped to 2^100 (x = 0 gives +-2^100 by
// the sign of the zero).  Bounded so that products with the quantised grid
// scale and with (p - o) stay finite for coordinates within +-2^26: an
// infinite factor would turn 0 * inf into NaN.
__device__ __forceinline__ float inv_safe(float x) {
    const float big = 0x1p100f;
    if (x == 0.0f) return __builtin_signbit(x) ? -big : big;
    const float r = 1.0f / x;
    return fminf(fmaxf(r, -big), big);
}

__device__ __forceinline__ float sel3(int k, float x, float y, float z) {
    return k == 0 ? x : (k == 1 ? y : z);
}

__device__ __forceinline__ RayPre ray_setup(float ox, float oy, float oz, float dx, float dy, float dz,
                                            float origin_limit) {
    RayPre p;
    p.ox = ox; p.oy = oy; p.oz = oz;
    p.dx = dx; p.dy = dy; p.dz = dz;
    p.idx = inv_safe(dx); p.idy = inv_safe(dy); p.idz = inv_safe(dz);
    p.nox = -ox * p.idx; p.noy = -oy * p.idy; p.noz = -oz * p.idz;
    const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    int kz = 0;
    float m = ax;
    if (ay > m) { kz = 1; m = ay; }
    if (az > m) { kz = 2; }
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    const float dkz = sel3(kz, dx, dy, dz);
    if (dkz < 0.0f) { const int t = kx; kx = ky; ky = t; }
    p.kx = kx; p.ky = ky; p.kz = kz;
    p.valid = dkz != 0.0f;
    p.sx = sel3(kx, dx, dy, dz) / dkz;
    p.sy = sel3(ky, dx, dy, dz) / dkz;
    p.sz = 1.0f / dkz;
    const float omax = fmaxf(fabsf(ox), fmaxf(fabsf(oy), fabsf(oz)));
    const float imax = fmaxf(fabsf(p.idx), fmaxf(fabsf(p.idy), fabsf(p.idz)));
    p.tpad = omax > origin_limit ? omax * kPadRel * imax : 0.0f;
    return p;
}

// The origin limit of a ray whose traversal starts at ref: its tree's own
// (kept by the root node; in a batch the context's limit is the smallest of
// its scenes'), or the context's when the ray starts at a leaf.
template <int W>
__device__ __forceinline__ float root_origin_limit(const void* __restrict__ nodes, uint32_t ref, float ctx_limit) {
    if ((int)ref < 0) return ctx_limit;
    uint32_t bits;
    if constexpr (W == 8) bits = reinterpret_cast<const GpuNode8*>(nodes)[ref].pad[0];
    else if constexpr (W == 4) bits = reinterpret_cast<const GpuNode4*>(nodes)[ref].pad[0];
    else bits = reinterpret_cast<const GpuNode2*>(nodes)[ref].d[2];
    return fmaxf(ctx_limit, __uint_as_float(bits));
}

// Watertight test; returns true when the ray's line crosses the triangle and
// writes t.  With full = true also writes u, v (weights of v1, v2) and flags.
// The arithmetic is written to round exactly like the CPU formula: no
// contraction into FMA, IEEE division, double fallback on zero edges.
template <bool full>
__device__ __forceinline__ bool tri_test(const RayPre& p, float4 va, float4 vb, float4 vc,
                                         float& t, float* u = nullptr, float* v = nullptr,
                                         uint32_t* flags = nullptr) {
#pragma clang fp contract(off)
    const float A0 = va.x - p.ox, A1 = va.y - p.oy, A2 = va.z - p.oz;
    const float B0 = vb.x - p.ox, B1 = vb.y - p.oy, B2 = vb.z - p.oz;
    const float C0 = vc.x - p.ox, C1 = vc.y - p.oy, C2 = vc.z - p.oz;
    const float Akx = sel3(p.kx, A0, A1, A2), Aky = sel3(p.ky, A0, A1, A2), Akz = sel3(p.kz, A0, A1, A2);
    const float Bkx = sel3(p.kx, B0, B1, B2), Bky = sel3(p.ky, B0, B1, B2), Bkz = sel3(p.kz, B0, B1, B2);
    const float Ckx = sel3(p.kx, C0, C1, C2), Cky = sel3(p.ky, C0, C1, C2), Ckz = sel3(p.kz, C0, C1, C2);
    const float Ax = Akx - p.sx * Akz;
    const float Ay = Aky - p.sy * Akz;
    const float Bx = Bkx - p.sx * Bkz;
    const float By = Bky - p.sy * Bkz;
    const float Cx = Ckx - p.sx * Ckz;
    const float Cy = Cky - p.sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        const double CxBy = (double)Cx * (double)By;
        const double CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        const double AxCy = (double)Ax * (double)Cy;
        const double AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        const double BxAy = (double)Bx * (double)Ay;
        const double ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = p.sz * Akz;
    const float Bz = p.sz * Bkz;
    const float Cz = p.sz * Ckz;
    const float T = U * Az + V * Bz + W * Cz;
    t = T / det;
    if constexpr (full) {
        *u = V / det;
        *v = W / det;
        uint32_t f = 0;
        if (U == 0.0f) f |= AR_HIT_W0_ZERO;
        if (V == 0.0f) f |= AR_HIT_W1_ZERO;
        if (W == 0.0f) f |= AR_HIT_W2_ZERO;
        if (det < 0.0f) f |= AR_HIT_BACKFACE;
        *flags = f;
    }
    return true;
}

// Slab test of one child box (FMA form, boxes are pre-padded).
__device__ __forceinline__ bool box_test(const RayPre& p, float lx, float hx, float ly, float hy,
                                         float lz, float hz, float tmin, float tmax, float& tnear) {
    const float x0 = fmaf(lx, p.idx, p.nox), x1 = fmaf(hx, p.idx, p.nox);
    const float y0 = fmaf(ly, p.idy, p.noy), y1 = fmaf(hy, p.idy, p.noy);
    const float z0 = fmaf(lz, p.idz, p.noz), z1 = fmaf(hz, p.idz, p.noz);
    const float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin)) - p.tpad;
    const float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax)) + p.tpad;
    tnear = tn;
    return !(tn > tf);   // NaN counts as a hit (conservative)
}

struct Stack {
    int* lds;        // this lane's column: lds[e * kTraceBlock]
    int* spill;      // this thread's column: spill[e * threads]
    uint32_t threads;
    int sp;
    __device__ __forceinline__ void push(uint32_t v) {
        if (sp < kLdsStack) lds[sp * kTraceBlock] = (int)v;
        else spill[(size_t)(sp - kLdsStack) * threads] = (int)v;
        ++sp;
    }
    __device__ __forceinline__ uint32_t pop() {
        --sp;
        return (uint32_t)(sp < kLdsStack ? lds[sp * kTraceBlock] : spill[(size_t)(sp - kLdsStack) * threads]);
    }
    // Pushes refs[c] for every set bit c of mask: in slot order 0..N-1 when
    // up is false (slot N-1 pops first), else N-1..0 (slot 0 pops first).
    // When all of them fit in LDS every child is written without a branch, to
    // its rank's entry or, if not set, to the lane's spare entry kLdsStack.
    template <int N>
    __device__ __forceinline__ void push_mask(const uint32_t (&refs)[N], uint32_t mask, bool up) {
        const int n = __popc(mask);
        if (sp + n <= kLdsStack) {
            int rank = 0;
            const int top = sp + n - 1;
#pragma unroll
            for (int c = 0; c < N; ++c) {
                const bool hit = (mask >> c) & 1u;
                const int pos = up ? top - rank : sp + rank;
                lds[(hit ? pos : kLdsStack) * kTraceBlock] = (int)refs[c];
                rank += hit ? 1 : 0;
            }
            sp += n;
        } else if (up) {
#pragma unroll
            for (int c = N - 1; c >= 0; --c)
                if ((mask >> c) & 1u) push(refs[c]);
        } else {
#pragma unroll
            for (int c = 0; c < N; ++c)
                if ((mask >> c) & 1u) push(refs[c]);
        }
    }
};

// Per-lane traversal state, kept in registers across refill rounds.
struct Trav {
    RayPre p;
    float tmin, tmax, best_t;
    float meps;            // SDF march tolerance (object scenes)
    uint32_t ref, best_p, best_slot, idx;
    bool have;
    bool best_in;          // the best hit is an object the ray started inside
};

__device__ __forceinline__ float byte_f(uint32_t w, int k) { return (float)((w >> (8 * k)) & 0xFFu); }

// Interior node of the binary layout: both child boxes from one 64 B node;
// the nearer hit child is next, the farther one is pushed.
__device__ __forceinline__ void interior2(const GpuNode2* __restrict__ nodes, Trav& s, Stack& st) {
    const GpuNode2* nd = nodes + s.ref;
    const float4 a = *reinterpret_cast<const float4*>(nd->a);
    const float4 b = *reinterpret_cast<const float4*>(nd->b);
    const float4 c = *reinterpret_cast<const float4*>(nd->c);
    const uint4 d = *reinterpret_cast<const uint4*>(nd->d);
    float t0, t1;
    const bool h0 = box_test(s.p, a.x, a.y, a.z, a.w, c.x, c.y, s.tmin, s.best_t, t0);
    const bool h1 = box_test(s.p, b.x, b.y, b.z, b.w, c.z, c.w, s.tmin, s.best_t, t1);
    if (h0 && h1) {
        uint32_t nr = d.x, fr = d.y;
        if (t1 < t0) { nr = d.y; fr = d.x; }
        st.push(fr);
        s.ref = nr;
    } else if (h0) {
        s.ref = d.x;
    } else if (h1) {
        s.ref = d.y;
    } else {
        s.ref = st.pop();
    }
}

// Interior node of the quantised 8-wide layout.  Hit children's entry
// distances stay in registers (static indices only); the nearest is next and
// the rest are pushed far-to-near (closest hit) or in slot order (any hit).
template <bool ANY>
__device__ __forceinline__ uint32_t interior8(const GpuNode8* __restrict__ nodes, Trav& s, Stack& st) {
    const float4* nd = reinterpret_cast<const float4*>(nodes + s.ref);
    const float4 h = nd[0];
    const uint4 r0 = *reinterpret_cast<const uint4*>(nd + 1);
    const uint4 r1 = *reinterpret_cast<const uint4*>(nd + 2);
    const uint4 q0 = *reinterpret_cast<const uint4*>(nd + 3);
    const uint4 q1 = *reinterpret_cast<const uint4*>(nd + 4);
    const uint4 q2 = *reinterpret_cast<const uint4*>(nd + 5);
    const uint32_t meta = __float_as_uint(h.w);
    const uint32_t nchild = (meta >> 24) & 15u;
    const RayPre& p = s.p;
    const float ax = __uint_as_float((meta & 0xFFu) << 23) * p.idx;
    const float ay = __uint_as_float(((meta >> 8) & 0xFFu) << 23) * p.idy;
    const float az = __uint_as_float(((meta >> 16) & 0xFFu) << 23) * p.idz;
    const float bx = (h.x - p.ox) * p.idx;
    const float by = (h.y - p.oy) * p.idy;
    const float bz = (h.z - p.oz) * p.idz;
    const uint32_t refs[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
    float tn[8];
    uint32_t mask = 0;
    // Near and far planes picked once per axis by the sign of the ray's
    // direction (the quantised lo byte is <= the hi byte, and the scale has
    // the sign of 1/d), so each child needs no min / max per axis.
    const bool nxs = p.idx < 0.0f, nys = p.idy < 0.0f, nzs = p.idz < 0.0f;
    const uint32_t nearx[2] = {nxs ? q1.z : q0.x, nxs ? q1.w : q0.y};
    const uint32_t farx[2] = {nxs ? q0.x : q1.z, nxs ? q0.y : q1.w};
    const uint32_t neary[2] = {nys ? q2.x : q0.z, nys ? q2.y : q0.w};
    const uint32_t fary[2] = {nys ? q0.z : q2.x, nys ? q0.w : q2.y};
    const uint32_t nearz[2] = {nzs ? q2.z : q1.x, nzs ? q2.w : q1.y};
    const uint32_t farz[2] = {nzs ? q1.x : q2.z, nzs ? q1.y : q2.w};
    const float lo_t = s.tmin, hi_t = s.best_t;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int w = c >> 2, k = c & 3;
        const float xn = fmaf(byte_f(nearx[w], k), ax, bx), xf = fmaf(byte_f(farx[w], k), ax, bx);
        const float yn = fmaf(byte_f(neary[w], k), ay, by), yf = fmaf(byte_f(fary[w], k), ay, by);
        const float zn = fmaf(byte_f(nearz[w], k), az, bz), zf = fmaf(byte_f(farz[w], k), az, bz);
        const float a = fmaxf(fmaxf(xn, yn), fmaxf(zn, lo_t)) - p.tpad;
        const float b = fminf(fminf(xf, yf), fminf(zf, hi_t)) + p.tpad;
        tn[c] = a;
        if (!(a > b)) mask |= 1u << c;   // NaN counts as a hit
    }
    mask &= (1u << nchild) - 1u;
    if (mask == 0) {
        s.ref = st.pop();
        return nchild;
    }
    if (ANY) {
#if AR_TRACE_PUSH_MASK
        // the lowest hit slot next, the others pushed so that the highest pops first
        const int ni = __ffs(mask) - 1;
        uint32_t next = refs[0];
#pragma unroll
        for (int c = 1; c < 8; ++c)
            if (c == ni) next = refs[c];
        st.push_mask(refs, mask & (mask - 1u), false);
#else
        uint32_t next = kRefEmpty;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if ((mask >> c) & 1u) {
                if (next == kRefEmpty) next = refs[c];
                else st.push(refs[c]);
            }
        }
#endif
        s.ref = next;
    } else {
        float nt = INFINITY;
        uint32_t next = kRefEmpty;
        int ni = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (((mask >> c) & 1u) && tn[c] < nt) { nt = tn[c]; next = refs[c]; ni = c; }
        mask &= ~(1u << ni);
#if AR_TRACE_PUSH_SLOT_ORDER
        // the rest without a sort: children are stored along the node's
        // largest axis, so popping them in slot order (ray going up that
        // axis) or reverse slot order (going down) is near to far
        const uint32_t paxis = (meta >> 28) & 3u;
        const float pdir = paxis == 0u ? p.idx : (paxis == 1u ? p.idy : p.idz);
#if AR_TRACE_PUSH_MASK
        st.push_mask(refs, mask, !(pdir < 0.0f));
#else
        if (pdir < 0.0f) {
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if ((mask >> c) & 1u) st.push(refs[c]);
        } else {
#pragma unroll
            for (int c = 7; c >= 0; --c)
                if ((mask >> c) & 1u) st.push(refs[c]);
        }
#endif
#else
        while (mask) {
            float ft = -INFINITY;
            uint32_t fr = 0;
            int fi = 0;
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (((mask >> c) & 1u) && tn[c] >= ft) { ft = tn[c]; fr = refs[c]; fi = c; }
            st.push(fr);
            mask &= ~(1u << fi);
        }
#endif
        s.ref = next;
    }
    return nchild;
}

// Interior node of the quantised 4-wide layout: four child boxes from one
// 64 B node, otherwise as interior8.
template <bool ANY>
__device__ __forceinline__ uint32_t interior4(const GpuNode4* __restrict__ nodes, Trav& s, Stack& st) {
    const float4* nd = reinterpret_cast<const float4*>(nodes + s.ref);
    const float4 h = nd[0];
    const uint4 r = *reinterpret_cast<const uint4*>(nd + 1);
    const uint4 q0 = *reinterpret_cast<const uint4*>(nd + 2);
    const uint4 q1 = *reinterpret_cast<const uint4*>(nd + 3);
    const uint32_t meta = __float_as_uint(h.w);
    const uint32_t nchild = (meta >> 24) & 15u;
    const RayPre& p = s.p;
    const float ax = __uint_as_float((meta & 0xFFu) << 23) * p.idx;
    const float ay = __uint_as_float(((meta >> 8) & 0xFFu) << 23) * p.idy;
    const float az = __uint_as_float(((meta >> 16) & 0xFFu) << 23) * p.idz;
    const float bx = (h.x - p.ox) * p.idx;
    const float by = (h.y - p.oy) * p.idy;
    const float bz = (h.z - p.oz) * p.idz;
    const uint32_t refs[4] = {r.x, r.y, r.z, r.w};
    float tn[4];
    uint32_t mask = 0;
    const bool nxs = p.idx < 0.0f, nys = p.idy < 0.0f, nzs = p.idz < 0.0f;
    const uint32_t nearx = nxs ? q0.w : q0.x, farx = nxs ? q0.x : q0.w;
    const uint32_t neary = nys ? q1.x : q0.y, fary = nys ? q0.y : q1.x;
    const uint32_t nearz = nzs ? q1.y : q0.z, farz = nzs ? q0.z : q1.y;
    const float lo_t = s.tmin, hi_t = s.best_t;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float xn = fmaf(byte_f(nearx, c), ax, bx), xf = fmaf(byte_f(farx, c), ax, bx);
        const float yn = fmaf(byte_f(neary, c), ay, by), yf = fmaf(byte_f(fary, c), ay, by);
        const float zn = fmaf(byte_f(nearz, c), az, bz), zf = fmaf(byte_f(farz, c), az, bz);
        const float a = fmaxf(fmaxf(xn, yn), fmaxf(zn, lo_t)) - p.tpad;
        const float b = fminf(fminf(xf, yf), fminf(zf, hi_t)) + p.tpad;
        tn[c] = a;
        if (!(a > b)) mask |= 1u << c;   // NaN counts as a hit
    }
    mask &= (1u << nchild) - 1u;
    if (mask == 0) {
        s.ref = st.pop();
        return nchild;
    }
    if (ANY) {
        const int ni = __ffs(mask) - 1;
        uint32_t next = refs[0];
#pragma unroll
        for (int c = 1; c < 4; ++c)
            if (c == ni) next = refs[c];
        st.push_mask(refs, mask & (mask - 1u), false);
        s.ref = next;
    } else {
        float nt = INFINITY;
        uint32_t next = kRefEmpty;
        int ni = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (((mask >> c) & 1u) && tn[c] < nt) { nt = tn[c]; next = refs[c]; ni = c; }
        mask &= ~(1u << ni);
        const uint32_t paxis = (meta >> 28) & 3u;
        const float pdir = paxis == 0u ? p.idx : (paxis == 1u ? p.idy : p.idz);
        st.push_mask(refs, mask, !(pdir < 0.0f));
        s.ref = next;
    }
    return nchild;
}

// Leaf: every primitive in it, hit rule "smallest t, then smallest id".
// Any hit ends the traversal on its first accepted primitive.  OBJ: the
// scene has analytic objects, whose slots are tested by the object
// intersector (prim id = obj_base + object index; vol = SDF voxel pool).
template <bool ANY, bool OBJ>
__device__ __forceinline__ uint32_t leaf_prims(const float4* __restrict__ tris, const float4* __restrict__ objs,
                                               uint32_t obj_base, const float* __restrict__ vol, Trav& s,
                                               uint32_t lref) {
    const uint32_t first = lref & kMaxTriSlots;
    const uint32_t cnt = ((lref >> 27) & 15u) + 1u;
    for (uint32_t k = 0; k < cnt; ++k) {
        const float4* tp = tris + 3 * (size_t)(first + k);
        const float4 va = tp[0];
        uint32_t prim = __float_as_uint(va.w);
        float t;
        bool inside = false;
        if (OBJ && (prim & kObjSlot)) {
            const uint32_t ok = prim & ~kObjSlot;
            if (!obj::test(objs + (size_t)obj::kRecord * ok, s.p.ox, s.p.oy, s.p.oz, s.p.dx, s.p.dy, s.p.dz, s.meps,
                           vol, t, inside))
                continue;
            prim = obj_base + ok;
        } else {
            const float4 vb = tp[1], vc = tp[2];
            if (!tri_test<false>(s.p, va, vb, vc, t)) continue;
        }
        const bool accept = t >= s.tmin &&
            (s.have ? (t < s.best_t || (t == s.best_t && prim < s.best_p)) : t < s.tmax);
        if (accept) {
            s.have = true;
            s.best_t = t;
            s.best_p = prim;
            s.best_slot = first + k;
            if (OBJ) s.best_in = inside;
            if (ANY) { s.ref = kRefEmpty; return k + 1; }
        }
    }
    return cnt;
}

template <bool ANY, bool OBJ>
__device__ __forceinline__ uint32_t leaf(const float4* __restrict__ tris, const float4* __restrict__ objs,
                                         uint32_t obj_base, const float* __restrict__ vol, Trav& s, Stack& st) {
    const uint32_t c = leaf_prims<ANY, OBJ>(tris, objs, obj_base, vol, s, s.ref);
    if (!ANY || s.ref != kRefEmpty) s.ref = st.pop();
    return c;
}

// traversal steps between two ray-refill rounds: closest-hit rays are longer,
// so more steps per round pay for the refill; any-hit rays end early.  A
// 4-wide step is about half an 8-wide one and a ray takes about twice as
// many, so its rounds are longer (measured on the 1M-triangle traversal
// workload: 12 / 10 against 6 / 4, 4% less time for both launches)
#ifndef AR_TRACE_STEP_BUDGET_CLOSEST
#define AR_TRACE_STEP_BUDGET_CLOSEST 6
#endif
#ifndef AR_TRACE_STEP_BUDGET_ANY
#define AR_TRACE_STEP_BUDGET_ANY 4
#endif
#ifndef AR_TRACE_STEP_BUDGET_CLOSEST4
#define AR_TRACE_STEP_BUDGET_CLOSEST4 12
#endif
#ifndef AR_TRACE_STEP_BUDGET_ANY4
#define AR_TRACE_STEP_BUDGET_ANY4 10
#endif
#ifndef AR_TRACE_CHUNK
#define AR_TRACE_CHUNK 128
#endif

// Persistent traversal with per-lane ray refill (Aila & Laine, HPG'09): a
// lane whose ray is finished takes the next ray index from a global counter
// (one atomic per wave and round), so a wave never idles on its slowest ray.
// ANY: sets bit idx of occ (zeroed beforehand) for occluded rays.
#ifndef AR_TRACE_WAVES_CLOSEST
#define AR_TRACE_WAVES_CLOSEST 7
#endif
#ifndef AR_TRACE_WAVES_ANY
#define AR_TRACE_WAVES_ANY 7
#endif
// kernels with analytic and SDF objects hold more state: at 5 waves per SIMD
// they keep it in registers (at 6 or 7 they spill 70-100 bytes per lane)
#ifndef AR_TRACE_WAVES_OBJ
#define AR_TRACE_WAVES_OBJ 5
#endif

template <int W, bool ANY, bool COUNT, bool OBJ>
__global__ __launch_bounds__(kTraceBlock)
__attribute__((amdgpu_waves_per_eu(OBJ ? AR_TRACE_WAVES_OBJ
                                       : ANY ? AR_TRACE_WAVES_ANY : COUNT ? 6 : AR_TRACE_WAVES_CLOSEST,
                                   8))) void k_trace(
    const void* __restrict__ nodes, const float4* __restrict__ tris, uint32_t root,
    const uint32_t* __restrict__ roots, float origin_limit, const float4* __restrict__ objs, uint32_t obj_base,
    const float* __restrict__ vol, ar_rays r, size_t n, const uint32_t* __restrict__ n_dev, ar_hits h, uint32_t* __restrict__ occ,
    int* __restrict__ spill, uint32_t threads, uint32_t* __restrict__ cnt_box, uint32_t* __restrict__ cnt_tri,
    uint32_t* __restrict__ fetch) {
    __shared__ int lds_stack[(kLdsStack + 1) * kTraceBlock];   // + one spare entry per lane (push_mask)
    if (n_dev) n = *n_dev;
    const uint32_t gtid = blockIdx.x * kTraceBlock + threadIdx.x;
    const uint32_t lane = __lane_id();
    Stack st;
    st.lds = lds_stack + threadIdx.x;
    st.spill = spill + gtid;
    st.threads = threads;
    st.sp = 0;
    Trav s;
    s.ref = kRefEmpty;
    bool active = false, exhausted = false;
    uint32_t nb = 0, nt = 0;
#if AR_TRACE_CHUNK
    uint32_t pool = 0, pool_end = 0;   // this wave's unclaimed ray indices (wave-uniform)
#endif
    while (true) {
        const bool need = !active && !exhausted;
        const uint64_t nm = __ballot(need);
        if (nm) {
#if AR_TRACE_CHUNK
            // rays come from a wave-private chunk; one atomic per chunk
            const uint32_t cnt = (uint32_t)__popcll(nm);
            const uint32_t left = pool_end - pool;
            uint32_t base2 = 0;
            if (left < cnt) {
                const uint32_t leader = (uint32_t)__ffsll((unsigned long long)nm) - 1u;
                if (lane == leader) base2 = atomicAdd(fetch, (uint32_t)AR_TRACE_CHUNK);
                base2 = __shfl(base2, (int)leader);
            }
            const uint32_t rk = (uint32_t)__popcll(nm & ((1ull << lane) - 1ull));
            const uint32_t pool0 = pool;
            if (left < cnt) { pool = base2 + (cnt - left); pool_end = base2 + (uint32_t)AR_TRACE_CHUNK; }
            else pool += cnt;
            if (need) {
                const uint32_t idx = rk < left ? pool0 + rk : base2 + (rk - left);
#else
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)nm) - 1u;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(fetch, (uint32_t)__popcll(nm));
            base = __shfl(base, (int)leader);
            if (need) {
                const uint32_t idx = base + (uint32_t)__popcll(nm & ((1ull << lane) - 1ull));
#endif
                if (idx < n) {
                    s.idx = idx;
                    const uint32_t rt = roots ? roots[idx] : root;
                    s.p = ray_setup(r.ox[idx], r.oy[idx], r.oz[idx], r.dx[idx], r.dy[idx], r.dz[idx],
                                    root_origin_limit<W>(nodes, rt, origin_limit));
                    s.tmin = r.tmin[idx];
                    s.tmax = r.tmax[idx];
                    s.best_t = s.tmax;
                    s.best_p = kRefEmpty;
                    s.best_slot = kRefEmpty;
                    s.have = false;
                    if (OBJ) {
                        s.best_in = false;
                        s.meps = r.meps ? r.meps[idx] : 0.5f * fmaxf(s.tmin, 1e-3f);
                    }
                    st.sp = 0;
                    st.push(kRefEmpty);
                    s.ref = (s.p.valid && s.tmax >= s.tmin) ? rt : kRefEmpty;
                    nb = 0;
                    nt = 0;
                    active = true;
                } else {
                    exhausted = true;
                }
            }
        }
        if (__ballot(active) == 0) break;
        if (active) {
            // Lanes that reach a leaf wait while the wave has few of them, so
            // the wave does not run the leaf branch for one or two lanes every
            // step; leaves run once at least 1 / kDefer of the live lanes are
            // at one (always when all are).  kDefer = 0: no waiting.
            constexpr uint32_t kDefer = ANY ? AR_TRACE_LEAF_DEFER_ANY : AR_TRACE_LEAF_DEFER_CLOSEST;
            constexpr int kStepBudget = W == 4 ? (ANY ? AR_TRACE_STEP_BUDGET_ANY4 : AR_TRACE_STEP_BUDGET_CLOSEST4)
                                               : (ANY ? AR_TRACE_STEP_BUDGET_ANY : AR_TRACE_STEP_BUDGET_CLOSEST);
            for (int it = 0; it < kStepBudget; ++it) {
                const bool live = s.ref != kRefEmpty;
                const bool at_leaf = live && (int)s.ref < 0;
                const uint64_t am = __ballot(live);
                if (am == 0) break;
                bool run_leaves = true;
                if constexpr (kDefer > 0) {
                    const uint64_t lm = __ballot(at_leaf);
                    run_leaves = (uint32_t)__popcll(lm) * kDefer >= (uint32_t)__popcll(am);
                }
                if (live && !at_leaf) {
                    if constexpr (W == 8) {
                        const uint32_t c = interior8<ANY>(reinterpret_cast<const GpuNode8*>(nodes), s, st);
                        if (COUNT) nb += c;
                    } else if constexpr (W == 4) {
                        const uint32_t c = interior4<ANY>(reinterpret_cast<const GpuNode4*>(nodes), s, st);
                        if (COUNT) nb += c;
                    } else {
                        interior2(reinterpret_cast<const GpuNode2*>(nodes), s, st);
                        if (COUNT) nb += 2;
                    }
                } else if (at_leaf && run_leaves) {
                    const uint32_t c = leaf<ANY, OBJ>(tris, objs, obj_base, vol, s, st);
                    if (COUNT) nt += c;
                }
            }
            if (s.ref == kRefEmpty) {
                const uint32_t i = s.idx;
                if constexpr (ANY) {
                    if (s.have) atomicOr(&occ[i >> 5], 1u << (i & 31u));
                    if (COUNT) { cnt_box[i] = nb; cnt_tri[i] = nt; }
                } else {
                    if (!s.have) {
                        h.prim[i] = AR_MISS;
                        h.t[i] = __int_as_float(0x7f800000);
                        h.u[i] = 0.0f; h.v[i] = 0.0f; h.flags[i] = 0u;
                    } else {
                        const float4* tp = tris + 3 * (size_t)s.best_slot;
                        const float4 va = tp[0];
                        float tt, u = 0.0f, v = 0.0f;
                        uint32_t flags = 0;
                        if (OBJ && (__float_as_uint(va.w) & kObjSlot)) {
                            flags = s.best_in ? AR_HIT_INSIDE : 0u;
                        } else {
                            const float4 vb = tp[1], vc = tp[2];
                            tri_test<true>(s.p, va, vb, vc, tt, &u, &v, &flags);
                        }
                        h.prim[i] = s.best_p;
                        h.t[i] = s.best_t;
                        h.u[i] = u; h.v[i] = v; h.flags[i] = flags;
                    }
                    if (COUNT) { cnt_box[i] = nb; cnt_tri[i] = nt; }
                }
                active = false;
            }
        }
    }
}

// Clears the refill counter and, for any-hit, the occlusion words of the
// rays about to be traced (n read from n_dev when given).
__global__ void k_trace_prepare(uint32_t* __restrict__ fetch, uint32_t* __restrict__ occ, size_t n,
                                const uint32_t* __restrict__ n_dev, size_t cap_words) {
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n_dev) n = *n_dev;
    if (w == 0) *fetch = 0u;
    if (occ && w < cap_words && w < (n + 31) / 32) occ[w] = 0u;
}

// Shadow segment from each closest hit toward a point (the traversal
// benchmark's any-hit workload).
__global__ void k_shadow_rays(ar_rays r, ar_hits h, size_t n, float tx, float ty, float tz,
                              float tmin, float tmax_scale, ar_rays out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float* ox = const_cast<float*>(out.ox);
    float* oy = const_cast<float*>(out.oy);
    float* oz = const_cast<float*>(out.oz);
    float* otmin = const_cast<float*>(out.tmin);
    float* dxo = const_cast<float*>(out.dx);
    float* dyo = const_cast<float*>(out.dy);
    float* dzo = const_cast<float*>(out.dz);
    float* otmax = const_cast<float*>(out.tmax);
    const uint32_t prim = h.prim[i];
    float px = 0.0f, py = 0.0f, pz = 0.0f, dx = 0.0f, dy = 0.0f, dz = 1.0f, tm = 0.0f;
    if (prim != AR_MISS) {
        const float t = h.t[i];
        px = r.ox[i] + t * r.dx[i];
        py = r.oy[i] + t * r.dy[i];
        pz = r.oz[i] + t * r.dz[i];
        const float ex = tx - px, ey = ty - py, ez = tz - pz;
        const float dist = sqrtf(ex * ex + ey * ey + ez * ez);
        if (dist > 0.0f) {
            const float inv = 1.0f / dist;
            dx = ex * inv; dy = ey * inv; dz = ez * inv;
            tm = dist * (1.0f - tmax_scale);
        }
    }
    ox[i] = px; oy[i] = py; oz[i] = pz; otmin[i] = tmin;
    dxo[i] = dx; dyo[i] = dy; dzo[i] = dz; otmax[i] = tm;
}

} // namespace

// Resident blocks per CU of a traversal kernel (0: the largest over every
// variant, used to size the stack spill buffer).
int trace_blocks_per_cu(int width, bool any) {
    int b = 0;
    hipError_t e;
    if (width == 8) e = any ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_trace<8, true, false, false>, kTraceBlock, 0)
                            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_trace<8, false, false, false>, kTraceBlock, 0);
    else if (width == 4) e = any ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_trace<4, true, false, false>, kTraceBlock, 0)
                                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_trace<4, false, false, false>, kTraceBlock, 0);
    else if (width == 2) e = any ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_trace<2, true, false, false>, kTraceBlock, 0)
                                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_trace<2, false, false, false>, kTraceBlock, 0);
    else {
        int m = 1;
        for (int w : {2, 4, 8})
            for (bool a : {false, true}) m = std::max(m, trace_blocks_per_cu(w, a));
        return m;
    }
    return (e == hipSuccess && b > 0) ? b : 1;
}

static hipError_t prepare(const TraceArgs& ta, uint32_t* occ, size_t n, const uint32_t* n_dev, hipStream_t s) {
    const size_t words = (n + 31) / 32;
    const size_t grid = std::max<size_t>(1, (words + 255) / 256);
    hipLaunchKernelGGL(k_trace_prepare, dim3((uint32_t)grid), dim3(256), 0, s, ta.fetch, occ, n, n_dev, words);
    return hipGetLastError();
}

hipError_t launch_trace_closest(const TraceArgs& ta, const ar_rays& r, size_t n, const ar_hits& h,
                                uint32_t* cnt_box, uint32_t* cnt_tri, uint32_t blocks, hipStream_t s,
                                const uint32_t* n_dev) {
    hipError_t e = prepare(ta, nullptr, n, n_dev, s);
    if (e != hipSuccess) return e;
    const void* nd = ta.nodes;
    blocks = std::max(1u, std::min(blocks, ta.full_closest));
#define AR_LAUNCH_CLOSEST(W, C, O)                                                                               \
    hipLaunchKernelGGL((k_trace<W, false, C, O>), dim3(blocks), dim3(kTraceBlock), 0, s, nd, ta.tris, ta.root,  \
                       ta.roots, ta.origin_limit, ta.objs, ta.obj_base, ta.vol, r, n, n_dev, h, nullptr, ta.spill, \
                       ta.threads, cnt_box, cnt_tri, ta.fetch)
    const bool ob = ta.objs != nullptr;
    if (ta.width == 8) {
        if (cnt_box) { if (ob) AR_LAUNCH_CLOSEST(8, true, true); else AR_LAUNCH_CLOSEST(8, true, false); }
        else { if (ob) AR_LAUNCH_CLOSEST(8, false, true); else AR_LAUNCH_CLOSEST(8, false, false); }
    } else if (ta.width == 4) {
        if (cnt_box) { if (ob) AR_LAUNCH_CLOSEST(4, true, true); else AR_LAUNCH_CLOSEST(4, true, false); }
        else { if (ob) AR_LAUNCH_CLOSEST(4, false, true); else AR_LAUNCH_CLOSEST(4, false, false); }
    } else {
        if (cnt_box) { if (ob) AR_LAUNCH_CLOSEST(2, true, true); else AR_LAUNCH_CLOSEST(2, true, false); }
        else { if (ob) AR_LAUNCH_CLOSEST(2, false, true); else AR_LAUNCH_CLOSEST(2, false, false); }
    }
#undef AR_LAUNCH_CLOSEST
    return hipGetLastError();
}

hipError_t launch_trace_any(const TraceArgs& ta, const ar_rays& r, size_t n, uint32_t* occ, uint32_t blocks,
                            hipStream_t s, const uint32_t* n_dev, uint32_t* cnt_box, uint32_t* cnt_tri) {
    hipError_t e = prepare(ta, occ, n, n_dev, s);
    if (e != hipSuccess) return e;
    const void* nd = ta.nodes;
    const ar_hits none{nullptr, nullptr, nullptr, nullptr, nullptr};
    blocks = std::max(1u, std::min(blocks, ta.full_any));
#define AR_LAUNCH_ANY(W, C, O)                                                                                   \
    hipLaunchKernelGGL((k_trace<W, true, C, O>), dim3(blocks), dim3(kTraceBlock), 0, s, nd, ta.tris, ta.root,     \
                       ta.roots, ta.origin_limit, ta.objs, ta.obj_base, ta.vol, r, n, n_dev, none, occ, ta.spill, \
                       ta.threads, cnt_box, cnt_tri, ta.fetch)
#define AR_LAUNCH_ANY_W(W)                                                                                       \
    if (cnt_box) { if (ob) AR_LAUNCH_ANY(W, true, true); else AR_LAUNCH_ANY(W, true, false); }                   \
    else { if (ob) AR_LAUNCH_ANY(W, false, true); else AR_LAUNCH_ANY(W, false, false); }
    const bool ob = ta.objs != nullptr;
    if (ta.width == 8) { AR_LAUNCH_ANY_W(8) }
    else if (ta.width == 4) { AR_LAUNCH_ANY_W(4) }
    else { AR_LAUNCH_ANY_W(2) }
#undef AR_LAUNCH_ANY_W
#undef AR_LAUNCH_ANY
    return hipGetLastError();
}

hipError_t launch_shadow_rays(const ar_rays& r, const ar_hits& h, size_t n, float tx, float ty, float tz,
                              float tmin, float tmax_scale, const ar_rays& out, hipStream_t s) {
    const uint32_t bs = 256;
    const size_t blocks = (n + bs - 1) / bs;
    hipLaunchKernelGGL(k_shadow_rays, dim3((uint32_t)blocks), dim3(bs), 0, s, r, h, n, tx, ty, tz, tmin, tmax_scale, out);
    return hipGetLastError();
}

// ---- closest hit without allocation ----------------------------------------
// The traversal's scratch (the 32-bit ray-refill counter and the per-thread
// stack spill) lives in this fixed array instead of context buffers, for
// callers that cannot allocate.  Layout: the counter in bytes 0..255, then
// the spill ints [entry][thread] for the launched grid.  Each device holds its
// own copy for the life of the process; every call on a device shares it.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr size_t kNoAllocHeader = 256;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};
hipError_t g_scratch_err[kNoAllocMaxDevices] = {};
bool g_scratch_done[kNoAllocMaxDevices] = {};
} // namespace

// Resolves the device address of `scratch` on `device`, which must be the
// current device; once per device (later calls return the first result).
hipError_t noalloc_scratch_setup(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_done[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        g_scratch_ptr[device] = e == hipSuccess ? ptr : nullptr;
        g_scratch_err[device] = e;
        g_scratch_done[device] = true;
    }
    return g_scratch_err[device];
}

static void* noalloc_scratch(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_done[device] ? g_scratch_ptr[device] : nullptr;
}

} // namespace ar

using namespace ar;

extern "C" int ar_trace_closest_noalloc(ar_ctx* ctx, const ar_rays* rays, size_t n, const ar_hits* hits,
                                        void* stream) {
    if (!ctx) return AR_E_INVALID;
    if (n == 0) return AR_OK;
    if (!rays || !rays->ox || !rays->oy || !rays->oz || !rays->tmin || !rays->dx || !rays->dy || !rays->dz ||
        !rays->tmax || !hits || !hits->prim || !hits->t || !hits->u || !hits->v || !hits->flags)
        return set_error(ctx, AR_E_INVALID, "NULL ray or hit array");
    if (!ctx->bvh_ready) return set_error(ctx, AR_E_NOSCENE, "no BVH: call ar_scene_set and ar_bvh_build first");
    // ray indices come from the 32-bit refill counter, advanced in chunks
    if (n > (size_t(1) << 31))
        return set_error(ctx, AR_E_CAPACITY, "ar_trace_closest_noalloc: at most 2^31 rays per call");
    unsigned char* base = static_cast<unsigned char*>(noalloc_scratch(ctx->device));
    if (!base)
        return set_error(ctx, AR_E_DEVICE, "ar_trace_closest_noalloc: the scratch array's device address is unknown "
                                           "(hipGetSymbolAddress failed when the context was created)");
    const uint32_t stack = ctx->bvh.stack_entries;
    const size_t deep = stack > (uint32_t)kLdsStack ? stack - (uint32_t)kLdsStack : 1u;
    const size_t per_block = deep * kTraceBlock * sizeof(int);
    const size_t fit_blocks = (kNoAllocScratchBytes - kNoAllocHeader) / per_block;
    if (fit_blocks < (size_t)ctx->num_cus)
        return set_error(ctx, AR_E_CAPACITY,
                         "ar_trace_closest_noalloc: the BVH's stack (" + std::to_string(stack) +
                             " entries) needs " + std::to_string(per_block * (size_t)ctx->num_cus + kNoAllocHeader) +
                             " bytes of scratch for one block per CU; the scratch array holds " +
                             std::to_string(kNoAllocScratchBytes));
    TraceArgs ta;
    ta.width = ctx->bvh.width;
    ta.stack_entries = stack;
    ta.nodes = ctx->bvh.nodes;
    ta.tris = reinterpret_cast<const float4*>(ctx->bvh.tris);
    ta.root = ctx->bvh.root;
    ta.origin_limit = ctx->bvh.origin_limit;
    ta.objs = ctx->scene.nobj ? ctx->scene.objs : nullptr;
    // objects follow the triangles (a batch's merged triangles: scene.ntri)
    ta.obj_base = ctx->batch.empty() ? (uint32_t)(ctx->idx.size() / 3) : ctx->scene.ntri;
    ta.vol = ctx->sdf.dev;
    const size_t resident = (size_t)ctx->num_cus * (size_t)trace_blocks_per_cu((int)ctx->bvh.width, false);
    const size_t need = (n + kTraceBlock - 1) / kTraceBlock;
    const uint32_t blocks = (uint32_t)std::max<size_t>(1, std::min(std::min(resident, fit_blocks), need));
    ta.full_closest = blocks;
    ta.full_any = blocks;
    ta.threads = blocks * (uint32_t)kTraceBlock;   // spill stride = launched threads
    ta.fetch = reinterpret_cast<uint32_t*>(base);
    ta.spill = reinterpret_cast<int*>(base + kNoAllocHeader);
    const int rc = order_stream(ctx, reinterpret_cast<hipStream_t>(stream));
    if (rc != AR_OK) return rc;
    return check_hip(ctx, launch_trace_closest(ta, *rays, n, *hits, nullptr, nullptr, blocks,
                                               reinterpret_cast<hipStream_t>(stream)),
                     "trace_closest_noalloc launch");
}
