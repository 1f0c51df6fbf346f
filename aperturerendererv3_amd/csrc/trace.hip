// Closest-hit and any-hit BVH2 traversal for gfx950.
//
// Replaces TraverseBVH<bAnyHit> / AnyHitSegmentBVH of the reference
// (core/render/intersection.cuh:148-310) for triangle meshes.  Design:
//  * persistent grid: each wave walks 64 consecutive rays at a time, so ray
//    and hit arrays (SoA) are read and written fully coalesced;
//  * 64-byte nodes holding both child boxes: one node fetch tests two boxes;
//  * per-lane traversal stack: the top kLdsStack entries in LDS (indexed
//    [entry][thread], conflict-free), deeper entries spill to an HBM buffer
//    laid out [entry][global thread];
//  * watertight triangle test (Woop, Benthin, Wald 2013) compiled without
//    FMA contraction and with IEEE division, so hit ids, t and barycentrics
//    are bit-identical to a CPU evaluation of the same formula.
#include "ar_internal.hpp"

#include <cfloat>

namespace ar {
namespace {

struct RayPre {
    float ox, oy, oz;
    float idx, idy, idz;   // 1/d, +-FLT_MAX for zero components
    float nox, noy, noz;   // -o/d
    int kx, ky, kz;
    float sx, sy, sz;
    float tpad;            // extra t padding for far-away origins
    bool valid;
};

__device__ __forceinline__ float inv_safe(float x) {
    return (x == 0.0f) ? (x >= 0.0f ? FLT_MAX : -FLT_MAX) : 1.0f / x;
}

__device__ __forceinline__ float sel3(int k, float x, float y, float z) {
    return k == 0 ? x : (k == 1 ? y : z);
}

__device__ __forceinline__ RayPre ray_setup(float ox, float oy, float oz, float dx, float dy, float dz,
                                            float origin_limit) {
    RayPre p;
    p.ox = ox; p.oy = oy; p.oz = oz;
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
    return tn <= tf;
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
};

// Traverses one ray.  ANY: stop at the first accepted hit.  Returns the leaf
// slot of the hit triangle (kRefEmpty on a miss); its t goes to best_t.
template <bool ANY, bool COUNT>
__device__ __forceinline__ uint32_t traverse(const GpuNode2* __restrict__ nodes, const float4* __restrict__ tris,
                                             uint32_t root, const RayPre& p, float tmin, float tmax, Stack& st,
                                             float& best_t, uint32_t& nbox, uint32_t& ntri) {
    uint32_t best_p = kRefEmpty;
    uint32_t best_slot = kRefEmpty;
    bool have = false;
    best_t = tmax;
    if (!p.valid || !(tmax >= tmin) || root == kRefEmpty) return kRefEmpty;
    st.sp = 0;
    st.push(kRefEmpty);
    uint32_t ref = root;
    while (ref != kRefEmpty) {
        // interior nodes until a leaf (or nothing) is next
        while ((int)ref >= 0) {
            const GpuNode2* nd = nodes + ref;
            const float4 a = *reinterpret_cast<const float4*>(nd->a);
            const float4 b = *reinterpret_cast<const float4*>(nd->b);
            const float4 c = *reinterpret_cast<const float4*>(nd->c);
            const uint4 d = *reinterpret_cast<const uint4*>(nd->d);
            float t0, t1;
            const bool h0 = box_test(p, a.x, a.y, a.z, a.w, c.x, c.y, tmin, best_t, t0);
            const bool h1 = box_test(p, b.x, b.y, b.z, b.w, c.z, c.w, tmin, best_t, t1);
            if (COUNT) nbox += 2;
            if (h0 && h1) {
                uint32_t nr = d.x, fr = d.y;
                if (t1 < t0) { nr = d.y; fr = d.x; }
                st.push(fr);
                ref = nr;
            } else if (h0) {
                ref = d.x;
            } else if (h1) {
                ref = d.y;
            } else {
                ref = st.pop();
            }
        }
        // leaves until an interior node (or the sentinel) is next
        while ((int)ref < 0 && ref != kRefEmpty) {
            const uint32_t first = ref & kMaxTriSlots;
            const uint32_t cnt = ((ref >> 27) & 15u) + 1u;
            for (uint32_t k = 0; k < cnt; ++k) {
                const float4* tp = tris + 3 * (size_t)(first + k);
                const float4 va = tp[0], vb = tp[1], vc = tp[2];
                const uint32_t prim = __float_as_uint(va.w);
                float t;
                if (COUNT) ++ntri;
                if (!tri_test<false>(p, va, vb, vc, t)) continue;
                const bool accept = t >= tmin &&
                    (have ? (t < best_t || (t == best_t && prim < best_p)) : t < tmax);
                if (accept) {
                    have = true;
                    best_t = t;
                    best_p = prim;
                    best_slot = first + k;
                    if (ANY) return best_slot;
                }
            }
            ref = st.pop();
        }
    }
    return have ? best_slot : kRefEmpty;
}

template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void k_trace_closest(
    const GpuNode2* __restrict__ nodes, const float4* __restrict__ tris, uint32_t root, float origin_limit,
    ar_rays r, size_t n, const uint32_t* __restrict__ n_dev, ar_hits h, int* __restrict__ spill, uint32_t threads,
    uint32_t* __restrict__ cnt_box, uint32_t* __restrict__ cnt_tri) {
    __shared__ int lds_stack[kLdsStack * kTraceBlock];
    if (n_dev) n = *n_dev;
    const uint32_t gtid = blockIdx.x * kTraceBlock + threadIdx.x;
    Stack st;
    st.lds = lds_stack + threadIdx.x;
    st.spill = spill + gtid;
    st.threads = threads;
    for (size_t i = gtid; i < n; i += threads) {
        const float ox = r.ox[i], oy = r.oy[i], oz = r.oz[i], tmin = r.tmin[i];
        const float dx = r.dx[i], dy = r.dy[i], dz = r.dz[i], tmax = r.tmax[i];
        const RayPre p = ray_setup(ox, oy, oz, dx, dy, dz, origin_limit);
        float t;
        uint32_t nb = 0, nt = 0;
        const uint32_t slot = traverse<false, COUNT>(nodes, tris, root, p, tmin, tmax, st, t, nb, nt);
        if (slot == kRefEmpty) {
            h.prim[i] = AR_MISS;
            h.t[i] = __int_as_float(0x7f800000);
            h.u[i] = 0.0f; h.v[i] = 0.0f; h.flags[i] = 0u;
        } else {
            const float4* tp = tris + 3 * (size_t)slot;
            const float4 va = tp[0], vb = tp[1], vc = tp[2];
            float tt, u = 0.0f, v = 0.0f;
            uint32_t flags = 0;
            tri_test<true>(p, va, vb, vc, tt, &u, &v, &flags);
            h.prim[i] = __float_as_uint(va.w);
            h.t[i] = t;
            h.u[i] = u; h.v[i] = v; h.flags[i] = flags;
        }
        if (COUNT) { cnt_box[i] = nb; cnt_tri[i] = nt; }
    }
}

// Any hit.  Each wave owns 64 consecutive rays per iteration (threads is a
// multiple of 64), so the wave's ballot is exactly two occlusion words.
__global__ __launch_bounds__(kTraceBlock) void k_trace_any(
    const GpuNode2* __restrict__ nodes, const float4* __restrict__ tris, uint32_t root, float origin_limit,
    ar_rays r, size_t n, const uint32_t* __restrict__ n_dev, uint32_t* __restrict__ occ, int* __restrict__ spill,
    uint32_t threads) {
    __shared__ int lds_stack[kLdsStack * kTraceBlock];
    if (n_dev) n = *n_dev;
    const uint32_t gtid = blockIdx.x * kTraceBlock + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    Stack st;
    st.lds = lds_stack + threadIdx.x;
    st.spill = spill + gtid;
    st.threads = threads;
    const size_t nwords = (n + 31) / 32;
    for (size_t base = gtid - lane; base < n; base += threads) {
        const size_t i = base + lane;
        bool hit = false;
        if (i < n) {
            const float ox = r.ox[i], oy = r.oy[i], oz = r.oz[i], tmin = r.tmin[i];
            const float dx = r.dx[i], dy = r.dy[i], dz = r.dz[i], tmax = r.tmax[i];
            const RayPre p = ray_setup(ox, oy, oz, dx, dy, dz, origin_limit);
            float t;
            uint32_t nb = 0, nt = 0;
            hit = traverse<true, false>(nodes, tris, root, p, tmin, tmax, st, t, nb, nt) != kRefEmpty;
        }
        const uint64_t m = __ballot(hit);
        const size_t w = base / 32;
        if (lane == 0) occ[w] = (uint32_t)m;
        if (lane == 32 && w + 1 < nwords) occ[w + 1] = (uint32_t)(m >> 32);
    }
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

int trace_blocks_per_cu() {
    int b0 = 0, b1 = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b0, k_trace_closest<false>, kTraceBlock, 0) != hipSuccess) b0 = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b1, k_trace_any, kTraceBlock, 0) != hipSuccess) b1 = 1;
    return std::max(1, std::max(b0, b1));
}

hipError_t launch_trace_closest(const TraceArgs& ta, const ar_rays& r, size_t n, const ar_hits& h,
                                uint32_t* cnt_box, uint32_t* cnt_tri, uint32_t blocks, hipStream_t s,
                                const uint32_t* n_dev) {
    if (cnt_box) {
        hipLaunchKernelGGL(k_trace_closest<true>, dim3(blocks), dim3(kTraceBlock), 0, s, ta.nodes, ta.tris, ta.root,
                           ta.origin_limit, r, n, n_dev, h, ta.spill, ta.threads, cnt_box, cnt_tri);
    } else {
        hipLaunchKernelGGL(k_trace_closest<false>, dim3(blocks), dim3(kTraceBlock), 0, s, ta.nodes, ta.tris, ta.root,
                           ta.origin_limit, r, n, n_dev, h, ta.spill, ta.threads, nullptr, nullptr);
    }
    return hipGetLastError();
}

hipError_t launch_trace_any(const TraceArgs& ta, const ar_rays& r, size_t n, uint32_t* occ, uint32_t blocks,
                            hipStream_t s, const uint32_t* n_dev) {
    hipLaunchKernelGGL(k_trace_any, dim3(blocks), dim3(kTraceBlock), 0, s, ta.nodes, ta.tris, ta.root,
                       ta.origin_limit, r, n, n_dev, occ, ta.spill, ta.threads);
    return hipGetLastError();
}

hipError_t launch_shadow_rays(const ar_rays& r, const ar_hits& h, size_t n, float tx, float ty, float tz,
                              float tmin, float tmax_scale, const ar_rays& out, hipStream_t s) {
    const uint32_t bs = 256;
    const size_t blocks = (n + bs - 1) / bs;
    hipLaunchKernelGGL(k_shadow_rays, dim3((uint32_t)blocks), dim3(bs), 0, s, r, h, n, tx, ty, tz, tmin, tmax_scale, out);
    return hipGetLastError();
}

} // namespace ar
