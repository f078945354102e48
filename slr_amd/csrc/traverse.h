// QBVH traversal on the device: closest hit and any hit.
//
// The visiting order is the reference's (libSLR/Accelerator/QBVH.h:294-341):
// one 4-wide slab test per node with the current [tmin, tmax], children
// ordered by the node's three split axes and the ray's direction signs,
// internal children pushed far-to-near, then this node's leaf children
// intersected in near-to-far order before the next pop. Triangles are tested
// with the reference's Moller-Trumbore form and its inclusive t <= tmax rule
// (TriangleSurfaceShape.cpp:205-257), so a later triangle at equal t replaces
// an earlier one exactly as it does there. The source is compiled with
// -ffp-contract=off and correctly rounded division, and max/min are written
// as the compare-selects SSE's maxps/minps perform, so each hit t and each
// accepted/rejected triangle is bit-identical to the x86 reference.
//
// Stack: the top kLdsDepth entries of each lane's stack live in LDS, laid out
// [depth][lane] so a wave's pushes and pops hit 64 distinct banks; deeper
// entries (rare) spill to a per-lane region of global memory, again
// interleaved by lane. Total depth is the reference's 64.
#pragma once

#include "device_common.h"

namespace slrh {

struct TraceStack {
    uint32_t* lds;       // base of this lane's column: lds[d * ldsStride]
    uint32_t ldsStride;
    uint32_t* spill;     // this lane's column in the spill area
    uint32_t spillStride;
    int ldsDepth;
};

__device__ __forceinline__ void stackPush(const TraceStack& st, int& sp, uint32_t v, uint32_t& overflow) {
    if (sp < st.ldsDepth) st.lds[sp * st.ldsStride] = v;
    else if (sp < kMaxStack) st.spill[(sp - st.ldsDepth) * st.spillStride] = v;
    else { overflow = 1; return; }
    ++sp;
}
__device__ __forceinline__ uint32_t stackPop(const TraceStack& st, int& sp) {
    --sp;
    return sp < st.ldsDepth ? st.lds[sp * st.ldsStride] : st.spill[(sp - st.ldsDepth) * st.spillStride];
}

// SSE semantics: _mm_max_ps(a, b) = a > b ? a : b; _mm_min_ps(a, b) = a < b ? a : b.
__device__ __forceinline__ float sseMax(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float sseMin(float a, float b) { return a < b ? a : b; }

struct RayIn {
    float3 o, d;
    float tmin, tmax;
};

struct HitOut {
    float t;
    int32_t prim;
    float b0, b1;
};

// Per-ray constants of QBVH::Node::intersect and the order-table lookup.
struct RayPre {
    float3 inv;
    uint32_t nearIsMin;   // bit a set: inv[a] > 0 -> near plane is the box min
    uint32_t dirPos;      // bit a set: dir[a] >= 0 (dirIsPositive)
};

__device__ __forceinline__ RayPre prepareRay(const RayIn& r) {
    RayPre p;
    p.inv = f3(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
    p.nearIsMin = (p.inv.x > 0.0f ? 1u : 0u) | (p.inv.y > 0.0f ? 2u : 0u) | (p.inv.z > 0.0f ? 4u : 0u);
    p.dirPos = (r.d.x >= 0 ? 1u : 0u) | (r.d.y >= 0 ? 2u : 0u) | (r.d.z >= 0 ? 4u : 0u);
    return p;
}

// Slab test of the four children (QBVH.h:54-75); returns a 4-bit hit mask.
__device__ __forceinline__ uint32_t slab4(const float4* node, const RayIn& r, const RayPre& p, float tmin,
                                          float tmax) {
    const float4 mnx = node[0], mny = node[1], mnz = node[2];
    const float4 mxx = node[3], mxy = node[4], mxz = node[5];
    const bool nx = p.nearIsMin & 1u, ny = p.nearIsMin & 2u, nz = p.nearIsMin & 4u;
    const float nxa[4] = {mnx.x, mnx.y, mnx.z, mnx.w}, xxa[4] = {mxx.x, mxx.y, mxx.z, mxx.w};
    const float nya[4] = {mny.x, mny.y, mny.z, mny.w}, xya[4] = {mxy.x, mxy.y, mxy.z, mxy.w};
    const float nza[4] = {mnz.x, mnz.y, mnz.z, mnz.w}, xza[4] = {mxz.x, mxz.y, mxz.z, mxz.w};
    uint32_t mask = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float tn = tmin, tf = tmax;
        tn = sseMax(((nx ? nxa[c] : xxa[c]) - r.o.x) * p.inv.x, tn);
        tn = sseMax(((ny ? nya[c] : xya[c]) - r.o.y) * p.inv.y, tn);
        tn = sseMax(((nz ? nza[c] : xza[c]) - r.o.z) * p.inv.z, tn);
        tf = sseMin(((nx ? xxa[c] : nxa[c]) - r.o.x) * p.inv.x, tf);
        tf = sseMin(((ny ? xya[c] : nya[c]) - r.o.y) * p.inv.y, tf);
        tf = sseMin(((nz ? xza[c] : nza[c]) - r.o.z) * p.inv.z, tf);
        mask |= (tn <= tf ? 1u : 0u) << c;
    }
    return mask;
}

// The reference's OrderTable (QBVH.h:309-312), indexed by
// 4*dirPos[top] + 2*dirPos[left] + dirPos[right]; nibble i = child visited i-th.
__device__ __forceinline__ uint32_t orderCode(uint32_t idx) {
    const uint32_t lo = idx & 3u;
    // 0x0123, 0x0132, 0x1023, 0x1032, 0x2301, 0x3201, 0x2310, 0x3210
    const uint32_t tbl0 = (lo == 0) ? 0x0123u : (lo == 1) ? 0x0132u : (lo == 2) ? 0x1023u : 0x1032u;
    const uint32_t tbl1 = (lo == 0) ? 0x2301u : (lo == 1) ? 0x3201u : (lo == 2) ? 0x2310u : 0x3210u;
    return (idx & 4u) ? tbl1 : tbl0;
}

// Moller-Trumbore as TriangleSurfaceShape::intersect writes it. Returns true
// and updates t/b1/b2 when the triangle is accepted within [tmin, tmax].
__device__ __forceinline__ bool triTest(const float4 a, const float4 e1v, const float4 e2v, const RayIn& r,
                                        float tmin, float tmax, float* tt, float* ob1, float* ob2) {
    const float3 v0 = xyz(a), e1 = xyz(e1v), e2 = xyz(e2v);
    const float3 p = cross3(r.d, e2);
    const float det = dot3(e1, p);
    if (det == 0.0f) return false;
    const float invDet = 1.0f / det;
    const float3 d = r.o - v0;
    const float b1 = dot3(d, p) * invDet;
    if (b1 < 0.0f || b1 > 1.0f) return false;
    const float3 q = cross3(d, e1);
    const float b2 = dot3(r.d, q) * invDet;
    if (b2 < 0.0f || b1 + b2 > 1.0f) return false;
    const float t = dot3(e2, q) * invDet;
    if (t < tmin || t > tmax) return false;
    *tt = t;
    *ob1 = b1;
    *ob2 = b2;
    return true;
}

// Closest hit (QBVH::intersect, QBVH.h:355-365). Returns nodes visited.
template <bool kAnyHit>
__device__ __forceinline__ uint32_t traceRay(const DevScene& sc, const RayIn& r, const TraceStack& st, HitOut* hit,
                                             uint32_t& overflow, uint32_t* trisTested) {
    const RayPre p = prepareRay(r);
    float tmax = r.tmax;
    int32_t prim = -1;
    float hb1 = 0.0f, hb2 = 0.0f;
    int sp = 0;
    uint32_t visited = 0, tested = 0;
    stackPush(st, sp, 0u, overflow);
    while (sp > 0) {
        const uint32_t ni = stackPop(st, sp);
        const float4* node = sc.nodes + 8 * (size_t)ni;
        ++visited;
        const uint32_t hitMask = slab4(node, r, p, r.tmin, tmax);
        if (hitMask == 0) continue;
        const float4 cw = node[6];
        const uint32_t axes = __float_as_uint(node[7].x);
        const uint32_t children[4] = {__float_as_uint(cw.x), __float_as_uint(cw.y), __float_as_uint(cw.z),
                                      __float_as_uint(cw.w)};
        const uint32_t top = axes & 3u, left = (axes >> 2) & 3u, right = (axes >> 4) & 3u;
        const uint32_t code = orderCode(4u * ((p.dirPos >> top) & 1u) + 2u * ((p.dirPos >> left) & 1u) +
                                        ((p.dirPos >> right) & 1u));
        // internal children far-to-near
#pragma unroll
        for (int i = 3; i >= 0; --i) {
            const uint32_t c = (code >> (4 * i)) & 0xFu;
            const uint32_t ch = children[c];
            if (((hitMask >> c) & 1u) && ch != kInvalidChild && !(ch & kLeafBit)) stackPush(st, sp, ch, overflow);
        }
        // leaf children near-to-far
        for (int i = 0; i < 4; ++i) {
            const uint32_t c = (code >> (4 * i)) & 0xFu;
            const uint32_t ch = children[c];
            if (!((hitMask >> c) & 1u) || ch == kInvalidChild || !(ch & kLeafBit)) continue;
            const uint32_t first = ch & 0x00FFFFFFu;
            const uint32_t count = (ch >> 24) & 0x7Fu;
            for (uint32_t k = 0; k < count; ++k) {
                const float4* tri = sc.tris + 3 * (size_t)(first + k);
                const float4 a = tri[0];
                float t, b1, b2;
                ++tested;
                if (triTest(a, tri[1], tri[2], r, r.tmin, tmax, &t, &b1, &b2)) {
                    if (kAnyHit) {
                        hit->t = t;
                        hit->prim = (int32_t)__float_as_uint(a.w);
                        hit->b0 = 1.0f - b1 - b2;
                        hit->b1 = b1;
                        if (trisTested) *trisTested = tested;
                        return visited;
                    }
                    tmax = t;
                    prim = (int32_t)__float_as_uint(a.w);
                    hb1 = b1;
                    hb2 = b2;
                }
            }
        }
    }
    hit->t = tmax;
    hit->prim = prim;
    hit->b0 = prim >= 0 ? 1.0f - hb1 - hb2 : 0.0f;
    hit->b1 = prim >= 0 ? hb1 : 0.0f;
    if (trisTested) *trisTested = tested;
    return visited;
}

}  // namespace slrh
