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
// -ffp-contract=off and correctly rounded division, and the slab test's
// max/min select what SSE's maxps/minps select (see slab4), so each hit t and
// each accepted/rejected triangle is bit-identical to the x86 reference.
//
// Stack: the top kLdsDepth entries of each lane's stack live in LDS, laid out
// [depth][lane] so a wave's pushes and pops hit 64 distinct banks; deeper
// entries (rare) spill to a per-lane region of global memory, again
// interleaved by lane. Total depth is the reference's 64.
#pragma once

#include "device_common.h"

namespace slrh {

typedef __attribute__((address_space(3))) uint32_t lds_u32;

struct TraceStack {
    lds_u32* lds;        // this lane's column: lds[d * ldsStride]
    uint32_t ldsStride;
    uint32_t* spill;     // this lane's column in the spill area
    uint32_t spillStride;
    int ldsDepth;
};

__device__ __forceinline__ lds_u32* ldsColumn(uint32_t* shared, uint32_t lane) {
    return (lds_u32*)(shared + lane);
}

__device__ __forceinline__ void stackPush(const TraceStack& st, int& sp, uint32_t v, uint32_t& overflow) {
    if (sp < st.ldsDepth) {
        st.lds[sp * st.ldsStride] = v;
    } else if (sp < kMaxStack) {
        st.spill[(sp - st.ldsDepth) * st.spillStride] = v;
    } else {
        overflow = 1;
        return;
    }
    ++sp;
}
__device__ __forceinline__ uint32_t stackPop(const TraceStack& st, int& sp) {
    --sp;
    uint32_t v;
    if (sp < st.ldsDepth) v = st.lds[sp * st.ldsStride];
    else v = st.spill[(sp - st.ldsDepth) * st.spillStride];
    return v;
}

struct RayIn {
    float3 o, d;
    float tmin, tmax;
};

struct HitOut {
    float t;
    int32_t prim;
    float b0, b1;
    uint32_t aux;   // the hit triangle's per-reference payload (material class)
};

// Per-ray constants of QBVH::Node::intersect and the order-table lookup.
struct RayPre {
    float3 inv;
    uint32_t nearOff[3], farOff[3];   // byte offsets of the near/far plane rows per axis
    uint32_t dirPos;                  // bit a set: dir[a] >= 0 (dirIsPositive)
};

__device__ __forceinline__ RayPre prepareRay(const RayIn& r) {
    RayPre p;
    p.inv = f3(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
    // QBVH.h:67-72 picks the min plane as the near plane when invRayDir > 0;
    // node rows: 0-2 min x/y/z, 3-5 max x/y/z, 16 bytes each
    const bool px = p.inv.x > 0.0f, py = p.inv.y > 0.0f, pz = p.inv.z > 0.0f;
    p.nearOff[0] = px ? 0u : 48u;  p.farOff[0] = px ? 48u : 0u;
    p.nearOff[1] = py ? 16u : 64u; p.farOff[1] = py ? 64u : 16u;
    p.nearOff[2] = pz ? 32u : 80u; p.farOff[2] = pz ? 80u : 32u;
    p.dirPos = (r.d.x >= 0 ? 1u : 0u) | (r.d.y >= 0 ? 2u : 0u) | (r.d.z >= 0 ? 4u : 0u);
    return p;
}

__device__ __forceinline__ float4 ldg4(const char* base, uint32_t off) {
    return *(const float4*)(base + off);
}

// Slab test of the four children (QBVH.h:54-75); returns a 4-bit hit mask.
// The reference accumulates with SSE maxps/minps, i.e. a > b ? a : b. The
// accumulator b starts at the finite segment bound and only ever takes
// non-NaN values, so fmaxf/fminf (IEEE maxNum/minNum: a NaN term yields the
// accumulator) select the same values up to the sign of a zero, which no
// later comparison distinguishes; the hit masks are identical.
__device__ __forceinline__ uint32_t slab4(const char* nodes, uint32_t nb, const RayIn& r, const RayPre& p,
                                          float tmin, float tmax) {
    const float4 nx = ldg4(nodes, nb + p.nearOff[0]), ny = ldg4(nodes, nb + p.nearOff[1]);
    const float4 nz = ldg4(nodes, nb + p.nearOff[2]);
    const float4 fx = ldg4(nodes, nb + p.farOff[0]), fy = ldg4(nodes, nb + p.farOff[1]);
    const float4 fz = ldg4(nodes, nb + p.farOff[2]);
    const float nxa[4] = {nx.x, nx.y, nx.z, nx.w}, fxa[4] = {fx.x, fx.y, fx.z, fx.w};
    const float nya[4] = {ny.x, ny.y, ny.z, ny.w}, fya[4] = {fy.x, fy.y, fy.z, fy.w};
    const float nza[4] = {nz.x, nz.y, nz.z, nz.w}, fza[4] = {fz.x, fz.y, fz.z, fz.w};
    uint32_t mask = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float tn = fmaxf((nxa[c] - r.o.x) * p.inv.x, tmin);
        tn = fmaxf((nya[c] - r.o.y) * p.inv.y, tn);
        tn = fmaxf((nza[c] - r.o.z) * p.inv.z, tn);
        float tf = fminf((fxa[c] - r.o.x) * p.inv.x, tmax);
        tf = fminf((fya[c] - r.o.y) * p.inv.y, tf);
        tf = fminf((fza[c] - r.o.z) * p.inv.z, tf);
        mask |= (tn <= tf ? 1u : 0u) << c;
    }
    return mask;
}

// The reference's OrderTable (QBVH.h:309-312) as selections: with t, l, r the
// direction-sign bits of the node's top, left and right axes, the children
// are visited as (left pair, right pair) when t is set, else (right pair,
// left pair); the left pair is (0, 1) when l is set, else (1, 0); the right
// pair is (2, 3) when r is set, else (3, 2). This reproduces all eight table
// entries (0x0123 ... 0x3210, nibble i = child visited i-th).
__device__ __forceinline__ void orderChildren(const uint32_t c[4], uint32_t t, uint32_t l, uint32_t r,
                                              uint32_t v[4]) {
    const uint32_t a0 = l ? c[0] : c[1], a1 = l ? c[1] : c[0];
    const uint32_t b0 = r ? c[2] : c[3], b1 = r ? c[3] : c[2];
    v[0] = t ? a0 : b0;
    v[1] = t ? a1 : b1;
    v[2] = t ? b0 : a0;
    v[3] = t ? b1 : a1;
}

// Moller-Trumbore as TriangleSurfaceShape::intersect writes it. Returns true
// and sets t/b1/b2 when the triangle is accepted within [tmin, tmax].
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

struct TraceCounters {
    uint32_t nodes = 0, tris = 0, rays = 0;          // all rays
    uint32_t nodesAny = 0, trisAny = 0, raysAny = 0; // the any-hit share of them
    uint32_t overflow = 0;
};

constexpr uint32_t kNoNode = 0xFFFFFFFFu;

// Persistent traversal loop over the queue entries 0..count-1.
//
// Lane stages: 0 idle; 1 queue entry requested; 2 ray requested; 3 tracing.
// A lane moves one stage per iteration, so the loads a stage issues (the
// queue entry, then the ray it names) are waited for together with the other
// lanes' node and triangle loads of the same iteration instead of stalling
// the wave on their own. Idle lanes are refilled once at least kRefill of the
// wave's lanes are idle: the wave claims a chunk of entries with one atomic
// on fetchCtr (zeroed before the launch) and hands them out in lane order,
// so waves balance across the chip and lanes do not wait for the wave's
// longest ray.
//
// Tracing lanes advance their ray by ONE step per iteration, either one node
// slab test or one triangle test, so lanes in a leaf and lanes between leaves
// both do useful work in every iteration (a node-then-all-its-leaves step
// would run the leaf loop at the width of the wave's longest leaf list). The
// nearest internal child is kept in a register as the next node instead of
// being pushed and popped back, which the reference's push-all-then-pop order
// makes equivalent. Lane state while tracing: the next node (ni, or kNoNode
// to pop), the ordered child words of the last node tested (v[4]) with a mask
// of its leaf children still to intersect (leafPend, bit i = v[i]), and the
// triangle range of the leaf being intersected [tcur, tend).
//
//   entry(i)          -> uint32_t, the i-th queue entry
//   load(e, RayIn& r) -> bool, fills the ray of entry e; true for any hit
//   finish(e, anyHit, const HitOut&)
template <int kRefill, typename Entry, typename Load, typename Finish>
__device__ __forceinline__ void traceLoop(const DevScene& sc, const TraceStack& st, uint32_t count,
                                          uint32_t* fetchCtr, Entry&& entry, Load&& load, Finish&& finish,
                                          TraceCounters& tc) {
    constexpr uint32_t kChunk = 64;
    const char* nodes = (const char*)sc.nodes;
    const char* tris = (const char*)sc.tris;
    uint32_t next = 0, end = 0;   // wave-uniform chunk cursor
    bool exhausted = false;       // wave-uniform
    uint32_t stage = 0;
    uint32_t ent = 0;
    bool anyHit = false;
    RayIn r;
    RayPre p;
    float tmax = 0.0f, hb1 = 0.0f, hb2 = 0.0f;
    int32_t prim = -1;
    uint32_t aux = 0;
    int sp = 0;
    uint32_t ni = kNoNode, tcur = 0, tend = 0, leafPend = 0;
    uint32_t v[4] = {kInvalidChild, kInvalidChild, kInvalidChild, kInvalidChild};
    // once the current leaf is exhausted, move to the next pending leaf child
    // in visiting order (skipping empty ones)
    auto nextLeaf = [&]() {
        while (tcur >= tend && leafPend) {
            const uint32_t i = (uint32_t)__builtin_ctz(leafPend);
            leafPend &= leafPend - 1u;
            const uint32_t w = i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3];
            tcur = w & 0x00FFFFFFu;
            tend = tcur + ((w >> 24) & 0x7Fu);
        }
    };
    for (;;) {
        if (stage == 2) {
            // the ray requested last iteration starts at the root
            p = prepareRay(r);
            tmax = r.tmax;
            prim = -1;
            aux = 0;
            sp = 0;
            ni = 0;
            tcur = tend = 0;
            leafPend = 0;
            stage = 3;
        } else if (stage == 1) {
            anyHit = load(ent, r);
            stage = 2;
        }
        const unsigned long long idle = __ballot(stage == 0);
        if (!exhausted && __popcll(idle) >= kRefill) {
            if (next >= end) {
                uint32_t base = 0;
                if (__lane_id() == 0) base = atomicAdd(fetchCtr, kChunk);
                base = __shfl(base, 0);
                next = base;
                end = base + kChunk < count ? base + kChunk : count;
                if (base >= count) { exhausted = true; next = end = count; }
            }
            const uint32_t avail = end - next;
            // rank of this lane among the idle lanes
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
            if (stage == 0 && rank < avail) {
                ent = entry(next + rank);
                stage = 1;
            }
            const uint32_t need = (uint32_t)__popcll(idle);
            next += need < avail ? need : avail;
        }
        if (__ballot(stage != 0) == 0ull) {
            if (exhausted) break;
            continue;
        }
        if (stage == 3) {
            bool done = false;
            if (tcur < tend) {
                // one triangle of the current leaf
                const uint32_t tb = tcur * 48u;
                const float4 a = ldg4(tris, tb), e1 = ldg4(tris, tb + 16u), e2 = ldg4(tris, tb + 32u);
                float t, b1, b2;
                ++tc.tris;
                tc.trisAny += anyHit ? 1u : 0u;
                if (triTest(a, e1, e2, r, r.tmin, tmax, &t, &b1, &b2)) {
                    tmax = t;
                    prim = (int32_t)__float_as_uint(a.w);
                    aux = __float_as_uint(e1.w);
                    hb1 = b1;
                    hb2 = b2;
                    done = anyHit;
                }
                ++tcur;
                nextLeaf();
            } else {
                if (ni == kNoNode) ni = stackPop(st, sp);
                const uint32_t nb = ni * 128u;
                ++tc.nodes;
                tc.nodesAny += anyHit ? 1u : 0u;
                const uint32_t hitMask = slab4(nodes, nb, r, p, r.tmin, tmax);
                ni = kNoNode;
                if (hitMask) {
                    const uint4 cw = *(const uint4*)(nodes + nb + 96u);
                    const uint32_t axes = *(const uint32_t*)(nodes + nb + 112u);
                    const uint32_t c[4] = {(hitMask & 1u) ? cw.x : kInvalidChild, (hitMask & 2u) ? cw.y : kInvalidChild,
                                           (hitMask & 4u) ? cw.z : kInvalidChild, (hitMask & 8u) ? cw.w : kInvalidChild};
                    orderChildren(c, (p.dirPos >> (axes & 3u)) & 1u, (p.dirPos >> ((axes >> 2) & 3u)) & 1u,
                                  (p.dirPos >> ((axes >> 4) & 3u)) & 1u, v);
                    // internal children (leaf bit clear; the invalid word has it set):
                    // far-to-near, each pushing the one held before it
#pragma unroll
                    for (int i = 3; i >= 0; --i) {
                        if (!(v[i] & kLeafBit)) {
                            if (ni != kNoNode) stackPush(st, sp, ni, tc.overflow);
                            ni = v[i];
                        }
                    }
                    leafPend = 0;
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        leafPend |= ((v[i] & kLeafBit) && v[i] != kInvalidChild ? 1u : 0u) << i;
                    nextLeaf();
                }
            }
            if (done || (tcur >= tend && ni == kNoNode && sp == 0)) {
                HitOut h;
                h.t = tmax;
                h.prim = prim;
                h.b0 = prim >= 0 ? 1.0f - hb1 - hb2 : 0.0f;
                h.b1 = prim >= 0 ? hb1 : 0.0f;
                h.aux = aux;
                finish(ent, anyHit, h);
                ++tc.rays;
                tc.raysAny += anyHit ? 1u : 0u;
                stage = 0;
            }
        }
    }
}

}  // namespace slrh
