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

// Queue entries a wave claims per atomic on the traversal's fetch cursor.
#ifndef SLRH_FETCH_CHUNK
#define SLRH_FETCH_CHUNK 128
#endif
// The queue can be split into parts with a cursor each (kFetchStride uint32
// apart, so on separate cache lines): block b starts on part b % parts, the
// blocks of one XCD on the same part, and moves on when its part runs dry.
#ifndef SLRH_FETCH_PARTS
#define SLRH_FETCH_PARTS 8
#endif
constexpr uint32_t kFetchStride = 32;

#include "device_common.h"

namespace slrh {

typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) const uint8_t lds_u8;

struct TraceStack {
    lds_u32* lds;        // this lane's column: lds[d * ldsStride]
    uint32_t ldsStride;
    uint32_t* spill;     // this lane's column in the spill area
    uint32_t spillStride;
    int ldsDepth;
    lds_u8* treelet;     // LDS copy of nodes 0..treeletNodes-1, kTreeletRows rows each
    uint32_t treeletNodes;
};

typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 ldsf4(lds_u8* base, uint32_t off) {
    const f32x4 v = *(__attribute__((address_space(3))) const f32x4*)(base + off);
    return make_float4(v.x, v.y, v.z, v.w);
}

// Copies the first n nodes' rows 0..kTreeletRows-1 into LDS (called by the
// whole block before the traversal loop; ends with a barrier).
__device__ __forceinline__ lds_u8* loadTreelet(float4* shared, const float4* nodes, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n * kTreeletRows; i += blockDim.x) {
        const uint32_t node = i / kTreeletRows, row = i - node * kTreeletRows;
        shared[i] = nodes[8 * node + row];
    }
    __syncthreads();
    return (lds_u8*)(const uint8_t*)shared;
}

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

// A ray as it is stored and requested from memory: origin + tmin, direction
// + tmax, rayInv(direction).
struct RayIn {
    float4 o, d, inv;
    __device__ float3 org() const { return xyz(o); }
    __device__ float3 dir() const { return xyz(d); }
    __device__ float tmin() const { return o.w; }
    __device__ float tmax() const { return d.w; }
    __device__ uint32_t signs() const { return __float_as_uint(inv.w); }
};

struct HitOut {
    float t;
    int32_t prim;
    float b0, b1;
    uint32_t aux;   // the hit triangle's per-reference payload (material class)
    int32_t inst;   // instance of the hit triangle, -1 in world space
};

// Two-level traversal (instances). A world-level reference whose first word
// has kRefInstBit set is an instance: word 3 of its first float4 holds the
// instance id, of the second the root node of its prototype's tree, of the
// third the material class that replaces the prototype's (kNoClassOverride:
// none). Entering one suspends the rest of the current node (the held next
// node, the pending leaf children, the rest of the current leaf) onto the
// stack as node and leaf words under a kInstMarker, takes the ray into
// object space and traverses the prototype's tree; popping the marker
// restores the world ray. Hit distances are the same in both spaces (the
// direction is transformed, not renormalised).
constexpr uint32_t kRefInstBit = 0x80000000u;
constexpr uint32_t kInstMarker = 0x7FFFFFFFu;   // stack word: back to the world ray
constexpr uint32_t kEnterInst = 0x7FFFFFFEu;    // next node: the instance transform is on its way
constexpr uint32_t kNoClassOverride = 0xFFFFFFFFu;

// Reciprocal direction of a ray and its sign bits, computed once by whoever
// produces the ray: xyz = 1 / d (QBVH::Node::intersect's invRayDir, correctly
// rounded), w = bits: bit a (0-2) set when 1/d[a] > 0, i.e. the min plane is
// the near plane (QBVH.h:67-72); bit 3 + a set when d[a] >= 0 (the traversal
// order's dirIsPositive, QBVH.h:296; differs from bit a for d[a] = -0).
__device__ __forceinline__ float4 rayInv(float3 d) {
    const float ix = 1.0f / d.x, iy = 1.0f / d.y, iz = 1.0f / d.z;
    const uint32_t bits = (ix > 0.0f ? 1u : 0u) | (iy > 0.0f ? 2u : 0u) | (iz > 0.0f ? 4u : 0u) |
                          (d.x >= 0.0f ? 8u : 0u) | (d.y >= 0.0f ? 16u : 0u) | (d.z >= 0.0f ? 32u : 0u);
    return make_float4(ix, iy, iz, __uint_as_float(bits));
}

// Per-ray constants of QBVH::Node::intersect and the order-table lookup.
// Node rows (16 B each): 0 min x, 1 max x, 2 min y, 3 max y, 4 min z, 5 max z,
// 6 child words (with the split axes); the far row of an axis is the near
// row ^ 16.
struct RayPre {
    uint32_t nearOff[3];   // byte offset of the near plane row per axis
    uint32_t dirPos;       // bit a set: dir[a] >= 0 (dirIsPositive)
};

__device__ __forceinline__ RayPre prepareRay(const RayIn& r) {
    RayPre p;
    const uint32_t b = r.signs();
    p.nearOff[0] = (b & 1u) ? 0u : 16u;
    p.nearOff[1] = (b & 2u) ? 32u : 48u;
    p.nearOff[2] = (b & 4u) ? 64u : 80u;
    p.dirPos = (b >> 3) & 7u;
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
__device__ __forceinline__ uint32_t slab4(const float4 nx, const float4 fx, const float4 ny, const float4 fy,
                                          const float4 nz, const float4 fz, const RayIn& r, float tmin, float tmax) {
    const float nxa[4] = {nx.x, nx.y, nx.z, nx.w}, fxa[4] = {fx.x, fx.y, fx.z, fx.w};
    const float nya[4] = {ny.x, ny.y, ny.z, ny.w}, fya[4] = {fy.x, fy.y, fy.z, fy.w};
    const float nza[4] = {nz.x, nz.y, nz.z, nz.w}, fza[4] = {fz.x, fz.y, fz.z, fz.w};
    uint32_t mask = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float tn = fmaxf((nxa[c] - r.o.x) * r.inv.x, tmin);
        tn = fmaxf((nya[c] - r.o.y) * r.inv.y, tn);
        tn = fmaxf((nza[c] - r.o.z) * r.inv.z, tn);
        float tf = fminf((fxa[c] - r.o.x) * r.inv.x, tmax);
        tf = fminf((fya[c] - r.o.y) * r.inv.y, tf);
        tf = fminf((fza[c] - r.o.z) * r.inv.z, tf);
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
    const float3 rd = r.dir();
    const float3 p = cross3(rd, e2);
    const float det = dot3(e1, p);
    if (det == 0.0f) return false;
    const float invDet = 1.0f / det;
    const float3 d = r.org() - v0;
    const float b1 = dot3(d, p) * invDet;
    if (b1 < 0.0f || b1 > 1.0f) return false;
    const float3 q = cross3(d, e1);
    const float b2 = dot3(rd, q) * invDet;
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
    uint32_t nodesLds = 0;                           // nodes read from the LDS copy of the tree top
    uint32_t insts = 0;                              // instances entered (two-level scenes)
    uint32_t overflow = 0;
};

constexpr uint32_t kNoNode = 0xFFFFFFFFu;

// Position of the k-th (0-based) set bit of m; m has more than k set bits.
__device__ __forceinline__ uint32_t kthSetBit(unsigned long long m, uint32_t k) {
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t w = 32; w >= 1; w >>= 1) {
        const uint32_t c = (uint32_t)__popcll((m >> pos) & ((1ull << w) - 1ull));
        if (k >= c) { k -= c; pos += w; }
    }
    return pos;
}

// Persistent traversal loop over the queue entries 0..count-1.
//
// Feeding lanes: every lane holds, besides the ray it traces, the queue
// entry of its next ray (entNext), claimed ahead of time. When a ray
// finishes, the lane issues the loads of the next ray at once and starts
// tracing it in the following iteration; those loads are waited for together
// with the other lanes' node and triangle loads. Lanes that lack a next entry
// are refilled once at least kRefill of them do: the wave claims a chunk of
// entries with one atomic on fetchCtr (zeroed before the launch) and hands
// them out in lane order, so waves balance across the chip. Once the queue
// is exhausted, idle lanes take the next entries that busy lanes still hold,
// so no ray waits behind another lane's long ray at the end of the launch.
//
// Tracing lanes advance their ray by one step per iteration, one node slab
// test or one triangle test (with kFuse, a node test whose node opens a leaf
// is followed by that leaf's first triangle test in the same iteration), so
// lanes in a leaf and lanes between leaves both do useful work in every
// iteration (a node-then-all-its-leaves step would run the leaf loop at the
// width of the wave's longest leaf list). The nearest internal child is kept
// in a register as the next node instead of being pushed and popped back,
// which the reference's push-all-then-pop order makes equivalent. Lane state
// while tracing: the next node (ni, or kNoNode to pop), the ordered child
// words of the last node tested (v[4]) with a mask of its leaf children still
// to intersect (leafPend, bit i = v[i]), and the triangle range of the leaf
// being intersected [tcur, tend).
//
//   entry(i)            -> uint32_t, the i-th queue entry
//   load(e, RayIn& r)   -> bool, requests the ray of entry e (r.inv too when
//                          kHasInv, else it is computed here); true for any hit
//   finish(e, anyHit, const HitOut&)
// chunk: entries a wave claims at a time (wave-uniform; smaller chunks spread
// a short queue over more waves)
template <int kRefill, bool kFuse, bool kHasInv, bool kInst = false, uint32_t kParts = 1, typename Entry,
          typename Load, typename Finish>
__device__ __forceinline__ void traceLoop(const DevScene& sc, const TraceStack& st, uint32_t count,
                                          uint32_t* fetchCtr, Entry&& entry, Load&& load, Finish&& finish,
                                          TraceCounters& tc, uint32_t chunk = SLRH_FETCH_CHUNK) {
    const uint32_t kChunk = chunk;
    uint32_t part = kParts > 1 ? blockIdx.x % kParts : 0;   // wave-uniform
    uint32_t partsLeft = kParts;
    const char* nodes = (const char*)sc.nodes;
    const char* tris = (const char*)sc.tris;
    const uint32_t lane = __lane_id();
    uint32_t next = 0, end = 0;   // wave-uniform chunk cursor
    bool exhausted = false;       // wave-uniform
    uint32_t stage = 0;           // 0 idle, 2 ray requested, 3 tracing
    uint32_t ent = 0, entNext = 0;
    bool hasNext = false;         // entNext holds a claimed entry
    bool anyHit = false;
    RayIn r;
    RayPre p;
    float tmax = 0.0f, hb1 = 0.0f, hb2 = 0.0f;
    int32_t prim = -1;
    uint32_t aux = 0;
    // instances: the one being traversed, its root and class override, the
    // hit's instance, and the world-to-object rows requested on entry
    int32_t inst = -1, hitInst = -1;
    uint32_t instRoot = 0, instCls = kNoClassOverride;
    float4 x0 = make_float4(0, 0, 0, 0), x1 = x0, x2 = x0, x3 = x0;
    int sp = 0;
    uint32_t ni = kNoNode, tcur = 0, tend = 0, leafPend = 0;
    uint32_t v[4] = {kNoChild, kNoChild, kNoChild, kNoChild};
    // once the current leaf is exhausted, move to the next pending leaf child
    // in visiting order (skipping empty ones)
    auto nextLeaf = [&]() {
        while (tcur >= tend && leafPend) {
            const uint32_t i = (uint32_t)__builtin_ctz(leafPend);
            leafPend &= leafPend - 1u;
            const uint32_t w = i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3];
            tcur = w & kLeafFirstMask;
            tend = tcur + ((w >> kLeafCountShift) & kLeafCountMask);
        }
    };
    auto startRay = [&](uint32_t e) {
        ent = e;
        anyHit = load(e, r);
        stage = 2;
    };
    for (;;) {
        if (stage == 2) {
            // the ray requested last iteration starts at the root
            if (!kHasInv) r.inv = rayInv(r.dir());
            p = prepareRay(r);
            tmax = r.tmax();
            prim = -1;
            aux = 0;
            sp = 0;
            ni = 0;
            tcur = tend = 0;
            leafPend = 0;
            if (kInst) {
                inst = hitInst = -1;
                instCls = kNoClassOverride;
            }
            stage = 3;
        }
        if (kInst && stage == 4) {
            // the world ray reloaded after an instance: traversal goes on
            if (!kHasInv) r.inv = rayInv(r.dir());
            p = prepareRay(r);
            stage = 3;
        }
        // idle lanes start the entry they claimed in an earlier iteration
        if (stage == 0 && hasNext) {
            startRay(entNext);
            hasNext = false;
        }
        if (exhausted) {
            // the k-th idle lane takes the next entry of the k-th lane holding one
            const unsigned long long idleM = __ballot(stage == 0);
            const unsigned long long donorM = __ballot(hasNext);
            if (idleM != 0ull && donorM != 0ull) {
                const uint32_t nIdle = (uint32_t)__popcll(idleM), nDonor = (uint32_t)__popcll(donorM);
                const uint32_t rI = __builtin_amdgcn_mbcnt_hi((uint32_t)(idleM >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)idleM, 0u));
                const uint32_t rD = __builtin_amdgcn_mbcnt_hi((uint32_t)(donorM >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)donorM, 0u));
                const bool takes = stage == 0 && rI < nDonor;
                const uint32_t src = takes ? kthSetBit(donorM, rI) : lane;
                const uint32_t got = (uint32_t)__shfl((int)entNext, (int)src);
                if (hasNext && rD < nIdle) hasNext = false;
                if (takes) startRay(got);
            }
        }
        // lanes without a next entry claim one; its load completes with this
        // iteration's traversal loads and is first used in a later step
        const unsigned long long lack = __ballot(!hasNext);
        if (!exhausted && __popcll(lack) >= kRefill) {
            if (next >= end) {
                if (kParts == 1) {
                    uint32_t base = 0;
                    if (lane == 0) base = atomicAdd(fetchCtr, kChunk);
                    base = __shfl(base, 0);
                    next = base;
                    end = base + kChunk < count ? base + kChunk : count;
                    if (base >= count) { exhausted = true; next = end = count; }
                } else {
                    for (;;) {
                        const uint32_t pb = (uint32_t)((uint64_t)count * part / kParts);
                        const uint32_t pe = (uint32_t)((uint64_t)count * (part + 1) / kParts);
                        uint32_t base = 0;
                        if (lane == 0) base = atomicAdd(fetchCtr + part * kFetchStride, kChunk);
                        base = __shfl(base, 0);
                        if (base < pe - pb) {
                            next = pb + base;
                            end = pe - next > kChunk ? next + kChunk : pe;
                            break;
                        }
                        if (--partsLeft == 0) {
                            exhausted = true;
                            next = end = count;
                            break;
                        }
                        part = part + 1 < kParts ? part + 1 : 0;
                    }
                }
            }
            const uint32_t avail = end - next;
            // rank of this lane among the lanes lacking a next entry
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(lack >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)lack, 0u));
            if (!hasNext && rank < avail) {
                entNext = entry(next + rank);
                hasNext = true;
            }
            const uint32_t need = (uint32_t)__popcll(lack);
            next += need < avail ? need : avail;
        }
        if (__ballot(stage != 0 || hasNext) == 0ull) {
            if (exhausted) break;
            continue;
        }
        if (stage == 3) {
            bool done = false;
            if (tcur >= tend) {
                // one node
                if (ni == kNoNode) ni = stackPop(st, sp);
                if (kInst) {
                    if (ni == kInstMarker) {
                        // the instance is done: request the world ray again
                        inst = -1;
                        instCls = kNoClassOverride;
                        ni = kNoNode;
                        load(ent, r);
                        stage = 4;
                        goto stepDone;
                    }
                    if (ni == kEnterInst) {
                        // object-space ray: origin by the affine rows, direction
                        // by their linear part (StaticTransform * Ray)
                        if (!(sc.compatInstShadow && anyHit)) {
                            const float3 o = r.org(), d = r.dir();
                            // Matrix4x4 * Point3D: rows dotted with (o, 1), then
                            // divided by w unless it is exactly one
                            float px = x0.x * o.x + x0.y * o.y + x0.z * o.z + x0.w;
                            float py = x1.x * o.x + x1.y * o.y + x1.z * o.z + x1.w;
                            float pz = x2.x * o.x + x2.y * o.y + x2.z * o.z + x2.w;
                            const float pw = x3.x * o.x + x3.y * o.y + x3.z * o.z + x3.w;
                            if (pw != 1.0f) {
                                const float rw = 1.0f / pw;
                                px *= rw;
                                py *= rw;
                                pz *= rw;
                            }
                            r.o = make_float4(px, py, pz, r.o.w);
                            r.d = make_float4(x0.x * d.x + x0.y * d.y + x0.z * d.z, x1.x * d.x + x1.y * d.y + x1.z * d.z,
                                              x2.x * d.x + x2.y * d.y + x2.z * d.z, r.d.w);
                            r.inv = rayInv(r.dir());
                            p = prepareRay(r);
                        }
                        ni = instRoot;
                    } else if (ni & kLeafBit) {
                        // a suspended leaf range: intersect it next
                        tcur = ni & kLeafFirstMask;
                        tend = tcur + ((ni >> kLeafCountShift) & kLeafCountMask);
                        leafPend = 0;
                        ni = kNoNode;
                        goto triStep;
                    }
                }
                ++tc.nodes;
                tc.nodesAny += anyHit ? 1u : 0u;
                // all rows the step needs are requested together: one round
                // trip, from the LDS copy of the top of the tree when the node
                // is there
                float4 nx, fx, ny, fy, nz, fz;
                uint4 cw;
                if (ni < st.treeletNodes) {
                    ++tc.nodesLds;
                    const lds_u8* q = st.treelet + ni * (kTreeletRows * 16u);
                    nx = ldsf4(q, p.nearOff[0]); fx = ldsf4(q, p.nearOff[0] ^ 16u);
                    ny = ldsf4(q, p.nearOff[1]); fy = ldsf4(q, p.nearOff[1] ^ 16u);
                    nz = ldsf4(q, p.nearOff[2]); fz = ldsf4(q, p.nearOff[2] ^ 16u);
                    const float4 w = ldsf4(q, 96u);
                    cw = make_uint4(__float_as_uint(w.x), __float_as_uint(w.y), __float_as_uint(w.z),
                                    __float_as_uint(w.w));
                } else {
                    const uint32_t nb = ni * 128u;
                    nx = ldg4(nodes, nb + p.nearOff[0]); fx = ldg4(nodes, nb + (p.nearOff[0] ^ 16u));
                    ny = ldg4(nodes, nb + p.nearOff[1]); fy = ldg4(nodes, nb + (p.nearOff[1] ^ 16u));
                    nz = ldg4(nodes, nb + p.nearOff[2]); fz = ldg4(nodes, nb + (p.nearOff[2] ^ 16u));
                    cw = *(const uint4*)(nodes + nb + 96u);
                }
                // keeps the child-word loads beside the plane loads (the
                // compiler would otherwise issue them after the slab test)
                asm volatile("" ::"v"(cw.x), "v"(cw.y), "v"(cw.z), "v"(cw.w));
                const uint32_t hitMask = slab4(nx, fx, ny, fy, nz, fz, r, r.tmin(), tmax);
                ni = kNoNode;
                if (hitMask) {
                    const uint32_t top = (cw.x >> kAxisShift) & 3u, left = (cw.y >> kAxisShift) & 3u,
                                   right = (cw.z >> kAxisShift) & 3u;
                    const uint32_t c[4] = {(hitMask & 1u) ? cw.x : kNoChild, (hitMask & 2u) ? cw.y : kNoChild,
                                           (hitMask & 4u) ? cw.z : kNoChild, (hitMask & 8u) ? cw.w : kNoChild};
                    orderChildren(c, (p.dirPos >> top) & 1u, (p.dirPos >> left) & 1u, (p.dirPos >> right) & 1u, v);
                    // internal children (leaf bit clear): far-to-near, each
                    // pushing the one held before it
#pragma unroll
                    for (int i = 3; i >= 0; --i) {
                        if (!(v[i] & kLeafBit)) {
                            if (ni != kNoNode) stackPush(st, sp, ni, tc.overflow);
                            ni = v[i] & kNodeIndexMask;
                        }
                    }
                    // leaf children with triangles, in visiting order
                    leafPend = 0;
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        leafPend |= ((v[i] & kLeafBit) && ((v[i] >> kLeafCountShift) & kLeafCountMask) ? 1u : 0u) << i;
                    nextLeaf();
                }
                // kFuse: a node that opens a leaf has its first triangle tested in
                // the same iteration; otherwise the next iteration does
                if (!kFuse) goto stepDone;
            }
        triStep:
            if (tcur < tend) {
                // one triangle of the current leaf
                const uint32_t tb = tcur * 48u;
                const float4 a = ldg4(tris, tb), e1 = ldg4(tris, tb + 16u), e2 = ldg4(tris, tb + 32u);
                // one round trip for the whole reference (the compiler would
                // otherwise request v0 only after the determinant test)
                asm volatile("" ::"v"(a.x), "v"(a.y), "v"(a.z), "v"(a.w), "v"(e1.x), "v"(e1.y), "v"(e1.z),
                             "v"(e1.w), "v"(e2.x), "v"(e2.y), "v"(e2.z));
                if (kInst && (__float_as_uint(a.w) & kRefInstBit)) {
                    // an instance: suspend this node's remaining work (held
                    // next node deepest, then pending leaves far to near, then
                    // the rest of this leaf) under the marker, and request the
                    // instance's world-to-object rows
                    ++tcur;
                    if (ni != kNoNode) stackPush(st, sp, ni, tc.overflow);
#pragma unroll
                    for (int i = 3; i >= 0; --i)
                        if ((leafPend >> i) & 1u) stackPush(st, sp, i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3],
                                                           tc.overflow);
                    if (tcur < tend) stackPush(st, sp, kLeafBit | ((tend - tcur) << kLeafCountShift) | tcur, tc.overflow);
                    stackPush(st, sp, kInstMarker, tc.overflow);
                    leafPend = 0;
                    tcur = tend;
                    inst = (int32_t)(__float_as_uint(a.w) & ~kRefInstBit);
                    ++tc.insts;
                    instRoot = __float_as_uint(e1.w);
                    instCls = __float_as_uint(e2.w);
                    const float4* xr = sc.instXf + kInstXfRows * (uint32_t)inst + 3u;
                    x0 = xr[0];
                    x1 = xr[1];
                    x2 = xr[2];
                    x3 = xr[3];
                    ni = kEnterInst;
                } else {
                    float t, b1, b2;
                    ++tc.tris;
                    tc.trisAny += anyHit ? 1u : 0u;
                    if (triTest(a, e1, e2, r, r.tmin(), tmax, &t, &b1, &b2)) {
                        tmax = t;
                        prim = (int32_t)__float_as_uint(a.w);
                        aux = __float_as_uint(e1.w);
                        if (kInst) {
                            if (instCls != kNoClassOverride) aux = instCls;
                            hitInst = inst;
                        }
                        hb1 = b1;
                        hb2 = b2;
                        done = anyHit;
                    }
                    ++tcur;
                    nextLeaf();
                }
            }
        stepDone:
            if (done || (tcur >= tend && ni == kNoNode && sp == 0)) {
                HitOut h;
                h.t = tmax;
                h.prim = prim;
                h.b0 = prim >= 0 ? 1.0f - hb1 - hb2 : 0.0f;
                h.b1 = prim >= 0 ? hb1 : 0.0f;
                h.aux = aux;
                h.inst = kInst ? hitInst : -1;
                finish(ent, anyHit, h);
                ++tc.rays;
                tc.raysAny += anyHit ? 1u : 0u;
                stage = 0;
                if (hasNext) {
                    startRay(entNext);
                    hasNext = false;
                }
            }
        }
    }
}

}  // namespace slrh
