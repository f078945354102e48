// Device functions shared by the integrators' kernels (render.hip, bpt.hip):
// wave-level queue appends and counters, the XORShift stream, spectrum
// evaluation at a path's 16 wavelengths, surface points with their shading
// frames, instance transforms, and triangle-light sampling.
#pragma once

#include <hip/hip_runtime.h>

#include "bsdf.h"
#include "device_common.h"

namespace slrh {

__device__ __forceinline__ uint32_t laneId() { return __lane_id(); }

// Appends the lanes with pred set to a queue; returns each such lane's slot.
// Must be reached by every lane of the wave.
__device__ __forceinline__ uint32_t waveAppend(uint32_t* counter, bool pred) {
    const unsigned long long mask = __ballot(pred);
    if (mask == 0) return 0;
    const uint32_t lane = laneId();
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    const unsigned long long below = lane == 0 ? 0ull : (mask & ((~0ull) >> (64 - lane)));
    return base + (uint32_t)__popcll(below);
}

__device__ __forceinline__ void waveAdd64(unsigned long long* counter, unsigned long long v) {
    // wave reduction then one atomic
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (laneId() == 0 && v) atomicAdd(counter, v);
}

// Render statistics are spread over kStatSpread copies per counter (waves
// pick one by their index; the host sums them): same-address atomics
// serialise, and one per wave of a persistent grid on a single address
// costs a visible share of a short kernel.
constexpr uint32_t kStatSpread = 32;
__device__ __forceinline__ void statAdd(unsigned long long* stat, uint32_t which, unsigned long long v) {
    const uint32_t w = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (kStatSpread - 1);
    waveAdd64(stat + which * kStatSpread + w, v);
}

// XORShift128 (libSLR/RNG/XORShiftRNG.cpp:21-36) and its float conversion
// (libSLR/Core/random_number_generator.cpp:12-15).
__device__ __forceinline__ uint32_t xsNext(uint4& s) {
    uint32_t t = s.x ^ (s.x << 11);
    s.x = s.y; s.y = s.z; s.z = s.w;
    s.w = (s.w ^ (s.w >> 19)) ^ (t ^ (t >> 8));
    return s.w;
}
__device__ __forceinline__ float xsFloat(uint4& s) {
    uint32_t bits = (xsNext(s) >> 9) | 0x3f800000u;
    return __uint_as_float(bits) - 1.0f;
}
// XORShiftRNG(seed): the seed is signed there (Types32bit::Int), so the
// shift is arithmetic (XORShiftRNG.cpp:21-27, random_number_generator.h:15-19).
// MurmurHash3's 32-bit finaliser.
__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
    h ^= h >> 16; h *= 0x85EBCA6Bu; h ^= h >> 13; h *= 0xC2B2AE35u; h ^= h >> 16;
    return h;
}

// A stream of its own for the transmittance estimate of one light sample:
// hashed from the slot's seed, its sample count and the path length, none of
// which change while the sample's rays are in flight, so the estimate can
// run beside the extension ray's free-path sampling, which owns the slot's
// stream.
__device__ __forceinline__ uint4 shadowStream(uint32_t slotSeed, uint4 info) {
    const uint32_t a = fmix32(slotSeed ^ fmix32(info.z * 0x9E3779B9u ^ (info.y >> 16)));
    uint4 st = make_uint4(fmix32(a ^ 0x1B873593u), fmix32(a ^ 0xCC9E2D51u), fmix32(a ^ 0xE6546B64u),
                          fmix32(a ^ 0x85EBCA6Bu));
    if ((st.x | st.y | st.z | st.w) == 0u) st.x = 1u;
    return st;
}

__device__ __forceinline__ uint4 xsSeed(uint32_t seed) {
    uint32_t st[4];
    int32_t si = (int32_t)seed;
    for (int i = 0; i < 4; ++i) {
        st[i] = 1812433253u * (uint32_t)(si ^ (si >> 30)) + (uint32_t)i;
        si = (int32_t)st[i];
    }
    uint4 s = make_uint4(st[0], st[1], st[2], st[3]);
    for (int i = 0; i < 50; ++i) xsNext(s);
    return s;
}

// x / y, correctly rounded, from r = 1 / y (correctly rounded): q = x r is
// within an ulp of x / y, the remainder x - q y is exact in one FMA, and one
// correction step rounds to the nearest quotient (Markstein's theorem) as
// long as the remainder does not underflow. Checked against IEEE division
// for every float x at y = 470 whose quotient is above 1e-30; on random
// pairs it differs only where the remainder is subnormal. The callers
// divide many values by one divisor (wavelength offsets by a spectrum's
// span, coordinates by the medium's extent, lattice distances by a ray's
// slope), far from that range: one division per divisor instead of one per
// value, with the bits of x / y.
__device__ __forceinline__ float quotient(float x, float y, float r) {
    const float q = x * r;
    return fmaf(fmaf(-q, y, x), r, q);
}

// wavelength i of a path: createWithEqualOffsets (libSLR/BasicTypes/spectrum_types.h:59-69)
__device__ __forceinline__ float lambdaOf(int i, float off) {
    return kLambdaLowD + (kLambdaHighD - kLambdaLowD) * (i + off) / kNumL;
}

// RegularContinuousSpectrum::evaluate (libSLR/BasicTypes/spectrum_types.cpp:82-100)
// at the path's 16 wavelengths, from the spectrum's n samples v over
// [lo, hi]: the first or last sample outside the range, else the linear
// interpolation of the two samples around the wavelength. Two loads per
// wavelength whichever case applies (v may be global memory or LDS).
__device__ __forceinline__ void spectrumValues(const float* v, float lo, float hi, uint32_t n, float off,
                                               float* out) {
    const float span = hi - lo;
    const float rspan = 1.0f / span;
    const float nm1 = (float)(n - 1);
    const int32_t last = (int32_t)n - 1;
#pragma unroll
    for (int i = 0; i < kNumL; ++i) {
        const float binF = quotient(lambdaOf(i, off) - lo, span, rspan) * nm1;
        const bool below = binF <= 0.0f, above = binF >= nm1;
        const int32_t b = below ? 0 : above ? last : (int32_t)binF;
        const float v0 = v[b], v1 = v[min(b + 1, last)];
        const float t = binF - b;
        out[i] = (below || above) ? v0 : (1 - t) * v0 + t * v1;
    }
}
__device__ __forceinline__ void evalSpectrumAt(const DevScene& sc, float lo, float hi, uint32_t offset, uint32_t n,
                                               float off, float* out) {
    spectrumValues(sc.specPool + offset, lo, hi, n, off, out);
}
__device__ __forceinline__ void evalSpectrum(const DevScene& sc, int32_t id, float off, float* out) {
    const DevSpectrum h = sc.spectra[id];
    evalSpectrumAt(sc, h.lo, h.hi, h.offset, h.n, off, out);
}
// spectrum j (0: spec0, 1: spec1, 2: emittance) of material m
__device__ __forceinline__ void evalMatSpectrum(const DevScene& sc, uint32_t m, uint32_t j, float off, float* out) {
    const DevSpectrum h = sc.matSpec[3 * m + j];
    evalSpectrumAt(sc, h.lo, h.hi, h.offset, h.n, off, out);
}

// Vector3D::makeCoordinateSystem (libSLR/BasicTypes/Vector3D.h:61-67): the
// frame a medium point's incoming direction spans (ReferenceFrame(z)).
__device__ __forceinline__ void coordinateSystem(float3 z, float3* x, float3* y) {
    const float sign = z.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sign + z.z);
    const float b = z.x * z.y * a;
    *x = f3(1.0f + sign * z.x * z.x * a, sign * b, -sign * z.x);
    *y = f3(b, sign + z.y * z.y * a, -z.y);
}

struct SurfPt {
    float3 p, gN;         // position, geometric normal (world)
    float3 sx, sy, sz;    // shading frame
    uint32_t mat;
    uint32_t prim;
};

__device__ __forceinline__ float3 toLocal(const SurfPt& sp, const float3& v) {
    return f3(dot3(sp.sx, v), dot3(sp.sy, v), dot3(sp.sz, v));
}
__device__ __forceinline__ float3 fromLocal(const SurfPt& sp, const float3& v) {
    return f3(dot3(f3(sp.sx.x, sp.sy.x, sp.sz.x), v), dot3(f3(sp.sx.y, sp.sy.y, sp.sz.y), v),
              dot3(f3(sp.sx.z, sp.sy.z, sp.sz.z), v));
}

// Instance transforms: object-to-world rows X[0..2], world-to-object rows
// X[3..5]. Vectors go by the linear part, normals by the inverse transpose
// (StaticTransform * Vector3D / Normal3D, libSLR/Core/transform.h:41-49).
__device__ __forceinline__ float3 xfPoint(const float4* X, float3 p) {
    return f3(X[0].x * p.x + X[0].y * p.y + X[0].z * p.z + X[0].w, X[1].x * p.x + X[1].y * p.y + X[1].z * p.z + X[1].w,
              X[2].x * p.x + X[2].y * p.y + X[2].z * p.z + X[2].w);
}
__device__ __forceinline__ float3 xfVector(const float4* X, float3 v) {
    return f3(X[0].x * v.x + X[0].y * v.y + X[0].z * v.z, X[1].x * v.x + X[1].y * v.y + X[1].z * v.z,
              X[2].x * v.x + X[2].y * v.y + X[2].z * v.z);
}
__device__ __forceinline__ float3 xfNormal(const float4* X, float3 n) {
    return f3(X[3].x * n.x + X[4].x * n.y + X[5].x * n.z, X[3].y * n.x + X[4].y * n.y + X[5].y * n.z,
              X[3].z * n.x + X[4].z * n.y + X[5].z * n.z);
}

// A triangle's vertices in world space (inst < 0: already there).
__device__ __forceinline__ void worldVerts(const DevScene& sc, uint32_t prim, int32_t inst, float3* q0, float3* q1,
                                           float3* q2, uint4* pvOut);

// A light sample on a triangle light (TriangleSurfaceShape::sample,
// TriangleSurfaceShape.cpp:357-413, then the instance's transform,
// TransformedSurfaceObject::selectLight): point, geometric normal, shading
// normal, area PDF and material.
struct LightPoint {
    float3 lp, lg, ln;
    float areaPDF;
    uint32_t mat;
    float4 emit;        // header of the light material's emittance spectrum (DevSpectrum bits)
};
__device__ __forceinline__ LightPoint lightPoint(const DevScene& sc, uint32_t li, float lu0, float lu1) {
    const float4* R = sc.lightRec + (size_t)kLightRecRows * li;
    const float4 r0 = R[0], r1 = R[1], r2 = R[2], r3 = R[3], n0 = R[4], n1 = R[5], n2 = R[6], r7 = R[7];
    const float su = sqrtf(lu0);
    const float lb0 = 1.0f - su, lb1 = lu1 * su, lb2 = 1.0f - lb0 - lb1;
    const int32_t linst = __float_as_int(r1.w);
    LightPoint L;
    L.ln = normalize3(lb0 * xyz(n0) + lb1 * xyz(n1) + lb2 * xyz(n2));
    if (linst >= 0) L.ln = normalize3(xfNormal(sc.instXf + kInstXfRows * (uint32_t)linst, L.ln));
    L.lp = lb0 * xyz(r0) + lb1 * xyz(r1) + lb2 * xyz(r2);
    L.lg = xyz(r3);
    L.areaPDF = r2.w;
    L.mat = __float_as_uint(r0.w);
    L.emit = r7;
    return L;
}

__device__ __forceinline__ void worldVerts(const DevScene& sc, uint32_t prim, int32_t inst, float3* q0, float3* q1,
                                           float3* q2, uint4* pvOut) {
    const uint4 pv = sc.primVerts[prim];
    *q0 = xyz(sc.vpos[pv.x]);
    *q1 = xyz(sc.vpos[pv.y]);
    *q2 = xyz(sc.vpos[pv.z]);
    if (inst >= 0) {
        const float4* X = sc.instXf + kInstXfRows * (uint32_t)inst;
        *q0 = xfPoint(X, *q0);
        *q1 = xfPoint(X, *q1);
        *q2 = xfPoint(X, *q2);
    }
    *pvOut = pv;
}

// TriangleSurfaceShape::calculateSurfacePoint (TriangleSurfaceShape.cpp:301-346);
// for a triangle of an instance the object-space point is then taken to
// world space as SurfacePoint::applyTransform does (libSLR/Core/geometry.cpp:
// 40-45, 73-77): frame vectors by the transform, the normal and geometric
// normal by its inverse transpose, each renormalised.
__device__ __forceinline__ SurfPt surfaceAt(const DevScene& sc, uint32_t prim, float b0, float b1, float3 o, float3 d,
                                            float t, int32_t inst = -1) {
    SurfPt s;
    // the triangle's shading record: one 128-B line (k_tri_shade)
    const float4* R = sc.triShade + (size_t)kTriShadeRows * prim;
    const float4 r0 = R[0], r1 = R[1], r2 = R[2], r3 = R[3], r4 = R[4], r5 = R[5], r6 = R[6];
    const float3 n0 = xyz(r0), n1 = xyz(r1), n2 = xyz(r2);
    const float3 t0 = xyz(r3), t1 = xyz(r4), t2 = xyz(r5);
    const float b2 = 1.0f - b0 - b1;
    s.sz = normalize3(b0 * n0 + b1 * n1 + b2 * n2);
    float3 x = b0 * t0 + b1 * t1 + b2 * t2;
    const float dNT = dot3(s.sz, x);
    s.sx = normalize3(x - dNT * s.sz);
    s.sy = cross3(s.sz, s.sx);
    s.p = o + d * t;
    s.gN = xyz(r6);
    s.mat = __float_as_uint(r0.w);
    s.prim = prim;
    if (inst >= 0) {
        const float4* X = sc.instXf + kInstXfRows * (uint32_t)inst;
        s.sx = normalize3(xfVector(X, s.sx));
        s.sy = normalize3(xfVector(X, s.sy));
        s.sz = normalize3(xfNormal(X, s.sz));
        s.gN = normalize3(xfNormal(X, s.gN));
        const int32_t om = sc.instMat[inst];
        if (om >= 0) s.mat = (uint32_t)om;
    }
    return s;
}

// World-space area (the reference's light area PDF uses the object-space
// area, which differs only under a scaling instance transform).
__device__ __forceinline__ float triArea(const DevScene& sc, uint32_t prim, int32_t inst = -1) {
    float3 p0, p1, p2;
    uint4 pv;
    worldVerts(sc, prim, inst, &p0, &p1, &p2, &pv);
    return 0.5f * length3(cross3(p1 - p0, p2 - p0));
}

// DiscreteDistribution1D::sample (libSLR/Core/distributions.cpp:138-161)
__device__ __forceinline__ uint32_t sampleLight(const DevScene& sc, float u, float* prob) {
    const uint32_t n = sc.numLights;
    int idx = (int)n;
    uint32_t d = 1;
    while ((d << 1) <= n) d <<= 1;   // prevPowerOf2
    for (; d > 0; d >>= 1)
        if (idx - (int)d > 0 && sc.lightCdf[idx - d] >= u) idx -= (int)d;
    --idx;
    *prob = sc.lightPmf[idx];
    return (uint32_t)idx;
}

// Probability of choosing the surface lights over other emitters
// (Scene::selectSurfaceLight, libSLR/Scene/Scene.cpp:137-155): 1 without an
// environment light.
__device__ __forceinline__ float lightSelectScale(const DevScene& sc) {
    if (!sc.hasEnv) return 1.0f;
    // InfiniteSphereSurfaceObject::importance() is 1 (surface_object.cpp:197-199)
    return sc.lightImportance / (sc.lightImportance + 1.0f);
}
// ... and of choosing the environment light.
__device__ __forceinline__ float envSelectProb(const DevScene& sc) { return 1.0f / (sc.lightImportance + 1.0f); }

}  // namespace slrh
