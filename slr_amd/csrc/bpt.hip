// Bidirectional path tracing on the device: BPTRenderer
// (libSLR/Renderer/BPTRenderer.cpp:25-432) as a streaming wavefront over
// slots. A slot owns one pixel and one film plane and renders its share of
// the pixel's samples one after another: light subpath, eye subpath with its
// connections, film add, next sample. Slots in either phase share every
// step, so the steps stay full until the last samples drain.
//
//   trace  : closest hit of the light and eye rays, any hit of the
//            light-to-lens rays and of the connection rays (one launch)
//   light  : each unoccluded light-to-lens ray adds its contribution to the
//            light image (t = 1; the pixel is where the ray crosses the image
//            plane, so these adds use FP64 atomics); each light-ray hit
//            becomes the subpath's next stored vertex, gets its own
//            light-to-lens ray, and samples the next direction with Russian
//            roulette; a subpath that ends starts the sample's eye subpath
//            (the camera ray)
//   gather : each slot adds its unoccluded connections (s >= 1) to the
//            sample's radiance; a sample whose eye subpath ended is filed
//            into the slot's film plane (the slot owns it: no atomics) and
//            the slot's next sample starts (light vertex, first light ray)
//   eye    : each eye-ray hit becomes the eye vertex: emission with its MIS
//            weight (s = 0), the next direction or the end of the subpath
//   connect: one connection per stored light vertex (a shadow ray each)
//
// A slot draws its samples from its own random stream in a fixed order, so
// the film planes do not depend on how the steps interleave the slots.
//
// MIS weights are the reference's power heuristic over every (s, t) with
// the subpaths' area PDFs and Russian-roulette probabilities
// (calculateMISWeight, BPTRenderer.cpp:383-431). For a connection the
// reference walks both subpaths from the connected ends down, multiplying
// ratios of reverse to forward PDFs. The two end vertices of each side need
// the connection's own PDFs; every term below them depends only on the
// subpath, so each stored vertex k carries
//   tail[k] = T[k-2],  T[j] = (Q[j] / P[j])^2 ([!delta_j && !delta_j+1] + T[j-1])
// with P[j] = areaPDF * RRProb of vertex j as sampled and Q[j] the same for
// the reverse direction (revAreaPDF * revRRProb), T below the side's
// shortest subpath (1 eye vertex, 0 light vertices) zero. The weight is then
// 1 / (1 + S_eye + S_light) with
//   S = (A1 / P[n-1])^2 ([!delta_n-1] + [n-1 > m] (A2 / P[n-2])^2 ([..] + tail[n-1]))
// where A1, A2 are the other side's extension PDFs at the connection.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "bsdf.h"
#include "bsdf_bidir.h"
#include "device_scene.h"
#include "engine.h"
#include "shading.h"
#include "traverse.h"

namespace slrh {

namespace {

constexpr uint32_t kBptBlock = 256;
// Light subpath vertices stored per sample (vertex 0 on the light). The
// reference keeps every vertex; a subpath that would need more stops there
// and the render statistics count it (BPT stat T_LIGHT_CAP).
constexpr uint32_t kLightCap = 64;
// Eye subpath vertices at most (the reference has no cap; Russian roulette
// ends paths long before; the statistics count any path stopped here).
constexpr uint32_t kEyeCap = 256;
constexpr uint32_t kVtxRows = 9;              // a stored vertex: 9 float4 (144 B)
constexpr uint32_t kEyeRows = kVtxRows + 2;   // the eye vertex, then its predecessor's position/PDF and normal/flags
constexpr uint32_t kConnRows = 6;             // contribution (4 float4), shadow ray (2 float4)
constexpr uint32_t kSplatRows = 7;            // the same and the pixel
// trace entries: bit 31 any hit, bit 30 a light-to-lens ray (else a
// connection ray); connection entries index slot * kLightCap + vertex, so
// slots are limited to 2^24
constexpr uint32_t kShadowBit = 0x80000000u;
constexpr uint32_t kSplatBit = 0x40000000u;
constexpr uint32_t kMaxBptSlots = 1u << 24;
constexpr int kBptLdsDepth = 16;
constexpr uint32_t kBptTreelet = 128;
constexpr uint32_t kTargetSlots = 1u << 22;   // slots when the caller leaves it open (about 64 GB)
constexpr uint32_t kStepsPerRead = 4;        // wavefront steps between reads of the queue lengths

// vertex flags (row 3 w)
constexpr uint32_t V_DELTA = 1u;   // sampled by a delta distribution
constexpr uint32_t V_EDF = 2u;     // the light vertex: its DDF is the light's diffuse EDF
// slot flags (aux.y bits 0-7; bits 8-15 the hero wavelength, 16-31 the
// number of samples the slot has started)
constexpr uint32_t B_SAMPLE = 1u;  // the slot holds a sample
constexpr uint32_t B_END = 2u;     // its eye subpath ended: file after the last connections
constexpr uint32_t B_CONNECT = 4u; // its eye vertex scatters: connect it to the light subpath
constexpr uint32_t kMaxSlotSamples = 0xFFFFu;
// carried flags (mis.y): delta of the last stored vertex and of the one being traced to
constexpr uint32_t M_DELTA_LAST = 1u, M_DELTA_NEXT = 2u;

enum BptCounter : uint32_t {
    K_LE = 0,     // +0/1: light-ray queue lengths (ping-pong by step parity)
    K_EE = 2,     // +0/1: eye-ray queue lengths
    K_SP = 4,     // +0/1: light-to-lens ray queue lengths
    K_CN = 6,     // +0/1: connection ray queue lengths
    K_GA = 8,     // +0/1: gather queue lengths (slots whose eye vertex was just made)
    K_NONFINITE = 11,
    K_OVERFLOW = 12,
    kBptCounters = 16
};
enum BptStatId : uint32_t {
    T_EXT_RAYS = 0, T_SH_RAYS, T_NODES_EXT, T_TRIS_EXT, T_NODES_SH, T_TRIS_SH, T_SAMPLES, T_LIGHT_VERTS,
    T_EYE_VERTS, T_CONNECT, T_SPLAT, T_LIGHT_CAP, T_EYE_CAP, T_NODES_LDS, T_INSTS, kBptStats
};

struct BptState {
    uint32_t n, pixels, planes, spp, maxEye;
    float4* gen;        // n x 4: ray origin + tmin, direction + tmax, hit (t, prim, b0, b1),
                        // carried (direction PDF, RR probability, cosine at the origin, P of the origin vertex)
    uint4* aux;         // n: pixel, flags | hero << 8, light vertices stored | eye vertex index << 8, hit instance
    float4* mis;        // n: T of the origin vertex's predecessor, carried flags, wavelength offset, 0
    uint4* rng;         // n: XORShift state
    float4* alpha;      // n x 4: throughput of the vertex being traced to
    float4* rad;        // n x 4: the sample's radiance for its own pixel
    float4* smp;        // n: pixel position x, y; lens point in the lens frame x, y
    float4* lens;       // n: lens point (world), lens area PDF
    float4* lv;         // n x kLightCap x kVtxRows: light subpath vertices
    float4* ev;         // n x kEyeRows: the eye vertex (kVtxRows), then its predecessor's
                        // position | P and geometric normal | flags
    float4* conn;       // n x kLightCap x kConnRows
    float4* splat;      // 2 x n x kSplatRows: light-to-lens rays, one buffer per step parity (a
                        // step reads the rays the trace before it tested and writes the next ones)
    uint8_t* vis;       // n x kLightCap: connection ray unoccluded
    uint8_t* visSplat;  // 2 x n: light-to-lens ray unoccluded, per step parity
    uint32_t* q;
    uint32_t* ctr;
    unsigned long long* stat;
    double* planeSums;  // [planes][pixels][16]
    double* lightImg;   // [pixels][16]
    uint32_t* spill;
    uint32_t spillStride;
    uint32_t* fetch;    // the trace's claim cursors, one per part, kFetchStride words apart
};

// queues of slot indices (connection entries: slot * kLightCap + light
// vertex), two of each kind, one per step parity
__device__ __forceinline__ uint32_t* qLight(const BptState& b, uint32_t p) { return b.q + (size_t)p * b.n; }
__device__ __forceinline__ uint32_t* qEye(const BptState& b, uint32_t p) { return b.q + (size_t)(2 + p) * b.n; }
__device__ __forceinline__ uint32_t* qGather(const BptState& b, uint32_t p) { return b.q + (size_t)(4 + p) * b.n; }
__device__ __forceinline__ uint32_t* qSplat(const BptState& b, uint32_t p) { return b.q + (size_t)(6 + p) * b.n; }
__device__ __forceinline__ uint32_t* qConn(const BptState& b, uint32_t p) {
    return b.q + 8 * (size_t)b.n + (size_t)p * b.n * kLightCap;
}
__host__ __device__ __forceinline__ size_t queueWords(uint32_t n) { return (size_t)n * (8 + 2 * kLightCap); }
__device__ __forceinline__ float4* splatRows(const BptState& b, uint32_t parity, uint32_t s) {
    return b.splat + ((size_t)parity * b.n + s) * kSplatRows;
}

// std::min(x, 1.0f) as the reference computes RR probabilities: NaN stays NaN
__device__ __forceinline__ float rrMin(float x) { return 1.0f < x ? 1.0f : x; }

struct Vtx {
    float3 p, gN, sz, sx, dir;
    float P, cosIn, tail;
    uint32_t flags, mat;
};

__device__ __forceinline__ void storeVtx(float4* r, const Vtx& v, const float* alpha) {
    r[0] = make_float4(v.p.x, v.p.y, v.p.z, v.P);
    r[1] = make_float4(v.gN.x, v.gN.y, v.gN.z, v.cosIn);
    r[2] = make_float4(v.sz.x, v.sz.y, v.sz.z, v.tail);
    r[3] = make_float4(v.sx.x, v.sx.y, v.sx.z, __uint_as_float(v.flags));
    r[4] = make_float4(v.dir.x, v.dir.y, v.dir.z, __uint_as_float(v.mat));
#pragma unroll
    for (int p = 0; p < 4; ++p)
        r[5 + p] = make_float4(alpha[4 * p], alpha[4 * p + 1], alpha[4 * p + 2], alpha[4 * p + 3]);
}
__device__ __forceinline__ Vtx loadVtx(const float4* r) {
    Vtx v;
    const float4 a = r[0], b = r[1], c = r[2], d = r[3], e = r[4];
    v.p = xyz(a); v.P = a.w;
    v.gN = xyz(b); v.cosIn = b.w;
    v.sz = xyz(c); v.tail = c.w;
    v.sx = xyz(d); v.flags = __float_as_uint(d.w);
    v.dir = xyz(e); v.mat = __float_as_uint(e.w);
    return v;
}
__device__ __forceinline__ void load16(const float4* r, float* v) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const float4 a = r[p];
        v[4 * p] = a.x; v[4 * p + 1] = a.y; v[4 * p + 2] = a.z; v[4 * p + 3] = a.w;
    }
}
__device__ __forceinline__ void store16(float4* r, const float* v) {
#pragma unroll
    for (int p = 0; p < 4; ++p) r[p] = make_float4(v[4 * p], v[4 * p + 1], v[4 * p + 2], v[4 * p + 3]);
}

__device__ __forceinline__ float3 vtxLocal(const Vtx& v, const float3& w) {
    const float3 sy = cross3(v.sz, v.sx);
    return f3(dot3(v.sx, w), dot3(sy, w), dot3(v.sz, w));
}

// The scattering model's spectra at the path's wavelengths: reflectance, or
// the two complex-IOR / IOR spectra.
__device__ __forceinline__ void materialSpectra(const DevScene& sc, uint32_t mat, uint32_t type, float off, float* A,
                                                float* B) {
    evalMatSpectrum(sc, mat, 0, off, A);
    if (type == MAT_GGX_CONDUCTOR || type == MAT_GGX_DIELECTRIC) evalMatSpectrum(sc, mat, 1, off, B);
}

// A stored vertex's DDF toward dl (its shading frame): value, PDF and the
// reverse PDF. The light vertex carries the light's diffuse EDF
// (DiffuseEDF::evaluate/evaluatePDF, libSLR/EDF/basic_edfs.cpp:20-30; its
// proxy leaves the reverse unset, BPTRenderer.h:30-37), other vertices their
// BSDF with the subpath's query (BSDFProxy, BPTRenderer.h:40-52).
template <bool kDiel>
__device__ __forceinline__ void vertexDdf(const DevScene& sc, const Vtx& v, bool adjoint, int hero, float off,
                                          const float3& dl, float* fs, float* pdf, float* rev,
                                          const float* spA = nullptr, const float* spB = nullptr) {
    *rev = 0.0f;
    if (v.flags & V_EDF) {
        const float e = dl.z > 0.0f ? (float)(1.0 / 3.14159265358979323846) : 0.0f;
#pragma unroll
        for (int k = 0; k < kNumL; ++k) fs[k] = e;
        *pdf = dl.z > 0.0f ? (float)(dl.z / 3.14159265358979323846) : 0.0f;
        return;
    }
    const DevMaterial m = sc.mats[v.mat];
    if (m.type == MAT_NONE) {
#pragma unroll
        for (int k = 0; k < kNumL; ++k) fs[k] = 0.0f;
        *pdf = 0.0f;
        return;
    }
    // the vertex's material spectra, unless the caller has them already
    float A[kNumL], B[kNumL];
    if (spA) {
#pragma unroll
        for (int k = 0; k < kNumL; ++k) {
            A[k] = spA[k];
            B[k] = spB[k];
        }
    } else {
        materialSpectra(sc, v.mat, m.type, off, A, B);
    }
    const BiBsdf<kDiel> b{m.type, A, B, m.p0, m.p1};
    BiQuery q;
    q.dir = v.dir;
    q.ng = vtxLocal(v, v.gN);
    q.hero = hero;
    q.adjoint = adjoint;
    b.evaluate(q, dl, fs);
    *pdf = b.pdf(q, dl, rev);
}

// One side's sum of squared PDF ratios for shortening it by its end vertex
// and the one below (extendAndShorten, BPTRenderer.cpp:386-418), the rest of
// the side carried in the end vertex's tail. a1, a2: the other side's
// extension PDFs (area measure times RR probability) at the end vertex and
// at the one below; below: whether the one below may be shortened too.
__device__ __forceinline__ float sideSum(float a1, float endP, uint32_t endFlags, bool below, float a2, float prevP,
                                         uint32_t prevFlags, float tail) {
    float r = a1 / endP;
    const bool dEnd = (endFlags & V_DELTA) != 0;
    float s = dEnd ? 0.0f : r * r;
    if (below) {
        r *= a2 / prevP;
        const bool dPrev = (prevFlags & V_DELTA) != 0;
        if (!dPrev && !dEnd) s += r * r;
        s += r * r * tail;
    }
    return s;
}

// SpectrumStorage::add's binning (libSLR/BasicTypes/spectrum_types.h:725-736,
// ImageSensor::add, libSLR/Core/ImageSensor.cpp:125-138): wavelength k goes
// to stratum k or k + 1, scaled by 16 / 470.
__device__ __forceinline__ void binStrata(float off, const float* c, float* add) {
    const float recBinWidth = kNumL / (kLambdaHighD - kLambdaLowD);
#pragma unroll
    for (int k = 0; k < kNumL; ++k) add[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < kNumL; ++k) {
        const float l = lambdaOf(k, off);
        const float pos =
            quotient(l - kLambdaLowD, kLambdaHighD - kLambdaLowD, 1.0f / (kLambdaHighD - kLambdaLowD));
        const uint32_t bin = min((uint32_t)(pos * kNumL), (uint32_t)(kNumL - 1));
        const float v = c[k] * recBinWidth;
        add[k] += bin == (uint32_t)k ? v : 0.0f;
        if (k + 1 < kNumL) add[k + 1] += bin == (uint32_t)(k + 1) ? v : 0.0f;
    }
}

struct LensFrame {
    float3 x, y, z;
};
// The lens's shading frame (PerspectiveCamera::sample, PerspectiveCamera.cpp:31-57)
__device__ __forceinline__ LensFrame lensFrame(const DevCamera& cam) {
    LensFrame f;
    const float* M = cam.l2w;
    const float* N = cam.nrm;
    f.z = normalize3(f3(N[2], N[5], N[8]));
    f.x = normalize3(f3(M[0], M[4], M[8]));
    f.y = cross3(f.z, f.x);
    return f;
}

// Connection of a light vertex to the lens (t = 1; BPTRenderer.cpp:173-273
// with the IDF proxy): the contribution for the pixel the direction crosses,
// and the shadow ray from the lens (Scene::testVisibility, Scene.cpp:96-112).
// Lp0/Lp1: rows 0 and 1 of the vertex below (when hasPrev), pFlags its flags.
template <bool kDiel>
__device__ __forceinline__ bool connectLens(const DevScene& sc, const DevCamera& cam, const LensFrame& lf,
                                            const float4& lens, float2 orgL, const Vtx& L, const float* lAlpha,
                                            bool hasPrev, float4 Lp0, float4 Lp1, uint32_t pFlags, int hero,
                                            float off, float* c, uint32_t* pixel, float4* ro, float4* rd,
                                            const float* spA = nullptr, const float* spB = nullptr) {
    const float3 pe = xyz(lens);
    const float3 dv = L.p - pe;
    const float dist2 = dot3(dv, dv);
    const float3 cv = divs(dv, sqrtf(dist2));
    const float cosL = fabsf(dot3(cv, L.gN));
    const float cosE = fabsf(dot3(cv, lf.z));
    const float G = cosE * cosL / dist2;
    // the lens's IDF (PerspectiveIDF::evaluate/evaluatePDF/calculatePixel,
    // PerspectiveCamera.cpp:98-131)
    const float3 e = f3(dot3(lf.x, cv), dot3(lf.y, cv), dot3(lf.z, cv));
    const float s = cam.objDist / e.z;
    const float pfx = e.x * s + orgL.x, pfy = e.y * s + orgL.y;
    const bool valid = pfx >= -cam.opWidth * 0.5f && pfx <= cam.opWidth * 0.5f && pfy >= -cam.opHeight * 0.5f &&
                       pfy <= cam.opHeight * 0.5f && e.z >= 0;
    if (!valid) return false;
    const float eExt1DirPDF = cam.imgDist * cam.imgDist / ((e.z * e.z * e.z) * cam.imgPlaneArea);
    float lDDF[kNumL], lExt1DirPDF, eExt2DirPDF;
    vertexDdf<kDiel>(sc, L, true, hero, off, vtxLocal(L, -cv), lDDF, &lExt1DirPDF, &eExt2DirPDF, spA, spB);
    bool any = false;
#pragma unroll
    for (int k = 0; k < kNumL; ++k) any = any || lDDF[k] * G != 0.0f;
    if (!any) return false;
    // light side: eye extension PDFs at the light vertex and the one below
    const float eExt1Area = eExt1DirPDF * cosL / dist2;
    float eExt2 = 0.0f;
    if (hasPrev) {
        const float3 d2 = xyz(Lp0) - L.p;
        const float dd2 = dot3(d2, d2);
        const float3 dir2 = divs(d2, sqrtf(dd2));
        const float eExt2Area = eExt2DirPDF * fabsf(dot3(dir2, xyz(Lp1))) / dd2;
        float t[kNumL];
#pragma unroll
        for (int k = 0; k < kNumL; ++k) t[k] = lDDF[k] * L.cosIn / eExt2DirPDF;
        eExt2 = eExt2Area * rrMin(importance16(t, hero));
    }
    const float sL = sideSum(eExt1Area, L.P, L.flags, hasPrev, eExt2, Lp0.w, pFlags, L.tail);
    const float w = 1.0f / (1.0f + sL);
    if (isnan(w) || isinf(w)) return false;
    // eye vertex 0: We0 / (lens area PDF * wavelength selection PDF), We0 = 1
    const float eAlpha = 1.0f / (lens.w * (kNumL / (kLambdaHighD - kLambdaLowD)));
    bool finite = true;
    any = false;
#pragma unroll
    for (int k = 0; k < kNumL; ++k) {
        c[k] = w * lAlpha[k] * (lDDF[k] * G) * eAlpha;
        finite = finite && isfinite(c[k]);
        any = any || c[k] != 0.0f;
    }
    if (!finite || !any) return false;
    const float hx = cam.width * (0.5f - pfx / cam.opWidth);
    const float hy = cam.height * (0.5f - pfy / cam.opHeight);
    const uint32_t ix = min((uint32_t)hx, cam.width - 1), iy = min((uint32_t)hy, cam.height - 1);
    *pixel = iy * cam.width + ix;
    const float dist = sqrtf(dist2);
    const float3 dir = divs(dv, dist);
    *ro = make_float4(pe.x, pe.y, pe.z, 1e-4f);
    *rd = make_float4(dir.x, dir.y, dir.z, dist * (1 - 1e-4f));
    return true;
}

// Connection of eye vertex E (t >= 2) with light vertex L (s >= 1),
// BPTRenderer.cpp:173-273. Ep0/Ep1 and Lp0/Lp1: rows 0 and 1 of the vertices
// below (E's always exists; L's when lHasPrev); eFlagsPrev, lFlagsPrev their
// flags; eAlpha, lAlpha the two vertices' throughputs (read at the end). The
// RR probabilities of the extensions are importances of the DDF values
// scaled by a cosine over a PDF; importance is linear, so each DDF's
// importance is taken once.
template <bool kDiel>
__device__ __forceinline__ bool connectVertices(const DevScene& sc, const Vtx& E, const float4* eAlpha, bool eBelow,
                                                float4 Ep0, float4 Ep1, uint32_t eFlagsPrev, const Vtx& L,
                                                const float4* lAlpha, bool lHasPrev, float4 Lp0, float4 Lp1,
                                                uint32_t lFlagsPrev, int hero, float off, float* c, float4* ro,
                                                float4* rd) {
    const float3 dv = L.p - E.p;
    const float dist2 = dot3(dv, dv);
    const float3 cv = divs(dv, sqrtf(dist2));
    const float cosL = fabsf(dot3(cv, L.gN));
    const float cosE = fabsf(dot3(cv, E.gN));
    const float G = cosE * cosL / dist2;
    float lExt1DirPDF, eExt2DirPDF, eExt1DirPDF, lExt2DirPDF;
    vertexDdf<kDiel>(sc, L, true, hero, off, vtxLocal(L, -cv), c, &lExt1DirPDF, &eExt2DirPDF);
    bool any = false;
#pragma unroll
    for (int k = 0; k < kNumL; ++k) any = any || c[k] != 0.0f;
    if (!any) return false;
    const float impL = importance16(c, hero);
    float eDDF[kNumL];
    vertexDdf<kDiel>(sc, E, false, hero, off, vtxLocal(E, cv), eDDF, &eExt1DirPDF, &lExt2DirPDF);
    any = false;
#pragma unroll
    for (int k = 0; k < kNumL; ++k) {
        c[k] = c[k] * G * eDDF[k];
        any = any || c[k] != 0.0f;
    }
    if (!any) return false;
    const float impE = importance16(eDDF, hero);
    // light subpath extended through the connection to E and below
    const float lExt1Area = lExt1DirPDF * cosE / dist2;
    const float lExt1RR = (L.flags & V_EDF) ? 1.0f : rrMin(impL * cosL / lExt1DirPDF);
    float lExt2;
    {
        const float3 d2 = E.p - xyz(Ep0);
        const float dd2 = dot3(d2, d2);
        const float3 dir2 = divs(d2, sqrtf(dd2));
        const float lExt2Area = lExt2DirPDF * fabsf(dot3(dir2, xyz(Ep1))) / dd2;
        lExt2 = lExt2Area * rrMin(impE * E.cosIn / lExt2DirPDF);
    }
    // eye subpath extended through the connection to L and below
    const float eExt1Area = eExt1DirPDF * cosL / dist2;
    const float eExt1RR = rrMin(impE * cosE / eExt1DirPDF);
    float eExt2 = 0.0f;
    if (lHasPrev) {
        const float3 d2 = xyz(Lp0) - L.p;
        const float dd2 = dot3(d2, d2);
        const float3 dir2 = divs(d2, sqrtf(dd2));
        const float eExt2Area = eExt2DirPDF * fabsf(dot3(dir2, xyz(Lp1))) / dd2;
        eExt2 = eExt2Area * rrMin(impL * L.cosIn / eExt2DirPDF);
    }
    const float sE = sideSum(lExt1Area * lExt1RR, E.P, E.flags, eBelow, lExt2, Ep0.w, eFlagsPrev, E.tail);
    const float sL = sideSum(eExt1Area * eExt1RR, L.P, L.flags, lHasPrev, eExt2, Lp0.w, lFlagsPrev, L.tail);
    const float w = 1.0f / (1.0f + sE + sL);
    if (isnan(w) || isinf(w)) return false;
    bool finite = true;
    any = false;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const float4 la = lAlpha[p], ea = eAlpha[p];
        const float l4[4] = {la.x, la.y, la.z, la.w}, e4[4] = {ea.x, ea.y, ea.z, ea.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = 4 * p + q;
            c[k] = w * l4[q] * c[k] * e4[q];
            finite = finite && isfinite(c[k]);
            any = any || c[k] != 0.0f;
        }
    }
    if (!finite || !any) return false;
    const float dist = sqrtf(dist2);
    const float3 dir = divs(dv, dist);
    *ro = make_float4(E.p.x, E.p.y, E.p.z, 1e-4f);
    *rd = make_float4(dir.x, dir.y, dir.z, dist * (1 - 1e-4f));
    return true;
}

// ------------------------------------------------------------------ kernels

__global__ void k_bpt_seed(BptState b, uint32_t seed) {
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < b.n; s += gridDim.x * blockDim.x) {
        uint32_t h = seed ^ (0x9E3779B9u * (s + 1u));
        h ^= h >> 16; h *= 0x85EBCA6Bu; h ^= h >> 13; h *= 0xC2B2AE35u; h ^= h >> 16;
        b.rng[s] = xsSeed(h);
        b.aux[s] = make_uint4(s % b.pixels, 0u, 0u, 0u);
        b.visSplat[s] = 0;
        b.visSplat[(size_t)b.n + s] = 0;
    }
}

// The slot's next sample while it has one left (slot s renders samples
// plane, plane + planes, ... of its pixel, plane = s / pixels): pixel
// position, wavelengths, lens point, the light vertex
// (Scene::selectSurfaceLight, SurfaceLight::sampleRay,
// BPTRenderer.cpp:121-148) with its connection to the lens (written for the
// trace of step parity `parity`) and the first light ray. Returns whether a
// sample started; *splat whether its lens ray is to be traced.
template <bool kDiel>
__device__ __forceinline__ bool startSample(const DevScene& sc, const BptState& b, const DevCamera& cam,
                                            const LensFrame& lf, uint32_t s, uint32_t parity, uint4& aux,
                                            bool* splat) {
    *splat = false;
    const uint32_t plane = s / b.pixels, pixel = aux.x;
    const uint32_t started = aux.y >> 16;
    const uint32_t quota = (b.spp - plane + b.planes - 1) / b.planes;
    if (started >= quota) {
        aux.y = started << 16;
        aux.z = 0;
        b.aux[s] = aux;
        return false;
    }
    uint4 rng = b.rng[s];
    const uint32_t px = pixel % cam.width, py = pixel / cam.width;
    const float fx = px + xsFloat(rng), fy = py + xsFloat(rng);
    const float off = xsFloat(rng);
    const float uL = xsFloat(rng);
    const uint32_t hero = min((uint32_t)(uint16_t)(kNumL * uL), (uint32_t)(kNumL - 1));
    // light vertex
    float lprob;
    const uint32_t li = sampleLight(sc, xsFloat(rng), &lprob);
    const float lu0 = xsFloat(rng), lu1 = xsFloat(rng);
    const LightPoint lp = lightPoint(sc, li, lu0, lu1);
    float Le[kNumL];
    spectrumValues(sc.specPool + __float_as_uint(lp.emit.z), lp.emit.x, lp.emit.y,
                   __float_as_uint(lp.emit.w), off, Le);
    const float P0 = lprob * lp.areaPDF;
    float alpha0[kNumL];
    const float rP0 = 1.0f / P0;
#pragma unroll
    for (int k = 0; k < kNumL; ++k) alpha0[k] = Le[k] * rP0;
    float3 fxv, fyv;
    coordinateSystem(lp.ln, &fxv, &fyv);
    const float3 d = cosineHemisphere(xsFloat(rng), xsFloat(rng));
    const float dirPDF = (float)(d.z / 3.14159265358979323846);
    const float3 dir = d.x * fxv + d.y * fyv + d.z * lp.ln;
    Vtx L;
    L.p = lp.lp; L.P = P0; L.gN = lp.lg; L.cosIn = 0.0f; L.sz = lp.ln; L.tail = 0.0f; L.sx = fxv;
    L.flags = V_EDF; L.dir = f3(0, 0, 0); L.mat = lp.mat;
    storeVtx(b.lv + (size_t)s * kLightCap * kVtxRows, L, alpha0);
    // the lens point (PerspectiveCamera::sample)
    float lx = 0, ly = 0;
    concentricDisk(xsFloat(rng), xsFloat(rng), &lx, &ly);
    const float2 orgL = make_float2(cam.lensRadius * lx, cam.lensRadius * ly);
    const float* M = cam.l2w;
    const float3 pe = f3(M[0] * orgL.x + M[1] * orgL.y + M[3], M[4] * orgL.x + M[5] * orgL.y + M[7],
                         M[8] * orgL.x + M[9] * orgL.y + M[11]);
    const float lensP = cam.lensRadius > 0.0f
                            ? (float)(1.0 / (3.14159265358979323846 * cam.lensRadius * cam.lensRadius))
                            : 1.0f;
    const float4 lens = make_float4(pe.x, pe.y, pe.z, lensP);
    b.lens[s] = lens;
    b.smp[s] = make_float4(fx, fy, orgL.x, orgL.y);
    // the first light ray and its throughput: alpha0 * Le1 * cos / dirPDF
    const float sc1 = (float)(1.0 / 3.14159265358979323846) * (fabsf(dot3(dir, lp.lg)) / dirPDF);
    float a1[kNumL];
#pragma unroll
    for (int k = 0; k < kNumL; ++k) a1[k] = alpha0[k] * sc1;
    store16(b.alpha + (size_t)s * 4, a1);
    float4* g = b.gen + (size_t)s * 4;
    g[0] = make_float4(lp.lp.x, lp.lp.y, lp.lp.z, 1e-4f);
    g[1] = make_float4(dir.x, dir.y, dir.z, INFINITY);
    g[3] = make_float4(dirPDF, 1.0f, d.z, P0);
    b.mis[s] = make_float4(0.0f, __uint_as_float(0u), off, 0.0f);
    aux = make_uint4(pixel, B_SAMPLE | (hero << 8) | ((started + 1) << 16), 1u, 0u);
    b.aux[s] = aux;
    float z[kNumL];
#pragma unroll
    for (int k = 0; k < kNumL; ++k) z[k] = 0.0f;
    store16(b.rad + (size_t)s * 4, z);
    // light vertex to lens (s = 1, t = 1)
    float c[kNumL];
    uint32_t pix = 0;
    float4 ro, rd;
    const bool sh = connectLens<kDiel>(sc, cam, lf, lens, orgL, L, alpha0, false, make_float4(0, 0, 0, 1),
                                make_float4(0, 0, 0, 0), 0u, (int)hero, off, c, &pix, &ro, &rd);
    if (sh) {
        float4* sp = splatRows(b, parity, s);
        store16(sp, c);
        sp[4] = ro;
        sp[5] = rd;
        sp[6] = make_float4(__uint_as_float(pix), 0, 0, 0);
    }
    b.rng[s] = rng;
    *splat = sh;
    return true;
}

// Every slot's first sample.
template <bool kDiel>
__global__ __launch_bounds__(kBptBlock) void k_bpt_start(DevScene sc, BptState b, DevCamera cam) {
    const LensFrame lf = lensFrame(cam);
    if (blockIdx.x == 0 && threadIdx.x < SLRH_FETCH_PARTS) b.fetch[threadIdx.x * kFetchStride] = 0;
    for (uint32_t base = blockIdx.x * kBptBlock; base < b.n; base += gridDim.x * kBptBlock) {
        const uint32_t s = base + threadIdx.x;
        bool has = false, sh = false;
        if (s < b.n) {
            uint4 aux = b.aux[s];
            has = startSample<kDiel>(sc, b, cam, lf, s, 0u, aux, &sh);
        }
        const uint32_t shPos = waveAppend(&b.ctr[K_SP + 0], sh);
        if (sh) qSplat(b, 0)[shPos] = s;
        const uint32_t ePos = waveAppend(&b.ctr[K_LE + 0], has);
        if (has) qLight(b, 0)[ePos] = s;
    }
}

// Closest hit for the light and eye rays, any hit for the light-to-lens and
// connection rays, in one persistent launch (the traversal of the path
// tracer).
template <bool I>
__global__ __launch_bounds__(kBptBlock) void k_bpt_trace(DevScene sc, BptState b, uint32_t cur) {
    __shared__ uint32_t lds[kBptLdsDepth * kBptBlock];
    __shared__ float4 treelet[kBptTreelet * kTreeletRows];
    const uint32_t nxt = cur ^ 1u;
    const uint32_t nl = b.ctr[K_LE + cur], ne = b.ctr[K_EE + cur], nsp = b.ctr[K_SP + cur],
                   ncn = b.ctr[K_CN + cur];
    const uint32_t gtid = blockIdx.x * kBptBlock + threadIdx.x;
    TraceStack st;
    st.lds = ldsColumn(lds, threadIdx.x);
    st.ldsStride = kBptBlock;
    st.spill = b.spill + gtid;
    st.spillStride = b.spillStride;
    st.ldsDepth = kBptLdsDepth;
    st.treeletNodes = kBptTreelet < sc.numNodes ? kBptTreelet : sc.numNodes;
    st.treelet = loadTreelet(treelet, sc.nodes, st.treeletNodes);
    // the next step appends to the other queues: they start empty (their
    // contents were consumed before this launch)
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        b.ctr[K_LE + nxt] = 0;
        b.ctr[K_EE + nxt] = 0;
        b.ctr[K_SP + nxt] = 0;
        b.ctr[K_CN + nxt] = 0;
        b.ctr[K_GA + nxt] = 0;
    }
    const uint32_t* ql = qLight(b, cur);
    const uint32_t* qe = qEye(b, cur);
    const uint32_t* qs = qSplat(b, cur);
    const uint32_t* qc = qConn(b, cur);
    const uint32_t n1 = nl, n2 = n1 + ne, n3 = n2 + nsp;
    // a short queue (the last steps of a frame) goes out in smaller chunks,
    // so its rays spread over more waves instead of queueing on few lanes
    const uint32_t waves = gridDim.x * (kBptBlock / 64);
    const uint32_t chunk = min(128u, max(16u, ((n3 + ncn) / waves + 15u) & ~15u));
    TraceCounters tc;
    traceLoop<16, true, false, I, SLRH_FETCH_PARTS>(
        sc, st, n3 + ncn, b.fetch,
        [&](uint32_t i) {
            return i < n1 ? ql[i] : i < n2 ? qe[i - n1] : i < n3 ? (qs[i - n2] | kShadowBit | kSplatBit)
                                                                 : (qc[i - n3] | kShadowBit);
        },
        [&](uint32_t e, RayIn& r) {
            if (e & kShadowBit) {
                const uint32_t idx = e & ~(kShadowBit | kSplatBit);
                const float4* R = (e & kSplatBit) ? splatRows(b, cur, idx) + 4 : b.conn + (size_t)idx * kConnRows + 4;
                r.o = R[0];
                r.d = R[1];
                return true;
            }
            r.o = b.gen[(size_t)e * 4];
            r.d = b.gen[(size_t)e * 4 + 1];
            return false;
        },
        [&](uint32_t e, bool isAny, const HitOut& h) {
            if (isAny) {
                const uint32_t idx = e & ~(kShadowBit | kSplatBit);
                const uint8_t v = h.prim >= 0 ? 0 : 1;
                if (e & kSplatBit) b.visSplat[(size_t)cur * b.n + idx] = v;
                else b.vis[idx] = v;
            } else {
                b.gen[(size_t)e * 4 + 2] = make_float4(h.t, __int_as_float(h.prim), h.b0, h.b1);
                if (I) b.aux[e].w = (uint32_t)h.inst;
            }
        },
        tc, chunk);
    statAdd(b.stat, T_EXT_RAYS, tc.rays - tc.raysAny);
    statAdd(b.stat, T_SH_RAYS, tc.raysAny);
    statAdd(b.stat, T_NODES_EXT, tc.nodes - tc.nodesAny);
    statAdd(b.stat, T_TRIS_EXT, tc.tris - tc.trisAny);
    statAdd(b.stat, T_NODES_SH, tc.nodesAny);
    statAdd(b.stat, T_TRIS_SH, tc.trisAny);
    statAdd(b.stat, T_NODES_LDS, tc.nodesLds);
    statAdd(b.stat, T_INSTS, tc.insts);
    if (tc.overflow) atomicOr(&b.ctr[K_OVERFLOW], 1u);
}

// A surface point where a subpath's ray hit: position, frames, the incoming
// direction in the shading frame and the vertex's P (BPTRenderer.cpp:302-313).
struct NewVertex {
    SurfPt sp;
    float3 dirOut, gL;
    float cosOut, dist2;
    int32_t inst;
};
__device__ __forceinline__ NewVertex hitVertex(const DevScene& sc, const BptState& b, uint32_t s, const float4* g) {
    const float4 o4 = g[0], d4 = g[1], h = g[2];
    const float3 ro = xyz(o4), rd = xyz(d4);
    NewVertex nv;
    nv.inst = sc.numInstances ? (int32_t)b.aux[s].w : -1;
    nv.sp = surfaceAt(sc, (uint32_t)__float_as_int(h.y), h.z, h.w, ro, rd, h.x, nv.inst);
    const float3 dv = nv.sp.p - ro;
    nv.dist2 = dot3(dv, dv);
    nv.dirOut = toLocal(nv.sp, -rd);
    nv.gL = toLocal(nv.sp, nv.sp.gN);
    nv.cosOut = fabsf(dot3(-rd, nv.sp.gN));
    return nv;
}

// The BSDF sample at a new vertex, Russian roulette and the carried state
// for the next one (BPTRenderer.cpp:344-373). Returns false when the subpath
// ends here. On success the next ray, throughput and carried PDFs are
// stored; T[k-1] (k the new vertex) is formed from the reverse PDF toward
// the vertex below.
template <bool kDiel>
__device__ __forceinline__ bool extendSubpath(const DevScene& sc, const BptState& b, uint32_t s, const NewVertex& nv,
                                              uint32_t mat, const DevMaterial& m, float P_k, bool deltaK, int hero,
                                              float off, bool adjoint, bool tailDefined, float* alpha, uint4& rng,
                                              const float* spA = nullptr, const float* spB = nullptr) {
    if (m.type == MAT_NONE) return false;
    float A[kNumL], B[kNumL];
    if (spA) {
#pragma unroll
        for (int k = 0; k < kNumL; ++k) {
            A[k] = spA[k];
            B[k] = spB[k];
        }
    } else {
        materialSpectra(sc, mat, m.type, off, A, B);
    }
    const BiBsdf<kDiel> bs{m.type, A, B, m.p0, m.p1};
    BiQuery q;
    q.dir = nv.dirOut;
    q.ng = nv.gL;
    q.hero = hero;
    q.adjoint = adjoint;
    float fs[kNumL];
    BiSample smp;
    const float uc = xsFloat(rng), u0 = xsFloat(rng), u1 = xsFloat(rng);
    if (!bs.sample(q, uc, u0, u1, fs, &smp)) return false;
    bool any = false;
#pragma unroll
    for (int k = 0; k < kNumL; ++k) any = any || fs[k] != 0.0f;
    if (!any || smp.pdf == 0.0f) return false;
    const float3 vecIn = fromLocal(nv.sp, smp.dir);
    const float cosIn = fabsf(dot3(vecIn, nv.sp.gN));
    float weight[kNumL];
    const float wsc = cosIn / smp.pdf;
#pragma unroll
    for (int k = 0; k < kNumL; ++k) weight[k] = fs[k] * wsc;
    const float rr = rrMin(importance16(weight, hero));
    if (!(xsFloat(rng) < rr)) return false;
    const float rrr = 1.0f / rr;
#pragma unroll
    for (int k = 0; k < kNumL; ++k) alpha[k] *= weight[k] * rrr;
    float4* g = b.gen + (size_t)s * 4;
    const float4 carried = g[3];   // direction PDF, RR, cosine at the vertex below, its P
    const float4 mis = b.mis[s];
    // reverse PDFs at the vertex below (revAreaPDF, revRRProb)
    const float revArea = smp.revPdf * carried.z / nv.dist2;
    float t[kNumL];
    const float rsc = nv.cosOut / smp.revPdf;
#pragma unroll
    for (int k = 0; k < kNumL; ++k) t[k] = fs[k] * rsc;
    const float revRR = rrMin(importance16(t, hero));
    const uint32_t cf = __float_as_uint(mis.y);
    float T = 0.0f;
    if (tailDefined) {
        const float ratio = revArea * revRR / carried.w;
        const bool dPrev = (cf & M_DELTA_LAST) != 0;
        T = ratio * ratio * ((!dPrev && !deltaK ? 1.0f : 0.0f) + mis.x);
    }
    g[0] = make_float4(nv.sp.p.x, nv.sp.p.y, nv.sp.p.z, 1e-4f);
    g[1] = make_float4(vecIn.x, vecIn.y, vecIn.z, INFINITY);
    g[3] = make_float4(smp.pdf, rr, cosIn, P_k);
    b.mis[s] = make_float4(T, __uint_as_float(deltaK ? M_DELTA_LAST : 0u), mis.z, 0.0f);
    store16(b.alpha + (size_t)s * 4, alpha);
    return true;
}

// The sample's camera ray once its light subpath has ended
// (PerspectiveCamera::sampleRay, BPTRenderer.cpp:151-169): eye vertex 0 on
// the lens, then the first eye ray.
__device__ __forceinline__ void startEye(const BptState& b, const DevCamera& cam, const LensFrame& lf, uint32_t s,
                                         uint4& aux) {
    const float4 sm = b.smp[s];
    const float4 lens = b.lens[s];
    const float3 orgL = f3(sm.z, sm.w, 0.0f);
    const float ux = sm.x / cam.width, uy = sm.y / cam.height;
    const float3 pFocus = f3(cam.opWidth * (0.5f - ux), cam.opHeight * (0.5f - uy), cam.objDist);
    const float3 dl = normalize3(pFocus - orgL);
    const float dirPDF = cam.imgDist * cam.imgDist / ((dl.z * dl.z * dl.z) * cam.imgPlaneArea);
    const float3 dir = dl.x * lf.x + dl.y * lf.y + dl.z * lf.z;
    const float cosT = fabsf(dot3(dir, lf.z));
    const float a0 = 1.0f / (lens.w * (kNumL / (kLambdaHighD - kLambdaLowD)));
    const float a1 = a0 * (cosT / dirPDF);
    float al[kNumL];
#pragma unroll
    for (int k = 0; k < kNumL; ++k) al[k] = a1;
    store16(b.alpha + (size_t)s * 4, al);
    float4* g = b.gen + (size_t)s * 4;
    g[0] = make_float4(lens.x, lens.y, lens.z, 0.0f);
    g[1] = make_float4(dir.x, dir.y, dir.z, INFINITY);
    g[3] = make_float4(dirPDF, 1.0f, dl.z, lens.w);
    const bool pinhole = !(cam.lensRadius > 0.0f);
    const float4 mis = b.mis[s];
    b.mis[s] = make_float4(0.0f, __uint_as_float(pinhole ? M_DELTA_LAST : 0u), mis.z, 0.0f);
    // eye vertex 0 as the predecessor of the first hit
    float4* ev = b.ev + (size_t)s * kEyeRows;
    ev[kVtxRows] = make_float4(lens.x, lens.y, lens.z, lens.w);
    ev[kVtxRows + 1] = make_float4(lf.z.x, lf.z.y, lf.z.z, __uint_as_float(pinhole ? V_DELTA : 0u));
    aux.z &= 0xFFu;   // eye vertex index 0
}

// Light step: t = 1 contributions of the traced light-to-lens rays into the
// light image, then the light rays' hits; a light subpath that ends starts
// the sample's eye subpath.
template <bool kDiel>
__global__ __launch_bounds__(kBptBlock) void k_bpt_light(DevScene sc, BptState b, DevCamera cam, uint32_t cur) {
    const uint32_t nxt = cur ^ 1u;
    const uint32_t ns = b.ctr[K_SP + cur], ne = b.ctr[K_LE + cur];
    const LensFrame lf = lensFrame(cam);
    unsigned long long splats = 0, verts = 0, capped = 0;
    for (uint32_t i = blockIdx.x * kBptBlock + threadIdx.x; i < ns; i += gridDim.x * kBptBlock) {
        const uint32_t s = qSplat(b, cur)[i];
        uint8_t& vs = b.visSplat[(size_t)cur * b.n + s];
        if (!vs) continue;
        vs = 0;
        const float4* sp = splatRows(b, cur, s);
        float c[kNumL], add[kNumL];
        load16(sp, c);
        const uint32_t pix = __float_as_uint(sp[6].x);
        binStrata(b.mis[s].z, c, add);
        double* f = b.lightImg + (size_t)pix * kNumL;
#pragma unroll
        for (int k = 0; k < kNumL; ++k)
            if (add[k] != 0.0f) unsafeAtomicAdd(f + k, (double)add[k]);
        ++splats;
    }
    const uint32_t iters = (ne + gridDim.x * kBptBlock - 1) / (gridDim.x * kBptBlock);
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t i = it * gridDim.x * kBptBlock + blockIdx.x * kBptBlock + threadIdx.x;
        bool cont = false, sh = false, eye = false;
        uint32_t s = 0;
        if (i < ne) {
            s = qLight(b, cur)[i];
            const float4* g = b.gen + (size_t)s * 4;
            uint4 aux = b.aux[s];
            const uint32_t k = aux.z & 0xFFu;
            const float4 h = g[2];
            if (__float_as_int(h.y) >= 0) {
                if (k >= kLightCap) {
                    ++capped;
                } else {
                    // every read of the slot's state is issued before the
                    // vertex and splat stores (which the compiler cannot
                    // move them past)
                    float4* lvS = b.lv + (size_t)s * kLightCap * kVtxRows;
                    const float4* pr = lvS + (size_t)(k - 1) * kVtxRows;
                    const float4 pr0 = pr[0], pr1 = pr[1], pr3 = pr[3];
                    const float4 sm = b.smp[s], lensS = b.lens[s];
                    uint4 rng = b.rng[s];
                    const NewVertex nv = hitVertex(sc, b, s, g);
                    const float4 carried = g[3];
                    const float4 mis = b.mis[s];
                    const uint32_t cf = __float_as_uint(mis.y);
                    const int hero = (int)((aux.y >> 8) & 0xFFu);
                    const float off = mis.z;
                    const DevMaterial m = sc.mats[nv.sp.mat];
                    float alpha[kNumL];
                    load16(b.alpha + (size_t)s * 4, alpha);
                    // the material's spectra, shared by the lens connection and
                    // the next direction
                    float spA[kNumL], spB[kNumL];
                    if (m.type != MAT_NONE) materialSpectra(sc, nv.sp.mat, m.type, off, spA, spB);
                    Vtx V;
                    V.p = nv.sp.p;
                    V.P = carried.x * nv.cosOut / nv.dist2 * carried.y;
                    V.gN = nv.sp.gN;
                    V.cosIn = nv.cosOut;
                    V.sz = nv.sp.sz;
                    V.tail = mis.x;
                    V.sx = nv.sp.sx;
                    const bool deltaK = (cf & M_DELTA_NEXT) != 0;
                    V.flags = deltaK ? V_DELTA : 0u;
                    V.dir = nv.dirOut;
                    V.mat = nv.sp.mat;
                    storeVtx(lvS + (size_t)k * kVtxRows, V, alpha);
                    ++verts;
                    aux.z = k + 1;
                    b.aux[s] = aux;
                    // light vertex k to the lens
                    float c[kNumL];
                    uint32_t pix = 0;
                    float4 ro, rd;
                    sh = connectLens<kDiel>(sc, cam, lf, lensS, make_float2(sm.z, sm.w), V, alpha, true, pr0, pr1,
                                            __float_as_uint(pr3.w), hero, off, c, &pix, &ro, &rd, spA, spB);
                    if (sh) {
                        float4* spl = splatRows(b, nxt, s);
                        store16(spl, c);
                        spl[4] = ro;
                        spl[5] = rd;
                        spl[6] = make_float4(__uint_as_float(pix), 0, 0, 0);
                    }
                    cont = extendSubpath<kDiel>(sc, b, s, nv, nv.sp.mat, m, V.P, deltaK, hero, off, true, true, alpha, rng,
                                                spA, spB);
                    b.rng[s] = rng;
                }
            }
            if (!cont) {
                startEye(b, cam, lf, s, aux);
                b.aux[s] = aux;
                eye = true;
            }
        }
        const uint32_t pe = waveAppend(&b.ctr[K_LE + nxt], cont);
        if (cont) qLight(b, nxt)[pe] = s;
        const uint32_t ps = waveAppend(&b.ctr[K_SP + nxt], sh);
        if (sh) qSplat(b, nxt)[ps] = s;
        const uint32_t pq = waveAppend(&b.ctr[K_EE + nxt], eye);
        if (eye) qEye(b, nxt)[pq] = s;
    }
    statAdd(b.stat, T_SPLAT, splats);
    statAdd(b.stat, T_LIGHT_VERTS, verts);
    statAdd(b.stat, T_LIGHT_CAP, capped);
}

// Gather: unoccluded connections of the last eye vertex join the sample's
// radiance; a sample whose eye subpath ended is filed into the slot's plane
// and the slot's next sample starts.
template <bool kDiel>
__global__ __launch_bounds__(kBptBlock) void k_bpt_gather(DevScene sc, BptState b, DevCamera cam, uint32_t cur) {
    const uint32_t nxt = cur ^ 1u;
    const uint32_t nc = b.ctr[K_GA + cur];
    const LensFrame lf = lensFrame(cam);
    unsigned long long samples = 0;
    const uint32_t iters = (nc + gridDim.x * kBptBlock - 1) / (gridDim.x * kBptBlock);
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t i = it * gridDim.x * kBptBlock + blockIdx.x * kBptBlock + threadIdx.x;
        bool next = false, sh = false;
        uint32_t s = 0;
        if (i < nc) {
            s = qGather(b, cur)[i];
            uint4 aux = b.aux[s];
            const uint32_t nL = aux.z & 0xFFu;
            float rad[kNumL];
            load16(b.rad + (size_t)s * 4, rad);
            bool changed = false;
            for (uint32_t j = 0; j < nL; ++j) {
                const size_t idx = (size_t)s * kLightCap + j;
                if (!b.vis[idx]) continue;
                b.vis[idx] = 0;
                float c[kNumL];
                load16(b.conn + idx * kConnRows, c);
#pragma unroll
                for (int k = 0; k < kNumL; ++k) rad[k] += c[k];
                changed = true;
            }
            if (aux.y & B_END) {
                bool finite = true, any = false;
#pragma unroll
                for (int k = 0; k < kNumL; ++k) {
                    finite = finite && isfinite(rad[k]);
                    any = any || rad[k] != 0.0f;
                }
                if (!finite) {
                    atomicAdd(&b.ctr[K_NONFINITE], 1u);
                } else if (any) {
                    float add[kNumL];
                    binStrata(b.mis[s].z, rad, add);
                    const uint32_t plane = s / b.pixels;
                    double* f = b.planeSums + ((size_t)plane * b.pixels + aux.x) * kNumL;
#pragma unroll
                    for (int k = 0; k < kNumL; k += 2) {
                        double2 v = *(double2*)(f + k);
                        v.x += (double)add[k];
                        v.y += (double)add[k + 1];
                        *(double2*)(f + k) = v;
                    }
                }
                ++samples;
                next = startSample<kDiel>(sc, b, cam, lf, s, nxt, aux, &sh);
            } else if (changed) {
                store16(b.rad + (size_t)s * 4, rad);
            }
        }
        const uint32_t pl = waveAppend(&b.ctr[K_LE + nxt], next);
        if (next) qLight(b, nxt)[pl] = s;
        const uint32_t ps = waveAppend(&b.ctr[K_SP + nxt], sh);
        if (sh) qSplat(b, nxt)[ps] = s;
    }
    statAdd(b.stat, T_SAMPLES, samples);
}

// Eye step: each eye-ray hit becomes eye vertex k; its emission (s = 0,
// BPTRenderer.cpp:316-336), its connections to the stored light vertices
// (s >= 1) and the next direction.
template <bool kDiel>
__global__ __launch_bounds__(kBptBlock) void k_bpt_eye(DevScene sc, BptState b, uint32_t cur) {
    const uint32_t nxt = cur ^ 1u;
    const uint32_t ne = b.ctr[K_EE + cur];
    unsigned long long verts = 0, capped = 0;
    const uint32_t iters = (ne + gridDim.x * kBptBlock - 1) / (gridDim.x * kBptBlock);
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t i = it * gridDim.x * kBptBlock + blockIdx.x * kBptBlock + threadIdx.x;
        bool cont = false, gather = false;
        uint32_t s = 0;
        if (i < ne) {
            s = qEye(b, cur)[i];
            gather = true;
            const float4* g = b.gen + (size_t)s * 4;
            uint4 aux = b.aux[s];
            const uint32_t k = ((aux.z >> 8) & 0xFFFFu) + 1;
            const float4 h = g[2];
            if (__float_as_int(h.y) < 0 || (b.maxEye && k > b.maxEye) || k >= kEyeCap) {
                if (__float_as_int(h.y) >= 0) ++capped;
                aux.y = (aux.y | B_END) & ~B_CONNECT;
                b.aux[s] = aux;
            } else {
                const NewVertex nv = hitVertex(sc, b, s, g);
                const float4 carried = g[3];
                const float4 mis = b.mis[s];
                const uint32_t cf = __float_as_uint(mis.y);
                const int hero = (int)((aux.y >> 8) & 0xFFu);
                const float off = mis.z;
                const DevMaterial m = sc.mats[nv.sp.mat];
                float alpha[kNumL];
                load16(b.alpha + (size_t)s * 4, alpha);
                Vtx E;
                E.p = nv.sp.p;
                E.P = carried.x * nv.cosOut / nv.dist2 * carried.y;
                E.gN = nv.sp.gN;
                E.cosIn = nv.cosOut;
                E.sz = nv.sp.sz;
                E.tail = mis.x;
                E.sx = nv.sp.sx;
                const bool deltaK = (cf & M_DELTA_NEXT) != 0;
                E.flags = deltaK ? V_DELTA : 0u;
                E.dir = nv.dirOut;
                E.mat = nv.sp.mat;
                ++verts;
                float4* ev = b.ev + (size_t)s * kEyeRows;
                if (k >= 2) {
                    // the vertex below is the record's current vertex (eye
                    // vertex 0, the lens, was written there by the camera ray)
                    const float4 r0 = ev[0], r1 = ev[1], r3 = ev[3];
                    ev[kVtxRows] = r0;
                    ev[kVtxRows + 1] = make_float4(r1.x, r1.y, r1.z, r3.w);
                }
                const float4 Ep0 = ev[kVtxRows], Ep1 = ev[kVtxRows + 1];
                const uint32_t eFlagsPrev = __float_as_uint(Ep1.w);
                const bool eBelow = k >= 2;
                float rad[kNumL];
                bool radChanged = false;
                // s = 0: the eye subpath reached an emitter
                if (m.emit >= 0 && nv.dirOut.z > 0.0f) {
                    float Le[kNumL];
                    evalMatSpectrum(sc, nv.sp.mat, 2, off, Le);
                    const uint32_t prim = nv.sp.prim;
                    const float lightProb = sc.lightPmf[nv.inst >= 0 ? 0 : sc.primToLight[prim]] * lightSelectScale(sc);
                    const float ext1 = lightProb * (1.0f / triArea(sc, prim, nv.inst));
                    const float ext2 = (float)(nv.dirOut.z / 3.14159265358979323846) * carried.z / nv.dist2;
                    const float sE = sideSum(ext1, E.P, E.flags, eBelow, ext2, Ep0.w, eFlagsPrev, E.tail);
                    const float w = 1.0f / (1.0f + sE);
                    if (!isnan(w) && !isinf(w)) {
                        load16(b.rad + (size_t)s * 4, rad);
                        const float le1 = (float)(1.0 / 3.14159265358979323846);
#pragma unroll
                        for (int j = 0; j < kNumL; ++j) rad[j] += w * alpha[j] * Le[j] * le1;
                        radChanged = true;
                    }
                }
                if (radChanged) store16(b.rad + (size_t)s * 4, rad);
                // s >= 1: the vertex is stored for k_bpt_connect
                storeVtx(ev, E, alpha);
                if (m.type != MAT_NONE) aux.y |= B_CONNECT;
                else aux.y &= ~B_CONNECT;
                // the next direction, or the end of the eye subpath
                uint4 rng = b.rng[s];
                cont = extendSubpath<kDiel>(sc, b, s, nv, nv.sp.mat, m, E.P, deltaK, hero, off, false, k >= 2, alpha, rng);
                b.rng[s] = rng;
                if (!cont) aux.y |= B_END;
                aux.z = (aux.z & 0xFFu) | (k << 8);
                b.aux[s] = aux;
            }
        }
        const uint32_t pe = waveAppend(&b.ctr[K_EE + nxt], cont);
        if (cont) qEye(b, nxt)[pe] = s;
        const uint32_t pc = waveAppend(&b.ctr[K_GA + nxt], gather);
        if (gather) qGather(b, nxt)[pc] = s;
    }
    statAdd(b.stat, T_EYE_VERTS, verts);
    statAdd(b.stat, T_EYE_CAP, capped);
}

// Connections of the new eye vertices (s >= 1, t >= 2) to their light
// subpaths, one (eye vertex, light vertex) pair per lane: each wave lays
// its entries' pairs end to end (a prefix sum of their light vertex
// counts) and deals them out 64 at a time, so a long light subpath does not
// hold the rest of the wave. An unoccluded pair's contribution joins the
// sample in k_bpt_gather.
template <bool kDiel>
__global__ __launch_bounds__(kBptBlock) void k_bpt_connect(DevScene sc, BptState b, uint32_t nxt) {
    __shared__ uint32_t incl[kBptBlock], slots[kBptBlock];
    const uint32_t nc = b.ctr[K_GA + nxt];
    const uint32_t lane = threadIdx.x & 63u, wbase = threadIdx.x & ~63u;
    // the next step's traversal starts its claim cursors at 0 (the last
    // kernel of a step resets them: no memset node in the step graph)
    if (blockIdx.x == 0 && threadIdx.x < SLRH_FETCH_PARTS) b.fetch[threadIdx.x * kFetchStride] = 0;
    // entries per wave: 64, or fewer when the queue is short (the last steps
    // of a frame), so the pairs of a few long light subpaths spread over the
    // chip instead of queueing in a few waves
    const uint32_t waves = gridDim.x * (kBptBlock / 64);
    const uint32_t per = min(64u, max(1u, (nc + waves - 1) / waves));
    const uint32_t wave = blockIdx.x * (kBptBlock / 64) + (threadIdx.x >> 6);
    unsigned long long conns = 0;
    for (uint32_t base = wave * per; base < nc; base += waves * per) {
        const uint32_t i = base + lane;
        uint32_t s = 0, np = 0;
        if (lane < per && i < nc) {
            s = qGather(b, nxt)[i];
            const uint4 aux = b.aux[s];
            np = (aux.y & B_CONNECT) ? (aux.z & 0xFFu) : 0u;
        }
        // inclusive prefix sum of the pair counts over the wave
        uint32_t sum = np;
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(sum, d);
            if (lane >= d) sum += o;
        }
        incl[threadIdx.x] = sum;
        slots[threadIdx.x] = s;
        __builtin_amdgcn_wave_barrier();
        const uint32_t total = __shfl(sum, 63);
        for (uint32_t p0 = 0; p0 < total; p0 += 64) {
            const uint32_t p = p0 + lane;
            bool ok = false;
            size_t idx = 0;
            if (p < total) {
                // the entry owning pair p: the first whose inclusive sum exceeds p
                uint32_t lo = 0, hi = 63;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (incl[wbase + mid] > p) hi = mid;
                    else lo = mid + 1;
                }
                const uint32_t slot = slots[wbase + lo];
                const uint32_t j = p - (lo ? incl[wbase + lo - 1] : 0u);
                const uint4 aux = b.aux[slot];
                const int hero = (int)((aux.y >> 8) & 0xFFu);
                const float off = b.mis[slot].z;
                const uint32_t k = (aux.z >> 8) & 0xFFFFu;
                const float4* ev = b.ev + (size_t)slot * kEyeRows;
                const Vtx E = loadVtx(ev);
                const float4 Ep0 = ev[kVtxRows], Ep1 = ev[kVtxRows + 1];
                const float4* lr = b.lv + ((size_t)slot * kLightCap + j) * kVtxRows;
                const Vtx L = loadVtx(lr);
                float c[kNumL];
                float4 Lp0 = make_float4(0, 0, 0, 1), Lp1 = make_float4(0, 0, 0, 0);
                uint32_t lpFlags = 0;
                if (j > 0) {
                    const float4* pr = lr - kVtxRows;
                    Lp0 = pr[0];
                    Lp1 = pr[1];
                    lpFlags = __float_as_uint(pr[3].w);
                }
                float4 ro, rd;
                idx = (size_t)slot * kLightCap + j;
                ok = connectVertices<kDiel>(sc, E, ev + 5, k >= 2, Ep0, Ep1, __float_as_uint(Ep1.w), L, lr + 5, j > 0, Lp0,
                                     Lp1, lpFlags, hero, off, c, &ro, &rd);
                b.vis[idx] = 0;
                if (ok) {
                    float4* cr = b.conn + idx * kConnRows;
                    store16(cr, c);
                    cr[4] = ro;
                    cr[5] = rd;
                    ++conns;
                }
            }
            const uint32_t pos = waveAppend(&b.ctr[K_CN + nxt], ok);
            if (ok) qConn(b, nxt)[pos] = (uint32_t)idx;
        }
        __builtin_amdgcn_wave_barrier();
    }
    statAdd(b.stat, T_CONNECT, conns);
}


// The film gets the planes' sums and the light image.
__global__ void k_bpt_reduce(const double* planes, uint32_t numPlanes, size_t perPlane, const double* light,
                             double* out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < perPlane; i += (size_t)gridDim.x * blockDim.x) {
        double v = out[i];
        for (uint32_t p = 0; p < numPlanes; ++p) v += planes[(size_t)p * perPlane + i];
        out[i] = v + light[i];
    }
}

}  // namespace

// The bidirectional wavefront's device buffers, kept with the scene
// (DeviceScene::bptPool) and reused by later renders that fit.
struct BptPool {
    uint32_t slots = 0, traceBlocks = 0;
    DevBuf<float4> gen, mis, alpha, rad, smp, lens, lv, ev, conn, splat;
    DevBuf<uint4> aux, rng;
    DevBuf<uint8_t> vis, visSplat;
    DevBuf<uint32_t> q, ctr, spill, fetch;
    DevBuf<unsigned long long> stat;
    DevBuf<double> planeSums, lightImg;
    hipStream_t stream = nullptr;   // the wavefront's stream (steps replay as a graph on it)
    ~BptPool() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

int renderBpt(DeviceScene* ds, const RenderConfig& cfg, const CameraParams& camera, DeviceFilm* film,
              RenderStats* stOut, std::string* err) {
    if (ds->view.hasEnv) {
        if (err) *err = "the bidirectional integrator takes surface lights only (the scene has an environment light)";
        return SLRH_ERR_STATE_CODE;
    }
    if (ds->view.numLights == 0) {
        if (err) *err = "the bidirectional integrator needs a light in the scene";
        return SLRH_ERR_STATE_CODE;
    }
    if (film->w != cfg.width || film->h != cfg.height) {
        if (err) *err = "film size differs from the render size";
        return SLRH_ERR_STATE_CODE;
    }
    const uint32_t W = cfg.width, H = cfg.height, spp = cfg.spp;
    const uint32_t pixels = W * H;
    if (pixels == 0 || spp == 0) return 0;
    uint32_t planes = cfg.slotsPerPixel ? std::min(cfg.slotsPerPixel, spp)
                                        : std::max(1u, std::min(spp, kTargetSlots / std::max(1u, pixels)));
    const size_t nSlots = (size_t)pixels * planes;
    if (nSlots > kMaxBptSlots) {
        if (err) *err = "too many slots for the bidirectional integrator (pixels x slots_per_pixel > 2^24)";
        return SLRH_ERR_CAPACITY_CODE;
    }
    if ((spp + planes - 1) / planes > kMaxSlotSamples) {
        if (err) *err = "too many samples per slot for the bidirectional integrator (spp / slots_per_pixel > 65535)";
        return SLRH_ERR_CAPACITY_CODE;
    }
    const uint32_t n = (uint32_t)nSlots;
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const uint32_t maxBlocks = (uint32_t)std::max(1, cus) * 8;
    auto blocksFor = [&](uint32_t work) { return std::max(1u, std::min(maxBlocks, (work + kBptBlock - 1) / kBptBlock)); };
    const uint32_t traceBlocks = maxBlocks;

    // the wavefront's buffers stay with the scene between renders and grow
    // when a render needs more slots or a larger image
    std::shared_ptr<BptPool>& pool = ds->bptPool;
    if (!pool || pool->slots < n || pool->traceBlocks != traceBlocks) {
        pool.reset();
        std::shared_ptr<BptPool> np = std::make_shared<BptPool>();
        int r = 0;
        if ((r = np->gen.alloc((size_t)n * 4, err)) || (r = np->mis.alloc(n, err)) ||
            (r = np->alpha.alloc((size_t)n * 4, err)) || (r = np->rad.alloc((size_t)n * 4, err)) ||
            (r = np->smp.alloc(n, err)) || (r = np->lens.alloc(n, err)) ||
            (r = np->lv.alloc((size_t)n * kLightCap * kVtxRows, err)) || (r = np->ev.alloc((size_t)n * kEyeRows, err)) ||
            (r = np->conn.alloc((size_t)n * kLightCap * kConnRows, err)) ||
            (r = np->splat.alloc((size_t)2 * n * kSplatRows, err)) || (r = np->aux.alloc(n, err)) ||
            (r = np->rng.alloc(n, err)) || (r = np->vis.alloc((size_t)n * kLightCap, err)) ||
            (r = np->visSplat.alloc((size_t)2 * n, err)) || (r = np->q.alloc(queueWords(n), err)) ||
            (r = np->ctr.alloc(kBptCounters, err)) || (r = np->stat.alloc((size_t)kBptStats * kStatSpread, err)) ||
            (r = np->spill.alloc((size_t)(kMaxStack - kBptLdsDepth) * traceBlocks * kBptBlock, err)) ||
            (r = np->fetch.alloc((size_t)SLRH_FETCH_PARTS * kFetchStride, err)))
            return r;
        np->slots = n;
        np->traceBlocks = traceBlocks;
        pool = np;
    }
    BptPool& P = *pool;
    if (P.planeSums.n < (size_t)planes * pixels * kNumL)
        if (int r = P.planeSums.alloc((size_t)planes * pixels * kNumL, err)) return r;
    if (P.lightImg.n < (size_t)pixels * kNumL)
        if (int r = P.lightImg.alloc((size_t)pixels * kNumL, err)) return r;
    auto& gen = P.gen; auto& mis = P.mis; auto& alpha = P.alpha; auto& rad = P.rad; auto& smp = P.smp;
    auto& lens = P.lens; auto& lv = P.lv; auto& ev = P.ev; auto& conn = P.conn; auto& splat = P.splat;
    auto& aux = P.aux; auto& rng = P.rng; auto& vis = P.vis; auto& visSplat = P.visSplat; auto& q = P.q;
    auto& ctr = P.ctr; auto& stat = P.stat; auto& planeSums = P.planeSums; auto& lightImg = P.lightImg;
    auto& spill = P.spill;
    if (!P.stream) HIP_TRY(hipStreamCreate(&P.stream));
    hipStream_t stream = P.stream;
    HIP_TRY(hipMemsetAsync(vis.p, 0, (size_t)n * kLightCap, stream));
    HIP_TRY(hipMemsetAsync(stat.p, 0, stat.n * sizeof(unsigned long long), stream));
    HIP_TRY(hipMemsetAsync(planeSums.p, 0, (size_t)planes * pixels * kNumL * sizeof(double), stream));
    HIP_TRY(hipMemsetAsync(lightImg.p, 0, (size_t)pixels * kNumL * sizeof(double), stream));
    HIP_TRY(hipMemsetAsync(ctr.p, 0, kBptCounters * sizeof(uint32_t), stream));

    BptState b;
    b.n = n; b.pixels = pixels; b.planes = planes; b.spp = spp; b.maxEye = cfg.maxDepth;
    b.gen = gen.p; b.aux = aux.p; b.mis = mis.p; b.rng = rng.p; b.alpha = alpha.p; b.rad = rad.p; b.smp = smp.p;
    b.lens = lens.p; b.lv = lv.p; b.ev = ev.p; b.conn = conn.p; b.splat = splat.p; b.vis = vis.p;
    b.visSplat = visSplat.p; b.q = q.p; b.ctr = ctr.p; b.stat = stat.p; b.planeSums = planeSums.p;
    b.lightImg = lightImg.p; b.spill = spill.p; b.spillStride = traceBlocks * kBptBlock; b.fetch = P.fetch.p;
    DevCamera cam;
    cameraDerived(camera, W, H, &cam);
    const DevScene& sc = ds->view;

    uint32_t* hostCtr = nullptr;
    HIP_TRY(hipHostMalloc((void**)&hostCtr, kBptCounters * sizeof(uint32_t), hipHostMallocDefault));
    struct HostFree { uint32_t* p; ~HostFree() { if (p) (void)hipHostFree(p); } } hostGuard{hostCtr};
    hipEvent_t ev0, ev1;
    HIP_TRY(hipEventCreate(&ev0));
    HIP_TRY(hipEventCreate(&ev1));
    std::vector<hipEvent_t> traceEvents;   // (start, end) per timed traversal
    struct EvFree {
        hipEvent_t a, b;
        std::vector<hipEvent_t>* v;
        ~EvFree() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            for (hipEvent_t e : *v) (void)hipEventDestroy(e);
        }
    } evGuard{ev0, ev1, &traceEvents};
    double traceSeconds = 0, traceLaunches = 0, iterations = 0;
    auto readCounters = [&]() -> int {
        HIP_TRY(hipMemcpyAsync(hostCtr, ctr.p, kBptCounters * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return 0;
    };
    // kernels without the GGX dielectric for scenes that have none
    const bool diel = (ds->matTypeMask & (1u << MAT_GGX_DIELECTRIC)) != 0;
    const uint32_t grid = blocksFor(n);
    const uint32_t traceGrid = std::min(traceBlocks, blocksFor(4 * n));
    auto launchStep = [&](uint32_t cur) -> int {
        // timed traversals: an event pair each, read after the frame (no
        // wait between steps)
        hipEvent_t ta = nullptr, tb = nullptr;
        if (cfg.timeKernels) {
            HIP_TRY(hipEventCreate(&ta));
            traceEvents.push_back(ta);
            HIP_TRY(hipEventCreate(&tb));
            traceEvents.push_back(tb);
            HIP_TRY(hipEventRecord(ta, stream));
        }
        if (sc.numInstances)
            hipLaunchKernelGGL(k_bpt_trace<true>, dim3(traceGrid), dim3(kBptBlock), 0, stream, sc, b, cur);
        else
            hipLaunchKernelGGL(k_bpt_trace<false>, dim3(traceGrid), dim3(kBptBlock), 0, stream, sc, b, cur);
        HIP_TRY(hipGetLastError());
        if (cfg.timeKernels) HIP_TRY(hipEventRecord(tb, stream));
        traceLaunches += 1;
        if (diel) {
            hipLaunchKernelGGL(k_bpt_light<true>, dim3(grid), dim3(kBptBlock), 0, stream, sc, b, cam, cur);
            hipLaunchKernelGGL(k_bpt_gather<true>, dim3(grid), dim3(kBptBlock), 0, stream, sc, b, cam, cur);
            hipLaunchKernelGGL(k_bpt_eye<true>, dim3(grid), dim3(kBptBlock), 0, stream, sc, b, cur);
            hipLaunchKernelGGL(k_bpt_connect<true>, dim3(grid), dim3(kBptBlock), 0, stream, sc, b, cur ^ 1u);
        } else {
            hipLaunchKernelGGL(k_bpt_light<false>, dim3(grid), dim3(kBptBlock), 0, stream, sc, b, cam, cur);
            hipLaunchKernelGGL(k_bpt_gather<false>, dim3(grid), dim3(kBptBlock), 0, stream, sc, b, cam, cur);
            hipLaunchKernelGGL(k_bpt_eye<false>, dim3(grid), dim3(kBptBlock), 0, stream, sc, b, cur);
            hipLaunchKernelGGL(k_bpt_connect<false>, dim3(grid), dim3(kBptBlock), 0, stream, sc, b, cur ^ 1u);
        }
        HIP_TRY(hipGetLastError());
        iterations += 1;
        return 0;
    };

    HIP_TRY(hipEventRecord(ev0, stream));
    hipLaunchKernelGGL(k_bpt_seed, dim3(grid), dim3(kBptBlock), 0, stream, b, cfg.seed);
    HIP_TRY(hipGetLastError());
    if (diel) hipLaunchKernelGGL(k_bpt_start<true>, dim3(grid), dim3(kBptBlock), 0, stream, sc, b, cam);
    else hipLaunchKernelGGL(k_bpt_start<false>, dim3(grid), dim3(kBptBlock), 0, stream, sc, b, cam);
    HIP_TRY(hipGetLastError());
    // Steps until every queue is empty (every slot has filed its share of
    // samples). The kernels take their queue lengths from the device
    // counters, so the host reads them only every few steps; the steps past
    // the end find empty queues.
    // An even number of steps between reads, so every block of steps starts
    // at parity 0 and one recorded graph of the block serves every replay
    // (kernel timing launches the steps one by one instead).
    static_assert(kStepsPerRead % 2 == 0, "step blocks start at parity 0");
    hipGraphExec_t block = nullptr;
    struct GraphFree { hipGraphExec_t* g; ~GraphFree() { if (*g) (void)hipGraphExecDestroy(*g); } } graphGuard{&block};
    if (!cfg.timeKernels) {
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int e = 0;
        for (uint32_t k = 0; k < kStepsPerRead && !e; ++k) e = launchStep(k & 1u);
        const hipError_t ce = hipStreamEndCapture(stream, &graph);
        if (e) {
            if (graph) (void)hipGraphDestroy(graph);
            return e;
        }
        HIP_TRY(ce);
        const hipError_t ie = hipGraphInstantiate(&block, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIP_TRY(ie);
        // the capture counted its steps; the replays count their own
        iterations -= kStepsPerRead;
        traceLaunches -= kStepsPerRead;
    }
    for (;;) {
        if (int e = readCounters()) return e;
        const uint32_t pending = hostCtr[K_LE] + hostCtr[K_EE] + hostCtr[K_SP] + hostCtr[K_CN] + hostCtr[K_GA];
        if (pending == 0) break;
        if (block) {
            HIP_TRY(hipGraphLaunch(block, stream));
            iterations += kStepsPerRead;
            traceLaunches += kStepsPerRead;
        } else {
            for (uint32_t k = 0; k < kStepsPerRead; ++k)
                if (int e = launchStep(k & 1u)) return e;
        }
    }
    hipLaunchKernelGGL(k_bpt_reduce, dim3(blocksFor(pixels * kNumL)), dim3(kBptBlock), 0, stream, planeSums.p, planes,
                       (size_t)pixels * kNumL, lightImg.p, film->data.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev1, stream));
    HIP_TRY(hipEventSynchronize(ev1));
    float totalMs = 0;
    HIP_TRY(hipEventElapsedTime(&totalMs, ev0, ev1));
    for (size_t k = 0; k + 1 < traceEvents.size(); k += 2) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, traceEvents[k], traceEvents[k + 1]));
        traceSeconds += ms * 1e-3;
    }
    if (int e = readCounters()) return e;
    std::vector<unsigned long long> hs(stat.n);
    HIP_TRY(hipMemcpy(hs.data(), stat.p, stat.n * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    auto S = [&](uint32_t id) {
        double v = 0;
        for (uint32_t k = 0; k < kStatSpread; ++k) v += (double)hs[(size_t)id * kStatSpread + k];
        return v;
    };
    if (hostCtr[K_OVERFLOW]) {
        if (err) *err = "traversal stack overflow (depth > 64)";
        return SLRH_ERR_CAPACITY_CODE;
    }
    if (stOut) {
        RenderStats& st = *stOut;
        st.extendRays += S(T_EXT_RAYS);
        st.shadowRays += S(T_SH_RAYS);
        st.raysTraced += S(T_EXT_RAYS) + S(T_SH_RAYS);
        st.nodesVisited += S(T_NODES_EXT);
        st.trisTested += S(T_TRIS_EXT);
        st.nodesShadow += S(T_NODES_SH);
        st.trisShadow += S(T_TRIS_SH);
        st.nodesLds += S(T_NODES_LDS);
        st.instEntered += S(T_INSTS);
        st.samples += S(T_SAMPLES);
        st.nonFinite += hostCtr[K_NONFINITE];
        st.kernelSeconds += totalMs * 1e-3;
        st.traceSeconds += traceSeconds;
        st.traceLaunches += traceLaunches;
        st.iterations += iterations;
        st.slotsPerPixel = planes;
        st.bptLightVertices += S(T_LIGHT_VERTS);
        st.bptEyeVertices += S(T_EYE_VERTS);
        st.bptConnections += S(T_CONNECT);
        st.bptSplats += S(T_SPLAT);
        st.bptCapped += S(T_LIGHT_CAP) + S(T_EYE_CAP);
    }
    return 0;
}

}  // namespace slrh
