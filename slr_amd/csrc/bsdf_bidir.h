// BSDF queries of the bidirectional path tracer: the models of bsdf.h with
// what BPTRenderer asks of them beyond the unidirectional path tracer
// (libSLR/Renderer/BPTRenderer.cpp:171-186, :337-373):
//   - adjoint queries for light subpaths: the shading-normal correction is
//     taken on the query direction, |q.z / dot(q, ng)|, instead of the
//     sampled or evaluated one, and GGX transmission is scaled by eta_exit^2
//     instead of eta_enter^2 (BSDF::sample/evaluate,
//     libSLR/Core/directional_distribution_functions.h:295-332;
//     MicrofacetBSDF, libSLR/BSDF/microfacet_bsdfs.cpp:200, :252);
//   - the reverse direction's PDF with every sample and PDF evaluation
//     (requestReverse): Lambert and Oren-Nayar |q.z| / pi
//     (basic_bsdfs.cpp:21-24, OrenNayerBRDF.cpp:27-30), GGX reflection the
//     half-vector PDF seen from the other side (microfacet_bsdfs.cpp:36-39,
//     :95-96), GGX transmission microfacet_bsdfs.cpp:204-207, :313-314.
// The reverse value of every model here equals the forward value (each sets
// rev_fs = fs before the shared correction), so only the forward value is
// returned.
#pragma once

#include "bsdf.h"

namespace slrh {

struct BiQuery {
    float3 dir;    // the vertex's other direction in its shading frame (dirLocal)
    float3 ng;     // geometric normal in the shading frame
    int hero;
    bool adjoint;  // light subpath vertex
};

struct BiSample {
    float3 dir;
    float pdf, revPdf;
    bool transmission;
};

__device__ __forceinline__ float biCorrection(const BiQuery& q, const float3& d) {
    return q.adjoint ? fabsf(q.dir.z / dot3(q.dir, q.ng)) : fabsf(d.z / dot3(d, q.ng));
}

// A vertex's scattering model: type (MAT_*), the 16-wavelength spectra A, B
// (reflectance | eta, k | eta_ext, eta_int) and the model parameters.
// kDiel: the GGX dielectric is compiled in (kernels for scenes without one
// leave it out and hold far fewer registers).
template <bool kDiel>
struct BiBsdf {
    uint32_t type;
    const float* A;
    const float* B;
    float p0, p1;   // Oren-Nayar A, B | GGX alpha_x, alpha_y

    __device__ GGX ggx() const { return GGX{p0, p1}; }

    __device__ float onFactor(const float3& wi, const float3& wo, bool sampled) const {
        Diffuse d{A, type == MAT_OREN_NAYAR, p0, p1};
        return d.onFactor(wi, wo, sampled);
    }
    __device__ void diffuseValue(float f, float* fs) const {
#pragma unroll
        for (int i = 0; i < kNumL; ++i)
            fs[i] = type == MAT_OREN_NAYAR ? A[i] * (float)(f / 3.14159265358979323846) : A[i] * kInvPiF;
    }
    // GGX transmission value per wavelength before the eta^2 scaling
    // (microfacet_bsdfs.cpp:186-198)
    __device__ void transmission(const float3& v, const float3& l, bool entering, float* fs) const {
        const GGX D = ggx();
#pragma unroll
        for (int w = 0; w < kNumL; ++w) {
            const float eEn = entering ? A[w] : B[w];
            const float eEx = entering ? B[w] : A[w];
            const float3 mw = normalize3(-(eEn * v + eEx * l));
            const float dv = dot3(v, mw), dl = dot3(l, mw);
            const float Fw = fresnelDielectric(A[w], B[w], dv);
            const float Gw = D.G1(v, mw) * D.G1(l, mw);
            const float Dw = D.D(mw);
            const float den = eEn * dv + eEx * dl;
            fs[w] = fabsf(dv * dl) * (1 - Fw) * Gw * Dw / (den * den);
        }
    }
    __device__ void transmissionScale(const BiQuery& q, bool entering, float nvnl, float* fs) const {
#pragma unroll
        for (int w = 0; w < kNumL; ++w) {
            const float eEn = entering ? A[w] : B[w];
            const float eEx = entering ? B[w] : A[w];
            const float e = q.adjoint ? eEx : eEn;
            fs[w] = fs[w] / fabsf(nvnl) * (e * e);
        }
    }

    // BSDF::evaluate with the correction (no reverse value: it equals fs)
    __device__ void evaluate(const BiQuery& q, const float3& d, float* fs) const {
#pragma unroll
        for (int i = 0; i < kNumL; ++i) fs[i] = 0.0f;
        const float sn = biCorrection(q, d);
        if (type == MAT_LAMBERT || type == MAT_OREN_NAYAR) {
            if (q.dir.z * d.z <= 0.0f) return;
            diffuseValue(onFactor(d, q.dir, false), fs);
        } else if (type == MAT_GGX_CONDUCTOR) {
            if (d.z * q.dir.z <= 0) return;
            const bool entering = q.dir.z >= 0.0f;
            const float sign = entering ? 1.0f : -1.0f;
            const GGX D = ggx();
            const float3 m = sign * halfVector(q.dir, d);
            const float dotHV = dot3(q.dir, m);
            const float dm = D.D(m);
            const float G = D.G1(q.dir, m) * D.G1(d, m);
            const float denom = 4 * q.dir.z * d.z;
#pragma unroll
            for (int i = 0; i < kNumL; ++i) fs[i] = fresnelConductor(A[i], B[i], dotHV) * dm * G / denom;
        } else if ((kDiel && type == MAT_GGX_DIELECTRIC)) {
            const bool entering = q.dir.z >= 0.0f;
            const float sign = entering ? 1.0f : -1.0f;
            const float nvnl = d.z * q.dir.z;
            const GGX D = ggx();
            if (nvnl > 0) {
                const float3 m = sign * halfVector(q.dir, d);
                const float dotHV = dot3(q.dir, m);
                const float dm = D.D(m);
                const float G = D.G1(q.dir, m) * D.G1(d, m);
#pragma unroll
                for (int i = 0; i < kNumL; ++i) fs[i] = fresnelDielectric(A[i], B[i], dotHV) * dm * G / (4 * nvnl);
            } else if (nvnl < 0) {
                transmission(q.dir, d, entering, fs);
                transmissionScale(q, entering, nvnl, fs);
            }
        } else {
            return;
        }
#pragma unroll
        for (int i = 0; i < kNumL; ++i) fs[i] *= sn;
    }

    // BSDF::evaluatePDF with the reverse PDF
    __device__ float pdf(const BiQuery& q, const float3& d, float* rev) const {
        *rev = 0.0f;
        if (type == MAT_LAMBERT || type == MAT_OREN_NAYAR) {
            if (q.dir.z * d.z <= 0.0f) return 0.0f;
            *rev = (float)(fabsf(q.dir.z) / 3.14159265358979323846);
            return (float)(fabsf(d.z) / 3.14159265358979323846);
        }
        const bool entering = q.dir.z >= 0.0f;
        const float sign = entering ? 1.0f : -1.0f;
        const GGX D = ggx();
        if (type == MAT_GGX_CONDUCTOR) {
            if (d.z * q.dir.z <= 0) return 0.0f;
            const float3 m = sign * halfVector(q.dir, d);
            const float dotHV = dot3(q.dir, m);
            if (dotHV * sign <= 0) return 0.0f;
            const float common = 1.0f / (4 * dotHV * sign);
            *rev = common * D.pdfVisible(sign * d, m);
            return common * D.pdfVisible(sign * q.dir, m);
        }
        if (!kDiel || type != MAT_GGX_DIELECTRIC) return 0.0f;
        const float nvnl = d.z * q.dir.z;
        if (nvnl == 0) return 0.0f;
        const float extH = pick16(A, q.hero), intH = pick16(B, q.hero);
        const float eEnterH = entering ? extH : intH, eExitH = entering ? intH : extH;
        const float3 m = nvnl > 0 ? sign * halfVector(q.dir, d) : normalize3(-(eEnterH * q.dir + eExitH * d));
        const float dotHV = dot3(q.dir, m);
        if (dotHV * sign <= 0) return 0.0f;
        const float mPdf = D.pdfVisible(sign * q.dir, m);
        float F[kNumL];
#pragma unroll
        for (int i = 0; i < kNumL; ++i) F[i] = fresnelDielectric(A[i], B[i], dotHV);
        const float reflectProb = importance16(F, q.hero);
        if (nvnl > 0) {
            const float common = reflectProb / (4 * dotHV * sign);
            *rev = common * D.pdfVisible(sign * d, m);
            return common * mPdf;
        }
        const float dotHL = dot3(d, m);
        const float den = eEnterH * dotHV + eExitH * dotHL;
        const float common = (1 - reflectProb) / (den * den);
        *rev = common * D.pdfVisible(-sign * d, m) * eEnterH * eEnterH * fabsf(dotHV);
        return common * mPdf * eExitH * eExitH * fabsf(dotHL);
    }

    // BSDF::sample with the reverse PDF; false (no sample) where the
    // reference returns a zero PDF
    __device__ bool sample(const BiQuery& q, float uc, float u0, float u1, float* fs, BiSample* s) const {
#pragma unroll
        for (int i = 0; i < kNumL; ++i) fs[i] = 0.0f;
        s->pdf = 0.0f;
        s->revPdf = 0.0f;
        s->transmission = false;
        if (type == MAT_LAMBERT || type == MAT_OREN_NAYAR) {
            float3 d = cosineHemisphere(u0, u1);
            s->pdf = (float)(d.z / 3.14159265358979323846);
            d.z *= q.dir.z > 0 ? 1 : -1;
            s->dir = d;
            s->revPdf = (float)(fabsf(q.dir.z) / 3.14159265358979323846);
            diffuseValue(onFactor(d, q.dir, true), fs);
        } else if (type == MAT_GGX_CONDUCTOR) {
            const bool entering = q.dir.z >= 0.0f;
            const float sign = entering ? 1.0f : -1.0f;
            const GGX D = ggx();
            float3 m;
            float mPdf;
            const float dm = D.sampleVisible(sign * q.dir, u0, u1, &m, &mPdf);
            const float dotHV = dot3(q.dir, m);
            if (dotHV * sign <= 0) return false;
            const float3 wi = 2 * dotHV * m - q.dir;
            if (wi.z * q.dir.z <= 0) return false;
            const float common = 1.0f / (4 * dotHV * sign);
            s->pdf = common * mPdf;
            s->dir = wi;
            const float G = D.G1(q.dir, m) * D.G1(wi, m);
            const float denom = 4 * q.dir.z * wi.z;
#pragma unroll
            for (int i = 0; i < kNumL; ++i) fs[i] = fresnelConductor(A[i], B[i], dotHV) * dm * G / denom;
            s->revPdf = common * D.pdfVisible(sign * wi, m);
        } else if ((kDiel && type == MAT_GGX_DIELECTRIC)) {
            const bool entering = q.dir.z >= 0.0f;
            const float sign = entering ? 1.0f : -1.0f;
            const GGX D = ggx();
            const float extH = pick16(A, q.hero), intH = pick16(B, q.hero);
            const float eEnterH = entering ? extH : intH, eExitH = entering ? intH : extH;
            float3 m;
            float mPdf;
            const float dm = D.sampleVisible(sign * q.dir, u0, u1, &m, &mPdf);
            const float dotHV = dot3(q.dir, m);
            if (dotHV * sign <= 0 || isnan(dm)) return false;
            float F[kNumL];
#pragma unroll
            for (int i = 0; i < kNumL; ++i) F[i] = fresnelDielectric(A[i], B[i], dotHV);
            const float reflectProb = importance16(F, q.hero);
            if (uc < reflectProb) {
                const float3 wi = 2 * dotHV * m - q.dir;
                if (wi.z * q.dir.z <= 0) return false;
                const float common = reflectProb / (4 * dotHV * sign);
                s->pdf = common * mPdf;
                s->dir = wi;
                const float G = D.G1(q.dir, m) * D.G1(wi, m);
                const float denom = 4 * q.dir.z * wi.z;
#pragma unroll
                for (int i = 0; i < kNumL; ++i) fs[i] = F[i] * dm * G / denom;
                s->revPdf = common * D.pdfVisible(sign * wi, m);
            } else {
                const float recRel = eEnterH / eExitH;
                const float inner = 1 + recRel * recRel * (dotHV * dotHV - 1);
                if (inner < 0) return false;
                const float3 wi = (recRel * dotHV - sign * sqrtf(inner)) * m - recRel * q.dir;
                if (wi.z * q.dir.z >= 0) return false;
                const float dotHL = dot3(wi, m);
                const float den = eEnterH * dotHV + eExitH * dotHL;
                const float common = (1 - reflectProb) / (den * den);
                s->pdf = common * mPdf * eExitH * eExitH * fabsf(dotHL);
                s->dir = wi;
                s->transmission = true;
                transmission(q.dir, wi, entering, fs);
                transmissionScale(q, entering, q.dir.z * wi.z, fs);
                s->revPdf = common * D.pdfVisible(-sign * wi, m) * eEnterH * eEnterH * fabsf(dotHV);
            }
        } else {
            return false;
        }
        const float sn = biCorrection(q, s->dir);
#pragma unroll
        for (int i = 0; i < kNumL; ++i) fs[i] *= sn;
        return true;
    }
};

}  // namespace slrh
