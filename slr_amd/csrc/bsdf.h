// BSDFs of the configurations, evaluated for all 16 wavelengths of a path.
//
// Each function restates the reference's model with the same sampling
// routines, so that sample/evaluate/PDF agree with it value for value up to
// libm rounding:
//   Lambert            libSLR/BSDF/basic_bsdfs.cpp:13-58
//   Oren-Nayar         libSLR/BSDF/OrenNayerBRDF.cpp:13-74 (A, B: OrenNayerBRDF.h:28-30)
//   GGX conductor      libSLR/BSDF/microfacet_bsdfs.cpp:11-113
//   GGX dielectric     libSLR/BSDF/microfacet_bsdfs.cpp:116-328
//   GGX D, G1, VNDF    libSLR/Core/directional_distribution_functions.cpp:247-315
//   Fresnel            directional_distribution_functions.cpp:118-211
//   sampling helpers   libSLR/Core/distributions.cpp:57-90, distributions.h:30-37
// and the shading-normal correction |d.z / dot(d, ng)| the BSDF front end
// applies to sample and evaluate (libSLR/Core/directional_distribution_functions.h:295-349).
#pragma once

#include "device_common.h"

namespace slrh {

constexpr float kPi = 3.14159265358979323846f;

struct Spec16 {
    float v[kNumL];
};

// v[i] for a lane-varying i, as an OR of masked values: indexing a register
// array with a runtime index would move the array to scratch memory. The
// empty asm keeps each mask opaque, so the optimiser cannot fold the masks
// back into an indexed load.
__device__ __forceinline__ float pick16(const float* v, int i) {
    uint32_t r = 0u;
#pragma unroll
    for (int k = 0; k < kNumL; ++k) {
        uint32_t m = i == k ? 0xFFFFFFFFu : 0u;
        asm volatile("" : "+v"(m));
        r |= __float_as_uint(v[k]) & m;
    }
    return __uint_as_float(r);
}

__device__ __forceinline__ float importance16(const float* v, int hero) {
    // SampledSpectrum::importance (libSLR/BasicTypes/spectrum_types.h:473-486)
    float sum = 0;
#pragma unroll
    for (int i = 0; i < kNumL; ++i) sum += v[i];
    const float primary = 0.9f;
    const float marginal = (1 - primary) / (kNumL - 1);
    return sum * marginal + pick16(v, hero) * (primary - marginal);
}

// concentricSampleDisk (distributions.cpp:57-90)
__device__ __forceinline__ void concentricDisk(float u0, float u1, float* dx, float* dy) {
    float r, theta;
    float sx = 2 * u0 - 1, sy = 2 * u1 - 1;
    if (sx == 0 && sy == 0) { *dx = 0; *dy = 0; return; }
    if (sx >= -sy) {
        if (sx > sy) { r = sx; theta = sy / sx; }
        else { r = sy; theta = 2 - sx / sy; }
    } else {
        if (sx > sy) { r = -sy; theta = 6 + sx / sy; }
        else { r = -sx; theta = 4 + sy / sx; }
    }
    theta = (float)(theta * 0.78539816339744830962);
    *dx = r * cosf(theta);
    *dy = r * sinf(theta);
}

__device__ __forceinline__ float3 cosineHemisphere(float u0, float u1) {
    float x, y;
    concentricDisk(u0, u1, &x, &y);
    return f3(x, y, sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y)));
}

struct BsdfQuery {
    float3 dir;   // outgoing direction in the shading frame (dirLocal)
    float3 ng;    // geometric normal in the shading frame
    int hero;
};

struct BsdfSample {
    float3 dir;
    float pdf;
    bool reflection;
};

// shading-normal correction for the non-adjoint query
__device__ __forceinline__ float snCorrection(const float3& d, const float3& ng) { return fabsf(d.z / dot3(d, ng)); }

// ---------------------------------------------------------------- GGX

struct GGX {
    float ax, ay;
    __device__ float D(const float3& m) const {
        if (m.z <= 0) return 0.0f;
        float temp = m.x * m.x / (ax * ax) + m.y * m.y / (ay * ay) + m.z * m.z;
        return 1.0f / ((float)(3.14159265358979323846 * ax) * ay * temp * temp);
    }
    __device__ float G1(const float3& v, const float3& m) const {
        float chi = (dot3(v, m) / v.z) > 0 ? 1.0f : 0.0f;
        float t2 = (v.x * v.x * ax * ax + v.y * v.y * ay * ay) / (v.z * v.z);
        return chi * 2 / (1 + sqrtf(1 + t2));
    }
    __device__ float pdfVisible(const float3& v, const float3& m) const {
        return G1(v, m) * fabsf(dot3(v, m)) * D(m) / fabsf(v.z);
    }
    // GGXMicrofacetDistribution::sample(v, u0, u1) (directional_distribution_functions.cpp:269-303)
    __device__ float sampleVisible(const float3& v, float u0, float u1, float3* m, float* pdf) const {
        float3 sv = normalize3(f3(ax * v.x, ay * v.y, v.z));
        float dist2D = sqrtf(sv.x * sv.x + sv.y * sv.y);
        float rec = 1.0f / dist2D;
        float3 T1 = (sv.z < 0.9999f) ? f3(sv.y * rec, -sv.x * rec, 0) : f3(1, 0, 0);
        float3 T2 = f3(T1.y * sv.z, -T1.x * sv.z, dist2D);
        float a = 1.0f / (1.0f + sv.z);
        float r = sqrtf(u0);
        float phi = kPi * ((u1 < a) ? u1 / a : 1 + (u1 - a) / (1.0f - a));
        float P1 = r * cosf(phi);
        float P2 = r * sinf(phi) * ((u1 < a) ? 1.0f : sv.z);
        float3 mm = P1 * T1 + P2 * T2 + sqrtf(1.0f - P1 * P1 - P2 * P2) * sv;
        mm = normalize3(f3(ax * mm.x, ay * mm.y, mm.z));
        *m = mm;
        float d = D(mm);
        *pdf = G1(v, mm) * fabsf(dot3(v, mm)) * d / fabsf(v.z);
        return d;
    }
};

// FresnelConductor::evaluate (directional_distribution_functions.cpp:118-128)
__device__ __forceinline__ float fresnelConductor(float eta, float k, float cosEnter) {
    cosEnter = fabsf(cosEnter);
    float c2 = cosEnter * cosEnter;
    float twoEtaCos = 2.0f * eta * cosEnter;
    float tmpf = eta * eta + k * k;
    float tmp = tmpf * c2;
    float rparl2 = (tmp - twoEtaCos + 1) / (tmp + twoEtaCos + 1);
    float rperp2 = (tmpf - twoEtaCos + c2) / (tmpf + twoEtaCos + c2);
    return (rparl2 + rperp2) / 2.0f;
}

// FresnelDielectric::evaluate for one wavelength (directional_distribution_functions.cpp:182-211)
__device__ __forceinline__ float fresnelDielectric(float etaExt, float etaInt, float cosEnter) {
    cosEnter = fminf(fmaxf(cosEnter, -1.0f), 1.0f);
    bool entering = cosEnter > 0.0f;
    float eEnter = entering ? etaExt : etaInt;
    float eExit = entering ? etaInt : etaExt;
    float sinExit = eEnter / eExit * sqrtf(fmaxf(0.0f, 1.0f - cosEnter * cosEnter));
    cosEnter = fabsf(cosEnter);
    if (sinExit >= 1.0f) return 1.0f;
    float cosExit = sqrtf(fmaxf(0.0f, 1.0f - sinExit * sinExit));
    float rparl = ((eExit * cosEnter) - (eEnter * cosExit)) / ((eExit * cosEnter) + (eEnter * cosExit));
    float rperp = ((eEnter * cosEnter) - (eExit * cosExit)) / ((eEnter * cosEnter) + (eExit * cosExit));
    return (rparl * rparl + rperp * rperp) / 2.0f;
}

__device__ __forceinline__ float3 halfVector(const float3& a, const float3& b) { return normalize3(a + b); }

// ---------------------------------------------------------------- Lambert / Oren-Nayar

// Lambert's m_R / M_PI: SampledSpectrum::operator/(float) multiplies by the
// float reciprocal of (float)M_PI (libSLR/BasicTypes/spectrum_types.h:351-357)
constexpr float kInvPiF = 1.0f / 3.14159265358979323846f;

struct Diffuse {
    const float* R;   // 16 reflectance values
    bool orenNayar;
    float A, B;

    __device__ float onFactor(const float3& wi, const float3& wo, bool sampled) const {
        if (!orenNayar) return 1.0f;
        float sinI = 1.0f - wi.z * wi.z;
        float sinO = 1.0f - wo.z * wo.z;
        float tanI = sinI / fabsf(wi.z);
        float tanO = sinO / fabsf(wo.z);
        float sinAlpha = fmaxf(sinI, sinO);
        float tanBeta = fminf(tanI, tanO);
        float cosDA = (wi.x * wo.x + wi.y * wo.y) / (sinI * sinO);
        if (sampled && !isfinite(cosDA)) cosDA = 0.0f;
        float mx = (0.0f < cosDA) ? cosDA : 0.0f;   // std::max(0.0f, x)
        return (A + B * mx * sinAlpha * tanBeta);
    }

    __device__ bool sample(const BsdfQuery& q, float /*uc*/, float u0, float u1, float* fs, BsdfSample* s) const {
        float3 d = cosineHemisphere(u0, u1);
        s->pdf = (float)(d.z / 3.14159265358979323846);
        d.z *= q.dir.z > 0 ? 1 : -1;
        s->dir = d;
        s->reflection = true;
        float f = onFactor(d, q.dir, true);
        float sn = snCorrection(d, q.ng);
#pragma unroll
        for (int i = 0; i < kNumL; ++i) fs[i] = (orenNayar ? R[i] * (float)(f / 3.14159265358979323846)
                                                           : R[i] * kInvPiF) * sn;
        return true;
    }
    __device__ void evaluate(const BsdfQuery& q, const float3& d, float* fs, float* pdf) const {
        if (q.dir.z * d.z <= 0.0f) {
#pragma unroll
            for (int i = 0; i < kNumL; ++i) fs[i] = 0.0f;
            *pdf = 0.0f;
            return;
        }
        float f = onFactor(d, q.dir, false);
        float sn = snCorrection(d, q.ng);
#pragma unroll
        for (int i = 0; i < kNumL; ++i) fs[i] = (orenNayar ? R[i] * (float)(f / 3.14159265358979323846)
                                                           : R[i] * kInvPiF) * sn;
        *pdf = (float)(fabsf(d.z) / 3.14159265358979323846);
    }
};

// ---------------------------------------------------------------- GGX conductor

struct Conductor {
    const float* eta;
    const float* k;
    GGX D;

    __device__ bool sample(const BsdfQuery& q, float /*uc*/, float u0, float u1, float* fs, BsdfSample* s) const {
        bool entering = q.dir.z >= 0.0f;
        float sign = entering ? 1.0f : -1.0f;
        float3 m;
        float mPdf;
        float d = D.sampleVisible(sign * q.dir, u0, u1, &m, &mPdf);
        float dotHV = dot3(q.dir, m);
        if (dotHV * sign <= 0) return false;
        float3 wi = 2 * dotHV * m - q.dir;
        if (wi.z * q.dir.z <= 0) return false;
        float common = 1.0f / (4 * dotHV * sign);
        s->pdf = common * mPdf;
        s->dir = wi;
        s->reflection = true;
        float G = D.G1(q.dir, m) * D.G1(wi, m);
        float sn = snCorrection(wi, q.ng);
        float denom = 4 * q.dir.z * wi.z;
#pragma unroll
        for (int i = 0; i < kNumL; ++i) fs[i] = fresnelConductor(eta[i], k[i], dotHV) * d * G / denom * sn;
        return true;
    }
    __device__ void evaluate(const BsdfQuery& q, const float3& wi, float* fs, float* pdf) const {
#pragma unroll
        for (int i = 0; i < kNumL; ++i) fs[i] = 0.0f;
        *pdf = 0.0f;
        if (wi.z * q.dir.z <= 0) return;
        bool entering = q.dir.z >= 0.0f;
        float sign = entering ? 1.0f : -1.0f;
        float3 m = sign * halfVector(q.dir, wi);
        float dotHV = dot3(q.dir, m);
        float d = D.D(m);
        float G = D.G1(q.dir, m) * D.G1(wi, m);
        float sn = snCorrection(wi, q.ng);
        float denom = 4 * q.dir.z * wi.z;
#pragma unroll
        for (int i = 0; i < kNumL; ++i) fs[i] = fresnelConductor(eta[i], k[i], dotHV) * d * G / denom * sn;
        if (dotHV * sign <= 0) return;
        float mPdf = D.pdfVisible(sign * q.dir, m);
        *pdf = 1.0f / (4 * dotHV * sign) * mPdf;
    }
};

// ---------------------------------------------------------------- GGX dielectric

struct Dielectric {
    const float* etaExt;
    const float* etaInt;
    GGX D;

    __device__ float reflectProbOf(float dotHV, int hero, float* F) const {
#pragma unroll
        for (int i = 0; i < kNumL; ++i) F[i] = fresnelDielectric(etaExt[i], etaInt[i], dotHV);
        return importance16(F, hero);
    }

    __device__ void transmissionValue(const BsdfQuery& q, const float3& wi, bool entering, float* fs) const {
#pragma unroll
        for (int w = 0; w < kNumL; ++w) {
            float eEn = entering ? etaExt[w] : etaInt[w];
            float eEx = entering ? etaInt[w] : etaExt[w];
            float3 mw = normalize3(-(eEn * q.dir + eEx * wi));
            float dv = dot3(q.dir, mw), dl = dot3(wi, mw);
            float Fw = fresnelDielectric(etaExt[w], etaInt[w], dv);
            float Gw = D.G1(q.dir, mw) * D.G1(wi, mw);
            float Dw = D.D(mw);
            float den = eEn * dv + eEx * dl;
            fs[w] = fabsf(dv * dl) * (1 - Fw) * Gw * Dw / (den * den);
        }
    }

    __device__ bool sample(const BsdfQuery& q, float uc, float u0, float u1, float* fs, BsdfSample* s) const {
        bool entering = q.dir.z >= 0.0f;
        float sign = entering ? 1.0f : -1.0f;
        const int h = q.hero;
        const float extH = pick16(etaExt, h), intH = pick16(etaInt, h);
        float eEnterH = entering ? extH : intH;
        float eExitH = entering ? intH : extH;
        float3 m;
        float mPdf;
        float d = D.sampleVisible(sign * q.dir, u0, u1, &m, &mPdf);
        float dotHV = dot3(q.dir, m);
        if (dotHV * sign <= 0 || isnan(d)) return false;
        float F[kNumL];
        float reflectProb = reflectProbOf(dotHV, h, F);
        if (uc < reflectProb) {
            float3 wi = 2 * dotHV * m - q.dir;
            if (wi.z * q.dir.z <= 0) return false;
            float common = reflectProb / (4 * dotHV * sign);
            s->pdf = common * mPdf;
            s->dir = wi;
            s->reflection = true;
            float G = D.G1(q.dir, m) * D.G1(wi, m);
            float sn = snCorrection(wi, q.ng);
            float denom = 4 * q.dir.z * wi.z;
#pragma unroll
            for (int i = 0; i < kNumL; ++i) fs[i] = F[i] * d * G / denom * sn;
            return true;
        }
        float recRel = eEnterH / eExitH;
        float inner = 1 + recRel * recRel * (dotHV * dotHV - 1);
        if (inner < 0) return false;
        float3 wi = (recRel * dotHV - sign * sqrtf(inner)) * m - recRel * q.dir;
        if (wi.z * q.dir.z >= 0) return false;
        float dotHL = dot3(wi, m);
        float den = eEnterH * dotHV + eExitH * dotHL;
        float common = (1 - reflectProb) / (den * den);
        s->pdf = common * mPdf * eExitH * eExitH * fabsf(dotHL);
        s->dir = wi;
        s->reflection = false;
        transmissionValue(q, wi, entering, fs);
        float sn = snCorrection(wi, q.ng);
        float inv = fabsf(q.dir.z * wi.z);
#pragma unroll
        for (int w = 0; w < kNumL; ++w) {
            float eEn = entering ? etaExt[w] : etaInt[w];
            fs[w] = fs[w] / inv * (eEn * eEn) * sn;
        }
        return true;
    }

    __device__ void evaluate(const BsdfQuery& q, const float3& wi, float* fs, float* pdf) const {
        bool entering = q.dir.z >= 0.0f;
        float sign = entering ? 1.0f : -1.0f;
        float nvnl = wi.z * q.dir.z;
        float sn = snCorrection(wi, q.ng);
        *pdf = 0.0f;
        if (nvnl > 0) {
            float3 m = sign * halfVector(q.dir, wi);
            float dotHV = dot3(q.dir, m);
            float d = D.D(m);
            float G = D.G1(q.dir, m) * D.G1(wi, m);
            float F[kNumL];
            float reflectProb = reflectProbOf(dotHV, q.hero, F);
#pragma unroll
            for (int i = 0; i < kNumL; ++i) fs[i] = F[i] * d * G / (4 * nvnl) * sn;
            if (dotHV * sign > 0) *pdf = reflectProb / (4 * dotHV * sign) * D.pdfVisible(sign * q.dir, m);
            return;
        }
        if (nvnl < 0) {
            transmissionValue(q, wi, entering, fs);
#pragma unroll
            for (int w = 0; w < kNumL; ++w) {
                float eEn = entering ? etaExt[w] : etaInt[w];
                fs[w] = fs[w] / fabsf(nvnl) * (eEn * eEn) * sn;
            }
            const int h = q.hero;
            const float extH = pick16(etaExt, h), intH = pick16(etaInt, h);
            float eEnterH = entering ? extH : intH;
            float eExitH = entering ? intH : extH;
            float3 m = normalize3(-(eEnterH * q.dir + eExitH * wi));
            float dotHV = dot3(q.dir, m);
            if (dotHV * sign <= 0) return;
            float F[kNumL];
            float reflectProb = reflectProbOf(dotHV, h, F);
            float dotHL = dot3(wi, m);
            float den = eEnterH * dotHV + eExitH * dotHL;
            float common = (1 - reflectProb) / (den * den);
            *pdf = common * D.pdfVisible(sign * q.dir, m) * eExitH * eExitH * fabsf(dotHL);
            return;
        }
#pragma unroll
        for (int i = 0; i < kNumL; ++i) fs[i] = 0.0f;
    }
};

}  // namespace slrh
