/*
 * slr_oracle.c - CPU restatement of SLR's spectral path tracer, the checker
 * for the HIP path.
 *
 * TEST INFRASTRUCTURE ONLY. Tests, the smoke check and the CPU-baseline leg
 * of bench.py load it (oracle/_build/liboracle.so) to check or time; the
 * product never calls it.
 *
 * Scalar, single-threaded-per-call C. Each function cites the reference code
 * it restates. The integrator is PTRenderer::Job::contribution
 * (libSLR/Renderer/PTRenderer.cpp:146-277) with the camera weighting of
 * Job::kernel (:103-144); the random-number plumbing is the renderer's own
 * design (one XORShift128 stream per path slot, a fixed draw order) so that
 * an image from this file and one from the GPU share their random numbers
 * path for path. Intersection is brute force over all triangles in index
 * order with the reference's Moller-Trumbore test and inclusive t <= tmax
 * rule (TriangleSurfaceShape.cpp:205-299), which yields the same closest hit
 * as any correct traversal except for exact ties in t.
 *
 * With orc_scene_compat bit 0 and no path-length cap the estimator is
 * VolumetricPTRenderer::Job::contribution (VolumetricPTRenderer.cpp:145-296)
 * over a scene without media, including its NaN for a light sample in the
 * shading plane.
 *
 * Pinned by: tests/test_oracle.py (XORShift streams, wavelength samples,
 * sampling helpers, film binning and BSDF sample/evaluate/PDF tuples against
 * the compiled reference; ray hits against the reference QBVH) and
 * tests/test_oracle_render.py (images against the reference PTRenderer and,
 * in the compat mode, VolumetricPTRenderer within 3 sigma; white-furnace
 * closed form).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NL 16
#define LLOW 360.0f
#define LHIGH 830.0f
#define PI_D 3.14159265358979323846

/* ------------------------------------------------------------------ RNG */

/* XORShift128 (libSLR/RNG/XORShiftRNG.cpp:29-36) */
uint32_t orc_xs_next(uint32_t s[4]) {
    uint32_t t = s[0] ^ (s[0] << 11);
    s[0] = s[1];
    s[1] = s[2];
    s[2] = s[3];
    s[3] = (s[3] ^ (s[3] >> 19)) ^ (t ^ (t >> 8));
    return s[3];
}

/* getFloat0cTo1o (libSLR/Core/random_number_generator.cpp:12-15) */
float orc_xs_float(uint32_t s[4]) {
    uint32_t bits = (orc_xs_next(s) >> 9) | 0x3f800000u;
    float f;
    memcpy(&f, &bits, 4);
    return f - 1.0f;
}

/* XORShiftRNG(seed) (XORShiftRNG.cpp:21-27). The seed is a signed 32-bit
 * integer there (Types32bit::Int), so seed >> 30 shifts arithmetically. */
void orc_xs_seed(uint32_t seed, uint32_t s[4]) {
    int32_t si = (int32_t)seed;
    for (int i = 0; i < 4; ++i) {
        uint32_t v = 1812433253u * (uint32_t)(si ^ (si >> 30)) + (uint32_t)i;
        s[i] = v;
        si = (int32_t)v;
    }
    for (int i = 0; i < 50; ++i) orc_xs_next(s);
}

/* The renderer's per-slot seed: a hash of (render seed, slot) fed to the
 * reference seeding procedure. */
void orc_slot_seed(uint32_t seed, uint32_t slot, uint32_t s[4]) {
    uint32_t h = seed ^ (0x9E3779B9u * (slot + 1u));
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    orc_xs_seed(h, s);
}

/* ------------------------------------------------------------------ vectors */

typedef struct { float x, y, z; } v3;
static v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 mul(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static v3 smul(float s, v3 a) { return V(s * a.x, s * a.y, s * a.z); }
static v3 neg(v3 a) { return V(-a.x, -a.y, -a.z); }
/* Vector3D::operator/ multiplies by the reciprocal (Vector3D.h:31) */
static v3 divs(v3 a, float s) { float r = 1.0f / s; return V(a.x * r, a.y * r, a.z * r); }
static float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 cross(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static float len(v3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
static v3 nrmz(v3 a) { return divs(a, len(a)); }

/* ------------------------------------------------------------------ spectra */

/* WavelengthSamples::createWithEqualOffsets (spectrum_types.h:59-69) */
static float lambda_of(int i, float off) { return LLOW + (LHIGH - LLOW) * (i + off) / NL; }

void orc_wavelengths(float off, float u, float* lambdas, int* hero, float* pdf) {
    for (int i = 0; i < NL; ++i) lambdas[i] = lambda_of(i, off);
    unsigned h = (unsigned)(uint16_t)(NL * u);
    *hero = (int)(h < NL - 1 ? h : NL - 1);
    *pdf = NL / (LHIGH - LLOW);
}

/* SampledSpectrum::importance (spectrum_types.h:473-486) */
static float importance(const float* v, int hero) {
    float sum = 0;
    for (int i = 0; i < NL; ++i) sum += v[i];
    const float primary = 0.9f;
    const float marginal = (1 - primary) / (NL - 1);
    return sum * marginal + v[hero] * (primary - marginal);
}

/* RegularContinuousSpectrum::evaluate (spectrum_types.cpp:82-100) */
void orc_spectrum_eval(float lo, float hi, int n, const float* vals, const float* lambdas, float* out) {
    for (int i = 0; i < NL; ++i) {
        float binF = (lambdas[i] - lo) / (hi - lo) * (n - 1);
        if (binF <= 0.0f) { out[i] = vals[0]; continue; }
        if (binF >= n - 1) { out[i] = vals[n - 1]; continue; }
        int b = (int)binF;
        float t = binF - b;
        out[i] = (1 - t) * vals[b] + t * vals[b + 1];
    }
}

/* SpectrumStorage::add of one sample (spectrum_types.h:725-736) */
void orc_film_bin(const float* lambdas, const float* value, float* strata) {
    const float rec = NL / (LHIGH - LLOW);
    for (int i = 0; i < NL; ++i) strata[i] = 0.0f;
    for (int i = 0; i < NL; ++i) {
        unsigned b = (unsigned)((lambdas[i] - LLOW) / (LHIGH - LLOW) * NL);
        if (b > NL - 1) b = NL - 1;
        strata[b] += value[i] * rec;
    }
}

/* ------------------------------------------------------------------ sampling */

/* concentricSampleDisk (distributions.cpp:57-90) */
static void concentric(float u0, float u1, float* dx, float* dy) {
    float r, th;
    float sx = 2 * u0 - 1, sy = 2 * u1 - 1;
    if (sx == 0 && sy == 0) { *dx = 0; *dy = 0; return; }
    if (sx >= -sy) {
        if (sx > sy) { r = sx; th = sy / sx; }
        else { r = sy; th = 2 - sx / sy; }
    } else {
        if (sx > sy) { r = -sy; th = 6 + sx / sy; }
        else { r = -sx; th = 4 + sy / sx; }
    }
    th = (float)(th * 0.78539816339744830962);
    *dx = r * cosf(th);
    *dy = r * sinf(th);
}

/* cosineSampleHemisphere (distributions.h:30-37) */
static v3 cos_hemi(float u0, float u1) {
    float x, y;
    concentric(u0, u1, &x, &y);
    return V(x, y, sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y)));
}

void orc_sampling(int n, const float* u, float* disk, float* hemi, float* tri) {
    for (int i = 0; i < n; ++i) {
        concentric(u[2 * i], u[2 * i + 1], &disk[2 * i], &disk[2 * i + 1]);
        v3 h = cos_hemi(u[2 * i], u[2 * i + 1]);
        hemi[3 * i] = h.x; hemi[3 * i + 1] = h.y; hemi[3 * i + 2] = h.z;
        /* uniformSampleTriangle (distributions.h:68-72) */
        float su = sqrtf(u[2 * i]);
        tri[2 * i] = 1.0f - su;
        tri[2 * i + 1] = u[2 * i + 1] * su;
    }
}

/* ------------------------------------------------------------------ BSDFs */

enum { K_LAMBERT = 0, K_OREN_NAYAR = 1, K_CONDUCTOR = 2, K_DIELECTRIC = 3, K_NONE = 4 };

typedef struct {
    int kind;
    float A[NL], B[NL];   /* reflectance | eta, k | etaExt, etaInt */
    float p0, p1;         /* ON A,B | alpha_x, alpha_y */
} bsdf_t;

/* GGXMicrofacetDistribution (directional_distribution_functions.cpp:247-315) */
static float ggx_D(const bsdf_t* b, v3 m) {
    if (m.z <= 0) return 0.0f;
    float ax = b->p0, ay = b->p1;
    float temp = m.x * m.x / (ax * ax) + m.y * m.y / (ay * ay) + m.z * m.z;
    return (float)(1.0 / (PI_D * ax * ay * temp * temp));
}
static float ggx_G1(const bsdf_t* b, v3 v, v3 m) {
    float ax = b->p0, ay = b->p1;
    float chi = (dot(v, m) / v.z) > 0 ? 1.0f : 0.0f;
    float t2 = (v.x * v.x * ax * ax + v.y * v.y * ay * ay) / (v.z * v.z);
    return chi * 2 / (1 + sqrtf(1 + t2));
}
static float ggx_pdf(const bsdf_t* b, v3 v, v3 m) { return ggx_G1(b, v, m) * fabsf(dot(v, m)) * ggx_D(b, m) / fabsf(v.z); }
static float ggx_sample(const bsdf_t* b, v3 v, float u0, float u1, v3* m, float* pdf) {
    float ax = b->p0, ay = b->p1;
    v3 sv = nrmz(V(ax * v.x, ay * v.y, v.z));
    float d2 = sqrtf(sv.x * sv.x + sv.y * sv.y);
    float rec = 1.0f / d2;
    v3 T1 = (sv.z < 0.9999f) ? V(sv.y * rec, -sv.x * rec, 0) : V(1, 0, 0);
    v3 T2 = V(T1.y * sv.z, -T1.x * sv.z, d2);
    float a = 1.0f / (1.0f + sv.z);
    float r = sqrtf(u0);
    float phi = (float)PI_D * ((u1 < a) ? u1 / a : 1 + (u1 - a) / (1.0f - a));
    float P1 = r * cosf(phi);
    float P2 = r * sinf(phi) * ((u1 < a) ? 1.0f : sv.z);
    v3 mm = add(add(smul(P1, T1), smul(P2, T2)), smul(sqrtf(1.0f - P1 * P1 - P2 * P2), sv));
    mm = nrmz(V(ax * mm.x, ay * mm.y, mm.z));
    *m = mm;
    float D = ggx_D(b, mm);
    *pdf = ggx_G1(b, v, mm) * fabsf(dot(v, mm)) * D / fabsf(v.z);
    return D;
}

/* FresnelConductor::evaluate (directional_distribution_functions.cpp:118-128) */
static float fr_cond(float eta, float k, float c) {
    c = fabsf(c);
    float c2 = c * c;
    float te = 2.0f * eta * c;
    float tf = eta * eta + k * k;
    float tm = tf * c2;
    float rpa = (tm - te + 1) / (tm + te + 1);
    float rpe = (tf - te + c2) / (tf + te + c2);
    return (rpa + rpe) / 2.0f;
}

/* FresnelDielectric::evaluate (directional_distribution_functions.cpp:182-211) */
static float fr_diel(float ext, float in, float c) {
    c = fminf(fmaxf(c, -1.0f), 1.0f);
    int entering = c > 0.0f;
    float en = entering ? ext : in, ex = entering ? in : ext;
    float sinEx = en / ex * sqrtf(fmaxf(0.0f, 1.0f - c * c));
    c = fabsf(c);
    if (sinEx >= 1.0f) return 1.0f;
    float cosEx = sqrtf(fmaxf(0.0f, 1.0f - sinEx * sinEx));
    float rpa = ((ex * c) - (en * cosEx)) / ((ex * c) + (en * cosEx));
    float rpe = ((en * c) - (ex * cosEx)) / ((en * c) + (ex * cosEx));
    return (rpa * rpa + rpe * rpe) / 2.0f;
}

static v3 halfv(v3 a, v3 b) { return nrmz(add(a, b)); }
/* BSDF front end's shading-normal correction, non-adjoint (directional_distribution_functions.h:295-331) */
static float sncorr(v3 d, v3 ng) { return fabsf(d.z / dot(d, ng)); }

/* Oren-Nayar factor (OrenNayerBRDF.cpp:13-58; the reference's "sin" is 1 - cos^2) */
static float on_factor(const bsdf_t* b, v3 wi, v3 wo, int sampled) {
    float sI = 1.0f - wi.z * wi.z, sO = 1.0f - wo.z * wo.z;
    float tI = sI / fabsf(wi.z), tO = sO / fabsf(wo.z);
    float sa = fmaxf(sI, sO), tb = fminf(tI, tO);
    float cda = (wi.x * wo.x + wi.y * wo.y) / (sI * sO);
    if (sampled && !isfinite(cda)) cda = 0.0f;
    float mx = (0.0f < cda) ? cda : 0.0f;
    return b->p0 + b->p1 * mx * sa * tb;
}

static void diffuse_value(const bsdf_t* b, v3 wi, v3 wo, int sampled, float* fs) {
    if (b->kind == K_OREN_NAYAR) {
        float f = on_factor(b, wi, wo, sampled);
        for (int i = 0; i < NL; ++i) fs[i] = b->A[i] * (float)(f / PI_D);
    } else {
        /* m_R / M_PI: SampledSpectrum::operator/(float) multiplies by the
           float reciprocal of (float)M_PI (spectrum_types.h:351-357) */
        const float r = 1.0f / (float)PI_D;
        for (int i = 0; i < NL; ++i) fs[i] = b->A[i] * r;
    }
}

static void diel_trans(const bsdf_t* b, v3 q, v3 wi, int entering, float* fs) {
    for (int w = 0; w < NL; ++w) {
        float en = entering ? b->A[w] : b->B[w], ex = entering ? b->B[w] : b->A[w];
        v3 mw = nrmz(neg(add(smul(en, q), smul(ex, wi))));
        float dv = dot(q, mw), dl = dot(wi, mw);
        float F = fr_diel(b->A[w], b->B[w], dv);
        float G = ggx_G1(b, q, mw) * ggx_G1(b, wi, mw);
        float D = ggx_D(b, mw);
        float den = en * dv + ex * dl;
        fs[w] = fabsf(dv * dl) * (1 - F) * G * D / (den * den);
    }
}

static float diel_reflect_prob(const bsdf_t* b, float dotHV, int hero, float* F) {
    for (int i = 0; i < NL; ++i) F[i] = fr_diel(b->A[i], b->B[i], dotHV);
    return importance(F, hero);
}

/* BSDF::sample -> *Internal (basic_bsdfs.cpp:13-26, OrenNayerBRDF.cpp:13-35,
 * microfacet_bsdfs.cpp:11-40 and 116-208). Returns 1 with fs (snCorrection
 * applied), dir, pdf, refl set; 0 when the sample is rejected (pdf 0). */
static int bsdf_sample(const bsdf_t* b, v3 q, v3 ng, int hero, float uc, float u0, float u1, float* fs, v3* dir,
                       float* pdf, int* refl) {
    for (int i = 0; i < NL; ++i) fs[i] = 0.0f;
    *pdf = 0.0f;
    *refl = 1;
    if (b->kind == K_LAMBERT || b->kind == K_OREN_NAYAR) {
        v3 d = cos_hemi(u0, u1);
        *pdf = (float)(d.z / PI_D);
        d.z *= q.z > 0 ? 1 : -1;
        *dir = d;
        diffuse_value(b, d, q, 1, fs);
        float sn = sncorr(d, ng);
        for (int i = 0; i < NL; ++i) fs[i] *= sn;
        return 1;
    }
    int entering = q.z >= 0.0f;
    float sign = entering ? 1.0f : -1.0f;
    v3 m;
    float mpdf;
    float D = ggx_sample(b, smul(sign, q), u0, u1, &m, &mpdf);
    float dotHV = dot(q, m);
    if (b->kind == K_CONDUCTOR) {
        if (dotHV * sign <= 0) return 0;
        v3 wi = sub(smul(2 * dotHV, m), q);
        if (wi.z * q.z <= 0) return 0;
        *pdf = 1.0f / (4 * dotHV * sign) * mpdf;
        *dir = wi;
        float G = ggx_G1(b, q, m) * ggx_G1(b, wi, m);
        float sn = sncorr(wi, ng);
        float den = 4 * q.z * wi.z;
        for (int i = 0; i < NL; ++i) fs[i] = fr_cond(b->A[i], b->B[i], dotHV) * D * G / den * sn;
        return 1;
    }
    /* dielectric */
    if (dotHV * sign <= 0 || isnan(D)) return 0;
    float F[NL];
    float rp = diel_reflect_prob(b, dotHV, hero, F);
    float enH = entering ? b->A[hero] : b->B[hero], exH = entering ? b->B[hero] : b->A[hero];
    if (uc < rp) {
        v3 wi = sub(smul(2 * dotHV, m), q);
        if (wi.z * q.z <= 0) return 0;
        *pdf = rp / (4 * dotHV * sign) * mpdf;
        *dir = wi;
        float G = ggx_G1(b, q, m) * ggx_G1(b, wi, m);
        float sn = sncorr(wi, ng);
        float den = 4 * q.z * wi.z;
        for (int i = 0; i < NL; ++i) fs[i] = F[i] * D * G / den * sn;
        return 1;
    }
    float rr = enH / exH;
    float inner = 1 + rr * rr * (dotHV * dotHV - 1);
    if (inner < 0) return 0;
    v3 wi = sub(smul(rr * dotHV - sign * sqrtf(inner), m), smul(rr, q));
    if (wi.z * q.z >= 0) return 0;
    float dotHL = dot(wi, m);
    float den = enH * dotHV + exH * dotHL;
    *pdf = (1 - rp) / (den * den) * mpdf * exH * exH * fabsf(dotHL);
    *dir = wi;
    *refl = 0;
    diel_trans(b, q, wi, entering, fs);
    float sn = sncorr(wi, ng);
    float iv = fabsf(q.z * wi.z);
    for (int w = 0; w < NL; ++w) {
        float en = entering ? b->A[w] : b->B[w];
        fs[w] = fs[w] / iv * (en * en) * sn;
    }
    return 1;
}

/* BSDF::evaluate + evaluatePDF */
static void bsdf_eval(const bsdf_t* b, v3 q, v3 ng, int hero, v3 wi, float* fs, float* pdf) {
    for (int i = 0; i < NL; ++i) fs[i] = 0.0f;
    *pdf = 0.0f;
    float sn = sncorr(wi, ng);
    if (b->kind == K_LAMBERT || b->kind == K_OREN_NAYAR) {
        if (q.z * wi.z <= 0.0f) return;
        diffuse_value(b, wi, q, 0, fs);
        for (int i = 0; i < NL; ++i) fs[i] *= sn;
        *pdf = (float)(fabsf(wi.z) / PI_D);
        return;
    }
    int entering = q.z >= 0.0f;
    float sign = entering ? 1.0f : -1.0f;
    float nvnl = wi.z * q.z;
    if (b->kind == K_CONDUCTOR) {
        if (nvnl <= 0) return;
        v3 m = smul(sign, halfv(q, wi));
        float dotHV = dot(q, m);
        float D = ggx_D(b, m);
        float G = ggx_G1(b, q, m) * ggx_G1(b, wi, m);
        float den = 4 * q.z * wi.z;
        for (int i = 0; i < NL; ++i) fs[i] = fr_cond(b->A[i], b->B[i], dotHV) * D * G / den * sn;
        if (dotHV * sign <= 0) return;
        *pdf = 1.0f / (4 * dotHV * sign) * ggx_pdf(b, smul(sign, q), m);
        return;
    }
    if (nvnl > 0) {
        v3 m = smul(sign, halfv(q, wi));
        float dotHV = dot(q, m);
        float D = ggx_D(b, m);
        float G = ggx_G1(b, q, m) * ggx_G1(b, wi, m);
        float F[NL];
        float rp = diel_reflect_prob(b, dotHV, hero, F);
        for (int i = 0; i < NL; ++i) fs[i] = F[i] * D * G / (4 * nvnl) * sn;
        if (dotHV * sign > 0) *pdf = rp / (4 * dotHV * sign) * ggx_pdf(b, smul(sign, q), m);
        return;
    }
    if (nvnl < 0) {
        diel_trans(b, q, wi, entering, fs);
        for (int w = 0; w < NL; ++w) {
            float en = entering ? b->A[w] : b->B[w];
            fs[w] = fs[w] / fabsf(nvnl) * (en * en) * sn;
        }
        float enH = entering ? b->A[hero] : b->B[hero], exH = entering ? b->B[hero] : b->A[hero];
        v3 m = nrmz(neg(add(smul(enH, q), smul(exH, wi))));
        float dotHV = dot(q, m);
        if (dotHV * sign <= 0) return;
        float F[NL];
        float rp = diel_reflect_prob(b, dotHV, hero, F);
        float dotHL = dot(wi, m);
        float den = enH * dotHV + exH * dotHL;
        *pdf = (1 - rp) / (den * den) * ggx_pdf(b, smul(sign, q), m) * exH * exH * fabsf(dotHL);
    }
}

/* Same interface as the reference harness's refd_bsdf. */
int orc_bsdf(int kind, const float* A, const float* B, const float* p, int n, const float* q, const int* wl,
             int mode, const float* uOrDir, float* out_dir, float* out_pdf, float* out_fs, int* out_type) {
    bsdf_t b;
    b.kind = kind;
    memcpy(b.A, A, sizeof(b.A));
    if (B) memcpy(b.B, B, sizeof(b.B));
    if (kind == K_OREN_NAYAR) {
        float s = p[0];
        b.p0 = (float)(1.0f - 0.5f * s * s / (s * s + 0.33));
        b.p1 = (float)(0.45 * s * s / (s * s + 0.09));
    } else {
        b.p0 = p[0];
        b.p1 = p[1];
    }
    for (int i = 0; i < n; ++i) {
        v3 qq = V(q[6 * i], q[6 * i + 1], q[6 * i + 2]);
        v3 ng = V(q[6 * i + 3], q[6 * i + 4], q[6 * i + 5]);
        if (mode == 0) {
            const float* u = uOrDir + 3 * i;
            v3 d = V(0, 0, 0);
            float pdf;
            int refl;
            int ok = bsdf_sample(&b, qq, ng, wl[i], u[0], u[1], u[2], out_fs + NL * i, &d, &pdf, &refl);
            out_dir[3 * i] = d.x; out_dir[3 * i + 1] = d.y; out_dir[3 * i + 2] = d.z;
            out_pdf[i] = ok ? pdf : 0.0f;
            out_type[i] = ok && pdf > 0 ? (refl ? 1 : 2) : 0;
        } else {
            const float* d = uOrDir + 3 * i;
            bsdf_eval(&b, qq, ng, wl[i], V(d[0], d[1], d[2]), out_fs + NL * i, &out_pdf[i]);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ scene */

typedef struct {
    int nv, nt, nmat, nspec;
    float *pos, *nrm, *tan;   /* 3 per vertex */
    uint32_t* idx;            /* 3 per triangle */
    uint32_t* tmat;
    int *mkind, *mspec0, *mspec1, *memit;
    float *mp0, *mp1;         /* ON: sigma | GGX: alpha_x, alpha_y */
    float *slo, *shi;
    int *soff, *sn;
    float* pool;
    int nlights;
    uint32_t* lights;
    float* lcdf;
    float* lpmf;
    int* prim2light;
    /* camera */
    float l2w[12], nrm3[9];
    float opW, opH, objDist, imgDist, area, lensR;
    int tangentNeeNaN;        /* see orc_scene_compat */
} scene_t;

static void* dupa(const void* p, size_t bytes) {
    void* r = malloc(bytes ? bytes : 1);
    if (p && bytes) memcpy(r, p, bytes);
    return r;
}

/* Materials: kind 0 Lambert(spec0), 1 Oren-Nayar(spec0, p0 = sigma),
 * 2 conductor(spec0 eta, spec1 k, p0/p1 alpha), 3 dielectric(spec0 ext,
 * spec1 int), 4 none; emit = emittance spectrum or -1. The camera matrix is
 * column-major 4x4 (Matrix4x4.h:42-46). */
void* orc_scene_new(int nv, const float* pos, const float* nrm, const float* tan, int nt, const uint32_t* idx,
                    const uint32_t* tmat, int nmat, const int* kind, const int* spec0, const int* spec1,
                    const int* emit, const float* p0, const float* p1, int nspec, const float* lo, const float* hi,
                    const int* off, const int* n, int npool, const float* pool, const float* l2w, float aspect,
                    float fovY, float lensR, float imgDist, float objDist) {
    scene_t* s = (scene_t*)calloc(1, sizeof(scene_t));
    s->nv = nv; s->nt = nt; s->nmat = nmat; s->nspec = nspec;
    s->pos = dupa(pos, 12ull * nv); s->nrm = dupa(nrm, 12ull * nv); s->tan = dupa(tan, 12ull * nv);
    s->idx = dupa(idx, 12ull * nt); s->tmat = dupa(tmat, 4ull * nt);
    s->mkind = dupa(kind, 4ull * nmat); s->mspec0 = dupa(spec0, 4ull * nmat); s->mspec1 = dupa(spec1, 4ull * nmat);
    s->memit = dupa(emit, 4ull * nmat); s->mp0 = dupa(p0, 4ull * nmat); s->mp1 = dupa(p1, 4ull * nmat);
    s->slo = dupa(lo, 4ull * nspec); s->shi = dupa(hi, 4ull * nspec); s->soff = dupa(off, 4ull * nspec);
    s->sn = dupa(n, 4ull * nspec); s->pool = dupa(pool, 4ull * npool);
    /* light list: emitting triangles in object order, importance 1 each,
     * compensated CDF (surface_object.cpp:379-398, distributions.cpp:99-123) */
    s->lights = malloc(4ull * (nt + 1));
    s->prim2light = malloc(4ull * (nt + 1));
    for (int t = 0; t < nt; ++t) {
        s->prim2light[t] = -1;
        if (s->memit[s->tmat[t]] >= 0) { s->prim2light[t] = s->nlights; s->lights[s->nlights++] = (uint32_t)t; }
    }
    s->lcdf = malloc(4ull * (s->nlights + 1));
    s->lpmf = malloc(4ull * (s->nlights + 1));
    float sum = 0, comp = 0;
    s->lcdf[0] = 0;
    for (int i = 0; i < s->nlights; ++i) {
        float in = 1.0f - comp, tmp = sum + in;
        comp = (tmp - sum) - in;
        sum = tmp;
        s->lcdf[i + 1] = sum;
    }
    for (int i = 0; i < s->nlights; ++i) { s->lpmf[i] = 1.0f / sum; s->lcdf[i + 1] /= sum; }
    /* PerspectiveCamera (PerspectiveCamera.cpp:16-27) */
    s->opH = 2.0f * objDist * tanf(fovY * 0.5f);
    s->opW = s->opH * aspect;
    s->area = (float)(s->opW * s->opH * pow(imgDist / objDist, 2));
    s->objDist = objDist; s->imgDist = imgDist; s->lensR = lensR;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) s->l2w[4 * r + c] = l2w[4 * c + r];
    double a[3][3], inv[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a[r][c] = l2w[4 * c + r];
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
    inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
    inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
    inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
    inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
    inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
    inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
    inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
    inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) s->nrm3[3 * r + c] = (float)inv[c][r];
    return s;
}

/* flags bit 0: the reference VolumetricPTRenderer's treatment of a light
 * sample in the shading point's plane (the sample becomes NaN and is not
 * filed); with it and an unlimited path length the integrator is
 * VolumetricPTRenderer::Job::contribution (VolumetricPTRenderer.cpp:145-293)
 * over a scene without media. */
void orc_scene_compat(void* p, int flags) { ((scene_t*)p)->tangentNeeNaN = flags & 1; }

void orc_scene_free(void* p) {
    scene_t* s = (scene_t*)p;
    if (!s) return;
    free(s->pos); free(s->nrm); free(s->tan); free(s->idx); free(s->tmat);
    free(s->mkind); free(s->mspec0); free(s->mspec1); free(s->memit); free(s->mp0); free(s->mp1);
    free(s->slo); free(s->shi); free(s->soff); free(s->sn); free(s->pool);
    free(s->lights); free(s->prim2light); free(s->lcdf); free(s->lpmf);
    free(s);
}

static v3 vtx(const float* a, uint32_t i) { return V(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }

/* Moller-Trumbore (TriangleSurfaceShape.cpp:205-257) */
static int tri_test(const scene_t* s, uint32_t t, v3 o, v3 d, float tmin, float tmax, float* tt, float* b1o,
                    float* b2o) {
    v3 p0 = vtx(s->pos, s->idx[3 * t]), p1 = vtx(s->pos, s->idx[3 * t + 1]), p2 = vtx(s->pos, s->idx[3 * t + 2]);
    v3 e1 = sub(p1, p0), e2 = sub(p2, p0);
    v3 p = cross(d, e2);
    float det = dot(e1, p);
    if (det == 0.0f) return 0;
    float inv = 1.0f / det;
    v3 dd = sub(o, p0);
    float b1 = dot(dd, p) * inv;
    if (b1 < 0.0f || b1 > 1.0f) return 0;
    v3 q = cross(dd, e1);
    float b2 = dot(d, q) * inv;
    if (b2 < 0.0f || b1 + b2 > 1.0f) return 0;
    float tv = dot(e2, q) * inv;
    if (tv < tmin || tv > tmax) return 0;
    *tt = tv; *b1o = b1; *b2o = b2;
    return 1;
}

/* closest hit, brute force; returns prim or -1 */
static int closest(const scene_t* s, v3 o, v3 d, float tmin, float tmax, float* t, float* b0, float* b1) {
    int prim = -1;
    float hb1 = 0, hb2 = 0;
    for (int i = 0; i < s->nt; ++i) {
        float tt, u, v;
        if (tri_test(s, (uint32_t)i, o, d, tmin, tmax, &tt, &u, &v)) { tmax = tt; prim = i; hb1 = u; hb2 = v; }
    }
    *t = tmax;
    *b0 = prim >= 0 ? 1.0f - hb1 - hb2 : 0.0f;
    *b1 = prim >= 0 ? hb1 : 0.0f;
    return prim;
}

static int occluded(const scene_t* s, v3 o, v3 d, float tmin, float tmax) {
    for (int i = 0; i < s->nt; ++i) {
        float tt, u, v;
        if (tri_test(s, (uint32_t)i, o, d, tmin, tmax, &tt, &u, &v)) return 1;
    }
    return 0;
}

int orc_intersect(void* p, int n, const float* rays, float* t, int32_t* prim, float* b0, float* b1) {
    const scene_t* s = (const scene_t*)p;
    for (int i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        prim[i] = closest(s, V(r[0], r[1], r[2]), V(r[3], r[4], r[5]), r[6], r[7], &t[i], &b0[i], &b1[i]);
    }
    return 0;
}

/* QBVH traversal with counters: QBVH::commonProcedure (libSLR/Accelerator/
 * QBVH.h:294-341) over the tree as slrh_scene_qbvh_dump writes it (per node
 * 24 floats: min x/y/z then max x/y/z, 4 children each; 7 words: the four
 * children as the reference's Children bitfield {idx:27, numLeaves:4,
 * isLeaf:1}, all ones when invalid, then the top, left and right split
 * axes), Node::intersect's slab test (:54-75) in scalar form with the same
 * operations, the triangles of each leaf tested in order with the segment's
 * far end narrowed to each hit (closest hit), or stopping at the first hit
 * (any). counts[0] += nodes popped (each tests its four child boxes),
 * counts[1] += triangles tested: the per-ray work the algorithmic bytes of
 * a traversal are priced on. Returns -1 on a stack overflow. */
int orc_qbvh_trace(void* p, const float* boxes, const uint32_t* info, const int32_t* refs, int n,
                   const float* rays, int any, float* tout, int32_t* pout, float* b0out, float* b1out,
                   float* vis, double* counts) {
    const scene_t* s = (const scene_t*)p;
    static const uint32_t order_table[8] = {0x0123, 0x0132, 0x1023, 0x1032, 0x2301, 0x3201, 0x2310, 0x3210};
    double nodes = 0, tris = 0;
    for (int i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        const v3 o = V(r[0], r[1], r[2]), d = V(r[3], r[4], r[5]);
        const float tmin = r[6];
        float tmax = r[7];
        const float inv[3] = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
        const float org[3] = {o.x, o.y, o.z};
        const int pos[3] = {d.x >= 0, d.y >= 0, d.z >= 0};
        uint32_t stack[64];
        int depth = 0, prim = -1, hit = 0;
        float hb1 = 0, hb2 = 0;
        stack[depth++] = 0;
        while (depth > 0 && !hit) {
            const uint32_t ni = stack[--depth];
            const float* bx = boxes + 24 * (size_t)ni;
            const uint32_t* in = info + 7 * (size_t)ni;
            nodes += 1;
            uint32_t flags = 0;
            for (int c = 0; c < 4; ++c) {
                float tn = tmin, tf = tmax;
                for (int a = 0; a < 3; ++a) {
                    const float lo = bx[4 * a + c], hi = bx[12 + 4 * a + c];
                    const float nearB = inv[a] > 0.0f ? lo : hi, farB = inv[a] > 0.0f ? hi : lo;
                    const float t0 = (nearB - org[a]) * inv[a], t1 = (farB - org[a]) * inv[a];
                    /* _mm_max_ps(x, tNear) / _mm_min_ps(x, tFar): the second
                     * operand when either is NaN */
                    tn = t0 > tn ? t0 : tn;
                    tf = t1 < tf ? t1 : tf;
                }
                if (tn <= tf) flags |= 1u << c;
            }
            if (flags == 0) continue;
            const uint32_t enc = order_table[4 * pos[in[4]] + 2 * pos[in[5]] + pos[in[6]]];
            uint32_t ch[4];
            int ord[4];
            for (int k = 0; k < 4; ++k) {
                ord[k] = (int)((enc >> (4 * k)) & 0xF);
                ch[k] = ((flags >> ord[k]) & 1u) ? in[ord[k]] : 0xFFFFFFFFu;
            }
            for (int k = 3; k >= 0; --k) {
                if (ch[k] == 0xFFFFFFFFu || (ch[k] >> 31)) continue;
                if (depth >= 64) return -1;
                stack[depth++] = ch[k] & 0x07FFFFFFu;
            }
            for (int k = 0; k < 4 && !hit; ++k) {
                if (ch[k] == 0xFFFFFFFFu || !(ch[k] >> 31)) continue;
                const uint32_t first = ch[k] & 0x07FFFFFFu, cnt = (ch[k] >> 27) & 0xFu;
                for (uint32_t j = 0; j < cnt; ++j) {
                    const int32_t t = refs[first + j];
                    float tt, u, v;
                    tris += 1;
                    if (tri_test(s, (uint32_t)t, o, d, tmin, tmax, &tt, &u, &v)) {
                        if (any) { hit = 1; break; }
                        tmax = tt; prim = t; hb1 = u; hb2 = v;
                    }
                }
            }
        }
        if (any) {
            if (vis) vis[i] = hit ? 0.0f : 1.0f;
        } else {
            tout[i] = tmax;
            pout[i] = prim;
            b0out[i] = prim >= 0 ? 1.0f - hb1 - hb2 : 0.0f;
            b1out[i] = prim >= 0 ? hb1 : 0.0f;
        }
    }
    counts[0] += nodes;
    counts[1] += tris;
    return 0;
}

int orc_visibility(void* p, int n, const float* rays, float* vis) {
    const scene_t* s = (const scene_t*)p;
    for (int i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        vis[i] = occluded(s, V(r[0], r[1], r[2]), V(r[3], r[4], r[5]), r[6], r[7]) ? 0.0f : 1.0f;
    }
    return 0;
}

static void spec(const scene_t* s, int id, const float* lam, float* out) {
    orc_spectrum_eval(s->slo[id], s->shi[id], s->sn[id], s->pool + s->soff[id], lam, out);
}

static void make_bsdf(const scene_t* s, int m, const float* lam, bsdf_t* b) {
    b->kind = s->mkind[m];
    if (b->kind == K_NONE) return;
    spec(s, s->mspec0[m], lam, b->A);
    if (b->kind >= K_CONDUCTOR) spec(s, s->mspec1[m], lam, b->B);
    if (b->kind == K_OREN_NAYAR) {
        float sg = s->mp0[m];
        b->p0 = (float)(1.0f - 0.5f * sg * sg / (sg * sg + 0.33));
        b->p1 = (float)(0.45 * sg * sg / (sg * sg + 0.09));
    } else {
        b->p0 = s->mp0[m];
        b->p1 = s->mp1[m];
    }
}

typedef struct { v3 p, gN, sx, sy, sz; int mat; } surf_t;
static v3 to_local(const surf_t* f, v3 v) { return V(dot(f->sx, v), dot(f->sy, v), dot(f->sz, v)); }
static v3 from_local(const surf_t* f, v3 v) {
    return V(dot(V(f->sx.x, f->sy.x, f->sz.x), v), dot(V(f->sx.y, f->sy.y, f->sz.y), v),
             dot(V(f->sx.z, f->sy.z, f->sz.z), v));
}

/* calculateSurfacePoint (TriangleSurfaceShape.cpp:301-346) */
static surf_t surface(const scene_t* s, int prim, float b0, float b1, v3 o, v3 d, float t) {
    surf_t f;
    uint32_t i0 = s->idx[3 * prim], i1 = s->idx[3 * prim + 1], i2 = s->idx[3 * prim + 2];
    float b2 = 1.0f - b0 - b1;
    f.sz = nrmz(add(add(smul(b0, vtx(s->nrm, i0)), smul(b1, vtx(s->nrm, i1))), smul(b2, vtx(s->nrm, i2))));
    v3 x = add(add(smul(b0, vtx(s->tan, i0)), smul(b1, vtx(s->tan, i1))), smul(b2, vtx(s->tan, i2)));
    float dnt = dot(f.sz, x);
    f.sx = nrmz(sub(x, smul(dnt, f.sz)));
    f.sy = cross(f.sz, f.sx);
    f.p = add(o, mul(d, t));
    v3 p0 = vtx(s->pos, i0), p1 = vtx(s->pos, i1), p2 = vtx(s->pos, i2);
    f.gN = nrmz(cross(sub(p1, p0), sub(p2, p0)));
    f.mat = (int)s->tmat[prim];
    return f;
}

static float tri_area(const scene_t* s, int prim) {
    v3 p0 = vtx(s->pos, s->idx[3 * prim]), p1 = vtx(s->pos, s->idx[3 * prim + 1]),
       p2 = vtx(s->pos, s->idx[3 * prim + 2]);
    return 0.5f * len(cross(sub(p1, p0), sub(p2, p0)));
}

/* DiscreteDistribution1D::sample (distributions.cpp:138-161) */
static int sample_light(const scene_t* s, float u, float* prob) {
    int n = s->nlights, idx = n;
    unsigned d = 1;
    while ((d << 1) <= (unsigned)n) d <<= 1;
    for (; d > 0; d >>= 1)
        if (idx - (int)d > 0 && s->lcdf[idx - d] >= u) idx -= (int)d;
    --idx;
    *prob = s->lpmf[idx];
    return idx;
}

/* One sample of one slot: camera ray then PTRenderer::Job::contribution.
 * Returns the weighted contribution C (16 values) and its wavelength offset. */
static void trace_sample(const scene_t* s, uint32_t st[4], uint32_t px, uint32_t py, uint32_t W, uint32_t H,
                         uint32_t maxDepth, float* C, float* offOut, long long* rays) {
    /* Job::kernel (PTRenderer.cpp:108-129), PerspectiveCamera::sample / PerspectiveIDF::sample
     * (PerspectiveCamera.cpp:31-82) */
    float fx = px + orc_xs_float(st);
    float fy = py + orc_xs_float(st);
    float off = orc_xs_float(st);
    float uL = orc_xs_float(st);
    unsigned hh = (unsigned)(uint16_t)(NL * uL);
    int hero = (int)(hh < NL - 1 ? hh : NL - 1);
    float lu0 = orc_xs_float(st), lu1 = orc_xs_float(st);
    float lam[NL];
    for (int i = 0; i < NL; ++i) lam[i] = lambda_of(i, off);
    *offOut = off;
    float lx = 0, ly = 0;
    concentric(lu0, lu1, &lx, &ly);
    v3 oL = V(s->lensR * lx, s->lensR * ly, 0.0f);
    const float* M = s->l2w;
    v3 org = V(M[0] * oL.x + M[1] * oL.y + M[2] * oL.z + M[3], M[4] * oL.x + M[5] * oL.y + M[6] * oL.z + M[7],
               M[8] * oL.x + M[9] * oL.y + M[10] * oL.z + M[11]);
    const float* N = s->nrm3;
    v3 fz = nrmz(V(N[2], N[5], N[8]));
    v3 fxv = nrmz(V(M[0], M[4], M[8]));
    v3 fyv = cross(fz, fxv);
    float ux = fx / W, uy = fy / H;
    v3 pf = V(s->opW * (0.5f - ux), s->opH * (0.5f - uy), s->objDist);
    v3 dl = nrmz(sub(pf, oL));
    float dirPDF = s->imgDist * s->imgDist / ((dl.z * dl.z * dl.z) * s->area);
    v3 d = V(dot(V(fxv.x, fyv.x, fz.x), dl), dot(V(fxv.y, fyv.y, fz.y), dl), dot(V(fxv.z, fyv.z, fz.z), dl));
    float areaPDF = s->lensR > 0.0f ? (float)(1.0 / (PI_D * s->lensR * s->lensR)) : 1.0f;
    float weight = fabsf(dot(d, fz)) / (areaPDF * dirPDF * (NL / (LHIGH - LLOW)));

    float alpha[NL], sp[NL];
    for (int i = 0; i < NL; ++i) { alpha[i] = 1.0f; sp[i] = 0.0f; }
    float initY = importance(alpha, hero);
    v3 o = org;
    float tmin = 0.0f;
    uint32_t pathLen = 0;
    float lastPdf = 0.0f;
    const float edf = (float)(1.0 / PI_D);
    for (;;) {
        float t, b0, b1;
        ++*rays;
        int prim = closest(s, o, d, tmin, INFINITY, &t, &b0, &b1);
        if (prim < 0) break;
        surf_t f = surface(s, prim, b0, b1, o, d, t);
        v3 dirOut = to_local(&f, neg(d));
        int em = s->memit[f.mat];
        if (em >= 0 && dirOut.z > 0.0f) {
            float Le[NL];
            spec(s, em, lam, Le);
            float mis = 1.0f;
            if (pathLen > 0) {
                v3 dv = sub(f.p, o);
                float dist2 = dot(dv, dv);
                float lightPDF = s->lpmf[s->prim2light[prim]] * (1.0f / tri_area(s, prim)) * dist2 /
                                 fabsf(dot(d, f.gN));
                mis = (lastPdf * lastPdf) / (lightPDF * lightPDF + lastPdf * lastPdf);
            }
            for (int k = 0; k < NL; ++k) sp[k] += alpha[k] * (Le[k] * edf) * mis;
        }
        if (pathLen > 0) {
            float p = fminf(importance(alpha, hero) / initY, 1.0f);
            if (orc_xs_float(st) < p) {
                /* alpha /= continueProb multiplies by 1 / p (spectrum_types.h:390-395) */
                const float r = 1.0f / p;
                for (int k = 0; k < NL; ++k) alpha[k] *= r;
            } else break;
        }
        ++pathLen;
        if (pathLen >= maxDepth || s->mkind[f.mat] == K_NONE) break;
        bsdf_t b;
        make_bsdf(s, f.mat, lam, &b);
        v3 gL = to_local(&f, f.gN);
        /* next event estimation (PTRenderer.cpp:179-218) */
        {
            float uSel = orc_xs_float(st), u0 = orc_xs_float(st), u1 = orc_xs_float(st);
            float lprob;
            int li = sample_light(s, uSel, &lprob);
            int lp_ = (int)s->lights[li];
            float su = sqrtf(u0);
            float lb0 = 1.0f - su, lb1 = u1 * su, lb2 = 1.0f - lb0 - lb1;
            uint32_t j0 = s->idx[3 * lp_], j1 = s->idx[3 * lp_ + 1], j2 = s->idx[3 * lp_ + 2];
            v3 q0 = vtx(s->pos, j0), q1 = vtx(s->pos, j1), q2 = vtx(s->pos, j2);
            v3 ln = nrmz(add(add(smul(lb0, vtx(s->nrm, j0)), smul(lb1, vtx(s->nrm, j1))), smul(lb2, vtx(s->nrm, j2))));
            v3 lpt = add(add(smul(lb0, q0), smul(lb1, q1)), smul(lb2, q2));
            v3 lg = nrmz(cross(sub(q1, q0), sub(q2, q0)));
            float aPDF = 1.0f / (0.5f * len(cross(sub(q1, q0), sub(q2, q0))));
            v3 dv = sub(lpt, f.p);
            float dist2 = dot(dv, dv);
            v3 sdir = divs(dv, sqrtf(dist2));
            float cosS = fabsf(dot(sdir, f.gN)), cosL = fabsf(dot(sdir, lg));
            float lz = dot(ln, neg(sdir));
            if (s->tangentNeeNaN && cosS == 0.0f) {
                /* VolumetricPTRenderer.cpp:198-230 has no cosine guard: a visible
                 * light point in the shading point's plane makes the BSDF's
                 * shading-normal correction 0/0 (directional_distribution_functions.h:325-327),
                 * a NaN that drops the whole sample (VolumetricPTRenderer.cpp:133-138) */
                float dist = len(sub(lpt, f.p));
                v3 vd = divs(sub(lpt, f.p), dist);
                ++*rays;
                if (!occluded(s, f.p, vd, 1e-4f, dist * (1 - 1e-4f)))
                    for (int k = 0; k < NL; ++k) sp[k] = NAN;
            } else if (cosS * cosL > 0 && lz > 0.0f) {
                float Mv[NL], fs[NL], bpdf;
                spec(s, s->memit[s->tmat[lp_]], lam, Mv);
                bsdf_eval(&b, dirOut, gL, hero, to_local(&f, sdir), fs, &bpdf);
                float lightPDF = lprob * aPDF;
                float bsdfPDF = bpdf * cosL / dist2;
                float mis = (lightPDF * lightPDF) / (lightPDF * lightPDF + bsdfPDF * bsdfPDF);
                float G = cosS * cosL / dist2;
                float sc = G * mis / lightPDF;
                float c[NL];
                int any = 0;
                for (int k = 0; k < NL; ++k) {
                    c[k] = alpha[k] * (Mv[k] * edf) * fs[k] * sc;
                    any |= c[k] != 0.0f;
                }
                if (any) {
                    /* Scene::testVisibility (Scene.cpp:96-112) */
                    float dist = len(sub(lpt, f.p));
                    v3 vd = divs(sub(lpt, f.p), dist);
                    ++*rays;
                    if (!occluded(s, f.p, vd, 1e-4f, dist * (1 - 1e-4f)))
                        for (int k = 0; k < NL; ++k) sp[k] += c[k];
                }
            }
        }
        /* BSDF sampling (PTRenderer.cpp:221-237) */
        {
            float uc = orc_xs_float(st), u0 = orc_xs_float(st), u1 = orc_xs_float(st);
            float fs[NL], pdf;
            v3 nd;
            int refl;
            int ok = bsdf_sample(&b, dirOut, gL, hero, uc, u0, u1, fs, &nd, &pdf, &refl);
            int nz = 0;
            for (int k = 0; k < NL; ++k) nz |= fs[k] != 0.0f;
            if (!(ok && nz && pdf != 0.0f)) break;
            float w = fabsf(dot(nd, gL)) / pdf;
            for (int k = 0; k < NL; ++k) alpha[k] *= fs[k] * w;
            o = f.p;
            d = from_local(&f, nd);
            tmin = 1e-4f;
            lastPdf = pdf;
        }
    }
    for (int k = 0; k < NL; ++k) C[k] = sp[k] * weight;
}

/* Renders slots [slot_begin, slot_end) of a W x H image with K slots per
 * pixel (slot s -> pixel s % (W*H), plane s / (W*H)), exactly the GPU's
 * assignment of samples to slots and random streams. film (H*W*16 doubles)
 * accumulates the binned contributions; returns rays traced. */
long long orc_render(void* p, uint32_t W, uint32_t H, uint32_t spp, uint32_t seed, uint32_t K, uint32_t maxDepth,
                     uint32_t slot_begin, uint32_t slot_end, double* film, long long* nonfinite) {
    const scene_t* s = (const scene_t*)p;
    const uint32_t P = W * H;
    long long rays = 0;
    if (maxDepth == 0) maxDepth = 100;
    for (uint32_t slot = slot_begin; slot < slot_end && slot < K * P; ++slot) {
        uint32_t pixel = slot % P, plane = slot / P;
        uint32_t total = spp > plane ? (spp - plane + K - 1) / K : 0;
        uint32_t st[4];
        orc_slot_seed(seed, slot, st);
        for (uint32_t k = 0; k < total; ++k) {
            float C[NL], off;
            trace_sample(s, st, pixel % W, pixel / W, W, H, maxDepth, C, &off, &rays);
            int fin = 1;
            for (int i = 0; i < NL; ++i) fin &= isfinite(C[i]) != 0;
            if (!fin) { if (nonfinite) ++*nonfinite; continue; }
            float lam[NL], bins[NL];
            for (int i = 0; i < NL; ++i) lam[i] = lambda_of(i, off);
            orc_film_bin(lam, C, bins);
            for (int i = 0; i < NL; ++i) film[(size_t)pixel * NL + i] += (double)bins[i];
        }
    }
    return rays;
}
