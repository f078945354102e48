// Host-side spectral helpers: RGB -> spectrum upsampling (Meng et al. 2015)
// and the colour-matching-function strata of the film.
//
// The upsampling follows UpsampledContinuousSpectrum
// (libSLR/BasicTypes/spectrum_types.cpp:408-543 for the construction and
// :611-651 for evaluation) step for step, in the same precision: the colour
// matrices and the xy -> uv map are evaluated in double and rounded to float
// like the reference's float instantiation does. The device does not walk the
// basis: mengToRegular() folds the (at most four) weighted basis spectra into
// one 95-sample table that the kernels interpolate.
#include "host_scene.h"

#include <algorithm>
#include <cmath>

#include "spectral_tables.inc"

namespace slrh {
namespace {

constexpr float kMinWL = 360.0f, kMaxWL = 830.0f;
constexpr uint32_t kNumWL = 95;
constexpr uint32_t kGridW = 12, kGridH = 14;
constexpr float kEqualEnergyReflectance = 0.009355121400914532f;

void computeAdjacents(float u, float v, MengSpectrum* m) {
    u = std::min(std::max(u, 0.0f), (float)kGridW);
    v = std::min(std::max(v, 0.0f), (float)kGridH);
    int32_t ui = (int32_t)u, vi = (int32_t)v;
    ui = std::min(ui, (int32_t)kGridW - 1);
    vi = std::min(vi, (int32_t)kGridH - 1);
    const tables::MengCell& cell = tables::kMengCells[ui + kGridW * vi];
    const uint8_t* idx = cell.idx;
    const uint8_t np = cell.numPoints;
    if (cell.inside) {
        m->adj = ((uint32_t)idx[3] << 24) | ((uint32_t)idx[2] << 16) | ((uint32_t)idx[1] << 8) | (uint32_t)idx[0];
        m->s = u - ui;
        m->t = v - vi;
        return;
    }
    const float ex = u - tables::kMengUV[idx[0]][0];
    const float ey = v - tables::kMengUV[idx[0]][1];
    float e0x = tables::kMengUV[idx[1]][0] - tables::kMengUV[idx[0]][0];
    float e0y = tables::kMengUV[idx[1]][1] - tables::kMengUV[idx[0]][1];
    float uu = e0x * ey - ex * e0y;
    for (int i = 1; i < np; ++i) {
        uint32_t k = idx[i % (np - 1) + 1];
        float e1x = tables::kMengUV[k][0] - tables::kMengUV[idx[0]][0];
        float e1y = tables::kMengUV[k][1] - tables::kMengUV[idx[0]][1];
        float vv = ex * e1y - e1x * ey;
        const float area = e0x * e1y - e1x * e0y;
        const float b0 = uu / area;
        const float b1 = vv / area;
        float b2 = 1.0f - b0 - b1;
        m->adj = (0xFFu << 24) | ((uint32_t)idx[0] << 16) | ((uint32_t)idx[i] << 8) | k;
        m->s = b0;
        m->t = b1;
        if (b0 < -1e-6 || b1 < -1e-6 || b2 < -1e-6) {
            uu = -vv;
            e0x = e1x;
            e0y = e1y;
            continue;
        }
        break;
    }
}

void weightsOf(const MengSpectrum& m, uint8_t idx[4], float w[4], int* n) {
    for (int k = 0; k < 4; ++k) idx[k] = (m.adj >> (8 * k)) & 0xFF;
    if (idx[3] != 0xFF) {
        w[0] = (1 - m.s) * (1 - m.t);
        w[1] = m.s * (1 - m.t);
        w[2] = (1 - m.s) * m.t;
        w[3] = m.s * m.t;
        *n = 4;
    } else {
        w[0] = m.s;
        w[1] = m.t;
        w[2] = 1.0f - m.s - m.t;
        w[3] = 0;
        *n = 3;
    }
}

}  // namespace

MengSpectrum mengFromRGB(int type, float r, float g, float b) {
    // linear sRGB -> XYZ: sRGB_E_to_XYZ for reflectances, sRGB_to_XYZ for
    // light sources (libSLR/BasicTypes/spectrum_base.h:61-79).
    float X, Y, Z;
    if (type == 0) {
        X = (float)(0.4969 * r + 0.3391 * g + 0.1640 * b);
        Y = (float)(0.2562 * r + 0.6782 * g + 0.0656 * b);
        Z = (float)(0.0233 * r + 0.1130 * g + 0.8637 * b);
    } else {
        X = (float)(0.4124564 * r + 0.3575761 * g + 0.1804375 * b);
        Y = (float)(0.2126729 * r + 0.7151522 * g + 0.0721750 * b);
        Z = (float)(0.0193339 * r + 0.1191920 * g + 0.9503041 * b);
    }
    MengSpectrum m;
    float brightness = X + Y + Z;
    if (brightness == 0) {
        m.scale = 0;
        computeAdjacents(6, 4, &m);
        return m;
    }
    float x = X / brightness;
    float y = Y / brightness;
    m.scale = brightness / kEqualEnergyReflectance;
    float u = (float)(16.730260708356887 * x + 7.7801960340706 * y - 2.170152247475828);
    float v = (float)(-7.530081094743006 * x + 16.192422314095225 * y + 1.1125529268825947);
    computeAdjacents(u, v, &m);
    return m;
}

float mengEvaluate(const MengSpectrum& m, float lambda) {
    uint8_t idx[4];
    float w[4];
    int n;
    weightsOf(m, idx, w, &n);
    float p = (lambda - kMinWL) / (kMaxWL - kMinWL);
    p = std::min(std::max(p, 0.0f), 1.0f);
    float binF = p * (kNumWL - 1);
    uint32_t bin = std::min((uint32_t)binF, kNumWL - 1);
    uint32_t next = std::min(bin + 1, kNumWL - 1);
    float t = binF - bin;
    float r = 0.0f;
    for (int j = 0; j < n; ++j) {
        const float* s = tables::kMengSpectra[idx[j]];
        r += w[j] * (s[bin] * (1 - t) + s[next] * t);
    }
    return r * m.scale;
}

HostSpectrum mengToRegular(const MengSpectrum& m) {
    uint8_t idx[4];
    float w[4];
    int n;
    weightsOf(m, idx, w, &n);
    HostSpectrum h;
    h.lo = kMinWL;
    h.hi = kMaxWL;
    h.values.resize(kNumWL);
    for (uint32_t k = 0; k < kNumWL; ++k) {
        double acc = 0;
        for (int j = 0; j < n; ++j) acc += (double)w[j] * tables::kMengSpectra[idx[j]][k];
        h.values[k] = (float)(acc * m.scale);
    }
    return h;
}

float truncToHalf(double x) {
    if (std::isnan(x)) return (float)x;
    const double ax = std::fabs(x);
    if (ax >= 65536.0) return std::copysign(INFINITY, (float)x);   // beyond the largest half: infinity
    if (ax == 0.0) return (float)x;
    int e;
    std::frexp(ax, &e);   // ax = m * 2^e, m in [0.5, 1)
    // half: 10 fraction bits for normals (exponent >= -14), fixed 2^-24 steps below
    const int shift = e - 1 >= -14 ? (e - 1) - 10 : -24;
    const double q = std::ldexp(std::floor(std::ldexp(ax, -shift)), shift);
    if (q >= 65520.0) return std::copysign(INFINITY, (float)x);
    return (float)std::copysign(q, x);
}

const float* mengBasisTable(uint32_t* numPoints, uint32_t* numLambdas) {
    *numPoints = tables::kMengNumPoints;
    *numLambdas = tables::kMengNumLambdas;
    return &tables::kMengSpectra[0][0];
}

namespace {
// RegularConstantContinuousDistribution1D's construction
// (libSLR/Core/distributions.cpp:172-191): densities over [0, 1) and a CDF
// built with a compensated sum, both normalised by the integral when positive.
float buildDist1D(const float* values, uint32_t n, float* pdf, float* cdf) {
    float sum = 0, comp = 0;
    cdf[0] = 0;
    for (uint32_t i = 0; i < n; ++i) {
        pdf[i] = values[i];
        const float in = values[i] / n - comp, tmp = sum + in;
        comp = (tmp - sum) - in;
        sum = tmp;
        cdf[i + 1] = sum;
    }
    if (sum > 0)
        for (uint32_t i = 0; i < n; ++i) {
            pdf[i] /= sum;
            cdf[i + 1] /= sum;
        }
    return sum;
}
}  // namespace

void buildEnvironment(const float* rgb, uint32_t w, uint32_t h, float scale, EnvData* env) {
    constexpr double kScaleFactor = 0.009355121400914532;   // UPSAMPLED_CONTINOUS_SPECTRUM_SCALE_FACTOR
    env->present = true;
    env->width = w;
    env->height = h;
    env->scale = scale;
    env->texels.resize((size_t)w * h);
    std::vector<float> lum((size_t)w * h);
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        // RGBA16F input, clamped at zero, to uvs (sRGB_to_uvs for a light
        // source, spectrum_types.h:219-242), stored as half
        float c[3];
        for (int k = 0; k < 3; ++k) c[k] = std::max(truncToHalf(rgb[3 * i + k]), 0.0f);
        const float X = (float)(0.4124564 * c[0] + 0.3575761 * c[1] + 0.1804375 * c[2]);
        const float Y = (float)(0.2126729 * c[0] + 0.7151522 * c[1] + 0.0721750 * c[2]);
        const float Z = (float)(0.0193339 * c[0] + 0.1191920 * c[1] + 0.9503041 * c[2]);
        const float b = X + Y + Z;
        float x = X / b, y = Y / b;
        if (b == 0) x = y = (float)(1.0f / 3.0);
        const float u = (float)(16.730260708356887 * x + 7.7801960340706 * y - 2.170152247475828);
        const float v = (float)(-7.530081094743006 * x + 16.192422314095225 * y + 1.1125529268825947);
        const float sv = (float)(b / kEqualEnergyReflectance);
        const float uh = truncToHalf(u), vh = truncToHalf(v), sh = truncToHalf((double)sv * kScaleFactor);
        // ImageSpectrumTexture::evaluate: UpsampledContinuousSpectrum(u, v, s / scale factor)
        MengSpectrum& m = env->texels[i];
        computeAdjacents(uh, vh, &m);
        m.scale = (float)((double)sh / kScaleFactor);
        // luminance Y of the stored texel (uvs_to_luminance, spectrum_types.h:267-272)
        const float yy = (float)(0.022853819546830627 * uh + 0.05077639329371236 * vh - 0.006895157122499944);
        lum[i] = yy * (float)(m.scale * kEqualEnergyReflectance);
    }
    // importance map at a quarter of the resolution (as createIBLImportanceMap,
    // image_textures.cpp:142-199): per cell the mean texel luminance times the
    // sine of the cell's polar angle
    const uint32_t mw = std::max(1u, w / 4), mh = std::max(1u, h / 4);
    env->mapW = mw;
    env->mapH = mh;
    std::vector<float> cell((size_t)mw * mh);
    for (uint32_t cy = 0; cy < mh; ++cy)
        for (uint32_t cx = 0; cx < mw; ++cx) {
            const uint32_t x0 = (uint32_t)((uint64_t)cx * w / mw), x1 = (uint32_t)((uint64_t)(cx + 1) * w / mw);
            const uint32_t y0 = (uint32_t)((uint64_t)cy * h / mh), y1 = (uint32_t)((uint64_t)(cy + 1) * h / mh);
            double acc = 0;
            for (uint32_t yy = y0; yy < y1; ++yy)
                for (uint32_t xx = x0; xx < x1; ++xx) acc += lum[(size_t)yy * w + xx];
            const double mean = acc / ((double)(x1 - x0) * (y1 - y0));
            cell[(size_t)cy * mw + cx] = (float)(std::sin(M_PI * (cy + 0.5f) / mh) * mean);
        }
    // RegularConstantContinuousDistribution2D (distributions.cpp:242-262):
    // one distribution per row over phi, and one over the rows' integrals
    env->condPdf.resize((size_t)mw * mh);
    env->condCdf.resize((size_t)(mw + 1) * mh);
    env->rowIntegral.resize(mh);
    for (uint32_t cy = 0; cy < mh; ++cy)
        env->rowIntegral[cy] = buildDist1D(&cell[(size_t)cy * mw], mw, &env->condPdf[(size_t)cy * mw],
                                           &env->condCdf[(size_t)cy * (mw + 1)]);
    env->margPdf.resize(mh);
    env->margCdf.resize(mh + 1);
    buildDist1D(env->rowIntegral.data(), mh, env->margPdf.data(), env->margCdf.data());
}

// DiscretizedSpectrum::initialize (libSLR/BasicTypes/spectrum_types.cpp:945-990):
// trapezoidal integration of the 1 nm CMF samples into 16 equal strata.
void cmfStrata(float xbar[16], float ybar[16], float zbar[16], float* integral) {
    const int N = tables::kNumCMFSamples;
    const float* X = tables::kCIE1931_X;
    const float* Y = tables::kCIE1931_Y;
    const float* Z = tables::kCIE1931_Z;
    uint32_t bin = 0;
    float nextP = float(bin + 1) / 16;
    float xs = 0, xp = X[0], ys = 0, yp = Y[0], zs = 0, zp = Z[0];
    const float interval = 1;
    for (int i = 1; i < N; ++i) {
        float curP = float(i) / (N - 1);
        float width = interval;
        float xc = X[i], yc = Y[i], zc = Z[i];
        if (curP >= nextP) {
            width = (curP - nextP) * (kLambdaHigh - kLambdaLow);
            float t = 1 - width / interval;
            float xi = xp * (1 - t) + xc * t;
            float yi = yp * (1 - t) + yc * t;
            float zi = zp * (1 - t) + zc * t;
            xs += (xp + xi) * (interval - width) * 0.5f;
            ys += (yp + yi) * (interval - width) * 0.5f;
            zs += (zp + zi) * (interval - width) * 0.5f;
            if (bin < 16) { xbar[bin] = xs; ybar[bin] = ys; zbar[bin] = zs; }
            xs = ys = zs = 0;
            xp = xi; yp = yi; zp = zi;
            ++bin;
            nextP = float(bin + 1) / 16;
        }
        xs += (xp + xc) * width * 0.5f;
        ys += (yp + yc) * width * 0.5f;
        zs += (zp + zc) * width * 0.5f;
        xp = xc; yp = yc; zp = zc;
    }
    float integ = 0.0f;
    for (int i = 0; i < 16; ++i) integ += ybar[i];
    *integral = integ;
}

// Supervoxel majorants. Cell c of an axis covers lattice intervals
// [kSuperVoxel c, kSuperVoxel (c + 1)] (clipped to the lattice); trilinear
// interpolation anywhere in it is a convex combination of the lattice points
// in that range, so their maximum bounds the density there.
void buildMajorants(MediumData* m) {
    for (int a = 0; a < 3; ++a) m->cells[a] = std::max(1u, (m->n[a] - 1 + kSuperVoxel - 1) / kSuperVoxel);
    m->majorant.assign((size_t)m->cells[0] * m->cells[1] * m->cells[2], 0.0f);
    const uint32_t nx = m->n[0], ny = m->n[1], nz = m->n[2];
    for (uint32_t cz = 0; cz < m->cells[2]; ++cz)
        for (uint32_t cy = 0; cy < m->cells[1]; ++cy)
            for (uint32_t cx = 0; cx < m->cells[0]; ++cx) {
                float mx = 0.0f;
                for (uint32_t z = cz * kSuperVoxel; z <= std::min(nz - 1, (cz + 1) * kSuperVoxel); ++z)
                    for (uint32_t y = cy * kSuperVoxel; y <= std::min(ny - 1, (cy + 1) * kSuperVoxel); ++y)
                        for (uint32_t x = cx * kSuperVoxel; x <= std::min(nx - 1, (cx + 1) * kSuperVoxel); ++x)
                            mx = std::max(mx, m->density[((size_t)z * ny + y) * nx + x]);
                m->majorant[((size_t)cz * m->cells[1] + cy) * m->cells[0] + cx] = mx;
            }
}

}  // namespace slrh
