// Host-side scene description, accelerator build and spectral helpers shared
// by the C ABI (capi.cpp), the builder (bvh_build.cpp) and the device
// orchestration (render.hip).
#pragma once

#include <cstdint>
#include <cmath>
#include <string>
#include <vector>

namespace slrh {

constexpr int kNumLambdas = 16;            // NumSpectralSamples (libSLR/declarations.h:50)
constexpr float kLambdaLow = 360.0f;       // WavelengthLowBound (libSLR/BasicTypes/spectrum_base.h:110)
constexpr float kLambdaHigh = 830.0f;      // WavelengthHighBound (spectrum_base.h:111)

struct V3 {
    float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }

// Axis-aligned box with the reference's conventions: empty = (+inf, -inf),
// union/intersection through fmin/fmax (libSLR/BasicTypes/BoundingBox3D.h:25-163).
struct Box {
    float mn[3], mx[3];
    Box() { mn[0] = mn[1] = mn[2] = INFINITY; mx[0] = mx[1] = mx[2] = -INFINITY; }
    static Box point(const V3& p) { Box b; b.mn[0] = b.mx[0] = p.x; b.mn[1] = b.mx[1] = p.y; b.mn[2] = b.mx[2] = p.z; return b; }
    void unify(const Box& o) {
        for (int a = 0; a < 3; ++a) { mn[a] = std::fmin(mn[a], o.mn[a]); mx[a] = std::fmax(mx[a], o.mx[a]); }
    }
    void unify(const V3& p) {
        const float q[3] = {p.x, p.y, p.z};
        for (int a = 0; a < 3; ++a) { mn[a] = std::fmin(mn[a], q[a]); mx[a] = std::fmax(mx[a], q[a]); }
    }
    float area() const {
        float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
        return 2 * (dx * dy + dy * dz + dz * dx);
    }
    bool valid() const { return (mx[0] - mn[0]) >= 0 && (mx[1] - mn[1]) >= 0 && (mx[2] - mn[2]) >= 0; }
    int widestAxis() const {
        float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
        if (dx > dy && dx > dz) return 0;
        if (dy > dz) return 1;
        return 2;
    }
    float center(int a) const { return (mn[a] + mx[a]) * 0.5f; }
    V3 centroid() const { return v3((mn[0] + mx[0]) * 0.5f, (mn[1] + mx[1]) * 0.5f, (mn[2] + mx[2]) * 0.5f); }
};
inline Box intersect(const Box& a, const Box& b) {
    Box r;
    for (int k = 0; k < 3; ++k) { r.mn[k] = std::fmax(a.mn[k], b.mn[k]); r.mx[k] = std::fmin(a.mx[k], b.mx[k]); }
    return r;
}

// Binary SBVH node in the reference's DFS pre-order layout (SBVH.h:21-43).
struct SbvhNode {
    Box box;
    uint32_t c0 = 0, c1 = 0, first = 0, count = 0;
    uint8_t axis = 0;
};

// Collapsed 4-wide node (QBVH.h:41-53), kept on the host in the reference's
// logical form; the device copy is packed by the renderer.
struct QbvhChild {
    // invalid: all bits set. Otherwise leaf/internal with an index and count.
    uint32_t raw = 0xFFFFFFFFu;   // the reference's bitfield {idx:27, numLeaves:4, isLeaf:1}
    bool valid() const { return raw != 0xFFFFFFFFu; }
};
struct QbvhNode {
    float bmin[3][4], bmax[3][4];
    bool isLeaf[4];
    uint32_t index[4];   // node index (internal) or first leaf ref (leaf)
    uint32_t count[4];   // leaf prim count (0 for internal)
    bool childValid[4];
    uint8_t topAxis = 0, leftAxis = 0, rightAxis = 0;
};

struct Accel {
    std::vector<SbvhNode> sbvh;
    std::vector<uint32_t> sbvhRefs;   // leaf refs (prim ids) in DFS order
    uint32_t sbvhDepth = 0;
    float sbvhCost = 0;
    uint32_t numFragments = 0;
    std::vector<QbvhNode> qbvh;
    std::vector<uint32_t> qbvhRefs;
    uint32_t qbvhDepth = 0;
    float qbvhCost = 0;
    Box bounds;
    double buildSeconds = 0;
    uint32_t maxLeafCount = 0;
    uint32_t deviceNodes = 0;   // SBVH nodes binned and partitioned on the device
};

// Binary BVH node of the binned-SAH builder (StandardBVH::Node,
// libSLR/Accelerator/StandardBVH.h:24-46): leaf when count > 0.
struct StdBvhNode {
    Box box;
    uint32_t c0 = 0, c1 = 0, first = 0, count = 0;
    uint8_t axis = 0;
};
struct StdBvh {
    std::vector<StdBvhNode> nodes;   // creation (pre-)order, root first
    std::vector<uint32_t> refs;      // leaf triangle ids in leaf order
    uint32_t depth = 0;
    float cost = 0;
};
// StandardBVH(objs, Partitioning::BinnedSAH) over triangles (9 floats each).
void buildStandardBvh(const std::vector<float>& triPos, StdBvh* out);

// Builds SBVH then QBVH over triangles given as 9 floats each (world space).
// Deterministic; identical decisions to the reference builder.
// With device set (a device is current on the calling thread), the nodes
// at the top of a large triangle tree are binned and partitioned there; the
// tree is the same.
void buildAccel(const std::vector<float>& triPos, Accel* out, int numThreads, bool device = false);
// The same over triangles followed by boxes (instances): primitive ids
// 0..ntris-1 are the triangles, ntris + i is boxes[i], whose intersection
// cost in the SAH is boxCosts[i] (a triangle's is 1).
void buildAccel(const std::vector<float>& triPos, const std::vector<Box>& boxes, const std::vector<float>& boxCosts,
                Accel* out, int numThreads, bool device = false);

// ---------------------------------------------------------------- spectra

// A spectrum as the device evaluates it: regular samples over [lo, hi].
struct HostSpectrum {
    float lo = kLambdaLow, hi = kLambdaHigh;
    std::vector<float> values;  // n >= 1
};
// Meng upsampling in the reference's representation: up to four data-point
// indices with bilinear (inside) or barycentric (fan) weights and a scale.
struct MengSpectrum {
    uint32_t adj = 0;     // packed indices, byte 3 == 0xFF for three points
    float s = 0, t = 0, scale = 0;
};
MengSpectrum mengFromRGB(int type, float r, float g, float b);
// Evaluates exactly as UpsampledContinuousSpectrum::evaluate does.
float mengEvaluate(const MengSpectrum& m, float lambda);
// Combines the basis into a 95-sample table for the device.
HostSpectrum mengToRegular(const MengSpectrum& m);
// CMF strata for the film (DiscretizedSpectrum::initialize).
void cmfStrata(float xbar[16], float ybar[16], float zbar[16], float* integral);

// Environment light (image-based): the texels as the reference stores an
// RGBA16F image for a light source (u, v and scaled s truncated to half
// precision, libSLR/Core/image_2d.h:270-289), each decoded into its Meng
// basis weights, and a piecewise-constant importance map over (phi, theta).
struct EnvData {
    bool present = false;
    uint32_t width = 0, height = 0;
    float scale = 1.0f;
    std::vector<MengSpectrum> texels;   // row-major, row = theta, column = phi
    uint32_t mapW = 0, mapH = 0;
    std::vector<float> condPdf, condCdf;   // per map row: mapW densities, mapW + 1 CDF values
    std::vector<float> margPdf, margCdf;   // over rows: mapH densities, mapH + 1 CDF values
    std::vector<float> rowIntegral;
};
// rgb: width * height linear sRGB triples (row 0 = theta 0, +y up).
void buildEnvironment(const float* rgb, uint32_t width, uint32_t height, float scale, EnvData* env);
// A heterogeneous medium on a density lattice (DensityGridMediumDistribution,
// libSLR/MediumDistribution/DensityGridMediumDistribution.h:16-60): region,
// nx*ny*nz densities (x fastest, then y, then z), base scattering and
// extinction spectra, Henyey-Greenstein asymmetry. Majorants: the maximum
// density over each supervoxel of kSuperVoxel^3 lattice cells (its
// (kSuperVoxel+1)^3 lattice points bound the trilinear density inside it).
constexpr uint32_t kSuperVoxel = 8;
struct MediumData {
    bool present = false;
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
    uint32_t n[3] = {0, 0, 0};
    std::vector<float> density;
    int32_t sigmaS = -1;      // spectrum ids
    int32_t sigmaE = -1;
    float sigmaEValue = 0;    // the extinction spectrum's (constant) value
    float g = 0;
    uint32_t cells[3] = {0, 0, 0};
    std::vector<float> majorant;   // cells[0]*cells[1]*cells[2] maximum densities
};
void buildMajorants(MediumData* m);

// Value of a float truncated toward zero to IEEE half precision (overflow to
// infinity), the conversion of the reference's half type (External/half.hpp
// 1.12 with its default rounding style).
float truncToHalf(double x);
// The Meng basis table and grid constants the device evaluates texels with.
const float* mengBasisTable(uint32_t* numPoints, uint32_t* numLambdas);

}  // namespace slrh
