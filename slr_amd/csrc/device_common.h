// Device-side data layout and math shared by the kernels.
//
// Layout in HBM (all arrays 16-byte aligned, structure-of-arrays where lanes
// of a wave touch consecutive elements):
//   QBVH node   : 8 x float4 = 128 B, one cache line. Rows 0-5 hold the four
//                 children's bounds per axis (min x, max x, min y, max y,
//                 min z, max z: the SoA-in-node form of
//                 libSLR/Accelerator/QBVH.h:41-53), row 6 the four child
//                 words, row 7 is unused. Nodes are stored breadth-first for
//                 the first levels (the part the traversal caches in LDS),
//                 then in build order.
//   child word  : bits 29-30 of child words 0, 1, 2 hold the node's top,
//                 left and right split axes. Bit 31 set = leaf with count in
//                 bits 24-28 and first triangle reference in bits 0-23 (count
//                 0: no child); otherwise an internal node index (bits 0-24).
//   triangle ref: 3 x float4 = 48 B {v0.xyz, prim id}, {v1-v0, class},
//                 {v2-v0, 0}; one per leaf reference, in leaf order, so a leaf
//                 is contiguous. The edges are the same FP32 differences the
//                 reference forms at each test (TriangleSurfaceShape.cpp:212-213).
//   instance ref: the same 48 B with w words {0x80000000 | instance id},
//                 {root node of the prototype's tree}, {class override}.
//                 With instances the node and reference arrays hold the
//                 world tree, then each prototype's tree.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace slrh {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kNoChild = kLeafBit;          // an empty leaf
constexpr uint32_t kAxisShift = 29;
constexpr uint32_t kNodeIndexMask = 0x01FFFFFFu;
constexpr uint32_t kLeafFirstMask = 0x00FFFFFFu;
constexpr uint32_t kLeafCountShift = 24;
constexpr uint32_t kLeafCountMask = 0x1Fu;
constexpr uint32_t kMaxLeafCount = 31;
constexpr uint32_t kMaxTriRefs = 1u << 24;
constexpr uint32_t kMaxNodes = 1u << 25;
constexpr uint32_t kTreeletRows = 7;             // node rows kept in the LDS copy (planes + child words)
constexpr int kMaxStack = 64;   // QBVH::commonProcedure's StackSize (QBVH.h:299)
constexpr int kNumL = 16;       // wavelengths per path
constexpr float kLambdaLowD = 360.0f;    // WavelengthLowBound
constexpr float kLambdaHighD = 830.0f;   // WavelengthHighBound

enum MaterialType : uint32_t {
    MAT_LAMBERT = 0,
    MAT_OREN_NAYAR = 1,
    MAT_GGX_CONDUCTOR = 2,
    MAT_GGX_DIELECTRIC = 3,
    MAT_NONE = 4,   // emitter without a scattering base
};
// Shading queues: one per scattering model plus misses.
enum ShadeClass : uint32_t {
    CLS_DIFFUSE = 0,     // Lambert and Oren-Nayar share a queue (same sampling)
    CLS_CONDUCTOR = 1,
    CLS_DIELECTRIC = 2,
    CLS_MISS = 3,        // no hit / emitter without base
    CLS_MEDIUM = 4,      // scattering in the medium (volumetric integrator)
    kNumClasses = 5,
};

struct DevMaterial {
    uint32_t type;
    int32_t spec0, spec1;   // reflectance | eta, k | etaExt, etaInt
    int32_t emit;           // emittance spectrum, -1 if not an emitter
    float p0, p1;           // OrenNayar A, B | GGX alpha_x, alpha_y
    uint32_t cls;
    uint32_t pad;
};

struct DevSpectrum {
    float lo, hi;
    uint32_t offset, n;
};
// LDS words a shading block gives to the spectrum pool and material headers
constexpr uint32_t kShadeLdsWords = 4096;

struct DevCamera {
    float l2w[12];        // rows of the 3x4 camera-to-world matrix
    float nrm[9];         // rows of the inverse-transpose (for normals)
    float opWidth, opHeight, objDist, imgDist, imgPlaneArea, lensRadius;
    uint32_t width, height;
};

constexpr uint32_t kNoInstance = 0xFFFFFFFFu;
constexpr uint32_t kLightRecRows = 8;
constexpr uint32_t kTriShadeRows = 8;   // per-triangle shading record: one 128-B line
constexpr uint32_t kInstXfRows = 8;

struct DevScene {
    const float4* nodes;
    const float4* tris;
    const uint4* primVerts;   // i0, i1, i2, material
    const float4* vpos;
    const float4* vnrm;
    const float4* vtan;
    const DevMaterial* mats;
    const DevSpectrum* spectra;
    const float* specPool;
    const uint2* lights;      // emitting triangles: global id, instance (kNoInstance: world space)
    const float4* lightRec;   // per light kLightRecRows float4: world vertices (material, instance and area
                              // PDF in w), geometric normal, the three object-space vertex normals, the
                              // header of the light material's emittance spectrum (DevSpectrum bits)
    const float4* triShade;   // per triangle kTriShadeRows float4: the three vertex normals (material id
                              // bits in n0.w), the three vertex tangents, the geometric normal (its own
                              // space: world, or object space for prototype triangles), unused
    const DevSpectrum* matSpec;   // per material 3 headers: spec0, spec1, emittance (n = 0: none)
    const float* lightCdf;    // numLights + 1 entries
    const float* lightPmf;
    const int32_t* primToLight;
    uint32_t numLights;
    uint32_t numPrims;
    float lightImportance;    // SurfaceObjectAggregate::importance()
    uint32_t numNodes;
    // the shading kernels' LDS copy of the spectrum pool followed by the
    // material spectrum headers: specLdsPool words of pool (a multiple of 4),
    // specLdsWords in all; 0 when they do not fit in kShadeLdsWords
    uint32_t specLdsPool, specLdsWords;
    uint32_t classMask;   // bit c: some material shades with class c (misses always)
    // instances: kInstXfRows float4 each: 3 rows object-to-world (affine,
    // row-major, translation in w), then the 4 rows of world-to-object
    const float4* instXf;
    const int32_t* instMat;      // material replacing the prototype's, -1: none
    uint32_t numInstances;
    uint32_t compatInstShadow;   // shadow rays test instances with the world ray
    uint32_t compatTangentNee;   // the scene asks for the reference's volumetric tangent-light NaN
    uint32_t tangentNeeNaN;      // set per render: compatTangentNee with the volumetric integrator
    DevCamera cam;
    // environment light (image-based); hasEnv = 0 without one
    const float4* envTexel;   // per texel: Meng basis indices (bits), s, t, scale
    const float2* mengBasis;  // [kMengPoints][95]: each bin with the next (the last with itself)
    const float *envCondPdf, *envCondCdf, *envMargPdf, *envMargCdf;
    uint32_t envW, envH, envMapW, envMapH;
    float envScale;
    uint32_t hasEnv;
    // heterogeneous medium (volumetric integrator)
    // lattice densities corner-packed: for lattice point (x, y, z), x fastest,
    // two float4 holding the densities at the corners (x|x+1, y|y+1, z|z+1),
    // each index clamped to the lattice, in the order z, y, x (x fastest), so
    // a trilinear lookup is one 32-byte read
    const float4* medCorners;
    const float* medMajorant;   // maximum density per supervoxel (kMedSuperVoxel^3 lattice cells)
    float medMin[3], medMax[3]; // region
    float medScale[3];          // (n - 1) / (max - min): lattice units per length
    float medInvExt[3];         // 1 / (max - min)
    uint32_t medN[3], medCells[3];
    float medSigmaE;            // extinction per unit density (constant over wavelength)
    int32_t medSigmaS;          // scattering per unit density (spectrum id)
    float medG;                 // Henyey-Greenstein asymmetry
    uint32_t hasMedium;
    uint32_t volume;            // set per render: the volumetric integrator with a medium
};
constexpr uint32_t kMedSuperVoxel = 8;

__device__ __forceinline__ float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float3 operator+(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 operator-(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 operator-(float3 a) { return f3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float3 operator*(float3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 operator*(float s, float3 a) { return f3(s * a.x, s * a.y, s * a.z); }
// Vector3D::operator/(s) multiplies by the reciprocal (libSLR/BasicTypes/Vector3D.h:31).
__device__ __forceinline__ float3 divs(float3 a, float s) { float r = 1.0f / s; return f3(a.x * r, a.y * r, a.z * r); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
    return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float length3(float3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ float3 normalize3(float3 a) { return divs(a, length3(a)); }
__device__ __forceinline__ float3 xyz(float4 a) { return f3(a.x, a.y, a.z); }

}  // namespace slrh
