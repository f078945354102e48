// Reader for a subset of SLR's scene language (libSLRSceneGraph/Parser/
// SceneParser.yy, libSLRSceneGraph/API.cpp:295-1325): enough to turn a scene
// file such as TestScenes/Original_Cornell_Box.txt into the spectra,
// materials, world-space meshes, camera and render settings the C ABI takes.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace slrh {
namespace scenefile {

enum SpectrumKind { kRgb = 0, kRegular = 1 };
enum SpectrumTypeId { kReflectance = 0, kLightSource = 1, kIndexOfRefraction = 2 };

struct Spectrum {
    int kind = kRgb;
    int type = kReflectance;
    float rgb[3] = {0, 0, 0};   // kRgb: linear sRGB
    float lo = 0, hi = 0;       // kRegular: samples over [lo, hi]
    std::vector<float> values;
};

enum MaterialKind { kMatte = 0, kMicrofacetMetal = 1, kMicrofacetGlass = 2, kEmitter = 3 };

struct Material {
    int kind = kMatte;
    int spec0 = -1, spec1 = -1;   // matte: reflectance; metal: eta, k; glass: etaExt, etaInt
    bool hasSigma = false;        // matte with a sigma texture: Oren-Nayar
    float sigma = 0, alphaX = 0, alphaY = 0;
    int base = -1, emittance = -1;   // emitter: scattering material, emittance spectrum
};

// A mesh in world space: the reference's recomputed vertex normals and
// tangents (TriangleMeshNode::recalculateNormalAndTangents), then the node
// chain's transform (libSLR TriangleMeshNode::createRenderingData).
struct Mesh {
    std::vector<float> pos, nrm, tan;   // 3 per vertex
    std::vector<uint32_t> idx, mat;     // 3 per triangle, 1 per triangle
};

struct Camera {
    bool present = false;
    float l2w[16] = {};   // column-major
    float sensitivity = 0, aspect = 1, fovY = 0.5235987756f, radius = 0, imgDist = 0.02f, objDist = 5.0f;
};

struct SceneFile {
    std::vector<Spectrum> spectra;
    std::vector<Material> materials;
    std::vector<Mesh> meshes;
    Camera camera;
    std::string method = "PT";
    uint32_t spp = 8, width = 1024, height = 1024, rngSeed = 1509761209u;
    float brightness = 1.0f;
    bool hasRenderer = false;
};

// Runs the scene text; false with a message (and its line) on an error or
// on a construct outside the subset.
bool readSceneText(const std::string& text, SceneFile* out, std::string* err);

}  // namespace scenefile

// Shared with the C ABI (capi.cpp): the reference's Matrix4x4 inverse and
// makeCoordinateSystem.
bool invertMatrix4(const float m[4][4], float out[4][4]);
void coordinateTangent(const float n[3], float t[3]);

}  // namespace slrh
