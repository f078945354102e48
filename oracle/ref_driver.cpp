// Test harness around the reference renderer (libSLR compiled from its own
// sources by oracle/ref.mk). It is part of the oracle: only tests, the smoke
// check and the CPU-baseline leg of the benchmark load it, and only as a
// checker or a timed baseline, never as the product path.
//
// It builds scenes directly from libSLR classes (the scene-language layer of
// the reference does not compile here) and exposes a small C interface:
//   - ray batches through an explicitly constructed SBVH -> QBVH
//     (libSLR/Accelerator/QBVH.h:252-376; the reference aggregate would build
//     a plain SBVH, libSLR/Core/surface_object.cpp:376),
//   - PTRenderer::render (libSLR/Renderer/PTRenderer.cpp:26-101) with the
//     resulting 16-stratum film,
//   - known-answer probes: XORShift streams, wavelength sampling, Meng
//     upsampling, CMF strata and BSDF sample/evaluate/PDF tuples.
#include <libSLR/defines.h>
#include <libSLR/declarations.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <vector>
#include <map>
#include <memory>
#include <chrono>
#include <unistd.h>
#include <fcntl.h>
#include <sys/stat.h>
#include <string>
#include <functional>
#include <algorithm>
#include <thread>
#include <mutex>
#include <condition_variable>
#include <atomic>
#include <limits>
#include <cmath>
#include <queue>
#include <set>
#include <unordered_map>
#include <nmmintrin.h>
// The accelerators keep their node arrays private; the harness reads them to
// report topology statistics. Everything the standard library provides is
// included above, so only libSLR's own classes see the change.
#define private public
#define protected public

#include <libSLR/BasicTypes/spectrum_base.h>
#include <libSLR/BasicTypes/spectrum_types.h>
#include <libSLR/BasicTypes/rgb_types.h>
#include <libSLR/BasicTypes/TexCoord2D.h>
#include <libSLR/Core/transform.h>
// SurfaceInteraction keeps its hit triangle in a member declared before any
// access label; spelling its classes as structs (same layout, same mangling)
// makes it readable here. Its dependencies are included above, so only
// geometry.h's own declarations see the change.
#define class struct
#include <libSLR/Core/geometry.h>
#undef class
#include <libSLR/MemoryAllocators/Allocator.h>
#include <libSLR/MemoryAllocators/ArenaAllocator.h>
#include <libSLR/Core/surface_object.h>
#include <libSLR/Core/light_path_sampler.h>
#include <libSLR/Core/RenderSettings.h>
#include <libSLR/Core/ImageSensor.h>
#include <libSLR/Core/camera.h>
#include <libSLR/Core/transform.h>
#include <libSLR/Core/surface_material.h>
#include <libSLR/Core/directional_distribution_functions.h>
#include <libSLR/Core/distributions.h>
#include <libSLR/Scene/Scene.h>
#include <libSLR/Scene/node.h>
#include <libSLR/Scene/TriangleMeshNode.h>
#include <libSLR/Scene/camera_nodes.h>
#include <libSLR/Texture/constant_textures.h>
#include <libSLR/Texture/image_textures.h>
#include <libSLR/Core/image_2d.h>
#include <libSLR/Core/textures.h>
#include <libSLR/SurfaceMaterial/IBLEmitterSurfaceProperty.h>
#include <libSLR/SurfaceMaterial/basic_surface_materials.h>
#include <libSLR/SurfaceMaterial/microfacet_surface_materials.h>
#include <libSLR/SurfaceMaterial/basic_emitter_surface_properties.h>
#include <libSLR/BSDF/basic_bsdfs.h>
#include <libSLR/BSDF/OrenNayerBRDF.h>
#include <libSLR/BSDF/microfacet_bsdfs.h>
#include <libSLR/Renderer/BPTRenderer.h>
#include <libSLR/Renderer/PTRenderer.h>
#include <libSLR/Renderer/VolumetricPTRenderer.h>
#include <libSLR/Scene/medium_nodes.h>
#include <libSLR/MediumDistribution/DensityGridMediumDistribution.h>
#include <libSLR/MediumMaterial/basic_medium_materials.h>
#include <libSLR/RNG/XORShiftRNG.h>
// SBVH and QBVH are declared with the class keyword and no access label, so
// their members default to private; spelling them as structs (same layout,
// same mangling) makes the node arrays readable here.
#define class struct
#include <libSLR/Accelerator/SBVH.h>
#include <libSLR/Accelerator/QBVH.h>
#include <libSLR/Accelerator/StandardBVH.h>
#undef class
#undef private
#undef protected

using namespace SLR;

// libSLR/BasicTypes/spectrum_types.cpp:922-925 (and the double copy below it)
// writes the CMF-strata statics as explicit specialisations without an
// initializer. ISO C++ reads those as declarations only, so g++ leaves the
// symbols undefined; these are the zero-initialised definitions the
// reference's own compilers produced from the same lines.
namespace SLR {
template <> std::unique_ptr<float[]> DiscretizedSpectrumTemplate<float, NumStrataForStorage>::xbar{};
template <> std::unique_ptr<float[]> DiscretizedSpectrumTemplate<float, NumStrataForStorage>::ybar{};
template <> std::unique_ptr<float[]> DiscretizedSpectrumTemplate<float, NumStrataForStorage>::zbar{};
template <> float DiscretizedSpectrumTemplate<float, NumStrataForStorage>::integralCMF{};
template <> std::unique_ptr<double[]> DiscretizedSpectrumTemplate<double, NumStrataForStorage>::xbar{};
template <> std::unique_ptr<double[]> DiscretizedSpectrumTemplate<double, NumStrataForStorage>::ybar{};
template <> std::unique_ptr<double[]> DiscretizedSpectrumTemplate<double, NumStrataForStorage>::zbar{};
template <> double DiscretizedSpectrumTemplate<double, NumStrataForStorage>::integralCMF{};
}  // namespace SLR

namespace {

struct RefScene {
    std::vector<std::unique_ptr<AssetSpectrum>> spectra;
    std::vector<std::unique_ptr<SpectrumTexture>> spTextures;
    std::vector<std::unique_ptr<FloatTexture>> flTextures;
    std::vector<std::unique_ptr<SVMicrofacetDistribution>> dists;
    std::vector<std::unique_ptr<EmitterSurfaceProperty>> emitters;
    std::vector<std::unique_ptr<SurfaceMaterial>> materials;
    std::vector<TriangleMeshNode*> meshes;
    std::vector<uint32_t> meshPrimBase;
    uint32_t numPrims = 0;
    // prototypes (meshes placed only through instances) and instances
    // (prototype, object-to-world rows)
    std::vector<TriangleMeshNode*> protos;
    std::vector<uint32_t> protoPrimBase;
    struct Inst { int proto; float rows[12]; };
    std::vector<Inst> insts;
    std::vector<std::unique_ptr<SurfaceObjectAggregate>> protoAggs;
    std::vector<std::unique_ptr<StaticTransform>> instTFs;
    std::vector<std::unique_ptr<TransformedSurfaceObject>> instObjs;
    std::map<const SurfaceObject*, int32_t> objToInst;

    // camera
    bool hasCamera = false;
    float l2w[16];
    float sensitivity = 1, aspect = 1, fovY = 0.5f, lensRadius = 0, imgDist = 1, objDist = 1;

    // environment: an RGBA16F equirectangular image stored as a light-source
    // uvs texture (image_2d.h:270-289) behind the default texture mapping
    TiledImage2D* envImage = nullptr;
    Texture2DMapping envMapping;
    std::unique_ptr<ImageSpectrumTexture> envTex;
    float envScale = 1.0f;

    // heterogeneous medium (DensityGridMediumNode) and the integrator
    bool hasMedium = false;
    float medBounds[6];
    uint32_t medN[3];
    std::vector<std::vector<float>> medGrid;   // per z slice, x fastest
    float medSigmaS = 0, medSigmaE = 0, medG = 0;
    int integrator = 0;   // 0 PTRenderer, 1 VolumetricPTRenderer, 2 BPTRenderer

    // accelerator for ray batches
    bool built = false;
    std::vector<SurfaceObject*> objs;
    std::map<const SurfaceObject*, int32_t> objToPrim;
    SBVH* sbvh = nullptr;
    QBVH* qbvh = nullptr;
    StandardBVH* stdbvh = nullptr;   // StandardBVH(objs, BinnedSAH), built beside the SBVH

    ~RefScene() {
        envTex.reset();
        if (envImage) DefaultAllocator::instance().destroy(envImage);
        delete qbvh;
        delete sbvh;
        instObjs.clear();
        protoAggs.clear();
        if (built) {
            for (auto m : meshes) m->destroyRenderingData(&DefaultAllocator::instance());
            for (auto m : protos) m->destroyRenderingData(&DefaultAllocator::instance());
        }
        for (auto m : meshes) delete m;
        for (auto m : protos) delete m;
    }
    // Matrix4x4 (column-major array) of instance k's object-to-world rows.
    Matrix4x4 instMatrix(size_t k) const {
        const float* r = insts[k].rows;
        float m[16] = {r[0], r[4], r[8], 0, r[1], r[5], r[9], 0, r[2], r[6], r[10], 0, r[3], r[7], r[11], 1};
        return Matrix4x4(m);
    }

    const SpectrumTexture* constTexRGB(SpectrumType type, float r, float g, float b) {
        spectra.emplace_back(new UpsampledContinuousSpectrum(type, ColorSpace::sRGB, r, g, b));
        spTextures.emplace_back(new ConstantSpectrumTexture(spectra.back().get()));
        return spTextures.back().get();
    }
    const SpectrumTexture* constTexValue(float v) {
        spectra.emplace_back(new RegularContinuousSpectrum(WavelengthLowBound, WavelengthHighBound, &v, 1));
        spTextures.emplace_back(new ConstantSpectrumTexture(spectra.back().get()));
        return spTextures.back().get();
    }
    const FloatTexture* constFloat(float v) {
        flTextures.emplace_back(new ConstantFloatTexture(v));
        return flTextures.back().get();
    }
    int addMaterial(SurfaceMaterial* m) {
        materials.emplace_back(m);
        return (int)materials.size() - 1;
    }
};

// Silences stdout (progress reporter, timing prints) while the reference runs.
struct QuietStdout {
    int saved = -1;
    explicit QuietStdout(bool on) {
        if (!on) return;
        fflush(stdout);
        saved = dup(1);
        int devnull = open("/dev/null", O_WRONLY);
        dup2(devnull, 1);
        close(devnull);
    }
    ~QuietStdout() {
        if (saved < 0) return;
        fflush(stdout);
        dup2(saved, 1);
        close(saved);
    }
};

bool g_colorInit = false;
void ensureColor() {
    if (!g_colorInit) {
        initializeColorSystem();
        g_colorInit = true;
    }
}

void loadSpectrum(SampledSpectrum* s, const float* v) {
    for (int i = 0; i < (int)NumSpectralSamples; ++i) (*s)[i] = v[i];
}

} // namespace

extern "C" {

void refd_init(void) { ensureColor(); }

void* refd_scene_new(void) { ensureColor(); return new RefScene(); }
void refd_scene_free(void* s) { delete (RefScene*)s; }

// Lambert (sigma == 0) or Oren-Nayar material with an sRGB reflectance,
// upsampled by the reference (basic_surface_materials.cpp:16-26).
int refd_add_material_diffuse_rgb(void* sp, float r, float g, float b, float sigma) {
    RefScene* s = (RefScene*)sp;
    const SpectrumTexture* R = s->constTexRGB(SpectrumType::Reflectance, r, g, b);
    const FloatTexture* sg = sigma > 0 ? s->constFloat(sigma) : nullptr;
    return s->addMaterial(new DiffuseReflectionSurfaceMaterial(R, sg));
}

// GGX conductor with constant (eta, k) (microfacet_surface_materials.cpp:15-20).
int refd_add_material_ggx_conductor(void* sp, float ax, float ay, float eta, float k) {
    RefScene* s = (RefScene*)sp;
    const SpectrumTexture* te = s->constTexValue(eta);
    const SpectrumTexture* tk = s->constTexValue(k);
    s->dists.emplace_back(new GGXSVMicrofacetDistribution(s->constFloat(ax), s->constFloat(ay)));
    return s->addMaterial(new MicrofacetReflectionSurfaceMaterial(te, tk, s->dists.back().get()));
}

// GGX dielectric with constant IORs (microfacet_surface_materials.cpp:24-30).
int refd_add_material_ggx_dielectric(void* sp, float ax, float ay, float etaExt, float etaInt) {
    RefScene* s = (RefScene*)sp;
    const SpectrumTexture* te = s->constTexValue(etaExt);
    const SpectrumTexture* ti = s->constTexValue(etaInt);
    s->dists.emplace_back(new GGXSVMicrofacetDistribution(s->constFloat(ax), s->constFloat(ay)));
    return s->addMaterial(new MicrofacetScatteringSurfaceMaterial(te, ti, s->dists.back().get()));
}

// Diffuse emitter with a light-source RGB emittance, optionally over a
// reflecting base material (surface_material.h:55-70).
int refd_add_material_emitter_rgb(void* sp, int baseMat, float r, float g, float b) {
    RefScene* s = (RefScene*)sp;
    const SpectrumTexture* Le = s->constTexRGB(SpectrumType::LightSource, r, g, b);
    s->emitters.emplace_back(new DiffuseEmitterSurfaceProperty(Le));
    const SurfaceMaterial* base = baseMat >= 0 ? s->materials[baseMat].get() : nullptr;
    return s->addMaterial(new EmitterSurfaceMaterial(base, s->emitters.back().get()));
}

}  // extern "C"

namespace {
TriangleMeshNode* makeMesh(RefScene* s, int nv, const float* pos, const float* nrm, const float* tan,
                           const float* uv, int nt, const uint32_t* idx, int material) {
    TriangleMeshNode* m = new TriangleMeshNode((uint32_t)nv, 1, false, -1);
    Vertex* vs = m->getVertexArray();
    for (int i = 0; i < nv; ++i) {
        vs[i] = Vertex(Point3D(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]),
                       Normal3D(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]),
                       Tangent3D(tan[3 * i], tan[3 * i + 1], tan[3 * i + 2]),
                       uv ? TexCoord2D(uv[2 * i], uv[2 * i + 1]) : TexCoord2D(0, 0));
    }
    MaterialGroupInTriangleMesh& g = m->getMaterialGroupArray()[0];
    g.material = s->materials.at(material).get();
    std::unique_ptr<Vertex*[]> refs(new Vertex*[3 * (size_t)nt]);
    for (int i = 0; i < 3 * nt; ++i) refs[i] = vs + idx[i];
    g.setTriangles(refs, (uint32_t)nt);
    return m;
}
}  // namespace

extern "C" {

// One triangle mesh with one material group. Positions, normals and tangents
// are world space (3 floats per vertex); uvs may be null.
int refd_add_mesh(void* sp, int nv, const float* pos, const float* nrm, const float* tan,
                  const float* uv, int nt, const uint32_t* idx, int material) {
    RefScene* s = (RefScene*)sp;
    if (s->built) return -1;
    s->meshes.push_back(makeMesh(s, nv, pos, nrm, tan, uv, nt, idx, material));
    s->meshPrimBase.push_back(s->numPrims);
    s->numPrims += (uint32_t)nt;
    return (int)s->meshes.size() - 1;
}

// A mesh in object space placed only through refd_add_instance (a
// TriangleMeshNode under ReferenceNodes). Its triangles take the next prim
// ids, as a mesh would.
int refd_add_prototype(void* sp, int nv, const float* pos, const float* nrm, const float* tan,
                       const float* uv, int nt, const uint32_t* idx, int material) {
    RefScene* s = (RefScene*)sp;
    if (s->built) return -1;
    s->protos.push_back(makeMesh(s, nv, pos, nrm, tan, uv, nt, idx, material));
    s->protoPrimBase.push_back(s->numPrims);
    s->numPrims += (uint32_t)nt;
    return (int)s->protos.size() - 1;
}

// Places prototype `proto` with the object-to-world affine rows (12 floats,
// row-major 3x4): InternalNode(StaticTransform) over the prototype's
// ReferenceNode when rendering, TransformedSurfaceObject for ray batches.
int refd_add_instance(void* sp, int proto, const float* rows) {
    RefScene* s = (RefScene*)sp;
    if (s->built || proto < 0 || (size_t)proto >= s->protos.size()) return -1;
    RefScene::Inst in;
    in.proto = proto;
    memcpy(in.rows, rows, sizeof(in.rows));
    s->insts.push_back(in);
    return (int)s->insts.size() - 1;
}

// Camera-to-world matrix, column-major 4x4 (Matrix4x4.h:42-46).
void refd_set_camera(void* sp, const float* l2w, float sensitivity, float aspect, float fovY,
                     float lensRadius, float imgDist, float objDist) {
    RefScene* s = (RefScene*)sp;
    memcpy(s->l2w, l2w, sizeof(s->l2w));
    s->sensitivity = sensitivity; s->aspect = aspect; s->fovY = fovY;
    s->lensRadius = lensRadius; s->imgDist = imgDist; s->objDist = objDist;
    s->hasCamera = true;
}

// Builds SBVH -> QBVH over all triangles (object order = mesh order, then
// triangle order), as the accelerator for the ray-batch entry points.
int refd_build_accel(void* sp, int quiet) {
    RefScene* s = (RefScene*)sp;
    if (s->built) return 0;
    QuietStdout q(quiet != 0);
    RenderingData rd(nullptr);
    for (auto m : s->meshes)
        m->createRenderingData(&DefaultAllocator::instance(), nullptr, &rd);
    s->objs = rd.surfObjs;
    for (size_t i = 0; i < s->objs.size(); ++i) s->objToPrim[s->objs[i]] = (int32_t)i;
    // prototypes: one aggregate each (ReferenceNode::createRenderingData),
    // instances as TransformedSurfaceObjects after the world triangles
    std::vector<SurfaceObject*> protoObj;
    for (size_t k = 0; k < s->protos.size(); ++k) {
        RenderingData pd(nullptr);
        s->protos[k]->createRenderingData(&DefaultAllocator::instance(), nullptr, &pd);
        for (size_t i = 0; i < pd.surfObjs.size(); ++i)
            s->objToPrim[pd.surfObjs[i]] = (int32_t)(s->protoPrimBase[k] + i);
        if (pd.surfObjs.size() > 1) {
            s->protoAggs.emplace_back(new SurfaceObjectAggregate(pd.surfObjs));
            protoObj.push_back(s->protoAggs.back().get());
        } else {
            protoObj.push_back(pd.surfObjs[0]);
        }
    }
    for (size_t k = 0; k < s->insts.size(); ++k) {
        s->instTFs.emplace_back(new StaticTransform(s->instMatrix(k)));
        s->instObjs.emplace_back(new TransformedSurfaceObject(protoObj[s->insts[k].proto], s->instTFs.back().get()));
        SurfaceObject* o = s->instObjs.back().get();
        s->objs.push_back(o);
        s->objToInst[o] = (int32_t)k;
        s->objToPrim[o] = (int32_t)(0x80000000u | (uint32_t)k);
    }
    s->sbvh = new SBVH(s->objs);
    s->qbvh = new QBVH(*s->sbvh);
    s->stdbvh = new StandardBVH(s->objs, StandardBVH::Partitioning::BinnedSAH);
    s->built = true;
    return 0;
}

// Dumps the StandardBVH as refd_sbvh_dump does the SBVH; stats: nodes, leaf
// refs, depth, SAH cost.
int refd_stdbvh_dump(void* sp, float* boxes, uint32_t* info, int32_t* refs, double* stats) {
    RefScene* s = (RefScene*)sp;
    if (!s->stdbvh) return -1;
    const auto& n = s->stdbvh->m_nodes;
    const auto& l = s->stdbvh->m_objLists;
    if (stats) {
        stats[0] = (double)n.size();
        stats[1] = (double)l.size();
        stats[2] = s->stdbvh->m_depth;
        stats[3] = s->stdbvh->m_cost;
    }
    for (size_t i = 0; boxes && info && i < n.size(); ++i) {
        const auto& b = n[i].bbox;
        float* o = boxes + 6 * i;
        o[0] = b.minP.x; o[1] = b.minP.y; o[2] = b.minP.z;
        o[3] = b.maxP.x; o[4] = b.maxP.y; o[5] = b.maxP.z;
        uint32_t* f = info + 5 * i;
        f[0] = n[i].c0; f[1] = n[i].c1; f[2] = n[i].offsetFirstLeaf; f[3] = n[i].numLeaves;
        f[4] = n[i].numLeaves ? 0xFF : (uint32_t)n[i].axis;
    }
    for (size_t i = 0; refs && i < l.size(); ++i) refs[i] = s->objToPrim.at(l[i]);
    return (int)n.size();
}

// stats: [0]=SBVH nodes, [1]=SBVH leaf refs, [2]=SBVH depth, [3]=SBVH SAH cost,
//        [4]=QBVH nodes, [5]=QBVH leaf refs, [6]=QBVH depth, [7]=QBVH SAH cost
void refd_accel_stats(void* sp, double* st) {
    RefScene* s = (RefScene*)sp;
    st[0] = (double)s->sbvh->m_nodes.size();
    st[1] = (double)s->sbvh->m_objLists.size();
    st[2] = s->sbvh->m_depth;
    st[3] = s->sbvh->m_cost;
    st[4] = (double)s->qbvh->m_nodes.size();
    st[5] = (double)s->qbvh->m_objLists.size();
    st[6] = s->qbvh->m_depth;
    st[7] = s->qbvh->m_cost;
}

// Dumps the SBVH: per node bbox[6] (min xyz, max xyz) into boxes, and
// {c0, c1, offsetFirstLeaf, numLeaves, axis} into info; leaf refs as prim ids.
int refd_sbvh_dump(void* sp, float* boxes, uint32_t* info, int32_t* refs) {
    RefScene* s = (RefScene*)sp;
    const auto& n = s->sbvh->m_nodes;
    for (size_t i = 0; i < n.size(); ++i) {
        const auto& b = n[i].bbox;
        float* o = boxes + 6 * i;
        o[0] = b.minP.x; o[1] = b.minP.y; o[2] = b.minP.z;
        o[3] = b.maxP.x; o[4] = b.maxP.y; o[5] = b.maxP.z;
        uint32_t* f = info + 5 * i;
        f[0] = n[i].c0; f[1] = n[i].c1; f[2] = n[i].offsetFirstLeaf; f[3] = n[i].numLeaves;
        f[4] = n[i].numLeaves ? 0xFF : (uint32_t)n[i].axis;
    }
    const auto& l = s->sbvh->m_objLists;
    for (size_t i = 0; i < l.size(); ++i) refs[i] = s->objToPrim.at(l[i]);
    return (int)n.size();
}

// Dumps the QBVH: 24 floats of child boxes (min_x[4] min_y[4] min_z[4] max_x[4]
// max_y[4] max_z[4]) and 7 uint32 (children[4] raw, top/left/right axis).
int refd_qbvh_dump(void* sp, float* boxes, uint32_t* info, int32_t* refs) {
    RefScene* s = (RefScene*)sp;
    const auto& n = s->qbvh->m_nodes;
    for (size_t i = 0; i < n.size(); ++i) {
        memcpy(boxes + 24 * i, &n[i].min_x, 24 * sizeof(float));
        uint32_t* f = info + 7 * i;
        for (int c = 0; c < 4; ++c) f[c] = n[i].children[c].asUInt;
        f[4] = n[i].topAxis; f[5] = n[i].leftAxis; f[6] = n[i].rightAxis;
    }
    const auto& l = s->qbvh->m_objLists;
    for (size_t i = 0; i < l.size(); ++i) refs[i] = s->objToPrim.at(l[i]);
    return (int)n.size();
}

// A DensityGridMediumNode over bounds[6] with nx*ny*nz densities (x fastest,
// then y, then z), constant base sigma_s / sigma_e and a Henyey-Greenstein
// phase function of asymmetry g (libSLR/Scene/medium_nodes.h:33-46).
int refd_set_medium(void* sp, const float* bounds, const float* density, int nx, int ny, int nz, float sigmaS,
                    float sigmaE, float g) {
    RefScene* s = (RefScene*)sp;
    if (s->built || s->hasMedium || nx < 2 || ny < 2 || nz < 2) return -1;
    memcpy(s->medBounds, bounds, sizeof(s->medBounds));
    s->medN[0] = nx; s->medN[1] = ny; s->medN[2] = nz;
    s->medGrid.assign(nz, std::vector<float>());
    for (int z = 0; z < nz; ++z)
        s->medGrid[z].assign(density + (size_t)z * nx * ny, density + (size_t)(z + 1) * nx * ny);
    s->medSigmaS = sigmaS; s->medSigmaE = sigmaE; s->medG = g;
    s->hasMedium = true;
    return 0;
}

// Probe: density of the scene's medium at n points given in the region's
// local coordinates (xyz in [0, 1]), through DensityGridMediumDistribution::
// evaluateExtinctionCoefficient with a unit extinction spectrum (its value is
// calcDensity, DensityGridMediumDistribution.cpp:369-395, 749-752).
int refd_medium_density(void* sp, int n, const float* params, float* out) {
    RefScene* s = (RefScene*)sp;
    if (!s->hasMedium) return -1;
    float one = 1.0f;
    RegularContinuousSpectrum unit(WavelengthLowBound, WavelengthHighBound, &one, 1);
    BoundingBox3D region(Point3D(s->medBounds[0], s->medBounds[1], s->medBounds[2]),
                         Point3D(s->medBounds[3], s->medBounds[4], s->medBounds[5]));
    DensityGridMediumDistribution dist(region, &unit, &unit, s->medGrid, s->medN[0], s->medN[1], s->medN[2]);
    float pdf;
    WavelengthSamples wls = WavelengthSamples::createWithEqualOffsets(0.5f, 0.5f, &pdf);
    for (int i = 0; i < n; ++i) {
        SampledSpectrum e = dist.evaluateExtinctionCoefficient(Point3D(params[3 * i], params[3 * i + 1], params[3 * i + 2]), wls);
        out[i] = e[0];
    }
    return 0;
}

// Probe: per ray (8 floats {org, dir, tmin, tmax}) the reference medium's
// transmittance estimate (mode 0, DensityGridMediumDistribution::
// evaluateTransmittance, hero wavelength entry) or free-path distance
// (mode 1, DensityGridMediumDistribution::interact: the scattering distance,
// tmax if none), with sigma_s / sigma_e of the scene's medium; one
// IndependentLightPathSampler stream for the whole batch.
int refd_medium_track(void* sp, int mode, int n, const float* rays, uint32_t seed, float* out) {
    RefScene* s = (RefScene*)sp;
    if (!s->hasMedium) return -1;
    RegularContinuousSpectrum ss(WavelengthLowBound, WavelengthHighBound, &s->medSigmaS, 1);
    RegularContinuousSpectrum se(WavelengthLowBound, WavelengthHighBound, &s->medSigmaE, 1);
    BoundingBox3D region(Point3D(s->medBounds[0], s->medBounds[1], s->medBounds[2]),
                         Point3D(s->medBounds[3], s->medBounds[4], s->medBounds[5]));
    DensityGridMediumDistribution dist(region, &ss, &se, s->medGrid, s->medN[0], s->medN[1], s->medN[2]);
    IndependentLightPathSampler sampler(seed);
    float pdf;
    WavelengthSamples wls = WavelengthSamples::createWithEqualOffsets(0.5f, 0.5f, &pdf);
    for (int i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        Ray ray(Point3D(r[0], r[1], r[2]), Vector3D(r[3], r[4], r[5]), 0.0f);
        RaySegment seg(r[6], r[7]);
        bool single = false;
        if (mode == 0) {
            SampledSpectrum T = dist.evaluateTransmittance(ray, seg, wls, sampler, &single);
            out[i] = T[wls.selectedLambdaIndex];
        } else {
            MediumInteraction mi;
            SampledSpectrum thr;
            out[i] = dist.interact(ray, seg, wls, sampler, &mi, &thr, &single) ? mi.getDistance() : r[7];
        }
    }
    return 0;
}

// 0: PTRenderer, 1: VolumetricPTRenderer, 2: BPTRenderer for refd_render.
void refd_set_integrator(void* sp, int integrator) { ((RefScene*)sp)->integrator = integrator; }

// World-to-object matrix (16 floats, row-major 4x4) of instance k as
// StaticTransform computes its inverse (libSLR/Core/transform.h:38).
int refd_instance_inverse(void* sp, int k, float* rows) {
    RefScene* s = (RefScene*)sp;
    if (k < 0 || (size_t)k >= s->insts.size()) return -1;
    const StaticTransform inv = invert(StaticTransform(s->instMatrix(k)));
    // column c of the inverse is the inverse applied to the unit vector e_c
    for (int c = 0; c < 4; ++c) {
        const Vector4D col = inv * Vector4D(c == 0 ? 1.0f : 0.0f, c == 1 ? 1.0f : 0.0f, c == 2 ? 1.0f : 0.0f,
                                            c == 3 ? 1.0f : 0.0f);
        rows[c] = col.x;
        rows[4 + c] = col.y;
        rows[8 + c] = col.z;
        rows[12 + c] = col.w;
    }
    return 0;
}

// Closest hit for n rays, each 8 floats {org xyz, dir xyz, tmin, tmax}.
// accel: 0 = QBVH, 1 = SBVH, 2 = StandardBVH. Miss: prim = -1, t = tmax. inst (may be null):
// the instance of the hit triangle, -1 in world space.
int refd_intersect(void* sp, int accel, int n, const float* rays, float* t, int32_t* prim,
                   float* b0, float* b1, int32_t* inst) {
    RefScene* s = (RefScene*)sp;
    IndependentLightPathSampler sampler(1u);
    for (int i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        Ray ray(Point3D(r[0], r[1], r[2]), Vector3D(r[3], r[4], r[5]), 0.0f);
        RaySegment seg(r[6], r[7]);
        SurfaceInteraction si;
        const SurfaceObject* hit = nullptr;
        bool ok = accel == 0   ? s->qbvh->intersect(ray, seg, sampler, &si, &hit)
                  : accel == 1 ? s->sbvh->intersect(ray, seg, sampler, &si, &hit)
                               : s->stdbvh->intersect(ray, seg, sampler, &si, &hit);
        if (ok) {
            t[i] = si.getDistance();
            prim[i] = s->objToPrim.at(si.m_obj);
            float u, v;
            si.getSurfaceParameter(&u, &v);
            b0[i] = u; b1[i] = v;
            if (inst) inst[i] = s->objToInst.count(hit) ? s->objToInst.at(hit) : -1;
        } else {
            t[i] = r[7]; prim[i] = -1; b0[i] = 0; b1[i] = 0;
            if (inst) inst[i] = -1;
        }
    }
    return 0;
}

// Fractional visibility for n segments (QBVH.h:367-376).
int refd_visibility(void* sp, int accel, int n, const float* rays, float* vis) {
    RefScene* s = (RefScene*)sp;
    for (int i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        Ray ray(Point3D(r[0], r[1], r[2]), Vector3D(r[3], r[4], r[5]), 0.0f);
        RaySegment seg(r[6], r[7]);
        vis[i] = accel == 0   ? s->qbvh->testVisibility(ray, seg)
                 : accel == 1 ? s->sbvh->testVisibility(ray, seg)
                              : s->stdbvh->testVisibility(ray, seg);
    }
    return 0;
}

// Equirectangular environment of w x h RGB floats (rows from the +y pole),
// converted to half as an RGBA16F image (alpha 1) and attached to the scene
// by InfiniteSphereNode (libSLR/Scene/node.cpp:201-206) with the given scale.
int refd_set_environment(void* sp, int w, int h, const float* rgb, float scale) {
    RefScene* s = (RefScene*)sp;
    if (w < 4 || h < 4 || s->envTex) return -1;
    std::vector<RGBA16Fx4> px((size_t)w * h);
    for (size_t i = 0; i < px.size(); ++i)
        px[i] = RGBA16Fx4{half(rgb[3 * i]), half(rgb[3 * i + 1]), half(rgb[3 * i + 2]), half(1.0f)};
    s->envImage = DefaultAllocator::instance().create<TiledImage2D>(
        (const void*)px.data(), (uint32_t)w, (uint32_t)h, ColorFormat::RGBA16Fx4, &DefaultAllocator::instance(),
        ImageStoreMode::AsIs, SpectrumType::LightSource);
    s->envTex.reset(new ImageSpectrumTexture(s->envImage, &s->envMapping));
    s->envScale = scale;
    return 0;
}

// PTRenderer::render over the scene; film_out receives h*w*16 floats (row-major
// pixels, 16 strata each) of the raw sensor sums. Returns render seconds
// (excluding Scene::build) or a negative value on error.
double refd_render_batches(void* sp, int w, int h, int spp, uint32_t seed, int batches, int threads,
                           float* film_out, int quiet, double* build_seconds) {
    RefScene* s = (RefScene*)sp;
    if (!s->hasCamera || s->built) return -1.0;
    QuietStdout q(quiet != 0);

    InternalNode root(new StaticTransform());
    for (auto m : s->meshes) root.addChildNode(m);
    // instances: InternalNode(transform) over the prototype's ReferenceNode,
    // one ReferenceNode per prototype shared by its instances
    std::vector<std::unique_ptr<ReferenceNode>> refNodes;
    for (auto m : s->protos) refNodes.emplace_back(new ReferenceNode(m));
    std::vector<std::unique_ptr<StaticTransform>> instTF;
    std::vector<std::unique_ptr<InternalNode>> instNodes;
    for (size_t k = 0; k < s->insts.size(); ++k) {
        instTF.emplace_back(new StaticTransform(s->instMatrix(k)));
        instNodes.emplace_back(new InternalNode(instTF.back().get()));
        instNodes.back()->addChildNode(refNodes[s->insts[k].proto].get());
        root.addChildNode(instNodes.back().get());
    }
    float mat[16];
    memcpy(mat, s->l2w, sizeof(mat));
    InternalNode camNode(new StaticTransform(Matrix4x4(mat)));
    PerspectiveCameraNode cam(s->sensitivity, s->aspect, s->fovY, s->lensRadius, s->imgDist, s->objDist);
    camNode.addChildNode(&cam);
    root.addChildNode(&camNode);

    std::unique_ptr<DensityGridMediumNode> medNode;
    std::unique_ptr<HenyeyGreensteinScatteringMediumMaterial> medMat;
    if (s->hasMedium) {
        BoundingBox3D region(Point3D(s->medBounds[0], s->medBounds[1], s->medBounds[2]),
                             Point3D(s->medBounds[3], s->medBounds[4], s->medBounds[5]));
        s->spectra.emplace_back(new RegularContinuousSpectrum(WavelengthLowBound, WavelengthHighBound, &s->medSigmaS, 1));
        const AssetSpectrum* ss = s->spectra.back().get();
        s->spectra.emplace_back(new RegularContinuousSpectrum(WavelengthLowBound, WavelengthHighBound, &s->medSigmaE, 1));
        const AssetSpectrum* se = s->spectra.back().get();
        medMat.reset(new HenyeyGreensteinScatteringMediumMaterial(s->constFloat(s->medG)));
        medNode.reset(new DensityGridMediumNode(region, ss, se, s->medGrid, s->medN[0], s->medN[1], s->medN[2],
                                                medMat.get()));
        root.addChildNode(medNode.get());
    }
    Scene scene(&root);
    std::unique_ptr<InfiniteSphereNode> envNode;
    if (s->envTex) {
        envNode.reset(new InfiniteSphereNode(&scene, s->envTex.get(), s->envScale));
        scene.setEnvironmentNode(envNode.get());
    }
    // BPTRenderer::render adds per-thread separated buffers to the camera's
    // sensor on every call (ImageSensor::addSeparatedBuffers, which cannot
    // add them a second time), so each BPT batch renders on a fresh build of
    // the scene: Scene::destory releases the camera with its sensor.
    const bool bpt = s->integrator == 2;
    auto t0 = std::chrono::steady_clock::now();
    scene.build(&DefaultAllocator::instance());
    auto t1 = std::chrono::steady_clock::now();
    if (build_seconds) *build_seconds = std::chrono::duration<double>(t1 - t0).count();

    RenderSettings settings;
    settings.addItem(RenderSettingItem::NumThreads, (int32_t)threads);
    settings.addItem(RenderSettingItem::ImageWidth, (int32_t)w);
    settings.addItem(RenderSettingItem::ImageHeight, (int32_t)h);
    settings.addItem(RenderSettingItem::TimeStart, 0.0f);
    settings.addItem(RenderSettingItem::TimeEnd, 0.0f);
    settings.addItem(RenderSettingItem::Brightness, 1.0f);
    settings.addItem(RenderSettingItem::RNGSeed, (int32_t)seed);

    // PTRenderer writes BMP snapshots at 2^k spp into the working directory.
    char cwd[4096];
    if (!getcwd(cwd, sizeof(cwd))) cwd[0] = 0;
    char tmpl[] = "/tmp/slr_ref_XXXXXX";
    char* dir = mkdtemp(tmpl);
    if (dir && chdir(dir) != 0) dir = nullptr;

    // batch b: the renderer again with seed + b (PTRenderer::render starts
    // with ImageSensor::init, so each batch's sensor sums start at zero)
    double renderSeconds = 0;
    for (int b = 0; b < batches; ++b) {
        settings.addItem(RenderSettingItem::RNGSeed, (int32_t)(seed + (uint32_t)b));
        if (bpt && b > 0) {
            scene.destory();
            scene.build(&DefaultAllocator::instance());
        }
        const ImageSensor* sensor = scene.getCamera()->getSensor();
        auto t2 = std::chrono::steady_clock::now();
        if (bpt) {
            BPTRenderer renderer((uint32_t)spp);
            renderer.render(scene, settings);
        } else if (s->integrator == 1) {
            VolumetricPTRenderer renderer((uint32_t)spp);
            renderer.render(scene, settings);
        } else {
            PTRenderer renderer((uint32_t)spp);
            renderer.render(scene, settings);
        }
        auto t3 = std::chrono::steady_clock::now();
        renderSeconds += std::chrono::duration<double>(t3 - t2).count();
        float* out = film_out + (size_t)b * w * h * NumStrataForStorage;
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                DiscretizedSpectrum px = sensor->pixel((uint32_t)x, (uint32_t)y);
                for (int k = 0; k < (int)NumStrataForStorage; ++k)
                    out[((size_t)y * w + x) * NumStrataForStorage + k] = px[k];
                // BPT: the light-to-lens sums of each thread's separated
                // buffer, which saveImage adds with the same scale
                // (ImageSensor.cpp:139-152)
                for (int t = 0; bpt && t < threads; ++t) {
                    DiscretizedSpectrum sp = sensor->pixel((uint32_t)t, (uint32_t)x, (uint32_t)y);
                    for (int k = 0; k < (int)NumStrataForStorage; ++k)
                        out[((size_t)y * w + x) * NumStrataForStorage + k] += sp[k];
                }
            }
    }

    if (dir) {
        char cmd[256];
        for (uint32_t k = 0; k < 32; ++k) {
            snprintf(cmd, sizeof(cmd), "%s/%03u.bmp", dir, k);
            unlink(cmd);
        }
        if (cwd[0] && chdir(cwd) != 0) {}
        rmdir(dir);
    }

    scene.destory();   // also releases the node tree's rendering data (Scene.cpp:40-46)
    return renderSeconds;
}

double refd_render(void* sp, int w, int h, int spp, uint32_t seed, int threads, float* film_out, int quiet,
                   double* build_seconds) {
    return refd_render_batches(sp, w, h, spp, seed, 1, threads, film_out, quiet, build_seconds);
}

// ImageSensor::saveImage (libSLR/Core/ImageSensor.cpp:139-195) of a sensor
// holding the given sums: w*h pixels of 16 strata (row-major), written as a
// BMP to `path` with the sensor's sensitivity and the scale PTRenderer passes
// (brightness / samples). Returns 0, or -1 if the file could not be read back.
int refd_save_image(int w, int h, const float* strata, float sensitivity, float scale, const char* path) {
    ImageSensor sensor((uint32_t)w, (uint32_t)h, sensitivity);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            sensor.pixel((uint32_t)x, (uint32_t)y) =
                SpectrumStorage(DiscretizedSpectrum(strata + ((size_t)y * w + x) * NumStrataForStorage));
    sensor.saveImage(path, scale);
    FILE* f = fopen(path, "rb");
    if (!f) return -1;
    fclose(f);
    return 0;
}

// ---------------------------------------------------------------- probes

// XORShift128 seeded as XORShiftRNG(seed) (XORShiftRNG.cpp:21-27); out gets
// n raw uint32 draws, fout n getFloat0cTo1o draws from a second identical stream.
void refd_rng_stream(uint32_t seed, int n, uint32_t* out, float* fout) {
    XORShiftRNG a(seed), b(seed);
    for (int i = 0; i < n; ++i) out[i] = a.getUInt();
    for (int i = 0; i < n; ++i) fout[i] = b.getFloat0cTo1o();
}

// WavelengthSamples::createWithEqualOffsets (spectrum_types.h:59-69).
void refd_wavelengths(float offset, float uLambda, float* lambdas, int* hero, float* pdf) {
    WavelengthSamples w = WavelengthSamples::createWithEqualOffsets(offset, uLambda, pdf);
    for (int i = 0; i < (int)NumSpectralSamples; ++i) lambdas[i] = w.lambdas[i];
    *hero = w.selectedLambdaIndex;
}

// Meng upsampling of an sRGB triple (type 0 = reflectance, 1 = light source)
// evaluated at n wavelengths (spectrum_types.cpp:482-543,611-651).
void refd_upsample(int type, float r, float g, float b, int n, const float* lambdas, float* out) {
    ensureColor();
    UpsampledContinuousSpectrum s(type == 0 ? SpectrumType::Reflectance : SpectrumType::LightSource,
                                  ColorSpace::sRGB, r, g, b);
    s.evaluate(lambdas, (uint32_t)n, out);
}

// Same, through the 16-wavelength path used by textures during rendering.
void refd_upsample16(int type, float r, float g, float b, const float* lambdas16, float* out16) {
    ensureColor();
    UpsampledContinuousSpectrum s(type == 0 ? SpectrumType::Reflectance : SpectrumType::LightSource,
                                  ColorSpace::sRGB, r, g, b);
    WavelengthSamples w(lambdas16);
    SampledSpectrum v = s.evaluate(w);
    for (int i = 0; i < (int)NumSpectralSamples; ++i) out16[i] = v[i];
}

// Film strata of the colour-matching functions (spectrum_types.cpp:945-990).
void refd_cmf_strata(float* xbar, float* ybar, float* zbar, float* integral) {
    ensureColor();
    for (int i = 0; i < (int)NumStrataForStorage; ++i) {
        xbar[i] = DiscretizedSpectrum::xbar[i];
        ybar[i] = DiscretizedSpectrum::ybar[i];
        zbar[i] = DiscretizedSpectrum::zbar[i];
    }
    *integral = DiscretizedSpectrum::integralCMF;
}

// SpectrumStorage::add of one sample (spectrum_types.h:725-736) into 16 strata.
void refd_film_add(const float* lambdas16, const float* value16, float* strata16) {
    SpectrumStorage st;
    WavelengthSamples w(lambdas16);
    SampledSpectrum v;
    loadSpectrum(&v, value16);
    st.add(w, v);
    DiscretizedSpectrum r = st.getValue().result;
    for (int i = 0; i < (int)NumStrataForStorage; ++i) strata16[i] = r[i];
}

// BSDF probes. kind: 0 Lambert(R), 1 Oren-Nayar(R, p0 = sigma),
// 2 GGX conductor(eta = A, k = B, ax = p0, ay = p1),
// 3 GGX dielectric(etaExt = A, etaInt = B, ax = p0, ay = p1).
// q = {dirLocal xyz, gNormalLocal xyz}; wl = hero index.
// mode 0: sample with u[3] -> out_dir[3], out_pdf, out_fs[16], out_type (bit0 = reflection)
// mode 1: evaluate + PDF at dir -> out_fs[16], out_pdf
static BSDF* makeBSDF(int kind, const float* A, const float* B, const float* p, ArenaAllocator& mem,
                      MicrofacetDistribution** dist) {
    SampledSpectrum a, b;
    loadSpectrum(&a, A);
    if (B) loadSpectrum(&b, B);
    switch (kind) {
    case 0: return mem.create<LambertianBRDF>(a);
    case 1: return mem.create<OrenNayerBRDF>(a, p[0]);
    case 2: *dist = mem.create<GGXMicrofacetDistribution>(p[0], p[1]); return mem.create<MicrofacetBRDF>(a, b, *dist);
    case 3: *dist = mem.create<GGXMicrofacetDistribution>(p[0], p[1]); return mem.create<MicrofacetBSDF>(a, b, *dist);
    }
    return nullptr;
}

int refd_bsdf(int kind, const float* A, const float* B, const float* p, int n, const float* q,
              const int* wl, int mode, const float* uOrDir, float* out_dir, float* out_pdf,
              float* out_fs, int* out_type) {
    ArenaAllocator mem;
    MicrofacetDistribution* dist = nullptr;
    BSDF* bsdf = makeBSDF(kind, A, B, p, mem, &dist);
    if (!bsdf) return -1;
    for (int i = 0; i < n; ++i) {
        const float* qq = q + 6 * i;
        BSDFQuery query(Vector3D(qq[0], qq[1], qq[2]), Normal3D(qq[3], qq[4], qq[5]), (int16_t)wl[i]);
        SampledSpectrum fs;
        if (mode == 0) {
            const float* u = uOrDir + 3 * i;
            BSDFQueryResult res;
            res.dirPDF = 0.0f;
            fs = bsdf->sample(query, BSDFSample(u[0], u[1], u[2]), &res);
            out_dir[3 * i + 0] = res.dirLocal.x;
            out_dir[3 * i + 1] = res.dirLocal.y;
            out_dir[3 * i + 2] = res.dirLocal.z;
            out_pdf[i] = res.dirPDF;
            out_type[i] = res.dirPDF > 0 ? (res.sampledType.isReflection() ? 1 : 2) : 0;
        } else {
            const float* d = uOrDir + 3 * i;
            Vector3D dir(d[0], d[1], d[2]);
            fs = bsdf->evaluate(query, dir);
            out_pdf[i] = bsdf->evaluatePDF(query, dir);
        }
        for (int k = 0; k < (int)NumSpectralSamples; ++k) out_fs[16 * i + k] = fs[k];
    }
    return 0;
}

// concentricSampleDisk / cosineSampleHemisphere / uniformSampleTriangle probes
// (distributions.cpp:57-90, distributions.h:30-72).
void refd_sampling(int n, const float* u, float* disk, float* hemi, float* tri) {
    for (int i = 0; i < n; ++i) {
        float dx, dy;
        concentricSampleDisk(u[2 * i], u[2 * i + 1], &dx, &dy);
        disk[2 * i] = dx; disk[2 * i + 1] = dy;
        Vector3D h = cosineSampleHemisphere(u[2 * i], u[2 * i + 1]);
        hemi[3 * i] = h.x; hemi[3 * i + 1] = h.y; hemi[3 * i + 2] = h.z;
        float b0, b1;
        uniformSampleTriangle(u[2 * i], u[2 * i + 1], &b0, &b1);
        tri[2 * i] = b0; tri[2 * i + 1] = b1;
    }
}

} // extern "C"
