// Reference-side adapter: an SLR::Renderer (libSLR/Core/renderer.h:16-20)
// whose render() runs this library's GPU integrators on a scene mirrored
// into it. A libSLR host program includes this header (with libSLR's own
// headers on its include path) and links libslr_hip.so; where the scene
// graph builds meshes, materials and the camera (libSLRSceneGraph/API.cpp:
// 795-868 createMesh, createSurfaceMaterial; camera_nodes.h:16-35), it hands
// the same data to the renderer, and where setRenderer picks "PT",
// "Volumetric PT" or "BPT" (API.cpp:1214-1300) it creates this renderer
// with the matching integrator. render() ignores the SLR::Scene it is given:
// the mirrored scene is the one rendered.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "slr_hip.h"

#include <libSLR/Core/RenderSettings.h>
#include <libSLR/Core/geometry.h>
#include <libSLR/Core/renderer.h>

namespace slrh_slr {

class HIPRenderer : public SLR::Renderer {
    slrh_context* m_ctx = nullptr;
    slrh_scene* m_scene = nullptr;
    mutable slrh_film* m_film = nullptr;
    mutable uint32_t m_w = 0, m_h = 0;
    uint32_t m_spp;
    uint32_t m_integrator;
    mutable bool m_built = false;

    void check(int r, const char* what) const {
        if (r != SLRH_OK) throw std::runtime_error(std::string(what) + ": " + slrh_last_error(m_ctx));
    }

public:
    // PTRenderer(spp) with SLRH_PT, VolumetricPTRenderer(spp) with
    // SLRH_VOL_PT, BPTRenderer(spp) with SLRH_BPT
    HIPRenderer(uint32_t spp, uint32_t integrator = SLRH_PT, int device = 0) : m_spp(spp), m_integrator(integrator) {
        if (slrh_create(device, &m_ctx) != SLRH_OK) throw std::runtime_error("no HIP device");
        check(slrh_scene_create(m_ctx, &m_scene), "slrh_scene_create");
    }
    ~HIPRenderer() override {
        if (m_film) slrh_film_destroy(m_film);
        if (m_scene) slrh_scene_destroy(m_scene);
        if (m_ctx) slrh_destroy(m_ctx);
    }
    HIPRenderer(const HIPRenderer&) = delete;
    HIPRenderer& operator=(const HIPRenderer&) = delete;

    // Spectra and materials as the scene language creates them: an sRGB
    // reflectance or light-source spectrum, matte (Lambert), and an emitter
    // over a scattering base (EmitterSurfaceMaterial with a diffuse EDF).
    int spectrumRGB(bool lightSource, float r, float g, float b) {
        int id;
        check(slrh_spectrum_rgb(m_scene, lightSource ? SLRH_SPECTRUM_LIGHT_SOURCE : SLRH_SPECTRUM_REFLECTANCE, r, g,
                                b, &id),
              "slrh_spectrum_rgb");
        return id;
    }
    int matte(int reflectance) {
        int id;
        check(slrh_material_lambert(m_scene, reflectance, &id), "slrh_material_lambert");
        return id;
    }
    int emitter(int base, int emittance) {
        int id;
        check(slrh_material_diffuse_emitter(m_scene, base, emittance, &id), "slrh_material_diffuse_emitter");
        return id;
    }
    // A mesh as TriangleMeshNode holds it (libSLR/Core/geometry.h:293-301
    // Vertex, 3 indices per triangle, one material per triangle), in world
    // space.
    int addMesh(const std::vector<SLR::Vertex>& v, const std::vector<uint32_t>& idx,
                const std::vector<uint32_t>& matPerTri) {
        std::vector<float> p, n, t;
        p.reserve(3 * v.size());
        n.reserve(3 * v.size());
        t.reserve(3 * v.size());
        for (const SLR::Vertex& x : v) {
            p.insert(p.end(), {x.position.x, x.position.y, x.position.z});
            n.insert(n.end(), {x.normal.x, x.normal.y, x.normal.z});
            t.insert(t.end(), {x.tangent.x, x.tangent.y, x.tangent.z});
        }
        int id;
        check(slrh_mesh_add(m_scene, (uint32_t)v.size(), p.data(), n.data(), t.data(), (uint32_t)(idx.size() / 3),
                            idx.data(), matPerTri.data(), 0, &id),
              "slrh_mesh_add");
        return id;
    }
    // PerspectiveCameraNode(sensitivity, aspect, fovY, lensRadius, imgDist,
    // objDist) under a node whose world transform is l2w (column-major, as
    // Matrix4x4(float[16]) takes it)
    void camera(const float l2w[16], float sensitivity, float aspect, float fovY, float lensRadius, float imgDist,
                float objDist) {
        check(slrh_camera_perspective(m_scene, l2w, sensitivity, aspect, fovY, lensRadius, imgDist, objDist),
              "slrh_camera_perspective");
    }

    // Renderer::render: ImageWidth, ImageHeight and RNGSeed from the
    // settings; builds the mirrored scene on the first call (SBVH -> QBVH,
    // upload) and renders into the film, which holds the sensor's sums.
    void render(const SLR::Scene&, const SLR::RenderSettings& settings) const override {
        const uint32_t w = (uint32_t)settings.getInt(SLR::RenderSettingItem::ImageWidth);
        const uint32_t h = (uint32_t)settings.getInt(SLR::RenderSettingItem::ImageHeight);
        if (!m_built) {
            check(slrh_scene_build(m_scene), "slrh_scene_build");
            m_built = true;
        }
        if (m_film) slrh_film_destroy(m_film);
        m_film = nullptr;
        check(slrh_film_create(m_ctx, w, h, &m_film), "slrh_film_create");
        m_w = w;
        m_h = h;
        slrh_render_settings rs = {w, h, m_spp, 0, (uint32_t)settings.getInt(SLR::RenderSettingItem::RNGSeed),
                                   m_integrator, 0, 0, 0, 0};
        check(slrh_render(m_scene, &rs, m_film, nullptr), "slrh_render");
    }

    // the film's sums: w * h * 16 doubles (pixel-major, 16 strata each)
    std::vector<double> strata() const {
        if (!m_film) return {};
        std::vector<double> out((size_t)m_w * m_h * 16);
        check(slrh_film_read(m_film, SLRH_FILM_STRATA16_F64, 1.0, out.data(), out.size() * sizeof(double)),
              "slrh_film_read");
        return out;
    }
    // the BMP ImageSensor::saveImage(path, brightness / spp) writes
    // (libSLR/Core/ImageSensor.cpp:139-195)
    std::vector<uint8_t> bmp(float brightness, float sensitivity) const {
        size_t size = 0;
        check(slrh_film_bmp(m_film, brightness / m_spp, sensitivity, nullptr, 0, &size), "slrh_film_bmp");
        std::vector<uint8_t> out(size);
        check(slrh_film_bmp(m_film, brightness / m_spp, sensitivity, out.data(), size, &size), "slrh_film_bmp");
        return out;
    }
    slrh_scene* scene() const { return m_scene; }
};

}  // namespace slrh_slr
