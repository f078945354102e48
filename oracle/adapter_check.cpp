// Test program for include/slr_hip_renderer.hpp, the reference-side adapter:
// a libSLR host program's view of this library. It reads a scene dump
// (written by tests/test_gpu_adapter.py), mirrors it into the adapter with
// the calls a scene-graph binding makes, constructs an SLR::Scene and
// SLR::RenderSettings as HostProgram/main.cpp does, calls the renderer
// through its SLR::Renderer base, and writes the film's sums.
//
//   adapter_check <scene dump> <film out> <integrator>
//
// Dump layout (little-endian): uint32 W, H, spp, seed, nmat; per material
// uint32 kind (0 matte, 1 emitter), uint32 base, float r, g, b; uint32 nmesh;
// per mesh uint32 nv, nt, then nv x (pos 3, nrm 3, tan 3) floats, nt x 3
// uint32 indices, nt uint32 materials; then the camera: 16 floats l2w
// (column-major) and sensitivity, aspect, fovY, lensRadius, imgDist, objDist.
#include <cstdio>
#include <cstring>
#include <vector>

#include <libSLR/Core/transform.h>
#include <libSLR/Scene/Scene.h>
#include <libSLR/Scene/node.h>

#include "slr_hip_renderer.hpp"

namespace {
struct Reader {
    FILE* f;
    uint32_t u() { uint32_t v = 0; if (fread(&v, 4, 1, f) != 1) throw std::runtime_error("short dump"); return v; }
    float fl() { float v = 0; if (fread(&v, 4, 1, f) != 1) throw std::runtime_error("short dump"); return v; }
};
}  // namespace

int main(int argc, char** argv) {
    if (argc != 4) {
        fprintf(stderr, "usage: adapter_check <scene dump> <film out> <integrator>\n");
        return 2;
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    try {
        Reader rd{f};
        const uint32_t W = rd.u(), H = rd.u(), spp = rd.u(), seed = rd.u();
        slrh_slr::HIPRenderer renderer(spp, (uint32_t)atoi(argv[3]));
        const uint32_t nmat = rd.u();
        std::vector<int> mats;
        for (uint32_t i = 0; i < nmat; ++i) {
            const uint32_t kind = rd.u(), base = rd.u();
            const float r = rd.fl(), g = rd.fl(), b = rd.fl();
            if (kind == 0) mats.push_back(renderer.matte(renderer.spectrumRGB(false, r, g, b)));
            else mats.push_back(renderer.emitter(mats.at(base), renderer.spectrumRGB(true, r, g, b)));
        }
        const uint32_t nmesh = rd.u();
        for (uint32_t m = 0; m < nmesh; ++m) {
            const uint32_t nv = rd.u(), nt = rd.u();
            std::vector<SLR::Vertex> v(nv);
            for (uint32_t i = 0; i < nv; ++i) {
                float x[9];
                for (float& c : x) c = rd.fl();
                v[i] = SLR::Vertex(SLR::Point3D(x[0], x[1], x[2]), SLR::Normal3D(x[3], x[4], x[5]),
                                   SLR::Tangent3D(x[6], x[7], x[8]), SLR::TexCoord2D(0, 0));
            }
            std::vector<uint32_t> idx(3 * nt), mt(nt);
            for (uint32_t& x : idx) x = rd.u();
            for (uint32_t& x : mt) x = (uint32_t)mats.at(rd.u());
            renderer.addMesh(v, idx, mt);
        }
        float l2w[16], cam[6];
        for (float& x : l2w) x = rd.fl();
        for (float& x : cam) x = rd.fl();
        renderer.camera(l2w, cam[0], cam[1], cam[2], cam[3], cam[4], cam[5]);
        fclose(f);

        // the host program's side: a scene and settings (HostProgram/main.cpp:42-66)
        SLR::InternalNode root(new SLR::StaticTransform());
        SLR::Scene scene(&root);
        SLR::RenderSettings settings;
        settings.addItem(SLR::RenderSettingItem::NumThreads, (int32_t)1);
        settings.addItem(SLR::RenderSettingItem::ImageWidth, (int32_t)W);
        settings.addItem(SLR::RenderSettingItem::ImageHeight, (int32_t)H);
        settings.addItem(SLR::RenderSettingItem::Brightness, 1.0f);
        settings.addItem(SLR::RenderSettingItem::RNGSeed, (int32_t)seed);
        const SLR::Renderer& base = renderer;
        base.render(scene, settings);
        const std::vector<double> sums = renderer.strata();
        FILE* o = fopen(argv[2], "wb");
        if (!o || fwrite(sums.data(), sizeof(double), sums.size(), o) != sums.size()) return 3;
        fclose(o);
    } catch (const std::exception& e) {
        fprintf(stderr, "adapter_check: %s\n", e.what());
        return 1;
    }
    return 0;
}
