// Host-side owners of the device-resident scene and films, shared by the
// path tracer (render.hip) and the bidirectional path tracer (bpt.hip):
// device buffers, the scene's arrays with the kernel-side view (DevScene)
// over them, and the FP64 film sums.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "engine.h"

namespace slrh {

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            if (err) *err = std::string(#expr) + ": " + hipGetErrorString(e_);     \
            return SLRH_ERR_DEVICE_CODE;                                           \
        }                                                                          \
    } while (0)

constexpr int SLRH_ERR_STATE_CODE = -2;
constexpr int SLRH_ERR_DEVICE_CODE = -3;
constexpr int SLRH_ERR_NOMEM_CODE = -4;
constexpr int SLRH_ERR_CAPACITY_CODE = -5;

// ------------------------------------------------------------------ buffers

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    int alloc(size_t count, std::string* err) {
        if (p) { (void)hipFree(p); p = nullptr; }
        n = count;
        if (count == 0) return 0;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            if (err) *err = std::string("hipMalloc: ") + hipGetErrorString(e);
            return SLRH_ERR_NOMEM_CODE;
        }
        return 0;
    }
    int upload(const std::vector<T>& v, std::string* err) {
        int r = alloc(v.size(), err);
        if (r) return r;
        if (!v.empty()) {
            hipError_t e = hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                if (err) *err = std::string("hipMemcpy: ") + hipGetErrorString(e);
                return SLRH_ERR_DEVICE_CODE;
            }
        }
        return 0;
    }
};

struct DeviceScene {
    DevBuf<float4> nodes, tris, vpos, vnrm, vtan;
    DevBuf<uint4> primVerts;
    DevBuf<DevMaterial> mats;
    DevBuf<DevSpectrum> spectra;
    DevBuf<float> specPool, lightCdf, lightPmf;
    DevBuf<uint2> lights;
    DevBuf<float4> lightRec, triShade;
    DevBuf<DevSpectrum> matSpec;
    DevBuf<int32_t> primToLight, instMat;
    DevBuf<float4> instXf;
    DevBuf<uint32_t> spill;      // traversal stack spill for ray batches
    DevBuf<float4> envTexel;
    DevBuf<float2> mengBasis;
    DevBuf<float> envCondPdf, envCondCdf, envMargPdf, envMargCdf;
    DevBuf<float> medMajorant;
    DevBuf<float4> medCorners;
    DevBuf<unsigned long long> counters;
    // binned-SAH binary BVH for ray-batch queries (when requested): 3 float4
    // per node, 3 float4 per leaf triangle reference (the QBVH's format)
    DevBuf<float4> bvh2Nodes, bvh2Tris;
    DevScene view;
    bool hasCamera = false;
    uint32_t matTypeMask = 0;   // bit t set: a material of type t (MAT_*) is present
    CameraParams camera;
    size_t bytes = 0;
    std::shared_ptr<struct SlotPool> pool;   // wavefront state kept between renders
    std::shared_ptr<struct BptPool> bptPool; // the bidirectional integrator's, likewise
    // traversal variant chosen by the autotuner (on a pool of tunedSlots slots)
    int tunedVariant = -1;
    int tunedShape = 0;   // index into kShadeShapes
    uint32_t tunedSlots = 0;
    uint32_t tunedPoolWidth = 0;   // live paths chosen by the pool-width pilot (0: not tuned)
};

struct DeviceFilm {
    uint32_t w = 0, h = 0;
    DevBuf<double> data;   // h*w*16
};

// PerspectiveCamera's derived quantities for a w x h render (PerspectiveCamera.cpp:16-27)
// and the camera-to-world rows with their inverse transpose.
void cameraDerived(const CameraParams& c, uint32_t w, uint32_t h, DevCamera* d);

}  // namespace slrh
