// Internal interface between the C ABI (capi.cpp) and the HIP side
// (render.hip): host scene description in, device-resident scene out.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "device_common.h"
#include "host_scene.h"

namespace slrh {

struct CameraParams {
    float l2w[16];   // column-major 4x4
    float sensitivity = 1, aspect = 1, fovY = 0.5f, lensRadius = 0, imgDist = 1, objDist = 1;
};

// A mesh placed only through instances: its triangles (a range of the
// global arrays, in object space) and its own accelerator over them, whose
// references are global triangle ids.
struct Prototype {
    uint32_t triFirst = 0, triCount = 0;
    Accel accel;
};

// A placed copy of a prototype: object-to-world affine transform (3 rows of
// 4, row-major), its inverse as the reference computes it (4 rows: the FP32
// elimination leaves rounding residue in the bottom row, and points are
// divided by the w it yields), and a material replacing the prototype's
// materials (-1: keep them).
struct Instance {
    uint32_t proto = 0;
    float l2w[12], w2l[16];
    int32_t material = -1;
};

// Accelerator references and hit ids: a world-space triangle by its global
// id, an instance by kRefInstance | instance id.
constexpr uint32_t kRefInstance = 0x80000000u;



struct SceneData {
    std::vector<float> vpos, vnrm, vtan;   // 3 floats per vertex (object space for prototype triangles)
    std::vector<uint32_t> idx;             // 3 per triangle (global vertex indices)
    std::vector<uint32_t> triMat;          // material per triangle
    std::vector<int32_t> triProto;         // per triangle: its prototype, or -1 in world space
    std::vector<HostSpectrum> spectra;
    std::vector<DevMaterial> materials;
    bool hasCamera = false;
    CameraParams camera;
    EnvData env;
    std::vector<Prototype> protos;
    std::vector<Instance> instances;
    // the reference's instance shadow test: the world ray against the
    // object-space prototype (TransformedSurfaceObject::testVisibility,
    // libSLR/Core/surface_object.cpp:363-369)
    bool compatInstanceShadow = false;
    bool compatTangentNee = false;   // volumetric integrator: NaN for a light sample in the shading plane
    MediumData medium;
    Accel accel;   // world level: world-space triangles and instances, refs as above
    bool wantStdBvh = false;   // also build the binned-SAH binary BVH (ray-batch queries)
    StdBvh stdBvh;
};

struct DeviceScene;
struct DeviceFilm;

struct TraceArgs {
    const float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
    float* t;
    int32_t* prim;
    float *b0, *b1;
    int32_t* inst;   // may be null
    float* vis;
};

struct RenderConfig {
    uint32_t width, height, spp, maxDepth, seed, slotsPerPixel;
    bool timeKernels;
    bool volumetric = false;   // VolumetricPTRenderer: the medium takes part
    int traceVariant = 0;   // traversal kernel variant (see traceVariant), -1: autotuned
    bool useGraph = true;
    bool parallelShade = false;
    int shadeShape = -1;
    bool tunePool = false;
};

struct RenderStats {
    double raysTraced = 0, samples = 0, kernelSeconds = 0, nonFinite = 0, iterations = 0;
    double traceSeconds = 0, traceLaunches = 0, extendRays = 0, shadowRays = 0;
    double shadeSeconds = 0, otherSeconds = 0;
    double nodesVisited = 0, trisTested = 0;   // extension rays
    double nodesShadow = 0, trisShadow = 0;    // shadow rays
    double overflow = 0;
    double shadeBlock = 0;   // workgroup size of the surface shading kernels
    double slotsPerPixel = 0;   // live paths per pixel of the wavefront
    double traceVariant = 0;   // traversal variant the render ran with
    double mediumEvents = 0, mediumShadows = 0;   // scattering events, light samples through the medium
    double mediumSeconds = 0;                     // k_medium time (with kernel timing)
    double mediumLookups = 0, mediumCells = 0, mediumRays = 0;   // k_medium work: density lookups, majorant cells, entries
    double mediumGuard = 0;                       // tracking walks stopped by the step guard
    double nodesLds = 0;                          // QBVH nodes read from the LDS copy of the tree top
    double instEntered = 0;                       // instances entered by the traversal
    // bidirectional integrator: light and eye subpath vertices, connections
    // traced (s >= 1, t >= 2), light-to-lens contributions filed (t = 1),
    // subpaths stopped by the vertex caps
    double bptLightVertices = 0, bptEyeVertices = 0, bptConnections = 0, bptSplats = 0, bptCapped = 0;
};

int deviceInit(int device, std::string* err);
int deviceSelect(int device, std::string* err);   // makes `device` current on the calling thread
int uploadScene(const SceneData& sd, DeviceScene** out, std::string* err);
void freeScene(DeviceScene* ds);
size_t sceneDeviceBytes(const DeviceScene* ds);

// any == false: closest hit into t/prim/b0/b1; any == true: visibility into vis.
int traceBatch(DeviceScene* ds, const TraceArgs& a, uint32_t n, bool any, bool devicePtrs, void* stream,
               double* nodesVisited, double* trisTested, std::string* err);

// The same on device pointers without allocating: counters and stack spill
// come from a static device array, the launch is asynchronous on `stream`,
// and n is at most kNoAllocMaxRays. The caller checks beforehand that the
// scene's trees cannot overflow the traversal stack (nothing is read back).
constexpr uint32_t kNoAllocMaxRays = 1u << 30;
int traceBatchNoAlloc(DeviceScene* ds, const TraceArgs& a, uint32_t n, bool any, void* stream, std::string* err);

// The same queries through the binned-SAH binary BVH (built with the scene
// when SceneData::wantStdBvh), synchronous, host or device pointers.
int traceBatchStdBvh(DeviceScene* ds, const TraceArgs& a, uint32_t n, bool any, bool devicePtrs, void* stream,
                     std::string* err);

// mode 0: a transmittance estimate per ray, 1: a free-path distance per ray.
int mediumQuery(DeviceScene* ds, const TraceArgs& a, uint32_t n, int mode, uint32_t seed, float* out,
                bool devicePtrs, void* stream, std::string* err);

int filmCreate(uint32_t w, uint32_t h, DeviceFilm** out, std::string* err);
void filmFree(DeviceFilm* f);
int filmClear(DeviceFilm* f, std::string* err);
int filmDownload(const DeviceFilm* f, std::vector<double>* strata, std::string* err);
int filmUpload(DeviceFilm* f, const double* strata, std::string* err);   // w*h*16 doubles, replacing the sums
void filmSize(const DeviceFilm* f, uint32_t* w, uint32_t* h);

int render(DeviceScene* ds, const RenderConfig& cfg, DeviceFilm* film, RenderStats* st, std::string* err);

// BPTRenderer::render (libSLR/Renderer/BPTRenderer.cpp:25-101) on the
// device: every pixel sample is a light subpath, an eye subpath and all
// their connections, weighted by the reference's MIS; accumulates into film.
// cfg.maxDepth caps the eye subpath's vertices (0: no cap).
int renderBpt(DeviceScene* ds, const RenderConfig& cfg, const CameraParams& camera, DeviceFilm* film, RenderStats* st,
              std::string* err);

// One frame of a batch: its camera, resolution, samples, depth cap, seed
// and the film it accumulates into.
struct JobSpec {
    CameraParams camera;
    uint32_t width = 0, height = 0, spp = 0, maxDepth = 0, seed = 0;
    DeviceFilm* film = nullptr;
};
struct RenderOptions {
    uint32_t slotsPerPixel = 0;   // 0: automatic
    bool timeKernels = false;
    bool useGraph = true;         // replay recorded iteration blocks (ignored when timing kernels)
    int traceVariant = 0;         // -1: chosen by the autotuner
    bool volumetric = false;      // VolumetricPTRenderer: media scatter and attenuate
    bool parallelShade = false;   // other shading classes on forked streams beside the diffuse kernel
    int shadeShape = -1;          // shading workgroup shape index; -1: tuned (automatic) or the first
    bool tunePool = false;        // automatic slots per pixel from the pool-width pilot
};
// The variant the autotuner chose for the scene (0 before any tuning).
int tunedTraceVariant(const DeviceScene* ds);
// Renders all frames in one wavefront (one slot pool, shared kernel launches).
int renderJobs(DeviceScene* ds, const JobSpec* specs, uint32_t n, const RenderOptions& opt, RenderStats* st,
               std::string* err);

// A slot pool that a scheduler admits jobs into while it runs (serving).
// Jobs occupy planes * width * height slots from a base that is a multiple
// of kPoolAlign; job indices are below kPoolMaxJobs. Between runs of
// iteration blocks the scheduler may admit and cancel jobs and must then
// call poolRecount; a job is finished once its active count reads 0.
struct SlotPool;
constexpr uint32_t kPoolAlign = 2048;
constexpr uint32_t kPoolMaxJobs = 256;
int poolOpen(DeviceScene* ds, uint32_t capacity, SlotPool** out, std::string* err);
void poolClose(SlotPool* p);
int poolBindThread(SlotPool* p);   // makes the pool's device current on the calling thread
uint32_t poolCapacity(const SlotPool* p);
int poolAdmit(SlotPool* p, uint32_t idx, const JobSpec& spec, uint32_t planes, uint32_t base, std::string* err);
int poolCancel(SlotPool* p, uint32_t idx, std::string* err);
int poolRecount(SlotPool* p, std::string* err);
int poolRunBlock(SlotPool* p, int traceVariant, std::string* err);
const uint32_t* poolActiveCounts(const SlotPool* p);
int poolFinish(SlotPool* p, uint32_t idx, std::string* err);
int poolSamples(SlotPool* p, double* samples, double* rays, std::string* err);
int numTraceVariants();
void traceVariant(int v, int* ldsDepth, int* refill, int* fused);

}  // namespace slrh
