// Device side of the renderer: scene upload, ray-batch queries and the
// wavefront spectral path tracer.
//
// Wavefront loop (one iteration = one bounce of every live path):
//   regen   : retire finished paths into the film, start the pixel's next
//             sample (camera ray) while samples remain.
//   extend  : QBVH closest hit for the extension-ray queue; each slot is
//             appended to the shading queue of the material it hit (wave
//             ballot + one atomic per wave per queue).
//   shade_k : one kernel per queue (diffuse, conductor, dielectric, miss):
//             emission with MIS, Russian roulette, light sample for NEE
//             (appended to the shadow queue) and BSDF sample (appended to the
//             next extension queue) or retirement.
//   shadow  : QBVH any hit for the shadow queue; unoccluded NEE contributions
//             are added to the path's radiance.
// A slot owns one pixel (and one film plane when several slots share a
// pixel), so film accumulation needs no atomics: only the slot's own thread
// ever adds into its film entry, in sample order, in FP64.
//
// The estimator is PTRenderer::Job::contribution
// (libSLR/Renderer/PTRenderer.cpp:146-277) and the camera/weight of
// PTRenderer::Job::kernel (:103-144).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "bsdf.h"
#include "device_scene.h"
#include "engine.h"
#include "shading.h"
#include "traverse.h"

namespace slrh {

constexpr int kBlock = 256;
constexpr int kLdsDepth = 16;       // LDS stack entries per lane for the ray-batch kernel
constexpr int kMinLdsDepth = 8;     // smallest LDS depth among the traversal variants

// Traversal kernel variants the autotuner chooses from: LDS stack depth per
// lane and the idle-lane threshold of the persistent refill.
// wavesPerEU: occupancy the register allocation is held to (0: compiler's
// choice); treelet: nodes of the top of the tree each block copies into LDS.
struct TraceVariant { int ldsDepth, refill; bool fuse; int wavesPerEU; int treelet; };
constexpr TraceVariant kVariants[] = {{16, 16, true, 0, 256}, {16, 16, true, 0, 0},   {16, 16, true, 0, 128},
                                      {16, 16, true, 0, 512}, {8, 16, true, 0, 256},  {8, 16, true, 0, 512},
                                      {16, 16, true, 6, 256}, {16, 32, true, 0, 256}, {16, 16, false, 0, 256},
                                      {8, 16, true, 0, 128},  {16, 16, true, 0, 64}};
constexpr int kBatchTreelet = 128;
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);
// refill threshold of the ray-batch entry points
constexpr int kBatchRefill = 16;
// nodes placed breadth-first at the front of the node array (the largest
// LDS treelet a traversal variant copies)
constexpr uint32_t kBfsPrefix = 512;

static int g_numCUs = 256;

// ------------------------------------------------------------------ scratch of the allocation-free entry point
//
// traceBatchNoAlloc (slrh_intersect_noalloc / slrh_occluded_noalloc) takes
// its device memory from this array instead of allocating: words
// [0, kNoAllocCounterWords) hold the launch's counters (nodes, triangles,
// overflow as 64-bit words, then the traversal's fetch cursor), the rest the
// traversal stack spill of a grid of at most kNoAllocBlocks blocks (the
// entries below the kLdsDepth that live in LDS, kMaxStack in all). Each device
// holds its own copy, from module load to process end; every call on a
// device shares it, so calls that overlap on one device race.
constexpr uint32_t kNoAllocBlocks = 1024;
constexpr size_t kNoAllocCounterWords = 64;
constexpr size_t kNoAllocSpillWords = (size_t)(kMaxStack - kLdsDepth) * kNoAllocBlocks * kBlock;
static __device__ __align__(256) uint32_t scratch[kNoAllocCounterWords + kNoAllocSpillWords];
constexpr int kMaxDevices = 64;
// device address of `scratch` per device (the host-side name is only a
// stand-in), looked up once per device in deviceInit
static uint32_t* g_scratchPtr[kMaxDevices] = {};
static std::mutex g_scratchMu;

int deviceInit(int device, std::string* err) {
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    g_numCUs = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    if (device < kMaxDevices) {
        std::lock_guard<std::mutex> lk(g_scratchMu);
        if (!g_scratchPtr[device]) {
            void* ptr = nullptr;
            HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
            g_scratchPtr[device] = (uint32_t*)ptr;
        }
    }
    return 0;
}

int deviceSelect(int device, std::string* err) {
    HIP_TRY(hipSetDevice(device));
    return 0;
}

size_t sceneDeviceBytes(const DeviceScene* ds) { return ds->bytes; }

void cameraDerived(const CameraParams& c, uint32_t w, uint32_t h, DevCamera* d) {
    // PerspectiveCamera::PerspectiveCamera (libSLR/Camera/PerspectiveCamera.cpp:16-27)
    float opH = 2.0f * c.objDist * std::tan(c.fovY * 0.5f);
    float opW = opH * c.aspect;
    d->opHeight = opH;
    d->opWidth = opW;
    d->objDist = c.objDist;
    d->imgDist = c.imgDist;
    d->imgPlaneArea = (float)(opW * opH * std::pow(c.imgDist / c.objDist, 2));
    d->lensRadius = c.lensRadius;
    d->width = w;
    d->height = h;
    const float* m = c.l2w;   // column-major: m[col*4 + row]
    for (int r = 0; r < 3; ++r)
        for (int col = 0; col < 4; ++col) d->l2w[r * 4 + col] = m[col * 4 + r];
    // inverse transpose of the upper 3x3 for normals
    double a[3][3];
    for (int r = 0; r < 3; ++r)
        for (int col = 0; col < 3; ++col) a[r][col] = m[col * 4 + r];
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    double inv[3][3];
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
        for (int col = 0; col < 3; ++col) d->nrm[r * 3 + col] = (float)inv[col][r];
}

static int buildLightRecords(DeviceScene* ds, std::string* err);
static int buildTriShade(DeviceScene* ds, std::string* err);

// Corner-packed lattice (see DevScene::medCorners): entry i = (x, y, z)
// holds g at (x|ux, y|uy, z|uz) with u = min(l + 1, n - 1), the eight values
// calcDensity weights for a point whose lower lattice indices are (x, y, z).
__global__ void k_corner_pack(const float* g, uint32_t nx, uint32_t ny, uint32_t nz, float4* out) {
    const size_t n = (size_t)nx * ny * nz;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(i % nx), y = (uint32_t)((i / nx) % ny), z = (uint32_t)(i / ((size_t)nx * ny));
        const uint32_t ux = min(x + 1, nx - 1), uy = min(y + 1, ny - 1), uz = min(z + 1, nz - 1);
        const size_t zl = (size_t)z * ny * nx, zu = (size_t)uz * ny * nx;
        out[2 * i] = make_float4(g[zl + (size_t)nx * y + x], g[zl + (size_t)nx * y + ux], g[zl + (size_t)nx * uy + x],
                                 g[zl + (size_t)nx * uy + ux]);
        out[2 * i + 1] = make_float4(g[zu + (size_t)nx * y + x], g[zu + (size_t)nx * y + ux],
                                     g[zu + (size_t)nx * uy + x], g[zu + (size_t)nx * uy + ux]);
    }
}

int uploadScene(const SceneData& sd, DeviceScene** out, std::string* err) {
    std::unique_ptr<DeviceScene> ds(new DeviceScene());
    // With instances the device holds one node array and one reference array:
    // the world tree first (its root is node 0), then each prototype's tree
    // with its child and reference indices shifted.
    std::vector<QbvhNode> mergedNodes;
    std::vector<uint32_t> mergedRefs, protoRoot(sd.protos.size(), 0);
    const bool twoLevel = !sd.instances.empty();
    if (twoLevel) {
        mergedNodes = sd.accel.qbvh;
        mergedRefs = sd.accel.qbvhRefs;
        for (size_t k = 0; k < sd.protos.size(); ++k) {
            const Accel& pa = sd.protos[k].accel;
            const uint32_t nodeBase = (uint32_t)mergedNodes.size(), refBase = (uint32_t)mergedRefs.size();
            protoRoot[k] = nodeBase;
            for (QbvhNode q : pa.qbvh) {
                for (int c = 0; c < 4; ++c)
                    if (q.childValid[c]) q.index[c] += q.isLeaf[c] ? refBase : nodeBase;
                mergedNodes.push_back(q);
            }
            mergedRefs.insert(mergedRefs.end(), pa.qbvhRefs.begin(), pa.qbvhRefs.end());
        }
    }
    const std::vector<QbvhNode>& qnodes = twoLevel ? mergedNodes : sd.accel.qbvh;
    const std::vector<uint32_t>& qrefs = twoLevel ? mergedRefs : sd.accel.qbvhRefs;
    // QBVH nodes; the traversal addresses nodes and triangle references with
    // 32-bit byte offsets (128 B per node, 48 B per reference)
    const size_t nn = qnodes.size();
    if (nn > kMaxNodes) {
        if (err) *err = "QBVH exceeds 2^25 nodes";
        return SLRH_ERR_CAPACITY_CODE;
    }
    // device order: breadth-first from the root for the first levels (so a
    // prefix of the array is the top of the tree, which the traversal copies
    // into LDS), then the remaining nodes in build order
    std::vector<uint32_t> order, newIndex(nn, UINT32_MAX);
    order.reserve(nn);
    if (nn) {
        order.push_back(0);
        newIndex[0] = 0;
        for (size_t h = 0; h < order.size() && order.size() < kBfsPrefix; ++h) {
            const QbvhNode& q = qnodes[order[h]];
            for (int c = 0; c < 4 && order.size() < kBfsPrefix; ++c)
                if (q.childValid[c] && !q.isLeaf[c] && newIndex[q.index[c]] == UINT32_MAX) {
                    newIndex[q.index[c]] = (uint32_t)order.size();
                    order.push_back(q.index[c]);
                }
        }
        for (size_t i = 0; i < nn; ++i)
            if (newIndex[i] == UINT32_MAX) {
                newIndex[i] = (uint32_t)order.size();
                order.push_back((uint32_t)i);
            }
    }
    std::vector<float4> nodes(nn * 8);
    for (size_t k = 0; k < nn; ++k) {
        const QbvhNode& q = qnodes[order[k]];
        float4* o = &nodes[8 * k];
        // rows 0-5: min x, max x, min y, max y, min z, max z of the four children
        for (int a = 0; a < 3; ++a) {
            o[2 * a] = make_float4(q.bmin[a][0], q.bmin[a][1], q.bmin[a][2], q.bmin[a][3]);
            o[2 * a + 1] = make_float4(q.bmax[a][0], q.bmax[a][1], q.bmax[a][2], q.bmax[a][3]);
        }
        uint32_t ch[4];
        for (int c = 0; c < 4; ++c) {
            if (!q.childValid[c]) ch[c] = kNoChild;
            else if (q.isLeaf[c]) {
                if (q.count[c] > kMaxLeafCount || q.index[c] + q.count[c] > kMaxTriRefs) {
                    if (err) *err = "QBVH leaf exceeds the device encoding (count <= 31, refs < 2^24)";
                    return SLRH_ERR_CAPACITY_CODE;
                }
                ch[c] = kLeafBit | (q.count[c] << kLeafCountShift) | q.index[c];
            } else ch[c] = newIndex[q.index[c]];
        }
        ch[0] |= (uint32_t)q.topAxis << kAxisShift;
        ch[1] |= (uint32_t)q.leftAxis << kAxisShift;
        ch[2] |= (uint32_t)q.rightAxis << kAxisShift;
        float4 cw;
        memcpy(&cw, ch, 16);
        o[6] = cw;
        o[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    // triangle references in leaf order
    std::vector<float4> tris(qrefs.size() * 3);
    for (size_t r = 0; r < qrefs.size(); ++r) {
        uint32_t p = qrefs[r];
        if (p & kRefInstance) {
            // an instance: id, root of its prototype's tree, class override
            const Instance& in = sd.instances[p & ~kRefInstance];
            const uint32_t root = newIndex[protoRoot[in.proto]];
            const uint32_t cls = in.material >= 0 ? sd.materials[in.material].cls : kNoClassOverride;
            float w[3];
            const uint32_t bits[3] = {p, root, cls};
            memcpy(w, bits, sizeof(w));
            tris[3 * r + 0] = make_float4(0.0f, 0.0f, 0.0f, w[0]);
            tris[3 * r + 1] = make_float4(0.0f, 0.0f, 0.0f, w[1]);
            tris[3 * r + 2] = make_float4(0.0f, 0.0f, 0.0f, w[2]);
            continue;
        }
        const uint32_t* v = &sd.idx[3 * (size_t)p];
        const float* p0 = &sd.vpos[3 * (size_t)v[0]];
        const float* p1 = &sd.vpos[3 * (size_t)v[1]];
        const float* p2 = &sd.vpos[3 * (size_t)v[2]];
        float pid;
        memcpy(&pid, &p, 4);
        tris[3 * r + 0] = make_float4(p0[0], p0[1], p0[2], pid);
        // e1.w: shading class of the triangle's material, handed out with a hit
        const uint32_t cls = sd.materials[sd.triMat[p]].cls;
        float clsF;
        memcpy(&clsF, &cls, 4);
        tris[3 * r + 1] = make_float4(p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2], clsF);
        tris[3 * r + 2] = make_float4(p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2], 0.0f);
    }
    const size_t nv = sd.vpos.size() / 3, nt = sd.idx.size() / 3;
    std::vector<float4> vp(nv), vn(nv), vt(nv);
    for (size_t i = 0; i < nv; ++i) {
        vp[i] = make_float4(sd.vpos[3 * i], sd.vpos[3 * i + 1], sd.vpos[3 * i + 2], 0);
        vn[i] = make_float4(sd.vnrm[3 * i], sd.vnrm[3 * i + 1], sd.vnrm[3 * i + 2], 0);
        vt[i] = make_float4(sd.vtan[3 * i], sd.vtan[3 * i + 1], sd.vtan[3 * i + 2], 0);
    }
    std::vector<uint4> pv(nt);
    for (size_t i = 0; i < nt; ++i) pv[i] = make_uint4(sd.idx[3 * i], sd.idx[3 * i + 1], sd.idx[3 * i + 2], sd.triMat[i]);
    // spectra
    std::vector<DevSpectrum> specs;
    std::vector<float> pool;
    for (const HostSpectrum& h : sd.spectra) {
        DevSpectrum d;
        d.lo = h.lo; d.hi = h.hi; d.offset = (uint32_t)pool.size(); d.n = (uint32_t)h.values.size();
        pool.insert(pool.end(), h.values.begin(), h.values.end());
        specs.push_back(d);
    }
    if (pool.empty()) pool.push_back(0.0f);
    while (pool.size() % 4) pool.push_back(0.0f);   // whole float4s for the shading kernels' LDS copy
    // per material the headers of its spectra, read beside the material so
    // evaluating one needs no lookup of the header by id
    std::vector<DevSpectrum> matSpec(std::max<size_t>(1, 3 * sd.materials.size()));
    for (size_t m = 0; m < sd.materials.size(); ++m) {
        const int32_t ids[3] = {sd.materials[m].spec0, sd.materials[m].spec1, sd.materials[m].emit};
        for (int j = 0; j < 3; ++j) {
            DevSpectrum h{0.0f, 0.0f, 0u, 0u};
            if (ids[j] >= 0 && (size_t)ids[j] < specs.size()) h = specs[ids[j]];
            matSpec[3 * m + j] = h;
        }
    }
    // lights: every emitting triangle with importance 1, in object order
    // (SurfaceObjectAggregate ctor, libSLR/Core/surface_object.cpp:379-398):
    // the world-space ones, then per instance its prototype's (an instance
    // carries its prototype's importance, TransformedSurfaceObject::importance,
    // surface_object.cpp:298-300)
    std::vector<uint2> lights;
    std::vector<int32_t> primToLight(nt, -1);
    for (size_t i = 0; i < nt; ++i) {
        const DevMaterial& m = sd.materials[sd.triMat[i]];
        if (sd.triProto[i] < 0 && m.emit >= 0) {
            primToLight[i] = (int32_t)lights.size();
            lights.push_back(make_uint2((uint32_t)i, kNoInstance));
        }
    }
    for (size_t k = 0; k < sd.instances.size(); ++k) {
        const Instance& in = sd.instances[k];
        const Prototype& pr = sd.protos[in.proto];
        for (uint32_t i = pr.triFirst; i < pr.triFirst + pr.triCount; ++i) {
            const uint32_t mat = in.material >= 0 ? (uint32_t)in.material : sd.triMat[i];
            if (sd.materials[mat].emit >= 0) lights.push_back(make_uint2(i, (uint32_t)k));
        }
    }
    // per instance: object-to-world rows, world-to-object rows; material override
    std::vector<float4> instXf(std::max<size_t>(1, kInstXfRows * sd.instances.size()), make_float4(0, 0, 0, 0));
    std::vector<int32_t> instMat(std::max<size_t>(1, sd.instances.size()), -1);
    for (size_t k = 0; k < sd.instances.size(); ++k) {
        const Instance& in = sd.instances[k];
        for (int r = 0; r < 3; ++r) {
            instXf[kInstXfRows * k + r] =
                make_float4(in.l2w[4 * r], in.l2w[4 * r + 1], in.l2w[4 * r + 2], in.l2w[4 * r + 3]);
        }
        for (int r = 0; r < 4; ++r)
            instXf[kInstXfRows * k + 3 + r] =
                make_float4(in.w2l[4 * r], in.w2l[4 * r + 1], in.w2l[4 * r + 2], in.w2l[4 * r + 3]);
        instMat[k] = in.material;
    }
    // DiscreteDistribution1D over unit importances with a compensated sum
    // (libSLR/Core/distributions.cpp:99-123)
    std::vector<float> cdf(lights.size() + 1), pmf(lights.size());
    {
        float sum = 0, comp = 0;
        cdf[0] = 0;
        for (size_t i = 0; i < lights.size(); ++i) {
            float in = 1.0f - comp, tmp = sum + in;
            comp = (tmp - sum) - in;
            sum = tmp;
            cdf[i + 1] = sum;
        }
        for (size_t i = 0; i < lights.size(); ++i) { pmf[i] = 1.0f / sum; cdf[i + 1] /= sum; }
        ds->view.lightImportance = lights.empty() ? 0.0f : sum;
    }
    if (cdf.empty()) cdf.push_back(0);
    if (pmf.empty()) pmf.push_back(0);
    if (lights.empty()) lights.push_back(make_uint2(0, kNoInstance));
    int r;
    if ((r = ds->nodes.upload(nodes, err)) || (r = ds->tris.upload(tris, err)) || (r = ds->vpos.upload(vp, err)) ||
        (r = ds->vnrm.upload(vn, err)) || (r = ds->vtan.upload(vt, err)) || (r = ds->primVerts.upload(pv, err)) ||
        (r = ds->mats.upload(sd.materials, err)) || (r = ds->spectra.upload(specs, err)) ||
        (r = ds->specPool.upload(pool, err)) || (r = ds->lights.upload(lights, err)) ||
        (r = ds->lightCdf.upload(cdf, err)) || (r = ds->lightPmf.upload(pmf, err)) ||
        (r = ds->primToLight.upload(primToLight, err)) || (r = ds->instXf.upload(instXf, err)) ||
        (r = ds->instMat.upload(instMat, err)) || (r = ds->matSpec.upload(matSpec, err)))
        return r;
    ds->matTypeMask = 0;
    for (const DevMaterial& m : sd.materials) ds->matTypeMask |= 1u << m.type;
    if ((r = ds->counters.alloc(16, err))) return r;
    DevScene& ve = ds->view;
    ve.hasEnv = sd.env.present ? 1u : 0u;
    ve.envTexel = nullptr;
    ve.mengBasis = nullptr;
    ve.envCondPdf = ve.envCondCdf = ve.envMargPdf = ve.envMargCdf = nullptr;
    ve.envW = ve.envH = ve.envMapW = ve.envMapH = 0;
    ve.envScale = 0.0f;
    if (sd.env.present) {
        const EnvData& e = sd.env;
        std::vector<float4> tex(e.texels.size());
        for (size_t i = 0; i < tex.size(); ++i) {
            float adj;
            memcpy(&adj, &e.texels[i].adj, 4);
            tex[i] = make_float4(adj, e.texels[i].s, e.texels[i].t, e.texels[i].scale);
        }
        uint32_t np = 0, nl = 0;
        const float* basis = mengBasisTable(&np, &nl);
        // each bin with its upper neighbour (the last bin with itself): the
        // linear interpolation between two bins is then one 8-byte load
        std::vector<float2> b((size_t)np * nl);
        for (uint32_t q = 0; q < np; ++q)
            for (uint32_t i = 0; i < nl; ++i)
                b[(size_t)q * nl + i] = make_float2(basis[(size_t)q * nl + i], basis[(size_t)q * nl + std::min(i + 1, nl - 1)]);
        if ((r = ds->envTexel.upload(tex, err)) || (r = ds->mengBasis.upload(b, err)) ||
            (r = ds->envCondPdf.upload(e.condPdf, err)) || (r = ds->envCondCdf.upload(e.condCdf, err)) ||
            (r = ds->envMargPdf.upload(e.margPdf, err)) || (r = ds->envMargCdf.upload(e.margCdf, err)))
            return r;
        ve.envTexel = ds->envTexel.p;
        ve.mengBasis = ds->mengBasis.p;
        ve.envCondPdf = ds->envCondPdf.p;
        ve.envCondCdf = ds->envCondCdf.p;
        ve.envMargPdf = ds->envMargPdf.p;
        ve.envMargCdf = ds->envMargCdf.p;
        ve.envW = e.width;
        ve.envH = e.height;
        ve.envMapW = e.mapW;
        ve.envMapH = e.mapH;
        ve.envScale = e.scale;
    }
    ve.medCorners = nullptr;
    ve.medMajorant = nullptr;
    ve.hasMedium = 0;
    ve.volume = 0;
    if (sd.medium.present) {
        static_assert(kMedSuperVoxel == kSuperVoxel, "supervoxel size");
        const MediumData& m = sd.medium;
        if ((r = ds->medMajorant.upload(m.majorant, err))) return r;
        {
            // the lattice goes up as given and is turned into the
            // corner-packed form on the device; the plain copy is dropped
            DevBuf<float> lattice;
            if ((r = lattice.upload(m.density, err))) return r;
            const size_t np = (size_t)m.n[0] * m.n[1] * m.n[2];
            if ((r = ds->medCorners.alloc(2 * np, err))) return r;
            hipLaunchKernelGGL(k_corner_pack, dim3(2048), dim3(256), 0, 0, lattice.p, m.n[0], m.n[1], m.n[2],
                               ds->medCorners.p);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipDeviceSynchronize());
        }
        ve.medCorners = ds->medCorners.p;
        ve.medMajorant = ds->medMajorant.p;
        for (int a = 0; a < 3; ++a) {
            ve.medMin[a] = m.bmin[a];
            ve.medMax[a] = m.bmax[a];
            ve.medN[a] = m.n[a];
            ve.medCells[a] = m.cells[a];
            // the device's own IEEE quotients, computed once
            ve.medScale[a] = (float)(m.n[a] - 1) / (m.bmax[a] - m.bmin[a]);
            ve.medInvExt[a] = 1.0f / (m.bmax[a] - m.bmin[a]);
        }
        ve.medSigmaE = m.sigmaEValue;
        ve.medSigmaS = m.sigmaS;
        ve.medG = m.g;
        ve.hasMedium = 1;
    }
    ds->bytes = nodes.size() * 16 + tris.size() * 16 + (vp.size() * 3) * 16 + pv.size() * 16 + pool.size() * 4;
    DevScene& v = ds->view;
    v.nodes = ds->nodes.p; v.tris = ds->tris.p; v.primVerts = ds->primVerts.p;
    v.vpos = ds->vpos.p; v.vnrm = ds->vnrm.p; v.vtan = ds->vtan.p;
    v.mats = ds->mats.p; v.spectra = ds->spectra.p; v.specPool = ds->specPool.p;
    v.matSpec = ds->matSpec.p;
    v.lights = ds->lights.p; v.lightCdf = ds->lightCdf.p; v.lightPmf = ds->lightPmf.p;
    v.primToLight = ds->primToLight.p;
    v.instXf = ds->instXf.p;
    v.instMat = ds->instMat.p;
    v.numInstances = (uint32_t)sd.instances.size();
    v.compatInstShadow = sd.compatInstanceShadow ? 1u : 0u;
    v.compatTangentNee = sd.compatTangentNee ? 1u : 0u;
    v.tangentNeeNaN = 0;
    v.numLights = (uint32_t)(cdf.size() - 1);
    v.numPrims = (uint32_t)nt;
    v.numNodes = (uint32_t)nn;
    v.classMask = 1u << CLS_MISS;
    for (const DevMaterial& m : sd.materials) v.classMask |= 1u << m.cls;
    {
        const size_t words = pool.size() + 4 * matSpec.size();
        v.specLdsPool = words <= kShadeLdsWords ? (uint32_t)pool.size() : 0u;
        v.specLdsWords = words <= kShadeLdsWords ? (uint32_t)words : 0u;
    }
    ds->hasCamera = sd.hasCamera;
    ds->camera = sd.camera;
    if (sd.hasCamera) cameraDerived(sd.camera, 1, 1, &v.cam);
    if (int rl = buildLightRecords(ds.get(), err)) return rl;
    if (int rt = buildTriShade(ds.get(), err)) return rt;
    if (sd.wantStdBvh && !sd.stdBvh.nodes.empty()) {
        // node rows: min xyz + (first child | first leaf ref), max xyz + (second
        // child | leaf count), {split axis, leaf flag}
        const StdBvh& b = sd.stdBvh;
        std::vector<float4> bn(3 * b.nodes.size()), bt(3 * std::max<size_t>(1, b.refs.size()));
        for (size_t k = 0; k < b.nodes.size(); ++k) {
            const StdBvhNode& q = b.nodes[k];
            const bool leaf = q.count > 0 || (q.c0 == 0 && q.c1 == 0);
            const uint32_t w0 = leaf ? q.first : q.c0, w1 = leaf ? q.count : q.c1;
            bn[3 * k] = make_float4(q.box.mn[0], q.box.mn[1], q.box.mn[2], __builtin_bit_cast(float, w0));
            bn[3 * k + 1] = make_float4(q.box.mx[0], q.box.mx[1], q.box.mx[2], __builtin_bit_cast(float, w1));
            bn[3 * k + 2] = make_float4(__builtin_bit_cast(float, (uint32_t)q.axis),
                                        __builtin_bit_cast(float, leaf ? 1u : 0u), 0.0f, 0.0f);
        }
        for (size_t r = 0; r < b.refs.size(); ++r) {
            const uint32_t p = b.refs[r];
            const uint32_t* v = &sd.idx[3 * (size_t)p];
            const float* p0 = &sd.vpos[3 * (size_t)v[0]];
            const float* p1 = &sd.vpos[3 * (size_t)v[1]];
            const float* p2 = &sd.vpos[3 * (size_t)v[2]];
            const uint32_t cls = sd.materials[sd.triMat[p]].cls;
            bt[3 * r] = make_float4(p0[0], p0[1], p0[2], __builtin_bit_cast(float, p));
            bt[3 * r + 1] = make_float4(p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2], __builtin_bit_cast(float, cls));
            bt[3 * r + 2] = make_float4(p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2], 0.0f);
        }
        if ((r = ds->bvh2Nodes.upload(bn, err)) || (r = ds->bvh2Tris.upload(bt, err))) return r;
    }
    *out = ds.release();
    return 0;
}

void freeScene(DeviceScene* ds) { delete ds; }

// ------------------------------------------------------------------ wave helpers

// ------------------------------------------------------------------ ray batches

template <bool I>
__global__ __launch_bounds__(kBlock) void k_trace_batch(DevScene sc, TraceArgs a, uint32_t n, int any,
                                                        uint32_t* spill, uint32_t spillStride,
                                                        unsigned long long* counters) {
    __shared__ uint32_t lds[kLdsDepth * kBlock];
    __shared__ float4 treelet[kBatchTreelet * kTreeletRows];
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    TraceStack st;
    st.lds = ldsColumn(lds, threadIdx.x);
    st.ldsStride = kBlock;
    st.spill = spill + gtid;
    st.spillStride = spillStride;
    st.ldsDepth = kLdsDepth;
    st.treeletNodes = (uint32_t)kBatchTreelet < sc.numNodes ? (uint32_t)kBatchTreelet : sc.numNodes;
    st.treelet = loadTreelet(treelet, sc.nodes, st.treeletNodes);
    TraceCounters tc;
    uint32_t* fetchCtr = (uint32_t*)&counters[3];
    const bool anyHit = any != 0;
    traceLoop<kBatchRefill, true, false, I>(sc, st, n, fetchCtr, [](uint32_t i) { return i; },
        [&](uint32_t i, RayIn& r) {
            r.o = make_float4(a.ox[i], a.oy[i], a.oz[i], a.tmin[i]);
            r.d = make_float4(a.dx[i], a.dy[i], a.dz[i], a.tmax[i]);
            return anyHit;
        },
        [&](uint32_t i, bool isAny, const HitOut& h) {
            if (isAny) {
                a.vis[i] = h.prim >= 0 ? 0.0f : 1.0f;
            } else {
                a.t[i] = h.t;
                a.prim[i] = h.prim;
                a.b0[i] = h.b0;
                a.b1[i] = h.b1;
                if (a.inst) a.inst[i] = h.inst;
            }
        }, tc);
    waveAdd64(&counters[0], tc.nodes);
    waveAdd64(&counters[1], tc.tris);
    if (tc.overflow) atomicOr((unsigned int*)&counters[2], 1u);
}

static uint32_t persistentBlocks(uint32_t work) {
    uint32_t maxBlocks = (uint32_t)g_numCUs * 8;
    uint32_t need = (work + kBlock - 1) / kBlock;
    return std::max(1u, std::min(maxBlocks, need));
}

int traceBatch(DeviceScene* ds, const TraceArgs& a, uint32_t n, bool any, bool devicePtrs, void* stream,
               double* nodesVisited, double* trisTested, std::string* err) {
    if (n == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    TraceArgs d = a;
    std::vector<void*> temps;
    auto cleanup = [&]() { for (void* p : temps) (void)hipFree(p); };
    if (!devicePtrs) {
        // stage host arrays through device memory
        const float* hin[8] = {a.ox, a.oy, a.oz, a.dx, a.dy, a.dz, a.tmin, a.tmax};
        const float** din[8] = {&d.ox, &d.oy, &d.oz, &d.dx, &d.dy, &d.dz, &d.tmin, &d.tmax};
        for (int k = 0; k < 8; ++k) {
            void* p = nullptr;
            if (hipMalloc(&p, n * sizeof(float)) != hipSuccess) { cleanup(); if (err) *err = "hipMalloc"; return SLRH_ERR_NOMEM_CODE; }
            temps.push_back(p);
            (void)hipMemcpyAsync(p, hin[k], n * sizeof(float), hipMemcpyHostToDevice, s);
            *din[k] = (const float*)p;
        }
        int nout = any ? 1 : (a.inst ? 5 : 4);
        for (int k = 0; k < nout; ++k) {
            void* p = nullptr;
            if (hipMalloc(&p, n * sizeof(float)) != hipSuccess) { cleanup(); if (err) *err = "hipMalloc"; return SLRH_ERR_NOMEM_CODE; }
            temps.push_back(p);
        }
        if (any) d.vis = (float*)temps[8];
        else {
            d.t = (float*)temps[8]; d.prim = (int32_t*)temps[9]; d.b0 = (float*)temps[10]; d.b1 = (float*)temps[11];
            d.inst = a.inst ? (int32_t*)temps[12] : nullptr;
        }
    }
    uint32_t blocks = persistentBlocks(n);
    size_t spillNeed = (size_t)(kMaxStack - kLdsDepth) * blocks * kBlock;
    if (ds->spill.n < spillNeed) {
        int r = ds->spill.alloc(spillNeed, err);
        if (r) { cleanup(); return r; }
    }
    (void)hipMemsetAsync(ds->counters.p, 0, 4 * sizeof(unsigned long long), s);
    if (ds->view.numInstances)
        hipLaunchKernelGGL(k_trace_batch<true>, dim3(blocks), dim3(kBlock), 0, s, ds->view, d, n, any ? 1 : 0,
                           ds->spill.p, blocks * kBlock, ds->counters.p);
    else
        hipLaunchKernelGGL(k_trace_batch<false>, dim3(blocks), dim3(kBlock), 0, s, ds->view, d, n, any ? 1 : 0,
                           ds->spill.p, blocks * kBlock, ds->counters.p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { cleanup(); if (err) *err = std::string("k_trace_batch: ") + hipGetErrorString(e); return SLRH_ERR_DEVICE_CODE; }
    unsigned long long cnt[4] = {0, 0, 0, 0};
    (void)hipMemcpyAsync(cnt, ds->counters.p, sizeof(cnt), hipMemcpyDeviceToHost, s);
    if (!devicePtrs) {
        if (any) (void)hipMemcpyAsync(a.vis, d.vis, n * sizeof(float), hipMemcpyDeviceToHost, s);
        else {
            (void)hipMemcpyAsync(a.t, d.t, n * sizeof(float), hipMemcpyDeviceToHost, s);
            (void)hipMemcpyAsync(a.prim, d.prim, n * sizeof(int32_t), hipMemcpyDeviceToHost, s);
            (void)hipMemcpyAsync(a.b0, d.b0, n * sizeof(float), hipMemcpyDeviceToHost, s);
            (void)hipMemcpyAsync(a.b1, d.b1, n * sizeof(float), hipMemcpyDeviceToHost, s);
            if (a.inst) (void)hipMemcpyAsync(a.inst, d.inst, n * sizeof(int32_t), hipMemcpyDeviceToHost, s);
        }
    }
    e = hipStreamSynchronize(s);
    cleanup();
    if (e != hipSuccess) { if (err) *err = std::string("trace sync: ") + hipGetErrorString(e); return SLRH_ERR_DEVICE_CODE; }
    if (nodesVisited) *nodesVisited = (double)cnt[0];
    if (trisTested) *trisTested = (double)cnt[1];
    if (cnt[2]) { if (err) *err = "traversal stack overflow (depth > 64)"; return SLRH_ERR_CAPACITY_CODE; }
    return 0;
}

// Ray batches through the binned-SAH binary BVH: StandardBVH::commonProcedure
// (libSLR/Accelerator/StandardBVH.h:265-293) per ray: a stack of node
// indices from the root, each popped node's box tested against the current
// segment with BoundingBox3D::intersect (libSLR/BasicTypes/BoundingBox3D.h:
// 100-116: near/far per axis from the reciprocal direction, swapped when
// reversed, the segment narrowed and given up when empty), an internal
// node's children pushed far then near (near = the first child when the ray
// points along the positive split axis), a leaf's triangles tested in order
// with the closest hit narrowing the segment, or, for visibility, the first
// hit ending the ray. The stack lives in LDS, [depth][lane]; the reference's
// depth of 64.
constexpr int kBvh2Block = 128;
constexpr int kBvh2Stack = 64;

__global__ __launch_bounds__(kBvh2Block) void k_trace_bvh2(DevScene sc, const float4* nodes, const float4* tris,
                                                          TraceArgs a, uint32_t n, int any,
                                                          unsigned long long* counters) {
    __shared__ uint32_t stack[kBvh2Stack * kBvh2Block];
    uint32_t* st = stack + threadIdx.x;
    unsigned long long nodeCount = 0, triCount = 0;
    uint32_t overflow = 0;
    for (uint32_t i = blockIdx.x * kBvh2Block + threadIdx.x; i < n; i += gridDim.x * kBvh2Block) {
        RayIn r;
        r.o = make_float4(a.ox[i], a.oy[i], a.oz[i], a.tmin[i]);
        r.d = make_float4(a.dx[i], a.dy[i], a.dz[i], a.tmax[i]);
        const float inv[3] = {1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z};   // Vector3D::reciprocal
        const float org[3] = {r.o.x, r.o.y, r.o.z};
        const bool pos[3] = {r.d.x >= 0.0f, r.d.y >= 0.0f, r.d.z >= 0.0f};
        const float tmin = r.o.w;
        float tmax = r.d.w, hb1 = 0.0f, hb2 = 0.0f;
        int32_t prim = -1;
        int sp = 0;
        st[0] = 0;
        sp = 1;
        bool done = false;
        while (sp > 0 && !done) {
            const uint32_t ni = st[--sp * kBvh2Block];
            const float4 lo = nodes[3 * ni], hi = nodes[3 * ni + 1], info = nodes[3 * ni + 2];
            ++nodeCount;
            const float mn[3] = {lo.x, lo.y, lo.z}, mx[3] = {hi.x, hi.y, hi.z};
            // BoundingBox3D::isValid, then the slab test
            if (!(mx[0] - mn[0] >= 0 && mx[1] - mn[1] >= 0 && mx[2] - mn[2] >= 0)) continue;
            float d0 = tmin, d1 = tmax;
            bool hit = true;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                float tn = (mn[k] - org[k]) * inv[k], tf = (mx[k] - org[k]) * inv[k];
                if (tn > tf) {
                    const float x = tn;
                    tn = tf;
                    tf = x;
                }
                d0 = tn > d0 ? tn : d0;
                d1 = tf < d1 ? tf : d1;
                if (d0 > d1) hit = false;
            }
            if (!hit) continue;
            const uint32_t w0 = __float_as_uint(lo.w), w1 = __float_as_uint(hi.w);
            if (__float_as_uint(info.y) == 0u) {
                const uint32_t axis = __float_as_uint(info.x);
                const bool positive = axis == 0 ? pos[0] : axis == 1 ? pos[1] : pos[2];
                if (sp + 2 > kBvh2Stack) { overflow = 1; break; }
                st[sp++ * kBvh2Block] = positive ? w1 : w0;
                st[sp++ * kBvh2Block] = positive ? w0 : w1;
            } else {
                for (uint32_t k = 0; k < w1; ++k) {
                    const float4 v0 = tris[3 * (w0 + k)], e1 = tris[3 * (w0 + k) + 1], e2 = tris[3 * (w0 + k) + 2];
                    ++triCount;
                    float t, b1, b2;
                    if (triTest(v0, e1, e2, r, tmin, tmax, &t, &b1, &b2)) {
                        tmax = t;
                        prim = (int32_t)__float_as_uint(v0.w);
                        hb1 = b1;
                        hb2 = b2;
                        if (any) {
                            done = true;
                            break;
                        }
                    }
                }
            }
        }
        if (any) {
            a.vis[i] = prim >= 0 ? 0.0f : 1.0f;
        } else {
            a.t[i] = tmax;
            a.prim[i] = prim;
            a.b0[i] = prim >= 0 ? 1.0f - hb1 - hb2 : 0.0f;
            a.b1[i] = prim >= 0 ? hb1 : 0.0f;
            if (a.inst) a.inst[i] = -1;
        }
    }
    waveAdd64(&counters[0], nodeCount);
    waveAdd64(&counters[1], triCount);
    if (overflow) atomicOr((unsigned int*)&counters[2], 1u);
}

int traceBatchStdBvh(DeviceScene* ds, const TraceArgs& a, uint32_t n, bool any, bool devicePtrs, void* stream,
                     std::string* err) {
    if (n == 0) return 0;
    if (!ds->bvh2Nodes.p) {
        if (err) *err = "the scene was built without the binned-SAH BVH";
        return SLRH_ERR_STATE_CODE;
    }
    hipStream_t s = (hipStream_t)stream;
    TraceArgs d = a;
    std::vector<DevBuf<float>> temps(13);
    if (!devicePtrs) {
        const float* hin[8] = {a.ox, a.oy, a.oz, a.dx, a.dy, a.dz, a.tmin, a.tmax};
        const float** din[8] = {&d.ox, &d.oy, &d.oz, &d.dx, &d.dy, &d.dz, &d.tmin, &d.tmax};
        for (int k = 0; k < 8; ++k) {
            if (int r = temps[k].alloc(n, err)) return r;
            HIP_TRY(hipMemcpyAsync(temps[k].p, hin[k], n * sizeof(float), hipMemcpyHostToDevice, s));
            *din[k] = temps[k].p;
        }
        const int nout = any ? 1 : (a.inst ? 5 : 4);
        for (int k = 0; k < nout; ++k)
            if (int r = temps[8 + k].alloc(n, err)) return r;
        if (any) {
            d.vis = temps[8].p;
        } else {
            d.t = temps[8].p;
            d.prim = (int32_t*)temps[9].p;
            d.b0 = temps[10].p;
            d.b1 = temps[11].p;
            d.inst = a.inst ? (int32_t*)temps[12].p : nullptr;
        }
    }
    HIP_TRY(hipMemsetAsync(ds->counters.p, 0, 4 * sizeof(unsigned long long), s));
    const uint32_t blocks = std::max(1u, std::min((uint32_t)g_numCUs * 8, (n + kBvh2Block - 1) / kBvh2Block));
    hipLaunchKernelGGL(k_trace_bvh2, dim3(blocks), dim3(kBvh2Block), 0, s, ds->view, ds->bvh2Nodes.p,
                       ds->bvh2Tris.p, d, n, any ? 1 : 0, ds->counters.p);
    HIP_TRY(hipGetLastError());
    unsigned long long cnt[4] = {0, 0, 0, 0};
    HIP_TRY(hipMemcpyAsync(cnt, ds->counters.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
    if (!devicePtrs) {
        if (any) {
            HIP_TRY(hipMemcpyAsync(a.vis, d.vis, n * sizeof(float), hipMemcpyDeviceToHost, s));
        } else {
            HIP_TRY(hipMemcpyAsync(a.t, d.t, n * sizeof(float), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipMemcpyAsync(a.prim, d.prim, n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipMemcpyAsync(a.b0, d.b0, n * sizeof(float), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipMemcpyAsync(a.b1, d.b1, n * sizeof(float), hipMemcpyDeviceToHost, s));
            if (a.inst) HIP_TRY(hipMemcpyAsync(a.inst, d.inst, n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        }
    }
    HIP_TRY(hipStreamSynchronize(s));
    if (cnt[2]) {
        if (err) *err = "binary BVH traversal stack overflow (depth > 64)";
        return SLRH_ERR_CAPACITY_CODE;
    }
    return 0;
}

// Fills n words with zero (the allocation-free path clears its counters with
// a kernel, not a memset node, so that it can be captured in a graph).
__global__ void k_zero_words(uint32_t* p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

int traceBatchNoAlloc(DeviceScene* ds, const TraceArgs& a, uint32_t n, bool any, void* stream, std::string* err) {
    if (n == 0) return 0;
    if (n > kNoAllocMaxRays) {
        if (err) *err = "allocation-free ray batch larger than 2^30 rays";
        return SLRH_ERR_CAPACITY_CODE;
    }
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    uint32_t* ptr = dev >= 0 && dev < kMaxDevices ? g_scratchPtr[dev] : nullptr;
    if (!ptr) {
        if (err) *err = "no scratch set up on the current device (create a context on it first)";
        return SLRH_ERR_STATE_CODE;
    }
    hipStream_t s = (hipStream_t)stream;
    const uint32_t blocks = std::min(persistentBlocks(n), kNoAllocBlocks);
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, ptr, (uint32_t)kNoAllocCounterWords);
    unsigned long long* counters = (unsigned long long*)ptr;
    uint32_t* spill = ptr + kNoAllocCounterWords;
    if (ds->view.numInstances)
        hipLaunchKernelGGL(k_trace_batch<true>, dim3(blocks), dim3(kBlock), 0, s, ds->view, a, n, any ? 1 : 0, spill,
                           blocks * kBlock, counters);
    else
        hipLaunchKernelGGL(k_trace_batch<false>, dim3(blocks), dim3(kBlock), 0, s, ds->view, a, n, any ? 1 : 0, spill,
                           blocks * kBlock, counters);
    HIP_TRY(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------------ films

int filmCreate(uint32_t w, uint32_t h, DeviceFilm** out, std::string* err) {
    std::unique_ptr<DeviceFilm> f(new DeviceFilm());
    f->w = w; f->h = h;
    int r = f->data.alloc((size_t)w * h * kNumL, err);
    if (r) return r;
    HIP_TRY(hipMemset(f->data.p, 0, (size_t)w * h * kNumL * sizeof(double)));
    HIP_TRY(hipDeviceSynchronize());
    *out = f.release();
    return 0;
}
void filmFree(DeviceFilm* f) { delete f; }
int filmClear(DeviceFilm* f, std::string* err) {
    HIP_TRY(hipMemset(f->data.p, 0, (size_t)f->w * f->h * kNumL * sizeof(double)));
    HIP_TRY(hipDeviceSynchronize());
    return 0;
}
int filmDownload(const DeviceFilm* f, std::vector<double>* strata, std::string* err) {
    strata->resize((size_t)f->w * f->h * kNumL);
    HIP_TRY(hipMemcpy(strata->data(), f->data.p, strata->size() * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}
int filmUpload(DeviceFilm* f, const double* strata, std::string* err) {
    HIP_TRY(hipMemcpy(f->data.p, strata, (size_t)f->w * f->h * kNumL * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}
void filmSize(const DeviceFilm* f, uint32_t* w, uint32_t* h) { *w = f->w; *h = f->h; }

// ------------------------------------------------------------------ path state

// Counter layout (uint32 each).
enum Ctr : uint32_t {
    C_TRACE = 0,          // trace queue length (extension and shadow entries)
    C_DONE = 1,           // done queue length
    C_CLS0 = 3,           // .. C_CLS0 + kNumClasses - 1: shading queue lengths
    C_PENDING = 8,        // slots with an extension or shadow ray pending after shading
    C_NONFINITE = 9,
    C_OVERFLOW = 10,
    kNumCtr = 16
};
// 64-bit statistics
enum Stat64 : uint32_t {
    S_RAYS_EXT = 0, S_RAYS_SHADOW = 1, S_NODES_EXT = 2, S_TRIS_EXT = 3, S_SAMPLES = 4, S_NODES_SH = 5, S_TRIS_SH = 6,
    S_MED_EVENTS = 7,     // scattering events in the medium
    S_MED_SHADOWS = 8,    // light samples whose segment crossed the medium
    S_MED_LOOKUPS = 9,    // tentative collisions (density lookups) in k_medium
    S_MED_CELLS = 10,     // majorant cells entered in k_medium
    S_MED_RAYS = 11,      // trace entries k_medium tracked through the medium
    S_MED_GUARD = 12,     // tracking walks stopped by the kMaxTrackSteps guard
    S_NODES_LDS = 13,     // QBVH nodes the traversal read from its LDS copy of the tree top
    S_INSTS = 14,         // instances entered by the traversal
    kNumStat = 15
};
// Per-slot work code, one byte per slot, rewritten by each stage and turned
// into slot-ordered queues by k_compact. After shading a slot holds EXT or
// DONE, plus SHADOW when it also emitted a shadow ray; the trace kernel turns
// EXT into CLS0 + shading class.
enum Code : uint8_t { CODE_IDLE = 0, CODE_EXT = 1, CODE_DONE = 2, CODE_CLS0 = 3, CODE_SHADOW = 0x80 };
// trace queue entries: slot id, bit 31 set for the slot's shadow ray
constexpr uint32_t kShadowEntry = 0x80000000u;

// flags in info.y
constexpr uint32_t F_HAS_SAMPLE = 1u;    // slot holds a sample whose radiance must be filed
constexpr uint32_t F_LAST_DELTA = 2u;    // last BSDF sample was a delta lobe
constexpr uint32_t F_RAD = 4u;           // SF_SP holds the path's radiance so far (zero when clear)

// Per-slot path record, 20 float4 (320 B) in three arrays, each entry
// aligned to its size so a slot's fields share cache lines (each kernel works
// on a sparse, slot-ordered subset of the slots; per-field arrays would fetch
// a whole line per 16-B field):
//   ray  (128 B, one line): everything the traversal and the medium walk
//        read and write: ray, hit, shadow direction, path info, misc, RNG,
//        job words
//   spec (128 B, one line): throughput alpha and radiance so far (16
//        wavelengths each)
//   nee  (64 B): the pending light-sample contribution
// A shadow ray starts where the extension ray does (the shading point, tmin
// 1e-4), so only its direction and length are stored; the traversal forms
// both rays' reciprocal directions itself.
enum SlotField : uint32_t {
    SF_RO = 0,      // ray origin.xyz + tmin (also the shadow ray's)
    SF_RD = 1,      // ray direction.xyz + tmax (infinite)
    SF_HIT = 2,     // t, prim, b0, b1
    SF_SD = 3,      // shadow ray direction.xyz + tmax
    SF_INFO = 4,    // pixel, flags | hero << 8 | pathLength << 16, samples started, samples total
    SF_MISC = 5,    // wl offset, initY, last BSDF pdf, camera weight
    SF_RNG = 6,     // XORShift state
    SF_JOB = 7,     // job index, film plane of the slot, seed of its light-sample transmittance
                    // streams, instance of the last hit (scenes with instances)
    SF_ALPHA = 8,   // 4 float4 (spec array)
    SF_SP = 12,     // 4 float4 (spec array)
    SF_NEE = 16,    // 4 float4 (nee array)
};
constexpr uint32_t kRayVec = 8, kSpecVec = 8, kNeeVec = 4;
constexpr uint32_t kSlotRecordBytes = 16 * (kRayVec + kSpecVec + kNeeVec);

// One render job (a frame or a served request) in the slot pool: its camera
// (resolution included), its film planes and its path-length cap. Jobs own
// contiguous slot ranges; slot base + plane * pixels + pixel.
struct DevJob {
    DevCamera cam;
    double* film;        // [planes][pixels][16]
    uint32_t pixels, planes, maxDepth, slotBase;
};
constexpr uint32_t kMaxJobs = 256;
constexpr uint32_t kNoJob = 0xFFFFFFFFu;
constexpr uint32_t kCompactSlots = 2048;   // slots per compaction block; jobs start at multiples of it
static_assert(kCompactSlots == kPoolAlign && kMaxJobs == kPoolMaxJobs, "pool constants");

struct PathState {
    uint32_t n;          // slots
    float4* ray;         // n x kRayVec: SF_RO .. SF_JOB
    float4* spec;        // n x kSpecVec: SF_ALPHA, SF_SP
    float4* nee;         // n x kNeeVec: SF_NEE
    const DevJob* jobs;  // kMaxJobs entries
    uint32_t* jobActive; // per job: slots not idle, counted by the post-shade compaction
    const uint32_t* blockJob;   // job owning compaction block b (slots 2048 b ..), kNoJob if none
    uint8_t* code;       // per-slot work code (Code)
    uint8_t* vis;        // 1: the slot's shadow ray was traced and is unoccluded
    uint32_t* q;         // queues: done (n entries), trace (2n), cls0..3 (n each)
    uint32_t* ctr;
    uint32_t* fetch;     // entry cursors, kFetchStride apart: k_trace's (one per queue part), then k_medium's
    unsigned long long* stat;
    uint32_t* spill;
    uint32_t spillStride;
    // f is a compile-time constant at every use, so the array choice folds away
    __device__ __forceinline__ float4& f4(uint32_t s, uint32_t f) const {
        return f < SF_ALPHA ? ray[(size_t)s * kRayVec + f]
                            : f < SF_NEE ? spec[(size_t)s * kSpecVec + (f - SF_ALPHA)] : nee[(size_t)s * kNeeVec + (f - SF_NEE)];
    }
    __device__ __forceinline__ uint4& u4(uint32_t s, uint32_t f) const { return *(uint4*)&f4(s, f); }
};

// queue offsets in units of n entries
enum QueueId : uint32_t { Q_DONE = 0, Q_TRACE = 1, Q_CLS0 = 3, kNumQueues = Q_CLS0 + kNumClasses };
static_assert(C_CLS0 + kNumClasses <= C_PENDING, "counter layout");

__device__ __forceinline__ uint32_t* queue(const PathState& ps, uint32_t id) { return ps.q + (size_t)id * ps.n; }

__device__ __forceinline__ void loadPlanes(const PathState& ps, uint32_t s, uint32_t f, float* v) {
    const float4* b = &ps.f4(s, f);
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        float4 a = b[p];
        v[4 * p] = a.x; v[4 * p + 1] = a.y; v[4 * p + 2] = a.z; v[4 * p + 3] = a.w;
    }
}
__device__ __forceinline__ void storePlanes(const PathState& ps, uint32_t s, uint32_t f, const float* v) {
    float4* b = &ps.f4(s, f);
#pragma unroll
    for (int p = 0; p < 4; ++p) b[p] = make_float4(v[4 * p], v[4 * p + 1], v[4 * p + 2], v[4 * p + 3]);
}

// An unoccluded shadow ray's contribution (already weighted by the shading
// kernel that emitted it) joins the path radiance; the shadow kernel only
// flags it. Called by the first kernel that reads the radiance afterwards,
// which keeps the summation order of libSLR/Renderer/PTRenderer.cpp:190-214
// (light sample before the next emission term).
__device__ __forceinline__ void resolveShadow(const PathState& ps, uint32_t s, float* rad) {
    if (ps.vis[s]) {
        float c[kNumL];
        loadPlanes(ps, s, SF_NEE, c);
#pragma unroll
        for (int k = 0; k < kNumL; ++k) rad[k] += c[k];
        ps.vis[s] = 0;
    }
}

// ------------------------------------------------------------------ environment light

constexpr double kPiD = 3.14159265358979323846;

// UpsampledContinuousSpectrum::evaluate (libSLR/BasicTypes/spectrum_types.cpp:
// 611-651) of one environment texel at the path's 16 wavelengths: the three
// or four weighted basis spectra, each interpolated linearly in wavelength.
__device__ __forceinline__ void envTexelSpectrum(const DevScene& sc, uint32_t texel, float off, float* out) {
    const float4 tx = sc.envTexel[texel];
    const uint32_t adj = __float_as_uint(tx.x);
    uint32_t idx[4];
    float w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) idx[k] = (adj >> (8 * k)) & 0xFFu;
    int n;
    if (idx[3] != 0xFFu) {
        w[0] = (1 - tx.y) * (1 - tx.z);
        w[1] = tx.y * (1 - tx.z);
        w[2] = (1 - tx.y) * tx.z;
        w[3] = tx.y * tx.z;
        n = 4;
    } else {
        w[0] = tx.y;
        w[1] = tx.z;
        w[2] = 1.0f - tx.y - tx.z;
        w[3] = 0.0f;
        n = 3;
    }
#pragma unroll
    for (int i = 0; i < kNumL; ++i) {
        float p = quotient(lambdaOf(i, off) - kLambdaLowD, kLambdaHighD - kLambdaLowD,
                           1.0f / (kLambdaHighD - kLambdaLowD));
        p = fminf(fmaxf(p, 0.0f), 1.0f);
        const float binF = p * 94;
        const uint32_t bin = min((uint32_t)binF, 94u);
        const float t = binF - bin;
        float r = 0.0f;
        for (int j = 0; j < n; ++j) {
            const float2 pr = sc.mengBasis[idx[j] * 95 + bin];   // bins bin and min(bin + 1, 94)
            r += w[j] * (pr.x * (1 - t) + pr.y * t);
        }
        out[i] = r * tx.w;
    }
}

// Radiance reaching a path from the environment at texture coordinate
// (tu, tv): IBLEmitterSurfaceProperty::emittance (pi * texture * scale,
// IBLEmitterSurfaceProperty.cpp:16-18) times IBLEDF's 1/pi (IBLEDF.cpp:18-22),
// the texel chosen as ImageSpectrumTexture::evaluate does
// (libSLR/Texture/image_textures.cpp:14-21).
__device__ __forceinline__ void envRadiance(const DevScene& sc, float tu, float tv, float off, float* Le) {
    float u = fmodf(tu, 1.0f), v = fmodf(tv, 1.0f);
    u += u < 0 ? 1.0f : 0.0f;
    v += v < 0 ? 1.0f : 0.0f;
    const uint32_t px = min((uint32_t)(sc.envW * u), sc.envW - 1);
    const uint32_t py = min((uint32_t)(sc.envH * v), sc.envH - 1);
    envTexelSpectrum(sc, py * sc.envW + px, off, Le);
    const float pi = (float)kPiD, invPi = (float)(1.0f / kPiD);
#pragma unroll
    for (int k = 0; k < kNumL; ++k) Le[k] = ((pi * Le[k]) * sc.envScale) * invPi;
}

// Vector3D::toPolarYUp (libSLR/BasicTypes/Vector3D.h:78-81).
__device__ __forceinline__ void toPolarYUp(float3 d, float* theta, float* phi) {
    *theta = acosf(fminf(fmaxf(d.y, -1.0f), 1.0f));
    *phi = fmodf((float)(atan2f(-d.x, d.z) + 2 * kPiD), (float)(2 * kPiD));
}

__device__ __forceinline__ uint32_t prevPowerOf2(uint32_t n) {
    uint32_t d = 1;
    while ((d << 1) <= n) d <<= 1;
    return d;
}

// RegularConstantContinuousDistribution1D::sample (libSLR/Core/distributions.cpp:212-223).
__device__ __forceinline__ float dist1DSample(const float* cdf, const float* pdf, uint32_t n, float u, float* p) {
    int idx = (int)n;
    for (uint32_t d = prevPowerOf2(n); d > 0; d >>= 1)
        if (idx - (int)d > 0 && cdf[idx - d] >= u) idx -= (int)d;
    --idx;
    *p = pdf[idx];
    const float t = (u - cdf[idx]) / (cdf[idx + 1] - cdf[idx]);
    return (idx + t) / n;
}

// RegularConstantContinuousDistribution2D::sample / evaluatePDF
// (distributions.cpp:264-280) over the environment's importance map:
// d1 (theta / pi) from the row marginal, d0 (phi / 2 pi) from the row.
__device__ __forceinline__ void envSample(const DevScene& sc, float u0, float u1, float* d0, float* d1, float* pdf) {
    float topPdf;
    *d1 = dist1DSample(sc.envMargCdf, sc.envMargPdf, sc.envMapH, u1, &topPdf);
    const uint32_t row = min((uint32_t)(sc.envMapH * *d1), sc.envMapH - 1);
    *d0 = dist1DSample(sc.envCondCdf + row * (sc.envMapW + 1), sc.envCondPdf + row * sc.envMapW, sc.envMapW, u0, pdf);
    *pdf *= topPdf;
}
__device__ __forceinline__ float envPdf(const DevScene& sc, float d0, float d1) {
    const uint32_t row = min((uint32_t)(sc.envMapH * d1), sc.envMapH - 1);
    const uint32_t col = min((uint32_t)(d0 * sc.envMapW), sc.envMapW - 1);
    const uint32_t r1 = min((uint32_t)(d1 * sc.envMapH), sc.envMapH - 1);
    return sc.envMargPdf[r1] * sc.envCondPdf[row * sc.envMapW + col];
}

// InfiniteSphereSurfaceObject::evaluateAreaPDF (libSLR/Core/surface_object.cpp:281-286).
__device__ __forceinline__ float envAreaPdf(const DevScene& sc, float phi, float theta) {
    const float uvPdf = envPdf(sc, (float)(phi / (2 * kPiD)), (float)(theta / kPiD));
    return (float)(uvPdf / (2 * kPiD * kPiD * sinf(theta)));
}

// ------------------------------------------------------------------ participating medium

// DensityGridMediumDistribution::calcDensity (libSLR/MediumDistribution/
// DensityGridMediumDistribution.cpp:369-395): the point's coordinates in the
// region (BoundingBox3D::calculateLocalCoordinates), zero outside it,
// trilinear between the lattice points around it, in the reference's term
// order.
__device__ __forceinline__ float mediumDensity(const DevScene& sc, float3 p) {
    // (p - min) / (max - min) per axis, the IEEE quotients by way of the
    // region's reciprocal extents
    const float px = quotient(p.x - sc.medMin[0], sc.medMax[0] - sc.medMin[0], sc.medInvExt[0]);
    const float py = quotient(p.y - sc.medMin[1], sc.medMax[1] - sc.medMin[1], sc.medInvExt[1]);
    const float pz = quotient(p.z - sc.medMin[2], sc.medMax[2] - sc.medMin[2], sc.medInvExt[2]);
    // outside the region (and any NaN coordinate) reads zero
    if (!(px >= 0 && py >= 0 && pz >= 0 && px <= 1 && py <= 1 && pz <= 1)) return 0.0f;
    const uint32_t nx = sc.medN[0], ny = sc.medN[1], nz = sc.medN[2];
    // lower lattice indices; the upper ones, min(l + 1, n - 1), are folded
    // into the corner-packed entry of (lx, ly, lz)
    const uint32_t lx = min((uint32_t)(px * (nx - 1)), nx - 1);
    const uint32_t ly = min((uint32_t)(py * (ny - 1)), ny - 1);
    const uint32_t lz = min((uint32_t)(pz * (nz - 1)), nz - 1);
    const float wux = px * (nx - 1) - lx, wlx = 1 - wux;
    const float wuy = py * (ny - 1) - ly, wly = 1 - wuy;
    const float wuz = pz * (nz - 1) - lz, wlz = 1 - wuz;
    // the eight corners in one 32-byte read (corner-packed lattice)
    const float4* c = sc.medCorners + 2 * (((size_t)lz * ny + ly) * nx + lx);
    const float4 a = c[0], b = c[1];
    return a.x * wlz * wly * wlx + a.y * wlz * wly * wux + a.z * wlz * wuy * wlx + a.w * wlz * wuy * wux +
           b.x * wuz * wly * wlx + b.y * wuz * wly * wux + b.z * wuz * wuy * wlx + b.w * wuz * wuy * wux;
}

// Clips [t0, t1] of the ray to the medium region; false if nothing is left.
__device__ __forceinline__ bool clipToMedium(const DevScene& sc, float3 o, float3 d, float* t0, float* t1) {
    const float oa[3] = {o.x, o.y, o.z}, da[3] = {d.x, d.y, d.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float inv = 1.0f / da[a];
        float ta = (sc.medMin[a] - oa[a]) * inv, tb = (sc.medMax[a] - oa[a]) * inv;
        if (ta > tb) { const float x = ta; ta = tb; tb = x; }
        *t0 = fmaxf(*t0, ta);   // a NaN (origin on a plane, zero direction) keeps the bound
        *t1 = fminf(*t1, tb);
    }
    return *t0 < *t1;
}

// Upper bound on the number of supervoxel steps plus tentative collisions
// one tracking walk takes before it gives up (a guard; optical depths in the
// configurations stay far below it; walks that hit it are counted).
constexpr uint32_t kMaxTrackSteps = 1u << 16;

// Work counters of the tracking walks (majorant cells entered, density lookups).
struct MedCount {
    uint32_t cells = 0, lookups = 0;
    uint32_t guard = 0;   // walks stopped by kMaxTrackSteps (reported; expected 0)
};

// A walk over the supervoxels that the part of [t0, t1] inside the medium
// crosses (3D DDA in lattice space, cells of kMedSuperVoxel lattice
// intervals), drawing tentative collisions against the piecewise-constant
// majorant mu = sigma_e * max density of the cell: exponential free paths
// whose optical depth carries across cells, as the reference's supervoxel
// traversal does with its trilinear majorant (DensityGridMediumDistribution.cpp:
// 208-367). One advance() is one step: a tentative collision in the current
// cell, or the move into the next cell.
struct MediumWalk {
    enum : uint32_t { kCollide = 0, kMoved = 1, kDone = 2, kGuard = 3 };
    float t, t1, tau;
    // per axis: supervoxel, step direction, next boundary and spacing in t
    // (named per axis, not arrays: a lane-varying axis index into an array
    // would move the walk state to scratch memory)
    int cx, cy, cz, sx, sy, sz;
    float nx, ny, nz, dx, dy, dz;
    uint32_t steps;

    // Clips [t0, t1] to the medium, sets up the DDA and draws the first free
    // path; false if the segment misses the medium (no draw then).
    __device__ __forceinline__ bool init(const DevScene& sc, float3 o, float3 d, float t0, float t1In, uint4& rng) {
        t1 = t1In;
        if (!clipToMedium(sc, o, d, &t0, &t1)) return false;
        const float oa[3] = {o.x, o.y, o.z}, da[3] = {d.x, d.y, d.z};
        int cell[3], step[3];
        float tNext[3], tDelta[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float scale = sc.medScale[a];   // lattice units per length
            const float l0 = (oa[a] + da[a] * t0 - sc.medMin[a]) * scale;
            int c = (int)floorf(l0 / kMedSuperVoxel);
            c = max(0, min((int)sc.medCells[a] - 1, c));
            cell[a] = c;
            const float dl = da[a] * scale;
            // one division per axis; the two quotients by dl from it
            const float rdl = 1.0f / dl;
            if (dl > 0.0f) {
                step[a] = 1;
                tNext[a] = t0 + quotient((float)((c + 1) * (int)kMedSuperVoxel) - l0, dl, rdl);
                tDelta[a] = quotient(kMedSuperVoxel, dl, rdl);
            } else if (dl < 0.0f) {
                step[a] = -1;
                tNext[a] = t0 + quotient((float)(c * (int)kMedSuperVoxel) - l0, dl, rdl);
                tDelta[a] = -quotient(kMedSuperVoxel, dl, rdl);
            } else {
                step[a] = 0;
                tNext[a] = INFINITY;
                tDelta[a] = INFINITY;
            }
        }
        cx = cell[0]; cy = cell[1]; cz = cell[2];
        sx = step[0]; sy = step[1]; sz = step[2];
        nx = tNext[0]; ny = tNext[1]; nz = tNext[2];
        dx = tDelta[0]; dy = tDelta[1]; dz = tDelta[2];
        t = t0;
        tau = -logf(1.0f - xsFloat(rng));
        steps = 0;
        return true;
    }
    __device__ __forceinline__ uint32_t cellIndex(const DevScene& sc) const {
        return ((uint32_t)cz * sc.medCells[1] + (uint32_t)cy) * sc.medCells[0] + (uint32_t)cx;
    }
    // One step with the current cell's majorant mu. kCollide: t is a
    // tentative collision (the caller decides, and redraws tau to go on);
    // kMoved: the walk entered the next cell (its majorant is needed next);
    // kDone: the walk left [t0, t1] or the lattice; kGuard: kMaxTrackSteps.
    __device__ __forceinline__ uint32_t advance(const DevScene& sc, float mu) {
        if (++steps > kMaxTrackSteps) return kGuard;
        const float tExit = fminf(fminf(nx, ny), fminf(nz, t1));
        const float span = fmaxf(tExit - t, 0.0f);
        if (mu > 0.0f && tau < mu * span) {
            t += tau / mu;
            return kCollide;
        }
        if (mu > 0.0f) tau -= mu * span;
        t = fmaxf(t, tExit);
        if (t >= t1) return kDone;
        if (nx < ny ? nx < nz : false) {
            cx += sx;
            if (cx < 0 || cx >= (int)sc.medCells[0]) return kDone;
            nx += dx;
        } else if (nx < ny ? false : ny < nz) {
            cy += sy;
            if (cy < 0 || cy >= (int)sc.medCells[1]) return kDone;
            ny += dy;
        } else {
            cz += sz;
            if (cz < 0 || cz >= (int)sc.medCells[2]) return kDone;
            nz += dz;
        }
        return kMoved;
    }
};

// The whole walk: collide(t, mu) is called at each tentative collision and
// returns true to stop the walk (the walk then returns true); false: the
// walk reached t1 (or the step guard, counted in mc).
template <typename Collide>
__device__ __forceinline__ bool trackMedium(const DevScene& sc, float3 o, float3 d, float t0, float t1, uint4& rng,
                                            Collide&& collide, MedCount* mc = nullptr) {
    MediumWalk w;
    if (!w.init(sc, o, d, t0, t1, rng)) return false;
    if (mc) ++mc->cells;
    float mu = sc.medSigmaE * sc.medMajorant[w.cellIndex(sc)];
    for (;;) {
        const uint32_t r = w.advance(sc, mu);
        if (r == MediumWalk::kCollide) {
            if (mc) ++mc->lookups;
            if (collide(w.t, mu)) return true;
            w.tau = -logf(1.0f - xsFloat(rng));
        } else if (r == MediumWalk::kMoved) {
            if (mc) ++mc->cells;
            mu = sc.medSigmaE * sc.medMajorant[w.cellIndex(sc)];
        } else {
            if (r == MediumWalk::kGuard && mc) ++mc->guard;
            return false;
        }
    }
}

// Free-path sampling for the hero wavelength by delta tracking (the
// medium's extinction is grey, so the free path is the same at every
// wavelength and the reference's ratio-tracked throughput of the others,
// DensityGridMediumDistribution.cpp:454-486, is one). Returns true with the
// scattering distance in *ts when a real collision happens before t1.
__device__ __forceinline__ bool sampleFreePath(const DevScene& sc, float3 o, float3 d, float t0, float t1, uint4& rng,
                                               float* ts, MedCount* mc = nullptr) {
    return trackMedium(sc, o, d, t0, t1, rng, [&](float t, float mu) {
        const float sigma = sc.medSigmaE * mediumDensity(sc, o + d * t);
        if (xsFloat(rng) < sigma / mu) {
            *ts = t;
            return true;
        }
        return false;
    }, mc);
}

// Transmittance of [t0, t1] by ratio tracking with Russian roulette below
// 0.1, as the reference's estimateTransmittance (DensityGridMediumDistribution.cpp:
// 618-655); one estimate serves all wavelengths (grey extinction).
__device__ __forceinline__ float mediumTransmittance(const DevScene& sc, float3 o, float3 d, float t0, float t1,
                                                     uint4& rng, MedCount* mc = nullptr) {
    float T = 1.0f;
    trackMedium(sc, o, d, t0, t1, rng, [&](float t, float mu) {
        const float sigma = sc.medSigmaE * mediumDensity(sc, o + d * t);
        T *= 1.0f - fminf(sigma / mu, 1.0f);
        if (T < 0.1f) {
            if (xsFloat(rng) < T) T = 1.0f;
            else {
                T = 0.0f;
                return true;
            }
        }
        return false;
    }, mc);
    return T;
}

// HenyeyGreensteinPhaseFunction (libSLR/PhaseFunction/basic_phase_functions.cpp:
// 36-64): cosTheta is taken against the direction of propagation; g = 0 is
// the isotropic limit (IsotropicPhaseFunction, :13-33).
__device__ __forceinline__ float hgValue(float g, float cosTheta) {
    if (g == 0.0f) return (float)(1.0 / (4.0 * kPiD));
    return (1 - g * g) / ((float)(4 * kPiD) * powf(1 + g * g - 2 * g * cosTheta, 1.5f));
}
__device__ __forceinline__ float3 hgSample(float g, float u0, float u1, float* value) {
    if (g == 0.0f) {
        // uniformSampleSphere (libSLR/Core/distributions.h:54-58)
        const float phi = (float)(2 * kPiD) * u1, theta = acosf(1 - 2 * u0);
        *value = (float)(1.0 / (4.0 * kPiD));
        return f3(cosf(phi) * sinf(theta), sinf(phi) * sinf(theta), cosf(theta));
    }
    const float sqTerm = (1 - g * g) / (1 - g + 2 * g * u1);
    const float cosTheta = fminf(fmaxf(0.5f / g * (1 + g * g - sqTerm * sqTerm), -1.0f), 1.0f);
    const float phi = (float)(2 * kPiD) * u0;
    *value = (1 - g * g) / ((float)(4 * kPiD) * powf(1 + g * g - 2 * g * cosTheta, 1.5f));
    const float sinTheta = sqrtf(1 - cosTheta * cosTheta);
    return f3(cosf(phi) * sinTheta, sinf(phi) * sinTheta, cosTheta);
}
// ------------------------------------------------------------------ kernels

// Puts job `job` (slots base .. base + planes * pixels - 1) into the pool:
// slot (plane, pixel) takes samples plane, plane + planes, ... of the pixel,
// and its XORShift stream is seeded from the job's seed and the slot's index
// within the job, so a job renders the same alone or beside other jobs.
__global__ void k_init_slots(PathState ps, uint32_t job, uint32_t base, uint32_t pixels, uint32_t planes, uint32_t spp,
                             uint32_t seed) {
    const uint32_t count = pixels * planes;
    for (uint32_t local = blockIdx.x * blockDim.x + threadIdx.x; local < count; local += gridDim.x * blockDim.x) {
        const uint32_t s = base + local;
        const uint32_t pixel = local % pixels;
        const uint32_t plane = local / pixels;
        const uint32_t total = spp > plane ? (spp - plane + planes - 1) / planes : 0;
        ps.u4(s, SF_INFO) = make_uint4(pixel, 0u, 0u, total);
        // one XORShift stream per slot, seeded like XORShiftRNG(seed')
        uint32_t h = seed ^ (0x9E3779B9u * (local + 1u));
        h ^= h >> 16; h *= 0x85EBCA6Bu; h ^= h >> 13; h *= 0xC2B2AE35u; h ^= h >> 16;
        ps.u4(s, SF_RNG) = xsSeed(h);
        ps.u4(s, SF_JOB) = make_uint4(job, plane, fmix32(h ^ 0x68E31DA4u), 0u);
        ps.code[s] = CODE_DONE;
        ps.vis[s] = 0;
    }
}

// Makes slots base .. base + count - 1 idle (a cancelled job).
__global__ void k_idle_slots(PathState ps, uint32_t base, uint32_t count) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        ps.code[base + i] = CODE_IDLE;
        ps.vis[base + i] = 0;
    }
}

// Zeroes what the post-shade compaction counts, so that it can be run again
// after jobs were admitted or cancelled between iterations.
__global__ void k_zero_counts(PathState ps) {
    if (threadIdx.x == 0) {
        ps.ctr[C_DONE] = 0;
        ps.ctr[C_PENDING] = 0;
    }
    for (uint32_t j = threadIdx.x; j < kMaxJobs; j += blockDim.x) ps.jobActive[j] = 0;
}

// Clears the pool (all slots idle) and its counters.
__global__ void k_reset_pool(PathState ps) {
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < ps.n; s += gridDim.x * blockDim.x) {
        ps.code[s] = CODE_IDLE;
        ps.vis[s] = 0;
    }
    if (blockIdx.x == 0) {
        for (uint32_t i = threadIdx.x; i < kNumCtr; i += blockDim.x) ps.ctr[i] = 0;
        for (uint32_t i = threadIdx.x; i < kNumStat * kStatSpread; i += blockDim.x) ps.stat[i] = 0;
        for (uint32_t i = threadIdx.x; i < kMaxJobs; i += blockDim.x) ps.jobActive[i] = 0;
    }
}

// C *= weight; film add (PTRenderer.cpp:129-138, ImageSensor.cpp:125-130,
// SpectrumStorage::add spectrum_types.h:725-736). The slot owns its film cell
// (plane s / pixels), so the add is a plain read-modify-write. A sample with
// no radiance leaves the sums as they are and is not written.
__device__ __forceinline__ void fileSample(const PathState& ps, uint32_t s, uint32_t pixel, const float4& misc,
                                           float* c) {
    const uint4 jp = ps.u4(s, SF_JOB);
    const DevJob& job = ps.jobs[jp.x];
    const float recBinWidth = kNumL / (kLambdaHighD - kLambdaLowD);
    bool finite = true, any = false;
#pragma unroll
    for (int k = 0; k < kNumL; ++k) {
        c[k] *= misc.w;
        finite = finite && isfinite(c[k]);
        any = any || c[k] != 0.0f;
    }
    if (!finite) {
        atomicAdd(&ps.ctr[C_NONFINITE], 1u);
        return;
    }
    if (!any) return;
    float add[kNumL];
#pragma unroll
    for (int k = 0; k < kNumL; ++k) add[k] = 0.0f;
    // the bin of wavelength k is k or k + 1 (k + 1 when rounding carries
    // (k + offset) / 16 over the next bin edge; checked exhaustively over every
    // offset the random stream yields), so the adds keep static indices and
    // the sums stay in registers, in the same order as indexing add[b]
#pragma unroll
    for (int k = 0; k < kNumL; ++k) {
        float l = lambdaOf(k, misc.x);
        const float pos = quotient(l - kLambdaLowD, kLambdaHighD - kLambdaLowD, 1.0f / (kLambdaHighD - kLambdaLowD));
        uint32_t b = min((uint32_t)(pos * kNumL), (uint32_t)(kNumL - 1));
        const float v = c[k] * recBinWidth;
        add[k] += b == (uint32_t)k ? v : 0.0f;
        if (k + 1 < kNumL) add[k + 1] += b == (uint32_t)(k + 1) ? v : 0.0f;
    }
    double* f = job.film + ((size_t)jp.y * job.pixels + pixel) * kNumL;
#pragma unroll
    for (int k = 0; k < kNumL; k += 2) {
        double2 v = *(double2*)(f + k);
        v.x += (double)add[k];
        v.y += (double)add[k + 1];
        *(double2*)(f + k) = v;
    }
}

// The slot's next camera sample (PTRenderer.cpp:108-124,
// PerspectiveCamera.cpp:31-82): draws pixel jitter, wavelength offset, hero
// wavelength and lens position from the slot's stream, writes the ray and
// misc, and sets info for a path of length 0. The throughput (all ones) and
// radiance (zero) of a fresh path are implied by path length 0 and not
// stored; the first shading kernel starts from them.
__device__ __forceinline__ void startCameraSample(const DevScene& sc, const PathState& ps, uint32_t s, uint4& info,
                                                  uint4& rng, float4& misc) {
    const DevCamera& cam = ps.jobs[ps.u4(s, SF_JOB).x].cam;
    const uint32_t px = info.x % cam.width, py = info.x / cam.width;
    const float fx = px + xsFloat(rng);
    const float fy = py + xsFloat(rng);
    const float off = xsFloat(rng);
    const float uL = xsFloat(rng);
    const uint32_t hero = min((uint32_t)(uint16_t)(kNumL * uL), (uint32_t)(kNumL - 1));
    const float lu0 = xsFloat(rng), lu1 = xsFloat(rng);
    float lx = 0, ly = 0;
    concentricDisk(lu0, lu1, &lx, &ly);
    const float3 orgL = f3(cam.lensRadius * lx, cam.lensRadius * ly, 0.0f);
    const float* M = cam.l2w;
    const float3 org = f3(M[0] * orgL.x + M[1] * orgL.y + M[2] * orgL.z + M[3],
                          M[4] * orgL.x + M[5] * orgL.y + M[6] * orgL.z + M[7],
                          M[8] * orgL.x + M[9] * orgL.y + M[10] * orgL.z + M[11]);
    const float* N = cam.nrm;
    const float3 fz = normalize3(f3(N[2], N[5], N[8]));
    const float3 fxv = normalize3(f3(M[0], M[4], M[8]));
    const float3 fyv = cross3(fz, fxv);
    const float ux = fx / cam.width, uy = fy / cam.height;
    const float3 pFocus = f3(cam.opWidth * (0.5f - ux), cam.opHeight * (0.5f - uy), cam.objDist);
    const float3 dl = normalize3(pFocus - orgL);
    const float dirPDF = cam.imgDist * cam.imgDist / ((dl.z * dl.z * dl.z) * cam.imgPlaneArea);
    const float3 dir = f3(dot3(f3(fxv.x, fyv.x, fz.x), dl), dot3(f3(fxv.y, fyv.y, fz.y), dl),
                          dot3(f3(fxv.z, fyv.z, fz.z), dl));
    const float areaPDF = cam.lensRadius > 0.0f
                              ? (float)(1.0 / (3.14159265358979323846 * cam.lensRadius * cam.lensRadius))
                              : 1.0f;
    const float selectWLPDF = kNumL / (kLambdaHighD - kLambdaLowD);
    const float cosT = fabsf(dot3(dir, fz));
    const float weight = cosT / (areaPDF * dirPDF * selectWLPDF);
    float one[kNumL];
#pragma unroll
    for (int k = 0; k < kNumL; ++k) one[k] = 1.0f;
    const float initY = importance16(one, (int)hero);
    ps.f4(s, SF_RO) = make_float4(org.x, org.y, org.z, 0.0f);
    ps.f4(s, SF_RD) = make_float4(dir.x, dir.y, dir.z, INFINITY);
    misc = make_float4(off, initY, 0.0f, weight);
    info.y = F_HAS_SAMPLE | (hero << 8);   // pathLength 0
    info.z += 1;
}

// A path that ended: its sample is filed and, while the slot has samples
// left, the next one starts. Returns the slot's new code (EXT or IDLE).
// info, misc and rng are written back here.
__device__ __forceinline__ uint8_t retireAndRestart(const DevScene& sc, const PathState& ps, uint32_t s, uint4 info,
                                                    float4 misc, uint4 rng, float* rad, unsigned long long& samples) {
    if (info.y & F_HAS_SAMPLE) {
        fileSample(ps, s, info.x, misc, rad);
        ++samples;
    }
    info.y &= ~(F_HAS_SAMPLE | F_RAD);
    uint8_t code = CODE_IDLE;
    if (info.z < info.w) {
        startCameraSample(sc, ps, s, info, rng, misc);
        ps.f4(s, SF_MISC) = misc;
        ps.u4(s, SF_RNG) = rng;
        code = CODE_EXT;
    }
    ps.u4(s, SF_INFO) = info;
    return code;
}

// Loads the radiance of a slot's current path (zero for a fresh path) with
// any resolved light-sample contribution added.
// The path's radiance so far plus its resolved light sample. The stored sum
// is read only when `read` is set (the caller adds to it or files it); a
// vertex that adds nothing leaves SF_SP untouched, and F_RAD says whether it
// holds anything.
__device__ __forceinline__ void loadRadiance(const PathState& ps, uint32_t s, uint32_t infoY, bool read, float* rad) {
    if (read && (infoY & F_RAD)) {
        loadPlanes(ps, s, SF_SP, rad);
    } else {
#pragma unroll
        for (int k = 0; k < kNumL; ++k) rad[k] = 0.0f;
    }
    resolveShadow(ps, s, rad);
}

// Slots whose path ended with a light sample still being traced (and the
// slots' first samples): file the finished sample, start the next one.
__global__ __launch_bounds__(kBlock) void k_regen(DevScene sc, PathState ps) {
    const uint32_t count = ps.ctr[C_DONE];
    const uint32_t* qin = queue(ps, Q_DONE);
    // the trace queue is rebuilt after this kernel
    if (blockIdx.x == 0 && threadIdx.x == 0) ps.ctr[C_TRACE] = 0;
    unsigned long long samples = 0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < count; i += gridDim.x * kBlock) {
        const uint32_t s = qin[i];
        const uint4 info = ps.u4(s, SF_INFO);
        const float4 misc = ps.f4(s, SF_MISC);
        const uint4 rng = ps.u4(s, SF_RNG);
        float rad[kNumL];
        if (info.y & F_HAS_SAMPLE) loadRadiance(ps, s, info.y, true, rad);
        ps.code[s] = retireAndRestart(sc, ps, s, info, misc, rng, rad, samples);
    }
    statAdd(ps.stat, S_SAMPLES, samples);
}

// One traversal launch per iteration for the whole trace queue: closest hit
// for extension entries (the hit and the hit material's shading class are
// written for the classification that follows), any hit for shadow entries
// (an unoccluded ray sets the slot's vis flag; its contribution is added by
// resolveShadow in the next kernel that reads the path radiance, outside this
// divergent loop). Mixing both kinds in one launch leaves one tail of long
// rays per iteration instead of two.
template <int D, int T, bool F, int W, int L, bool I>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(W > 0 ? W : 1, W > 0 ? W : 10)))
void k_trace(DevScene sc, PathState ps) {
    __shared__ uint32_t lds[D * kBlock];
    __shared__ float4 treelet[L > 0 ? L * kTreeletRows : 1];
    const uint32_t count = ps.ctr[C_TRACE];
    const uint32_t* qin = queue(ps, Q_TRACE);
    // the shading queues are rebuilt after this kernel
    if (blockIdx.x == 0 && threadIdx.x < kNumClasses) ps.ctr[C_CLS0 + threadIdx.x] = 0;
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    TraceStack st;
    st.lds = ldsColumn(lds, threadIdx.x);
    st.ldsStride = kBlock;
    st.spill = ps.spill + gtid;
    st.spillStride = ps.spillStride;
    st.ldsDepth = D;
    st.treeletNodes = L < sc.numNodes ? (uint32_t)L : sc.numNodes;
    st.treelet = loadTreelet(treelet, sc.nodes, st.treeletNodes);
    TraceCounters tc;
    traceLoop<T, F, false, I, SLRH_FETCH_PARTS>(sc, st, count, ps.fetch, [&](uint32_t i) { return qin[i]; },
        [&](uint32_t e, RayIn& r) {
            const uint32_t s = e & ~kShadowEntry;
            const bool shadow = (e & kShadowEntry) != 0;
            // origin + tmin (shared by both rays), then the ray's direction + tmax;
            // the reciprocal direction is formed when the ray starts
            r.o = ps.f4(s, SF_RO);
            r.d = ps.f4(s, shadow ? SF_SD : SF_RD);
            return shadow;
        },
        [&](uint32_t e, bool shadow, const HitOut& h) {
            const uint32_t s = e & ~kShadowEntry;
            if (shadow) {
                if (h.prim < 0) ps.vis[s] = 1;
            } else {
                ps.f4(s, SF_HIT) = make_float4(h.t, __int_as_float(h.prim), h.b0, h.b1);
                if (I) ps.u4(s, SF_JOB).w = (uint32_t)h.inst;
                ps.code[s] = (uint8_t)(CODE_CLS0 + (h.prim >= 0 ? h.aux : (uint32_t)CLS_MISS));
            }
        }, tc);
    statAdd(ps.stat, S_RAYS_EXT, tc.rays - tc.raysAny);
    statAdd(ps.stat, S_NODES_EXT, tc.nodes - tc.nodesAny);
    statAdd(ps.stat, S_TRIS_EXT, tc.tris - tc.trisAny);
    statAdd(ps.stat, S_RAYS_SHADOW, tc.raysAny);
    statAdd(ps.stat, S_NODES_SH, tc.nodesAny);
    statAdd(ps.stat, S_TRIS_SH, tc.trisAny);
    statAdd(ps.stat, S_NODES_LDS, tc.nodesLds);
    if (I) statAdd(ps.stat, S_INSTS, tc.insts);
    if (tc.overflow) atomicOr(&ps.ctr[C_OVERFLOW], 1u);
}

// The per-light records the next-event estimate reads (one gather instead of
// the light list, triangle, vertex and normal chain), computed once per
// scene on the device with the expressions the estimate used inline, so
// the values are the same bits.
__global__ void k_light_records(DevScene sc, float4* out) {
    for (uint32_t li = blockIdx.x * blockDim.x + threadIdx.x; li < sc.numLights; li += gridDim.x * blockDim.x) {
        const uint2 lent = sc.lights[li];
        const int32_t linst = lent.y == kNoInstance ? -1 : (int32_t)lent.y;
        float3 q0, q1, q2;
        uint4 lv;
        worldVerts(sc, lent.x, linst, &q0, &q1, &q2, &lv);
        uint32_t mat = lv.w;
        if (linst >= 0 && sc.instMat[linst] >= 0) mat = (uint32_t)sc.instMat[linst];
        const float3 lg = normalize3(cross3(q1 - q0, q2 - q0));
        const float areaPDF = 1.0f / (0.5f * length3(cross3(q1 - q0, q2 - q0)));
        float4* R = out + (size_t)kLightRecRows * li;
        R[0] = make_float4(q0.x, q0.y, q0.z, __uint_as_float(mat));
        R[1] = make_float4(q1.x, q1.y, q1.z, __int_as_float(linst));
        R[2] = make_float4(q2.x, q2.y, q2.z, areaPDF);
        R[3] = make_float4(lg.x, lg.y, lg.z, 0.0f);
        R[4] = sc.vnrm[lv.x];
        R[5] = sc.vnrm[lv.y];
        R[6] = sc.vnrm[lv.z];
        const DevSpectrum e = sc.matSpec[3 * mat + 2];
        R[7] = make_float4(e.lo, e.hi, __uint_as_float(e.offset), __uint_as_float(e.n));
    }
}

// Per-triangle shading records (DevScene::triShade), with the expressions
// calculateSurfacePoint uses, so the values are the same bits.
__global__ void k_tri_shade(DevScene sc, float4* out) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < sc.numPrims; t += gridDim.x * blockDim.x) {
        const uint4 pv = sc.primVerts[t];
        const float3 p0 = xyz(sc.vpos[pv.x]), p1 = xyz(sc.vpos[pv.y]), p2 = xyz(sc.vpos[pv.z]);
        const float3 g = normalize3(cross3(p1 - p0, p2 - p0));
        float4* R = out + (size_t)kTriShadeRows * t;
        const float4 n0 = sc.vnrm[pv.x];
        R[0] = make_float4(n0.x, n0.y, n0.z, __uint_as_float(pv.w));
        R[1] = sc.vnrm[pv.y];
        R[2] = sc.vnrm[pv.z];
        R[3] = sc.vtan[pv.x];
        R[4] = sc.vtan[pv.y];
        R[5] = sc.vtan[pv.z];
        R[6] = make_float4(g.x, g.y, g.z, 0.0f);
        R[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

static int buildTriShade(DeviceScene* ds, std::string* err) {
    const uint32_t n = ds->view.numPrims;
    ds->view.triShade = nullptr;
    if (n == 0) return 0;
    if (int r = ds->triShade.alloc((size_t)n * kTriShadeRows, err)) return r;
    hipLaunchKernelGGL(k_tri_shade, dim3(std::min(4096u, (n + 255) / 256)), dim3(256), 0, 0, ds->view,
                       ds->triShade.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    ds->view.triShade = ds->triShade.p;
    return 0;
}

static int buildLightRecords(DeviceScene* ds, std::string* err) {
    const uint32_t n = ds->view.numLights;
    ds->view.lightRec = nullptr;
    if (n == 0) return 0;
    if (int r = ds->lightRec.alloc((size_t)n * kLightRecRows, err)) return r;
    hipLaunchKernelGGL(k_light_records, dim3((n + 255) / 256), dim3(256), 0, 0, ds->view, ds->lightRec.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    ds->view.lightRec = ds->lightRec.p;
    return 0;
}

template <uint32_t CLS>
__device__ __forceinline__ void evalBsdf(const DevMaterial& m, const float* A, const float* B, const BsdfQuery& q,
                                         const float3& d, float* fs, float* pdf) {
    if (CLS == CLS_DIFFUSE) {
        Diffuse b{A, m.type == MAT_OREN_NAYAR, m.p0, m.p1};
        b.evaluate(q, d, fs, pdf);
    } else if (CLS == CLS_CONDUCTOR) {
        Conductor b{A, B, GGX{m.p0, m.p1}};
        b.evaluate(q, d, fs, pdf);
    } else {
        Dielectric b{A, B, GGX{m.p0, m.p1}};
        b.evaluate(q, d, fs, pdf);
    }
}

template <uint32_t CLS>
__device__ __forceinline__ bool sampleBsdf(const DevMaterial& m, const float* A, const float* B, const BsdfQuery& q,
                                           float uc, float u0, float u1, float* fs, BsdfSample* s) {
#pragma unroll
    for (int k = 0; k < kNumL; ++k) fs[k] = 0.0f;
    s->pdf = 0.0f;
    if (CLS == CLS_DIFFUSE) {
        Diffuse b{A, m.type == MAT_OREN_NAYAR, m.p0, m.p1};
        return b.sample(q, uc, u0, u1, fs, s);
    } else if (CLS == CLS_CONDUCTOR) {
        Conductor b{A, B, GGX{m.p0, m.p1}};
        return b.sample(q, uc, u0, u1, fs, s);
    } else {
        Dielectric b{A, B, GGX{m.p0, m.p1}};
        return b.sample(q, uc, u0, u1, fs, s);
    }
}

// Shading of one queue. CLS selects the scattering model compiled in.
// Register budget of the shading kernels: the diffuse kernel (most of the
// work) is held to 3 waves per SIMD; the others use what they need.
#ifndef SLRH_DIFFUSE_WAVES
#define SLRH_DIFFUSE_WAVES 3
#endif
#ifndef SLRH_SHADE_MIN_WAVES
#define SLRH_SHADE_MIN_WAVES 1
#endif
// BS is the workgroup size (the autotuner picks one of kShadeShapes).
template <uint32_t CLS, uint32_t BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(
    CLS == CLS_DIFFUSE ? SLRH_DIFFUSE_WAVES : SLRH_SHADE_MIN_WAVES, CLS == CLS_DIFFUSE ? SLRH_DIFFUSE_WAVES : 10)))
void k_shade(DevScene sc, PathState ps) {
    const uint32_t count = ps.ctr[C_CLS0 + CLS];
    const uint32_t* qin = queue(ps, Q_CLS0 + CLS);
    // the done queue and the pending count of the previous iteration were
    // consumed before the shading kernels; the compaction that follows them
    // rebuilds both
    if (CLS == 0 && blockIdx.x == 0) {
        if (threadIdx.x == 0) {
            ps.ctr[C_DONE] = 0;
            ps.ctr[C_PENDING] = 0;
        }
        for (uint32_t j = threadIdx.x; j < kMaxJobs; j += blockDim.x) ps.jobActive[j] = 0;
    }
    // the spectrum pool and the material spectrum headers, copied into LDS
    // when they fit: a spectrum evaluation is 32 scattered reads per path
    // vertex, which the vector memory path serves a lane at a time
    __shared__ float4 specLds4[kShadeLdsWords / 4];
    const float* specLds = (const float*)specLds4;
    const bool inLds = CLS != CLS_MISS && sc.specLdsWords != 0;
    if (inLds) {
        const uint32_t poolQuads = sc.specLdsPool / 4, quads = sc.specLdsWords / 4;
        const float4* pool4 = (const float4*)sc.specPool;
        const float4* hdr4 = (const float4*)sc.matSpec;
        for (uint32_t j = threadIdx.x; j < quads; j += blockDim.x)
            specLds4[j] = j < poolQuads ? pool4[j] : hdr4[j - poolQuads];
        __syncthreads();
    }
    auto spectrumAt = [&](float lo, float hi, uint32_t offset, uint32_t n, float off, float* out) {
        if (inLds) spectrumValues(specLds + offset, lo, hi, n, off, out);
        else spectrumValues(sc.specPool + offset, lo, hi, n, off, out);
    };
    // spectrum j (0: spec0, 1: spec1, 2: emittance) of material m
    auto matSpectrum = [&](uint32_t m, uint32_t j, float off, float* out) {
        if (inLds) {
            const float4 h = specLds4[sc.specLdsPool / 4 + 3 * m + j];
            spectrumValues(specLds + __float_as_uint(h.z), h.x, h.y, __float_as_uint(h.w), off, out);
        } else {
            const DevSpectrum h = sc.matSpec[3 * m + j];
            spectrumValues(sc.specPool + h.offset, h.lo, h.hi, h.n, off, out);
        }
    };
    unsigned long long samples = 0;
    for (uint32_t base = blockIdx.x * BS; base < count; base += gridDim.x * BS) {
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < count;
        uint8_t code = CODE_IDLE;
        uint32_t s = 0;
        if (valid) {
            s = qin[i];
            if (CLS == CLS_MISS) {
                // the extension ray left the scene: the path ends here, after
                // the environment's emission when there is one (Scene::intersect's
                // fallback, libSLR/Scene/Scene.cpp:57-62; PTRenderer.cpp:160-166
                // for a camera ray, :247-263 with MIS otherwise)
                const uint4 info = ps.u4(s, SF_INFO);
                const float4 misc = ps.f4(s, SF_MISC);
                const uint4 rng = ps.u4(s, SF_RNG);
                const uint32_t pathLen = info.y >> 16;
                float rad[kNumL];
                loadRadiance(ps, s, info.y, true, rad);
                if (sc.hasEnv) {
                    const float3 rd = xyz(ps.f4(s, SF_RD));
                    float theta, phi;
                    toPolarYUp(rd, &theta, &phi);
                    float Le[kNumL], alpha[kNumL];
                    envRadiance(sc, (float)(phi / (2 * kPiD)), (float)(theta / kPiD), misc.x, Le);
                    float misW = 1.0f;
                    if (pathLen > 0 && !(info.y & F_LAST_DELTA)) {
                        const float bsdfPDF = misc.z;
                        const float lightProb = envSelectProb(sc);
                        const float lightPDF = lightProb * envAreaPdf(sc, phi, theta) * 1.0f / fabsf(dot3(rd, -rd));
                        misW = (bsdfPDF * bsdfPDF) / (lightPDF * lightPDF + bsdfPDF * bsdfPDF);
                    }
                    if (pathLen == 0) {
#pragma unroll
                        for (int k = 0; k < kNumL; ++k) alpha[k] = 1.0f;
                    } else {
                        loadPlanes(ps, s, SF_ALPHA, alpha);
                    }
#pragma unroll
                    for (int k = 0; k < kNumL; ++k) rad[k] += alpha[k] * Le[k] * misW;
                }
                code = retireAndRestart(sc, ps, s, info, misc, rng, rad, samples);
            } else {
                bool cont = false, shadow = false;
                uint4 info = ps.u4(s, SF_INFO);
                float4 misc = ps.f4(s, SF_MISC);
                uint4 rng = ps.u4(s, SF_RNG);
                const float4 o4 = ps.f4(s, SF_RO), d4 = ps.f4(s, SF_RD), h = ps.f4(s, SF_HIT);
                const float3 ro = xyz(o4), rd = xyz(d4);
                const uint32_t prim = (uint32_t)__float_as_int(h.y);
                const int32_t hinst = sc.numInstances ? (int32_t)ps.u4(s, SF_JOB).w : -1;
                const SurfPt sp = surfaceAt(sc, prim, h.z, h.w, ro, rd, h.x, hinst);
                const DevMaterial mat = sc.mats[sp.mat];
                const int hero = (int)((info.y >> 8) & 0xFFu);
                uint32_t pathLen = info.y >> 16;
                float alpha[kNumL], rad[kNumL];
                if (pathLen == 0) {
                    // a fresh path: throughput one (PTRenderer.cpp:115)
#pragma unroll
                    for (int k = 0; k < kNumL; ++k) alpha[k] = 1.0f;
                } else {
                    loadPlanes(ps, s, SF_ALPHA, alpha);
                }
                const float3 dirOut = toLocal(sp, -rd);
                const bool emits = mat.emit >= 0 && dirOut.z > 0.0f;
                // the stored radiance is read here only when this vertex adds to it
                const bool adds = emits || ps.vis[s] != 0;
                loadRadiance(ps, s, info.y, adds, rad);
                // emission (PTRenderer.cpp:162-166 first hit, :247-263 implicit with MIS)
                if (emits) {
                    float Le[kNumL];
                    matSpectrum(sp.mat, 2, misc.x, Le);
                    float misW = 1.0f;
                    if (pathLen > 0 && !(info.y & F_LAST_DELTA)) {
                        const float bsdfPDF = misc.z;
                        const float3 dv = sp.p - ro;
                        const float dist2 = dot3(dv, dv);
                        // every light has importance one: an instance's triangles
                        // share the world triangles' selection probability
                        const float lightProb =
                            sc.lightPmf[hinst >= 0 ? 0 : sc.primToLight[prim]] * lightSelectScale(sc);
                        const float lightPDF =
                            lightProb * (1.0f / triArea(sc, prim, hinst)) * dist2 / fabsf(dot3(rd, sp.gN));
                        misW = (bsdfPDF * bsdfPDF) / (lightPDF * lightPDF + bsdfPDF * bsdfPDF);
                    }
                    const float edf = (float)(1.0 / 3.14159265358979323846);
#pragma unroll
                    for (int k = 0; k < kNumL; ++k) rad[k] += alpha[k] * (Le[k] * edf) * misW;
                }
                // the radiance so far is complete for this vertex: it goes to the
                // record now (and is read back if the path ends here), so its 16
                // values are not held in registers through light and BSDF sampling
                if (adds) {
                    storePlanes(ps, s, SF_SP, rad);
                    info.y |= F_RAD;
                }
                bool alive = true;
                // Russian roulette (PTRenderer.cpp:267-273), from the second vertex on
                if (pathLen > 0) {
                    const float p = fminf(importance16(alpha, hero) / misc.y, 1.0f);
                    if (xsFloat(rng) < p) {
                        // alpha /= continueProb multiplies by 1 / p
                        // (libSLR/BasicTypes/spectrum_types.h:390-395)
                        const float rp = 1.0f / p;
#pragma unroll
                        for (int k = 0; k < kNumL; ++k) alpha[k] *= rp;
                    } else alive = false;
                }
                if (alive) {
                    ++pathLen;
                    if (pathLen >= ps.jobs[ps.u4(s, SF_JOB).x].maxDepth || mat.type == MAT_NONE) alive = false;
                }
                if (alive) {
                    const float3 gL = toLocal(sp, sp.gN);
                    BsdfQuery q;
                    q.dir = dirOut;
                    q.ng = gL;
                    q.hero = hero;
                    // the material's spectra (reflectance, or the two indices of
                    // refraction) are evaluated where each block needs them, so
                    // they are not live across both blocks
                    const uint32_t matId = sp.mat;
                    const int32_t spec0 = mat.spec0, spec1 = mat.spec1;
                    const float wlOff = misc.x;
                    auto materialSpectra = [&, matId, spec0, spec1, wlOff](float* A, float* B) {
                        if (CLS == CLS_DIFFUSE) {
                            matSpectrum(matId, 0, wlOff, A);
                        } else if (inLds) {
                            matSpectrum(matId, 0, wlOff, A);
                            matSpectrum(matId, 1, wlOff, B);
                        } else {
                            // (through the material's spectrum ids: the header table
                            // here leaves an unused stack object in these kernels)
                            evalSpectrum(sc, spec0, wlOff, A);
                            evalSpectrum(sc, spec1, wlOff, B);
                        }
                    };
                    // ---- next event estimation (PTRenderer.cpp:179-218)
                    const float uSel = xsFloat(rng);
                    const float lu0 = xsFloat(rng), lu1 = xsFloat(rng);
                    // Scene::selectSurfaceLight (libSLR/Scene/Scene.cpp:137-155):
                    // surface lights or the environment by importance
                    bool envLight = false;
                    float uSurf = uSel;
                    // VolumetricPTRenderer.cpp:190-231 (with SLRH_COMPAT_VOLPT_TANGENT_NEE):
                    // no cosine guard, so a light sample in the shading plane
                    // gives 0/0 in the BSDF's shading-normal correction
                    // (directional_distribution_functions.h:325-327); if it is
                    // visible the NaN reaches the radiance and the sample is
                    // not filed
                    auto tangentNaN = [&](float3 dir, float tmax) {
                        float c[kNumL];
#pragma unroll
                        for (int k = 0; k < kNumL; ++k) c[k] = __builtin_nanf("");
                        storePlanes(ps, s, SF_NEE, c);
                        ps.f4(s, SF_RO) = make_float4(sp.p.x, sp.p.y, sp.p.z, 1e-4f);
                        ps.f4(s, SF_SD) = make_float4(dir.x, dir.y, dir.z, tmax);
                        shadow = true;
                    };
                    if (sc.hasEnv) {
                        const float surfImp = sc.lightImportance, sumImps = surfImp + 1.0f;
                        if (sumImps * uSel < surfImp) uSurf = uSel / (surfImp / sumImps);
                        else envLight = true;
                    }
                    if (envLight) {
                        // InfiniteSphereSurfaceObject::sample (surface_object.cpp:201-238)
                        float d0, d1, uvPdf;
                        envSample(sc, lu0, lu1, &d0, &d1, &uvPdf);
                        const float phi = (float)(d0 * (2 * kPiD)), theta = (float)(d1 * kPiD);
                        const float sinT = sinf(theta);
                        const float3 pos = f3(-sinf(phi) * sinT, cosf(theta), cosf(phi) * sinT);
                        const float areaPDF = (float)(uvPdf / (2 * kPiD * kPiD * sinT));
                        // a point at infinity: unit distance, geometric normal -pos
                        // (SurfacePoint::getDirectionFrom, geometry.h:147-151)
                        const float3 sdir = normalize3(pos);
                        const float cosS = fabsf(dot3(sdir, sp.gN));
                        const float cosL = fabsf(dot3(sdir, -pos));
                        if (sc.tangentNeeNaN && cosS == 0.0f) {
                            tangentNaN(sdir, FLT_MAX);
                        } else if (cosS * cosL > 0) {
                            float Le[kNumL], fs[kNumL];
                            envRadiance(sc, (float)(phi / (2 * kPiD)), (float)(theta / kPiD), misc.x, Le);
                            const float3 sdL = toLocal(sp, sdir);
                            float bpdf = 0.0f;
                            float A[kNumL], B[kNumL];
                            materialSpectra(A, B);
                            evalBsdf<CLS>(mat, A, B, q, sdL, fs, &bpdf);
                            const float lightPDF = envSelectProb(sc) * areaPDF;
                            const float bsdfPDF = bpdf * cosL;
                            const float mis = isinf(areaPDF) ? 1.0f
                                                             : (lightPDF * lightPDF) / (lightPDF * lightPDF + bsdfPDF * bsdfPDF);
                            const float scal = (cosS * cosL) * mis / lightPDF;
                            float c[kNumL];
                            bool any = false;
#pragma unroll
                            for (int k = 0; k < kNumL; ++k) {
                                c[k] = alpha[k] * Le[k] * fs[k] * scal;
                                any = any || c[k] != 0.0f;
                            }
                            if (any) {
                                // Scene::testVisibility toward infinity: [Epsilon, FLT_MAX]
                                storePlanes(ps, s, SF_NEE, c);
                                ps.f4(s, SF_RO) = make_float4(sp.p.x, sp.p.y, sp.p.z, 1e-4f);
                                ps.f4(s, SF_SD) = make_float4(sdir.x, sdir.y, sdir.z, FLT_MAX);
                                shadow = true;
                            }
                        }
                    } else {
                        float lprob;
                        const uint32_t li = sampleLight(sc, uSurf, &lprob);
                        const LightPoint L = lightPoint(sc, li, lu0, lu1);
                        const float3 lp = L.lp, lg = L.lg, ln = L.ln;
                        const float areaPDF = L.areaPDF;
                        const float3 dv = lp - sp.p;
                        const float dist2 = dot3(dv, dv);
                        const float3 sdir = divs(dv, sqrtf(dist2));
                        const float cosS = fabsf(dot3(sdir, sp.gN));
                        const float cosL = fabsf(dot3(sdir, lg));
                        // EDF is evaluated in the light's shading frame: only its z matters
                        const float lz = dot3(ln, -sdir);
                        if (sc.tangentNeeNaN && cosS == 0.0f) {
                            const float dist = length3(lp - sp.p);
                            tangentNaN(divs(lp - sp.p, dist), dist * (1 - 1e-4f));
                        } else if (cosS * cosL > 0 && lz > 0.0f) {
                            float M[kNumL], fs[kNumL];
                            spectrumAt(L.emit.x, L.emit.y, __float_as_uint(L.emit.z), __float_as_uint(L.emit.w),
                                       misc.x, M);
                            const float3 sdL = toLocal(sp, sdir);
                            float bpdf = 0.0f;
                            float A[kNumL], B[kNumL];
                            materialSpectra(A, B);
                            evalBsdf<CLS>(mat, A, B, q, sdL, fs, &bpdf);
                            const float lightPDF = lprob * lightSelectScale(sc) * areaPDF;
                            const float bsdfPDF = bpdf * cosL / dist2;
                            const float mis = (lightPDF * lightPDF) / (lightPDF * lightPDF + bsdfPDF * bsdfPDF);
                            const float G = cosS * cosL / dist2;
                            const float scal = G * mis / lightPDF;
                            const float edf = (float)(1.0 / 3.14159265358979323846);
                            float c[kNumL];
                            bool any = false;
#pragma unroll
                            for (int k = 0; k < kNumL; ++k) {
                                c[k] = alpha[k] * (M[k] * edf) * fs[k] * scal;
                                any = any || c[k] != 0.0f;
                            }
                            if (any) {
                                // Scene::testVisibility segment (libSLR/Scene/Scene.cpp:96-112)
                                const float dist = length3(lp - sp.p);
                                const float3 vdir = divs(lp - sp.p, dist);
                                storePlanes(ps, s, SF_NEE, c);
                                ps.f4(s, SF_RO) = make_float4(sp.p.x, sp.p.y, sp.p.z, 1e-4f);
                                ps.f4(s, SF_SD) = make_float4(vdir.x, vdir.y, vdir.z, dist * (1 - 1e-4f));
                                shadow = true;
                            }
                        }
                    }
                    // ---- BSDF sampling (PTRenderer.cpp:221-237)
                    {
                        const float uc = xsFloat(rng), u0 = xsFloat(rng), u1 = xsFloat(rng);
                        float fs[kNumL];
                        BsdfSample bs;
                        float A[kNumL], B[kNumL];
                        materialSpectra(A, B);
                        bool ok = sampleBsdf<CLS>(mat, A, B, q, uc, u0, u1, fs, &bs);
                        bool nonzero = false;
#pragma unroll
                        for (int k = 0; k < kNumL; ++k) nonzero = nonzero || fs[k] != 0.0f;
                        if (ok && nonzero && bs.pdf != 0.0f) {
                            const float w = fabsf(dot3(bs.dir, gL)) / bs.pdf;
#pragma unroll
                            for (int k = 0; k < kNumL; ++k) alpha[k] *= fs[k] * w;
                            const float3 nd = fromLocal(sp, bs.dir);
                            ps.f4(s, SF_RO) = make_float4(sp.p.x, sp.p.y, sp.p.z, 1e-4f);
                            ps.f4(s, SF_RD) = make_float4(nd.x, nd.y, nd.z, INFINITY);
                            misc.z = bs.pdf;
                            info.y &= ~F_LAST_DELTA;
                            cont = true;
                        }
                    }
                }
                info.y = (info.y & 0xFFFFu) | (pathLen << 16);
                if (cont || shadow) {
                    // the path goes on, or it ended with a light sample still to
                    // trace: its sample is filed once the shadow ray is resolved
                    if (cont) storePlanes(ps, s, SF_ALPHA, alpha);
                    ps.u4(s, SF_INFO) = info;
                    ps.f4(s, SF_MISC) = misc;
                    ps.u4(s, SF_RNG) = rng;
                    code = (uint8_t)((cont ? CODE_EXT : CODE_DONE) | (shadow ? CODE_SHADOW : 0));
                } else {
                    // the sample's radiance is the stored sum
                    float fin[kNumL];
                    if (info.y & F_RAD) loadPlanes(ps, s, SF_SP, fin);
                    else {
#pragma unroll
                        for (int k = 0; k < kNumL; ++k) fin[k] = 0.0f;
                    }
                    code = retireAndRestart(sc, ps, s, info, misc, rng, fin, samples);
                }
            }
        }
        if (valid) ps.code[s] = code;
    }
    statAdd(ps.stat, S_SAMPLES, samples);
}

// Slot-ordered compaction of the per-slot codes into queues (one pass over
// all slots; wave prefix sums, a block scan in LDS and one atomic per block
// and queue). Each block's output is a sorted run of slot ids, so the
// kernels that read the queues touch path state in near-sequential order.
//   MODE 0 (after shading): done queue (DONE without a pending shadow ray),
//          count of slots with an extension or shadow ray pending, and per
//          job the count of slots that are not idle.
//   MODE 1 (after regeneration): trace queue, the slot's extension entry
//          and then its shadow entry; clears the shadow bit.
//   MODE 2 (after tracing): one queue per shading class.
// Hands out queue entries to the lanes of a wave that ask for one: chunks of
// kFeedChunk entries per atomic on one of kFeedParts cursors (kFetchStride
// words apart; block b starts on part b % kFeedParts and moves on when its
// part runs dry), given out in lane order. Wave-uniform state; claim() must
// be reached by every lane of the wave.
constexpr uint32_t kFeedParts = SLRH_FETCH_PARTS;
constexpr uint32_t kFeedChunk = 128;
struct WaveFeeder {
    uint32_t next, end, part, partsLeft;
    bool exhausted;
    __device__ __forceinline__ void init() {
        next = end = 0;
        part = blockIdx.x % kFeedParts;
        partsLeft = kFeedParts;
        exhausted = false;
    }
    __device__ __forceinline__ bool claim(uint32_t* cursors, uint32_t count, bool want, uint32_t* idx) {
        const unsigned long long lack = __ballot(want);
        if (exhausted || lack == 0ull) return false;
        if (next >= end) {
            for (;;) {
                const uint32_t pb = (uint32_t)((uint64_t)count * part / kFeedParts);
                const uint32_t pe = (uint32_t)((uint64_t)count * (part + 1) / kFeedParts);
                uint32_t base = 0;
                if (__lane_id() == 0) base = atomicAdd(cursors + part * kFetchStride, kFeedChunk);
                base = __shfl(base, 0);
                if (base < pe - pb) {
                    next = pb + base;
                    end = pe - next > kFeedChunk ? next + kFeedChunk : pe;
                    break;
                }
                if (--partsLeft == 0) {
                    exhausted = true;
                    return false;
                }
                part = part + 1 < kFeedParts ? part + 1 : 0;
            }
        }
        const uint32_t avail = end - next;
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(lack >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)lack, 0u));
        const uint32_t need = (uint32_t)__popcll(lack);
        const bool got = want && rank < avail;
        if (got) *idx = next + rank;
        next += need < avail ? need : avail;
        return got;
    }
};

// Volumetric integrator, after the traversal: the extension rays' free
// paths through the medium (Scene::interact, libSLR/Scene/Scene.cpp:66-94:
// the medium over [tmin, surface hit]) by delta tracking, and the light
// samples' transmittance (Scene::testVisibility, :114-135) by ratio tracking,
// which scales the stored contribution before it is resolved. Each trace
// queue entry is its own walk: the extension ray advances the slot's
// stream, the light sample draws from a stream hashed for it (shadowStream),
// so a slot's two walks run side by side.
//
// Persistent: a lane advances its walk by one step per loop iteration (one
// tentative collision with its density lookup, or the move into the next
// supervoxel, whose majorant is requested then and used in the next
// iteration) and starts its next queue entry, claimed ahead of time, as soon
// as the walk ends, so a wave's lanes stay busy while walk lengths differ.
// Stages: 0 idle, 1 the entry's first words requested, 2 a visible light
// sample's ray requested, 3 walking.
// kLdsMajorant: the majorant grid (at most kMajorantLdsCells cells) is
// copied into LDS by each block of BS threads, one block per CU, so a walk
// step waits for an LDS read instead of a cache miss; otherwise it is read
// from global memory.
constexpr uint32_t kMajorantLdsCells = 32768;   // 128 KB: a 256^3 lattice in 8^3 supervoxels
constexpr uint32_t kMediumLdsBlock = 1024;      // 16 waves per CU share one copy
#ifndef SLRH_MEDIUM_INNER
#define SLRH_MEDIUM_INNER 2
#endif
constexpr uint32_t kMediumInnerSteps = SLRH_MEDIUM_INNER;
#ifndef SLRH_MEDIUM_BATCH
#define SLRH_MEDIUM_BATCH 16
#endif
constexpr int kMediumBatch = SLRH_MEDIUM_BATCH;
template <bool kLdsMajorant, uint32_t BS>
__global__ __launch_bounds__(BS) void k_medium(DevScene sc, PathState ps) {
    const uint32_t count = ps.ctr[C_TRACE];
    const uint32_t* qin = queue(ps, Q_TRACE);
    uint32_t* cursors = ps.fetch + kFeedParts * kFetchStride;   // zeroed by the trace compaction
    __shared__ float majLds[kLdsMajorant ? kMajorantLdsCells : 1];
    if (kLdsMajorant) {
        const uint32_t cells = sc.medCells[0] * sc.medCells[1] * sc.medCells[2];
        for (uint32_t j = threadIdx.x; j < cells; j += BS) majLds[j] = sc.medMajorant[j];
        __syncthreads();
    }
    auto majorant = [&](uint32_t cell) { return kLdsMajorant ? majLds[cell] : sc.medMajorant[cell]; };
    WaveFeeder feed;
    feed.init();
    unsigned long long events = 0, shadows = 0, rays = 0;
    MedCount mc;
    uint32_t stage = 0, s = 0, entNext = 0;
    bool hasNext = false, shadow = false;
    uint8_t vis = 0;
    float4 o4 = make_float4(0, 0, 0, 0), d4 = o4, h = o4;
    uint4 rng = make_uint4(0, 0, 0, 0), info = rng;
    uint32_t slotSeed = 0;
    float T = 1.0f, mu = 0.0f;
    MediumWalk w;
    // the end of a walk: the extension ray's stream, or the light sample's
    // contribution scaled by its transmittance
    auto finish = [&]() {
        if (!shadow) {
            ps.u4(s, SF_RNG) = rng;
        } else if (T != 1.0f) {
            float c[kNumL];
            loadPlanes(ps, s, SF_NEE, c);
#pragma unroll
            for (int k = 0; k < kNumL; ++k) c[k] *= T;
            storePlanes(ps, s, SF_NEE, c);
        }
        stage = 0;
    };
    for (;;) {
        // lanes without a next entry claim one once 16 of them lack one, or
        // at once for an idle lane (every lane of the wave reaches the claim)
        uint32_t idx = 0;
        const unsigned long long lackM = __ballot(!hasNext);
        const bool ask = !hasNext && (stage == 0 || __popcll(lackM) >= 16);
        if (feed.claim(cursors, count, ask, &idx)) {
            entNext = qin[idx];
            hasNext = true;
        }
        // an idle lane starts its next entry; its words are used from the
        // next iteration on
        bool fresh = false;
        if (stage == 0 && hasNext) {
            hasNext = false;
            s = entNext & ~kShadowEntry;
            shadow = (entNext & kShadowEntry) != 0;
            if (shadow) {
                vis = ps.vis[s];   // an occluded light sample needs no transmittance
                stage = 1;
            } else {
                o4 = ps.f4(s, SF_RO);
                d4 = ps.f4(s, SF_RD);
                h = ps.f4(s, SF_HIT);
                rng = ps.u4(s, SF_RNG);
                stage = 2;
            }
            fresh = true;
        }
        // wave-uniform exit: every lane idle, nothing claimed, queue drained
        if (__ballot(stage != 0 || hasNext) == 0ull) {
            if (feed.exhausted) break;
            continue;
        }
        if (stage == 1 && !fresh) {
            // a light sample: its segment once it is known to be unoccluded
            if (vis == 0) {
                stage = 0;
            } else {
                o4 = ps.f4(s, SF_RO);
                d4 = ps.f4(s, SF_SD);
                slotSeed = ps.u4(s, SF_JOB).z;
                info = ps.u4(s, SF_INFO);
                stage = 2;
                fresh = true;
            }
        }
        // The set-up and the collision test are long code paths that a wave
        // runs for all its lanes at once: lanes that reach one wait until
        // kMediumBatch lanes of the wave are there or no lane is walking, so
        // the wave runs each path for many lanes instead of a few every
        // iteration. A lane's own arithmetic and random draws are unchanged.
        const bool walking = __ballot(stage == 3) != 0ull;
        const bool setUp = !walking || __popcll(__ballot(stage == 2 && !fresh)) >= kMediumBatch;
        if (stage == 2 && !fresh && setUp) {
            // one code path for both kinds of entry: the extension ray up to
            // its hit with the slot's stream, the light sample's segment with
            // its own stream
            float t1 = h.x;
            if (shadow) {
                rng = shadowStream(slotSeed, info);
                T = 1.0f;
                t1 = d4.w;
            }
            if (w.init(sc, xyz(o4), xyz(d4), o4.w, t1, rng)) {
                ++rays;
                shadows += shadow ? 1 : 0;
                ++mc.cells;
                mu = sc.medSigmaE * majorant(w.cellIndex(sc));
                stage = 3;
            } else {
                stage = 0;   // misses the medium: nothing drawn, nothing written
            }
        }
        // walk steps up to the next tentative collision (stage 4)
        for (uint32_t inner = 0; inner < kMediumInnerSteps && stage == 3; ++inner) {
            const uint32_t r = w.advance(sc, mu);
            if (r == MediumWalk::kCollide) {
                ++mc.lookups;
                stage = 4;
            } else if (r == MediumWalk::kMoved) {
                ++mc.cells;
                mu = sc.medSigmaE * majorant(w.cellIndex(sc));
            } else {
                if (r == MediumWalk::kGuard) ++mc.guard;
                finish();
            }
        }
        const bool stillWalking = __ballot(stage == 3) != 0ull;
        const bool test = !stillWalking || __popcll(__ballot(stage == 4)) >= kMediumBatch;
        if (stage == 4 && test) {
            bool done = false;
            const float sigma = sc.medSigmaE * mediumDensity(sc, xyz(o4) + xyz(d4) * w.t);
            if (!shadow) {
                // delta tracking (sampleFreePath)
                if (xsFloat(rng) < sigma / mu) {
                    ps.f4(s, SF_HIT) = make_float4(w.t, __int_as_float(-2), 0.0f, 0.0f);
                    ps.code[s] = (uint8_t)(CODE_CLS0 + CLS_MEDIUM);
                    ++events;
                    done = true;
                }
            } else {
                // ratio tracking with Russian roulette (mediumTransmittance)
                T *= 1.0f - fminf(sigma / mu, 1.0f);
                if (T < 0.1f) {
                    if (xsFloat(rng) < T) T = 1.0f;
                    else {
                        T = 0.0f;
                        done = true;
                    }
                }
            }
            if (done) {
                finish();
            } else {
                w.tau = -logf(1.0f - xsFloat(rng));
                stage = 3;
            }
        }
    }
    statAdd(ps.stat, S_MED_EVENTS, events);
    statAdd(ps.stat, S_MED_SHADOWS, shadows);
    statAdd(ps.stat, S_MED_LOOKUPS, mc.lookups);
    statAdd(ps.stat, S_MED_CELLS, mc.cells);
    statAdd(ps.stat, S_MED_RAYS, rays);
    statAdd(ps.stat, S_MED_GUARD, mc.guard);
}

// Shading at a scattering point in the medium (VolumetricPTRenderer::Job::
// contribution, libSLR/Renderer/VolumetricPTRenderer.cpp:179-296, for a
// MediumPoint): Russian roulette, next-event estimation with the phase
// function (its value is the reference's VolumetricBSDF value albedo *
// phase, the geometry term has no cosine at the medium point, the light
// sample's transmittance is applied by k_medium), then a new direction
// sampled from the phase function, which multiplies the throughput by the
// albedo sigma_s / sigma_e. The free-path estimator's sigma_e / sigma_e
// (hero) factor is one for grey extinction.
__global__ __launch_bounds__(kBlock) void k_shade_medium(DevScene sc, PathState ps) {
    const uint32_t count = ps.ctr[C_CLS0 + CLS_MEDIUM];
    const uint32_t* qin = queue(ps, Q_CLS0 + CLS_MEDIUM);
    unsigned long long samples = 0;
    for (uint32_t base = blockIdx.x * kBlock; base < count; base += gridDim.x * kBlock) {
        const uint32_t i = base + threadIdx.x;
        if (i >= count) continue;
        const uint32_t s = qin[i];
        uint8_t code = CODE_IDLE;
        bool cont = false, shadow = false;
        uint4 info = ps.u4(s, SF_INFO);
        float4 misc = ps.f4(s, SF_MISC);
        uint4 rng = ps.u4(s, SF_RNG);
        const float4 o4 = ps.f4(s, SF_RO), d4 = ps.f4(s, SF_RD), h = ps.f4(s, SF_HIT);
        const float3 rd = xyz(d4);
        const float3 pos = xyz(o4) + rd * h.x;
        const int hero = (int)((info.y >> 8) & 0xFFu);
        uint32_t pathLen = info.y >> 16;
        float alpha[kNumL], albedo[kNumL];
        if (pathLen == 0) {
#pragma unroll
            for (int k = 0; k < kNumL; ++k) alpha[k] = 1.0f;
        } else {
            loadPlanes(ps, s, SF_ALPHA, alpha);
        }
        const bool adds = ps.vis[s] != 0;
        if (adds) {
            // the resolved light sample joins the stored radiance now
            float rad[kNumL];
            loadRadiance(ps, s, info.y, true, rad);
            storePlanes(ps, s, SF_SP, rad);
            info.y |= F_RAD;
        }
        bool alive = true;
        if (pathLen > 0) {
            const float p = fminf(importance16(alpha, hero) / misc.y, 1.0f);
            if (xsFloat(rng) < p) {
                const float rp = 1.0f / p;   // alpha /= continueProb (spectrum_types.h:390-395)
#pragma unroll
                for (int k = 0; k < kNumL; ++k) alpha[k] *= rp;
            } else alive = false;
        }
        if (alive) {
            ++pathLen;
            if (pathLen >= ps.jobs[ps.u4(s, SF_JOB).x].maxDepth) alive = false;
        }
        if (alive) {
            // albedo: sigma_s / sigma_e (DensityGridMediumDistribution::evaluateAlbedo)
            evalSpectrum(sc, sc.medSigmaS, misc.x, albedo);
            const float invE = 1.0f / sc.medSigmaE;
#pragma unroll
            for (int k = 0; k < kNumL; ++k) albedo[k] = albedo[k] * invE;
            // ---- next event estimation
            const float uSel = xsFloat(rng);
            const float lu0 = xsFloat(rng), lu1 = xsFloat(rng);
            bool envLight = false;
            float uSurf = uSel;
            if (sc.hasEnv) {
                const float surfImp = sc.lightImportance, sumImps = surfImp + 1.0f;
                if (sumImps * uSel < surfImp) uSurf = uSel / (surfImp / sumImps);
                else envLight = true;
            }
            float3 sdir = f3(0, 0, 1);
            float tmax = 0.0f, lightPDF = 0.0f, pdfConv = 1.0f, geo = 0.0f;
            float Le[kNumL];
            bool ok = false;
            if (envLight) {
                float d0, d1, uvPdf;
                envSample(sc, lu0, lu1, &d0, &d1, &uvPdf);
                const float phi = (float)(d0 * (2 * kPiD)), theta = (float)(d1 * kPiD);
                const float sinT = sinf(theta);
                const float3 lpos = f3(-sinf(phi) * sinT, cosf(theta), cosf(phi) * sinT);
                const float areaPDF = (float)(uvPdf / (2 * kPiD * kPiD * sinT));
                sdir = normalize3(lpos);
                envRadiance(sc, (float)(phi / (2 * kPiD)), (float)(theta / kPiD), misc.x, Le);
                lightPDF = envSelectProb(sc) * areaPDF;
                pdfConv = 1.0f;   // unit distance, cosine one at infinity
                geo = 1.0f;
                tmax = FLT_MAX;
                ok = areaPDF > 0.0f && isfinite(areaPDF);
            } else if (sc.numLights > 0 && sc.lightImportance > 0.0f) {
                float lprob;
                const uint32_t li = sampleLight(sc, uSurf, &lprob);
                const LightPoint L = lightPoint(sc, li, lu0, lu1);
                const float3 lp = L.lp, lg = L.lg, ln = L.ln;
                const float areaPDF = L.areaPDF;
                const float3 dv = lp - pos;
                const float dist2 = dot3(dv, dv);
                const float dist = sqrtf(dist2);
                sdir = divs(dv, dist);
                const float cosL = fabsf(dot3(sdir, lg));
                if (cosL > 0.0f && dot3(ln, -sdir) > 0.0f) {
                    evalSpectrumAt(sc, L.emit.x, L.emit.y, __float_as_uint(L.emit.z), __float_as_uint(L.emit.w),
                                   misc.x, Le);
                    const float edf = (float)(1.0 / 3.14159265358979323846);
#pragma unroll
                    for (int k = 0; k < kNumL; ++k) Le[k] *= edf;
                    lightPDF = lprob * lightSelectScale(sc) * areaPDF;
                    pdfConv = cosL / dist2;
                    geo = cosL / dist2;
                    tmax = dist * (1 - 1e-4f);
                    ok = true;
                }
            }
            if (ok) {
                const float ph = hgValue(sc.medG, dot3(rd, sdir));
                const float phPDF = ph * pdfConv;
                const float mis = (lightPDF * lightPDF) / (lightPDF * lightPDF + phPDF * phPDF);
                const float scal = ph * geo * mis / lightPDF;
                float c[kNumL];
                bool any = false;
#pragma unroll
                for (int k = 0; k < kNumL; ++k) {
                    c[k] = alpha[k] * Le[k] * albedo[k] * scal;
                    any = any || c[k] != 0.0f;
                }
                if (any) {
                    storePlanes(ps, s, SF_NEE, c);
                    ps.f4(s, SF_RO) = make_float4(pos.x, pos.y, pos.z, 1e-4f);
                    ps.f4(s, SF_SD) = make_float4(sdir.x, sdir.y, sdir.z, tmax);
                    shadow = true;
                }
            }
            // ---- phase-function sampling
            {
                const float u0 = xsFloat(rng), u1 = xsFloat(rng);
                float value;
                const float3 local = hgSample(sc.medG, u0, u1, &value);
                float3 fx, fy;
                coordinateSystem(rd, &fx, &fy);
                const float3 nd = normalize3(fx * local.x + fy * local.y + rd * local.z);
                bool nonzero = false;
#pragma unroll
                for (int k = 0; k < kNumL; ++k) {
                    alpha[k] *= albedo[k];   // albedo * phase / pdf, the phase value being its pdf
                    nonzero = nonzero || alpha[k] != 0.0f;
                }
                if (nonzero && value > 0.0f) {
                    ps.f4(s, SF_RO) = make_float4(pos.x, pos.y, pos.z, 1e-4f);
                    ps.f4(s, SF_RD) = make_float4(nd.x, nd.y, nd.z, INFINITY);
                    misc.z = value;
                    info.y &= ~F_LAST_DELTA;
                    cont = true;
                }
            }
        }
        info.y = (info.y & 0xFFFFu) | (pathLen << 16);
        if (cont || shadow) {
            if (cont) storePlanes(ps, s, SF_ALPHA, alpha);
            ps.u4(s, SF_INFO) = info;
            ps.f4(s, SF_MISC) = misc;
            ps.u4(s, SF_RNG) = rng;
            code = (uint8_t)((cont ? CODE_EXT : CODE_DONE) | (shadow ? CODE_SHADOW : 0));
        } else {
            float fin[kNumL];
            if (info.y & F_RAD) loadPlanes(ps, s, SF_SP, fin);
            else {
#pragma unroll
                for (int k = 0; k < kNumL; ++k) fin[k] = 0.0f;
            }
            code = retireAndRestart(sc, ps, s, info, misc, rng, fin, samples);
        }
        ps.code[s] = code;
    }
    statAdd(ps.stat, S_SAMPLES, samples);
}

// Ray batches through the medium: mode 0 one ratio-tracking transmittance
// estimate of [tmin, tmax] per ray (MediumObjectAggregate::evaluateTransmittance,
// libSLR/Core/medium_object.cpp:348-399, over the density grid), mode 1 one
// delta-tracking free-path distance (the scattering distance, or tmax when
// the ray leaves [tmin, tmax] unscattered; MediumObjectAggregate::interact,
// :300-347). Ray i draws from its own XORShift stream seeded from (seed, i).
__global__ __launch_bounds__(kBlock) void k_medium_query(DevScene sc, TraceArgs a, uint32_t n, int mode,
                                                         uint32_t seed, float* out) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const float3 o = f3(a.ox[i], a.oy[i], a.oz[i]), d = f3(a.dx[i], a.dy[i], a.dz[i]);
        const float t0 = a.tmin[i], t1 = a.tmax[i];
        uint32_t h = seed ^ (0x9E3779B9u * (i + 1u));
        h ^= h >> 16; h *= 0x85EBCA6Bu; h ^= h >> 13; h *= 0xC2B2AE35u; h ^= h >> 16;
        uint4 rng = xsSeed(h);
        if (mode == 0) {
            out[i] = mediumTransmittance(sc, o, d, t0, t1, rng);
        } else {
            float ts = t1;
            if (!sampleFreePath(sc, o, d, t0, t1, rng, &ts)) ts = t1;
            out[i] = ts;
        }
    }
}

int mediumQuery(DeviceScene* ds, const TraceArgs& a, uint32_t n, int mode, uint32_t seed, float* out,
                bool devicePtrs, void* stream, std::string* err) {
    if (n == 0) return 0;
    if (!ds->view.hasMedium) { if (err) *err = "scene has no medium"; return SLRH_ERR_STATE_CODE; }
    hipStream_t s = (hipStream_t)stream;
    TraceArgs d = a;
    float* dout = out;
    std::vector<void*> temps;
    auto cleanup = [&]() { for (void* p : temps) (void)hipFree(p); };
    if (!devicePtrs) {
        const float* hin[8] = {a.ox, a.oy, a.oz, a.dx, a.dy, a.dz, a.tmin, a.tmax};
        const float** din[8] = {&d.ox, &d.oy, &d.oz, &d.dx, &d.dy, &d.dz, &d.tmin, &d.tmax};
        for (int k = 0; k < 9; ++k) {
            void* p = nullptr;
            if (hipMalloc(&p, n * sizeof(float)) != hipSuccess) { cleanup(); if (err) *err = "hipMalloc"; return SLRH_ERR_NOMEM_CODE; }
            temps.push_back(p);
            if (k < 8) {
                (void)hipMemcpyAsync(p, hin[k], n * sizeof(float), hipMemcpyHostToDevice, s);
                *din[k] = (const float*)p;
            } else {
                dout = (float*)p;
            }
        }
    }
    const uint32_t blocks = std::max(1u, std::min((uint32_t)g_numCUs * 8, (n + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(k_medium_query, dim3(blocks), dim3(kBlock), 0, s, ds->view, d, n, mode, seed, dout);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { cleanup(); if (err) *err = std::string("k_medium_query: ") + hipGetErrorString(e); return SLRH_ERR_DEVICE_CODE; }
    if (!devicePtrs) (void)hipMemcpyAsync(out, dout, n * sizeof(float), hipMemcpyDeviceToHost, s);
    e = hipStreamSynchronize(s);
    cleanup();
    if (e != hipSuccess) { if (err) *err = std::string("medium query: ") + hipGetErrorString(e); return SLRH_ERR_DEVICE_CODE; }
    return 0;
}

// 2048-slot groups per compaction block: each block takes one counter
// atomic per queue, and same-address atomics serialise, so fewer and larger
// blocks shorten the pass.
constexpr uint32_t kCompactGroups = 4;

template <int MODE>
__global__ __launch_bounds__(256) void k_compact(PathState ps, uint32_t groups) {
    constexpr int PER = 8;
    constexpr int G = (int)kCompactGroups;
    constexpr int NQ = MODE == 0 ? 3 : (MODE == 1 ? 1 : kNumClasses);
    __shared__ uint32_t waveTot[G][NQ][4];
    __shared__ uint32_t groupBase[G][NQ];
    __shared__ uint32_t blockBase[NQ];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    // the traversal that consumes this compaction's queue starts its cursor at 0
    if (MODE == 1 && blockIdx.x == 0 && threadIdx.x < 2 * SLRH_FETCH_PARTS) ps.fetch[threadIdx.x * kFetchStride] = 0;
    // entries slot s contributes to queue q
    auto entries = [](int q, uint8_t v) -> uint32_t {
        const bool ext = (v & 0x7F) == CODE_EXT, sh = (v & CODE_SHADOW) != 0;
        if (MODE == 0) return q == 0 ? (v == CODE_DONE ? 1u : 0u) : q == 1 ? (ext || sh ? 1u : 0u) : (v != CODE_IDLE ? 1u : 0u);
        if (MODE == 1) return (ext ? 1u : 0u) + (sh ? 1u : 0u);
        return v == CODE_CLS0 + q ? 1u : 0u;
    };
    auto codeOf = [](const uint2& w, int k) -> uint8_t {
        return (uint8_t)(((k < 4 ? w.x : w.y) >> (8 * (k & 3))) & 0xFFu);
    };
    const uint32_t g0 = blockIdx.x * (uint32_t)G;
    uint2 w[G];
    uint32_t cnt[G][NQ], incl[G][NQ];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        // the pool is a whole number of groups: one 8-byte load per thread
        const uint32_t first = ((g0 + g) * 256u + threadIdx.x) * PER;
        w[g] = g0 + g < groups ? *(const uint2*)(ps.code + first) : make_uint2(0u, 0u);   // 0: CODE_IDLE
#pragma unroll
        for (int q = 0; q < NQ; ++q) cnt[g][q] = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint8_t c = codeOf(w[g], k);
#pragma unroll
            for (int q = 0; q < NQ; ++q) cnt[g][q] += entries(q, c);
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            uint32_t x = cnt[g][q];
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                uint32_t y = __shfl_up(x, off);
                if ((int)lane >= off) x += y;
            }
            incl[g][q] = x;
            if (lane == 63) waveTot[g][q][wave] = x;
        }
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        const int q = threadIdx.x;
        uint32_t tot = 0;
        uint32_t gt[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            gt[g] = waveTot[g][q][0] + waveTot[g][q][1] + waveTot[g][q][2] + waveTot[g][q][3];
            groupBase[g][q] = tot;
            tot += gt[g];
        }
        if (MODE == 0 && q == 2) {
            // live slots per job: one atomic per run of groups of the same job
            uint32_t runJob = kNoJob, runTot = 0;
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (g0 + g >= groups) break;
                const uint32_t job = ps.blockJob[g0 + g];
                if (job != runJob) {
                    if (runTot && runJob != kNoJob) atomicAdd(&ps.jobActive[runJob], runTot);
                    runJob = job;
                    runTot = 0;
                }
                runTot += gt[g];
            }
            if (runTot && runJob != kNoJob) atomicAdd(&ps.jobActive[runJob], runTot);
            blockBase[q] = 0;
        } else {
            const uint32_t ctr = MODE == 0 ? (q == 0 ? C_DONE : C_PENDING) : MODE == 1 ? C_TRACE : C_CLS0 + q;
            blockBase[q] = tot ? atomicAdd(&ps.ctr[ctr], tot) : 0u;
        }
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const uint32_t first = ((g0 + g) * 256u + threadIdx.x) * PER;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if (MODE == 0 && q >= 1) continue;   // counted, not queued
            uint32_t off = blockBase[q] + groupBase[g][q] + incl[g][q] - cnt[g][q];
            for (uint32_t v = 0; v < wave; ++v) off += waveTot[g][q][v];
            const uint32_t qid = MODE == 0 ? Q_DONE : MODE == 1 ? Q_TRACE : Q_CLS0 + q;
            uint32_t* out = queue(ps, qid);
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const uint8_t c = codeOf(w[g], k);
                const uint32_t s = first + k;
                if (MODE == 1) {
                    if ((c & 0x7F) == CODE_EXT) out[off++] = s;
                    if (c & CODE_SHADOW) out[off++] = s | kShadowEntry;
                } else if (entries(q, c)) {
                    out[off++] = s;
                }
            }
        }
        if (MODE == 1 && g0 + g < groups) {
            // shadow bits cleared in the codes: one 8-byte store when any was set
            const uint2 nw = make_uint2(w[g].x & 0x7F7F7F7Fu, w[g].y & 0x7F7F7F7Fu);
            if (nw.x != w[g].x || nw.y != w[g].y) *(uint2*)(ps.code + first) = nw;
        }
    }
}

// Blocks of a compaction pass over a pool of n slots (a multiple of kCompactSlots).
static inline uint32_t compactBlocks(uint32_t n) {
    return (n / kCompactSlots + kCompactGroups - 1) / kCompactGroups;
}

// Sums film planes into the output film in plane order.
__global__ void k_film_reduce(const double* planes, uint32_t numPlanes, size_t perPlane, double* out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < perPlane; i += (size_t)gridDim.x * blockDim.x) {
        double acc = out[i];
        for (uint32_t p = 0; p < numPlanes; ++p) acc += planes[(size_t)p * perPlane + i];
        out[i] = acc;
    }
}

// ------------------------------------------------------------------ render loop

// ------------------------------------------------------------------ slot pool

// Workgroup sizes of the surface shading kernels the autotuner chooses from.
constexpr uint32_t kShadeShapes[2] = {256, 128};

// A job in flight: its description, slot range and film planes.
struct JobRun {
    DevCamera cam;
    DeviceFilm* out = nullptr;
    uint32_t spp = 0, maxDepth = 100, seed = 0, planes = 1, base = 0, count = 0;
    DevBuf<double> planeFilm;   // planes > 1: per-plane sums, reduced into out at the end
};


// The wavefront's path state: a pool of slots that jobs are admitted into,
// the queues, counters and job table, and the stream (and graph) the
// per-iteration kernels run on.
struct SlotPool {
    uint32_t n = 0, blocks = 0;
    // grid of each grid-stride kernel: the blocks that can be resident at
    // once (more would wait for a second round and leave a tail)
    uint32_t gridShade[kNumClasses] = {}, gridRegen = 0, gridMedium = 0, gridMediumLds = 0;
    uint32_t gridShade1[kNumClasses] = {};   // the surface shading kernels at kShadeShapes[1]
    int shadeShape = 0;                      // index into kShadeShapes, set per render
    int device = 0;
    DevBuf<float4> ray, spec, nee;
    DevBuf<uint32_t> q, ctr, fetch, spill, jobActive, blockJob;
    DevBuf<uint8_t> code, vis;
    DevBuf<unsigned long long> stat;
    DevBuf<DevJob> jobs;
    std::vector<uint32_t> hostBlockJob;
    uint32_t* hostCtr = nullptr;   // pinned: kNumCtr counters, then kMaxJobs active counts
    hipStream_t stream = nullptr;
    // the surface shading kernels other than the diffuse one run beside it,
    // on these streams (forked after classification, joined before the last
    // compaction); every stream has the default priority
    hipStream_t side[3] = {nullptr, nullptr, nullptr};
    hipEvent_t fork = nullptr, join[3] = {nullptr, nullptr, nullptr};
    bool parallelShade = false;   // set per render (RenderOptions::parallelShade)
    hipGraphExec_t graph = nullptr;
    int graphVariant = -1;
    uint32_t graphMode = 0;     // volume | tangentNeeNaN << 1 | parallelShade << 2 | shadeShape << 3, as recorded
    PathState ps;
    DevScene sc;
    JobRun runs[kMaxJobs];
    ~SlotPool() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (stream) (void)hipStreamDestroy(stream);
        for (int i = 0; i < 3; ++i) {
            if (side[i]) (void)hipStreamDestroy(side[i]);
            if (join[i]) (void)hipEventDestroy(join[i]);
        }
        if (fork) (void)hipEventDestroy(fork);
        if (hostCtr) (void)hipHostFree(hostCtr);
    }
};

// Iterations recorded per graph launch; the host looks at the counters
// between launches.
constexpr uint32_t kGraphIters = 8;

// Blocks of kBlock threads of `kernel` that fit on the device at once
// (registers and LDS as compiled), at most `cap`.
static uint32_t residentGrid(const void* kernel, uint32_t cap, uint32_t blockSize = kBlock) {
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, (int)blockSize, 0) != hipSuccess || per <= 0)
        return cap;
    return std::max(1u, std::min(cap, (uint32_t)g_numCUs * (uint32_t)per));
}

static int poolCreate(const DeviceScene* ds, uint32_t capacity, std::unique_ptr<SlotPool>* out, std::string* err) {
    if (capacity % kCompactSlots) capacity += kCompactSlots - capacity % kCompactSlots;
    if (capacity == 0 || capacity > (1u << 26)) { if (err) *err = "slot pool size out of range"; return SLRH_ERR_CAPACITY_CODE; }
    std::unique_ptr<SlotPool> p(new SlotPool());
    const uint32_t blocks = (uint32_t)g_numCUs * 8;
    int r;
    if ((r = p->ray.alloc((size_t)kRayVec * capacity, err)) || (r = p->spec.alloc((size_t)kSpecVec * capacity, err)) ||
        (r = p->nee.alloc((size_t)kNeeVec * capacity, err)) || (r = p->q.alloc((size_t)kNumQueues * capacity, err)) ||
        (r = p->ctr.alloc(kNumCtr, err)) || (r = p->fetch.alloc(2 * SLRH_FETCH_PARTS * kFetchStride, err)) ||
        (r = p->code.alloc(capacity, err)) || (r = p->vis.alloc(capacity, err)) ||
        (r = p->stat.alloc(kNumStat * kStatSpread, err)) || (r = p->jobActive.alloc(kMaxJobs, err)) ||
        (r = p->blockJob.alloc(capacity / kCompactSlots, err)) || (r = p->jobs.alloc(kMaxJobs, err)) ||
        (r = p->spill.alloc((size_t)(kMaxStack - kMinLdsDepth) * blocks * kBlock, err)))
        return r;
    if (hipHostMalloc((void**)&p->hostCtr, (kNumCtr + kMaxJobs) * sizeof(uint32_t)) != hipSuccess) {
        if (err) *err = "hipHostMalloc";
        return SLRH_ERR_NOMEM_CODE;
    }
    HIP_TRY(hipStreamCreate(&p->stream));
    HIP_TRY(hipEventCreateWithFlags(&p->fork, hipEventDisableTiming));
    for (int i = 0; i < 3; ++i) {
        HIP_TRY(hipStreamCreateWithFlags(&p->side[i], hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&p->join[i], hipEventDisableTiming));
    }
    HIP_TRY(hipGetDevice(&p->device));
    p->n = capacity;
    p->blocks = blocks;
    p->gridShade[CLS_DIFFUSE] = residentGrid((const void*)&k_shade<CLS_DIFFUSE, kShadeShapes[0]>, blocks);
    p->gridShade[CLS_CONDUCTOR] = residentGrid((const void*)&k_shade<CLS_CONDUCTOR, kShadeShapes[0]>, blocks);
    p->gridShade[CLS_DIELECTRIC] = residentGrid((const void*)&k_shade<CLS_DIELECTRIC, kShadeShapes[0]>, blocks);
    p->gridShade[CLS_MISS] = residentGrid((const void*)&k_shade<CLS_MISS, kShadeShapes[0]>, blocks);
    p->gridShade[CLS_MEDIUM] = residentGrid((const void*)&k_shade_medium, blocks);
    const uint32_t cap1 = blocks * (kBlock / kShadeShapes[1]);
    p->gridShade1[CLS_DIFFUSE] = residentGrid((const void*)&k_shade<CLS_DIFFUSE, kShadeShapes[1]>, cap1, kShadeShapes[1]);
    p->gridShade1[CLS_CONDUCTOR] =
        residentGrid((const void*)&k_shade<CLS_CONDUCTOR, kShadeShapes[1]>, cap1, kShadeShapes[1]);
    p->gridShade1[CLS_DIELECTRIC] =
        residentGrid((const void*)&k_shade<CLS_DIELECTRIC, kShadeShapes[1]>, cap1, kShadeShapes[1]);
    p->gridShade1[CLS_MISS] = residentGrid((const void*)&k_shade<CLS_MISS, kShadeShapes[1]>, cap1, kShadeShapes[1]);
    p->gridRegen = residentGrid((const void*)&k_regen, blocks);
    p->gridMedium = residentGrid((const void*)&k_medium<false, kBlock>, blocks);
    p->gridMediumLds = residentGrid((const void*)&k_medium<true, kMediumLdsBlock>, (uint32_t)g_numCUs * 2,
                                    kMediumLdsBlock);
    p->hostBlockJob.assign(capacity / kCompactSlots, kNoJob);
    PathState& ps = p->ps;
    ps.n = capacity;
    ps.ray = p->ray.p;
    ps.spec = p->spec.p;
    ps.nee = p->nee.p;
    ps.jobs = p->jobs.p;
    ps.jobActive = p->jobActive.p;
    ps.blockJob = p->blockJob.p;
    ps.code = p->code.p;
    ps.vis = p->vis.p;
    ps.q = p->q.p;
    ps.ctr = p->ctr.p;
    ps.fetch = p->fetch.p;
    ps.stat = p->stat.p;
    ps.spill = p->spill.p;
    ps.spillStride = blocks * kBlock;
    p->sc = ds->view;
    *out = std::move(p);
    return 0;
}

// All slots idle, counters and statistics zero, no job owns a block.
static int poolReset(SlotPool& p, std::string* err) {
    hipLaunchKernelGGL(k_reset_pool, dim3(p.blocks), dim3(kBlock), 0, p.stream, p.ps);
    HIP_TRY(hipGetLastError());
    std::fill(p.hostBlockJob.begin(), p.hostBlockJob.end(), kNoJob);
    HIP_TRY(hipMemcpyAsync(p.blockJob.p, p.hostBlockJob.data(), p.hostBlockJob.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice, p.stream));
    HIP_TRY(hipStreamSynchronize(p.stream));
    return 0;
}

static uint32_t alignSlots(uint64_t n) { return (uint32_t)((n + kCompactSlots - 1) / kCompactSlots * kCompactSlots); }

static int admitJob(SlotPool& p, uint32_t idx, JobRun& j, std::string* err) {
    const uint32_t P = j.cam.width * j.cam.height;
    DevJob dj;
    dj.cam = j.cam;
    dj.pixels = P;
    dj.planes = j.planes;
    dj.maxDepth = j.maxDepth;
    dj.slotBase = j.base;
    if (j.planes > 1) {
        int r = j.planeFilm.alloc((size_t)j.planes * P * kNumL, err);
        if (r) return r;
        HIP_TRY(hipMemsetAsync(j.planeFilm.p, 0, (size_t)j.planes * P * kNumL * sizeof(double), p.stream));
        dj.film = j.planeFilm.p;
    } else {
        dj.film = j.out->data.p;
    }
    HIP_TRY(hipStreamSynchronize(p.stream));
    HIP_TRY(hipMemcpy(p.jobs.p + idx, &dj, sizeof(dj), hipMemcpyHostToDevice));
    const uint32_t b0 = j.base / kCompactSlots, b1 = alignSlots((uint64_t)j.base + j.count) / kCompactSlots;
    for (uint32_t b = b0; b < b1; ++b) p.hostBlockJob[b] = idx;
    HIP_TRY(hipMemcpy(p.blockJob.p + b0, p.hostBlockJob.data() + b0, (b1 - b0) * sizeof(uint32_t),
                      hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_init_slots, dim3(p.blocks), dim3(kBlock), 0, p.stream, p.ps, idx, j.base, P, j.planes, j.spp,
                       j.seed);
    HIP_TRY(hipGetLastError());
    return 0;
}

// The job's planes are summed into its output film and its blocks freed.
static int finishJob(SlotPool& p, JobRun& j, std::string* err) {
    if (j.planes > 1) {
        const uint32_t P = j.cam.width * j.cam.height;
        hipLaunchKernelGGL(k_film_reduce, dim3(p.blocks), dim3(kBlock), 0, p.stream, j.planeFilm.p, j.planes,
                           (size_t)P * kNumL, j.out->data.p);
        HIP_TRY(hipGetLastError());
    }
    const uint32_t b0 = j.base / kCompactSlots, b1 = alignSlots((uint64_t)j.base + j.count) / kCompactSlots;
    for (uint32_t b = b0; b < b1; ++b) p.hostBlockJob[b] = kNoJob;
    HIP_TRY(hipMemcpyAsync(p.blockJob.p + b0, p.hostBlockJob.data() + b0, (b1 - b0) * sizeof(uint32_t),
                           hipMemcpyHostToDevice, p.stream));
    return 0;
}

template <int V>
static void launchTraceV(uint32_t blocks, hipStream_t s, const DevScene& sc, const PathState& ps) {
    if (sc.numInstances)
        hipLaunchKernelGGL((k_trace<kVariants[V].ldsDepth, kVariants[V].refill, kVariants[V].fuse,
                                    kVariants[V].wavesPerEU, kVariants[V].treelet, true>),
                           dim3(blocks), dim3(kBlock), 0, s, sc, ps);
    else
        hipLaunchKernelGGL((k_trace<kVariants[V].ldsDepth, kVariants[V].refill, kVariants[V].fuse,
                                    kVariants[V].wavesPerEU, kVariants[V].treelet, false>),
                           dim3(blocks), dim3(kBlock), 0, s, sc, ps);
}
typedef void (*LaunchFn)(uint32_t, hipStream_t, const DevScene&, const PathState&);
static const LaunchFn kTraceLaunch[kNumVariants] = {launchTraceV<0>, launchTraceV<1>, launchTraceV<2>,
                                                     launchTraceV<3>, launchTraceV<4>, launchTraceV<5>,
                                                     launchTraceV<6>, launchTraceV<7>, launchTraceV<8>,
                                                     launchTraceV<9>, launchTraceV<10>};

int numTraceVariants() { return kNumVariants; }
void traceVariant(int v, int* ldsDepth, int* refill, int* fused) {
    *ldsDepth = kVariants[v].ldsDepth;
    *refill = kVariants[v].refill;
    *fused = kVariants[v].fuse ? 1 : 0;
}

// Events around the kernel groups of one iteration (kernel timing only).
// [0] start, [1] after regeneration + compaction, [2] after the traversal,
// [3] after shading, [4] after the last compaction; [5], [6] around k_medium
// (recorded only with a medium)
typedef std::array<hipEvent_t, 7> IterEvents;

// One wavefront iteration: retire/restart, trace, classify, shade, compact.
static void launchIteration(SlotPool& p, int tv, IterEvents* ev) {
    const uint32_t groups = p.n / kCompactSlots;
    const uint32_t cblocks = (groups + kCompactGroups - 1) / kCompactGroups;
    hipStream_t s = p.stream;
    if (ev) (void)hipEventRecord((*ev)[0], s);
    hipLaunchKernelGGL(k_regen, dim3(p.gridRegen), dim3(kBlock), 0, s, p.sc, p.ps);
    hipLaunchKernelGGL(k_compact<1>, dim3(cblocks), dim3(256), 0, s, p.ps, groups);
    if (ev) (void)hipEventRecord((*ev)[1], s);
    kTraceLaunch[tv](p.blocks, s, p.sc, p.ps);
    if (ev) (void)hipEventRecord((*ev)[2], s);
    if (p.sc.volume) {
        if (ev) (void)hipEventRecord((*ev)[5], s);
        if (p.sc.medCells[0] * p.sc.medCells[1] * p.sc.medCells[2] <= kMajorantLdsCells)
            hipLaunchKernelGGL((k_medium<true, kMediumLdsBlock>), dim3(p.gridMediumLds), dim3(kMediumLdsBlock), 0, s,
                               p.sc, p.ps);
        else
            hipLaunchKernelGGL((k_medium<false, kBlock>), dim3(p.gridMedium), dim3(kBlock), 0, s, p.sc, p.ps);
        if (ev) (void)hipEventRecord((*ev)[6], s);
    }
    hipLaunchKernelGGL(k_compact<2>, dim3(cblocks), dim3(256), 0, s, p.ps, groups);
    // one shading kernel per class the scene has (each reads only its own
    // queue and slots; the diffuse one also resets the counters that the
    // last compaction rebuilds, so it always runs); the others on side
    // streams when parallel, else in order on the main stream
    const uint32_t mask = p.sc.classMask;
    const int nside = p.parallelShade ? 3 : 0;
    if (nside) (void)hipEventRecord(p.fork, s);
    int used = 0;
    auto shadeOn = [&]() -> hipStream_t {
        if (!nside) return s;
        hipStream_t q = p.side[used];
        (void)hipStreamWaitEvent(q, p.fork, 0);
        return q;
    };
    auto joinSide = [&](hipStream_t q) {
        if (q != s) (void)hipEventRecord(p.join[used++], q);
    };
    auto shade = [&](auto cls, hipStream_t q) {
        constexpr uint32_t C = decltype(cls)::value;
        if (p.shadeShape == 0)
            hipLaunchKernelGGL((k_shade<C, kShadeShapes[0]>), dim3(p.gridShade[C]), dim3(kShadeShapes[0]), 0, q,
                               p.sc, p.ps);
        else
            hipLaunchKernelGGL((k_shade<C, kShadeShapes[1]>), dim3(p.gridShade1[C]), dim3(kShadeShapes[1]), 0, q,
                               p.sc, p.ps);
    };
    if (mask & (1u << CLS_CONDUCTOR)) {
        hipStream_t q = shadeOn();
        shade(std::integral_constant<uint32_t, CLS_CONDUCTOR>(), q);
        joinSide(q);
    }
    if (mask & (1u << CLS_DIELECTRIC)) {
        hipStream_t q = shadeOn();
        shade(std::integral_constant<uint32_t, CLS_DIELECTRIC>(), q);
        joinSide(q);
    }
    {
        hipStream_t q = shadeOn();
        shade(std::integral_constant<uint32_t, CLS_MISS>(), q);
        joinSide(q);
    }
    shade(std::integral_constant<uint32_t, CLS_DIFFUSE>(), s);
    if (p.sc.volume)
        hipLaunchKernelGGL(k_shade_medium, dim3(p.gridShade[CLS_MEDIUM]), dim3(kBlock), 0, s, p.sc, p.ps);
    for (int i = 0; i < used; ++i) (void)hipStreamWaitEvent(s, p.join[i], 0);
    if (ev) (void)hipEventRecord((*ev)[3], s);
    hipLaunchKernelGGL(k_compact<0>, dim3(cblocks), dim3(256), 0, s, p.ps, groups);
    if (ev) (void)hipEventRecord((*ev)[4], s);
}

// kGraphIters iterations recorded once per pool and variant, then replayed.
// The kernels take their sizes from device counters, so one recording
// serves every replay.
static int launchGraphBlock(SlotPool& p, int tv, std::string* err) {
    const uint32_t mode = p.sc.volume | p.sc.tangentNeeNaN << 1 | (p.parallelShade ? 4u : 0u) |
                          (uint32_t)p.shadeShape << 3;
    if (!p.graph || p.graphVariant != tv || p.graphMode != mode) {
        if (p.graph) { (void)hipGraphExecDestroy(p.graph); p.graph = nullptr; }
        hipGraph_t g;
        HIP_TRY(hipStreamBeginCapture(p.stream, hipStreamCaptureModeThreadLocal));
        for (uint32_t i = 0; i < kGraphIters; ++i) launchIteration(p, tv, nullptr);
        HIP_TRY(hipStreamEndCapture(p.stream, &g));
        hipError_t e = hipGraphInstantiate(&p.graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIP_TRY(e);
        p.graphVariant = tv;
        p.graphMode = mode;
    }
    HIP_TRY(hipGraphLaunch(p.graph, p.stream));
    return 0;
}

// Copies the counters and per-job active counts to the host (synchronous).
static int readCounters(SlotPool& p, std::string* err) {
    HIP_TRY(hipMemcpyAsync(p.hostCtr, p.ctr.p, kNumCtr * sizeof(uint32_t), hipMemcpyDeviceToHost, p.stream));
    HIP_TRY(hipMemcpyAsync(p.hostCtr + kNumCtr, p.jobActive.p, kMaxJobs * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           p.stream));
    HIP_TRY(hipStreamSynchronize(p.stream));
    return 0;
}

int poolOpen(DeviceScene* ds, uint32_t capacity, SlotPool** out, std::string* err) {
    std::unique_ptr<SlotPool> p;
    int r = poolCreate(ds, capacity, &p, err);
    if (r) return r;
    if ((r = poolReset(*p, err))) return r;
    p->shadeShape = ds->tunedShape;
    *out = p.release();
    return 0;
}
void poolClose(SlotPool* p) { delete p; }
int poolBindThread(SlotPool* p) { return hipSetDevice(p->device) == hipSuccess ? 0 : SLRH_ERR_DEVICE_CODE; }
uint32_t poolCapacity(const SlotPool* p) { return p->n; }

int poolAdmit(SlotPool* p, uint32_t idx, const JobSpec& spec, uint32_t planes, uint32_t base, std::string* err) {
    if (idx >= kMaxJobs || planes == 0 || base % kCompactSlots ||
        (uint64_t)base + (uint64_t)planes * spec.width * spec.height > p->n) {
        if (err) *err = "job does not fit the slot pool";
        return SLRH_ERR_CAPACITY_CODE;
    }
    JobRun& j = p->runs[idx];
    j.planeFilm.alloc(0, err);
    cameraDerived(spec.camera, spec.width, spec.height, &j.cam);
    j.out = spec.film;
    j.spp = spec.spp;
    j.maxDepth = spec.maxDepth == 0 ? 100u : spec.maxDepth;
    j.seed = spec.seed;
    j.planes = planes;
    j.base = base;
    j.count = planes * spec.width * spec.height;
    return admitJob(*p, idx, j, err);
}

int poolCancel(SlotPool* p, uint32_t idx, std::string* err) {
    JobRun& j = p->runs[idx];
    hipLaunchKernelGGL(k_idle_slots, dim3(p->blocks), dim3(kBlock), 0, p->stream, p->ps, j.base, j.count);
    HIP_TRY(hipGetLastError());
    j.planes = 1;   // nothing to reduce
    return finishJob(*p, j, err);
}

int poolRecount(SlotPool* p, std::string* err) {
    hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(256), 0, p->stream, p->ps);
    hipLaunchKernelGGL(k_compact<0>, dim3(compactBlocks(p->n)), dim3(256), 0, p->stream, p->ps, p->n / kCompactSlots);
    HIP_TRY(hipGetLastError());
    return 0;
}

int poolRunBlock(SlotPool* p, int traceVariant, std::string* err) {
    const int tv = traceVariant >= 0 && traceVariant < kNumVariants ? traceVariant : 0;
    int r = launchGraphBlock(*p, tv, err);
    if (r) return r;
    if ((r = readCounters(*p, err))) return r;
    if (p->hostCtr[C_OVERFLOW]) { if (err) *err = "traversal stack overflow (depth > 64)"; return SLRH_ERR_CAPACITY_CODE; }
    return 0;
}

const uint32_t* poolActiveCounts(const SlotPool* p) { return p->hostCtr + kNumCtr; }

int poolFinish(SlotPool* p, uint32_t idx, std::string* err) {
    int r = finishJob(*p, p->runs[idx], err);
    if (r) return r;
    HIP_TRY(hipStreamSynchronize(p->stream));
    p->runs[idx].planeFilm.alloc(0, err);
    return 0;
}

// The render statistics summed over their spread copies (synchronous).
static int readStats(const SlotPool* p, unsigned long long* stat, std::string* err) {
    std::vector<unsigned long long> all((size_t)kNumStat * kStatSpread);
    HIP_TRY(hipMemcpy(all.data(), p->stat.p, all.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < kNumStat; ++i) {
        stat[i] = 0;
        for (uint32_t j = 0; j < kStatSpread; ++j) stat[i] += all[(size_t)i * kStatSpread + j];
    }
    return 0;
}

int poolSamples(SlotPool* p, double* samples, double* rays, std::string* err) {
    unsigned long long stat[kNumStat];
    if (int r = readStats(p, stat, err)) return r;
    *samples = (double)stat[S_SAMPLES];
    *rays = (double)(stat[S_RAYS_EXT] + stat[S_RAYS_SHADOW]);
    return 0;
}

// Live paths the wavefront aims for: the traversal's per-launch tail of long
// rays and each launch's fixed costs are amortised over enough rays
// (measured on the 1080p/64 spp workload with 2, 4, 8 and 16 slots per
// pixel, i.e. 4M to 33M paths: 164, 138, 125 and 119 ms per frame); path
// state is 354 B per slot, 11.7 GB at 33M.
constexpr uint32_t kTargetSlots = 32u << 20;

// Traversal autotuner. For each variant a pilot renders one sample per slot
// of every job of the render (their cameras and resolutions, as many
// samples per pixel as each job has slot planes, into scratch films): the
// same pool, the same
// wavefront width and the same scene, so the traversal launches see the
// real mix of camera, bounce and shadow rays. Each variant's traversal
// kernel time (HIP events around every launch) per traced ray is compared
// and the fastest is kept for the scene and pool size. Pilots change no
// film the caller owns; the variants give identical hits, so the choice
// changes timing only. Pools under kTuneMinSlots keep variant 0.
constexpr uint32_t kTuneMinSlots = 1u << 20;

static int tuneTrace(DeviceScene* ds, SlotPool& p, const JobSpec* specs, uint32_t n, const uint32_t* planes,
                     const uint32_t* bases, std::string* err) {
    // every job of the render takes part (one sample per slot, scratch
    // films), so the pilot's wavefront is as wide as the render's
    std::vector<std::unique_ptr<DeviceFilm, void (*)(DeviceFilm*)>> scratch;
    std::vector<JobSpec> pilots(specs, specs + n);
    int r = 0;
    for (uint32_t i = 0; i < n; ++i) {
        DeviceFilm* f = nullptr;
        if ((r = filmCreate(specs[i].width, specs[i].height, &f, err))) return r;
        scratch.emplace_back(f, filmFree);
        pilots[i].spp = planes[i];
        pilots[i].film = f;
    }
    std::vector<hipEvent_t> evs;
    auto cleanup = [&]() { for (hipEvent_t e : evs) (void)hipEventDestroy(e); };
    // one pilot with traversal variant v and shading shape sh: kernel
    // milliseconds of the traversal and of the shading group, traced rays
    // and extension rays (one shaded vertex each)
    struct Pilot { double traceMs = 0, shadeMs = 0, rays = 0, ext = 0; };
    auto pilot = [&](int v, int sh, Pilot* out) -> int {
        int r;
        p.shadeShape = sh;
        if ((r = poolReset(p, err))) return r;
        for (uint32_t i = 0; i < n; ++i)
            if ((r = poolAdmit(&p, i, pilots[i], planes[i], bases[i], err))) return r;
        hipLaunchKernelGGL(k_compact<0>, dim3(compactBlocks(p.n)), dim3(256), 0, p.stream, p.ps, p.n / kCompactSlots);
        std::vector<IterEvents> itEv;
        for (;;) {
            for (uint32_t k = 0; k < kGraphIters; ++k) {
                itEv.push_back({});
                for (auto& x : itEv.back()) {
                    (void)hipEventCreate(&x);
                    evs.push_back(x);
                }
                launchIteration(p, v, &itEv.back());
            }
            if ((r = readCounters(p, err))) return r;
            if (p.hostCtr[C_OVERFLOW] || (p.hostCtr[C_DONE] == 0 && p.hostCtr[C_PENDING] == 0)) break;
        }
        for (uint32_t i = 0; i < n; ++i)
            if ((r = finishJob(p, p.runs[i], err))) return r;
        HIP_TRY(hipStreamSynchronize(p.stream));
        unsigned long long stat[kNumStat];
        if ((r = readStats(&p, stat, err))) return r;
        *out = Pilot();
        for (auto& e : itEv) {
            float t = 0;
            (void)hipEventElapsedTime(&t, e[1], e[2]);
            out->traceMs += t;
            (void)hipEventElapsedTime(&t, e[2], e[3]);
            out->shadeMs += t;
        }
        out->rays = (double)(stat[S_RAYS_EXT] + stat[S_RAYS_SHADOW]);
        out->ext = (double)stat[S_RAYS_EXT];
        for (hipEvent_t e : evs) (void)hipEventDestroy(e);
        evs.clear();
        return 0;
    };
    // traversal variants by traversal time per traced ray
    double best = 0;
    int bestV = 0;
    for (int v = 0; v < kNumVariants; ++v) {
        Pilot pl;
        if ((r = pilot(v, 0, &pl))) { cleanup(); return r; }
        const double perRay = pl.rays > 0 ? pl.traceMs / pl.rays : 0;
        if (v == 0 || perRay < best) { best = perRay; bestV = v; }
    }
    // shading workgroup shape by shading time per shaded vertex
    int bestS = 0;
    double bestShade = 0;
    for (int sh = 0; sh < 2; ++sh) {
        Pilot pl;
        if ((r = pilot(bestV, sh, &pl))) { cleanup(); return r; }
        const double perVertex = pl.ext > 0 ? pl.shadeMs / pl.ext : 0;
        if (sh == 0 || perVertex < bestShade) { bestShade = perVertex; bestS = sh; }
    }
    ds->tunedVariant = bestV;
    ds->tunedShape = bestS;
    ds->tunedSlots = p.n;
    return 0;
}

int tunedTraceVariant(const DeviceScene* ds) { return ds->tunedVariant >= 0 ? ds->tunedVariant : 0; }

// Slots per pixel for a wavefront of totalPixels aiming at `target` live
// paths: at least 1, at most 16.
static uint32_t slotsPerPixelFor(uint64_t target, uint64_t totalPixels) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(16, (target + totalPixels - 1) / totalPixels));
}

// Pool-width tuning (opt-in, RenderOptions::tunePool): for each candidate
// wavefront width a pilot renders two samples per slot of every job of the
// render into scratch films, and the width with the most samples per second
// of device time is kept for the scene. Films depend on the slots per pixel
// (each slot draws its own random stream), so a tuned width makes them
// depend on this timing: hence opt-in.
static const uint32_t kPoolWidths[3] = {8u << 20, 16u << 20, 32u << 20};

static int tunePoolWidth(DeviceScene* ds, const JobSpec* specs, uint32_t n, const RenderOptions& opt,
                         std::string* err) {
    uint64_t totalPixels = 0;
    for (uint32_t i = 0; i < n; ++i) totalPixels += (uint64_t)specs[i].width * specs[i].height;
    std::vector<std::unique_ptr<DeviceFilm, void (*)(DeviceFilm*)>> scratch;
    std::vector<JobSpec> pilots(specs, specs + n);
    int r = 0;
    for (uint32_t i = 0; i < n; ++i) {
        DeviceFilm* f = nullptr;
        if ((r = filmCreate(specs[i].width, specs[i].height, &f, err))) return r;
        scratch.emplace_back(f, filmFree);
        pilots[i].film = f;
    }
    double best = 0;
    uint32_t bestW = kTargetSlots, lastK = 0;
    for (uint32_t w : kPoolWidths) {
        const uint32_t K = slotsPerPixelFor(w, totalPixels);
        if (K == lastK) continue;   // the same wavefront as a smaller width
        lastK = K;
        for (uint32_t i = 0; i < n; ++i) pilots[i].spp = std::min(specs[i].spp, 2 * K);
        RenderOptions po = opt;
        po.tunePool = false;
        po.slotsPerPixel = K;
        po.timeKernels = false;
        RenderStats ps;
        if ((r = renderJobs(ds, pilots.data(), n, po, &ps, err))) return r;
        const double rate = ps.kernelSeconds > 0 ? ps.samples / ps.kernelSeconds : 0;
        if (rate > best) { best = rate; bestW = w; }
    }
    ds->tunedPoolWidth = bestW;
    return 0;
}

int renderJobs(DeviceScene* ds, const JobSpec* specs, uint32_t n, const RenderOptions& opt, RenderStats* st,
               std::string* err) {
    if (n == 0) return 0;
    if (n > kMaxJobs) { if (err) *err = "too many frames in one batch"; return SLRH_ERR_CAPACITY_CODE; }
    std::vector<uint32_t> planes(n), bases(n);
    uint64_t totalPixels = 0;
    for (uint32_t i = 0; i < n; ++i) totalPixels += (uint64_t)specs[i].width * specs[i].height;
    if (totalPixels == 0) return 0;
    uint32_t K = opt.slotsPerPixel;
    if (K == 0) {
        uint32_t target = kTargetSlots;
        if (opt.tunePool) {
            if (ds->tunedPoolWidth == 0) {
                const int r = tunePoolWidth(ds, specs, n, opt, err);
                if (r) return r;
            }
            target = ds->tunedPoolWidth;
        }
        K = slotsPerPixelFor(target, totalPixels);
    }
    uint64_t capacity = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const JobSpec& js = specs[i];
        planes[i] = std::max(1u, std::min(K, js.spp));
        bases[i] = (uint32_t)std::min<uint64_t>(capacity, UINT32_MAX);
        capacity += alignSlots((uint64_t)planes[i] * js.width * js.height);
    }
    if (capacity > (1u << 26)) { if (err) *err = "too many path slots"; return SLRH_ERR_CAPACITY_CODE; }
    // the pool is kept between renders; it is rebuilt when too small, or when
    // much larger than needed (every iteration scans all of its slots)
    if (!ds->pool || ds->pool->n < capacity || ds->pool->n > 2 * capacity + 64 * kCompactSlots) {
        ds->pool.reset();
        std::unique_ptr<SlotPool> np;
        int r = poolCreate(ds, (uint32_t)capacity, &np, err);
        if (r) return r;
        ds->pool = std::shared_ptr<SlotPool>(np.release());
    }
    SlotPool& p = *ds->pool;
    int r = 0;
    // the volumetric integrator: media take part, and the path length has no
    // fixed cap (VolumetricPTRenderer.cpp:184-293 ends paths by Russian
    // roulette only) unless the settings set one
    p.sc.volume = opt.volumetric && ds->view.hasMedium ? 1u : 0u;
    p.sc.tangentNeeNaN = opt.volumetric && ds->view.compatTangentNee ? 1u : 0u;
    p.parallelShade = opt.parallelShade;
    std::vector<JobSpec> jobSpecs(specs, specs + n);
    if (opt.volumetric)
        for (JobSpec& js : jobSpecs)
            if (js.maxDepth == 0) js.maxDepth = 0xFFFFu;
    specs = jobSpecs.data();
    int tv = opt.traceVariant >= 0 && opt.traceVariant < kNumVariants ? opt.traceVariant : 0;
    if (opt.traceVariant < 0) {
        // automatic: tuned once per scene, by the first render whose pool is
        // large enough for the pilot to time the traversal at full occupancy;
        // later renders of any size reuse the choice (a server's requests
        // vary in size, and a pilot per pool size would cost more than it saves)
        if (p.n >= kTuneMinSlots && ds->tunedVariant < 0)
            if ((r = tuneTrace(ds, p, specs, n, planes.data(), bases.data(), err))) return r;
        tv = ds->tunedVariant >= 0 ? ds->tunedVariant : 0;
    }
    p.shadeShape = opt.shadeShape >= 0 ? opt.shadeShape : opt.traceVariant < 0 ? ds->tunedShape : 0;
    if ((r = poolReset(p, err))) return r;

    std::vector<hipEvent_t> evs;
    auto newEvent = [&]() { hipEvent_t e; (void)hipEventCreate(&e); evs.push_back(e); return e; };
    std::vector<IterEvents> itEv;
    hipEvent_t tStart = newEvent(), tEnd = newEvent();
    for (uint32_t i = 0; i < n; ++i)
        if ((r = poolAdmit(&p, i, specs[i], planes[i], bases[i], err))) return r;
    HIP_TRY(hipEventRecord(tStart, p.stream));
    hipLaunchKernelGGL(k_compact<0>, dim3(compactBlocks(p.n)), dim3(256), 0, p.stream, p.ps, p.n / kCompactSlots);
    HIP_TRY(hipGetLastError());
    uint32_t it = 0;
    for (;;) {
        if (opt.timeKernels || !opt.useGraph) {
            for (uint32_t k = 0; k < kGraphIters; ++k) {
                IterEvents* e = nullptr;
                if (opt.timeKernels) {
                    itEv.push_back({});
                    for (auto& x : itEv.back()) x = newEvent();
                    e = &itEv.back();
                }
                launchIteration(p, tv, e);
            }
            HIP_TRY(hipGetLastError());
        } else {
            if ((r = launchGraphBlock(p, tv, err))) return r;
        }
        it += kGraphIters;
        if ((r = readCounters(p, err))) return r;
        if (p.hostCtr[C_OVERFLOW]) { if (err) *err = "traversal stack overflow (depth > 64)"; break; }
        if (p.hostCtr[C_DONE] == 0 && p.hostCtr[C_PENDING] == 0) break;
    }
    for (uint32_t i = 0; i < n; ++i)
        if ((r = finishJob(p, p.runs[i], err))) return r;
    HIP_TRY(hipEventRecord(tEnd, p.stream));
    HIP_TRY(hipStreamSynchronize(p.stream));
    unsigned long long stat[kNumStat];
    if (int rs = readStats(&p, stat, err)) return rs;
    HIP_TRY(hipMemcpy(p.hostCtr, p.ctr.p, kNumCtr * sizeof(uint32_t), hipMemcpyDeviceToHost));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, tStart, tEnd);
    if (st) {
        st->raysTraced = (double)(stat[S_RAYS_EXT] + stat[S_RAYS_SHADOW]);
        st->extendRays = (double)stat[S_RAYS_EXT];
        st->shadowRays = (double)stat[S_RAYS_SHADOW];
        st->samples = (double)stat[S_SAMPLES];
        st->nodesVisited = (double)stat[S_NODES_EXT];
        st->trisTested = (double)stat[S_TRIS_EXT];
        st->nodesShadow = (double)stat[S_NODES_SH];
        st->trisShadow = (double)stat[S_TRIS_SH];
        st->mediumEvents = (double)stat[S_MED_EVENTS];
        st->mediumShadows = (double)stat[S_MED_SHADOWS];
        st->mediumLookups = (double)stat[S_MED_LOOKUPS];
        st->mediumCells = (double)stat[S_MED_CELLS];
        st->mediumRays = (double)stat[S_MED_RAYS];
        st->mediumGuard = (double)stat[S_MED_GUARD];
        st->nodesLds = (double)stat[S_NODES_LDS];
        st->instEntered = (double)stat[S_INSTS];
        st->kernelSeconds = ms * 1e-3;
        st->nonFinite = p.hostCtr[C_NONFINITE];
        st->overflow = p.hostCtr[C_OVERFLOW];
        st->iterations = it;
        double tr = 0, sh = 0, ot = 0, md = 0;
        auto el = [](hipEvent_t a, hipEvent_t b) { float t = 0; (void)hipEventElapsedTime(&t, a, b); return (double)t; };
        for (auto& e : itEv) {
            ot += el(e[0], e[1]) + el(e[3], e[4]);
            tr += el(e[1], e[2]);
            sh += el(e[2], e[3]);
            if (p.sc.volume) md += el(e[5], e[6]);
        }
        st->mediumSeconds = md * 1e-3;
        st->traceSeconds = tr * 1e-3;
        st->shadeSeconds = sh * 1e-3;
        st->otherSeconds = ot * 1e-3;
        st->traceLaunches = (double)itEv.size();
        st->traceVariant = tv;
        st->shadeBlock = kShadeShapes[p.shadeShape];
        st->slotsPerPixel = K;
    }
    for (hipEvent_t e : evs) (void)hipEventDestroy(e);
    if (p.hostCtr[C_OVERFLOW]) return SLRH_ERR_CAPACITY_CODE;
    return 0;
}

int render(DeviceScene* ds, const RenderConfig& cfg, DeviceFilm* film, RenderStats* st, std::string* err) {
    if (!ds->hasCamera) { if (err) *err = "scene has no camera"; return -2; }
    if (film->w != cfg.width || film->h != cfg.height) { if (err) *err = "film size differs from settings"; return -1; }
    if (cfg.width * cfg.height == 0 || cfg.spp == 0) return 0;
    JobSpec js;
    js.camera = ds->camera;
    js.width = cfg.width;
    js.height = cfg.height;
    js.spp = cfg.spp;
    js.maxDepth = cfg.maxDepth;
    js.seed = cfg.seed;
    js.film = film;
    RenderOptions opt;
    opt.slotsPerPixel = cfg.slotsPerPixel;
    opt.timeKernels = cfg.timeKernels;
    opt.traceVariant = cfg.traceVariant;
    opt.useGraph = cfg.useGraph;
    opt.parallelShade = cfg.parallelShade;
    opt.shadeShape = cfg.shadeShape;
    opt.tunePool = cfg.tunePool;
    opt.volumetric = cfg.volumetric;
    return renderJobs(ds, &js, 1, opt, st, err);
}

}  // namespace slrh
