// C ABI of libslr_hip.so (declared in include/slr_hip.h).
//
// Host-side responsibilities: argument checking, scene assembly (meshes,
// spectra, materials, camera), the SBVH/QBVH build, and turning device
// results into the film formats. Errors are reported as status codes with a
// message kept on the context; no exception leaves this file.
#include "../../include/slr_hip.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"
#include "scene_reader.h"
#include "server.h"

using namespace slrh;

// Calls on the scenes and films of one context serialise on its lock
// (recursive: an entry point may call another); the last error is kept per
// calling thread and context, like errno.
struct slrh_context {
    int device = 0;
    std::recursive_mutex mu;
};

struct slrh_scene {
    slrh_context* ctx = nullptr;
    SceneData data;
    DeviceScene* dev = nullptr;
    bool built = false;
    uint32_t numMeshes = 0;
    int queryAccel = SLRH_QUERY_QBVH;   // accelerator slrh_intersect / slrh_occluded use
    double lastNodes = 0, lastTris = 0;   // traversal work of the last QBVH ray batch
};

struct slrh_film {
    slrh_context* ctx = nullptr;
    DeviceFilm* dev = nullptr;
    uint32_t w = 0, h = 0;
};

struct slrh_server {
    slrh_context* ctx = nullptr;
    std::unique_ptr<Server> srv;
};

namespace {

thread_local std::unordered_map<const slrh_context*, std::string> tlsLastError;

int fail(slrh_context* c, int code, const std::string& msg) {
    if (c) tlsLastError[c] = msg;
    return code;
}

typedef std::unique_lock<std::recursive_mutex> CtxLock;
CtxLock lockOf(const slrh_context* c) { return c ? CtxLock(const_cast<slrh_context*>(c)->mu) : CtxLock(); }
template <typename T>
CtxLock lockOf(const T* obj) { return obj ? lockOf(obj->ctx) : CtxLock(); }

// makeCoordinateSystem of a unit normal (libSLR/BasicTypes/Vector3D.h:61-67)
void tangentOf(const float n[3], float t[3]) {
    float sign = n[2] >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sign + n[2]);
    const float b = n[0] * n[1] * a;
    (void)b;
    t[0] = 1.0f + sign * n[0] * n[0] * a;
    t[1] = sign * b;
    t[2] = -sign * n[0];
}

int addSpectrum(slrh_scene* s, HostSpectrum&& h, int* id) {
    s->data.spectra.push_back(std::move(h));
    *id = (int)s->data.spectra.size() - 1;
    return SLRH_OK;
}

bool validSpectrum(const slrh_scene* s, int id) { return id >= 0 && id < (int)s->data.spectra.size(); }

int addMaterial(slrh_scene* s, const DevMaterial& m, int* id) {
    s->data.materials.push_back(m);
    *id = (int)s->data.materials.size() - 1;
    return SLRH_OK;
}

#define CHECK_SCENE(s)                                                                  \
    do {                                                                                \
        if (!(s)) return SLRH_ERR_INVALID;                                              \
        if ((s)->built) return fail((s)->ctx, SLRH_ERR_STATE, "scene is already built"); \
    } while (0)

}  // namespace

extern "C" {

int slrh_version(void) { return 100; }

int slrh_create(int device, slrh_context** out) {
    if (!out) return SLRH_ERR_INVALID;
    std::unique_ptr<slrh_context> c(new (std::nothrow) slrh_context());
    if (!c) return SLRH_ERR_NOMEM;
    c->device = device;
    if (device >= 0) {   // device < 0: host-only context (scene assembly and accelerator inspection)
        std::string err;
        int r = deviceInit(device, &err);
        if (r) return r;
    }
    *out = c.release();
    return SLRH_OK;
}

int slrh_destroy(slrh_context* ctx) {
    tlsLastError.erase(ctx);
    delete ctx;
    return SLRH_OK;
}

const char* slrh_last_error(const slrh_context* ctx) {
    if (!ctx) return "null context";
    auto it = tlsLastError.find(ctx);
    return it == tlsLastError.end() ? "" : it->second.c_str();
}

int slrh_scene_create(slrh_context* ctx, slrh_scene** out) {
    if (!ctx || !out) return SLRH_ERR_INVALID;
    slrh_scene* s = new (std::nothrow) slrh_scene();
    if (!s) return SLRH_ERR_NOMEM;
    s->ctx = ctx;
    *out = s;
    return SLRH_OK;
}

int slrh_scene_destroy(slrh_scene* s) {
    CtxLock lk = lockOf(s);
    if (!s) return SLRH_ERR_INVALID;
    if (s->dev) freeScene(s->dev);
    delete s;
    return SLRH_OK;
}

int slrh_spectrum_rgb(slrh_scene* s, int type, float r, float g, float b, int* id) {
    CtxLock lk = lockOf(s);
    CHECK_SCENE(s);
    if (!id || (type != SLRH_SPECTRUM_REFLECTANCE && type != SLRH_SPECTRUM_LIGHT_SOURCE))
        return fail(s->ctx, SLRH_ERR_INVALID, "slrh_spectrum_rgb: bad type");
    if (!std::isfinite(r) || !std::isfinite(g) || !std::isfinite(b))
        return fail(s->ctx, SLRH_ERR_INVALID, "slrh_spectrum_rgb: non-finite value");
    MengSpectrum m = mengFromRGB(type, r, g, b);
    return addSpectrum(s, mengToRegular(m), id);
}

int slrh_spectrum_regular(slrh_scene* s, float lo, float hi, const float* values, int n, int* id) {
    CtxLock lk = lockOf(s);
    CHECK_SCENE(s);
    if (!id || !values || n < 1 || !(hi > lo) || (n > 1 && !(hi > lo)))
        return fail(s->ctx, SLRH_ERR_INVALID, "slrh_spectrum_regular: bad arguments");
    HostSpectrum h;
    h.lo = lo;
    h.hi = hi;
    h.values.assign(values, values + n);
    return addSpectrum(s, std::move(h), id);
}

int slrh_spectrum_evaluate(const slrh_scene* s, int id, int n, const float* lambdas, float* out) {
    CtxLock lk = lockOf(s);
    if (!s || !lambdas || !out || n < 0) return SLRH_ERR_INVALID;
    if (!validSpectrum(s, id)) return fail(s->ctx, SLRH_ERR_INVALID, "spectrum_evaluate: bad id");
    const HostSpectrum& h = s->data.spectra[id];
    const int m = (int)h.values.size();
    for (int i = 0; i < n; ++i) {
        float binF = (lambdas[i] - h.lo) / (h.hi - h.lo) * (m - 1);
        if (binF <= 0.0f) { out[i] = h.values[0]; continue; }
        if (binF >= m - 1) { out[i] = h.values[m - 1]; continue; }
        int b = (int)binF;
        float t = binF - b;
        out[i] = (1 - t) * h.values[b] + t * h.values[b + 1];
    }
    return SLRH_OK;
}

int slrh_cmf_strata(float* x, float* y, float* z, float* integral) {
    if (!x || !y || !z || !integral) return SLRH_ERR_INVALID;
    cmfStrata(x, y, z, integral);
    return SLRH_OK;
}

int slrh_material_lambert(slrh_scene* s, int refl, int* id) {
    CtxLock lk = lockOf(s);
    CHECK_SCENE(s);
    if (!id || !validSpectrum(s, refl)) return fail(s->ctx, SLRH_ERR_INVALID, "lambert: bad spectrum");
    DevMaterial m{};
    m.type = MAT_LAMBERT; m.spec0 = refl; m.spec1 = -1; m.emit = -1; m.cls = CLS_DIFFUSE;
    return addMaterial(s, m, id);
}

int slrh_material_oren_nayar(slrh_scene* s, int refl, float sigma, int* id) {
    CtxLock lk = lockOf(s);
    CHECK_SCENE(s);
    if (!id || !validSpectrum(s, refl) || !(sigma >= 0)) return fail(s->ctx, SLRH_ERR_INVALID, "oren_nayar: bad args");
    DevMaterial m{};
    m.type = MAT_OREN_NAYAR; m.spec0 = refl; m.spec1 = -1; m.emit = -1; m.cls = CLS_DIFFUSE;
    // OrenNayerBRDF's A and B (libSLR/BSDF/OrenNayerBRDF.h:28-30), double literals as there
    m.p0 = (float)(1.0f - 0.5f * sigma * sigma / (sigma * sigma + 0.33));
    m.p1 = (float)(0.45 * sigma * sigma / (sigma * sigma + 0.09));
    return addMaterial(s, m, id);
}

int slrh_material_ggx_conductor(slrh_scene* s, int eta, int k, float ax, float ay, int* id) {
    CtxLock lk = lockOf(s);
    CHECK_SCENE(s);
    if (!id || !validSpectrum(s, eta) || !validSpectrum(s, k) || !(ax > 0) || !(ay > 0))
        return fail(s->ctx, SLRH_ERR_INVALID, "ggx_conductor: bad args");
    DevMaterial m{};
    m.type = MAT_GGX_CONDUCTOR; m.spec0 = eta; m.spec1 = k; m.emit = -1; m.p0 = ax; m.p1 = ay; m.cls = CLS_CONDUCTOR;
    return addMaterial(s, m, id);
}

int slrh_material_ggx_dielectric(slrh_scene* s, int etaExt, int etaInt, float ax, float ay, int* id) {
    CtxLock lk = lockOf(s);
    CHECK_SCENE(s);
    if (!id || !validSpectrum(s, etaExt) || !validSpectrum(s, etaInt) || !(ax > 0) || !(ay > 0))
        return fail(s->ctx, SLRH_ERR_INVALID, "ggx_dielectric: bad args");
    DevMaterial m{};
    m.type = MAT_GGX_DIELECTRIC; m.spec0 = etaExt; m.spec1 = etaInt; m.emit = -1; m.p0 = ax; m.p1 = ay;
    m.cls = CLS_DIELECTRIC;
    return addMaterial(s, m, id);
}

int slrh_material_diffuse_emitter(slrh_scene* s, int base, int emittance, int* id) {
    CtxLock lk = lockOf(s);
    CHECK_SCENE(s);
    if (!id || !validSpectrum(s, emittance) || base >= (int)s->data.materials.size() || base < -1)
        return fail(s->ctx, SLRH_ERR_INVALID, "diffuse_emitter: bad args");
    DevMaterial m{};
    if (base >= 0) {
        m = s->data.materials[base];
        if (m.emit >= 0) return fail(s->ctx, SLRH_ERR_INVALID, "diffuse_emitter: base material already emits");
    } else {
        m.type = MAT_NONE; m.spec0 = m.spec1 = -1; m.cls = CLS_DIFFUSE;   // emits, then the path ends
    }
    m.emit = emittance;
    return addMaterial(s, m, id);
}

}  // extern "C"

namespace {
// Appends a mesh's vertices and triangles to the scene's global arrays:
// in world space (proto < 0) or as prototype `proto` in object space.
int addMesh(slrh_scene* s, uint32_t nverts, const float* pos, const float* nrm, const float* tan, uint32_t ntris,
            const uint32_t* idx, const uint32_t* mat, int material, int32_t proto) {
    if (!pos || !idx || nverts == 0 || ntris == 0) return fail(s->ctx, SLRH_ERR_INVALID, "mesh_add: empty mesh");
    const uint32_t nm = (uint32_t)s->data.materials.size();
    if (!mat && (material < 0 || (uint32_t)material >= nm)) return fail(s->ctx, SLRH_ERR_INVALID, "mesh_add: bad material");
    for (uint64_t i = 0; i < 3ull * ntris; ++i)
        if (idx[i] >= nverts) return fail(s->ctx, SLRH_ERR_INVALID, "mesh_add: vertex index out of range");
    if (mat)
        for (uint32_t i = 0; i < ntris; ++i)
            if (mat[i] >= nm) return fail(s->ctx, SLRH_ERR_INVALID, "mesh_add: material id out of range");
    for (uint64_t i = 0; i < 3ull * nverts; ++i)
        if (!std::isfinite(pos[i])) return fail(s->ctx, SLRH_ERR_INVALID, "mesh_add: non-finite position");
    SceneData& d = s->data;
    const uint32_t base = (uint32_t)(d.vpos.size() / 3);
    if ((uint64_t)base + nverts > 0xFFFFFFFFull) return fail(s->ctx, SLRH_ERR_CAPACITY, "mesh_add: too many vertices");
    d.vpos.insert(d.vpos.end(), pos, pos + 3ull * nverts);
    std::vector<float> n(3ull * nverts, 0.0f);
    if (nrm) {
        n.assign(nrm, nrm + 3ull * nverts);
    } else {
        // smooth normals: sum of unit face normals (libSLRSceneGraph/Scene/TriangleMeshNode.cpp:106-131)
        for (uint32_t t = 0; t < ntris; ++t) {
            const float* p0 = pos + 3ull * idx[3 * t];
            const float* p1 = pos + 3ull * idx[3 * t + 1];
            const float* p2 = pos + 3ull * idx[3 * t + 2];
            float e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
            float e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
            float c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            float l = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            if (!(l > 0)) continue;
            for (int k = 0; k < 3; ++k)
                for (int j = 0; j < 3; ++j) n[3ull * idx[3 * t + k] + j] += c[j] / l;
        }
        for (uint32_t v = 0; v < nverts; ++v) {
            float* q = &n[3ull * v];
            float l = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
            if (l > 0) { q[0] /= l; q[1] /= l; q[2] /= l; } else { q[0] = 0; q[1] = 0; q[2] = 1; }
        }
    }
    d.vnrm.insert(d.vnrm.end(), n.begin(), n.end());
    if (tan) {
        d.vtan.insert(d.vtan.end(), tan, tan + 3ull * nverts);
    } else {
        for (uint32_t v = 0; v < nverts; ++v) {
            float t[3];
            tangentOf(&n[3ull * v], t);
            d.vtan.insert(d.vtan.end(), t, t + 3);
        }
    }
    for (uint64_t i = 0; i < 3ull * ntris; ++i) d.idx.push_back(base + idx[i]);
    for (uint32_t i = 0; i < ntris; ++i) d.triMat.push_back(mat ? mat[i] : (uint32_t)material);
    d.triProto.insert(d.triProto.end(), ntris, proto);
    return SLRH_OK;
}

// StaticTransform * Point3D for an affine matrix (Matrix4x4::operator*,
// libSLR/BasicTypes/Matrix4x4.h:68-74): each row dotted with (p, 1) in order.
V3 affinePoint(const float* m, const V3& p) {
    float o[3];
    for (int r = 0; r < 3; ++r) o[r] = m[4 * r] * p.x + m[4 * r + 1] * p.y + m[4 * r + 2] * p.z + m[4 * r + 3] * 1.0f;
    return v3(o[0], o[1], o[2]);
}
// Matrix4x4 invert (libSLR/BasicTypes/Matrix4x4.cpp:36-87) restated in
// FP32 step by step: Gauss-Jordan elimination with full pivoting over the
// columns not yet done, then the recorded row swaps undone as column swaps
// in reverse order. The instance's world-to-object rows are this inverse,
// so rays reach object space with exactly the reference's operands.
// m, out: [column][row]. Returns false for a zero pivot.
bool invertMatrix(const float m[4][4], float out[4][4]) {
    float a[4][4];
    memcpy(a, m, sizeof(a));
    bool colDone[4] = {false, false, false, false};
    int swapA[4] = {0, 0, 0, 0}, swapB[4] = {0, 0, 0, 0};
    for (int pass = 0; pass < 4; ++pass) {
        int pvCol = 0, pvRow = 0;
        float maxPivot = -1.0f;
        for (int c = 0; c < 4; ++c) {
            if (colDone[c]) continue;
            for (int r = 0; r < 4; ++r) {
                if (colDone[r]) continue;
                const float v = std::fabs(a[c][r]);
                if (v > maxPivot) { pvCol = c; pvRow = r; maxPivot = v; }
            }
        }
        if (pvRow != pvCol)
            for (int k = 0; k < 4; ++k) std::swap(a[k][pvRow], a[k][pvCol]);
        swapA[pass] = pvRow;
        swapB[pass] = pvCol;
        const float pivot = a[pvCol][pvCol];
        if (pivot == 0.0f) return false;
        a[pvCol][pvCol] = 1.0f;
        const float scale = 1.0f / pivot;
        for (int k = 0; k < 4; ++k) a[k][pvCol] *= scale;
        const float addend[4] = {a[0][pvCol], a[1][pvCol], a[2][pvCol], a[3][pvCol]};
        for (int r = 0; r < 4; ++r) {
            if (r == pvCol) continue;
            const float sv = a[pvCol][r];
            a[pvCol][r] = 0.0f;
            for (int k = 0; k < 4; ++k) a[k][r] += -sv * addend[k];
        }
        colDone[pvCol] = true;
    }
    for (int pass = 3; pass >= 0; --pass)
        if (swapA[pass] != swapB[pass])
            for (int k = 0; k < 4; ++k) std::swap(a[swapA[pass]][k], a[swapB[pass]][k]);
    memcpy(out, a, sizeof(a));
    return true;
}
}  // namespace

namespace slrh {
bool invertMatrix4(const float m[4][4], float out[4][4]) { return invertMatrix(m, out); }
void coordinateTangent(const float n[3], float t[3]) { tangentOf(n, t); }
}  // namespace slrh

extern "C" {

int slrh_mesh_add(slrh_scene* s, uint32_t nverts, const float* pos, const float* nrm, const float* tan,
                  uint32_t ntris, const uint32_t* idx, const uint32_t* mat, int material, int* mesh_id) {
    CtxLock lk = lockOf(s);
    CHECK_SCENE(s);
    int r = addMesh(s, nverts, pos, nrm, tan, ntris, idx, mat, material, -1);
    if (r) return r;
    if (mesh_id) *mesh_id = (int)s->numMeshes;
    ++s->numMeshes;
    return SLRH_OK;
}

int slrh_prototype_add(slrh_scene* s, uint32_t nverts, const float* pos, const float* nrm, const float* tan,
                       uint32_t ntris, const uint32_t* idx, const uint32_t* mat, int material, int* proto_id) {
    CtxLock lk = lockOf(s);
    CHECK_SCENE(s);
    SceneData& d = s->data;
    if (d.protos.size() >= (1u << 20)) return fail(s->ctx, SLRH_ERR_CAPACITY, "prototype_add: too many prototypes");
    Prototype p;
    p.triFirst = (uint32_t)d.triMat.size();
    p.triCount = ntris;
    int r = addMesh(s, nverts, pos, nrm, tan, ntris, idx, mat, material, (int32_t)d.protos.size());
    if (r) return r;
    d.protos.push_back(std::move(p));
    if (proto_id) *proto_id = (int)d.protos.size() - 1;
    return SLRH_OK;
}

int slrh_instance_add(slrh_scene* s, int proto_id, const float* l2w, int material, int* inst_id) {
    CtxLock lk = lockOf(s);
    CHECK_SCENE(s);
    SceneData& d = s->data;
    if (proto_id < 0 || (size_t)proto_id >= d.protos.size()) return fail(s->ctx, SLRH_ERR_INVALID, "instance_add: bad prototype");
    if (!l2w) return fail(s->ctx, SLRH_ERR_INVALID, "instance_add: null transform");
    if (material < -1 || (material >= 0 && (size_t)material >= d.materials.size()))
        return fail(s->ctx, SLRH_ERR_INVALID, "instance_add: bad material");
    if (d.instances.size() >= (1u << 24)) return fail(s->ctx, SLRH_ERR_CAPACITY, "instance_add: too many instances");
    Instance in;
    in.proto = (uint32_t)proto_id;
    in.material = material;
    float m[4][4], inv[4][4];   // [column][row]
    for (int k = 0; k < 12; ++k) {
        if (!std::isfinite(l2w[k])) return fail(s->ctx, SLRH_ERR_INVALID, "instance_add: non-finite transform");
        in.l2w[k] = l2w[k];
        m[k % 4][k / 4] = l2w[k];
    }
    m[0][3] = m[1][3] = m[2][3] = 0.0f;
    m[3][3] = 1.0f;
    if (!invertMatrix(m, inv)) return fail(s->ctx, SLRH_ERR_INVALID, "instance_add: singular transform");
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            if (!std::isfinite(inv[c][r])) return fail(s->ctx, SLRH_ERR_INVALID, "instance_add: singular transform");
            in.w2l[4 * r + c] = inv[c][r];
        }
    d.instances.push_back(in);
    if (inst_id) *inst_id = (int)d.instances.size() - 1;
    return SLRH_OK;
}

int slrh_instance_get(const slrh_scene* s, int inst_id, float* l2w, float* w2l) {
    CtxLock lk = lockOf(s);
    if (!s) return SLRH_ERR_INVALID;
    if (inst_id < 0 || (size_t)inst_id >= s->data.instances.size()) return SLRH_ERR_INVALID;
    const Instance& in = s->data.instances[inst_id];
    if (l2w) memcpy(l2w, in.l2w, sizeof(in.l2w));
    if (w2l) memcpy(w2l, in.w2l, sizeof(in.w2l));
    return SLRH_OK;
}

int slrh_medium_density_grid(slrh_scene* s, const float* bounds, const float* density, uint32_t nx, uint32_t ny,
                             uint32_t nz, int sigma_s, int sigma_e, float g, int* medium_id) {
    CtxLock lk = lockOf(s);
    CHECK_SCENE(s);
    SceneData& d = s->data;
    if (d.medium.present) return fail(s->ctx, SLRH_ERR_CAPACITY, "medium: one medium per scene");
    if (!bounds || !density || nx < 2 || ny < 2 || nz < 2)
        return fail(s->ctx, SLRH_ERR_INVALID, "medium: bounds, density and at least 2 lattice points per axis");
    if ((uint64_t)nx * ny * nz > (1ull << 30)) return fail(s->ctx, SLRH_ERR_CAPACITY, "medium: lattice too large");
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(bounds[a]) || !std::isfinite(bounds[3 + a]) || !(bounds[3 + a] > bounds[a]))
            return fail(s->ctx, SLRH_ERR_INVALID, "medium: bad bounds");
    const int ns = (int)d.spectra.size();
    if (sigma_s < 0 || sigma_s >= ns || sigma_e < 0 || sigma_e >= ns)
        return fail(s->ctx, SLRH_ERR_INVALID, "medium: bad spectrum id");
    const HostSpectrum& se = d.spectra[sigma_e];
    for (float v : se.values)
        if (v != se.values[0])
            return fail(s->ctx, SLRH_ERR_INVALID, "medium: the extinction spectrum must be constant over wavelength");
    if (!(se.values[0] >= 0) || !std::isfinite(se.values[0])) return fail(s->ctx, SLRH_ERR_INVALID, "medium: bad extinction");
    for (float v : d.spectra[sigma_s].values)
        if (!(v >= 0 && v <= se.values[0])) return fail(s->ctx, SLRH_ERR_INVALID, "medium: scattering must be in [0, extinction]");
    if (!(g > -1.0f && g < 1.0f)) return fail(s->ctx, SLRH_ERR_INVALID, "medium: g must be in (-1, 1)");
    const size_t cnt = (size_t)nx * ny * nz;
    for (size_t i = 0; i < cnt; ++i)
        if (!(density[i] >= 0) || !std::isfinite(density[i])) return fail(s->ctx, SLRH_ERR_INVALID, "medium: bad density");
    MediumData& m = d.medium;
    for (int a = 0; a < 3; ++a) { m.bmin[a] = bounds[a]; m.bmax[a] = bounds[3 + a]; }
    m.n[0] = nx; m.n[1] = ny; m.n[2] = nz;
    m.density.assign(density, density + cnt);
    m.sigmaS = sigma_s;
    m.sigmaE = sigma_e;
    m.sigmaEValue = se.values[0];
    m.g = g;
    buildMajorants(&m);
    m.present = true;
    if (medium_id) *medium_id = 0;
    return SLRH_OK;
}

int slrh_scene_compat(slrh_scene* s, uint32_t flags) {
    CtxLock lk = lockOf(s);
    CHECK_SCENE(s);
    if (flags & ~(uint32_t)(SLRH_COMPAT_INSTANCE_SHADOW | SLRH_COMPAT_VOLPT_TANGENT_NEE))
        return fail(s->ctx, SLRH_ERR_INVALID, "scene_compat: unknown flag");
    s->data.compatInstanceShadow = (flags & SLRH_COMPAT_INSTANCE_SHADOW) != 0;
    s->data.compatTangentNee = (flags & SLRH_COMPAT_VOLPT_TANGENT_NEE) != 0;
    return SLRH_OK;
}

int slrh_camera_perspective(slrh_scene* s, const float* l2w, float sensitivity, float aspect, float fovY,
                            float lensRadius, float imgDist, float objDist) {
    CtxLock lk = lockOf(s);
    CHECK_SCENE(s);
    if (!l2w || !(aspect > 0) || !(fovY > 0) || !(lensRadius >= 0) || !(imgDist > 0) || !(objDist > 0))
        return fail(s->ctx, SLRH_ERR_INVALID, "camera_perspective: bad args");
    CameraParams& c = s->data.camera;
    memcpy(c.l2w, l2w, sizeof(c.l2w));
    c.sensitivity = sensitivity; c.aspect = aspect; c.fovY = fovY;
    c.lensRadius = lensRadius; c.imgDist = imgDist; c.objDist = objDist;
    s->data.hasCamera = true;
    return SLRH_OK;
}

int slrh_scene_load_text(slrh_scene* s, const char* text, slrh_scene_file_info* info) {
    CtxLock lk = lockOf(s);
    CHECK_SCENE(s);
    if (!text) return fail(s->ctx, SLRH_ERR_INVALID, "scene_load_text: null text");
    scenefile::SceneFile sf;
    std::string err;
    if (!scenefile::readSceneText(text, &sf, &err)) return fail(s->ctx, SLRH_ERR_INVALID, "scene file: " + err);
    std::vector<int> specId(sf.spectra.size()), matId(sf.materials.size());
    int r;
    for (size_t k = 0; k < sf.spectra.size(); ++k) {
        const scenefile::Spectrum& sp = sf.spectra[k];
        if (sp.kind == scenefile::kRgb)
            r = slrh_spectrum_rgb(s, sp.type == scenefile::kLightSource ? SLRH_SPECTRUM_LIGHT_SOURCE
                                                                         : SLRH_SPECTRUM_REFLECTANCE,
                                  sp.rgb[0], sp.rgb[1], sp.rgb[2], &specId[k]);
        else
            r = slrh_spectrum_regular(s, sp.lo, sp.hi, sp.values.data(), (int)sp.values.size(), &specId[k]);
        if (r) return r;
    }
    for (size_t k = 0; k < sf.materials.size(); ++k) {
        const scenefile::Material& m = sf.materials[k];
        switch (m.kind) {
            case scenefile::kMatte:
                r = m.hasSigma ? slrh_material_oren_nayar(s, specId[m.spec0], m.sigma, &matId[k])
                               : slrh_material_lambert(s, specId[m.spec0], &matId[k]);
                break;
            case scenefile::kMicrofacetMetal:
                r = slrh_material_ggx_conductor(s, specId[m.spec0], specId[m.spec1], m.alphaX, m.alphaY, &matId[k]);
                break;
            case scenefile::kMicrofacetGlass:
                r = slrh_material_ggx_dielectric(s, specId[m.spec0], specId[m.spec1], m.alphaX, m.alphaY, &matId[k]);
                break;
            default:
                r = slrh_material_diffuse_emitter(s, matId[m.base], specId[m.emittance], &matId[k]);
                break;
        }
        if (r) return r;
    }
    uint32_t tris = 0;
    for (const scenefile::Mesh& mesh : sf.meshes) {
        std::vector<uint32_t> mats(mesh.mat.size());
        for (size_t t = 0; t < mats.size(); ++t) mats[t] = (uint32_t)matId[mesh.mat[t]];
        int id;
        r = slrh_mesh_add(s, (uint32_t)(mesh.pos.size() / 3), mesh.pos.data(), mesh.nrm.data(), mesh.tan.data(),
                          (uint32_t)mats.size(), mesh.idx.data(), mats.data(), -1, &id);
        if (r) return r;
        tris += (uint32_t)mats.size();
    }
    const scenefile::Camera& c = sf.camera;
    if ((r = slrh_camera_perspective(s, c.l2w, c.sensitivity, c.aspect, c.fovY, c.radius, c.imgDist, c.objDist)))
        return r;
    if (info) {
        info->width = sf.width;
        info->height = sf.height;
        info->spp = sf.spp;
        info->integrator = sf.method == "Volumetric PT" ? SLRH_VOL_PT : sf.method == "BPT" ? SLRH_BPT : SLRH_PT;
        info->rng_seed = sf.rngSeed;
        info->brightness = sf.brightness;
        info->sensitivity = c.sensitivity > 0 ? c.sensitivity : (float)(1.0f / (M_PI * c.radius * c.radius));
        info->num_meshes = (uint32_t)sf.meshes.size();
        info->num_triangles = tris;
        info->num_materials = (uint32_t)sf.materials.size();
        info->num_spectra = (uint32_t)sf.spectra.size();
    }
    return SLRH_OK;
}

int slrh_scene_geometry(const slrh_scene* s, uint32_t* nverts, uint32_t* ntris, float* pos, float* nrm, float* tan,
                        uint32_t* idx, uint32_t* mat) {
    CtxLock lk = lockOf(s);
    if (!s || !nverts || !ntris) return SLRH_ERR_INVALID;
    const SceneData& d = s->data;
    const uint32_t nv = (uint32_t)(d.vpos.size() / 3), nt = (uint32_t)d.triMat.size();
    const uint32_t cv = *nverts, ct = *ntris;
    *nverts = nv;
    *ntris = nt;
    if ((pos || nrm || tan) && cv < nv) return fail(s->ctx, SLRH_ERR_CAPACITY, "scene_geometry: vertex arrays too small");
    if ((idx || mat) && ct < nt) return fail(s->ctx, SLRH_ERR_CAPACITY, "scene_geometry: triangle arrays too small");
    if (pos) memcpy(pos, d.vpos.data(), d.vpos.size() * sizeof(float));
    if (nrm) memcpy(nrm, d.vnrm.data(), d.vnrm.size() * sizeof(float));
    if (tan) memcpy(tan, d.vtan.data(), d.vtan.size() * sizeof(float));
    if (idx) memcpy(idx, d.idx.data(), d.idx.size() * sizeof(uint32_t));
    if (mat) memcpy(mat, d.triMat.data(), d.triMat.size() * sizeof(uint32_t));
    return SLRH_OK;
}

int slrh_scene_camera_get(const slrh_scene* s, float* l2w, float* params) {
    CtxLock lk = lockOf(s);
    if (!s) return SLRH_ERR_INVALID;
    if (!s->data.hasCamera) return fail(s->ctx, SLRH_ERR_STATE, "scene has no camera");
    const CameraParams& c = s->data.camera;
    if (l2w) memcpy(l2w, c.l2w, sizeof(c.l2w));
    if (params) {
        const float p[6] = {c.sensitivity, c.aspect, c.fovY, c.lensRadius, c.imgDist, c.objDist};
        memcpy(params, p, sizeof(p));
    }
    return SLRH_OK;
}

int slrh_material_info(const slrh_scene* s, int id, int* model, int* spec0, int* spec1, int* emittance, float* p0,
                       float* p1) {
    CtxLock lk = lockOf(s);
    if (!s) return SLRH_ERR_INVALID;
    if (id < 0 || (size_t)id >= s->data.materials.size()) return fail(s->ctx, SLRH_ERR_INVALID, "material_info: bad id");
    const DevMaterial& m = s->data.materials[(size_t)id];
    if (model) *model = (int)m.type;
    if (spec0) *spec0 = m.spec0;
    if (spec1) *spec1 = m.spec1;
    if (emittance) *emittance = m.emit;
    if (p0) *p0 = m.p0;
    if (p1) *p1 = m.p1;
    return SLRH_OK;
}

int slrh_environment(slrh_scene* s, const float* rgb, uint32_t width, uint32_t height, float scale) {
    CtxLock lk = lockOf(s);
    CHECK_SCENE(s);
    if (!rgb || width == 0 || height == 0 || !(scale >= 0) || (uint64_t)width * height > (1ull << 28))
        return fail(s->ctx, SLRH_ERR_INVALID, "environment: bad args");
    buildEnvironment(rgb, width, height, scale, &s->data.env);
    return SLRH_OK;
}

int slrh_scene_build(slrh_scene* s) {
    CHECK_SCENE(s);
    CtxLock lk = lockOf(s);
    try {
        SceneData& d = s->data;
        const size_t nt = d.idx.size() / 3;
        auto positions = [&](const std::vector<uint32_t>& ids) {
            std::vector<float> tp(9 * ids.size());
            for (size_t i = 0; i < ids.size(); ++i)
                for (int k = 0; k < 3; ++k)
                    for (int j = 0; j < 3; ++j) tp[9 * i + 3 * k + j] = d.vpos[3ull * d.idx[3 * (size_t)ids[i] + k] + j];
            return tp;
        };
        int threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        // with a device, the top of each large triangle tree is binned there
        const bool onDevice = s->ctx->device >= 0;
        if (onDevice) {
            std::string err;
            if (int r = deviceSelect(s->ctx->device, &err)) return fail(s->ctx, r, "scene build: " + err);
        }
        // one accelerator per prototype, over its object-space triangles
        for (Prototype& p : d.protos) {
            std::vector<uint32_t> ids(p.triCount);
            for (uint32_t i = 0; i < p.triCount; ++i) ids[i] = p.triFirst + i;
            buildAccel(positions(ids), &p.accel, threads, onDevice);
            for (uint32_t& r : p.accel.sbvhRefs) r = ids[r];
            for (uint32_t& r : p.accel.qbvhRefs) r = ids[r];
        }
        // the world level over world-space triangles and the instances'
        // world bounds (TransformedSurfaceObject::bounds, surface_object.cpp:290-292:
        // the eight transformed corners of the prototype's bounds)
        std::vector<uint32_t> world;
        for (size_t t = 0; t < nt; ++t)
            if (d.triProto[t] < 0) world.push_back((uint32_t)t);
        std::vector<Box> boxes;
        std::vector<float> costs;
        for (const Instance& in : d.instances) {
            // a one-triangle prototype is referenced as the triangle itself
            // (ReferenceNode::createRenderingData), otherwise as its aggregate
            const Prototype& pr = d.protos[in.proto];
            costs.push_back(pr.triCount > 1 ? pr.accel.sbvhCost : 1.0f);
            const Box& b = d.protos[in.proto].accel.bounds;
            Box w;
            for (int c = 0; c < 8; ++c)
                w.unify(affinePoint(in.l2w, v3((c & 4) ? b.mx[0] : b.mn[0], (c & 2) ? b.mx[1] : b.mn[1],
                                               (c & 1) ? b.mx[2] : b.mn[2])));
            boxes.push_back(w);
        }
        buildAccel(positions(world), boxes, costs, &d.accel, threads, onDevice);
        const uint32_t nw = (uint32_t)world.size();
        auto code = [&](uint32_t r) { return r < nw ? world[r] : (kRefInstance | (r - nw)); };
        for (uint32_t& r : d.accel.sbvhRefs) r = code(r);
        for (uint32_t& r : d.accel.qbvhRefs) r = code(r);
        if (d.wantStdBvh) {
            // the binned-SAH binary BVH over the world triangles
            // (StandardBVH(objs, Partitioning::BinnedSAH); single-level scenes)
            buildStandardBvh(positions(world), &d.stdBvh);
            for (uint32_t& r : d.stdBvh.refs) r = world[r];
        }
        if (s->ctx->device >= 0) {
            std::string err;
            int r = uploadScene(s->data, &s->dev, &err);
            if (r) return fail(s->ctx, r, "scene upload: " + err);
        }
    } catch (const std::bad_alloc&) {
        return fail(s->ctx, SLRH_ERR_NOMEM, "scene build: out of host memory");
    } catch (const std::exception& e) {
        return fail(s->ctx, SLRH_ERR_INVALID, std::string("scene build: ") + e.what());
    }
    s->built = true;
    return SLRH_OK;
}

int slrh_scene_accelerators(slrh_scene* s, uint32_t flags) {
    CtxLock lk = lockOf(s);
    CHECK_SCENE(s);
    if (flags & ~(uint32_t)SLRH_ACCEL_STANDARD_BVH) return fail(s->ctx, SLRH_ERR_INVALID, "unknown accelerator flag");
    if ((flags & SLRH_ACCEL_STANDARD_BVH) && !s->data.instances.empty())
        return fail(s->ctx, SLRH_ERR_STATE, "the binned-SAH BVH is built for scenes without instances");
    s->data.wantStdBvh = (flags & SLRH_ACCEL_STANDARD_BVH) != 0;
    return SLRH_OK;
}

int slrh_scene_query_accel(slrh_scene* s, int accel) {
    CtxLock lk = lockOf(s);
    if (!s) return SLRH_ERR_INVALID;
    if (accel != SLRH_QUERY_QBVH && accel != SLRH_QUERY_STANDARD_BVH)
        return fail(s->ctx, SLRH_ERR_INVALID, "unknown query accelerator");
    if (accel == SLRH_QUERY_STANDARD_BVH && !s->data.wantStdBvh)
        return fail(s->ctx, SLRH_ERR_STATE, "the binned-SAH BVH was not requested (slrh_scene_accelerators)");
    s->queryAccel = accel;
    return SLRH_OK;
}

int slrh_scene_standard_bvh_dump(const slrh_scene* s, float* boxes, uint32_t* info, int32_t* refs, uint32_t* nn,
                                 uint32_t* nr, double* stats) {
    CtxLock lk = lockOf(s);
    if (!s) return SLRH_ERR_INVALID;
    if (!s->built || !s->data.wantStdBvh) return SLRH_ERR_STATE;
    const StdBvh& b = s->data.stdBvh;
    if (nn) *nn = (uint32_t)b.nodes.size();
    if (nr) *nr = (uint32_t)b.refs.size();
    if (stats) {
        stats[0] = b.depth;
        stats[1] = b.cost;
    }
    for (size_t i = 0; boxes && info && i < b.nodes.size(); ++i) {
        const StdBvhNode& n = b.nodes[i];
        for (int k = 0; k < 3; ++k) { boxes[6 * i + k] = n.box.mn[k]; boxes[6 * i + 3 + k] = n.box.mx[k]; }
        info[5 * i + 0] = n.c0; info[5 * i + 1] = n.c1; info[5 * i + 2] = n.first; info[5 * i + 3] = n.count;
        info[5 * i + 4] = n.count ? 0xFFu : n.axis;
    }
    for (size_t i = 0; refs && i < b.refs.size(); ++i) refs[i] = (int32_t)b.refs[i];
    return SLRH_OK;
}

int slrh_scene_accel_stats(const slrh_scene* s, double* st) {
    CtxLock lk = lockOf(s);
    if (!s || !st) return SLRH_ERR_INVALID;
    if (!s->built) return SLRH_ERR_STATE;
    const Accel& a = s->data.accel;
    st[0] = (double)a.sbvh.size();
    st[1] = (double)a.sbvhRefs.size();
    st[2] = a.sbvhDepth;
    st[3] = a.sbvhCost;
    st[4] = (double)a.qbvh.size();
    st[5] = (double)a.qbvhRefs.size();
    st[6] = a.qbvhDepth;
    st[7] = a.qbvhCost;
    st[8] = a.buildSeconds;
    return SLRH_OK;
}

int slrh_scene_build_stats(const slrh_scene* s, double* st) {
    CtxLock lk = lockOf(s);
    if (!s || !st) return SLRH_ERR_INVALID;
    if (!s->built) return SLRH_ERR_STATE;
    const SceneData& d = s->data;
    double secs = d.accel.buildSeconds, dev = d.accel.deviceNodes;
    for (const Prototype& p : d.protos) {
        secs += p.accel.buildSeconds;
        dev += p.accel.deviceNodes;
    }
    st[0] = secs;
    st[1] = dev;
    return SLRH_OK;
}

int slrh_scene_sbvh_dump(const slrh_scene* s, float* boxes, uint32_t* info, int32_t* refs, uint32_t* nn,
                         uint32_t* nr) {
    CtxLock lk = lockOf(s);
    if (!s) return SLRH_ERR_INVALID;
    if (!s->built) return SLRH_ERR_STATE;
    const Accel& a = s->data.accel;
    if (nn) *nn = (uint32_t)a.sbvh.size();
    if (nr) *nr = (uint32_t)a.sbvhRefs.size();
    for (size_t i = 0; boxes && info && i < a.sbvh.size(); ++i) {
        const SbvhNode& n = a.sbvh[i];
        for (int k = 0; k < 3; ++k) { boxes[6 * i + k] = n.box.mn[k]; boxes[6 * i + 3 + k] = n.box.mx[k]; }
        info[5 * i + 0] = n.c0; info[5 * i + 1] = n.c1; info[5 * i + 2] = n.first; info[5 * i + 3] = n.count;
        info[5 * i + 4] = n.count ? 0xFFu : n.axis;
    }
    for (size_t i = 0; refs && i < a.sbvhRefs.size(); ++i) refs[i] = (int32_t)a.sbvhRefs[i];
    return SLRH_OK;
}

int slrh_scene_qbvh_dump(const slrh_scene* s, float* boxes, uint32_t* info, int32_t* refs, uint32_t* nn,
                         uint32_t* nr) {
    CtxLock lk = lockOf(s);
    if (!s) return SLRH_ERR_INVALID;
    if (!s->built) return SLRH_ERR_STATE;
    const Accel& a = s->data.accel;
    if (nn) *nn = (uint32_t)a.qbvh.size();
    if (nr) *nr = (uint32_t)a.qbvhRefs.size();
    for (size_t i = 0; boxes && info && i < a.qbvh.size(); ++i) {
        const QbvhNode& n = a.qbvh[i];
        for (int ax = 0; ax < 3; ++ax)
            for (int c = 0; c < 4; ++c) { boxes[24 * i + 4 * ax + c] = n.bmin[ax][c]; boxes[24 * i + 12 + 4 * ax + c] = n.bmax[ax][c]; }
        for (int c = 0; c < 4; ++c) {
            uint32_t w = 0xFFFFFFFFu;
            if (n.childValid[c]) {
                // the reference's Children bitfield {idx:27, numLeaves:4, isLeaf:1} (QBVH.h:26-39)
                w = (n.index[c] & 0x07FFFFFFu) | ((n.count[c] & 0xFu) << 27) | ((n.isLeaf[c] ? 1u : 0u) << 31);
            }
            info[7 * i + c] = w;
        }
        info[7 * i + 4] = n.topAxis; info[7 * i + 5] = n.leftAxis; info[7 * i + 6] = n.rightAxis;
    }
    for (size_t i = 0; refs && i < a.qbvhRefs.size(); ++i) refs[i] = (int32_t)a.qbvhRefs[i];
    return SLRH_OK;
}

static int doTrace(slrh_scene* s, const slrh_rays* rays, const slrh_hits* hits, float* vis, uint32_t n, int mem,
                   void* stream) {
    if (!s || !rays) return SLRH_ERR_INVALID;
    if (!s->built) return fail(s->ctx, SLRH_ERR_STATE, "trace before slrh_scene_build");
    if (!s->dev) return fail(s->ctx, SLRH_ERR_STATE, "host-only context has no device scene");
    if (mem != SLRH_MEM_HOST && mem != SLRH_MEM_DEVICE) return fail(s->ctx, SLRH_ERR_INVALID, "bad mem kind");
    if (n == 0) return SLRH_OK;
    TraceArgs a{};
    a.ox = rays->ox; a.oy = rays->oy; a.oz = rays->oz; a.dx = rays->dx; a.dy = rays->dy; a.dz = rays->dz;
    a.tmin = rays->tmin; a.tmax = rays->tmax;
    if (!a.ox || !a.oy || !a.oz || !a.dx || !a.dy || !a.dz || !a.tmin || !a.tmax)
        return fail(s->ctx, SLRH_ERR_INVALID, "null ray array");
    if (hits) {
        a.t = hits->t; a.prim = hits->prim; a.b0 = hits->b0; a.b1 = hits->b1; a.inst = hits->inst;
        if (!a.t || !a.prim || !a.b0 || !a.b1) return fail(s->ctx, SLRH_ERR_INVALID, "null hit array");
    } else {
        a.vis = vis;
        if (!a.vis) return fail(s->ctx, SLRH_ERR_INVALID, "null visibility array");
    }
    CtxLock lk = lockOf(s);
    std::string err;
    s->lastNodes = s->lastTris = 0;
    int r = s->queryAccel == SLRH_QUERY_STANDARD_BVH
                ? traceBatchStdBvh(s->dev, a, n, hits == nullptr, mem == SLRH_MEM_DEVICE, stream, &err)
                : traceBatch(s->dev, a, n, hits == nullptr, mem == SLRH_MEM_DEVICE, stream, &s->lastNodes,
                             &s->lastTris, &err);
    if (r) return fail(s->ctx, r, err);
    return SLRH_OK;
}

int slrh_scene_query_counts(const slrh_scene* s, double* nodes, double* tris) {
    if (!s) return SLRH_ERR_INVALID;
    CtxLock lk = lockOf(s);
    if (nodes) *nodes = s->lastNodes;
    if (tris) *tris = s->lastTris;
    return SLRH_OK;
}

int slrh_intersect(slrh_scene* s, const slrh_rays* rays, const slrh_hits* hits, uint32_t n, int mem, void* stream) {
    CtxLock lk = lockOf(s);
    if (!hits) return s ? fail(s->ctx, SLRH_ERR_INVALID, "null hits") : SLRH_ERR_INVALID;
    return doTrace(s, rays, hits, nullptr, n, mem, stream);
}

int slrh_occluded(slrh_scene* s, const slrh_rays* rays, float* vis, uint32_t n, int mem, void* stream) {
    CtxLock lk = lockOf(s);
    return doTrace(s, rays, nullptr, vis, n, mem, stream);
}

// Deepest traversal stack the scene's trees can need: a QBVH node step
// leaves at most three children pending per level (the nearest one is held
// in a register), and entering an instance suspends at most six words (the
// held node, three pending leaves, the rest of the leaf, the marker) before
// the prototype's tree adds its own levels.
static uint32_t stackBound(const SceneData& d) {
    uint32_t b = 3 * d.accel.qbvhDepth;
    if (!d.instances.empty()) {
        uint32_t pd = 0;
        for (const Prototype& p : d.protos) pd = std::max(pd, p.accel.qbvhDepth);
        b += 6 + 3 * pd;
    }
    return b;
}

int slrh_intersect_noalloc(slrh_scene* s, const slrh_rays* rays, const slrh_hits* hits, uint32_t n, void* stream) {
    if (!s || !rays || !hits) return s ? fail(s->ctx, SLRH_ERR_INVALID, "null rays or hits") : SLRH_ERR_INVALID;
    if (!s->built) return fail(s->ctx, SLRH_ERR_STATE, "trace before slrh_scene_build");
    if (!s->dev) return fail(s->ctx, SLRH_ERR_STATE, "host-only context has no device scene");
    if (n > kNoAllocMaxRays)
        return fail(s->ctx, SLRH_ERR_CAPACITY, "slrh_intersect_noalloc: more than 2^30 rays in one call");
    if (n == 0) return SLRH_OK;
    TraceArgs a{};
    a.ox = rays->ox; a.oy = rays->oy; a.oz = rays->oz; a.dx = rays->dx; a.dy = rays->dy; a.dz = rays->dz;
    a.tmin = rays->tmin; a.tmax = rays->tmax;
    a.t = hits->t; a.prim = hits->prim; a.b0 = hits->b0; a.b1 = hits->b1; a.inst = hits->inst;
    if (!a.ox || !a.oy || !a.oz || !a.dx || !a.dy || !a.dz || !a.tmin || !a.tmax)
        return fail(s->ctx, SLRH_ERR_INVALID, "null ray array");
    if (!a.t || !a.prim || !a.b0 || !a.b1) return fail(s->ctx, SLRH_ERR_INVALID, "null hit array");
    if (stackBound(s->data) > 64)
        return fail(s->ctx, SLRH_ERR_CAPACITY, "slrh_intersect_noalloc: the scene's trees may need a traversal "
                                                "stack deeper than 64 entries; use slrh_intersect");
    CtxLock lk = lockOf(s);
    std::string err;
    int r = deviceSelect(s->ctx->device, &err);
    if (!r) r = traceBatchNoAlloc(s->dev, a, n, false, stream, &err);
    if (r) return fail(s->ctx, r, err);
    return SLRH_OK;
}

int slrh_medium_query(slrh_scene* s, const slrh_rays* rays, uint32_t n, int mode, uint32_t seed, float* out,
                      int mem, void* stream) {
    CtxLock lk = lockOf(s);
    if (!s || !rays || !out) return SLRH_ERR_INVALID;
    if (!s->built) return fail(s->ctx, SLRH_ERR_STATE, "medium query before slrh_scene_build");
    if (!s->dev) return fail(s->ctx, SLRH_ERR_STATE, "host-only context has no device scene");
    if (!s->data.medium.present) return fail(s->ctx, SLRH_ERR_STATE, "scene has no medium");
    if (mode != SLRH_MEDIUM_TRANSMITTANCE && mode != SLRH_MEDIUM_FREE_PATH)
        return fail(s->ctx, SLRH_ERR_INVALID, "medium query: unknown mode");
    if (mem != SLRH_MEM_HOST && mem != SLRH_MEM_DEVICE) return fail(s->ctx, SLRH_ERR_INVALID, "bad mem kind");
    if (n == 0) return SLRH_OK;
    TraceArgs a{};
    a.ox = rays->ox; a.oy = rays->oy; a.oz = rays->oz; a.dx = rays->dx; a.dy = rays->dy; a.dz = rays->dz;
    a.tmin = rays->tmin; a.tmax = rays->tmax;
    if (!a.ox || !a.oy || !a.oz || !a.dx || !a.dy || !a.dz || !a.tmin || !a.tmax)
        return fail(s->ctx, SLRH_ERR_INVALID, "null ray array");
    std::string err;
    int r = mediumQuery(s->dev, a, n, mode, seed, out, mem == SLRH_MEM_DEVICE, stream, &err);
    if (r) return fail(s->ctx, r, err);
    return SLRH_OK;
}

int slrh_film_create(slrh_context* ctx, uint32_t w, uint32_t h, slrh_film** out) {
    if (!ctx || !out || w == 0 || h == 0) return SLRH_ERR_INVALID;
    if (ctx->device < 0) return fail(ctx, SLRH_ERR_STATE, "host-only context has no device");
    if ((uint64_t)w * h > (1ull << 27)) return fail(ctx, SLRH_ERR_CAPACITY, "film too large");
    std::unique_ptr<slrh_film> f(new (std::nothrow) slrh_film());
    if (!f) return SLRH_ERR_NOMEM;
    f->ctx = ctx; f->w = w; f->h = h;
    std::string err;
    int r = filmCreate(w, h, &f->dev, &err);
    if (r) return fail(ctx, r, err);
    *out = f.release();
    return SLRH_OK;
}

int slrh_film_destroy(slrh_film* f) {
    CtxLock lk = lockOf(f);
    if (!f) return SLRH_ERR_INVALID;
    filmFree(f->dev);
    delete f;
    return SLRH_OK;
}

int slrh_film_clear(slrh_film* f) {
    CtxLock lk = lockOf(f);
    if (!f) return SLRH_ERR_INVALID;
    std::string err;
    int r = filmClear(f->dev, &err);
    return r ? fail(f->ctx, r, err) : SLRH_OK;
}

namespace {
// One pixel through ImageSensor::saveImage's arithmetic
// (libSLR/Core/ImageSensor.cpp:139-195) with the reference's types: the
// stratum sums as FP32 (DiscretizedSpectrum), times scale in FP32, XYZ by the
// CMF strata accumulated in FP32 in stratum order and divided by their
// integral (spectrum_types.h:666-690), XYZ_to_sRGB with its double constants
// (spectrum_base.h:68-72), clamp at zero, the 1 - exp(-x) tone curve in FP32,
// sRGB_gamma (spectrum_base.cpp:16-21, double constants, FP32 result) and
// the byte min(255, uint32(256 * gamma)). out: R, G, B.
void saveImagePixel(const double* st, float scale, const float* xb, const float* yb, const float* zb, float integ,
                    uint8_t out[3]) {
    float v[16];
    for (int k = 0; k < 16; ++k) v[k] = (float)st[k] * scale;
    float XYZ[3] = {0.0f, 0.0f, 0.0f};
    for (int k = 0; k < 16; ++k) {
        XYZ[0] += xb[k] * v[k];
        XYZ[1] += yb[k] * v[k];
        XYZ[2] += zb[k] * v[k];
    }
    XYZ[0] /= integ;
    XYZ[1] /= integ;
    XYZ[2] /= integ;
    float RGB[3];
    RGB[0] = (float)(3.2404542 * XYZ[0] - 1.5371385 * XYZ[1] - 0.4985314 * XYZ[2]);
    RGB[1] = (float)(-0.9692660 * XYZ[0] + 1.8760108 * XYZ[1] + 0.0415560 * XYZ[2]);
    RGB[2] = (float)(0.0556434 * XYZ[0] - 0.2040259 * XYZ[1] + 1.0572252 * XYZ[2]);
    for (int c = 0; c < 3; ++c) {
        float x = RGB[c] < 0.0f ? 0.0f : RGB[c];
        x = 1.0f - std::exp(-x);
        const float g = x <= 0.0031308 ? (float)(12.92 * x) : (float)(1.055 * std::pow((double)x, 1.0 / 2.4) - 0.055);
        const float q = 256 * g;
        out[c] = (uint8_t)std::min<uint32_t>(255u, (uint32_t)q);
    }
}
}  // namespace

int slrh_film_read(const slrh_film* f, int format, double scale, void* dst, size_t bytes) {
    CtxLock lk = lockOf(f);
    if (!f || !dst) return SLRH_ERR_INVALID;
    const size_t px = (size_t)f->w * f->h;
    size_t need = format == SLRH_FILM_STRATA16_F64 ? px * 16 * 8 : format == SLRH_FILM_XYZ_F64 ? px * 3 * 8
                : format == SLRH_FILM_SRGB8 ? px * 3 : 0;
    if (need == 0) return fail(f->ctx, SLRH_ERR_INVALID, "film_read: bad format");
    if (bytes < need) return fail(f->ctx, SLRH_ERR_INVALID, "film_read: destination too small");
    std::vector<double> st;
    std::string err;
    int r = filmDownload(f->dev, &st, &err);
    if (r) return fail(f->ctx, r, err);
    if (format == SLRH_FILM_STRATA16_F64) {
        double* o = (double*)dst;
        for (size_t i = 0; i < st.size(); ++i) o[i] = st[i] * scale;
        return SLRH_OK;
    }
    float xb[16], yb[16], zb[16], integ;
    cmfStrata(xb, yb, zb, &integ);
    for (size_t p = 0; p < px; ++p) {
        if (format == SLRH_FILM_XYZ_F64) {
            // DiscretizedSpectrum::getRGB's XYZ step (libSLR/BasicTypes/spectrum_types.h:666-690), in FP64
            double X = 0, Y = 0, Z = 0;
            for (int k = 0; k < 16; ++k) {
                double v = st[16 * p + k] * scale;
                X += xb[k] * v; Y += yb[k] * v; Z += zb[k] * v;
            }
            double* o = (double*)dst + 3 * p;
            o[0] = X / integ; o[1] = Y / integ; o[2] = Z / integ;
        } else {
            saveImagePixel(&st[16 * p], (float)scale, xb, yb, zb, integ, (uint8_t*)dst + 3 * p);
        }
    }
    return SLRH_OK;
}

int slrh_film_load(slrh_film* f, const double* strata, size_t bytes) {
    CtxLock lk = lockOf(f);
    if (!f || !strata) return SLRH_ERR_INVALID;
    if (bytes < (size_t)f->w * f->h * 16 * sizeof(double))
        return fail(f->ctx, SLRH_ERR_INVALID, "film_load: source smaller than the film");
    std::string err;
    int r = filmUpload(f->dev, strata, &err);
    if (r) return fail(f->ctx, r, err);
    return SLRH_OK;
}

int slrh_film_bmp(const slrh_film* f, float scale, float sensitivity, void* dst, size_t dst_bytes, size_t* size) {
    CtxLock lk = lockOf(f);
    if (!f || !size) return SLRH_ERR_INVALID;
    // saveBMP's layout (libSLR/Helper/bmp_exporter.cpp:15-54): 54-byte header,
    // rows bottom-up, BGR, each row 3 * width + width % 4 bytes
    const uint32_t w = f->w, h = f->h;
    const uint32_t byteWidth = 3 * w + w % 4;
    const size_t total = 54 + (size_t)byteWidth * h;
    *size = total;
    if (!dst) return SLRH_OK;
    if (dst_bytes < total) return fail(f->ctx, SLRH_ERR_INVALID, "film_bmp: destination too small");
    std::vector<double> st;
    std::string err;
    int r = filmDownload(f->dev, &st, &err);
    if (r) return fail(f->ctx, r, err);
    uint8_t* o = (uint8_t*)dst;
    memset(o, 0, total);
    const uint32_t fileSize = (uint32_t)total, dataOffset = 54, infoSize = 40, compress = 0, ppm = 1;
    const uint32_t dataSize = byteWidth * h;
    const uint16_t planes = 1, bits = 24;
    o[0] = 'B';
    o[1] = 'M';
    memcpy(o + 2, &fileSize, 4);
    memcpy(o + 10, &dataOffset, 4);
    memcpy(o + 14, &infoSize, 4);
    memcpy(o + 18, &w, 4);
    memcpy(o + 22, &h, 4);
    memcpy(o + 26, &planes, 2);
    memcpy(o + 28, &bits, 2);
    memcpy(o + 30, &compress, 4);
    memcpy(o + 34, &dataSize, 4);
    memcpy(o + 38, &ppm, 4);
    memcpy(o + 42, &ppm, 4);
    // ImageSensor::saveImage: an infinite sensitivity counts as 1
    const float sens = std::isinf(sensitivity) ? 1.0f : sensitivity;
    const float sc = scale * sens;
    float xb[16], yb[16], zb[16], integ;
    cmfStrata(xb, yb, zb, &integ);
    for (uint32_t i = 0; i < h; ++i)
        for (uint32_t j = 0; j < w; ++j) {
            uint8_t rgb[3];
            saveImagePixel(&st[16 * ((size_t)i * w + j)], sc, xb, yb, zb, integ, rgb);
            uint8_t* px = o + 54 + (size_t)(h - i - 1) * byteWidth + 3 * j;
            px[0] = rgb[2];
            px[1] = rgb[1];
            px[2] = rgb[0];
        }
    return SLRH_OK;
}

int slrh_trace_variants(int* n, int* ldsDepth, int* refill, int* fused) {
    if (!n) return SLRH_ERR_INVALID;
    const int m = numTraceVariants();
    for (int v = 0; v < std::min(*n, m); ++v) {
        int d, r, f;
        traceVariant(v, &d, &r, &f);
        if (ldsDepth) ldsDepth[v] = d;
        if (refill) refill[v] = r;
        if (fused) fused[v] = f;
    }
    *n = m;
    return SLRH_OK;
}

namespace {
constexpr int kNumRenderStats = 27;   // the doubles slrh_render documents
void copyStats(const RenderStats& st, double* stats) {
    if (!stats) return;
    stats[0] = st.raysTraced; stats[1] = st.samples; stats[2] = st.kernelSeconds; stats[3] = st.nonFinite;
    stats[4] = st.iterations; stats[5] = st.traceSeconds; stats[6] = st.traceLaunches; stats[7] = st.extendRays;
    stats[8] = st.shadeSeconds; stats[9] = st.otherSeconds; stats[10] = st.shadowRays;
    stats[11] = st.nodesVisited; stats[12] = st.trisTested;
    stats[13] = st.nodesShadow; stats[14] = st.trisShadow;
    stats[15] = st.traceVariant;
    stats[16] = st.mediumEvents; stats[17] = st.mediumShadows;
    stats[18] = st.mediumSeconds; stats[19] = st.mediumLookups; stats[20] = st.mediumCells;
    stats[21] = st.mediumRays;
    stats[22] = st.mediumGuard;
    stats[23] = st.nodesLds;
    stats[24] = st.instEntered;
    stats[25] = st.shadeBlock;
    stats[26] = st.slotsPerPixel;
    if (st.bptLightVertices + st.bptEyeVertices > 0) {
        // the bidirectional integrator has no medium: its counters take those entries
        stats[16] = st.bptLightVertices;
        stats[17] = st.bptEyeVertices;
        stats[19] = st.bptConnections;
        stats[20] = st.bptSplats;
        stats[22] = st.bptCapped;
    }
}
bool knownIntegrator(uint32_t i) { return i == SLRH_PT || i == SLRH_VOL_PT || i == SLRH_BPT; }
}  // namespace

int slrh_render(slrh_scene* s, const slrh_render_settings* set, slrh_film* film, double* stats) {
    if (!s || !set || !film) return SLRH_ERR_INVALID;
    if (!s->built) return fail(s->ctx, SLRH_ERR_STATE, "render before slrh_scene_build");
    if (!s->dev) return fail(s->ctx, SLRH_ERR_STATE, "host-only context has no device scene");
    if (!s->data.hasCamera) return fail(s->ctx, SLRH_ERR_STATE, "scene has no camera");
    if (!knownIntegrator(set->integrator)) return fail(s->ctx, SLRH_ERR_INVALID, "unknown integrator");
    if (set->width != film->w || set->height != film->h) return fail(s->ctx, SLRH_ERR_INVALID, "film size mismatch");
    if (s->data.materials.empty() || s->data.idx.empty()) return fail(s->ctx, SLRH_ERR_STATE, "empty scene");
    CtxLock lk = lockOf(s);
    RenderConfig cfg{set->width, set->height, set->spp, set->max_depth, set->seed, set->slots_per_pixel,
                     (set->flags & 1u) != 0};
    if (set->reserved != 0) return fail(s->ctx, SLRH_ERR_INVALID, "reserved settings field must be 0");
    cfg.traceVariant = set->tune_trace == 0 ? -1 : (int)set->tune_trace - 1;
    if (cfg.traceVariant >= numTraceVariants()) return fail(s->ctx, SLRH_ERR_INVALID, "unknown traversal variant");
    cfg.useGraph = (set->flags & SLRH_FLAG_NO_GRAPH) == 0;
    cfg.volumetric = set->integrator == SLRH_VOL_PT;
    cfg.parallelShade = (set->flags & SLRH_FLAG_PARALLEL_SHADE) != 0;
    cfg.shadeShape = (set->flags & SLRH_FLAG_SHADE_BLOCK_128) ? 1 : -1;
    cfg.tunePool = (set->flags & SLRH_FLAG_TUNE_POOL) != 0;
    RenderStats st;
    std::string err;
    int r = set->integrator == SLRH_BPT ? renderBpt(s->dev, cfg, s->data.camera, film->dev, &st, &err)
                                        : slrh::render(s->dev, cfg, film->dev, &st, &err);
    if (r) return fail(s->ctx, r, err);
    copyStats(st, stats);
    return SLRH_OK;
}

int slrh_render_progressive(slrh_scene* s, const slrh_render_settings* set, slrh_film* film,
                            slrh_snapshot_fn snapshot, void* user, double* stats) {
    CtxLock lk = lockOf(s);
    if (!s || !set || !film) return SLRH_ERR_INVALID;
    if (set->spp == 0) return SLRH_OK;
    // passes of 1, 1, 2, 4, ... samples per pixel (the last cut to spp), so
    // the film holds 2^k samples per pixel after pass k, where
    // PTRenderer::render saves its snapshots (PTRenderer.cpp:77-94)
    double sum[kNumRenderStats] = {0};
    uint32_t done = 0, pass = 0;
    while (done < set->spp) {
        slrh_render_settings ps = *set;
        ps.spp = std::min(set->spp - done, std::max(1u, done));
        ps.seed = set->seed + 0x9E3779B9u * pass;
        double st[kNumRenderStats] = {0};
        int r = slrh_render(s, &ps, film, st);
        if (r) return r;
        for (int k = 0; k < kNumRenderStats; ++k) sum[k] += st[k];
        done += ps.spp;
        ++pass;
        if (snapshot && snapshot(user, film, done) != 0) break;
    }
    sum[15] /= pass;   // the traversal variant, not a sum
    sum[25] /= pass;   // the shading workgroup size
    sum[26] /= pass;   // slots per pixel
    if (stats) memcpy(stats, sum, sizeof(sum));
    return SLRH_OK;
}

namespace {
int frameSpec(slrh_context* ctx, const slrh_frame& f, uint32_t maxDepth, JobSpec* js) {
    if (!f.film) return fail(ctx, SLRH_ERR_INVALID, "frame without a film");
    if (f.width != f.film->w || f.height != f.film->h) return fail(ctx, SLRH_ERR_INVALID, "film size mismatch");
    memcpy(js->camera.l2w, f.l2w, sizeof(js->camera.l2w));
    js->camera.sensitivity = f.sensitivity;
    js->camera.aspect = f.aspect;
    js->camera.fovY = f.fov_y;
    js->camera.lensRadius = f.lens_radius;
    js->camera.imgDist = f.img_dist;
    js->camera.objDist = f.obj_dist;
    js->width = f.width;
    js->height = f.height;
    js->spp = f.spp;
    js->seed = f.seed;
    js->maxDepth = maxDepth;
    js->film = f.film->dev;
    return SLRH_OK;
}
}  // namespace

int slrh_render_batch(slrh_scene* s, const slrh_frame* frames, uint32_t n, const slrh_render_settings* set,
                      double* stats) {
    if (!s || !set || (n && !frames)) return SLRH_ERR_INVALID;
    if (!s->built) return fail(s->ctx, SLRH_ERR_STATE, "render before slrh_scene_build");
    if (!s->dev) return fail(s->ctx, SLRH_ERR_STATE, "host-only context has no device scene");
    if (!knownIntegrator(set->integrator)) return fail(s->ctx, SLRH_ERR_INVALID, "unknown integrator");
    if (set->reserved != 0) return fail(s->ctx, SLRH_ERR_INVALID, "reserved settings field must be 0");
    if (s->data.materials.empty() || s->data.idx.empty()) return fail(s->ctx, SLRH_ERR_STATE, "empty scene");
    std::vector<JobSpec> specs(n);
    for (uint32_t i = 0; i < n; ++i) {
        int r = frameSpec(s->ctx, frames[i], set->max_depth, &specs[i]);
        if (r) return r;
        // a slot adds into its film cell with a plain read-modify-write; two
        // frames of one wavefront sharing a film would race on its cells
        for (uint32_t j = 0; j < i; ++j)
            if (frames[j].film == frames[i].film)
                return fail(s->ctx, SLRH_ERR_INVALID, "two frames of a batch share a film");
    }
    RenderOptions opt;
    opt.slotsPerPixel = set->slots_per_pixel;
    opt.timeKernels = (set->flags & SLRH_FLAG_TIME_KERNELS) != 0;
    opt.useGraph = (set->flags & SLRH_FLAG_NO_GRAPH) == 0;
    opt.traceVariant = set->tune_trace == 0 ? -1 : (int)set->tune_trace - 1;
    opt.volumetric = set->integrator == SLRH_VOL_PT;
    opt.parallelShade = (set->flags & SLRH_FLAG_PARALLEL_SHADE) != 0;
    opt.shadeShape = (set->flags & SLRH_FLAG_SHADE_BLOCK_128) ? 1 : -1;
    opt.tunePool = (set->flags & SLRH_FLAG_TUNE_POOL) != 0;
    if (opt.traceVariant >= numTraceVariants()) return fail(s->ctx, SLRH_ERR_INVALID, "unknown traversal variant");
    CtxLock lk = lockOf(s);
    RenderStats st;
    std::string err;
    if (set->integrator == SLRH_BPT) {
        // the bidirectional integrator renders the frames one after another
        for (uint32_t i = 0; i < n; ++i) {
            RenderConfig cfg{specs[i].width, specs[i].height, specs[i].spp, set->max_depth, specs[i].seed,
                             set->slots_per_pixel, opt.timeKernels};
            int r = renderBpt(s->dev, cfg, specs[i].camera, specs[i].film, &st, &err);
            if (r) return fail(s->ctx, r, err);
        }
        copyStats(st, stats);
        return SLRH_OK;
    }
    int r = renderJobs(s->dev, specs.data(), n, opt, &st, &err);
    if (r) return fail(s->ctx, r, err);
    copyStats(st, stats);
    return SLRH_OK;
}

int slrh_server_create(slrh_scene* s, const slrh_server_config* cfg, slrh_server** out) {
    CtxLock lk = lockOf(s);
    if (!s || !cfg || !out) return SLRH_ERR_INVALID;
    if (!s->built) return fail(s->ctx, SLRH_ERR_STATE, "server before slrh_scene_build");
    if (!s->dev) return fail(s->ctx, SLRH_ERR_STATE, "host-only context has no device scene");
    if (cfg->reserved != 0) return fail(s->ctx, SLRH_ERR_INVALID, "reserved server field must be 0");
    if (s->data.materials.empty() || s->data.idx.empty()) return fail(s->ctx, SLRH_ERR_STATE, "empty scene");
    // 0: the variant the autotuner chose for this scene in a render (0 if none)
    const int tv = cfg->tune_trace == 0 ? tunedTraceVariant(s->dev) : (int)cfg->tune_trace - 1;
    if (tv >= numTraceVariants()) return fail(s->ctx, SLRH_ERR_INVALID, "unknown traversal variant");
    std::unique_ptr<slrh_server> sv(new (std::nothrow) slrh_server());
    if (!sv) return fail(s->ctx, SLRH_ERR_NOMEM, "out of host memory");
    sv->ctx = s->ctx;
    sv->srv.reset(new Server());
    std::string err;
    int r = sv->srv->start(s->dev, cfg->path_state_bytes, tv, &err);
    if (r) return fail(s->ctx, r, err);
    *out = sv.release();
    return SLRH_OK;
}

int slrh_server_destroy(slrh_server* sv) {
    delete sv;
    return SLRH_OK;
}

int slrh_submit(slrh_server* sv, const slrh_request* req, uint64_t* ticket) {
    if (!sv || !req || !ticket) return SLRH_ERR_INVALID;
    JobSpec js;
    int r = frameSpec(sv->ctx, req->frame, req->max_depth, &js);
    if (r) return r;
    std::string err;
    if ((r = sv->srv->submit(js, req->slots_per_pixel, ticket, &err))) return fail(sv->ctx, r, err);
    return SLRH_OK;
}

int slrh_poll(slrh_server* sv, uint64_t ticket, int* state) {
    if (!sv || !state) return SLRH_ERR_INVALID;
    if (sv->srv->state(ticket, state)) return fail(sv->ctx, SLRH_ERR_INVALID, "unknown ticket");
    return SLRH_OK;
}

int slrh_wait(slrh_server* sv, uint64_t ticket, int timeout_ms, int* state) {
    if (!sv || !state) return SLRH_ERR_INVALID;
    if (sv->srv->wait(ticket, timeout_ms, state)) return fail(sv->ctx, SLRH_ERR_INVALID, "unknown ticket");
    return SLRH_OK;
}

int slrh_cancel(slrh_server* sv, uint64_t ticket) {
    if (!sv) return SLRH_ERR_INVALID;
    if (sv->srv->cancel(ticket)) return fail(sv->ctx, SLRH_ERR_INVALID, "unknown ticket");
    return SLRH_OK;
}

int slrh_server_stats(slrh_server* sv, double* stats) {
    if (!sv || !stats) return SLRH_ERR_INVALID;
    sv->srv->stats(stats);
    return SLRH_OK;
}

}  // extern "C"
