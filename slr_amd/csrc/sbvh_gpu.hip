// Device half of the SBVH build: the binning and partitioning of the large
// nodes at the top of the tree, whose fragments (a triangle reference with
// its possibly clipped bounds) stay in device memory, while the host keeps
// the recursion, the SAH decisions and the small subtrees (bvh_build.cpp).
// Each step reproduces exactly what the host builder computes for the same
// node (SBVH::buildRecursive, libSLR/Accelerator/SBVH.h:57-349):
//   - node and centroid bounds, bin boxes: unions of boxes, i.e. per
//     component minima and maxima, which do not depend on the order they
//     are taken in; taken here as atomic minima/maxima of an order-keeping
//     integer encoding of the floats, first in LDS per block;
//   - bin indices, clipped triangle bounds (TriangleSurfaceShape::
//     choppedBounds / splitBounds, TriangleSurfaceShape.cpp:44-153) and
//     partition tests: the host's float expressions per fragment (compiled
//     without contraction, correctly rounded division);
//   - bin counts and, with unit triangle costs, the bins' cost sums;
//   - the order of the fragments after a split: an object split leaves them
//     as std::partition does (the k-th failing fragment from the left swaps
//     with the k-th passing one from the right, computed from prefix sums),
//     a spatial split keeps their order on both sides (stable compaction).
// So the tree equals the host builder's node for node and reference for
// reference.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_scene.h"
#include "sbvh_gpu.h"

namespace slrh {

namespace {

constexpr int kSbBlock = 256;
constexpr uint32_t kObjBins = 32, kSpBins = 16;

// fragment: bounds min xyz + primitive id, bounds max xyz + 0
struct __align__(16) GFrag {
    float mn[3];
    uint32_t prim;
    float mx[3];
    uint32_t pad;
};

#define SB_TRY(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// order-keeping encoding of a float as an unsigned integer
__host__ __device__ __forceinline__ uint32_t enc(float f) {
    uint32_t b;
    memcpy(&b, &f, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ __device__ __forceinline__ float dec(uint32_t u) {
    const uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    float f;
    memcpy(&f, &b, 4);
    return f;
}
constexpr uint32_t kEncPosInf = 0xFF800000u;   // enc(+inf)
constexpr uint32_t kEncNegInf = 0x007FFFFFu;   // enc(-inf)

struct DBox {
    float mn[3], mx[3];
};
__device__ __forceinline__ DBox emptyBox() {
    DBox b;
    for (int a = 0; a < 3; ++a) { b.mn[a] = INFINITY; b.mx[a] = -INFINITY; }
    return b;
}
__device__ __forceinline__ void unifyP(DBox& b, const float* p) {
    for (int a = 0; a < 3; ++a) { b.mn[a] = fminf(b.mn[a], p[a]); b.mx[a] = fmaxf(b.mx[a], p[a]); }
}
__device__ __forceinline__ DBox pointBox(const float* p) {
    DBox b;
    for (int a = 0; a < 3; ++a) { b.mn[a] = p[a]; b.mx[a] = p[a]; }
    return b;
}
__device__ __forceinline__ bool validBox(const DBox& b) {
    return (b.mx[0] - b.mn[0]) >= 0 && (b.mx[1] - b.mn[1]) >= 0 && (b.mx[2] - b.mn[2]) >= 0;
}
__device__ __forceinline__ DBox isectBox(const DBox& a, const float* mn, const float* mx) {
    DBox r;
    for (int k = 0; k < 3; ++k) { r.mn[k] = fmaxf(a.mn[k], mn[k]); r.mx[k] = fminf(a.mx[k], mx[k]); }
    return r;
}

struct TriP {
    float p[3][3];
};
__device__ __forceinline__ TriP loadTri(const float* tris, uint32_t prim) {
    TriP t;
    for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) t.p[k][a] = tris[9 * (size_t)prim + 3 * k + a];
    return t;
}
// the three vertices ordered by one coordinate as std::sort orders three
// elements (insertion sort, strict less: equal keys keep their order)
__device__ __forceinline__ void sortByAxis(TriP& t, int axis) {
    for (int i = 1; i < 3; ++i) {
        float v[3] = {t.p[i][0], t.p[i][1], t.p[i][2]};
        int j = i;
        while (j > 0 && v[axis] < t.p[j - 1][axis]) {
            for (int a = 0; a < 3; ++a) t.p[j][a] = t.p[j - 1][a];
            --j;
        }
        for (int a = 0; a < 3; ++a) t.p[j][a] = v[a];
    }
}
__device__ __forceinline__ void lerpP(const float* a, const float* b, float t, float* out) {
    const float s = 1 - t;
    for (int k = 0; k < 3; ++k) out[k] = s * a[k] + t * b[k];
}
__device__ __forceinline__ DBox triBoundsOf(const TriP& t) {
    DBox b = pointBox(t.p[0]);
    unifyP(b, t.p[1]);
    unifyP(b, t.p[2]);
    return b;
}

// TriangleSurfaceShape::choppedBounds (TriangleSurfaceShape.cpp:44-99), as
// the host builder computes it
__device__ DBox choppedTri(const TriP& tri0, int axis, float minChop, float maxChop) {
    const float chop[2] = {minChop, maxChop};
    TriP tri = tri0;
    sortByAxis(tri, axis);
    const float minPos = tri.p[0][axis], maxPos = tri.p[2][axis];
    if (minPos >= maxChop || maxPos <= minChop) return emptyBox();
    if (minPos >= minChop && maxPos <= maxChop) return triBoundsOf(tri0);
    uint32_t n = 0;
    float isect[4][3];
    for (int from = 0; from < 2; ++from) {
        const float* pa = tri.p[from];
        for (int to = from + 1; to < 3; ++to) {
            const float* pb = tri.p[to];
            const float dAB = pb[axis] - pa[axis];
            for (int c = 0; c < 2; ++c) {
                const float dAP = chop[c] - pa[axis];
                const float dPB = chop[c] - pb[axis];
                const float t = dAP / dAB;
                if (dAP > 0 && dPB <= 0 && n < 4) lerpP(pa, pb, t, isect[n++]);
            }
        }
    }
    DBox r = emptyBox();
    const float mid = tri.p[1][axis];
    if (mid >= minChop && mid < maxChop) unifyP(r, tri.p[1]);
    for (uint32_t i = 0; i < n; ++i) unifyP(r, isect[i]);
    if (n == 2) unifyP(r, maxPos < maxChop ? tri.p[2] : tri.p[0]);
    return r;
}

// TriangleSurfaceShape::splitBounds (TriangleSurfaceShape.cpp:101-153)
__device__ void splitTri(const TriP& tri0, int axis, float pos, DBox* left, DBox* right) {
    TriP tri = tri0;
    sortByAxis(tri, axis);
    const float minPos = tri.p[0][axis], maxPos = tri.p[2][axis];
    if (pos <= minPos) { *left = emptyBox(); *right = triBoundsOf(tri0); return; }
    if (pos >= maxPos) { *left = triBoundsOf(tri0); *right = emptyBox(); return; }
    uint32_t n = 0;
    float isect[2][3];
    for (int from = 0; from < 2; ++from) {
        const float* pa = tri.p[from];
        for (int to = from + 1; to < 3; ++to) {
            const float* pb = tri.p[to];
            const float dAB = pb[axis] - pa[axis];
            const float dAP = pos - pa[axis];
            const float dPB = pos - pb[axis];
            const float t = dAP / dAB;
            if (dAP > 0 && dPB <= 0 && n < 2) lerpP(pa, pb, t, isect[n++]);
        }
    }
    *left = pointBox(tri.p[0]);
    *right = pointBox(tri.p[2]);
    if (tri.p[1][axis] < pos) unifyP(*left, tri.p[1]);
    else unifyP(*right, tri.p[1]);
    for (uint32_t i = 0; i < n; ++i) { unifyP(*left, isect[i]); unifyP(*right, isect[i]); }
}

__device__ __forceinline__ void atomicBox(uint32_t* slot, const DBox& b) {
    for (int a = 0; a < 3; ++a) {
        atomicMin(slot + a, enc(b.mn[a]));
        atomicMax(slot + 3 + a, enc(b.mx[a]));
    }
}
__device__ __forceinline__ void initBoxSlots(uint32_t* slot) {
    for (int a = 0; a < 3; ++a) { slot[a] = kEncPosInf; slot[3 + a] = kEncNegInf; }
}

__global__ void k_sb_init(const float* tris, uint32_t n, GFrag* out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const DBox b = triBoundsOf(loadTri(tris, i));
        GFrag f;
        for (int a = 0; a < 3; ++a) { f.mn[a] = b.mn[a]; f.mx[a] = b.mx[a]; }
        f.prim = i;
        f.pad = 0;
        out[i] = f;
    }
}

// node bounds and centroid bounds: out[0..5] the box, out[6..11] the centroids' box
__global__ __launch_bounds__(kSbBlock) void k_sb_bounds(const GFrag* f, uint32_t n, uint32_t* out) {
    __shared__ uint32_t s[12];
    if (threadIdx.x < 12) s[threadIdx.x] = (threadIdx.x % 6) < 3 ? kEncPosInf : kEncNegInf;
    __syncthreads();
    DBox b = emptyBox(), c = emptyBox();
    for (uint32_t i = blockIdx.x * kSbBlock + threadIdx.x; i < n; i += gridDim.x * kSbBlock) {
        const GFrag g = f[i];
        float cen[3];
        for (int a = 0; a < 3; ++a) {
            b.mn[a] = fminf(b.mn[a], g.mn[a]);
            b.mx[a] = fmaxf(b.mx[a], g.mx[a]);
            cen[a] = (g.mn[a] + g.mx[a]) * 0.5f;
        }
        unifyP(c, cen);
    }
    atomicBox(s, b);
    atomicBox(s + 6, c);
    __syncthreads();
    if (threadIdx.x < 12) {
        if ((threadIdx.x % 6) < 3) atomicMin(out + threadIdx.x, s[threadIdx.x]);
        else atomicMax(out + threadIdx.x, s[threadIdx.x]);
    }
}

// object bins: out[7 * b] count, out[7 * b + 1 ..] box
__global__ __launch_bounds__(kSbBlock) void k_sb_obj(const GFrag* f, uint32_t n, int axis, float lo, float hi,
                                                     uint32_t* out) {
    __shared__ uint32_t s[kObjBins * 7];
    for (uint32_t i = threadIdx.x; i < kObjBins; i += kSbBlock) {
        s[7 * i] = 0;
        initBoxSlots(s + 7 * i + 1);
    }
    __syncthreads();
    for (uint32_t i = blockIdx.x * kSbBlock + threadIdx.x; i < n; i += gridDim.x * kSbBlock) {
        const GFrag g = f[i];
        const float c = (g.mn[axis] + g.mx[axis]) * 0.5f;
        uint32_t b = (uint32_t)(kObjBins * ((c - lo) / (hi - lo)));
        b = min(b, kObjBins - 1);
        atomicAdd(s + 7 * b, 1u);
        DBox bx;
        for (int a = 0; a < 3; ++a) { bx.mn[a] = g.mn[a]; bx.mx[a] = g.mx[a]; }
        atomicBox(s + 7 * b + 1, bx);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kObjBins * 7; i += kSbBlock) {
        const uint32_t k = i % 7;
        if (k == 0) atomicAdd(out + i, s[i]);
        else if (k <= 3) atomicMin(out + i, s[i]);
        else atomicMax(out + i, s[i]);
    }
}

// spatial bins: out[8 * b] entering count, out[8 * b + 1] leaving count,
// out[8 * b + 2 ..] box of the clipped fragments
__global__ __launch_bounds__(kSbBlock) void k_sb_sp(const GFrag* f, uint32_t n, const float* tris, int axis, float lo,
                                                    float hi, float binWidth, uint32_t* out) {
    __shared__ uint32_t s[kSpBins * 8];
    for (uint32_t i = threadIdx.x; i < kSpBins; i += kSbBlock) {
        s[8 * i] = s[8 * i + 1] = 0;
        initBoxSlots(s + 8 * i + 2);
    }
    __syncthreads();
    for (uint32_t i = blockIdx.x * kSbBlock + threadIdx.x; i < n; i += gridDim.x * kSbBlock) {
        const GFrag g = f[i];
        uint32_t eb = (uint32_t)(kSpBins * ((g.mn[axis] - lo) / (hi - lo)));
        uint32_t xb = (uint32_t)(kSpBins * ((g.mx[axis] - lo) / (hi - lo)));
        eb = min(eb, kSpBins - 1);
        xb = min(xb, kSpBins - 1);
        atomicAdd(s + 8 * eb, 1u);
        atomicAdd(s + 8 * xb + 1, 1u);
        const TriP t = loadTri(tris, g.prim);
        for (int b = (int)eb; b <= (int)xb; ++b) {
            const float blo = b * binWidth + lo;
            const DBox c = choppedTri(t, axis, blo, blo + binWidth);
            atomicBox(s + 8 * b + 2, isectBox(c, g.mn, g.mx));
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kSpBins * 8; i += kSbBlock) {
        const uint32_t k = i % 8;
        if (k < 2) atomicAdd(out + i, s[i]);
        else if (k <= 4) atomicMin(out + i, s[i]);
        else atomicMax(out + i, s[i]);
    }
}

__global__ void k_sb_pred(const GFrag* f, uint32_t n, int axis, float pivot, uint32_t* flag) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const GFrag g = f[i];
        flag[i] = ((g.mn[axis] + g.mx[axis]) * 0.5f) < pivot ? 1u : 0u;
    }
}

// std::partition's pairing: the k-th failing fragment in [0, m) and the
// k-th passing one from the right in [m, n) trade places
__global__ void k_sb_pairs(const uint32_t* flag, const uint32_t* before, uint32_t n, uint32_t m, uint32_t* fpos,
                           uint32_t* tpos) {
    const uint32_t tm = before[m < n ? m : n - 1] + (m < n ? 0u : flag[n - 1]);   // passing ones in [0, m)
    const uint32_t s = m - tm;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (i < m && !flag[i]) fpos[i - before[i]] = i;
        else if (i >= m && flag[i]) tpos[s - 1 - (before[i] - tm)] = i;
    }
}
__global__ void k_sb_permute(const GFrag* in, uint32_t n, const uint32_t* flag, const uint32_t* before, uint32_t m,
                             const uint32_t* fpos, const uint32_t* tpos, GFrag* out) {
    const uint32_t tm = before[m < n ? m : n - 1] + (m < n ? 0u : flag[n - 1]);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t src = i;
        if (i < m && !flag[i]) src = tpos[i - before[i]];
        else if (i >= m && flag[i]) src = fpos[(m - tm) - 1 - (before[i] - tm)];
        out[i] = in[src];
    }
}

// spatial split: 1 in fl / fr when the fragment (or its clipped part) goes left / right
__global__ void k_sb_spflags(const GFrag* f, uint32_t n, const float* tris, int axis, float lo, float hi,
                             uint32_t plane, float splitPos, uint32_t* fl, uint32_t* fr) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const GFrag g = f[i];
        uint32_t eb = (uint32_t)(kSpBins * ((g.mn[axis] - lo) / (hi - lo)));
        uint32_t xb = (uint32_t)(kSpBins * ((g.mx[axis] - lo) / (hi - lo)));
        eb = min(eb, kSpBins - 1);
        xb = min(xb, kSpBins - 1);
        uint32_t l = 0, r = 0;
        if (xb <= plane) l = 1;
        else if (eb > plane) r = 1;
        else {
            DBox lb, rb;
            splitTri(loadTri(tris, g.prim), axis, splitPos, &lb, &rb);
            l = validBox(lb) ? 1u : 0u;
            r = validBox(rb) ? 1u : 0u;
        }
        fl[i] = l;
        fr[i] = r;
    }
}
__global__ void k_sb_spscatter(const GFrag* f, uint32_t n, const float* tris, int axis, float lo, float hi,
                               uint32_t plane, float splitPos, const uint32_t* fl, const uint32_t* pl,
                               const uint32_t* fr, const uint32_t* pr, GFrag* left, GFrag* right) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const GFrag g = f[i];
        uint32_t eb = (uint32_t)(kSpBins * ((g.mn[axis] - lo) / (hi - lo)));
        uint32_t xb = (uint32_t)(kSpBins * ((g.mx[axis] - lo) / (hi - lo)));
        eb = min(eb, kSpBins - 1);
        xb = min(xb, kSpBins - 1);
        if (xb <= plane || eb > plane) {
            if (fl[i]) left[pl[i]] = g;
            if (fr[i]) right[pr[i]] = g;
            continue;
        }
        DBox lb, rb;
        splitTri(loadTri(tris, g.prim), axis, splitPos, &lb, &rb);
        if (fl[i]) {
            const DBox c = isectBox(lb, g.mn, g.mx);
            GFrag o = g;
            for (int a = 0; a < 3; ++a) { o.mn[a] = c.mn[a]; o.mx[a] = c.mx[a]; }
            left[pl[i]] = o;
        }
        if (fr[i]) {
            const DBox c = isectBox(rb, g.mn, g.mx);
            GFrag o = g;
            for (int a = 0; a < 3; ++a) { o.mn[a] = c.mn[a]; o.mx[a] = c.mx[a]; }
            right[pr[i]] = o;
        }
    }
}

template <typename T>
struct DArr {
    T* p = nullptr;
    size_t n = 0;
    DArr() {}
    explicit DArr(size_t count) { alloc(count); }
    DArr(const DArr&) = delete;
    DArr& operator=(const DArr&) = delete;
    ~DArr() { if (p) (void)hipFree(p); }
    void alloc(size_t count) {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = count;
        if (count) SB_TRY(hipMalloc((void**)&p, count * sizeof(T)));
    }
};

Box boxFromEnc(const uint32_t* e) {
    Box b;
    for (int a = 0; a < 3; ++a) { b.mn[a] = dec(e[a]); b.mx[a] = dec(e[3 + a]); }
    return b;
}

}  // namespace

struct DevFrags {
    DArr<GFrag> f;
    uint32_t count = 0;
};

struct SbvhGpu::Impl {
    DArr<float> tris;
    uint32_t ntris = 0;
    int blocks = 1024;
    hipStream_t stream = 0;
    DArr<uint32_t> small;           // bins and bounds results
    DArr<uint32_t> flag, before, flag2, before2, fpos, tpos;
    DArr<unsigned char> scanTemp;
    size_t scanTempBytes = 0;
    std::vector<uint32_t> host;

    uint32_t grid(uint32_t n) const { return std::max(1u, std::min((uint32_t)blocks, (n + kSbBlock - 1) / kSbBlock)); }
    void ensure(size_t n) {
        if (flag.n >= n) return;
        flag.alloc(n); before.alloc(n); flag2.alloc(n); before2.alloc(n); fpos.alloc(n); tpos.alloc(n);
        size_t bytes = 0;
        SB_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, flag.p, before.p, (int)n, stream));
        scanTemp.alloc(bytes);
        scanTempBytes = bytes;
    }
    void scan(const uint32_t* in, uint32_t* out, uint32_t n) {
        size_t bytes = scanTempBytes;
        SB_TRY(hipcub::DeviceScan::ExclusiveSum(scanTemp.p, bytes, in, out, (int)n, stream));
    }
    uint32_t total(const uint32_t* flagArr, const uint32_t* beforeArr, uint32_t n) {
        uint32_t a = 0, b = 0;
        SB_TRY(hipMemcpy(&a, beforeArr + n - 1, 4, hipMemcpyDeviceToHost));
        SB_TRY(hipMemcpy(&b, flagArr + n - 1, 4, hipMemcpyDeviceToHost));
        return a + b;
    }
};

SbvhGpu::SbvhGpu() : m(new Impl()) {}
SbvhGpu::~SbvhGpu() = default;

std::unique_ptr<SbvhGpu> SbvhGpu::create(const std::vector<float>& triPos) {
    std::unique_ptr<SbvhGpu> g(new SbvhGpu());
    Impl& m = *g->m;
    m.ntris = (uint32_t)(triPos.size() / 9);
    m.tris.alloc(triPos.size());
    SB_TRY(hipMemcpy(m.tris.p, triPos.data(), triPos.size() * sizeof(float), hipMemcpyHostToDevice));
    int dev = 0, cus = 256;
    SB_TRY(hipGetDevice(&dev));
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    m.blocks = std::max(1, cus) * 4;
    m.small.alloc(256);
    m.host.resize(256);
    return g;
}

DevFrags* SbvhGpu::root() {
    Impl& m = *this->m;
    std::unique_ptr<DevFrags> d(new DevFrags());
    d->f.alloc(m.ntris);
    d->count = m.ntris;
    hipLaunchKernelGGL(k_sb_init, dim3(m.grid(m.ntris)), dim3(kSbBlock), 0, m.stream, m.tris.p, m.ntris, d->f.p);
    SB_TRY(hipGetLastError());
    return d.release();
}

void SbvhGpu::release(DevFrags* d) { delete d; }
uint32_t SbvhGpu::count(const DevFrags* d) { return d->count; }

void SbvhGpu::bounds(DevFrags* d, Box* parent, Box* centroids) {
    Impl& m = *this->m;
    uint32_t init[12];
    for (int i = 0; i < 12; ++i) init[i] = (i % 6) < 3 ? kEncPosInf : kEncNegInf;
    SB_TRY(hipMemcpy(m.small.p, init, sizeof(init), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_sb_bounds, dim3(m.grid(d->count)), dim3(kSbBlock), 0, m.stream, d->f.p, d->count, m.small.p);
    SB_TRY(hipGetLastError());
    SB_TRY(hipMemcpy(m.host.data(), m.small.p, sizeof(init), hipMemcpyDeviceToHost));
    *parent = boxFromEnc(m.host.data());
    *centroids = boxFromEnc(m.host.data() + 6);
}

void SbvhGpu::objectBins(DevFrags* d, int axis, float lo, float hi, uint32_t* n, Box* box) {
    Impl& m = *this->m;
    std::vector<uint32_t> init(kObjBins * 7);
    for (uint32_t b = 0; b < kObjBins; ++b) {
        init[7 * b] = 0;
        for (int a = 0; a < 3; ++a) { init[7 * b + 1 + a] = kEncPosInf; init[7 * b + 4 + a] = kEncNegInf; }
    }
    SB_TRY(hipMemcpy(m.small.p, init.data(), init.size() * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_sb_obj, dim3(m.grid(d->count)), dim3(kSbBlock), 0, m.stream, d->f.p, d->count, axis, lo, hi,
                       m.small.p);
    SB_TRY(hipGetLastError());
    SB_TRY(hipMemcpy(init.data(), m.small.p, init.size() * 4, hipMemcpyDeviceToHost));
    for (uint32_t b = 0; b < kObjBins; ++b) {
        n[b] = init[7 * b];
        box[b] = boxFromEnc(&init[7 * b + 1]);
    }
}

void SbvhGpu::spatialBins(DevFrags* d, int axis, float lo, float hi, float binWidth, uint32_t* nIn, uint32_t* nOut,
                          Box* box) {
    Impl& m = *this->m;
    std::vector<uint32_t> init(kSpBins * 8);
    for (uint32_t b = 0; b < kSpBins; ++b) {
        init[8 * b] = init[8 * b + 1] = 0;
        for (int a = 0; a < 3; ++a) { init[8 * b + 2 + a] = kEncPosInf; init[8 * b + 5 + a] = kEncNegInf; }
    }
    SB_TRY(hipMemcpy(m.small.p, init.data(), init.size() * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_sb_sp, dim3(m.grid(d->count)), dim3(kSbBlock), 0, m.stream, d->f.p, d->count, m.tris.p, axis,
                       lo, hi, binWidth, m.small.p);
    SB_TRY(hipGetLastError());
    SB_TRY(hipMemcpy(init.data(), m.small.p, init.size() * 4, hipMemcpyDeviceToHost));
    for (uint32_t b = 0; b < kSpBins; ++b) {
        nIn[b] = init[8 * b];
        nOut[b] = init[8 * b + 1];
        box[b] = boxFromEnc(&init[8 * b + 2]);
    }
}

void SbvhGpu::partitionObject(DevFrags* d, int axis, float pivot, DevFrags** left, DevFrags** right) {
    Impl& m = *this->m;
    const uint32_t n = d->count;
    m.ensure(n);
    hipLaunchKernelGGL(k_sb_pred, dim3(m.grid(n)), dim3(kSbBlock), 0, m.stream, d->f.p, n, axis, pivot, m.flag.p);
    SB_TRY(hipGetLastError());
    m.scan(m.flag.p, m.before.p, n);
    const uint32_t mid = m.total(m.flag.p, m.before.p, n);
    DArr<GFrag> perm(n);
    hipLaunchKernelGGL(k_sb_pairs, dim3(m.grid(n)), dim3(kSbBlock), 0, m.stream, m.flag.p, m.before.p, n, mid, m.fpos.p,
                       m.tpos.p);
    SB_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_sb_permute, dim3(m.grid(n)), dim3(kSbBlock), 0, m.stream, d->f.p, n, m.flag.p, m.before.p, mid,
                       m.fpos.p, m.tpos.p, perm.p);
    SB_TRY(hipGetLastError());
    // SBVH.h: split = max(mid, 1); a partition leaving nothing on the right is invalid
    const uint32_t split = std::max(mid, 1u);
    if (split >= n) throw std::runtime_error("SBVH: invalid object partition");
    std::unique_ptr<DevFrags> l(new DevFrags()), r(new DevFrags());
    l->f.alloc(split);
    l->count = split;
    r->f.alloc(n - split);
    r->count = n - split;
    SB_TRY(hipMemcpy(l->f.p, perm.p, (size_t)split * sizeof(GFrag), hipMemcpyDeviceToDevice));
    SB_TRY(hipMemcpy(r->f.p, perm.p + split, (size_t)(n - split) * sizeof(GFrag), hipMemcpyDeviceToDevice));
    *left = l.release();
    *right = r.release();
}

void SbvhGpu::partitionSpatial(DevFrags* d, int axis, float lo, float hi, uint32_t plane, float splitPos,
                               DevFrags** left, DevFrags** right) {
    Impl& m = *this->m;
    const uint32_t n = d->count;
    m.ensure(n);
    hipLaunchKernelGGL(k_sb_spflags, dim3(m.grid(n)), dim3(kSbBlock), 0, m.stream, d->f.p, n, m.tris.p, axis, lo, hi,
                       plane, splitPos, m.flag.p, m.flag2.p);
    SB_TRY(hipGetLastError());
    m.scan(m.flag.p, m.before.p, n);
    m.scan(m.flag2.p, m.before2.p, n);
    const uint32_t nl = m.total(m.flag.p, m.before.p, n), nr = m.total(m.flag2.p, m.before2.p, n);
    if (nl == 0 || nr == 0) throw std::runtime_error("SBVH: invalid spatial partition");
    std::unique_ptr<DevFrags> l(new DevFrags()), r(new DevFrags());
    l->f.alloc(nl);
    l->count = nl;
    r->f.alloc(nr);
    r->count = nr;
    hipLaunchKernelGGL(k_sb_spscatter, dim3(m.grid(n)), dim3(kSbBlock), 0, m.stream, d->f.p, n, m.tris.p, axis, lo, hi,
                       plane, splitPos, m.flag.p, m.before.p, m.flag2.p, m.before2.p, l->f.p, r->f.p);
    SB_TRY(hipGetLastError());
    *left = l.release();
    *right = r.release();
}

void SbvhGpu::download(DevFrags* d, std::vector<uint32_t>* prims, std::vector<Box>* boxes) {
    std::vector<GFrag> h(d->count);
    SB_TRY(hipMemcpy(h.data(), d->f.p, h.size() * sizeof(GFrag), hipMemcpyDeviceToHost));
    prims->resize(h.size());
    boxes->resize(h.size());
    for (size_t i = 0; i < h.size(); ++i) {
        (*prims)[i] = h[i].prim;
        Box& b = (*boxes)[i];
        for (int a = 0; a < 3; ++a) { b.mn[a] = h[i].mn[a]; b.mx[a] = h[i].mx[a]; }
    }
}

}  // namespace slrh
