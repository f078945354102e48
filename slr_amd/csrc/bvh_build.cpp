// SBVH build and QBVH collapse on the host.
//
// The builder makes the same decisions as the reference builder so that the
// trees (and with them hit IDs, traversal order and tie-breaking) are the
// same: 32 centroid bins on the widest centroid axis, 16 spatial bins on the
// widest bounds axis tried only when the object split's children overlap by
// more than 1e-5 of the root area, traversal cost 1.2 and unit triangle cost
// (libSLR/Accelerator/SBVH.h:57-349), triangles clipped against bins exactly
// as TriangleSurfaceShape::choppedBounds / splitBounds do
// (libSLR/SurfaceShape/TriangleSurfaceShape.cpp:44-153), then grandchildren
// collapsed into 4-wide nodes (libSLR/Accelerator/QBVH.h:84-201). Every float
// expression is evaluated in the reference's order; the file is compiled with
// -ffp-contract=off so no multiply-add is fused.
//
// What differs is the machinery: fragment lists are per-node vectors rather
// than one shared array shifted with copy_backward (whose 5x budget the
// reference never checks), independent subtrees are built on separate
// threads, then stitched back into DFS pre-order, and with a device the
// nodes of at least kDeviceNodeFrags fragments at the top of a triangle tree
// keep their fragments in device memory, where their bounds, bins and
// partitions are computed (sbvh_gpu.hip) with the same results.
#include "host_scene.h"
#include "sbvh_gpu.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <future>
#include <memory>
#include <stdexcept>

namespace slrh {
namespace {

struct Frag {
    uint32_t prim;
    Box box;
};

struct Tri {
    V3 p[3];
};

inline float axisOf(const V3& p, int a) { return a == 0 ? p.x : (a == 1 ? p.y : p.z); }
inline V3 lerpPoint(const V3& a, const V3& b, float t) {
    float s = 1 - t;
    return v3(s * a.x + t * b.x, s * a.y + t * b.y, s * a.z + t * b.z);
}

Box triBounds(const Tri& t) {
    Box b = Box::point(t.p[0]);
    b.unify(t.p[1]);
    b.unify(t.p[2]);
    return b;
}

// TriangleSurfaceShape::choppedBounds (TriangleSurfaceShape.cpp:44-99).
Box choppedBounds(const Tri& tri, int axis, float minChop, float maxChop) {
    const float chop[2] = {minChop, maxChop};
    V3 p[3] = {tri.p[0], tri.p[1], tri.p[2]};
    std::sort(p, p + 3, [axis](const V3& a, const V3& b) { return axisOf(a, axis) < axisOf(b, axis); });
    float minPos = axisOf(p[0], axis);
    float maxPos = axisOf(p[2], axis);
    if (minPos >= maxChop || maxPos <= minChop) return Box();
    if (minPos >= minChop && maxPos <= maxChop) return triBounds(tri);
    uint32_t n = 0;
    V3 isect[4];
    for (int from = 0; from < 2; ++from) {
        const V3& pa = p[from];
        for (int to = from + 1; to < 3; ++to) {
            const V3& pb = p[to];
            float dAB = axisOf(pb, axis) - axisOf(pa, axis);
            for (int c = 0; c < 2; ++c) {
                float dAP = chop[c] - axisOf(pa, axis);
                float dPB = chop[c] - axisOf(pb, axis);
                float t = dAP / dAB;
                if (dAP > 0 && dPB <= 0 && n < 4) isect[n++] = lerpPoint(pa, pb, t);
            }
        }
    }
    Box r;
    float mid = axisOf(p[1], axis);
    if (mid >= minChop && mid < maxChop) r.unify(p[1]);
    for (uint32_t i = 0; i < n; ++i) r.unify(isect[i]);
    if (n == 2) r.unify(maxPos < maxChop ? p[2] : p[0]);
    return r;
}

// TriangleSurfaceShape::splitBounds (TriangleSurfaceShape.cpp:101-153).
void splitBounds(const Tri& tri, int axis, float pos, Box* left, Box* right) {
    V3 p[3] = {tri.p[0], tri.p[1], tri.p[2]};
    std::sort(p, p + 3, [axis](const V3& a, const V3& b) { return axisOf(a, axis) < axisOf(b, axis); });
    float minPos = axisOf(p[0], axis);
    float maxPos = axisOf(p[2], axis);
    if (pos <= minPos) { *left = Box(); *right = triBounds(tri); return; }
    if (pos >= maxPos) { *left = triBounds(tri); *right = Box(); return; }
    uint32_t n = 0;
    V3 isect[2];
    for (int from = 0; from < 2; ++from) {
        const V3& pa = p[from];
        for (int to = from + 1; to < 3; ++to) {
            const V3& pb = p[to];
            float dAB = axisOf(pb, axis) - axisOf(pa, axis);
            float dAP = pos - axisOf(pa, axis);
            float dPB = pos - axisOf(pb, axis);
            float t = dAP / dAB;
            if (dAP > 0 && dPB <= 0 && n < 2) isect[n++] = lerpPoint(pa, pb, t);
        }
    }
    *left = Box::point(p[0]);
    *right = Box::point(p[2]);
    if (axisOf(p[1], axis) < pos) left->unify(p[1]);
    else right->unify(p[1]);
    for (uint32_t i = 0; i < n; ++i) { left->unify(isect[i]); right->unify(isect[i]); }
}

// Object::choppedBounds / splitBounds (libSLR/Core/object.h:50-80): the
// default for objects that are not triangles, here the world bounds of an
// instance.
Box choppedBoxBounds(const Box& b, int axis, float minChop, float maxChop) {
    if (maxChop < b.mn[axis]) return Box();
    if (minChop > b.mx[axis]) return Box();
    if (minChop < b.mn[axis] && maxChop > b.mx[axis]) return b;
    Box r = b;
    r.mn[axis] = std::max(minChop, r.mn[axis]);
    r.mx[axis] = std::min(maxChop, r.mx[axis]);
    return r;
}
void splitBoxBounds(const Box& b, int axis, float pos, Box* left, Box* right) {
    if (pos < b.mn[axis]) { *left = Box(); *right = b; return; }
    if (pos > b.mx[axis]) { *left = b; *right = Box(); return; }
    *left = b;
    left->mx[axis] = std::min(left->mx[axis], pos);
    *right = b;
    right->mn[axis] = std::max(right->mn[axis], pos);
}

// Nodes with at least this many fragments are binned and partitioned on
// the device when the build has one.
constexpr size_t kDeviceNodeFrags = 1u << 16;

// One node's fragments: a host vector, or (dev set) device memory.
struct NodeFrags {
    std::vector<Frag> host;
    DevFrags* dev = nullptr;
    NodeFrags() {}
    explicit NodeFrags(std::vector<Frag>&& h) : host(std::move(h)) {}
    NodeFrags(NodeFrags&& o) noexcept : host(std::move(o.host)), dev(o.dev) { o.dev = nullptr; }
    NodeFrags& operator=(NodeFrags&& o) noexcept {
        if (this != &o) {
            if (dev) SbvhGpu::release(dev);
            host = std::move(o.host);
            dev = o.dev;
            o.dev = nullptr;
        }
        return *this;
    }
    NodeFrags(const NodeFrags&) = delete;
    ~NodeFrags() { if (dev) SbvhGpu::release(dev); }
    size_t size() const { return dev ? SbvhGpu::count(dev) : host.size(); }
};

// A subtree in DFS pre-order with local indices.
struct Subtree {
    std::vector<SbvhNode> nodes;
    std::vector<uint32_t> refs;
    uint32_t depth = 0;
    uint32_t fragments = 0;
};

class Builder {
public:
    Builder(const std::vector<Tri>& tris, const std::vector<Box>& boxes, const std::vector<float>& boxCosts,
            float rootArea, int parallelDepth, SbvhGpu* gpu = nullptr)
        : m_tris(tris), m_boxes(boxes), m_boxCosts(boxCosts), m_rootArea(rootArea), m_parDepth(parallelDepth),
          m_gpu(gpu) {}

    void build(NodeFrags&& frags, uint32_t depth, Subtree* out);
    uint32_t deviceNodes() const { return m_deviceNodes; }
    // After a build from device fragments: the host subtrees the device
    // nodes left behind, built on `threads` threads, put in place of their
    // placeholders (the tree in DFS pre-order, as one recursive build makes it).
    void finishDeviceBuild(Subtree* skeleton, int threads);

private:
    // primitives 0..nt-1 are triangles, nt.. the boxes
    Box chopped(uint32_t prim, int axis, float lo, float hi) const {
        return prim < m_tris.size() ? choppedBounds(m_tris[prim], axis, lo, hi)
                                    : choppedBoxBounds(m_boxes[prim - m_tris.size()], axis, lo, hi);
    }
    void split(uint32_t prim, int axis, float pos, Box* l, Box* r) const {
        if (prim < m_tris.size()) splitBounds(m_tris[prim], axis, pos, l, r);
        else splitBoxBounds(m_boxes[prim - m_tris.size()], axis, pos, l, r);
    }
    // costForIntersect: 1 for a triangle (TriangleSurfaceShape.h:25), the
    // prototype tree's SAH cost for an instance (SBVH::costForIntersect)
    float cost(uint32_t prim) const { return prim < m_tris.size() ? 1.0f : m_boxCosts[prim - m_tris.size()]; }
    const std::vector<Tri>& m_tris;
    const std::vector<Box>& m_boxes;
    const std::vector<float>& m_boxCosts;
    float m_rootArea;
    int m_parDepth;
    SbvhGpu* m_gpu;
    uint32_t m_deviceNodes = 0;   // touched only by the thread that runs device nodes
    // Subtrees below the device nodes: their fragments wait in a list while
    // the device nodes are processed (in DFS order, one thread), and a
    // placeholder node (count kPlaceholder, first = list index) holds their
    // place; finishDeviceBuild builds them all at once on the host threads.
    static constexpr uint32_t kPlaceholder = 0xFFFFFFFFu;
    struct Pending {
        std::vector<Frag> frags;
        uint32_t depth = 0;
        Subtree tree;
    };
    std::vector<std::unique_ptr<Pending>> m_pending;

    // the two children of a device node: device children are built now,
    // host children are left as placeholders
    void recurseDevice(NodeFrags&& left, NodeFrags&& right, uint32_t depth, Subtree* out, uint32_t nodeIdx,
                       const Box& parent, int axis) {
        auto child = [&](NodeFrags&& c) -> uint32_t {
            const uint32_t idx = (uint32_t)out->nodes.size();
            if (c.dev) {
                build(std::move(c), depth, out);
            } else {
                out->nodes.emplace_back();
                SbvhNode& ph = out->nodes.back();
                ph.count = kPlaceholder;
                ph.first = (uint32_t)m_pending.size();
                std::unique_ptr<Pending> p(new Pending());
                p->frags = std::move(c.host);
                p->depth = depth;
                m_pending.push_back(std::move(p));
            }
            return idx;
        };
        const uint32_t c0 = child(std::move(left));
        const uint32_t c1 = child(std::move(right));
        SbvhNode& n = out->nodes[nodeIdx];
        n.box = parent; n.c0 = c0; n.c1 = c1; n.axis = (uint8_t)axis; n.first = n.count = 0;
    }

    // a device node's fragments brought to the host, in order
    void toHost(NodeFrags& nf) const {
        if (!nf.dev) return;
        std::vector<uint32_t> prims;
        std::vector<Box> boxes;
        m_gpu->download(nf.dev, &prims, &boxes);
        SbvhGpu::release(nf.dev);
        nf.dev = nullptr;
        nf.host.resize(prims.size());
        for (size_t i = 0; i < prims.size(); ++i) nf.host[i] = Frag{prims[i], boxes[i]};
    }
    // a child below the device threshold continues on the host
    NodeFrags settle(NodeFrags&& nf) const {
        if (nf.dev && nf.size() < kDeviceNodeFrags) toHost(nf);
        return std::move(nf);
    }

    static void append(Subtree* dst, const Subtree& src) {
        uint32_t nodeBase = (uint32_t)dst->nodes.size();
        uint32_t refBase = (uint32_t)dst->refs.size();
        for (const SbvhNode& n : src.nodes) {
            SbvhNode m = n;
            if (m.count == 0) { m.c0 += nodeBase; m.c1 += nodeBase; }
            else m.first += refBase;
            dst->nodes.push_back(m);
        }
        dst->refs.insert(dst->refs.end(), src.refs.begin(), src.refs.end());
        dst->depth = std::max(dst->depth, src.depth);
        dst->fragments += src.fragments;
    }

    void recurse(NodeFrags&& left, NodeFrags&& right, uint32_t depth, Subtree* out, uint32_t nodeIdx,
                 const Box& parent, int axis) {
        bool parallel = (int)depth < m_parDepth && left.size() + right.size() > 20000;
        if (parallel) {
            Subtree l, r;
            auto fut = std::async(std::launch::async, [&]() { build(std::move(right), depth, &r); });
            build(std::move(left), depth, &l);
            fut.get();
            uint32_t c0 = (uint32_t)out->nodes.size();
            append(out, l);
            uint32_t c1 = (uint32_t)out->nodes.size();
            append(out, r);
            SbvhNode& n = out->nodes[nodeIdx];
            n.box = parent; n.c0 = c0; n.c1 = c1; n.axis = (uint8_t)axis; n.first = n.count = 0;
        } else {
            uint32_t c0 = (uint32_t)out->nodes.size();
            build(std::move(left), depth, out);
            uint32_t c1 = (uint32_t)out->nodes.size();
            build(std::move(right), depth, out);
            SbvhNode& n = out->nodes[nodeIdx];
            n.box = parent; n.c0 = c0; n.c1 = c1; n.axis = (uint8_t)axis; n.first = n.count = 0;
        }
    }
};

// SBVH::buildRecursive (SBVH.h:57-349), writing nodes of this subtree into
// `out` starting at out->nodes.size(). A node whose fragments are in device
// memory gets its bounds, bins and partition from the device (triangles
// only, unit costs: the cost sums are the counts) and the same decisions.
void Builder::build(NodeFrags&& nf, uint32_t depth, Subtree* out) {
    uint32_t nodeIdx = (uint32_t)out->nodes.size();
    out->nodes.emplace_back();
    ++depth;
    out->depth = std::max(out->depth, depth);
    const bool onDev = nf.dev != nullptr;
    std::vector<Frag>& frags = nf.host;

    Box parent, centroids;
    float leafCost = 0.0f;
    if (onDev) {
        ++m_deviceNodes;
        m_gpu->bounds(nf.dev, &parent, &centroids);
        leafCost = (float)nf.size();   // 1.0f added once per fragment: exact below 2^24
    } else {
        for (const Frag& f : frags) {
            parent.unify(f.box);
            centroids.unify(f.box.centroid());
            leafCost += cost(f.prim);
        }
    }
    const int axisOP = centroids.widestAxis();
    const int axisSP = parent.widestAxis();
    const float areaParent = parent.area();
    const float pcMin = centroids.mn[axisOP], pcMax = centroids.mx[axisOP];
    const float pMin = parent.mn[axisSP], pMax = parent.mx[axisSP];
    const uint32_t numObjs = (uint32_t)nf.size();

    auto makeLeaf = [&]() {
        toHost(nf);
        SbvhNode& n = out->nodes[nodeIdx];
        n.box = parent;
        n.c0 = n.c1 = 0;
        n.first = (uint32_t)out->refs.size();
        n.count = numObjs;
        for (const Frag& f : frags) out->refs.push_back(f.prim);
        out->fragments += numObjs;
    };
    if (numObjs == 1) { makeLeaf(); return; }

    const float travCost = 1.2f;
    constexpr uint32_t kObjBins = 32, kSpBins = 16;

    struct ObjBin { Box box; uint32_t n = 0; float cost = 0.0f; };
    ObjBin ob[kObjBins];
    uint32_t planeOP = 0;
    float minCostOP = INFINITY;
    if ((pcMax - pcMin) > 0) {
        if (onDev) {
            uint32_t cnt[kObjBins];
            Box box[kObjBins];
            m_gpu->objectBins(nf.dev, axisOP, pcMin, pcMax, cnt, box);
            for (uint32_t b = 0; b < kObjBins; ++b) { ob[b].n = cnt[b]; ob[b].cost = (float)cnt[b]; ob[b].box = box[b]; }
        } else {
            for (const Frag& f : frags) {
                uint32_t b = (uint32_t)(kObjBins * ((f.box.center(axisOP) - pcMin) / (pcMax - pcMin)));
                b = std::min(b, kObjBins - 1);
                ++ob[b].n;
                ob[b].cost += cost(f.prim);
                ob[b].box.unify(f.box);
            }
        }
        for (uint32_t i = 0; i < kObjBins - 1; ++i) {
            Box b0, b1;
            float c0 = 0.0f, c1 = 0.0f;
            for (uint32_t j = 0; j <= i; ++j) { b0.unify(ob[j].box); c0 += ob[j].cost; }
            for (uint32_t j = i + 1; j < kObjBins; ++j) { b1.unify(ob[j].box); c1 += ob[j].cost; }
            float cost = travCost + (b0.area() * c0 + b1.area() * c1) / areaParent;
            if (cost < minCostOP) { minCostOP = cost; planeOP = i; }
        }
    }

    Box leftOP, rightOP;
    for (uint32_t j = 0; j <= planeOP; ++j) leftOP.unify(ob[j].box);
    for (uint32_t j = planeOP + 1; j < kObjBins; ++j) rightOP.unify(ob[j].box);
    Box overlap = intersect(leftOP, rightOP);
    float overlapArea = 0;
    if (overlap.valid()) overlapArea = overlap.area();

    struct SpBin { Box box; uint32_t nIn = 0, nOut = 0; float cIn = 0.0f, cOut = 0.0f; };
    const float binWidth = (pMax - pMin) / kSpBins;
    SpBin sb[kSpBins];
    uint32_t planeSP = 0;
    float minCostSP = INFINITY;
    const float alpha = 1e-5f;
    if (overlapArea / m_rootArea > alpha) {
        if (onDev) {
            uint32_t nIn[kSpBins], nOut[kSpBins];
            Box box[kSpBins];
            m_gpu->spatialBins(nf.dev, axisSP, pMin, pMax, binWidth, nIn, nOut, box);
            for (uint32_t b = 0; b < kSpBins; ++b) {
                sb[b].nIn = nIn[b]; sb[b].nOut = nOut[b];
                sb[b].cIn = (float)nIn[b]; sb[b].cOut = (float)nOut[b];
                sb[b].box = box[b];
            }
        } else {
            for (const Frag& f : frags) {
                uint32_t eb = (uint32_t)(kSpBins * ((f.box.mn[axisSP] - pMin) / (pMax - pMin)));
                uint32_t xb = (uint32_t)(kSpBins * ((f.box.mx[axisSP] - pMin) / (pMax - pMin)));
                eb = std::min(eb, kSpBins - 1);
                xb = std::min(xb, kSpBins - 1);
                ++sb[eb].nIn;
                ++sb[xb].nOut;
                sb[eb].cIn += cost(f.prim);
                sb[xb].cOut += cost(f.prim);
                for (int b = (int)eb; b <= (int)xb; ++b) {
                    float lo = b * binWidth + pMin;
                    Box c = chopped(f.prim, axisSP, lo, lo + binWidth);
                    sb[b].box.unify(intersect(c, f.box));
                }
            }
        }
        for (uint32_t i = 0; i < kSpBins - 1; ++i) {
            Box b0, b1;
            float c0 = 0.0f, c1 = 0.0f;
            for (uint32_t j = 0; j <= i; ++j) { b0.unify(sb[j].box); c0 += sb[j].cIn; }
            for (uint32_t j = i + 1; j < kSpBins; ++j) { b1.unify(sb[j].box); c1 += sb[j].cOut; }
            float cost = travCost + (b0.area() * c0 + b1.area() * c1) / areaParent;
            if (cost < minCostSP) { minCostSP = cost; planeSP = i; }
        }
    }

    if (leafCost < minCostOP && leafCost < minCostSP) { makeLeaf(); return; }

    if (minCostOP < minCostSP) {
        float pivot = pcMin + (pcMax - pcMin) / kObjBins * (planeOP + 1);
        NodeFrags left, right;
        if (onDev) {
            m_gpu->partitionObject(nf.dev, axisOP, pivot, &left.dev, &right.dev);
            nf = NodeFrags();
            left = settle(std::move(left));
            right = settle(std::move(right));
            recurseDevice(std::move(left), std::move(right), depth, out, nodeIdx, parent, axisOP);
            return;
        } else {
            auto mid = std::partition(frags.begin(), frags.end(),
                                      [axisOP, pivot](const Frag& f) { return f.box.center(axisOP) < pivot; });
            size_t split = std::max<size_t>((size_t)(mid - frags.begin()), 1);
            if (split >= frags.size()) throw std::runtime_error("SBVH: invalid object partition");
            left.host.assign(frags.begin(), frags.begin() + split);
            right.host.assign(frags.begin() + split, frags.end());
            std::vector<Frag>().swap(frags);
        }
        recurse(std::move(left), std::move(right), depth, out, nodeIdx, parent, axisOP);
        return;
    }

    const float splitPos = (planeSP + 1) * binWidth + pMin;
    NodeFrags left, right;
    if (onDev) {
        m_gpu->partitionSpatial(nf.dev, axisSP, pMin, pMax, planeSP, splitPos, &left.dev, &right.dev);
        nf = NodeFrags();
        left = settle(std::move(left));
        right = settle(std::move(right));
        recurseDevice(std::move(left), std::move(right), depth, out, nodeIdx, parent, axisSP);
        return;
    } else {
        uint32_t nLeft = 0, nRight = 0;
        for (uint32_t j = 0; j <= planeSP; ++j) nLeft += sb[j].nIn;
        for (uint32_t j = planeSP + 1; j < kSpBins; ++j) nRight += sb[j].nOut;
        left.host.reserve(nLeft);
        right.host.reserve(nRight);
        for (const Frag& f : frags) {
            uint32_t eb = (uint32_t)(kSpBins * ((f.box.mn[axisSP] - pMin) / (pMax - pMin)));
            uint32_t xb = (uint32_t)(kSpBins * ((f.box.mx[axisSP] - pMin) / (pMax - pMin)));
            eb = std::min(eb, kSpBins - 1);
            xb = std::min(xb, kSpBins - 1);
            if (xb <= planeSP) {
                left.host.push_back(f);
            } else if (eb > planeSP) {
                right.host.push_back(f);
            } else {
                Box l, r;
                split(f.prim, axisSP, splitPos, &l, &r);
                if (l.valid()) left.host.push_back(Frag{f.prim, intersect(l, f.box)});
                if (r.valid()) right.host.push_back(Frag{f.prim, intersect(r, f.box)});
            }
        }
        if (left.host.empty() || right.host.empty()) throw std::runtime_error("SBVH: invalid spatial partition");
        std::vector<Frag>().swap(frags);
    }
    recurse(std::move(left), std::move(right), depth, out, nodeIdx, parent, axisSP);
}

void Builder::finishDeviceBuild(Subtree* skeleton, int threads) {
    // largest subtrees first, so the threads finish together
    std::vector<size_t> order(m_pending.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::sort(order.begin(), order.end(),
              [&](size_t a, size_t b) { return m_pending[a]->frags.size() > m_pending[b]->frags.size(); });
    std::atomic<size_t> next(0);
    auto worker = [&]() {
        for (;;) {
            const size_t k = next.fetch_add(1);
            if (k >= order.size()) return;
            Pending& p = *m_pending[order[k]];
            build(NodeFrags(std::move(p.frags)), p.depth, &p.tree);
        }
    };
    std::vector<std::future<void>> futs;
    for (int t = 1; t < std::max(1, threads); ++t) futs.push_back(std::async(std::launch::async, worker));
    worker();
    for (auto& f : futs) f.get();
    // DFS over the skeleton, each placeholder replaced by its subtree
    Subtree out;
    out.nodes.reserve(skeleton->nodes.size() + 2 * skeleton->refs.size());
    out.depth = skeleton->depth;
    out.fragments = skeleton->fragments;
    // iterative pre-order emit: (skeleton node, parent slot to patch, which child)
    struct Item { uint32_t sk; int32_t parent; int side; };
    std::vector<Item> st;
    st.push_back(Item{0, -1, 0});
    while (!st.empty()) {
        const Item it = st.back();
        st.pop_back();
        const SbvhNode& n = skeleton->nodes[it.sk];
        uint32_t idx = (uint32_t)out.nodes.size();
        if (n.count == kPlaceholder) {
            append(&out, m_pending[n.first]->tree);
        } else {
            SbvhNode m = n;
            if (m.count > 0) {
                m.first = (uint32_t)out.refs.size();
                out.refs.insert(out.refs.end(), skeleton->refs.begin() + n.first,
                                skeleton->refs.begin() + n.first + n.count);
            }
            out.nodes.push_back(m);
            if (m.count == 0) {
                // right child below left on the stack: left is emitted first
                st.push_back(Item{n.c1, (int32_t)idx, 1});
                st.push_back(Item{n.c0, (int32_t)idx, 0});
            }
        }
        if (it.parent >= 0) {
            if (it.side == 0) out.nodes[it.parent].c0 = idx;
            else out.nodes[it.parent].c1 = idx;
        }
    }
    *skeleton = std::move(out);
}

// SBVH::calcSAHCost (SBVH.h:351-377); primCost(ref) is the referenced
// primitive's intersection cost.
template <typename PrimCost>
float sbvhCost(const std::vector<SbvhNode>& nodes, const std::vector<uint32_t>& refs, PrimCost&& primCost) {
    const float Ci = 1.2f, Cl = 0.0f;
    float costInt = 0.0f, costLeaf = 0.0f, costObj = 0.0f;
    for (const SbvhNode& n : nodes) {
        float a = n.box.area();
        if (n.count == 0) {
            costInt += a;
        } else {
            costLeaf += a;
            float prims = 0.0f;
            for (uint32_t j = 0; j < n.count; ++j) prims += primCost(refs[n.first + j]);
            costObj += a * prims;
        }
    }
    float rootA = nodes[0].box.area();
    costInt *= Ci / rootA;
    costLeaf *= Cl / rootA;
    costObj /= rootA;
    return costInt + costLeaf + costObj;
}

// QBVH::collapseBBVH (QBVH.h:84-201).
class Collapser {
public:
    Collapser(const Accel& a, Accel* out) : m_in(a), m_out(out) {}

    struct Res { bool leaf; uint32_t idx; uint32_t count; };

    Res collapse(uint32_t gp, uint32_t depth) {
        const SbvhNode& root = m_in.sbvh[gp];
        if (root.count > 0) {
            uint32_t base = (uint32_t)m_out->qbvhRefs.size();
            for (uint32_t i = 0; i < root.count; ++i) m_out->qbvhRefs.push_back(m_in.sbvhRefs[root.first + i]);
            m_out->maxLeafCount = std::max(m_out->maxLeafCount, root.count);
            return Res{true, base, root.count};
        }
        if (++depth > m_out->qbvhDepth) m_out->qbvhDepth = depth;
        const uint32_t ci[2] = {root.c0, root.c1};
        const SbvhNode* ch[2] = {&m_in.sbvh[ci[0]], &m_in.sbvh[ci[1]]};

        uint32_t nodeIdx = (uint32_t)m_out->qbvh.size();
        m_out->qbvh.emplace_back();
        {
            QbvhNode& n = m_out->qbvh.back();
            n.leftAxis = n.rightAxis = 0;
            n.topAxis = root.axis;
            for (int c = 0; c < 4; ++c) {
                n.childValid[c] = false; n.isLeaf[c] = false; n.index[c] = 0; n.count[c] = 0;
                for (int a = 0; a < 3; ++a) { n.bmin[a][c] = INFINITY; n.bmax[a][c] = -INFINITY; }
            }
        }
        // slot pairs (0,1) come from the left child, (2,3) from the right one;
        // a leaf child occupies the first slot of its pair.
        uint32_t slotNode[4] = {UINT32_MAX, UINT32_MAX, UINT32_MAX, UINT32_MAX};
        for (int side = 0; side < 2; ++side) {
            const SbvhNode* c = ch[side];
            if (c->count == 0) {
                slotNode[2 * side + 0] = c->c0;
                slotNode[2 * side + 1] = c->c1;
                if (side == 0) m_out->qbvh[nodeIdx].leftAxis = c->axis;
                else m_out->qbvh[nodeIdx].rightAxis = c->axis;
            } else {
                slotNode[2 * side + 0] = ci[side];
            }
        }
        for (int s = 0; s < 4; ++s) {
            if (slotNode[s] == UINT32_MAX) continue;
            const Box& b = m_in.sbvh[slotNode[s]].box;
            QbvhNode& n = m_out->qbvh[nodeIdx];
            for (int a = 0; a < 3; ++a) { n.bmin[a][s] = b.mn[a]; n.bmax[a][s] = b.mx[a]; }
        }
        for (int s = 0; s < 4; ++s) {
            if (slotNode[s] == UINT32_MAX) continue;
            Res r = collapse(slotNode[s], depth);
            QbvhNode& n = m_out->qbvh[nodeIdx];
            n.childValid[s] = true;
            n.isLeaf[s] = r.leaf;
            n.index[s] = r.idx;
            n.count[s] = r.count;
        }
        return Res{false, nodeIdx, 0};
    }

private:
    const Accel& m_in;
    Accel* m_out;
};

// QBVH::calcSAHCost (QBVH.h:203-249).
template <typename PrimCost>
float qbvhCost(const Accel& a, PrimCost&& primCost) {
    const float Ci = 1.2f;
    float costInt = 0.0f, costObj = 0.0f;
    for (const QbvhNode& n : a.qbvh) {
        Box nodeBox;
        Box cb[4];
        for (int c = 0; c < 4; ++c) {
            if (!n.childValid[c]) continue;
            for (int k = 0; k < 3; ++k) { cb[c].mn[k] = n.bmin[k][c]; cb[c].mx[k] = n.bmax[k][c]; }
            nodeBox.unify(cb[c]);
        }
        costInt += nodeBox.area();
        for (int c = 0; c < 4; ++c) {
            if (!n.childValid[c] || !n.isLeaf[c]) continue;
            float prims = 0.0f;
            for (uint32_t j = 0; j < n.count[c]; ++j) prims += primCost(a.qbvhRefs[n.index[c] + j]);
            costObj += cb[c].area() * prims;
        }
    }
    float rootA = a.bounds.area();
    costInt *= Ci / rootA;
    costObj /= rootA;
    return costInt + costObj;
}

}  // namespace

void buildAccel(const std::vector<float>& triPos, Accel* out, int numThreads, bool device) {
    buildAccel(triPos, std::vector<Box>(), std::vector<float>(), out, numThreads, device);
}

void buildAccel(const std::vector<float>& triPos, const std::vector<Box>& boxes, const std::vector<float>& boxCosts,
                Accel* out, int numThreads, bool device) {
    auto t0 = std::chrono::steady_clock::now();
    const size_t nt = triPos.size() / 9, np = nt + boxes.size();
    std::vector<Tri> tris(nt);
    for (size_t i = 0; i < nt; ++i)
        for (int k = 0; k < 3; ++k)
            tris[i].p[k] = v3(triPos[9 * i + 3 * k], triPos[9 * i + 3 * k + 1], triPos[9 * i + 3 * k + 2]);

    *out = Accel();
    if (np == 0) {
        // SBVH(objs) with no objects: one empty leaf (SBVH.h:386-393).
        out->sbvh.emplace_back();
        out->sbvhDepth = 1;
        out->qbvh.emplace_back();
        QbvhNode& q = out->qbvh.back();
        for (int c = 0; c < 4; ++c) {
            q.childValid[c] = false; q.isLeaf[c] = false; q.index[c] = 0; q.count[c] = 0;
            for (int a = 0; a < 3; ++a) { q.bmin[a][c] = INFINITY; q.bmax[a][c] = -INFINITY; }
        }
        return;
    }

    std::vector<Frag> frags(np);
    for (size_t i = 0; i < np; ++i) {
        frags[i].prim = (uint32_t)i;
        frags[i].box = i < nt ? triBounds(tris[i]) : boxes[i - nt];
        out->bounds.unify(frags[i].box);
    }
    int parDepth = 0;
    for (int t = std::max(1, numThreads); t > 1; t >>= 1) ++parDepth;
    // the device takes the top of a large triangle tree: its root fragments
    // are the same triangle bounds, made there
    std::unique_ptr<SbvhGpu> gpu;
    NodeFrags rootFrags;
    if (device && boxes.empty() && np >= kDeviceNodeFrags) {
        gpu = SbvhGpu::create(triPos);
        rootFrags.dev = gpu->root();
    } else {
        rootFrags.host = std::move(frags);
    }
    Builder b(tris, boxes, boxCosts, out->bounds.area(), parDepth + 1, gpu.get());
    auto primCost = [&](uint32_t prim) { return prim < nt ? 1.0f : boxCosts[prim - nt]; };
    Subtree st;
    st.nodes.reserve(2 * np);
    b.build(std::move(rootFrags), 0, &st);
    if (gpu) b.finishDeviceBuild(&st, numThreads);
    out->deviceNodes = b.deviceNodes();
    out->sbvh = std::move(st.nodes);
    out->sbvhRefs = std::move(st.refs);
    out->sbvhDepth = st.depth;
    out->numFragments = st.fragments;
    out->sbvhCost = sbvhCost(out->sbvh, out->sbvhRefs, primCost);

    Collapser c(*out, out);
    Collapser::Res r = c.collapse(0, 0);
    if (r.leaf) {
        // The root itself is a leaf (QBVH.h:261-277).
        out->qbvh.emplace_back();
        QbvhNode& q = out->qbvh.back();
        for (int k = 0; k < 4; ++k) {
            q.childValid[k] = false; q.isLeaf[k] = false; q.index[k] = 0; q.count[k] = 0;
            for (int a = 0; a < 3; ++a) { q.bmin[a][k] = INFINITY; q.bmax[a][k] = -INFINITY; }
        }
        for (int a = 0; a < 3; ++a) { q.bmin[a][0] = out->sbvh[0].box.mn[a]; q.bmax[a][0] = out->sbvh[0].box.mx[a]; }
        q.childValid[0] = true; q.isLeaf[0] = true; q.index[0] = r.idx; q.count[0] = r.count;
        q.topAxis = q.leftAxis = q.rightAxis = 0;
    }
    out->qbvhCost = qbvhCost(*out, primCost);
    out->buildSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace slrh
