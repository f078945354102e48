// Binned-SAH binary BVH over triangles: the accelerator the reference
// builds as StandardBVH(objs, Partitioning::BinnedSAH)
// (libSLR/Accelerator/StandardBVH.h:62-193), node for node.
//
// Each decision follows the reference so that the tree is the same: the
// centroid box's widest axis, 16 bins over the centroid extent of that axis
// (bin = uint32(16 * (c - min) / (max - min)), clamped), the SAH cost
// 1.2 + (A0 * C0 + A1 * C1) / A for the 15 planes between bins with the
// first minimum kept, a leaf when the best cost is not below the leaf's
// cost (the sum of the triangles' intersection costs, 1 each), and a median
// split by std::nth_element when the centroids do not spread along the
// widest axis. Splitting reorders the index range with std::partition, and
// the median with std::nth_element: the same standard-library algorithms on
// the same input order give the same order, which fixes the leaf order.
// Nodes are stored in creation (pre-)order, as m_nodes is.
#include <algorithm>
#include <cmath>
#include <vector>

#include "host_scene.h"

namespace slrh {

namespace {

struct StdBuilder {
    const std::vector<Box>& boxes;
    const std::vector<V3>& centroids;
    std::vector<uint32_t>& idx;
    StdBvh* out;

    static float coord(const V3& p, int a) { return a == 0 ? p.x : a == 1 ? p.y : p.z; }

    uint32_t build(uint32_t start, uint32_t end, uint32_t depth) {
        const uint32_t node = (uint32_t)out->nodes.size();
        out->nodes.emplace_back();
        if (++depth > out->depth) out->depth = depth;
        Box bbox, cbox;
        for (uint32_t i = start; i < end; ++i) {
            bbox.unify(boxes[idx[i]]);
            cbox.unify(centroids[idx[i]]);
        }
        const int axis = cbox.widestAxis();
        const float cmin = cbox.mn[axis], cmax = cbox.mx[axis];
        const uint32_t n = end - start;
        auto leaf = [&]() {
            StdBvhNode& nd = out->nodes[node];
            nd.box = bbox;
            nd.first = (uint32_t)out->refs.size();
            nd.count = n;
            for (uint32_t i = start; i < end; ++i) out->refs.push_back(idx[i]);
            return node;
        };
        if (n == 1) return leaf();
        uint32_t split;
        if ((cmax - cmin) <= 0) {
            // no centroid spread: equal halves by a median selection
            split = (start + end) / 2;
            std::nth_element(idx.begin() + start, idx.begin() + split, idx.begin() + end,
                             [&](uint32_t a, uint32_t b) { return coord(centroids[a], axis) < coord(centroids[b], axis); });
        } else {
            constexpr uint32_t kBins = 16;
            const float travCost = 1.2f;
            Box binBox[kBins];
            float binCost[kBins] = {};
            float leafCost = 0.0f;
            for (uint32_t i = start; i < end; ++i) {
                const uint32_t o = idx[i];
                const float c = 1.0f;   // a triangle's intersection cost
                leafCost += c;
                uint32_t bin = (uint32_t)(kBins * ((coord(centroids[o], axis) - cmin) / (cmax - cmin)));
                bin = std::min(bin, kBins - 1);
                binCost[bin] += c;
                binBox[bin].unify(boxes[o]);
            }
            uint32_t plane = 0;
            float best = INFINITY;
            const float parentArea = bbox.area();
            for (uint32_t i = 0; i < kBins - 1; ++i) {
                Box b0, b1;
                float c0 = 0.0f, c1 = 0.0f;
                for (uint32_t j = 0; j <= i; ++j) {
                    b0.unify(binBox[j]);
                    c0 += binCost[j];
                }
                for (uint32_t j = i + 1; j < kBins; ++j) {
                    b1.unify(binBox[j]);
                    c1 += binCost[j];
                }
                const float cost = travCost + (b0.area() * c0 + b1.area() * c1) / parentArea;
                if (cost < best) {
                    best = cost;
                    plane = i;
                }
            }
            if (!(best < leafCost)) return leaf();
            const float pivot = cbox.mn[axis] + (cmax - cmin) / kBins * (plane + 1);
            auto mid = std::partition(idx.begin() + start, idx.begin() + end,
                                      [&](uint32_t o) { return coord(centroids[o], axis) < pivot; });
            split = std::max((uint32_t)std::distance(idx.begin() + start, mid), 1u) + start;
        }
        const uint32_t c0 = build(start, split, depth);
        const uint32_t c1 = build(split, end, depth);
        StdBvhNode& nd = out->nodes[node];
        nd.box = bbox;
        nd.c0 = c0;
        nd.c1 = c1;
        nd.axis = (uint8_t)axis;
        nd.count = 0;
        return node;
    }
};

// StandardBVH::calcSAHCost (StandardBVH.h:175-199): internal areas times 1.2,
// leaf areas times their triangle count, over the root's area.
float sahCost(const StdBvh& b) {
    float costInt = 0.0f, costObj = 0.0f, costLeaf = 0.0f;
    for (const StdBvhNode& n : b.nodes) {
        const float sa = n.box.area();
        if (n.count == 0) {
            costInt += sa;
        } else {
            costLeaf += sa;
            float prims = 0.0f;
            for (uint32_t j = 0; j < n.count; ++j) prims += 1.0f;
            costObj += sa * prims;
        }
    }
    const float rootSA = b.nodes[0].box.area();
    costInt *= 1.2f / rootSA;
    costLeaf *= 0.0f / rootSA;
    costObj /= rootSA;
    return costInt + costLeaf + costObj;
}

}  // namespace

void buildStandardBvh(const std::vector<float>& triPos, StdBvh* out) {
    out->nodes.clear();
    out->refs.clear();
    out->depth = 0;
    out->cost = 0.0f;
    const size_t nt = triPos.size() / 9;
    if (nt == 0) {
        // an empty leaf, as the reference's constructor makes for no objects
        out->nodes.emplace_back();
        out->depth = 1;
        return;
    }
    std::vector<Box> boxes(nt);
    std::vector<V3> centroids(nt);
    std::vector<uint32_t> idx(nt);
    for (size_t i = 0; i < nt; ++i) {
        const float* p = &triPos[9 * i];
        Box b = Box::point(v3(p[0], p[1], p[2]));
        b.unify(v3(p[3], p[4], p[5]));
        b.unify(v3(p[6], p[7], p[8]));
        boxes[i] = b;
        centroids[i] = b.centroid();
        idx[i] = (uint32_t)i;
    }
    StdBuilder sb{boxes, centroids, idx, out};
    sb.build(0, (uint32_t)nt, 0);
    out->cost = sahCost(*out);
}

}  // namespace slrh
