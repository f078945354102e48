// Device half of the SBVH build (sbvh_gpu.hip): the fragments of a large
// node in device memory, and the node's bounds, bins and partitions computed
// there with the host builder's results (bvh_build.cpp drives it).
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "host_scene.h"

namespace slrh {

struct DevFrags;   // one node's fragments in device memory

class SbvhGpu {
public:
    // Uploads the triangles (9 floats each); throws on a device error.
    static std::unique_ptr<SbvhGpu> create(const std::vector<float>& triPos);
    ~SbvhGpu();

    DevFrags* root();   // one fragment per triangle: its bounds
    static void release(DevFrags* d);
    static uint32_t count(const DevFrags* d);
    // the node's box and the box of its fragments' centroids
    void bounds(DevFrags* d, Box* parent, Box* centroids);
    // 32 centroid bins over [lo, hi] on axis: counts and boxes
    void objectBins(DevFrags* d, int axis, float lo, float hi, uint32_t* n, Box* box);
    // 16 spatial bins of width binWidth from lo: entering / leaving counts,
    // boxes of the fragments clipped to each bin
    void spatialBins(DevFrags* d, int axis, float lo, float hi, float binWidth, uint32_t* nIn, uint32_t* nOut,
                     Box* box);
    // object split at pivot in std::partition's order; d is left as it was
    void partitionObject(DevFrags* d, int axis, float pivot, DevFrags** left, DevFrags** right);
    // spatial split after bin `plane` at splitPos; straddling triangles split
    void partitionSpatial(DevFrags* d, int axis, float lo, float hi, uint32_t plane, float splitPos, DevFrags** left,
                          DevFrags** right);
    // the fragments in order (primitive ids and boxes)
    void download(DevFrags* d, std::vector<uint32_t>* prims, std::vector<Box>* boxes);

private:
    SbvhGpu();
    struct Impl;
    std::unique_ptr<Impl> m;
};

}  // namespace slrh
