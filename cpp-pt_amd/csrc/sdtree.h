// SD-tree of the reference's practical path guiding (src/algorithms/sdtree.h),
// host side: the spatial binary tree and the directional quadtrees are built
// and refined on the host between iterations and flattened into plain arrays
// for the device, which samples them, evaluates their pdf and splats into them.
#pragma once
#include <cstdint>
#include <vector>

namespace cpt {

// quadtree over (phi in [-pi, pi], cos(theta) in [-1, 1]) (sdtree.h:120-233)
struct SdQuad {
    int32_t child = -1;  // >= 0: four children at child .. child+3 (quadrant order of sdtree.h:306-326)
    float value = 0.0f;  // leaf: radiance splatted this iteration
};

struct SdRect {  // DTree::Leaf (sdtree.h:124-128): begin (phi, cos theta), size
    float bx, by, sx, sy;
};

struct SdDTree {
    std::vector<SdQuad> q;       // q[0] is the root
    std::vector<SdRect> leaves;  // sampling leaves, in the order adapt_and_build_leaves emits them
    int64_t count = 0;           // splats this iteration
    float radiance = 0.0f;       // their sum
    SdDTree();
    // DTree::adapt_and_build_leaves (sdtree.h:172-233): prune below 1% of the
    // total, split leaves above it, rebuild the sampling leaves, clear the sums
    void adapt();
    float prune(int n, float threshold);
    void compact();
};

struct SdSNode {  // spatial binary tree, split at the midpoint, axis = depth % 3
    int32_t child = -1;  // >= 0: children at child, child+1 (lower half first)
    int32_t dtree = -1;  // leaf: its directional tree
};

// Flattened image for the device.
struct SdFlat {
    std::vector<int32_t> snode;  // >= 0: children at v, v+1; < 0: leaf of dtree -1-v
    std::vector<int32_t> dinfo;  // per dtree: quad base, leaf base, leaf count, pad
    std::vector<int32_t> quad;   // per quad node: >= 0 absolute index of the first child; < 0: accumulator -1-v
    std::vector<float> rect;     // 4 floats per sampling leaf
    std::vector<int32_t> slot_dtree, slot_quad;  // accumulator -> (dtree, quad node)
};

class SdTree {
public:
    SdTree(const float bmin[3], const float bmax[3]);
    // SDTree::update (sdtree.h:454-457): subdivide(c, iteration), iteration + 1
    void update();
    void flatten(SdFlat& f) const;
    // add the device's per-accumulator splat sums and counts (layout of the
    // last flatten) to the quadtree leaves and the dtree totals
    void add_splats(const SdFlat& f, const float* sums, const uint32_t* counts);

    std::vector<SdSNode> s;  // s[0] is the root
    std::vector<SdDTree> d;
    float bmin[3], bmax[3];
    int iteration = 1;
    int c = 12000;

private:
    void subdivide(int c, int iteration);
};

}  // namespace cpt
