// Host-side acceleration structure construction for libcpt.
//
//  * build_bvh4: binned-SAH binary BVH collapsed to a 4-wide BVH whose nodes are
//    one 128-byte cache line each (child boxes in SoA order), and a triangle
//    array reordered into leaf order (48 bytes per triangle: a, b-a, c-a plus the
//    triangle id and its tie rank).  This replaces the reference's binary
//    5-bin BVH (src/primitive/bvh.h:151-217) for traversal.
//  * reference_tie_rank: the position at which the reference's own BVH
//    (bvh.h:57-217) tests each triangle.  The reference keeps the LAST tested
//    triangle among those reaching the minimal t (bvh.h:33-34,
//    simple_group.h:24-27, triangle.h:55), so the GPU breaks exact-t ties by
//    the highest rank.
#pragma once
#include <cstdint>
#include <vector>

namespace cpt {

struct alignas(16) Node4 {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    int32_t child[4];  // >= 0 internal node; leaf: LEAF_BIT | (count-1) << 27 | first; EMPTY_CHILD
    int32_t pad[4];
};
static_assert(sizeof(Node4) == 128, "Node4 must be one 128-byte line");

constexpr int32_t LEAF_BIT = (int32_t)0x80000000u;
constexpr int32_t EMPTY_CHILD = 0x7fffffff;
constexpr int LEAF_MAX = 4;

struct TriRec {  // 48 bytes, leaf order
    float ax, ay, az;
    int32_t id;
    float e1x, e1y, e1z;
    int32_t rank;
    float e2x, e2y, e2z;
    int32_t pad;
};
static_assert(sizeof(TriRec) == 48, "TriRec is three float4");

struct Bvh4 {
    std::vector<Node4> nodes;
    std::vector<TriRec> tris;
    int64_t leaves = 0;
    int max_depth = 0;
};

// positions: nv x 3, indices: nt x 3, rank: nt (may be empty -> rank = -id)
void build_bvh4(const float* positions, const int32_t* indices, int nt, const std::vector<int32_t>& rank,
                Bvh4& out);

// Test order of every triangle in the reference BVH (bvh.h:57-217 restated on
// index arrays; identical partitions, identical order).
void reference_tie_rank(const float* positions, const int32_t* indices, int nt, std::vector<int32_t>& rank);

}  // namespace cpt
