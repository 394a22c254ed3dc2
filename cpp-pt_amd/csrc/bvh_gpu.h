// Device-side BVH construction for libcpt (the GPU counterpart of build_bvh4).
//
// PLOC, parallel locally-ordered clustering (Meister and Bittner, "Parallel
// Locally-Ordered Clustering for Bounding Volume Hierarchy Construction", IEEE
// TVCG 2018): triangles sorted along a 63-bit Morton curve are merged bottom-up,
// each round joining every pair of clusters that are each other's nearest
// neighbour (surface area of the merged box) within a window of +-radius
// positions.  The binary tree is then collapsed level by level into the same
// 4-wide layout build_bvh4 produces (128-byte nodes with widened child boxes,
// 48-byte triangle records in leaf order), so traversal is unchanged.  It
// replaces the reference's BVH::build (src/primitive/bvh.h:151-217) when the
// scene is too large for a host build to be quick.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "bvh_build.h"

namespace cpt {

struct GpuBvh {
    Node4* nodes = nullptr;  // device, n_nodes entries (caller frees)
    TriRec* tris = nullptr;  // device, nt entries in leaf order (caller frees)
    int64_t n_nodes = 0, leaves = 0;
    int max_depth = 0;
    int rounds = 0;          // clustering rounds
    double ms = 0.0;         // wall time of the build (device work included)
};

// vpos: per vertex xyz (w unused); tri: per triangle v0, v1, v2 (w unused);
// rank: per triangle tie rank on the device, or null (rank = -id).
// leaf_max: largest leaf (1..16); a subtree that small becomes one leaf when the
// surface-area cost says so;
// radius: neighbour window of the clustering (1..32).  nt >= 1.
hipError_t build_bvh4_gpu(const float4* vpos, const int4* tri, const int32_t* rank, int nt, int leaf_max,
                          int radius, hipStream_t stream, GpuBvh* out, std::string* err);

}  // namespace cpt
