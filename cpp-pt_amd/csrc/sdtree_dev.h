// SD-tree on the device: spatial lookup, directional sampling, pdf and splats of
// the reference's practical path guiding (src/algorithms/sdtree.h:268-538),
// over the flattened arrays the host builds each iteration (sdtree.cpp).
#pragma once
#include "pt_device.h"

namespace cpt {

struct SdDev {
    const int32_t* snode;  // >= 0: children at v, v+1; < 0: dtree -1-v
    const int4* dinfo;     // quad base, leaf base, leaf count, pad
    const int32_t* quad;   // >= 0: first child; < 0: accumulator -1-v
    const float4* rect;    // sampling leaves (begin phi, begin cos, size phi, size cos)
    float* sum;            // per accumulator: splatted radiance
    uint32_t* cnt;         // per accumulator: splats
    float bmin[3], bmax[3];
};

// SDTree::get_leaf(Vec3) (sdtree.h:345-396): midpoint splits, axis = depth % 3,
// `p > mid` goes to the upper child
__device__ __forceinline__ int sd_dtree(const SdDev& t, V3 p) {
    float lo[3] = {t.bmin[0], t.bmin[1], t.bmin[2]}, hi[3] = {t.bmax[0], t.bmax[1], t.bmax[2]};
    const float pc[3] = {p.x, p.y, p.z};
    int n = 0, depth = 0;
    while (true) {
        const int v = t.snode[n];
        if (v < 0) return -1 - v;
        const int ax = depth % 3;
        const float mid = (lo[ax] + hi[ax]) * 0.5f;
        if (pc[ax] > mid) {
            lo[ax] = mid;
            n = v + 1;
        } else {
            hi[ax] = mid;
            n = v;
        }
        depth++;
    }
}

// DTree descent to the quadtree leaf holding (x = phi, y = cos theta)
// (get_leaf / get_area, sdtree.h:268-343); returns the accumulator, and the
// leaf's extent from the descent's own bounds
__device__ __forceinline__ int sd_quad_leaf(const SdDev& t, int dt, float x, float y, float& ax, float& ay) {
    float mnx = (float)-CPT_M_PI, mny = -1.0f, mxx = (float)CPT_M_PI, mxy = 1.0f;
    int n = t.dinfo[dt].x;
    while (true) {
        const float midx = (mnx + mxx) * 0.5f, midy = (mny + mxy) * 0.5f;
        const int v = t.quad[n];
        if (v < 0) {
            ax = mxx - mnx;
            ay = mxy - mny;
            return -1 - v;
        }
        int child;
        if (x < midx) {
            if (y < midy) {
                child = 0;
                mxx = midx;
                mxy = midy;
            } else {
                child = 2;
                mxx = midx;
                mny = midy;
            }
        } else {
            if (y < midy) {
                child = 1;
                mnx = midx;
                mxy = midy;
            } else {
                child = 3;
                mnx = midx;
                mny = midy;
            }
        }
        n = v + child;
    }
}

// SDTree::pdf (sdtree.h:525-538): 1 / (number of leaves * area of the leaf
// holding the direction), in (phi, cos theta) = solid angle measure
__device__ __forceinline__ float sd_pdf(const SdDev& t, V3 p, V3 w) {
    const int dt = sd_dtree(t, p);
    const float phi = atan2f(w.y, w.x);
    float ax, ay;
    sd_quad_leaf(t, dt, phi, w.z, ax, ay);
    return (float)((1.0 / (double)t.dinfo[dt].z) * (1.0 / (double)(ax * ay)));
}

// SDTree::sample_direction (sdtree.h:400-423): a leaf uniformly, a point
// uniformly in it
__device__ __forceinline__ V3 sd_sample(const SdDev& t, V3 p, uint64_t& rng, float& pdf) {
    const int dt = sd_dtree(t, p);
    const int4 di = t.dinfo[dt];
    const int sz = di.z;
    const int i = (int)(pcg_float(rng) * (float)sz);
    const float u = pcg_float(rng);
    const float v = pcg_float(rng);
    const float4 r = t.rect[di.y + i % sz];
    const float area = r.z * r.w;
    const float phi = r.x + u * r.z;
    const float ct = r.y + v * r.w;
    const float st = (float)sqrt(1.0 - (double)(ct * ct));
    pdf = (float)(1.0 / (double)((float)sz * area));
    return {cosf(phi) * st, sinf(phi) * st, ct};
}

// SDTree::splat (sdtree.h:514-524)
__device__ __forceinline__ void sd_splat(const SdDev& t, float radiance, V3 p, V3 w) {
    const int dt = sd_dtree(t, p);
    const float phi = atan2f(w.y, w.x);
    float ax, ay;
    const int slot = sd_quad_leaf(t, dt, phi, w.z, ax, ay);
    unsafeAtomicAdd(&t.sum[slot], radiance);
    atomicAdd(&t.cnt[slot], 1u);
}

}  // namespace cpt
