// Device-side BVH construction: PLOC clustering, then a level-by-level collapse
// into the 4-wide node layout of build_bvh4 (bvh_build.cpp).
//
// Buffers (binary tree): node ids 0..nt-1 are the triangles, nt.. the merges.
// Per node: box lo/hi (float4), triangle count, children.  Clusters are node ids
// kept in Morton order; a round computes every cluster's nearest neighbour in a
// +-radius window (boxes staged in LDS), marks mutual pairs, and one exclusive
// scan over packed (merge, keep) flags gives both the new node ids and the
// compacted cluster positions, so the tree does not depend on thread timing.
#include "bvh_gpu.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

namespace cpt {
namespace {

constexpr int BLK = 256;
constexpr int MAX_RADIUS = 32;
constexpr int MAX_LEAF = 16;

// float <-> unsigned with the same order, for min/max through integer atomics
__device__ __forceinline__ unsigned ord(float f) {
    unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord(unsigned u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__device__ __forceinline__ float half_area(float4 lo, float4 hi) {  // Box::half_area
    float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    if (dx < 0.0f) return 0.0f;
    return dx * dy + dy * dz + dz * dx;
}

// Per triangle: its box (the leaf node), centroid, and block-reduced bounds:
// bounds[0..2] centroid min, [3..5] centroid max, [6] max |coordinate| (ordered).
// Grid-stride over the triangles with a capped grid, so that the seven
// same-address atomics per block stay few.
__global__ void __launch_bounds__(BLK) k_leaves(const float4* __restrict__ vpos, const int4* __restrict__ tri, int nt,
                                                float4* lo, float4* hi, int* cnt, int2* child, float4* cen,
                                                unsigned* bounds) {
    using BR = hipcub::BlockReduce<unsigned, BLK>;
    __shared__ typename BR::TempStorage tmp;
    unsigned mn[3] = {~0u, ~0u, ~0u}, mx[3] = {0u, 0u, 0u}, sm = 0u;
    for (int t = blockIdx.x * BLK + threadIdx.x; t < nt; t += gridDim.x * BLK) {
        const int4 v = tri[t];
        const float4 a = vpos[v.x], b = vpos[v.y], c = vpos[v.z];
        const float l[3] = {fminf(fminf(a.x, b.x), c.x), fminf(fminf(a.y, b.y), c.y), fminf(fminf(a.z, b.z), c.z)};
        const float h[3] = {fmaxf(fmaxf(a.x, b.x), c.x), fmaxf(fmaxf(a.y, b.y), c.y), fmaxf(fmaxf(a.z, b.z), c.z)};
        lo[t] = make_float4(l[0], l[1], l[2], 0.0f);
        hi[t] = make_float4(h[0], h[1], h[2], 0.0f);
        cnt[t] = 1;
        child[t] = make_int2(-1, -1);
        float cc[3], s = 0.0f;
        for (int k = 0; k < 3; k++) {
            cc[k] = 0.5f * (l[k] + h[k]);
            s = fmaxf(s, fmaxf(fabsf(l[k]), fabsf(h[k])));
        }
        cen[t] = make_float4(cc[0], cc[1], cc[2], 0.0f);
        for (int k = 0; k < 3; k++) {
            mn[k] = min(mn[k], ord(cc[k]));
            mx[k] = max(mx[k], ord(cc[k]));
        }
        sm = max(sm, ord(s));
    }
    for (int k = 0; k < 3; k++) {
        unsigned r = BR(tmp).Reduce(mn[k], hipcub::Min());
        __syncthreads();
        if (threadIdx.x == 0) atomicMin(&bounds[k], r);
        r = BR(tmp).Reduce(mx[k], hipcub::Max());
        __syncthreads();
        if (threadIdx.x == 0) atomicMax(&bounds[3 + k], r);
    }
    unsigned r = BR(tmp).Reduce(sm, hipcub::Max());
    if (threadIdx.x == 0) atomicMax(&bounds[6], r);
}

__device__ __forceinline__ uint64_t spread3(uint32_t x) {  // 21 bits -> every third bit
    uint64_t v = x & 0x1fffffu;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

__global__ void __launch_bounds__(BLK) k_morton(const float4* __restrict__ cen, int nt, const unsigned* __restrict__ bounds,
                                                uint64_t* keys, int* ids) {
    const int t = blockIdx.x * BLK + threadIdx.x;
    if (t >= nt) return;
    const float4 c = cen[t];
    const float cv[3] = {c.x, c.y, c.z};
    // one scale for all axes (the largest centroid extent), so that the curve
    // keeps the scene's proportions
    float ext = 0.0f;
    for (int k = 0; k < 3; k++) ext = fmaxf(ext, unord(bounds[3 + k]) - unord(bounds[k]));
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
        float f = ext > 0.0f ? (cv[k] - unord(bounds[k])) / ext * 2097152.0f : 0.0f;
        f = fminf(fmaxf(f, 0.0f), 2097151.0f);
        q[k] = (uint32_t)f;
    }
    keys[t] = spread3(q[0]) << 2 | spread3(q[1]) << 1 | spread3(q[2]);
    ids[t] = t;
}

// Nearest neighbour of every cluster among the `radius` clusters on each side:
// the one whose merged box has the least surface area.  Pairs are ordered by
// (area, distance in the curve order, parity of the lower position, lower
// position): a strict order that both ends of a pair see alike, so the globally
// best pair is always mutual and every round merges.  The parity term pairs
// (2k, 2k+1) when areas and distances tie (coincident triangles), which halves
// such runs per round instead of growing one cluster by one triangle a round.
__global__ void __launch_bounds__(BLK) k_nearest(const int* __restrict__ cl, int m, int radius,
                                                 const float4* __restrict__ lo, const float4* __restrict__ hi,
                                                 int* nn) {
    __shared__ float slo[3][BLK + 2 * MAX_RADIUS], shi[3][BLK + 2 * MAX_RADIUS];
    const int base = blockIdx.x * BLK - radius;
    for (int k = threadIdx.x; k < BLK + 2 * radius; k += BLK) {
        const int g = base + k;
        if (g >= 0 && g < m) {
            const int id = cl[g];
            const float4 a = lo[id], b = hi[id];
            slo[0][k] = a.x;
            slo[1][k] = a.y;
            slo[2][k] = a.z;
            shi[0][k] = b.x;
            shi[1][k] = b.y;
            shi[2][k] = b.z;
        }
    }
    __syncthreads();
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= m) return;
    const int li = threadIdx.x + radius;
    const float ax = slo[0][li], ay = slo[1][li], az = slo[2][li];
    const float bx = shi[0][li], by = shi[1][li], bz = shi[2][li];
    float best = INFINITY;
    int bj = -1, bd = 0, bmn = 0;
    const int j0 = max(i - radius, 0), j1 = min(i + radius, m - 1);
    for (int j = j0; j <= j1; j++) {
        if (j == i) continue;
        const int lj = li + (j - i);
        const float dx = fmaxf(bx, shi[0][lj]) - fminf(ax, slo[0][lj]);
        const float dy = fmaxf(by, shi[1][lj]) - fminf(ay, slo[1][lj]);
        const float dz = fmaxf(bz, shi[2][lj]) - fminf(az, slo[2][lj]);
        const float sa = dx * dy + dy * dz + dz * dx;
        const int d = j > i ? j - i : i - j, mn = min(i, j);
        bool better = sa < best;
        if (!better && sa == best && bj >= 0) {
            if (d != bd) better = d < bd;
            else if ((mn & 1) != (bmn & 1)) better = (mn & 1) < (bmn & 1);
            else better = mn < bmn;
        }
        if (better) {
            best = sa;
            bj = j;
            bd = d;
            bmn = mn;
        }
    }
    nn[i] = bj;
}

// flags: bit 0 = the cluster survives the round (unmerged, or the lower of a
// mutual pair, which becomes the merged cluster), bit 32 = it creates a node
__global__ void __launch_bounds__(BLK) k_pairs(const int* __restrict__ nn, int m, uint64_t* flags) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= m) return;
    const int j = nn[i];
    const bool mutual = j >= 0 && nn[j] == i;
    const uint64_t merge = (mutual && i < j) ? 1ull : 0ull;
    const uint64_t keep = (mutual && i > j) ? 0ull : 1ull;
    flags[i] = merge << 32 | keep;
}

__global__ void __launch_bounds__(BLK) k_merge(const int* __restrict__ cl, const int* __restrict__ nn, int m, int node_base,
                                               const uint64_t* __restrict__ flags, const uint64_t* __restrict__ scan,
                                               float4* lo, float4* hi, int* cnt, int2* child, int* cl_out,
                                               int* totals) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= m) return;
    const uint64_t f = flags[i], s = scan[i];
    if (i == m - 1) {
        totals[0] = (int)((s & 0xffffffffu) + (f & 0xffffffffu));  // clusters after the round
        totals[1] = (int)((s >> 32) + (f >> 32));                   // nodes created
    }
    if (!(f & 1u)) return;
    const int pos = (int)(s & 0xffffffffu);
    if (f >> 32) {
        const int id = node_base + (int)(s >> 32);
        const int a = cl[i], b = cl[nn[i]];
        const float4 la = lo[a], lb = lo[b], ha = hi[a], hb = hi[b];
        lo[id] = make_float4(fminf(la.x, lb.x), fminf(la.y, lb.y), fminf(la.z, lb.z), 0.0f);
        hi[id] = make_float4(fmaxf(ha.x, hb.x), fmaxf(ha.y, hb.y), fmaxf(ha.z, hb.z), 0.0f);
        cnt[id] = cnt[a] + cnt[b];
        child[id] = make_int2(a, b);
        cl_out[pos] = id;
    } else {
        cl_out[pos] = cl[i];
    }
}

// Leaf choice, the rule build_bvh4 applies: a subtree of at most leaf_max
// triangles becomes one leaf when testing all of its triangles costs no more
// than one box test plus its two children's triangles (half areas times counts).
__global__ void __launch_bounds__(BLK) k_leaf_choice(int nt, int leaf_max, const float4* __restrict__ lo,
                                                     const float4* __restrict__ hi, const int* __restrict__ cnt,
                                                     const int2* __restrict__ child, uint8_t* leaf) {
    const int k = blockIdx.x * BLK + threadIdx.x;
    if (k >= 2 * nt - 1) return;
    if (k < nt) {
        leaf[k] = 1;
        return;
    }
    const int c = cnt[k];
    uint8_t v = 0;
    if (c <= leaf_max) {
        const int2 ch = child[k];
        const float a = half_area(lo[k], hi[k]);
        const float split = a + half_area(lo[ch.x], hi[ch.x]) * cnt[ch.x] + half_area(lo[ch.y], hi[ch.y]) * cnt[ch.y];
        v = a * c <= split ? 1 : 0;
    }
    leaf[k] = v;
}

// One level of the collapse.  Each item (binary node, 4-wide node) gathers up
// to four children by opening the largest-area child that is not a leaf (the
// rule build_bvh4 uses); leaf children have their triangles written
// contiguously, the others become the next level's items.  Counters: ctr[0] nodes, ctr[1] triangles, ctr[2] leaves; one
// atomic per block and counter (a block scan hands out the offsets).
__global__ void __launch_bounds__(BLK) k_collapse(const int2* __restrict__ fin, int nin, int2* fout, int level_base,
                                                  int* ctr, Node4* nodes, TriRec* tris,
                                                  const float4* __restrict__ lo, const float4* __restrict__ hi,
                                                  const int* __restrict__ cnt, const int2* __restrict__ child, int nt,
                                                  const uint8_t* __restrict__ is_leaf, float smax,
                                                  const float4* __restrict__ vpos,
                                                  const int4* __restrict__ tri, const int32_t* __restrict__ rank) {
    using BS = hipcub::BlockScan<uint64_t, BLK>;
    __shared__ typename BS::TempStorage tmp;
    __shared__ int s_node, s_tri;
    const int w = blockIdx.x * BLK + threadIdx.x;
    int kids[4];
    int nk = 0, ni = 0, nl = 0, ntri = 0;
    int2 it = make_int2(0, 0);
    if (w < nin) {
        it = fin[w];
        const int b = it.x;
        if (is_leaf[b]) {
            kids[nk++] = b;  // the root is itself a leaf
        } else {
            const int2 ch = child[b];
            kids[nk++] = ch.x;
            kids[nk++] = ch.y;
            while (nk < 4) {
                int bi = -1;
                float ba = -1.0f;
                for (int k = 0; k < nk; k++) {
                    if (is_leaf[kids[k]]) continue;
                    const float a = half_area(lo[kids[k]], hi[kids[k]]);
                    if (a > ba) {
                        ba = a;
                        bi = k;
                    }
                }
                if (bi < 0) break;
                const int2 oc = child[kids[bi]];
                kids[bi] = oc.x;
                kids[nk++] = oc.y;
            }
        }
        for (int k = 0; k < nk; k++) {
            if (is_leaf[kids[k]]) {
                nl++;
                ntri += cnt[kids[k]];
            } else {
                ni++;
            }
        }
    }
    // packed exclusive scan: triangles (bits 0-31), internal children (32-47), leaves (48-63)
    const uint64_t mine = (uint64_t)ntri | (uint64_t)ni << 32 | (uint64_t)nl << 48;
    uint64_t off, total;
    BS(tmp).ExclusiveSum(mine, off, total);
    if (threadIdx.x == 0) {
        s_node = atomicAdd(&ctr[0], (int)((total >> 32) & 0xffffu));
        s_tri = atomicAdd(&ctr[1], (int)(total & 0xffffffffu));
        atomicAdd(&ctr[2], (int)(total >> 48));
    }
    __syncthreads();
    if (w >= nin) return;
    int next_node = s_node + (int)((off >> 32) & 0xffffu);
    int next_tri = s_tri + (int)(off & 0xffffffffu);
    Node4 node;
    for (int k = 0; k < 4; k++) {
        node.lox[k] = node.loy[k] = node.loz[k] = 1e30f;
        node.hix[k] = node.hiy[k] = node.hiz[k] = 1e30f;
        node.child[k] = EMPTY_CHILD;
        node.pad[k] = 0;
    }
    for (int k = 0; k < nk; k++) {
        const int kid = kids[k];
        const float4 l = lo[kid], h = hi[kid];
        // widened by a few ulps plus a scene-scale term, as build_bvh4 does: the
        // reference's triangle test accepts points slightly outside the triangle
        const float lv[3] = {l.x, l.y, l.z}, hv[3] = {h.x, h.y, h.z};
        float wl[3], wh[3];
        for (int a = 0; a < 3; a++) {
            const float mm = fmaxf(fabsf(lv[a]), fabsf(hv[a]));
            const float dl = mm * 0x1p-20f + smax * 0x1p-22f;
            wl[a] = lv[a] - dl;
            wh[a] = hv[a] + dl;
        }
        node.lox[k] = wl[0];
        node.loy[k] = wl[1];
        node.loz[k] = wl[2];
        node.hix[k] = wh[0];
        node.hiy[k] = wh[1];
        node.hiz[k] = wh[2];
        const int c = cnt[kid];
        if (is_leaf[kid]) {
            const int first = next_tri;
            next_tri += c;
            int stack[MAX_LEAF + 1];
            int sp = 0, out = first;
            stack[sp++] = kid;
            while (sp > 0) {
                const int n = stack[--sp];
                if (n < nt) {
                    const int4 v = tri[n];
                    const float4 a = vpos[v.x], bb = vpos[v.y], cc = vpos[v.z];
                    TriRec r;
                    r.ax = a.x;
                    r.ay = a.y;
                    r.az = a.z;
                    r.id = n;
                    r.e1x = bb.x - a.x;  // b - a and c - a as the reference forms them (triangle.h:36-37)
                    r.e1y = bb.y - a.y;
                    r.e1z = bb.z - a.z;
                    r.rank = rank ? rank[n] : -n;
                    r.e2x = cc.x - a.x;
                    r.e2y = cc.y - a.y;
                    r.e2z = cc.z - a.z;
                    r.pad = 0;
                    tris[out++] = r;
                } else {
                    const int2 ch = child[n];
                    stack[sp++] = ch.y;
                    stack[sp++] = ch.x;
                }
            }
            node.child[k] = LEAF_BIT | ((c - 1) << 27) | first;
        } else {
            const int id = next_node++;
            node.child[k] = id;
            fout[id - level_base] = make_int2(kid, id);
        }
    }
    nodes[it.y] = node;
}

struct Scratch {
    std::vector<void*> ptrs;
    ~Scratch() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    hipError_t alloc(T*& p, size_t n) {
        p = nullptr;
        hipError_t e = hipMalloc((void**)&p, sizeof(T) * (n ? n : 1));
        if (e == hipSuccess) ptrs.push_back(p);
        return e;
    }
    void release(void* p) {  // hand ownership to the caller
        for (auto& q : ptrs)
            if (q == p) q = nullptr;
    }
};

}  // namespace

hipError_t build_bvh4_gpu(const float4* vpos, const int4* tri, const int32_t* rank, int nt, int leaf_max, int radius,
                          hipStream_t st, GpuBvh* out, std::string* err) {
    auto t0 = std::chrono::steady_clock::now();
    *out = GpuBvh();
    if (nt < 1 || leaf_max < 1 || leaf_max > MAX_LEAF || radius < 1 || radius > MAX_RADIUS) {
        *err = "build_bvh4_gpu: bad arguments";
        return hipErrorInvalidValue;
    }
#define BVH_CHK(x)                                              \
    do {                                                        \
        hipError_t e_ = (x);                                    \
        if (e_ != hipSuccess) {                                 \
            *err = std::string(#x) + ": " + hipGetErrorString(e_); \
            return e_;                                          \
        }                                                       \
    } while (0)
    Scratch s;
    const size_t nn_total = 2 * (size_t)nt - 1;
    float4 *lo, *hi, *cen;
    int *cnt, *ids_a, *ids_b, *nn, *totals, *ctr;
    int2 *child, *fa, *fb;
    uint64_t *keys_a, *keys_b, *flags, *scan;
    unsigned* bounds;
    Node4* nodes;
    TriRec* tris;
    BVH_CHK(s.alloc(lo, nn_total));
    BVH_CHK(s.alloc(hi, nn_total));
    BVH_CHK(s.alloc(cnt, nn_total));
    BVH_CHK(s.alloc(child, nn_total));
    BVH_CHK(s.alloc(cen, nt));
    BVH_CHK(s.alloc(keys_a, nt));
    BVH_CHK(s.alloc(keys_b, nt));
    BVH_CHK(s.alloc(ids_a, nt));
    BVH_CHK(s.alloc(ids_b, nt));
    BVH_CHK(s.alloc(nn, nt));
    BVH_CHK(s.alloc(flags, nt));
    BVH_CHK(s.alloc(scan, nt));
    BVH_CHK(s.alloc(totals, 2));
    BVH_CHK(s.alloc(ctr, 4));
    BVH_CHK(s.alloc(bounds, 8));
    BVH_CHK(s.alloc(tris, nt));
    const unsigned binit[8] = {~0u, ~0u, ~0u, 0u, 0u, 0u, 0u, 0u};
    BVH_CHK(hipMemcpyAsync(bounds, binit, sizeof(binit), hipMemcpyHostToDevice, st));
    const int gb = (nt + BLK - 1) / BLK;
    k_leaves<<<std::min(gb, 1024), BLK, 0, st>>>(vpos, tri, nt, lo, hi, cnt, child, cen, bounds);
    BVH_CHK(hipGetLastError());
    k_morton<<<gb, BLK, 0, st>>>(cen, nt, bounds, keys_a, ids_a);
    BVH_CHK(hipGetLastError());
    // sort along the curve (63 key bits)
    size_t sort_bytes = 0, scan_bytes = 0;
    BVH_CHK(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys_a, keys_b, ids_a, ids_b, nt, 0, 63, st));
    BVH_CHK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, flags, scan, nt, st));
    uint8_t* temp;
    BVH_CHK(s.alloc(temp, std::max(sort_bytes, scan_bytes)));
    BVH_CHK(hipcub::DeviceRadixSort::SortPairs(temp, sort_bytes, keys_a, keys_b, ids_a, ids_b, nt, 0, 63, st));
    unsigned hb[8];
    BVH_CHK(hipMemcpyAsync(hb, bounds, sizeof(hb), hipMemcpyDeviceToHost, st));
    // clustering rounds: ids_b holds the clusters in curve order
    int* cl = ids_b;
    int* cl_next = ids_a;
    int m = nt, node_base = nt, rounds = 0;
    while (m > 1) {
        const int g = (m + BLK - 1) / BLK;
        k_nearest<<<g, BLK, 0, st>>>(cl, m, radius, lo, hi, nn);
        k_pairs<<<g, BLK, 0, st>>>(nn, m, flags);
        size_t bytes = scan_bytes;
        BVH_CHK(hipcub::DeviceScan::ExclusiveSum(temp, bytes, flags, scan, m, st));
        k_merge<<<g, BLK, 0, st>>>(cl, nn, m, node_base, flags, scan, lo, hi, cnt, child, cl_next, totals);
        BVH_CHK(hipGetLastError());
        int tot[2];
        BVH_CHK(hipMemcpyAsync(tot, totals, sizeof(tot), hipMemcpyDeviceToHost, st));
        BVH_CHK(hipStreamSynchronize(st));
        if (tot[1] < 1 || tot[0] >= m) {
            *err = "build_bvh4_gpu: a clustering round merged nothing";
            return hipErrorUnknown;
        }
        m = tot[0];
        node_base += tot[1];
        std::swap(cl, cl_next);
        rounds++;
    }
    if (node_base != (int)nn_total) {
        *err = "build_bvh4_gpu: tree size mismatch";
        return hipErrorUnknown;
    }
    int root = 0;
    BVH_CHK(hipMemcpyAsync(&root, cl, sizeof(int), hipMemcpyDeviceToHost, st));
    BVH_CHK(hipStreamSynchronize(st));
    float smax = 0.0f;
    {
        unsigned u = hb[6];
        u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
        std::memcpy(&smax, &u, 4);
    }
    // collapse: at most nt - 1 four-wide nodes below the root
    BVH_CHK(s.alloc(nodes, (size_t)nt));
    BVH_CHK(s.alloc(fa, (size_t)nt));
    BVH_CHK(s.alloc(fb, (size_t)nt));
    uint8_t* is_leaf;
    BVH_CHK(s.alloc(is_leaf, nn_total));
    k_leaf_choice<<<(int)((nn_total + BLK - 1) / BLK), BLK, 0, st>>>(nt, leaf_max, lo, hi, cnt, child, is_leaf);
    BVH_CHK(hipGetLastError());
    const int init_ctr[4] = {1, 0, 0, 0};
    BVH_CHK(hipMemcpyAsync(ctr, init_ctr, sizeof(init_ctr), hipMemcpyHostToDevice, st));
    const int2 first_item = make_int2(root, 0);
    BVH_CHK(hipMemcpyAsync(fa, &first_item, sizeof(int2), hipMemcpyHostToDevice, st));
    int nin = 1, level_base = 1, depth = 0;
    int c4[4] = {1, 0, 0, 0};
    while (nin > 0) {
        depth++;
        k_collapse<<<(nin + BLK - 1) / BLK, BLK, 0, st>>>(fa, nin, fb, level_base, ctr, nodes, tris, lo, hi, cnt, child,
                                                           nt, is_leaf, smax, vpos, tri, rank);
        BVH_CHK(hipGetLastError());
        BVH_CHK(hipMemcpyAsync(c4, ctr, sizeof(c4), hipMemcpyDeviceToHost, st));
        BVH_CHK(hipStreamSynchronize(st));
        nin = c4[0] - level_base;
        level_base = c4[0];
        std::swap(fa, fb);
        if (depth > 64) {
            *err = "build_bvh4_gpu: tree deeper than 64 levels";
            return hipErrorUnknown;
        }
    }
    if (c4[1] != nt) {
        *err = "build_bvh4_gpu: triangle count mismatch";
        return hipErrorUnknown;
    }
    // exact-size node array
    Node4* exact = nullptr;
    BVH_CHK(hipMalloc((void**)&exact, sizeof(Node4) * (size_t)c4[0]));
    hipError_t e = hipMemcpyAsync(exact, nodes, sizeof(Node4) * (size_t)c4[0], hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        (void)hipFree(exact);
        *err = std::string("build_bvh4_gpu: ") + hipGetErrorString(e);
        return e;
    }
    s.release(tris);
    out->nodes = exact;
    out->tris = tris;
    out->n_nodes = c4[0];
    out->leaves = c4[2];
    out->max_depth = depth + 1;
    out->rounds = rounds;
    out->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return hipSuccess;
#undef BVH_CHK
}

}  // namespace cpt
