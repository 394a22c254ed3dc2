// Device-side scene records, BVH4 traversal and the PT-NEE-MIS shading math.
//
// Float semantics: the reference mixes float and double (EPS = 1e-4 is a double
// literal, src/math/math.h:11; Vec3::operator/(T) multiplies by 1.0/a computed
// in double, math.h:103-105; fabs(float) returns double, ...).  Every helper
// below reproduces those roundings, and the library is compiled with
// -ffp-contract=off so that no multiply-add is fused (the box test's explicit
// fmaf is the one exception: it is conservative, not a restatement of the
// reference's arithmetic; see slab_offsets).  Where the reference's
// double result is provably equal to a float operation (1.0/x and 1.0-x rounded
// to float, sqrt) the float form is used.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bvh_build.h"
#include "bvh_quant.h"

namespace cpt {

// smallest float t with (double)t >= 1e-4: `t < EPS` (double compare,
// triangle.h:55) <=> t < EPS_T
#define CPT_EPS_T 1.00000005e-4f
#define CPT_EPSF 9.99999975e-05f  // (float)1e-4: Vec3 * EPS
#define CPT_PI_F 3.14159274f      // (float)M_PI
#define CPT_M_PI 3.14159265358979323846

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float a, float b, float c) { return {a, b, c}; }
__device__ __forceinline__ V3 v3s(float a) { return {a, a, a}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 mul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 mulf(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
// Vec3 / float: v * float(1.0/double(a)) == v * (1.0f/a) (correctly rounded)
__device__ __forceinline__ V3 divs(V3 a, float s) { return mulf(a, 1.0f / s); }
__device__ __forceinline__ V3 negv(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len3(V3 a) { return (float)sqrt((double)dot(a, a)); }
__device__ __forceinline__ V3 normalized(V3 a) { return divs(a, len3(a)); }
__device__ __forceinline__ V3 ld3(const float4& f) { return {f.x, f.y, f.z}; }
__device__ __forceinline__ float dabs_f(float x) { return (float)fabs((double)x); }

// Mat3 with columns c0,c1,c2 times v (math.h:238-248,319-325)
__device__ __forceinline__ V3 mat_cols_mul(V3 c0, V3 c1, V3 c2, V3 a) {
    return {c0.x * a.x + c1.x * a.y + c2.x * a.z, c0.y * a.x + c1.y * a.y + c2.y * a.z,
            c0.z * a.x + c1.z * a.y + c2.z * a.z};
}
// transpose(Mat3(c0,c1,c2)) * a: rows are the columns
__device__ __forceinline__ V3 mat_cols_tmul(V3 c0, V3 c1, V3 c2, V3 a) {
    return {c0.x * a.x + c0.y * a.y + c0.z * a.z, c1.x * a.x + c1.y * a.y + c1.z * a.z,
            c2.x * a.x + c2.y * a.y + c2.z * a.z};
}

// ---------------------------------------------------------------- RNG
// PCG32 (O'Neill 2014, XSH-RR), one stream per path: initstate =
// splitmix64(splitmix64(seed) ^ path_id), fixed increment (distinct seeds give
// unrelated streams for every path).
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
#define CPT_PCG_MULT 6364136223846793005ull
#define CPT_PCG_INC ((0x5851F42D4C957F2Dull << 1) | 1ull)
__device__ __forceinline__ uint32_t pcg_next(uint64_t& s) {
    uint64_t old = s;
    s = old * CPT_PCG_MULT + CPT_PCG_INC;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
}
__device__ __forceinline__ uint64_t pcg_init(uint64_t seed, uint64_t path_id) {
    uint64_t s = 0;
    pcg_next(s);
    s += splitmix64(splitmix64(seed) ^ path_id);
    pcg_next(s);
    return s;
}
__device__ __forceinline__ float pcg_float(uint64_t& s) { return (float)(pcg_next(s) >> 8) * (1.0f / 16777216.0f); }

// ---------------------------------------------------------------- scene
struct DevScene {
    const QNode* qnodes;     // the BVH4 as 64-byte quantized nodes
    const float4* tris;      // 3 per triangle, leaf order (TriRec)
    const int4* tri_info;    // v0, v1, v2, material (by triangle id)
    const float4* shade;     // 6 per triangle id: g_normal + material, then n0 n1 n2 b0 b1 b2 packed
    const float4* vpos;      // xyz
    const float4* vnrm;
    const float4* vbit;
    const float* mats;       // 16 per material
    const int4* mtex;        // per material: albedo, roughness, metal, normal-map texture (-1: constant)
    const float2* uv;        // per vertex texture coordinates
    const int4* tex;         // per texture: width, height, channels, first texel
    const float* texels;     // linear, row 0 at the top
    const int* lights;       // LightGroup order: triangle id, or -1-k for point light k
    const float4* plights;   // 2 per point light: position, radiance
    int n_lights;
    int n_tris;
    float env[3];
    float cam[16];
    float bmin[3], bext[3];  // scene bounds (guide features)
};

struct Hit {
    float t, beta, gamma;
    int tri;
};

// ---------------------------------------------------------------- traversal
// Ordered BVH4 traversal, one ray per lane.  Stack: LDS_SS entries in LDS
// (laid out entry-major so that lanes hit distinct banks) then a per-thread
// overflow region in global memory.
// 24 LDS entries (a wave takes the overflow path when any lane's stack is deeper):
// 1-2% less closest-hit time than 16 on the 1080p terrain and the 4K interior grid,
// and ten 128-thread blocks per CU still fit the LDS
#ifndef CPT_LDS_SS
#define CPT_LDS_SS 24
#endif
#define CPT_GSTACK (128 - CPT_LDS_SS)
// nodes are addressed by 32-bit byte offsets (index << 6 for 64-byte nodes)
#define CPT_MAX_NODES ((1u << 25) - 1u)
// conservative culling: a box may be entered up to this relative distance
// beyond the current best t, so that triangles tying with the best t in other
// leaves are still tested (ties are decided by rank, not by visit order)
#define CPT_TIE_SLACK 1.000244140625f  // 1 + 2^-12
#define CPT_BOX_SLACK 1.00000048f      // 1 + 2*gamma_3

__device__ __forceinline__ unsigned gthreads_() { return gridDim.x * blockDim.x; }
__device__ __forceinline__ unsigned gtid_() { return blockIdx.x * blockDim.x + threadIdx.x; }

// LDS rows of the traversal stack: CPT_LDS_SS entries plus one row that takes
// the writes of absent children, so that pushes need no branch.
#define CPT_LDS_ROWS (CPT_LDS_SS + 1)

template <int BLOCK>
struct TravStack {
    int* lds;  // CPT_LDS_ROWS x BLOCK
    int* gmem;
    size_t gstride;
    int tid;
    __device__ __forceinline__ void put(int sp, int v) {
        if (sp < CPT_LDS_SS) lds[sp * BLOCK + tid] = v;
        else gmem[(size_t)(sp - CPT_LDS_SS) * gstride] = v;
    }
    __device__ __forceinline__ int get(int sp) {
        return sp < CPT_LDS_SS ? lds[sp * BLOCK + tid] : gmem[(size_t)(sp - CPT_LDS_SS) * gstride];
    }
};

// Reference triangle test (triangle.h:36-77).  Returns true with t, beta, gamma
// when t in [EPS, tmax] and all barycentrics in [0,1].  A ray in the triangle's
// plane (determinant 0, t = 0 * inf = NaN) misses; the reference would accept it
// with a NaN t that turns off every later comparison and all box culling.
__device__ __forceinline__ bool tri_test(const float4& A, const float4& E1, const float4& E2, V3 o, V3 d, float tmax,
                                         float& t, float& beta, float& gamma) {
    V3 v1 = ld3(E1), v2 = ld3(E2);
    V3 v3_ = {-d.x, -d.y, -d.z};
    V3 vs = sub(o, ld3(A));
    float den = v1.x * v2.y * v3_.z + v2.x * v3_.y * v1.z + v3_.x * v1.y * v2.z - v3_.x * v2.y * v1.z -
                v3_.y * v2.z * v1.x - v3_.z * v2.x * v1.y;
    float inv = 1.0f / den;
    t = (v1.x * v2.y * vs.z + v2.x * vs.y * v1.z + vs.x * v1.y * v2.z - vs.x * v2.y * v1.z - vs.y * v2.z * v1.x -
         vs.z * v2.x * v1.y) * inv;
    if (!(t >= CPT_EPS_T && t <= tmax)) return false;
    beta = (vs.x * v2.y * v3_.z + v2.x * v3_.y * vs.z + v3_.x * vs.y * v2.z - v3_.x * v2.y * vs.z -
            v3_.y * v2.z * vs.x - v3_.z * v2.x * vs.y) * inv;
    gamma = (v1.x * vs.y * v3_.z + vs.x * v3_.y * v1.z + v3_.x * v1.y * vs.z - v3_.x * vs.y * v1.z -
             v3_.y * vs.z * v1.x - v3_.z * vs.x * v1.y) * inv;
    float alpha = 1.0f - (gamma + beta);
    return !(alpha < 0 || alpha > 1 || beta < 0 || beta > 1 || gamma < 0 || gamma > 1);
}

// Slab offsets of one ray for the box test.  A box plane P is crossed at
// t = (P - o) * inv; with c = -(o * inv) a test can evaluate fma(P, inv, c), one
// instruction per plane (the quantized node test adds its own offset on top).  That form has an absolute error of up to one rounding
// of o * inv (plus the rounding of c), so c is moved by e = 2^-21 |o * inv| (four
// times that error) toward smaller t for the plane that is nearer along the axis
// (c_near) and toward larger t for the other (c_far): the computed entry never
// exceeds the exact entry and the computed exit never falls short of the exact
// exit, beyond one relative rounding of their own, which CPT_BOX_SLACK and
// CPT_TIE_SLACK cover.  Culling stays conservative; only boxes within e of the
// ray's slabs are entered that the exact test would skip.
__device__ __forceinline__ void slab_offsets(float o, float inv, float& c_near, float& c_far) {
    const float oi = o * inv;
    const float e = fabsf(oi) * 4.76837158203125e-7f;  // 2^-21
    c_near = -oi - e;
    c_far = -oi + e;
}

// Chunked persistent traversal.  Wave w of the grid owns rays
// [w*per, (w+1)*per) of the queue and keeps its 64 lanes busy: whenever at
// least `refill` lanes (or all of them) are idle it hands them the next rays of
// its chunk (refill: a launch parameter), so a wave no longer waits for its slowest ray before starting new
// ones, and no atomics are needed.  Each iteration a lane may do one node step
// (four slab tests on a 64-byte quantized node, near-first push) and one
// triangle test of a pending leaf:
// a leaf met during the node step is parked in `leaf` until the previous leaf's
// triangles are done.  Culling stays conservative and ties are decided by rank,
// so the result does not depend on this visiting order.  Pushes and pops use the LDS rows directly while every lane of the
// wave stays within them; a wave with a lane beyond them takes the slower path
// through the global overflow region.
//   load(i, o, d, tmax)   fetches ray i of the queue
//   done(i, hit)          receives the closest hit (ANY: h.tri >= 0 iff occluded)
// no node pending.  As a leaf reference -2 would be 16 triangles from index
// 2^27-2, which no scene of fewer than 2^27 triangles has; EMPTY_CHILD is never
// stored in `node`.
#define CPT_NONE (-2)
template <int BLOCK, bool ANY, bool COUNT, class Load, class Done>
__device__ __forceinline__ void trace_chunked(const DevScene& sc, unsigned count, unsigned refill, TravStack<BLOCK>& st,
                                              Load load, Done done, unsigned& nvis, unsigned& ntest) {
    const unsigned lane = __lane_id();
    const unsigned nw = gthreads_() >> 6, w = gtid_() >> 6;
    const unsigned per = (count + nw - 1) / nw;
    unsigned nextr = min(count, w * per);
    const unsigned endr = min(count, nextr + per);
    const unsigned long long lt = lane ? ((~0ull) >> (64 - lane)) : 0ull;
    const char* nbase = reinterpret_cast<const char*>(sc.qnodes);
    bool has = false, fresh = false;
    unsigned ray = 0;
    V3 o = {0, 0, 0}, d = {0, 0, 0};
    float ix = 0, iy = 0, iz = 0, best = 0;
    float cxn = 0, cxf = 0, cyn = 0, cyf = 0, czn = 0, czf = 0;  // slab offsets
    bool sgx = false, sgy = false, sgz = false;                  // direction signs: near planes high
    int best_rank = INT_MIN;
    Hit h;
    h.t = 0.0f;
    h.beta = 0.0f;
    h.gamma = 0.0f;
    h.tri = -1;
    int sp = 0, node = CPT_NONE, leaf = CPT_NONE;
    int tcur = 0, tend = 0;
    while (true) {
        // The reciprocal direction of a fresh ray is formed inside its first
        // node step, after the root's loads are issued, so that the ray loads
        // and the node loads are in flight together.
        const unsigned long long idle = __ballot(!has);
        const unsigned nidle = (unsigned)__popcll(idle);
        if (nextr < endr && (nidle >= refill || nidle == 64u)) {
            const unsigned r = nextr + (unsigned)__popcll(idle & lt);
            if (!has && r < endr) {
                load(r, o, d, best);
                best_rank = INT_MIN;
                // a miss reports zeros, not the lane's previous ray's hit
                h.t = 0.0f;
                h.beta = 0.0f;
                h.gamma = 0.0f;
                h.tri = -1;
                sp = 0;
                node = 0;  // root
                leaf = CPT_NONE;
                tcur = tend = 0;
                ray = r;
                has = true;
                fresh = true;
                sgx = (__float_as_uint(d.x) >> 31) != 0;
                sgy = (__float_as_uint(d.y) >> 31) != 0;
                sgz = (__float_as_uint(d.z) >> 31) != 0;
            }
            nextr = min(endr, nextr + nidle);
        }
        if (__ballot(has) == 0ull) break;
        if (!has) continue;
        // This iteration's loads are issued together (the pending triangle and
        // the node), then the triangle is tested and the node's boxes culled
        // against the possibly improved best.
        const bool do_tri = tcur < tend, do_node = node >= 0;
        float4 A, E1, E2;
        int4 ch;
        float4 qo;  // quantized node: low corner and scale exponents
        uint4 qa;   // its x and y plane bytes
        uint2 qb;   // its z plane bytes
        if (do_tri) {
            const float4* tp = sc.tris + 3 * (size_t)tcur;
            A = tp[0];
            E1 = tp[1];
            E2 = tp[2];
        }
        if (do_node) {
            const uint32_t no = (uint32_t)node << 6;
            qo = *reinterpret_cast<const float4*>(nbase + no);
            ch = *reinterpret_cast<const int4*>(nbase + (no + 16u));
            qa = *reinterpret_cast<const uint4*>(nbase + (no + 32u));
            qb = *reinterpret_cast<const uint2*>(nbase + (no + 48u));
        }
        if (fresh) {
            // 1/d with zero components replaced by a tiny signed value so that
            // slab products never form 0*inf
            float dx = fabsf(d.x) < 1e-20f ? copysignf(1e-20f, d.x) : d.x;
            float dy = fabsf(d.y) < 1e-20f ? copysignf(1e-20f, d.y) : d.y;
            float dz = fabsf(d.z) < 1e-20f ? copysignf(1e-20f, d.z) : d.z;
            ix = 1.0f / dx;
            iy = 1.0f / dy;
            iz = 1.0f / dz;
            slab_offsets(o.x, ix, cxn, cxf);
            slab_offsets(o.y, iy, cyn, cyf);
            slab_offsets(o.z, iz, czn, czf);
            fresh = false;
        }
        bool finished = false;
        // ---- one triangle of the pending leaf
        if (do_tri) {
            tcur++;
            float t, b, g;
            if (COUNT) ntest++;
            if (tri_test(A, E1, E2, o, d, best, t, b, g)) {
                if (ANY) {
                    h.tri = __float_as_int(A.w);
                    finished = true;
                } else {
                    int rk = __float_as_int(E1.w);
                    if (t < best || rk > best_rank) {
                        best = t;
                        best_rank = rk;
                        h.t = t;
                        h.beta = b;
                        h.gamma = g;
                        h.tri = __float_as_int(A.w);
                    }
                }
            }
        }
        // ---- node step
        if (do_node && !finished) {
            if (COUNT) nvis++;
            float lim = ANY ? best * CPT_BOX_SLACK : best * CPT_TIE_SLACK;
            float key[4];
            int cid[4];
            {
                // A plane of axis a is P = o_a + q 2^k_a (q the child's byte), crossed at
                // t = q (2^k_a inv_a) + (o_a inv_a + c): the first term is exact (a power
                // of two times inv, a normal float for the stored k), the node's offset
                // takes one rounding, so it is moved outward by 2^-21 of its size
                // (below the near planes, above the far ones; the near and far planes
                // are the low or high bytes by the direction's sign, as for float
                // nodes).  Quantized planes lie outside the exact ones, so culling stays
                // conservative.
                const uint32_t ex = __float_as_uint(qo.w);
                const float sx = __uint_as_float((ex & 0xffu) << 23) * ix;
                const float sy = __uint_as_float(((ex >> 8) & 0xffu) << 23) * iy;
                const float sz = __uint_as_float(((ex >> 16) & 0xffu) << 23) * iz;
                float bxn = fmaf(qo.x, ix, cxn), bxf = fmaf(qo.x, ix, cxf);
                float byn = fmaf(qo.y, iy, cyn), byf = fmaf(qo.y, iy, cyf);
                float bzn = fmaf(qo.z, iz, czn), bzf = fmaf(qo.z, iz, czf);
                constexpr float R = 4.76837158203125e-7f;  // 2^-21
                bxn = fmaf(-fabsf(bxn), R, bxn);
                byn = fmaf(-fabsf(byn), R, byn);
                bzn = fmaf(-fabsf(bzn), R, bzn);
                bxf = fmaf(fabsf(bxf), R, bxf);
                byf = fmaf(fabsf(byf), R, byf);
                bzf = fmaf(fabsf(bzf), R, bzf);
                const uint32_t nqx = sgx ? qa.y : qa.x, fqx = sgx ? qa.x : qa.y;
                const uint32_t nqy = sgy ? qa.w : qa.z, fqy = sgy ? qa.z : qa.w;
                const uint32_t nqz = sgz ? qb.y : qb.x, fqz = sgz ? qb.x : qb.y;
#define CPT_QB(v, c) ((float)(((v) >> (8 * (c))) & 0xffu))
#define CPT_QBOX(c, CH)                                                                                      \
    {                                                                                                         \
        float tn = fmaxf(fmaxf(fmaf(CPT_QB(nqx, c), sx, bxn), fmaf(CPT_QB(nqy, c), sy, byn)),                 \
                         fmaxf(fmaf(CPT_QB(nqz, c), sz, bzn), 0.0f));                                           \
        float tf = fminf(fminf(fmaf(CPT_QB(fqx, c), sx, bxf), fmaf(CPT_QB(fqy, c), sy, byf)),                 \
                         fmaf(CPT_QB(fqz, c), sz, bzf)) * CPT_BOX_SLACK;                                        \
        bool hit = tn <= tf && tn <= lim;                                                                     \
        key[c] = hit ? tn : INFINITY;                                                                         \
        cid[c] = hit ? CH : EMPTY_CHILD;                                                                      \
    }
                CPT_QBOX(0, ch.x)
                CPT_QBOX(1, ch.y)
                CPT_QBOX(2, ch.z)
                CPT_QBOX(3, ch.w)
#undef CPT_QBOX
#undef CPT_QB
            }
#define CPT_CSWAP(a, b)                                  \
    if (key[b] < key[a]) {                               \
        float tk = key[a]; key[a] = key[b]; key[b] = tk; \
        int tc = cid[a]; cid[a] = cid[b]; cid[b] = tc;   \
    }
#ifndef CPT_ANY_UNSORTED
#define CPT_ANY_UNSORTED 0
#endif
            if (!(ANY && CPT_ANY_UNSORTED)) {
                CPT_CSWAP(0, 1)
                CPT_CSWAP(2, 3)
                CPT_CSWAP(0, 2)
                CPT_CSWAP(1, 3)
                CPT_CSWAP(1, 2)
            }
#undef CPT_CSWAP
            // push cid[3], cid[2], cid[1] of the children present, deepest first
            // (absent ones sort last; an absent child of a quantized node on a flat
            // axis can pass the slab test but is still not pushed)
            const int v3 = cid[3] != EMPTY_CHILD, v2 = cid[2] != EMPTY_CHILD, v1 = cid[1] != EMPTY_CHILD;
            if (__ballot(sp + 3 > CPT_LDS_SS) == 0ull) {
                const int s2 = sp + v3, s1 = s2 + v2;
                st.lds[(v3 ? sp : CPT_LDS_SS) * BLOCK + st.tid] = cid[3];
                st.lds[(v2 ? s2 : CPT_LDS_SS) * BLOCK + st.tid] = cid[2];
                st.lds[(v1 ? s1 : CPT_LDS_SS) * BLOCK + st.tid] = cid[1];
                sp = s1 + v1;
            } else {
                if (v3) st.put(sp++, cid[3]);
                if (v2) st.put(sp++, cid[2]);
                if (v1) st.put(sp++, cid[1]);
            }
            node = cid[0] != EMPTY_CHILD ? cid[0] : CPT_NONE;
        }
        if (!finished) {
            // park a leaf met by the node step; start it once the previous
            // leaf's triangles are done; keep the node stream going
            if (node != CPT_NONE && node < 0 && leaf == CPT_NONE) {
                leaf = node;
                node = CPT_NONE;
            }
            if (leaf != CPT_NONE && tcur >= tend) {
                tcur = leaf & ((1 << 27) - 1);
                tend = tcur + ((leaf >> 27) & 15) + 1;
                leaf = CPT_NONE;
            }
            const bool pop = node == CPT_NONE && sp > 0;
            if (__ballot(pop && sp > CPT_LDS_SS) == 0ull) {
                if (pop) node = st.lds[(sp - 1) * BLOCK + st.tid];
            } else if (pop) {
                node = st.get(sp - 1);
            }
            sp -= pop ? 1 : 0;
            finished = tcur >= tend && leaf == CPT_NONE && node == CPT_NONE;
        }
        if (finished) {
            done(ray, h);
            has = false;
        }
    }
}

}  // namespace cpt
