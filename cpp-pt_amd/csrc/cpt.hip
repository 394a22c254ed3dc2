// libcpt: wavefront PT-NEE-MIS on one MI355X (gfx950) behind the C ABI of include/cpt.h.
//
// Design (one GPU, one HIP stream per context):
//  * Paths live in fixed slots of SoA state arrays for one pass.  A pass is a
//    rectangle (pixel range x sample range); slot = sample_local*npix + pixel_local.
//  * Per bounce three grid-stride kernels run over compacted queues:
//      extend  : closest hit for the extension-ray queue (BVH4, LDS stack)
//      shade   : emitter MIS, BSDF sample, NEE light sample, RR; appends the
//                next extension ray and one shadow ray per path with
//                wave64 ballot + popcount and one atomic per wave
//      shadow  : any hit for the shadow queue; unoccluded -> add the precomputed
//                NEE contribution to the path's radiance
//    Queue counts live in device memory; kernels read them and exit early, so
//    the whole pass (raygen, max_depth bounces, film accumulation) is captured
//    once in a hipGraph and replayed per pass.
//  * The film is a W*H*3 FP64 accumulator; each pass adds its slots' radiance
//    in sample order, so results are deterministic and identical in summation
//    order to a sequential per-pixel loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cpt.h"
#include "bvh_build.h"
#include "bvh_gpu.h"
#include "sdtree.h"
#include "sdtree_dev.h"
#include "pt_device.h"
#include "pt_shade.h"
#include "guide_mlp.h"

using namespace cpt;

namespace {

// A camera view of a batch (cpt_render renders a batch of one view).
struct ViewDesc {
    float cam[16];
    int W, H, S;
    unsigned id;                     // mixed into the path id: (id << 40) + pixel*spp + sample
    unsigned long long path_begin;   // streaming mode: first global path index of the view
    unsigned long long n_paths;      // W*H*spp
    unsigned long long film_off;     // first pixel of the view in the concatenated films
    unsigned spp;                    // samples per pixel: S*S, or any count when not stratified
    int strat;                       // 1: sample s covers stratum (s/S, s%S) of an SxS grid
                                     // (pathtracerMIS.h:147-153); 0: the whole pixel (ppg.h:178-181)
};

// film coordinates of sample `sample` of pixel (px, py) of a view
__device__ __forceinline__ void view_coords(const ViewDesc& vd, int px, int py, unsigned sample, float r1, float r2,
                                            float& cx, float& cy) {
    if (vd.strat) film_coords(px, py, (int)sample / vd.S, (int)sample % vd.S, vd.S, vd.W, vd.H, r1, r2, cx, cy);
    else film_coords(px, py, 0, 0, 1, vd.W, vd.H, r1, r2, cx, cy);
}

// One pass = up to CPT_MAX_SEG rectangles (view, pixel range x sample range).
// Slots of a segment: slot0 + sample_local*npix + pixel_local.  A pass never
// holds two segments of one view that share pixels, so film accumulation is
// race-free and adds every pixel's samples in ascending sample order.
#define CPT_MAX_SEG 256
struct Segment {
    uint32_t view, p0, npix, s0, ns, slot0, pix0, pad;
};
struct PassDesc {
    uint32_t n_seg, n_paths, n_pix, pass_id;
    uint64_t seed;     // render seed: copied in per pass, so a captured graph serves every seed
    float guide_frac;  // guided renders: probability of sampling the guide in this pass
    uint32_t pad;
    Segment seg[CPT_MAX_SEG];
};

enum {
    ST_EXT = 0,
    ST_SHADOW = 1,
    ST_NODES = 2,
    ST_TRIS = 3,
    ST_SNODES = 4,
    ST_STRIS = 5,
    ST_PATHS = 6,
    ST_N = 8
};

struct WfArgs {
    DevScene sc;
    float4* mul_p;     // mul.xyz, p
    float4* col;       // radiance
    uint64_t* rng;
    uint32_t* meta;    // depth | sampled_delta << 8
    uint32_t* pix;     // pixel index (streaming mode: index into the concatenated films)
    uint32_t* vid;     // streaming mode: view of the slot's path
    uint32_t* done;    // streaming mode: slots whose path ended this bounce
    float4* qo[2];     // o.xyz, tmax
    float4* qd[2];     // d.xyz, slot
    float4* hits;      // t, beta, gamma, tri
    float4* so;        // shadow o.xyz, tmax
    float4* sd;        // shadow d.xyz, slot
    float4* sc_rgb;    // shadow contribution
    // queue counters.  Words 2k and 2k+1 form one 64-bit word per extension queue
    // k: its ray count (low) and the count of the shadow queue filled by the same
    // shading launch (high), so that one atomic reserves space in both.  Word 4
    // counts the ended-path queue (streaming mode).
    uint32_t* counters;
    unsigned long long* stats;
    int* gstack;
    const PassDesc* pass;
    const ViewDesc* views;
    unsigned long long* pass_tend;  // s_memrealtime at the end of each pass
    double* film;      // per-pixel radiance sums
    double* film2;     // per-pixel sums of squared radiance
    // learned guide (guided renders only)
    float* lobes;            // stored mixtures, structure of arrays: value v (0..3 L - 1: half2 (kappa mu.x,
                             // kappa mu.y)[L], half2 (kappa mu.z, weight)[L], log term c[L], L = guide_lobes)
                             // of queue entry e at lobes[v * lobes_stride + e] (see gmix_pdf)
    size_t lobes_stride;
    float* rec;              // training records: [slot][depth][REC_FLOATS]
    uint8_t* rec_flag;       // per record: 0 empty, 1 vertex recorded, 2 radiance after it known
    uint32_t n_rec;          // slots that record: every rec_stride-th slot, n_rec of them
    uint32_t rec_stride;
    int rec_depth;
    int guide_depth;         // vertices at depth >= this are not guided
    int guide_target;        // 0: train on L_i f cos, 1: on L_i
    unsigned refill;         // idle lanes that make a traversal wave fetch new rays
    int integrator;          // 0 PathtracerMIS, 1 Pathtracer (NEE only), 2 PurePt, 3 PPG, 4 Lighttracer
    int corrected;           // CPT_RENDER_CORRECTED: the reference's biased quirks switched off
    // SD-tree guided arm (PPG): the tree, per path its vertices (origin, direction,
    // radiance factor: 9 floats each, max_depth per path) and vertex count
    SdDev sdt;
    float* ppg_vtx;
    uint32_t* ppg_nv;
    float* gsamples;         // training samples out (11 floats)
    uint32_t* gcount;
    const uint16_t* gwt;
    const uint8_t* gwt8;     // FP8 weights (k_quant8) and their inverse scales
    const float* ginv8;
    int guide_fp8;           // guide inference in FP8
    int guide_lobes;         // 8 (width-64 network) or 16 (width 128); lobes holds 5 x this per entry
    const float* gparams;
    int W, H, S, max_depth;
    uint64_t seed;
    uint32_t view;
};

// training record of one guided path vertex: inputs[7], dir[3], pdf, path throughput
// before the vertex [3], radiance snapshot[3], state (0 empty, 1 direction traced,
// 2 complete), path throughput after the vertex [3]
// record: features [0, 7), sampled direction [7, 10), its pdf [10], throughput before
// the vertex [11, 14), path radiance when the vertex was recorded [14, 17), throughput
// after it [17, 20)
#define REC_FLOATS 20

// waves per SIMD the closest-hit traversal is compiled for: caps its VGPRs at 96
// (its loop over quantized nodes needs 90)
#ifndef CPT_TRAV_WAVES
#define CPT_TRAV_WAVES 5
#endif
// ... and the any-hit kernel, whose loop needs fewer registers: 6 (80 VGPRs, 2
// spilled) is 7% faster for shadow rays than 5, while closest hit loses at 6
#ifndef CPT_SHADOW_WAVES
#define CPT_SHADOW_WAVES 6
#endif

// waves per SIMD the unguided shading kernel is compiled for (caps its VGPRs at
// 128: 7% faster on the 1080p terrain; the guided variant spills under that
// cap and runs uncapped)
#ifndef CPT_SHADE_WAVES
#define CPT_SHADE_WAVES 4
#endif

// idle lanes (of 64) that make a traversal wave fetch new rays of its chunk
#define CPT_DEFAULT_REFILL 32u

// device BVH build: largest leaf and the clustering's neighbour window
#ifndef CPT_GPU_BVH_LEAF
#define CPT_GPU_BVH_LEAF 4
#endif
#ifndef CPT_GPU_BVH_RADIUS
#define CPT_GPU_BVH_RADIUS 16
#endif

__device__ __forceinline__ unsigned gthreads() { return gridDim.x * blockDim.x; }
__device__ __forceinline__ unsigned gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }

// wave64 queue append: one atomic per wave, order-preserving within the wave
__device__ __forceinline__ unsigned wave_alloc(bool pred, uint32_t* counter) {
    unsigned long long mask = __ballot(pred);
    unsigned lane = __lane_id();
    unsigned base = 0;
    if (mask) {
        int leader = __ffsll((unsigned long long)mask) - 1;
        if ((int)lane == leader) base = atomicAdd(counter, (unsigned)__popcll(mask));
        base = __shfl(base, leader);
    }
    unsigned long long lt = lane ? (mask & ((~0ull) >> (64 - lane))) : 0ull;
    return base + (unsigned)__popcll(lt);
}

// Block-wide append to the two queues whose counts share the 64-bit word *ctr
// (low: q0, high: q1): one atomic per block, order-preserving within a wave.
// Every thread of the block must call it (it holds two barriers); sh is a
// 2*B/64+2 word LDS buffer that callers alternate between consecutive calls.
template <int B>
__device__ __forceinline__ void block_alloc2(bool p0, bool p1, unsigned long long* ctr, unsigned* sh, unsigned& o0,
                                             unsigned& o1) {
    constexpr int NW = B / 64;
    const unsigned lane = __lane_id(), w = threadIdx.x >> 6;
    const unsigned long long m0 = __ballot(p0), m1 = __ballot(p1);
    if (lane == 0) {
        sh[2 * w] = (unsigned)__popcll(m0);
        sh[2 * w + 1] = (unsigned)__popcll(m1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t0 = 0, t1 = 0;
        for (int k = 0; k < NW; k++) {
            t0 += sh[2 * k];
            t1 += sh[2 * k + 1];
        }
        unsigned long long old = 0;
        if (t0 | t1) old = atomicAdd(ctr, ((unsigned long long)t1 << 32) | t0);
        sh[2 * NW] = (unsigned)old;
        sh[2 * NW + 1] = (unsigned)(old >> 32);
    }
    __syncthreads();
    unsigned b0 = sh[2 * NW], b1 = sh[2 * NW + 1];
    for (unsigned k = 0; k < w; k++) {
        b0 += sh[2 * k];
        b1 += sh[2 * k + 1];
    }
    const unsigned long long lt = lane ? ((~0ull) >> (64 - lane)) : 0ull;
    o0 = b0 + (unsigned)__popcll(m0 & lt);
    o1 = b1 + (unsigned)__popcll(m1 & lt);
}

__device__ __forceinline__ unsigned wave_sum(unsigned v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// ------------------------------------------------------------------ raygen
// Training record of path slot `slot`: recording slots are spread evenly over the
// pass (every rec_stride-th slot), so the samples cover the whole image
__device__ __forceinline__ bool rec_index(const WfArgs& a, unsigned slot, unsigned& ri) {
    ri = slot / a.rec_stride;
    return ri < a.n_rec && ri * a.rec_stride == slot;
}

__device__ __forceinline__ int find_segment(const PassDesc* pd, unsigned i, bool by_pixel) {
    int lo = 0, hi = (int)pd->n_seg - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        unsigned start = by_pixel ? pd->seg[mid].pix0 : pd->seg[mid].slot0;
        if (start <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The traversal's 64-byte copy of the tree, one node per thread
__global__ void k_quantize_nodes(const Node4* nodes, int64_t n, QNode* q) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        quantize_node(nodes[i], q[i]);
}

// The tmax of k_raygen's camera rays: the largest finite float, which culls and
// accepts exactly as an infinite one (every t and box distance is finite), and marks
// a queue entry whose path state is the initial one: throughput 1, pdf 1, radiance
// 0, depth 0, camera ray.  Raygen writes only the path's random state; the shading
// kernels take the rest from this mark instead of reading 36 bytes per path.
#define CPT_FRESH_TMAX 3.40282347e+38f

__global__ void k_raygen(WfArgs a) {
    const PassDesc* pd = a.pass;
    const unsigned n_paths = pd->n_paths;
    if (gtid() == 0) {
        a.counters[0] = n_paths;
        a.counters[1] = 0;
        a.counters[2] = 0;
        a.counters[3] = 0;
        a.counters[4] = 0;
        atomicAdd(&a.stats[ST_PATHS], (unsigned long long)n_paths);
    }
    for (unsigned i = gtid(); i < n_paths; i += gthreads()) {
        const Segment sg = pd->seg[find_segment(pd, i, false)];
        const ViewDesc& vd = a.views[sg.view];
        unsigned local = i - sg.slot0;
        unsigned sl = local / sg.npix, pl = local - sl * sg.npix;
        unsigned pixel = sg.p0 + pl;
        unsigned sample = sg.s0 + sl;
        int px = (int)(pixel % (unsigned)vd.W), py = (int)(pixel / (unsigned)vd.W);
        uint64_t path_id = ((uint64_t)vd.id << 40) + (uint64_t)pixel * (uint64_t)vd.spp + sample;
        uint64_t rng = pcg_init(pd->seed, path_id);
        float r1 = pcg_float(rng);
        float r2 = pcg_float(rng);
        float cx, cy;
        view_coords(vd, px, py, sample, r1, r2, cx, cy);
        V3 o, d;
        camera_ray(vd.cam, cx, cy, o, d);
        a.qo[0][i] = make_float4(o.x, o.y, o.z, CPT_FRESH_TMAX);
        a.qd[0][i] = make_float4(d.x, d.y, d.z, __int_as_float((int)i));
        a.rng[i] = rng;
        unsigned ri;
        if (a.n_rec && rec_index(a, i, ri))
            for (int k = 0; k < a.rec_depth; k++) a.rec_flag[(size_t)ri * a.rec_depth + k] = 0;
    }
}

// Light paths (Lighttracer, lighttracer.h:27-35,138-148): one per camera sample
// of the view, with the same per-path stream; no film coordinates are drawn.
// Paths whose start fails get an empty ray (tmax -1) that ends at the first step.
__global__ void k_raygen_lt(WfArgs a) {
    const PassDesc* pd = a.pass;
    const unsigned n_paths = pd->n_paths;
    if (gtid() == 0) {
        a.counters[0] = n_paths;
        a.counters[1] = 0;
        a.counters[2] = 0;
        a.counters[3] = 0;
        a.counters[4] = 0;
        atomicAdd(&a.stats[ST_PATHS], (unsigned long long)n_paths);
    }
    for (unsigned i = gtid(); i < n_paths; i += gthreads()) {
        const Segment sg = pd->seg[find_segment(pd, i, false)];
        const ViewDesc& vd = a.views[sg.view];
        unsigned local = i - sg.slot0;
        unsigned sl = local / sg.npix, pl = local - sl * sg.npix;
        unsigned pixel = sg.p0 + pl;
        unsigned sample = sg.s0 + sl;
        uint64_t path_id = ((uint64_t)vd.id << 40) + (uint64_t)pixel * (uint64_t)vd.spp + sample;
        uint64_t rng = pcg_init(pd->seed, path_id);
        V3 o, d, rad;
        float pdf;
        V3 mulv = {0.0f, 0.0f, 0.0f};
        float tmax = INFINITY;
        if (light_start(a.sc, rng, o, d, rad, pdf)) {
            mulv = divs(rad, pdf);
        } else {
            o = v3(0.0f, 0.0f, 0.0f);
            d = v3(1.0f, 0.0f, 0.0f);
            tmax = -1.0f;
        }
        a.qo[0][i] = make_float4(o.x, o.y, o.z, tmax);
        a.qd[0][i] = make_float4(d.x, d.y, d.z, __int_as_float((int)i));
        a.mul_p[i] = make_float4(mulv.x, mulv.y, mulv.z, 1.0f);
        a.col[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        a.rng[i] = rng;
        a.meta[i] = 0u;
    }
}

// ------------------------------------------------------------------ extend
template <int B, bool COUNT>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(CPT_TRAV_WAVES))) void k_extend(WfArgs a, int cur) {
    __shared__ int lds[CPT_LDS_ROWS * B];
    const unsigned count = a.counters[2 * cur];
    if (gtid() == 0) {
        a.counters[2 * (cur ^ 1)] = 0;
        a.counters[2 * (cur ^ 1) + 1] = 0;
        a.counters[4] = 0;
        atomicAdd(&a.stats[ST_EXT], (unsigned long long)count);
    }
    TravStack<B> st{lds, a.gstack + gtid(), (size_t)gthreads(), (int)threadIdx.x};
    unsigned nv = 0, nt = 0;
    const float4* qo = a.qo[cur];
    const float4* qd = a.qd[cur];
    float4* hits = a.hits;
    trace_chunked<B, false, COUNT>(
        a.sc, count, a.refill, st,
        [&](unsigned i, V3& o, V3& d, float& tmax) {
            float4 O = qo[i], D = qd[i];
            o = v3(O.x, O.y, O.z);
            d = v3(D.x, D.y, D.z);
            tmax = O.w;
        },
        [&](unsigned i, const Hit& h) { hits[i] = make_float4(h.t, h.beta, h.gamma, __int_as_float(h.tri)); }, nv, nt);
    if (COUNT) {
        nv = wave_sum(nv);
        nt = wave_sum(nt);
        if (__lane_id() == 0) {
            atomicAdd(&a.stats[ST_NODES], (unsigned long long)nv);
            atomicAdd(&a.stats[ST_TRIS], (unsigned long long)nt);
        }
    }
}


// ------------------------------------------------------------------ learned guide in the path

// Stored mixture of one queue entry (lob = lobes + entry), NL lobes, 3 values per
// lobe k at lob[(v NL + k) st] (st: the structure-of-arrays stride, so a wave's
// loads of one value for consecutive entries are contiguous): half2 (kappa mu_x,
// kappa mu_y), half2 (kappa mu_z, weight w~), and c = log(w~ / S C(kappa')) -
// kappa' with kappa' = |kappa mu| (the guide network's decoders write them).  The
// pdf is sum_k exp(kappa mu_k . w + c_k); the sampler below draws from exactly that
// density (up to float rounding), from the same stored values.
template <int NL>
__device__ __forceinline__ float gmix_pdf(const float* lob, size_t st, V3 w) {
    float p = 0.0f;
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const float2 a = guide::unpack_h2(__float_as_uint(lob[k * st]));
        const float2 b = guide::unpack_h2(__float_as_uint(lob[(NL + k) * st]));
        p += __expf(a.x * w.x + a.y * w.y + b.x * w.z + lob[(2 * NL + k) * st]);
    }
    return p;
}

// The mixture pdf at two directions in one pass over the stored lobes (each sum
// formed exactly as gmix_pdf forms it)
template <int NL>
__device__ __forceinline__ void gmix_pdf2(const float* lob, size_t st, V3 w1, V3 w2, float& p1, float& p2) {
    p1 = 0.0f;
    p2 = 0.0f;
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const float2 a = guide::unpack_h2(__float_as_uint(lob[k * st]));
        const float2 b = guide::unpack_h2(__float_as_uint(lob[(NL + k) * st]));
        const float c = lob[(2 * NL + k) * st];
        p1 += __expf(a.x * w1.x + a.y * w1.y + b.x * w1.z + c);
        p2 += __expf(a.x * w2.x + a.y * w2.y + b.x * w2.z + c);
    }
}

template <int NL>
__device__ __forceinline__ V3 gmix_sample(const float* lob, size_t st, float u0, float u1, float u2) {
    // lobe k with probability w~_k / sum w~ (the decoder's S, summed in another
    // order, differs by rounding only), then a vMF sample of concentration kappa' = |kappa mu_k| about
    // kappa mu_k / kappa' (Jakob 2012, "Numerically stable sampling of the von
    // Mises Fisher distribution"; orthonormal basis after Duff et al. 2017); values
    // read from memory: a local array indexed by the chosen lobe would live in
    // scratch
    float S = 0.0f;
#pragma unroll
    for (int j = 0; j < NL; j++) S += guide::unpack_h2(__float_as_uint(lob[(NL + j) * st])).y;
    const float t = u0 * S;
    int k = NL - 1;
    float c = 0.0f;
    for (int j = 0; j < NL; j++) {
        c += guide::unpack_h2(__float_as_uint(lob[(NL + j) * st])).y;
        if (t < c) {
            k = j;
            break;
        }
    }
    const guide::MixLobe m = guide::mix_lobe(__float_as_uint(lob[k * st]), __float_as_uint(lob[(NL + k) * st]));
    const float kap = m.kap;
    float w = 1.0f + __logf(u1 + (1.0f - u1) * __expf(-2.0f * kap)) / kap;
    w = fminf(fmaxf(w, -1.0f), 1.0f);
    float r = sqrtf(fmaxf(0.0f, 1.0f - w * w));
    float sn, co;
    __sincosf(guide::TWO_PI * u2, &sn, &co);
    const float ik = 1.0f / kap;
    float mx = m.kx * ik, my = m.ky * ik, mz = m.kz * ik;
    float sign = copysignf(1.0f, mz);
    float a = -1.0f / (sign + mz);
    float b = mx * my * a;
    V3 t1 = {1.0f + sign * mx * mx * a, sign * b, -sign * mx};
    V3 t2 = {b, sign + my * my * a, -my};
    V3 d = {w * mx + r * (co * t1.x + sn * t2.x), w * my + r * (co * t1.y + sn * t2.y), w * mz + r * (co * t1.z + sn * t2.z)};
    float n = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
    return {d.x / n, d.y / n, d.z / n};
}

// the guide's lobe count is 8 (width-64 network) or 16 (width 128)
__device__ __forceinline__ float gmix_pdf_n(const float* lob, size_t st, int nl, V3 w) {
    return nl == 16 ? gmix_pdf<16>(lob, st, w) : gmix_pdf<8>(lob, st, w);
}
__device__ __forceinline__ void gmix_pdf2_n(const float* lob, size_t st, int nl, V3 w1, V3 w2, float& p1, float& p2) {
    if (nl == 16) gmix_pdf2<16>(lob, st, w1, w2, p1, p2);
    else gmix_pdf2<8>(lob, st, w1, w2, p1, p2);
}
__device__ __forceinline__ V3 gmix_sample_n(const float* lob, size_t st, int nl, float u0, float u1, float u2) {
    return nl == 16 ? gmix_sample<16>(lob, st, u0, u1, u2) : gmix_sample<8>(lob, st, u0, u1, u2);
}

__device__ __forceinline__ bool guidable(const Mat& m) {
    return m.type == CPT_MAT_DIFFUSE || (m.type == CPT_MAT_GLOSSY && m.rough >= 0.01f);
}

// Per bounce: network inputs of every queued hit, one MLP forward per tile of 64,
// decoded mixture written per queue entry (6 x lobes values, structure of arrays).
// The guide's inputs at a queued hit (the hit point, the shading normal the
// shading kernel will use, -d) without the full hit record: the interpolated
// normal in float (the features need no reference rounding).  Where the shading
// kernel re-frames the hit (`frame`: first hits, or every hit in corrected mode)
// and the material bends the normal (a normal-map image, or a constant texel
// other than the flat (0.5, 0.5, 1)), the full record and frame are taken.
__device__ __forceinline__ void guide_hit_features(const DevScene& sc, V3 o, V3 d, const Hit& h, bool frame, float* in) {
    const float4* sr = sc.shade + 6 * (size_t)h.tri;
    const float4 r0 = sr[0], r1 = sr[1], r2 = sr[2], r3 = sr[3];
    const int matid = __float_as_int(r0.w);
    bool full = false;
    if (frame) {
        const float* m = sc.mats + 16 * matid;
        full = sc.mtex[matid].w >= 0 || (m[12] != 0.0f && !(m[9] == 0.5f && m[10] == 0.5f && m[11] == 1.0f));
    }
    V3 p, n;
    if (full) {
        Isect it;
        int mid;
        fill_isect(sc, o, d, h, it, mid);
        first_hit_frame(hit_mat(sc, mid, h), it);
        p = it.hitpoint;
        n = it.normal;
    } else {
        const float al = 1.0f - (h.gamma + h.beta);
        const V3 nn = {r1.x * al + r1.w * h.beta + r2.z * h.gamma, r1.y * al + r2.x * h.beta + r2.w * h.gamma,
                       r1.z * al + r2.y * h.beta + r3.x * h.gamma};
        n = mulf(nn, rsqrtf(dot(nn, nn)));
        p = add(o, mulf(d, h.t));
    }
    guide::features_fast(sc.bmin, sc.bext, p.x, p.y, p.z, n.x, n.y, n.z, -d.x, -d.y, -d.z, in);
}

// The guide's inputs for queue entry e (zeros for a miss)
__device__ __forceinline__ void guide_entry_features(const WfArgs& a, int cur, unsigned e, float* in) {
    for (int k = 0; k < guide::NIN; k++) in[k] = 0.0f;
    const float4 HH = a.hits[e];
    Hit h;
    h.t = HH.x;
    h.beta = HH.y;
    h.gamma = HH.z;
    h.tri = __float_as_int(HH.w);
    if (h.tri < 0) return;
    const float4 O = a.qo[cur][e], D = a.qd[cur][e];
    const V3 o = v3(O.x, O.y, O.z), d = v3(D.x, D.y, D.z);
    const int slot = __float_as_int(D.w);
    const bool frame = O.w == CPT_FRESH_TMAX || (a.meta[slot] & 0xffu) == 0 || a.corrected;
    guide_hit_features(a.sc, o, d, h, frame, in);
}

// Per-bounce BF16 inference.  WIDTH: 64 (namespace guide) or 128 (guide128).
template <int WIDTH>
__global__ __launch_bounds__(256) void k_guide_infer(WfArgs a, int cur) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int NOUT = WIDTH == 128 ? guide128::NOUT : guide::NOUT;
    __shared__ __align__(16) float o_tile[guide::TILE * NOUT];
    // features of a group of 4 tiles, one sample per thread: the hit record,
    // material and feature maths are latency-bound, so all four waves take part
    __shared__ float in_s[4 * guide::TILE * guide::NIN];
    constexpr unsigned GROUP = 4 * guide::TILE;
    const unsigned count = a.counters[2 * cur];
    if (blockIdx.x * GROUP >= count) return;
    if constexpr (WIDTH == 128) guide128::load_weights(*reinterpret_cast<guide128::SmemFwd*>(smem_raw), a.gwt, a.gparams);
    else guide::load_weights(*reinterpret_cast<guide::SmemFwd*>(smem_raw), a.gwt, a.gparams);
    for (unsigned g = blockIdx.x; g * GROUP < count; g += gridDim.x) {
        const unsigned eg = g * GROUP;
        const int ng = (int)min(GROUP, count - eg);
        {
            float* in = in_s + threadIdx.x * guide::NIN;
            if ((int)threadIdx.x < ng) guide_entry_features(a, cur, eg + threadIdx.x, in);
            else
                for (int k = 0; k < guide::NIN; k++) in[k] = 0.0f;
        }
        __syncthreads();
        for (int q = 0; q * guide::TILE < ng; q++) {
            const unsigned e0 = eg + q * guide::TILE;
            const int nt = min(guide::TILE, ng - q * guide::TILE);
            const float* in = in_s + q * guide::TILE * guide::NIN;
            float* dst = a.lobes + e0;
            if constexpr (WIDTH == 128)
                guide128::infer_tile(*reinterpret_cast<guide128::SmemFwd*>(smem_raw), a.gwt, o_tile, in, nt, dst,
                                     a.lobes_stride);
            else guide::infer_tile(*reinterpret_cast<guide::SmemFwd*>(smem_raw), a.gwt, o_tile, in, nt, dst,
                                   a.lobes_stride);
        }
    }
}

// Per-bounce FP8 inference (the e4m3 network of k_quant8): one copy of the weights
// per block, then every wave on its own chunks of 64 queue entries: one entry's
// inputs per lane, then four tiles of 16 (N::wave_infer8), with no block barrier.
template <class N>
__global__ __launch_bounds__(64 * N::kWaves8) void k_guide_infer8(WfArgs a, int cur) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    auto& sm = *reinterpret_cast<typename N::Smem8*>(smem_raw);
    const unsigned count = a.counters[2 * cur];
    constexpr unsigned W8 = N::kWaves8;
    if (blockIdx.x * W8 * 64 >= count) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    auto& ws = sm.w[wave];
    N::load8(sm, a.gwt8, a.gparams, a.ginv8);
    __syncthreads();
    for (unsigned c = blockIdx.x * W8 + wave; c * 64 < count; c += gridDim.x * W8) {
        const unsigned e0 = c * 64;
        const int nin = (int)min(64u, count - e0);
        if (lane < nin) guide_entry_features(a, cur, e0 + lane, ws.in + lane * guide::NIN);
        else
            for (int k = 0; k < guide::NIN; k++) ws.in[lane * guide::NIN + k] = 0.0f;
        N::sync8();
        N::infer8(sm, ws, nin, lane, a.lobes + e0, a.lobes_stride);
    }
}

// Per-bounce BF16 inference at width 64 (weights in LDS), structured as
// k_guide_infer8: one wave per chunk of 64 queue entries, no block barrier after
// the weight load (N::infer16).
template <class N>
__global__ __launch_bounds__(64 * N::kWaves16) void k_guide_infer16(WfArgs a, int cur) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    auto& sm = *reinterpret_cast<typename N::Smem16*>(smem_raw);
    const unsigned count = a.counters[2 * cur];
    constexpr unsigned W16 = N::kWaves16;
    if (blockIdx.x * W16 * 64 >= count) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    auto& ws = sm.w[wave];
    N::load16(sm, a.gwt, a.gparams);
    __syncthreads();
    for (unsigned c = blockIdx.x * W16 + wave; c * 64 < count; c += gridDim.x * W16) {
        const unsigned e0 = c * 64;
        const int nin = (int)min(64u, count - e0);
        if (lane < nin) guide_entry_features(a, cur, e0 + lane, ws.in + lane * guide::NIN);
        else
            for (int k = 0; k < guide::NIN; k++) ws.in[lane * guide::NIN + k] = 0.0f;
        N::sync8();
        N::infer16(sm, ws, nin, lane, a.lobes + e0, a.lobes_stride);
    }
}

// After a pass: every complete record of a recording slot becomes a training
// sample (inputs, direction, weight = luminance of the radiance that came back
// along the direction / the pdf it was sampled with).
__global__ void k_guide_collect(WfArgs a) {
    const unsigned n_slots = min(a.n_rec, (a.pass->n_paths + a.rec_stride - 1) / a.rec_stride);
    const unsigned total = n_slots * (unsigned)a.rec_depth;
    for (unsigned base = (gtid() >> 6) * 64; base < total; base += (gthreads() >> 6) * 64) {
        unsigned idx = base + __lane_id();
        bool emit = false;
        float smp[guide::SAMPLE_FLOATS];
        if (idx < total && a.rec_flag[idx] == 2) {
            const float* r = a.rec + (size_t)idx * REC_FLOATS;
            const unsigned slot = idx / (unsigned)a.rec_depth * a.rec_stride;
            float4 c = a.col[slot];
            // radiance gathered through the sampled direction divided by the
            // throughput after the vertex is L_i; divided by the throughput
            // before it, L_i f cos / p.  Weight = target / sampling density.
            const int m = a.guide_target ? 17 : 11;
            float lr = r[m] > 0.0f ? (c.x - r[14]) / r[m] : 0.0f;
            float lg = r[m + 1] > 0.0f ? (c.y - r[15]) / r[m + 1] : 0.0f;
            float lb = r[m + 2] > 0.0f ? (c.z - r[16]) / r[m + 2] : 0.0f;
            float w = fmaxf(0.2126f * lr + 0.7152f * lg + 0.0722f * lb, 0.0f);
            if (a.guide_target) w = r[10] > 0.0f ? w / r[10] : 0.0f;
            if (w > 0.0f && isfinite(w)) {
                emit = true;
                for (int k = 0; k < 10; k++) smp[k] = r[k];
                smp[10] = w;
            }
        }
        unsigned q = wave_alloc(emit, a.gcount);
        if (emit)
            for (int k = 0; k < guide::SAMPLE_FLOATS; k++) a.gsamples[(size_t)q * guide::SAMPLE_FLOATS + k] = smp[k];
    }
}

// ------------------------------------------------------------------ shade
// pathtracerMIS.h:37-126, one loop iteration per queue entry.
template <int B, bool STREAM, bool GUIDED>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(GUIDED ? 1 : CPT_SHADE_WAVES))) void k_shade(
    WfArgs a, int cur, int depth_cap) {
    __shared__ unsigned qsh[2][2 * (B / 64) + 2];
    const unsigned count = a.counters[2 * cur];
    unsigned long long* const out_ctr = reinterpret_cast<unsigned long long*>(a.counters + 2 * (cur ^ 1));
    int par = 0;
    // block-uniform trip count: block_alloc2 holds barriers
    for (unsigned base = blockIdx.x * B; base < count; base += gridDim.x * B, par ^= 1) {
        const unsigned i = base + threadIdx.x;
        const bool active = i < count;
        bool push_ext = false, push_sh = false;
        V3 no = {0, 0, 0}, nd = {0, 0, 0}, sho = {0, 0, 0}, shd = {0, 0, 0}, contrib = {0, 0, 0};
        float shmax = 0.0f;
        int slot = 0;
        if (active) {
            float4 O = a.qo[cur][i], D = a.qd[cur][i], HH = a.hits[i];
            slot = __float_as_int(D.w);
            V3 o = v3(O.x, O.y, O.z), d = v3(D.x, D.y, D.z);
            // a camera ray's path state is the initial one (depth 0, sampled_delta)
            float4 MP = make_float4(1.0f, 1.0f, 1.0f, 1.0f), C = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            uint32_t meta = 1u << 8;
            if (O.w != CPT_FRESH_TMAX) {
                MP = a.mul_p[slot];
                C = a.col[slot];
                meta = a.meta[slot];
            }
            uint64_t rng = a.rng[slot];
            int depth = (int)(meta & 0xffu);
            bool sampled_delta = (meta >> 8) & 1u;
            V3 mulv = v3(MP.x, MP.y, MP.z), colv = v3(C.x, C.y, C.z);
            float p = MP.w;
            Hit h;
            h.t = HH.x;
            h.beta = HH.y;
            h.gamma = HH.z;
            h.tri = __float_as_int(HH.w);
            bool alive = false;
            float* rec = nullptr;
            uint8_t* rflag = nullptr;
            unsigned ri;
            if (GUIDED && a.n_rec && rec_index(a, (unsigned)slot, ri)) {
                rec = a.rec + (size_t)ri * a.rec_depth * REC_FLOATS;
                rflag = a.rec_flag + (size_t)ri * a.rec_depth;
                uint8_t* fl = rflag;
                if (depth > 0 && fl[depth - 1] == 1) {
                    float* r = rec + (depth - 1) * REC_FLOATS;
                    r[14] = colv.x;
                    r[15] = colv.y;
                    r[16] = colv.z;
                    fl[depth - 1] = 2;
                }
            }
            if (h.tri < 0) {
                // render_sky (pathtracerMIS.h:38-40), constant environment
                colv = add(colv, mul(mulv, v3(a.sc.env[0], a.sc.env[1], a.sc.env[2])));
            } else {
                Isect it;
                int matid;
                fill_isect(a.sc, o, d, h, it, matid);
                Mat m = hit_mat(a.sc, matid, h);
                if (depth == 0 || a.corrected) first_hit_frame(m, it);
                V3 wo = mulf(d, -1.0f);
                if (m.type == CPT_MAT_EMISSIVE) {
                    if (a.integrator == 2 || sampled_delta) {
                        // pure_pt.h:43-46 counts every hit, pathtracer.h:44-48 and
                        // pathtracerMIS.h:53-57 those after delta or camera bounces
                        colv = add(colv, mul(mulv, m.emission));
                    } else if (a.integrator == 1) {
                        // pathtracer.h: NEE already counted this light
                    } else {
                        float pbb = p;
                        float r = h.t;
                        float cos_theta = (float)(fabs((double)dot(it.normal, d)) + 1e-4);
                        float pbn = light_pdf(a.sc, h.tri) * r * r / cos_theta;
                        float w = pbb / (pbb + pbn);
                        colv = add(colv, mulf(mul(mulv, m.emission), w));
                    }
                } else {
                    sampled_delta = false;
                    const float* lob = GUIDED ? a.lobes + i : nullptr;
                    const bool gv = GUIDED && depth < a.guide_depth && guidable(m);
                    const float alpha = a.pass->guide_frac;
                    DSample ds;
                    if (gv) {
                        // one-sample MIS of guide and BSDF (balance heuristic, ppg.h:74-103)
                        float u = pcg_float(rng);
                        if (u < alpha) {
                            float u1 = pcg_float(rng);
                            float u2 = pcg_float(rng);
                            ds.wi = gmix_sample_n(lob, a.lobes_stride, a.guide_lobes, u / alpha, u1, u2);
                            ds.delta = false;
                            ds.pdf = bsdf_pdf(m, it, wo, ds.wi);
                        } else {
                            ds = bsdf_sample(m, it, wo, rng);
                        }
                    } else {
                        ds = bsdf_sample(m, it, wo, rng);
                    }
                    // A guided vertex evaluates the mixture at the sampled direction and
                    // at NEE's light direction in one pass over its lobes (below).
                    float gq_wi = 0.0f;
                    bool have_gq = false;
                    if (!ds.delta) {
                        if (a.sc.n_lights > 0 && a.integrator != 2) {
                            LSample ls = light_sample(a.sc, rng, it.hitpoint);
                            V3 sdir = sub(ls.position, it.hitpoint);
                            float len = len3(sdir);
                            sdir = divs(sdir, len);
                            if (bsdf_non_zero(m, it, wo, sdir)) {
                                float pnb = bsdf_pdf(m, it, wo, sdir);
                                if (gv) {
                                    float gq_l;
                                    gmix_pdf2_n(lob, a.lobes_stride, a.guide_lobes, ds.wi, sdir, gq_wi, gq_l);
                                    have_gq = true;
                                    pnb = alpha * gq_l + (1.0f - alpha) * pnb;
                                }
                                float ct = (float)(fabs((double)dot(ls.normal, sdir)) + 1e-4);
                                float pnn = ls.pdf * len * len / ct;
                                float w = pnn / (pnb + pnn);
                                if (a.corrected && ls.delta) w = 1.0f;  // no BSDF sample reaches a point
                                V3 ev = bsdf_eval(m, it, wo, sdir);
                                if (a.integrator == 1) w = 1.0f;  // pathtracer.h:70-76: no MIS weight
                                contrib = mulf(divs(mul(mul(mulv, ev), ls.intensity), ls.pdf * len * len), w);
                                sho = add(it.hitpoint, mulf(sdir, CPT_EPSF));
                                shd = sdir;
                                shmax = (float)((double)len - 2.0 * 1e-4);
                                push_sh = true;
                            }
                        }
                    } else {
                        sampled_delta = true;
                    }
                    if (gv) {
                        if (!have_gq) gq_wi = gmix_pdf_n(lob, a.lobes_stride, a.guide_lobes, ds.wi);
                        ds.pdf = alpha * gq_wi + (1.0f - alpha) * ds.pdf;
                    }
                    p = ds.pdf;
                    const float p_sampled = p;
                    if (bsdf_non_zero(m, it, wo, ds.wi)) {
                        V3 ev = bsdf_eval(m, it, wo, ds.wi);
                        bool survive = true;
                        if (depth > 2) {
                            float rr = rr_prob(ev, a.corrected);
                            if (pcg_float(rng) > rr) survive = false;
                            else p = p * rr;
                        }
                        if (survive) {
                            const V3 mul_prev = mulv;
                            mulv = divs(mul(mulv, ev), p);
                            // the next emitter hit's MIS weight: the reference keeps the
                            // pdf after Russian roulette, the corrected mode the BSDF's own
                            if (a.corrected) p = p_sampled;
                            if (depth + 1 < depth_cap) {
                                no = add(it.hitpoint, mulf(ds.wi, CPT_EPSF));
                                nd = ds.wi;
                                alive = true;
                                if (GUIDED && rec && gv) {
                                    float* r = rec + depth * REC_FLOATS;
                                    guide::features(a.sc.bmin, a.sc.bext, it.hitpoint.x, it.hitpoint.y, it.hitpoint.z,
                                                    it.normal.x, it.normal.y, it.normal.z, wo.x, wo.y, wo.z, r);
                                    r[7] = ds.wi.x;
                                    r[8] = ds.wi.y;
                                    r[9] = ds.wi.z;
                                    r[10] = p_sampled;
                                    r[11] = mul_prev.x;
                                    r[12] = mul_prev.y;
                                    r[13] = mul_prev.z;
                                    r[17] = mulv.x;
                                    r[18] = mulv.y;
                                    r[19] = mulv.z;
                                    rflag[depth] = 1;
                                }
                            }
                        }
                    }
                }
            }
            a.col[slot] = make_float4(colv.x, colv.y, colv.z, 0.0f);
            if (alive || push_sh) {
                a.mul_p[slot] = make_float4(mulv.x, mulv.y, mulv.z, p);
                a.rng[slot] = rng;
                a.meta[slot] = (uint32_t)(depth + 1) | ((sampled_delta ? 1u : 0u) << 8);
            }
            push_ext = alive;
        }
        unsigned qe, qs;
        block_alloc2<B>(push_ext, push_sh, out_ctr, qsh[par], qe, qs);
        if (push_ext) {
            a.qo[cur ^ 1][qe] = make_float4(no.x, no.y, no.z, INFINITY);
            a.qd[cur ^ 1][qe] = make_float4(nd.x, nd.y, nd.z, __int_as_float(slot));
        }
        if (STREAM) {
            bool ended = active && !push_ext;
            unsigned qd = wave_alloc(ended, &a.counters[4]);
            if (ended) a.done[qd] = (uint32_t)slot;
        }
        if (push_sh) {
            a.so[qs] = make_float4(sho.x, sho.y, sho.z, shmax);
            a.sd[qs] = make_float4(shd.x, shd.y, shd.z, __int_as_float(slot));
            a.sc_rgb[qs] = make_float4(contrib.x, contrib.y, contrib.z, 0.0f);
        }
    }
}

// ------------------------------------------------------------------ light tracer
// Lighttracer::integrate (lighttracer.h:37-110), one loop iteration per queue
// entry: emitters end the path; after a non-delta BSDF sample the vertex is
// connected to the camera (shadow ray from the hit point itself, to the camera
// point, contribution mul * f(normalized wo, normalized wi) * (1/spp) * the
// camera factor, splatted by k_shadow); Russian roulette from depth 3 on; the
// next ray leaves from the hit point itself (the reference offsets neither ray).
template <int B>
__global__ __launch_bounds__(B) void k_shade_lt(WfArgs a, int cur, int depth_cap) {
    __shared__ unsigned qsh[2][2 * (B / 64) + 2];
    const unsigned count = a.counters[2 * cur];
    unsigned long long* const out_ctr = reinterpret_cast<unsigned long long*>(a.counters + 2 * (cur ^ 1));
    const ViewDesc& vd = a.views[0];
    const float factor = (float)(1.0 / (double)(vd.S * vd.S));
    int par = 0;
    for (unsigned base = blockIdx.x * B; base < count; base += gridDim.x * B, par ^= 1) {
        const unsigned i = base + threadIdx.x;
        const bool active = i < count;
        bool push_ext = false, push_sh = false;
        V3 nd = {0, 0, 0}, ho = {0, 0, 0}, shd = {0, 0, 0}, contrib = {0, 0, 0};
        float shmax = 0.0f;
        int slot = 0, pix = 0;
        if (active) {
            float4 O = a.qo[cur][i], D = a.qd[cur][i], HH = a.hits[i];
            slot = __float_as_int(D.w);
            V3 o = v3(O.x, O.y, O.z), d = v3(D.x, D.y, D.z);
            float4 MP = a.mul_p[slot];
            uint64_t rng = a.rng[slot];
            const int depth = (int)(a.meta[slot] & 0xffu);
            V3 mulv = v3(MP.x, MP.y, MP.z);
            Hit h;
            h.t = HH.x;
            h.beta = HH.y;
            h.gamma = HH.z;
            h.tri = __float_as_int(HH.w);
            if (h.tri >= 0) {
                Isect it;
                int matid;
                fill_isect(a.sc, o, d, h, it, matid);
                Mat m = hit_mat(a.sc, matid, h);
                if (depth == 0 || a.corrected) first_hit_frame(m, it);
                if (m.type != CPT_MAT_EMISSIVE) {
                    const V3 wo = mulf(d, -1.0f);
                    DSample ds = bsdf_sample(m, it, wo, rng);
                    const V3 ev = bsdf_eval(m, it, wo, ds.wi);
                    float p = ds.pdf;
                    if (!ds.delta) {
                        int ci, cj;
                        float cf;
                        camera_connect(vd.cam, vd.W, vd.H, it.hitpoint, ci, cj, cf, a.corrected != 0);
                        const V3 cp = v3(vd.cam[0], vd.cam[1], vd.cam[2]);
                        const V3 sdir = normalized(sub(cp, it.hitpoint));
                        if (bsdf_non_zero(m, it, wo, sdir) && ci > -1) {
                            const V3 f = bsdf_eval(m, it, normalized(wo), normalized(sdir));
                            contrib = mulf(mulf(mul(mulv, f), factor), cf);
                            ho = it.hitpoint;
                            shd = sdir;
                            shmax = len3(sub(cp, it.hitpoint));
                            pix = (int)(vd.film_off + (unsigned long long)cj * (unsigned)vd.W + (unsigned)ci);
                            push_sh = true;
                        }
                    }
                    bool survive = true;
                    if (depth > 2) {
                        const float rr = rr_prob(ev, a.corrected);
                        if (pcg_float(rng) > rr) survive = false;
                        else p = p * rr;
                    }
                    if (survive && depth + 1 < depth_cap) {
                        mulv = divs(mul(mulv, ev), p);
                        ho = it.hitpoint;
                        nd = ds.wi;
                        push_ext = true;
                    }
                }
            }
            if (push_ext) {
                a.mul_p[slot] = make_float4(mulv.x, mulv.y, mulv.z, 1.0f);
                a.rng[slot] = rng;
                a.meta[slot] = (uint32_t)(depth + 1);
            }
        }
        unsigned qe, qs;
        block_alloc2<B>(push_ext, push_sh, out_ctr, qsh[par], qe, qs);
        if (push_ext) {
            a.qo[cur ^ 1][qe] = make_float4(ho.x, ho.y, ho.z, INFINITY);
            a.qd[cur ^ 1][qe] = make_float4(nd.x, nd.y, nd.z, __int_as_float(slot));
        }
        if (push_sh) {
            a.so[qs] = make_float4(ho.x, ho.y, ho.z, shmax);
            a.sd[qs] = make_float4(shd.x, shd.y, shd.z, __int_as_float(pix));
            a.sc_rgb[qs] = make_float4(contrib.x, contrib.y, contrib.z, 0.0f);
        }
    }
}

// ------------------------------------------------------------------ SD-tree guided arm
// PPG::integrate (ppg.h:41-143), one loop iteration per queue entry: no NEE;
// with probability 1/2 the BSDF, else the SD-tree; one-sample MIS weight
// p_s / (p_s + p_other) over p = p_s / 2.  Vertex d-1 (the ray that reached this
// hit) records its origin, direction and radiance factor: Le or the sky if the
// ray ended there, eval * weight / p if the path went on, 1 if it was cut by the
// BSDF support or Russian roulette (as the reference leaves it).
template <int B>
__global__ __launch_bounds__(B) void k_shade_ppg(WfArgs a, int cur, int depth_cap) {
    __shared__ unsigned qsh[2][2 * (B / 64) + 2];
    const unsigned count = a.counters[2 * cur];
    unsigned long long* const out_ctr = reinterpret_cast<unsigned long long*>(a.counters + 2 * (cur ^ 1));
    int par = 0;
    for (unsigned base = blockIdx.x * B; base < count; base += gridDim.x * B, par ^= 1) {
        const unsigned i = base + threadIdx.x;
        bool push_ext = false;
        V3 no = {0, 0, 0}, nd = {0, 0, 0};
        int slot = 0;
        if (i < count) {
            float4 O = a.qo[cur][i], D = a.qd[cur][i], HH = a.hits[i];
            slot = __float_as_int(D.w);
            V3 o = v3(O.x, O.y, O.z), d = v3(D.x, D.y, D.z);
            const bool fresh = O.w == CPT_FRESH_TMAX;  // a camera ray: initial path state
            const float4 MP = fresh ? make_float4(1.0f, 1.0f, 1.0f, 1.0f) : a.mul_p[slot];
            const float4 C = fresh ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : a.col[slot];
            uint64_t rng = a.rng[slot];
            const int depth = fresh ? 0 : (int)(a.meta[slot] & 0xffu);
            V3 mulv = v3(MP.x, MP.y, MP.z), colv = v3(C.x, C.y, C.z);
            Hit h;
            h.t = HH.x;
            h.beta = HH.y;
            h.gamma = HH.z;
            h.tri = __float_as_int(HH.w);
            V3 vrad = v3s(1.0f);
            bool alive = false;
            if (h.tri < 0) {
                V3 sky = v3(a.sc.env[0], a.sc.env[1], a.sc.env[2]);
                colv = add(colv, mul(mulv, sky));
                vrad = sky;
            } else {
                Isect it;
                int matid;
                fill_isect(a.sc, o, d, h, it, matid);
                Mat m = hit_mat(a.sc, matid, h);
                if (depth == 0 || a.corrected) first_hit_frame(m, it);
                V3 wo = mulf(d, -1.0f);
                if (m.type == CPT_MAT_EMISSIVE) {
                    colv = add(colv, mul(mulv, m.emission));
                    vrad = m.emission;
                } else {
                    DSample ds;
                    float weight = 1.0f;
                    bool ok = true;
                    if (pcg_float(rng) > 0.5f) {
                        ds = bsdf_sample(m, it, wo, rng);
                        weight = ds.pdf / (ds.pdf + sd_pdf(a.sdt, it.hitpoint, ds.wi));
                    } else {
                        ds.wi = sd_sample(a.sdt, it.hitpoint, rng, ds.pdf);
                        ds.delta = false;
                        if (!bsdf_non_zero(m, it, wo, ds.wi)) ok = false;
                        else weight = ds.pdf / (ds.pdf + bsdf_pdf(m, it, wo, ds.wi));
                    }
                    if (ok && bsdf_non_zero(m, it, wo, ds.wi)) {
                        float p = ds.pdf * 0.5f;
                        V3 ev = bsdf_eval(m, it, wo, ds.wi);
                        bool survive = true;
                        if (depth > 2) {
                            float rr = rr_prob(ev, a.corrected);
                            if (pcg_float(rng) > rr) survive = false;
                            else p = p * rr;
                        }
                        if (survive) {
                            V3 nm = mulf(ev, weight / p);
                            vrad = nm;
                            mulv = mul(mulv, nm);
                            if (depth + 1 < depth_cap) {
                                no = add(it.hitpoint, mulf(ds.wi, CPT_EPSF));
                                nd = ds.wi;
                                alive = true;
                            }
                        }
                    }
                }
            }
            if (depth > 0) {
                float* v = a.ppg_vtx + ((size_t)slot * a.max_depth + (size_t)(depth - 1)) * 9;
                v[0] = o.x;
                v[1] = o.y;
                v[2] = o.z;
                v[3] = d.x;
                v[4] = d.y;
                v[5] = d.z;
                v[6] = vrad.x;
                v[7] = vrad.y;
                v[8] = vrad.z;
            }
            a.col[slot] = make_float4(colv.x, colv.y, colv.z, 0.0f);
            if (alive) {
                a.mul_p[slot] = make_float4(mulv.x, mulv.y, mulv.z, 1.0f);
                a.rng[slot] = rng;
                a.meta[slot] = (uint32_t)(depth + 1);
            } else {
                a.ppg_nv[slot] = (uint32_t)depth;  // vertices of the ended path
            }
            push_ext = alive;
        }
        unsigned qe, qs;
        block_alloc2<B>(push_ext, false, out_ctr, qsh[par], qe, qs);
        if (push_ext) {
            a.qo[cur ^ 1][qe] = make_float4(no.x, no.y, no.z, INFINITY);
            a.qd[cur ^ 1][qe] = make_float4(nd.x, nd.y, nd.z, __int_as_float(slot));
        }
    }
}

// After a pass: every path whose last vertex carries radiance splats each vertex
// with the product of the factors from it to the end (ppg.h:130-141), its
// largest channel (rF, ppg.h:37-39).
__global__ void k_sd_splat(WfArgs a, unsigned n_paths) {
    for (unsigned slot = gtid(); slot < n_paths; slot += gthreads()) {
        const int nv = (int)a.ppg_nv[slot];
        if (nv <= 0) continue;
        const float* vb = a.ppg_vtx + (size_t)slot * a.max_depth * 9;
        const float* last = vb + (size_t)(nv - 1) * 9;
        if (!(last[6] > 0.0f || last[7] > 0.0f || last[8] > 0.0f)) continue;
        V3 acc = v3s(1.0f);
        for (int k = nv - 1; k >= 0; k--) {
            const float* v = vb + (size_t)k * 9;
            V3 r = v3(v[6], v[7], v[8]);
            if (k < nv - 1) r = mul(r, acc);
            acc = r;
            sd_splat(a.sdt, fmaxf(r.x, fmaxf(r.y, r.z)), v3(v[0], v[1], v[2]), v3(v[3], v[4], v[5]));
        }
    }
}

// ------------------------------------------------------------------ shadow
template <int B, bool COUNT>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(CPT_SHADOW_WAVES))) void k_shadow(WfArgs a, int nxt) {
    __shared__ int lds[CPT_LDS_ROWS * B];
    const unsigned count = a.counters[2 * nxt + 1];
    if (gtid() == 0) atomicAdd(&a.stats[ST_SHADOW], (unsigned long long)count);
    TravStack<B> st{lds, a.gstack + gtid(), (size_t)gthreads(), (int)threadIdx.x};
    unsigned nv = 0, nt = 0;
    trace_chunked<B, true, COUNT>(
        a.sc, count, a.refill, st,
        [&](unsigned i, V3& o, V3& d, float& tmax) {
            float4 O = a.so[i], D = a.sd[i];
            o = v3(O.x, O.y, O.z);
            d = v3(D.x, D.y, D.z);
            tmax = O.w;
        },
        [&](unsigned i, const Hit& h) {
            if (h.tri < 0) {
                int slot = __float_as_int(a.sd[i].w);
                float4 add_ = a.sc_rgb[i];
                if (a.integrator == 4) {
                    // light tracing: the camera sees the vertex; splat into the film
                    // pixel (the reference adds to RgbImage::get_pixel, lighttracer.h:77-91).
                    // The film holds spp times the image, as for camera paths.
                    const double sc = (double)a.views[0].spp;
                    double* f = a.film + 3 * (size_t)(unsigned)slot;
                    unsafeAtomicAdd(f, (double)add_.x * sc);
                    unsafeAtomicAdd(f + 1, (double)add_.y * sc);
                    unsafeAtomicAdd(f + 2, (double)add_.z * sc);
                } else {
                    float4 c = a.col[slot];
                    a.col[slot] = make_float4(c.x + add_.x, c.y + add_.y, c.z + add_.z, 0.0f);
                }
            }
        },
        nv, nt);
    if (COUNT) {
        nv = wave_sum(nv);
        nt = wave_sum(nt);
        if (__lane_id() == 0) {
            atomicAdd(&a.stats[ST_SNODES], (unsigned long long)nv);
            atomicAdd(&a.stats[ST_STRIS], (unsigned long long)nt);
        }
    }
}

// ------------------------------------------------------------------ film
__global__ void k_accum(WfArgs a) {
    const PassDesc* pd = a.pass;
    for (unsigned w = gtid(); w < pd->n_pix; w += gthreads()) {
        const Segment sg = pd->seg[find_segment(pd, w, true)];
        unsigned pl = w - sg.pix0;
        size_t fp = (size_t)a.views[sg.view].film_off + sg.p0 + pl;
        double* f = a.film + 3 * fp;
        double* f2 = a.film2 + 3 * fp;
        double x = f[0], y = f[1], z = f[2];
        double x2 = f2[0], y2 = f2[1], z2 = f2[2];
        for (unsigned s = 0; s < sg.ns; s++) {
            float4 c = a.col[(size_t)sg.slot0 + (size_t)s * sg.npix + pl];
            x += (double)c.x;
            y += (double)c.y;
            z += (double)c.z;
            x2 += (double)c.x * (double)c.x;
            y2 += (double)c.y * (double)c.y;
            z2 += (double)c.z * (double)c.z;
        }
        f[0] = x;
        f[1] = y;
        f[2] = z;
        f2[0] = x2;
        f2[1] = y2;
        f2[2] = z2;
    }
}

// timestamp (100 MHz constant clock) after the pass's last kernel
__global__ void k_stamp(WfArgs a, int fixed_slot) {
    if (gtid() == 0) a.pass_tend[fixed_slot >= 0 ? (unsigned)fixed_slot : a.pass->pass_id] = __builtin_amdgcn_s_memrealtime();
}

// ------------------------------------------------------------------ streaming (multi-view) mode
// Paths of all views form one index space [0, total): view v owns
// [path_begin, path_begin + W*H*S*S), sample-major within the view.  A slot
// whose path ended is refilled with the next path index by k_regen, so the
// queues stay full until the last path of the last view has been started.
struct StreamArgs {
    const ViewDesc* views;
    int n_views;
    unsigned long long total;
    unsigned long long* next_path;     // global path counter
    unsigned long long* view_done;     // accumulated paths per view
    unsigned long long* view_tend;     // s_memrealtime when the view completed
    unsigned long long* t_start;
};

#define CPT_NO_PATH 0xffffffffu

__global__ void k_stream_init(WfArgs a, StreamArgs sa, unsigned cap) {
    if (gtid() == 0) {
        *sa.next_path = 0;
        *sa.t_start = __builtin_amdgcn_s_memrealtime();
        a.counters[0] = a.counters[1] = a.counters[2] = a.counters[3] = 0;
        a.counters[4] = cap;
    }
    for (unsigned i = gtid(); i < cap; i += gthreads()) {
        a.done[i] = i;
        a.pix[i] = CPT_NO_PATH;
    }
    for (unsigned v = gtid(); v < (unsigned)sa.n_views; v += gthreads()) {
        sa.view_done[v] = 0;
        sa.view_tend[v] = 0;
    }
}

// Accumulate the radiance of ended paths (col already holds this bounce's NEE)
// and start new camera paths in their slots.
template <int B>
__global__ __launch_bounds__(B) void k_regen(WfArgs a, StreamArgs sa, int nxt) {
    const unsigned count = a.counters[4];
    const unsigned lane = __lane_id();
    const unsigned wave = gtid() >> 6, nwaves = gthreads() >> 6;
    for (unsigned base = wave * 64; base < count; base += nwaves * 64) {
        const unsigned i = base + lane;
        const bool active = i < count;
        unsigned slot = active ? a.done[i] : 0u;
        if (active) {
            unsigned fp = a.pix[slot];
            if (fp != CPT_NO_PATH) {
                float4 c = a.col[slot];
                double* f = a.film + 3 * (size_t)fp;
                double* f2 = a.film2 + 3 * (size_t)fp;
                unsafeAtomicAdd(f, (double)c.x);
                unsafeAtomicAdd(f + 1, (double)c.y);
                unsafeAtomicAdd(f + 2, (double)c.z);
                unsafeAtomicAdd(f2, (double)c.x * (double)c.x);
                unsafeAtomicAdd(f2 + 1, (double)c.y * (double)c.y);
                unsafeAtomicAdd(f2 + 2, (double)c.z * (double)c.z);
                unsigned v = a.vid[slot];
                unsigned long long old = atomicAdd(&sa.view_done[v], 1ull);
                if (old + 1 == sa.views[v].n_paths) sa.view_tend[v] = __builtin_amdgcn_s_memrealtime();
            }
        }
        // next path indices for the active lanes: one atomic per wave
        unsigned long long mask = __ballot(active);
        unsigned long long g0 = 0;
        if (mask) {
            int leader = __ffsll(mask) - 1;
            if ((int)lane == leader) g0 = atomicAdd(sa.next_path, (unsigned long long)__popcll(mask));
            g0 = __shfl(g0, leader);
        }
        unsigned long long lt = lane ? (mask & ((~0ull) >> (64 - lane))) : 0ull;
        unsigned long long g = g0 + (unsigned long long)__popcll(lt);
        bool start = active && g < sa.total;
        V3 o = {0, 0, 0}, d = {0, 0, 0};
        if (start) {
            int lo = 0, hi = sa.n_views - 1;
            while (lo < hi) {  // last view with path_begin <= g
                int mid = (lo + hi + 1) >> 1;
                if (sa.views[mid].path_begin <= g) lo = mid;
                else hi = mid - 1;
            }
            const ViewDesc& vd = sa.views[lo];
            unsigned long long local = g - vd.path_begin;
            unsigned long long P = (unsigned long long)vd.W * (unsigned long long)vd.H;
            unsigned sample = (unsigned)(local / P);
            unsigned pixel = (unsigned)(local - (unsigned long long)sample * P);
            int px = (int)(pixel % (unsigned)vd.W), py = (int)(pixel / (unsigned)vd.W);
            uint64_t path_id = ((uint64_t)vd.id << 40) + (uint64_t)pixel * (uint64_t)vd.spp + sample;
            uint64_t rng = pcg_init(a.seed, path_id);
            float r1 = pcg_float(rng);
            float r2 = pcg_float(rng);
            float cx, cy;
            view_coords(vd, px, py, sample, r1, r2, cx, cy);
            camera_ray(vd.cam, cx, cy, o, d);
            a.mul_p[slot] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
            a.col[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            a.rng[slot] = rng;
            a.meta[slot] = 1u << 8;
            a.pix[slot] = (unsigned)(vd.film_off + pixel);
            a.vid[slot] = (unsigned)lo;
        } else if (active) {
            a.pix[slot] = CPT_NO_PATH;
        }
        unsigned qe = wave_alloc(start, &a.counters[2 * nxt]);
        if (start) {
            a.qo[nxt][qe] = make_float4(o.x, o.y, o.z, INFINITY);
            a.qd[nxt][qe] = make_float4(d.x, d.y, d.z, __int_as_float((int)slot));
        }
    }
}

__global__ void k_fill_u64(unsigned long long* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// ------------------------------------------------------------------ ray-batch parity entries
template <int B>
__global__ __launch_bounds__(B) void k_trace_closest(DevScene sc, const float4* rays, int n, float4* hits, int* gstack,
                                                     unsigned refill) {
    __shared__ int lds[CPT_LDS_ROWS * B];
    TravStack<B> st{lds, gstack + gtid(), (size_t)gthreads(), (int)threadIdx.x};
    unsigned nv = 0, nt = 0;
    trace_chunked<B, false, false>(
        sc, (unsigned)n, refill, st,
        [&](unsigned i, V3& o, V3& d, float& tmax) {
            float4 O = rays[2 * i], D = rays[2 * i + 1];
            o = v3(O.x, O.y, O.z);
            d = v3(D.x, D.y, D.z);
            tmax = O.w;
        },
        [&](unsigned i, const Hit& h) { hits[i] = make_float4(h.t, h.beta, h.gamma, __int_as_float(h.tri)); }, nv, nt);
}

template <int B>
__global__ __launch_bounds__(B) void k_trace_any(DevScene sc, const float4* rays, int n, uint8_t* occ, int* gstack,
                                                 unsigned refill) {
    __shared__ int lds[CPT_LDS_ROWS * B];
    TravStack<B> st{lds, gstack + gtid(), (size_t)gthreads(), (int)threadIdx.x};
    unsigned nv = 0, nt = 0;
    trace_chunked<B, true, false>(
        sc, (unsigned)n, refill, st,
        [&](unsigned i, V3& o, V3& d, float& tmax) {
            float4 O = rays[2 * i], D = rays[2 * i + 1];
            o = v3(O.x, O.y, O.z);
            d = v3(D.x, D.y, D.z);
            tmax = O.w;
        },
        [&](unsigned i, const Hit& h) { occ[i] = h.tri >= 0 ? 1 : 0; }, nv, nt);
}

// Traversal overflow stacks of cpt_trace_closest_device_noalloc, which may not
// allocate: one file-scope array per GPU, sized for its fixed grid of
// NOALLOC_BLOCKS x 128 threads x CPT_GSTACK entries (52 MiB), held until the
// process ends.  Host code never uses this name as an address; it uses the
// per-device address from hipGetSymbolAddress (noalloc_scratch).
constexpr int NOALLOC_BLOCKS = 1024;
constexpr int NOALLOC_B = 128;
constexpr size_t NOALLOC_SCRATCH_INTS = (size_t)NOALLOC_BLOCKS * NOALLOC_B * CPT_GSTACK;
static __device__ __align__(256) int scratch[NOALLOC_SCRATCH_INTS];

}  // namespace

// ====================================================================== host side

struct cpt_ctx {
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // device address of this GPU's `scratch` array (cpt_trace_closest_device_noalloc)
    int* noalloc_scratch = nullptr;
    // fault injection for tests (cpt_debug_fail_alloc): device allocations
    // left before one fails; -1 = off
    int fail_alloc_in = -1;

    // scene
    bool has_scene = false;
    uint64_t scene_gen = 0;  // bumped by every upload
    // launch configuration chosen by cpt_autotune, valid for one key
    struct Tuned {
        bool valid = false;
        uint64_t scene_gen;
        int W, H, depth, guided;
        int B, G, SB, refill;
    } tuned;
    int nt = 0, nv = 0, nl = 0, nm = 0;
    int64_t bvh_nodes = 0, bvh_leaves = 0, bvh_build_us = 0;
    int bvh_depth = 0;
    bool bvh_gpu = false;
    std::vector<int32_t> rank;
    Node4* d_nodes = nullptr;
    QNode* d_qnodes = nullptr;  // the traversal's quantized tree (d_nodes: the build's float nodes, freed after)
    float4* d_tris = nullptr;
    int4* d_tri_info = nullptr;
    float4* d_shade = nullptr;
    float4 *d_vpos = nullptr, *d_vnrm = nullptr, *d_vbit = nullptr;
    float* d_mats = nullptr;
    int4* d_mtex = nullptr;     // per material: texture ids
    float2* d_uv = nullptr;     // per vertex texture coordinates
    int4* d_tex = nullptr;      // texture table
    float* d_texels = nullptr;
    float4* d_plights = nullptr;
    int* d_lights = nullptr;
    float cam[16];
    float env[3];
    float bmin[3], bext[3];
    float tb_lo[3], tb_hi[3];  // bounds of the triangles (the reference BVH's root box)

    // wavefront buffers
    size_t cap = 0;        // path slots
    size_t gcap = 0;       // threads covered by the overflow stack
    float4 *mul_p = nullptr, *col = nullptr, *qo[2] = {nullptr, nullptr}, *qd[2] = {nullptr, nullptr};
    float4 *hits = nullptr, *so = nullptr, *sd = nullptr, *sc_rgb = nullptr;
    uint64_t* rng = nullptr;
    uint32_t *meta = nullptr, *pix = nullptr, *vid = nullptr, *done = nullptr, *counters = nullptr;
    unsigned long long* stats = nullptr;
    int* gstack = nullptr;
    PassDesc* d_pass = nullptr;
    PassDesc* h_pass = nullptr;  // pinned
    size_t h_pass_cap = 0;
    ViewDesc* d_views = nullptr;
    size_t views_cap = 0;
    unsigned long long* pass_tend = nullptr;
    size_t tend_cap = 0;
    double* film = nullptr;
    double* film2 = nullptr;
    size_t film_px = 0;
    int film_w = 0, film_h = 0;
    unsigned film_spp = 1;  // samples per pixel of the last single-view render
    // (width, height) of every view in the film of the last completed render;
    // empty while the film holds nothing a render can add to (just allocated,
    // or a render failed part way)
    std::vector<std::pair<int, int>> film_layout;

    // learned guide
    bool has_guide = false;
    float* g_params = nullptr;   // FP32 master weights
    float* g_m = nullptr;        // Adam moments
    float* g_v = nullptr;
    float* g_grad = nullptr;
    uint16_t* g_wt = nullptr;    // transposed BF16 weights
    uint8_t* g_wt8 = nullptr;    // transposed FP8 weights (inference) and inverse layer scales
    float* g_inv8 = nullptr;
    float* g_samples = nullptr;  // training samples (11 floats each)
    size_t g_samples_cap = 0;
    uint16_t* g_rows = nullptr;  // per-sample activation / gradient rows
    size_t g_rows_cap = 0;
    float* g_partial = nullptr;  // per-block dW partials
    float* g_loss_part = nullptr;
    float* g_wscale = nullptr;
    float* g_wpart = nullptr;    // per-block partial sums of the sample weights
    int g_step = 0;
    cpt_guide_params g_cfg{};
    int g_width = 0;             // network of the allocated guide buffers: 64 or 128 (0: none)
    float* g_lobes = nullptr;    // per queue entry mixture (5 floats per lobe)
    size_t g_lobes_cap = 0;      // floats
    float* g_rec = nullptr;      // training records
    uint8_t* g_rec_flag = nullptr;
    size_t g_rec_cap = 0;
    uint32_t* g_count = nullptr;
    uint32_t* h_count = nullptr; // pinned
    int64_t g_samples_total = 0; // training samples used by the last guided render
    int g_steps_total = 0;

    // SD-tree guided arm: host tree, its flattened device copy, splat
    // accumulators and per-path vertex records
    std::unique_ptr<cpt::SdTree> sd_tree;
    cpt::SdFlat sd_flat;
    bool sd_dirty = true;  // host tree changed since the last upload
    int32_t *sd_snode = nullptr, *sd_quad = nullptr;
    int4* sd_dinfo = nullptr;
    float4* sd_rect = nullptr;
    float* sd_sum = nullptr;
    uint32_t* sd_cnt = nullptr;
    size_t sd_snode_cap = 0, sd_quad_cap = 0, sd_dinfo_cap = 0, sd_rect_cap = 0, sd_slot_cap = 0;
    float* ppg_vtx = nullptr;
    uint32_t* ppg_nv = nullptr;
    size_t ppg_cap = 0;  // floats of ppg_vtx

    // captured pass graph
    hipGraphExec_t gexec = nullptr;
    std::string graph_key;
};

namespace {

int fail(cpt_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(ctx, expr)                                                                           \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return fail(ctx, CPT_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

void free_scene(cpt_ctx* c) {
    dfree(c->d_nodes);
    dfree(c->d_qnodes);
    dfree(c->d_tris);
    dfree(c->d_tri_info);
    dfree(c->d_shade);
    dfree(c->d_vpos);
    dfree(c->d_vnrm);
    dfree(c->d_vbit);
    dfree(c->d_mats);
    dfree(c->d_mtex);
    dfree(c->d_uv);
    dfree(c->d_tex);
    dfree(c->d_texels);
    dfree(c->d_plights);
    dfree(c->d_lights);
    c->has_scene = false;
}

void free_wavefront(cpt_ctx* c) {
    dfree(c->mul_p);
    dfree(c->col);
    dfree(c->qo[0]);
    dfree(c->qo[1]);
    dfree(c->qd[0]);
    dfree(c->qd[1]);
    dfree(c->hits);
    dfree(c->so);
    dfree(c->sd);
    dfree(c->sc_rgb);
    dfree(c->rng);
    dfree(c->meta);
    dfree(c->pix);
    dfree(c->vid);
    dfree(c->done);
    c->cap = 0;
}

void drop_graph(cpt_ctx* c) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->graph_key.clear();
}

DevScene dev_scene(const cpt_ctx* c) {
    DevScene s;
    s.qnodes = c->d_qnodes;
    s.tris = c->d_tris;
    s.tri_info = c->d_tri_info;
    s.shade = c->d_shade;
    s.vpos = c->d_vpos;
    s.vnrm = c->d_vnrm;
    s.vbit = c->d_vbit;
    s.mats = c->d_mats;
    s.mtex = c->d_mtex;
    s.uv = c->d_uv;
    s.tex = c->d_tex;
    s.texels = c->d_texels;
    s.lights = c->d_lights;
    s.plights = c->d_plights;
    s.n_lights = c->nl;
    s.n_tris = c->nt;
    for (int k = 0; k < 3; k++) s.env[k] = c->env[k];
    for (int k = 0; k < 16; k++) s.cam[k] = c->cam[k];
    for (int k = 0; k < 3; k++) {
        s.bmin[k] = c->bmin[k];
        s.bext[k] = c->bext[k];
    }
    return s;
}

// Free p and allocate count elements.  On failure p is null: every caller
// keeps the capacity of a buffer group at 0 until all of the group's buffers
// are allocated, so a failed allocation is retried by the next call instead of
// leaving a stale capacity in front of a null or short buffer.
template <typename T>
int dalloc(cpt_ctx* c, T*& p, size_t count) {
    dfree(p);
    if (count == 0) count = 1;
    if (c->fail_alloc_in == 0) {
        c->fail_alloc_in = -1;
        return fail(c, CPT_E_HIP, "hipMalloc of %zu bytes failed (injected)", count * sizeof(T));
    }
    if (c->fail_alloc_in > 0) c->fail_alloc_in--;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
    if (e != hipSuccess) {
        p = nullptr;
        return fail(c, CPT_E_HIP, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
    }
    return CPT_OK;
}

int ensure_gstack(cpt_ctx* c, size_t threads) {
    if (c->gcap >= threads && c->gstack) return CPT_OK;
    drop_graph(c);
    c->gcap = 0;
    int r = dalloc(c, c->gstack, threads * (size_t)CPT_GSTACK);
    if (r) return r;
    c->gcap = threads;
    return CPT_OK;
}

int ensure_wavefront(cpt_ctx* c, size_t n) {
    if (c->cap >= n && c->mul_p) return CPT_OK;
    drop_graph(c);
    free_wavefront(c);
    int r = 0;
    r |= dalloc(c, c->mul_p, n);
    r |= dalloc(c, c->col, n);
    r |= dalloc(c, c->qo[0], n);
    r |= dalloc(c, c->qo[1], n);
    r |= dalloc(c, c->qd[0], n);
    r |= dalloc(c, c->qd[1], n);
    r |= dalloc(c, c->hits, n);
    r |= dalloc(c, c->so, n);
    r |= dalloc(c, c->sd, n);
    r |= dalloc(c, c->sc_rgb, n);
    r |= dalloc(c, c->rng, n);
    r |= dalloc(c, c->meta, n);
    r |= dalloc(c, c->pix, n);
    r |= dalloc(c, c->vid, n);
    r |= dalloc(c, c->done, n);
    if (r) {
        free_wavefront(c);
        return CPT_E_HIP;
    }
    c->cap = n;
    return CPT_OK;
}

// FP64 film sums and sums of squares for px pixels.  A new film holds nothing
// to add to, so its layout record is cleared.
int ensure_film(cpt_ctx* c, size_t px) {
    if (c->film_px >= px && c->film && c->film2) return CPT_OK;
    drop_graph(c);
    c->film_px = 0;
    c->film_layout.clear();
    int r = dalloc(c, c->film, 3 * px);
    if (!r) r = dalloc(c, c->film2, 3 * px);
    if (r) return r;
    c->film_px = px;
    return CPT_OK;
}

WfArgs make_args(cpt_ctx* c, int W, int H, int S, int depth, uint64_t seed, uint32_t view) {
    WfArgs a;
    a.sc = dev_scene(c);
    a.mul_p = c->mul_p;
    a.col = c->col;
    a.rng = c->rng;
    a.meta = c->meta;
    a.pix = c->pix;
    a.vid = c->vid;
    a.done = c->done;
    a.qo[0] = c->qo[0];
    a.qo[1] = c->qo[1];
    a.qd[0] = c->qd[0];
    a.qd[1] = c->qd[1];
    a.hits = c->hits;
    a.so = c->so;
    a.sd = c->sd;
    a.sc_rgb = c->sc_rgb;
    a.counters = c->counters;
    a.stats = c->stats;
    a.gstack = c->gstack;
    a.pass = c->d_pass;
    a.views = c->d_views;
    a.pass_tend = c->pass_tend;
    a.film = c->film;
    a.film2 = c->film2;
    a.lobes = c->g_lobes;
    a.lobes_stride = c->g_lobes_cap / (3 * (c->g_width == 128 ? 16 : 8));
    a.rec = c->g_rec;
    a.rec_flag = c->g_rec_flag;
    a.n_rec = 0;
    a.rec_stride = 1;
    a.rec_depth = depth;
    a.guide_depth = c->g_cfg.guide_depth > 0 ? c->g_cfg.guide_depth : 256;
    a.guide_target = c->g_cfg.guide_target;
    a.sdt.snode = c->sd_snode;
    a.sdt.dinfo = c->sd_dinfo;
    a.sdt.quad = c->sd_quad;
    a.sdt.rect = c->sd_rect;
    a.sdt.sum = c->sd_sum;
    a.sdt.cnt = c->sd_cnt;
    for (int k = 0; k < 3; k++) {
        a.sdt.bmin[k] = c->sd_tree ? c->sd_tree->bmin[k] : 0.0f;
        a.sdt.bmax[k] = c->sd_tree ? c->sd_tree->bmax[k] : 0.0f;
    }
    a.ppg_vtx = c->ppg_vtx;
    a.ppg_nv = c->ppg_nv;
    a.refill = CPT_DEFAULT_REFILL;
    a.integrator = 0;
    a.corrected = 0;
    a.gsamples = c->g_samples;
    a.gcount = c->g_count;
    a.gwt = c->g_wt;
    a.gwt8 = c->g_wt8;
    a.ginv8 = c->g_inv8;
    a.guide_fp8 = c->g_cfg.infer_fp8 != 0;
    a.guide_lobes = c->g_width == 128 ? 16 : 8;
    a.gparams = c->g_params;
    a.W = W;
    a.H = H;
    a.S = S;
    a.max_depth = depth;
    a.seed = seed;
    a.view = view;
    return a;
}

using KernelVoid = void (*)(WfArgs);

struct LaunchCfg {
    int B;
    int G;
    bool count;
    bool stream = false;
    bool guided = false;
    int infer_blocks = 0;
    int SB = 256;  // shading kernel block
    // grids of the traversal kernels (0: G).  Their waves split the queue into
    // equal chunks, so the grid should be what the device holds at once: a
    // second round of waves would double the time of the CUs that run it.
    int GE = 0, GS = 0;
    bool ppg = false;  // SD-tree guided arm: k_shade_ppg, no shadow rays
    bool lt = false;   // light tracer: k_raygen_lt, k_shade_lt, splatting shadow rays
};

template <int B>
void resident_grids(int cus, int& ge, int& gs) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)k_extend<B, false>, B, 0) != hipSuccess || nb <= 0)
        nb = 2048 / B;
    ge = nb * cus;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)k_shadow<B, false>, B, 0) != hipSuccess || nb <= 0)
        nb = 2048 / B;
    gs = nb * cus;
}

// launch helpers dispatching on the block size
template <int B>
void launch_extend(const LaunchCfg& L, const WfArgs& a, int cur, hipStream_t s) {
    const unsigned g = (unsigned)(L.GE ? L.GE : L.G);
    if (L.count) hipLaunchKernelGGL((k_extend<B, true>), dim3(g), dim3(B), 0, s, a, cur);
    else hipLaunchKernelGGL((k_extend<B, false>), dim3(g), dim3(B), 0, s, a, cur);
}
template <int B>
void launch_shade(const LaunchCfg& L, const WfArgs& a, int cur, int cap, hipStream_t s) {
    // same thread count as the other grid-stride kernels
    const unsigned G = (unsigned)std::max(1, (int)(((long long)L.G * L.B) / B));
    if (L.stream) hipLaunchKernelGGL((k_shade<B, true, false>), dim3(G), dim3(B), 0, s, a, cur, cap);
    else if (L.guided) hipLaunchKernelGGL((k_shade<B, false, true>), dim3(G), dim3(B), 0, s, a, cur, cap);
    else hipLaunchKernelGGL((k_shade<B, false, false>), dim3(G), dim3(B), 0, s, a, cur, cap);
}
template <int B>
void launch_shadow(const LaunchCfg& L, const WfArgs& a, int nxt, hipStream_t s) {
    const unsigned g = (unsigned)(L.GS ? L.GS : L.G);
    if (L.count) hipLaunchKernelGGL((k_shadow<B, true>), dim3(g), dim3(B), 0, s, a, nxt);
    else hipLaunchKernelGGL((k_shadow<B, false>), dim3(g), dim3(B), 0, s, a, nxt);
}

void L_extend(const LaunchCfg& L, const WfArgs& a, int cur, hipStream_t s) {
    switch (L.B) {
        case 64: launch_extend<64>(L, a, cur, s); break;
        case 256: launch_extend<256>(L, a, cur, s); break;
        default: launch_extend<128>(L, a, cur, s); break;
    }
}
void L_shade(const LaunchCfg& L, const WfArgs& a, int cur, int cap, hipStream_t s) {
    if (L.lt) {
        const unsigned G = (unsigned)std::max(1, (int)(((long long)L.G * L.B) / 256));
        hipLaunchKernelGGL((k_shade_lt<256>), dim3(G), dim3(256), 0, s, a, cur, cap);
        return;
    }
    if (L.ppg) {
        const unsigned G = (unsigned)std::max(1, (int)(((long long)L.G * L.B) / 256));
        hipLaunchKernelGGL((k_shade_ppg<256>), dim3(G), dim3(256), 0, s, a, cur, cap);
        return;
    }
    switch (L.SB) {
        case 64: launch_shade<64>(L, a, cur, cap, s); break;
        case 128: launch_shade<128>(L, a, cur, cap, s); break;
        case 512: launch_shade<512>(L, a, cur, cap, s); break;
        case 1024: launch_shade<1024>(L, a, cur, cap, s); break;
        default: launch_shade<256>(L, a, cur, cap, s); break;
    }
}
void L_shadow(const LaunchCfg& L, const WfArgs& a, int nxt, hipStream_t s) {
    switch (L.B) {
        case 64: launch_shadow<64>(L, a, nxt, s); break;
        case 256: launch_shadow<256>(L, a, nxt, s); break;
        default: launch_shadow<128>(L, a, nxt, s); break;
    }
}

// Issue one pass worth of kernels on stream s.
struct Timers {
    std::vector<hipEvent_t> ev;
    std::vector<int> kind;  // 0 extend, 1 shade, 2 shadow, 3 other, 4 guide inference, 5 guide training
    // interval ending here is of `kind` (-1: not counted)
    void mark(int k, hipStream_t s) {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        (void)hipEventRecord(e, s);
        ev.push_back(e);
        kind.push_back(k);
    }
};

int issue_pass(cpt_ctx* c, const LaunchCfg& L, const WfArgs& a, hipStream_t s, Timers* tm) {
    auto mark = [&](int kind) {
        if (tm) tm->mark(kind, s);
    };
    mark(-1);
    if (L.lt) hipLaunchKernelGGL(k_raygen_lt, dim3(L.G), dim3(L.B), 0, s, a);
    else hipLaunchKernelGGL(k_raygen, dim3(L.G), dim3(L.B), 0, s, a);
    mark(3);
    for (int b = 0; b < a.max_depth; b++) {
        int cur = b & 1;
        L_extend(L, a, cur, s);
        mark(0);
        if (L.guided && b < a.guide_depth) {
            // FP8: one block per CU (16 waves at width 64, 12 at 128)
            if (a.guide_lobes == 16) {
                if (a.guide_fp8)
                    hipLaunchKernelGGL(k_guide_infer8<guide128::Net>, dim3(L.infer_blocks),
                                       dim3(64 * guide128::Net::kWaves8), guide128::Net::kSmemFwd8, s, a, cur);
                else
                    hipLaunchKernelGGL(k_guide_infer<128>, dim3(L.infer_blocks), dim3(256),
                                       sizeof(guide128::SmemFwd), s, a, cur);
            } else if (a.guide_fp8) {
                hipLaunchKernelGGL(k_guide_infer8<guide::Net>, dim3(L.infer_blocks), dim3(64 * guide::Net::kWaves8),
                                   guide::Net::kSmemFwd8, s, a, cur);
            } else {
                hipLaunchKernelGGL(k_guide_infer16<guide::Net>, dim3(L.infer_blocks),
                                   dim3(64 * guide::Net::kWaves16), guide::Net::kSmemInf16, s, a, cur);
            }
            mark(4);
        }
        L_shade(L, a, cur, a.max_depth, s);
        mark(1);
        if (!L.ppg) {
            L_shadow(L, a, cur ^ 1, s);
            mark(2);
        }
    }
    hipLaunchKernelGGL(k_accum, dim3(L.G), dim3(L.B), 0, s, a);
    hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, s, a, -1);
    mark(3);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, CPT_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
    return CPT_OK;
}

int check_device(cpt_ctx* c) {
    hipDeviceProp_t prop;
    HIPCHK(c, hipGetDeviceProperties(&prop, c->device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(c, CPT_E_NODEVICE, "device %d is %s, libcpt is built for gfx950", c->device, prop.gcnArchName);
    c->cus = prop.multiProcessorCount;
    return CPT_OK;
}

}  // namespace

// Device address of `scratch` on the current device, looked up once per device.
static int* noalloc_scratch_of(int device, hipError_t* err) {
    static std::mutex mu;
    static std::map<int, int*> addr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(device);
    if (it != addr.end()) return it->second;
    void* p = nullptr;
    *err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (*err != hipSuccess) return nullptr;
    addr[device] = (int*)p;
    return (int*)p;
}

extern "C" {

int cpt_version(void) { return 1001; }

int cpt_create(cpt_ctx** out, int device) {
    if (!out) return CPT_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return CPT_E_NODEVICE;
    cpt_ctx* c = new cpt_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return CPT_E_NODEVICE;
    }
    int r = check_device(c);
    if (r) {
        delete c;
        return r;
    }
    hipError_t se = hipSuccess;
    c->noalloc_scratch = noalloc_scratch_of(device, &se);
    if (!c->noalloc_scratch) {
        delete c;
        return CPT_E_HIP;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void**)&c->counters, 8 * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void**)&c->stats, ST_N * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc((void**)&c->d_pass, sizeof(PassDesc)) != hipSuccess) {
        delete c;
        return CPT_E_HIP;
    }
    *out = c;
    return CPT_OK;
}

void cpt_destroy(cpt_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    free_scene(c);
    free_wavefront(c);
    dfree(c->gstack);
    dfree(c->counters);
    dfree(c->stats);
    dfree(c->d_pass);
    dfree(c->d_views);
    dfree(c->pass_tend);
    dfree(c->film);
    dfree(c->film2);
    dfree(c->g_params);
    dfree(c->g_m);
    dfree(c->g_v);
    dfree(c->g_grad);
    dfree(c->g_wt);
    dfree(c->g_wt8);
    dfree(c->g_inv8);
    dfree(c->g_samples);
    dfree(c->g_rows);
    dfree(c->g_partial);
    dfree(c->g_loss_part);
    dfree(c->g_wscale);
    dfree(c->g_wpart);
    dfree(c->g_lobes);
    dfree(c->sd_snode);
    dfree(c->sd_quad);
    dfree(c->sd_dinfo);
    dfree(c->sd_rect);
    dfree(c->sd_sum);
    dfree(c->sd_cnt);
    dfree(c->ppg_vtx);
    dfree(c->ppg_nv);
    dfree(c->g_rec);
    dfree(c->g_rec_flag);
    dfree(c->g_count);
    if (c->h_count) (void)hipHostFree(c->h_count);
    if (c->h_pass) (void)hipHostFree(c->h_pass);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* cpt_last_error(cpt_ctx* c) { return c ? c->err.c_str() : "null context"; }

int cpt_debug_fail_alloc(cpt_ctx* c, int32_t after) {
    if (!c) return CPT_E_INVALID;
    c->fail_alloc_in = after < 0 ? -1 : after;
    return CPT_OK;
}

int cpt_debug_quantize_node(const uint32_t* node4, uint32_t* qnode_out) {
    if (!node4 || !qnode_out) return CPT_E_INVALID;
    Node4 n;
    QNode q;
    std::memcpy(&n, node4, sizeof(n));
    quantize_node(n, q);
    std::memcpy(qnode_out, &q, sizeof(q));
    return CPT_OK;
}

int cpt_upload_scene(cpt_ctx* c, const cpt_scene_desc* d) {
    if (!c || !d) return CPT_E_INVALID;
    if (d->n_triangles < 0 || d->n_vertices < 0 || d->n_materials <= 0 || d->n_lights < 0)
        return fail(c, CPT_E_INVALID, "negative sizes or no materials");
    if (d->n_triangles > 0 && (!d->positions || !d->normals || !d->bitangents || !d->indices || !d->tri_material))
        return fail(c, CPT_E_INVALID, "missing vertex or triangle arrays");
    if (d->n_triangles >= (1 << 27)) return fail(c, CPT_E_LIMIT, "at most 2^27-1 triangles");
    for (int t = 0; t < d->n_triangles; t++) {
        for (int k = 0; k < 3; k++) {
            int v = d->indices[3 * (size_t)t + k];
            if (v < 0 || v >= d->n_vertices) return fail(c, CPT_E_INVALID, "triangle %d: vertex %d out of range", t, v);
        }
        int m = d->tri_material[t];
        if (m < 0 || m >= d->n_materials) return fail(c, CPT_E_INVALID, "triangle %d: material %d out of range", t, m);
    }
    // image textures (ImageTexture, image_texture.h:24-40)
    if (d->n_textures < 0) return fail(c, CPT_E_INVALID, "negative texture count");
    if (d->n_textures > 0 && (!d->textures || !d->texels || !d->uvs))
        return fail(c, CPT_E_INVALID, "textures need the texture table, texels and uvs");
    size_t n_texels = 0;
    for (int t = 0; t < d->n_textures; t++) {
        const int32_t* e = d->textures + 4 * (size_t)t;
        if (e[0] <= 0 || e[1] <= 0 || (e[2] != 1 && e[2] != 3) || e[3] < 0)
            return fail(c, CPT_E_INVALID, "texture %d: bad size, channels or offset", t);
        n_texels = std::max(n_texels, (size_t)e[3] + (size_t)e[0] * e[1] * e[2]);
    }
    for (int m = 0; m < d->n_materials; m++) {
        const float type = d->materials[16 * (size_t)m];
        if (!(type >= 0.0f && type <= (float)CPT_MAT_STANDARD) || type != (float)(int)type)
            return fail(c, CPT_E_INVALID, "material %d: unknown type", m);
        for (int k = 0; k < 4; k++) {
            const int tx = d->material_textures ? d->material_textures[4 * (size_t)m + k] : -1;
            if (tx < -1 || tx >= d->n_textures) return fail(c, CPT_E_INVALID, "material %d: texture %d out of range", m, tx);
        }
        if ((int)type == CPT_MAT_STANDARD && (!d->material_textures || d->material_textures[4 * (size_t)m + 2] < 0))
            return fail(c, CPT_E_INVALID, "material %d: CPT_MAT_STANDARD needs a metal texture", m);
    }
    if (d->n_point_lights < 0 || (d->n_point_lights > 0 && !d->point_lights))
        return fail(c, CPT_E_INVALID, "bad point-light table");
    if (d->n_lights > 0 && !d->light_triangles) return fail(c, CPT_E_INVALID, "missing light table");
    for (int i = 0; i < d->n_lights; i++) {
        int t = d->light_triangles[i];
        if (t >= d->n_triangles || (t < 0 && -1 - t >= d->n_point_lights))
            return fail(c, CPT_E_INVALID, "light %d: entry %d out of range", i, t);
    }
    HIPCHK(c, hipSetDevice(c->device));
    drop_graph(c);
    free_scene(c);
    c->nt = d->n_triangles;
    c->nv = d->n_vertices;
    c->nl = d->n_lights;
    c->nm = d->n_materials;
    std::memcpy(c->cam, d->camera, sizeof(c->cam));
    std::memcpy(c->env, d->env, sizeof(c->env));

    c->rank.clear();
    if (!(d->flags & CPT_SCENE_NO_TIE_RANK)) reference_tie_rank(d->positions, d->indices, c->nt, c->rank);
    Bvh4 bvh;
    const bool gpu_bvh = (d->flags & CPT_SCENE_GPU_BVH) && c->nt > 0;
    if (!gpu_bvh) {
        auto tb0 = std::chrono::steady_clock::now();
        build_bvh4(d->positions, d->indices, c->nt, c->rank, bvh);
        c->bvh_build_us = (int64_t)std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - tb0).count();
        if (3 * bvh.max_depth + 1 > CPT_LDS_SS + CPT_GSTACK)
            return fail(c, CPT_E_LIMIT, "BVH depth %d exceeds the traversal stack", bvh.max_depth);
        if (bvh.nodes.size() > CPT_MAX_NODES)  // traversal addresses nodes by 32-bit byte offsets
            return fail(c, CPT_E_LIMIT, "BVH of %zu nodes exceeds %u", bvh.nodes.size(), CPT_MAX_NODES);
        c->bvh_nodes = (int64_t)bvh.nodes.size();
        c->bvh_leaves = bvh.leaves;
        c->bvh_depth = bvh.max_depth;
    }
    c->bvh_gpu = gpu_bvh;

    for (int k = 0; k < 3; k++) {
        c->bmin[k] = INFINITY;
        c->bext[k] = -INFINITY;
    }
    for (int i = 0; i < c->nv; i++)
        for (int k = 0; k < 3; k++) {
            c->bmin[k] = std::min(c->bmin[k], d->positions[3 * (size_t)i + k]);
            c->bext[k] = std::max(c->bext[k], d->positions[3 * (size_t)i + k]);
        }
    for (int k = 0; k < 3; k++) {
        if (c->nv == 0) c->bmin[k] = c->bext[k] = 0.0f;
        c->bext[k] = std::max(c->bext[k] - c->bmin[k], 1e-6f);
        c->tb_lo[k] = INFINITY;
        c->tb_hi[k] = -INFINITY;
    }
    for (size_t i = 0; i < 3 * (size_t)c->nt; i++)
        for (int k = 0; k < 3; k++) {
            const float v = d->positions[3 * (size_t)d->indices[i] + k];
            c->tb_lo[k] = std::min(c->tb_lo[k], v);
            c->tb_hi[k] = std::max(c->tb_hi[k], v);
        }
    c->sd_tree.reset();
    std::vector<float4> vp(std::max(1, c->nv)), vn(std::max(1, c->nv)), vb(std::max(1, c->nv));
    for (int i = 0; i < c->nv; i++) {
        vp[i] = make_float4(d->positions[3 * (size_t)i], d->positions[3 * (size_t)i + 1], d->positions[3 * (size_t)i + 2], 0.f);
        vn[i] = make_float4(d->normals[3 * (size_t)i], d->normals[3 * (size_t)i + 1], d->normals[3 * (size_t)i + 2], 0.f);
        vb[i] = make_float4(d->bitangents[3 * (size_t)i], d->bitangents[3 * (size_t)i + 1], d->bitangents[3 * (size_t)i + 2], 0.f);
    }
    std::vector<int4> ti(std::max(1, c->nt));
    for (int t = 0; t < c->nt; t++)
        ti[t] = make_int4(d->indices[3 * (size_t)t], d->indices[3 * (size_t)t + 1], d->indices[3 * (size_t)t + 2], d->tri_material[t]);
    // shading records: geometric normal with the reference's float operations
    // (normalized(cross(b-a, c-a)), triangle.h:109; math.h:132-140), then vertex
    // normals and bitangents
    std::vector<float4> sh(6 * (size_t)std::max(1, c->nt), make_float4(0, 0, 0, 0));
    for (int t = 0; t < c->nt; t++) {
        const float* P = d->positions;
        int i0 = d->indices[3 * (size_t)t], i1 = d->indices[3 * (size_t)t + 1], i2 = d->indices[3 * (size_t)t + 2];
        float e1[3], e2[3];
        for (int k = 0; k < 3; k++) {
            e1[k] = P[3 * (size_t)i1 + k] - P[3 * (size_t)i0 + k];
            e2[k] = P[3 * (size_t)i2 + k] - P[3 * (size_t)i0 + k];
        }
        float cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
        float dd = cx * cx + cy * cy + cz * cz;
        float len = (float)std::sqrt((double)dd);
        float inv = (float)(1.0 / (double)len);
        float v[24];
        v[0] = cx * inv;
        v[1] = cy * inv;
        v[2] = cz * inv;
        int32_t m = d->tri_material[t];
        std::memcpy(&v[3], &m, 4);
        const int vi[3] = {i0, i1, i2};
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                v[4 + 3 * j + k] = d->normals[3 * (size_t)vi[j] + k];
                v[13 + 3 * j + k] = d->bitangents[3 * (size_t)vi[j] + k];
            }
        v[22] = v[23] = 0.0f;
        std::memcpy(&sh[6 * (size_t)t], v, sizeof(v));
    }
    int r = 0;
    if (!gpu_bvh) {
        r |= dalloc(c, c->d_nodes, bvh.nodes.size());
        r |= dalloc(c, c->d_tris, 3 * std::max<size_t>(1, bvh.tris.size()));
    }
    r |= dalloc(c, c->d_tri_info, ti.size());
    r |= dalloc(c, c->d_shade, sh.size());
    r |= dalloc(c, c->d_vpos, vp.size());
    r |= dalloc(c, c->d_vnrm, vn.size());
    r |= dalloc(c, c->d_vbit, vb.size());
    r |= dalloc(c, c->d_mats, 16 * (size_t)c->nm);
    r |= dalloc(c, c->d_mtex, (size_t)c->nm);
    r |= dalloc(c, c->d_uv, std::max<size_t>(1, d->uvs ? (size_t)d->n_vertices : 1));
    r |= dalloc(c, c->d_tex, std::max<size_t>(1, (size_t)d->n_textures));
    r |= dalloc(c, c->d_texels, std::max<size_t>(1, n_texels));
    r |= dalloc(c, c->d_lights, std::max(1, c->nl));
    r |= dalloc(c, c->d_plights, 2 * (size_t)std::max(1, d->n_point_lights));
    if (r) return CPT_E_HIP;
    if (!gpu_bvh) {
        HIPCHK(c, hipMemcpy(c->d_nodes, bvh.nodes.data(), bvh.nodes.size() * sizeof(Node4), hipMemcpyHostToDevice));
        if (!bvh.tris.empty())
            HIPCHK(c, hipMemcpy(c->d_tris, bvh.tris.data(), bvh.tris.size() * sizeof(TriRec), hipMemcpyHostToDevice));
    }
    HIPCHK(c, hipMemcpy(c->d_tri_info, ti.data(), ti.size() * sizeof(int4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_shade, sh.data(), sh.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_vpos, vp.data(), vp.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_vnrm, vn.data(), vn.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_vbit, vb.data(), vb.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_mats, d->materials, 16 * sizeof(float) * (size_t)c->nm, hipMemcpyHostToDevice));
    {
        std::vector<int4> mt((size_t)c->nm, make_int4(-1, -1, -1, -1));
        if (d->material_textures)
            for (int m = 0; m < c->nm; m++) {
                const int32_t* e = d->material_textures + 4 * (size_t)m;
                mt[m] = make_int4(e[0], e[1], e[2], e[3]);
            }
        HIPCHK(c, hipMemcpy(c->d_mtex, mt.data(), mt.size() * sizeof(int4), hipMemcpyHostToDevice));
        if (d->uvs && d->n_vertices > 0)
            HIPCHK(c, hipMemcpy(c->d_uv, d->uvs, 2 * sizeof(float) * (size_t)d->n_vertices, hipMemcpyHostToDevice));
        if (d->n_textures > 0) {
            HIPCHK(c, hipMemcpy(c->d_tex, d->textures, 4 * sizeof(int32_t) * (size_t)d->n_textures, hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy(c->d_texels, d->texels, sizeof(float) * n_texels, hipMemcpyHostToDevice));
        }
    }
    if (c->nl > 0)
        HIPCHK(c, hipMemcpy(c->d_lights, d->light_triangles, sizeof(int) * (size_t)c->nl, hipMemcpyHostToDevice));
    if (gpu_bvh) {
        int32_t* d_rank = nullptr;
        if (!c->rank.empty()) {
            if (dalloc(c, d_rank, c->rank.size())) return CPT_E_HIP;
            HIPCHK(c, hipMemcpy(d_rank, c->rank.data(), sizeof(int32_t) * c->rank.size(), hipMemcpyHostToDevice));
        }
        GpuBvh gb;
        std::string berr;
        hipError_t be = build_bvh4_gpu(c->d_vpos, c->d_tri_info, d_rank, c->nt, CPT_GPU_BVH_LEAF, CPT_GPU_BVH_RADIUS,
                                       c->stream, &gb, &berr);
        dfree(d_rank);
        if (be != hipSuccess) return fail(c, CPT_E_HIP, "%s", berr.c_str());
        c->d_nodes = gb.nodes;
        c->d_tris = reinterpret_cast<float4*>(gb.tris);
        c->bvh_nodes = gb.n_nodes;
        c->bvh_leaves = gb.leaves;
        c->bvh_depth = gb.max_depth;
        c->bvh_build_us = (int64_t)(gb.ms * 1000.0);
        if (3 * gb.max_depth + 1 > CPT_LDS_SS + CPT_GSTACK)
            return fail(c, CPT_E_LIMIT, "BVH depth %d exceeds the traversal stack", gb.max_depth);
        if (gb.n_nodes > (int64_t)CPT_MAX_NODES)  // traversal addresses nodes by 32-bit byte offsets
            return fail(c, CPT_E_LIMIT, "BVH of %lld nodes exceeds %u", (long long)gb.n_nodes, CPT_MAX_NODES);
    }
    if (d->n_point_lights > 0) {
        std::vector<float4> pl(2 * (size_t)d->n_point_lights);
        for (int k = 0; k < d->n_point_lights; k++) {
            const float* e = d->point_lights + 6 * (size_t)k;
            pl[2 * k] = make_float4(e[0], e[1], e[2], 0.0f);
            pl[2 * k + 1] = make_float4(e[3], e[4], e[5], 0.0f);
        }
        HIPCHK(c, hipMemcpy(c->d_plights, pl.data(), pl.size() * sizeof(float4), hipMemcpyHostToDevice));
    }
    // the traversal's quantized nodes (their scales cover extents up to 255 x 2^60)
    for (int k = 0; k < 3; k++)
        if ((double)c->bext[k] > QNODE_MAX_EXTENT)
            return fail(c, CPT_E_LIMIT, "scene extent %g exceeds the quantized BVH's range", (double)c->bext[k]);
    if (c->bvh_nodes < 1) return fail(c, CPT_E_INVALID, "empty BVH");
    if (dalloc(c, c->d_qnodes, (size_t)c->bvh_nodes)) return CPT_E_HIP;
    hipLaunchKernelGGL(k_quantize_nodes, dim3((unsigned)std::min<int64_t>((c->bvh_nodes + 255) / 256, 4096)), dim3(256),
                       0, c->stream, c->d_nodes, c->bvh_nodes, c->d_qnodes);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c->d_nodes);  // the float nodes were only the conversion's input
    c->has_scene = true;
    c->scene_gen++;
    return CPT_OK;
}

int cpt_bvh_info(cpt_ctx* c, int64_t* info) {
    if (!c || !info) return CPT_E_INVALID;
    if (!c->has_scene) return fail(c, CPT_E_NOSCENE, "no scene uploaded");
    info[0] = c->bvh_nodes;
    info[1] = c->bvh_leaves;
    info[2] = c->bvh_depth;
    info[3] = c->nt;
    info[4] = c->bvh_build_us;
    info[5] = c->bvh_gpu ? 1 : 0;
    return CPT_OK;
}

int cpt_tie_rank(cpt_ctx* c, int32_t* out) {
    if (!c || !out) return CPT_E_INVALID;
    if (!c->has_scene) return fail(c, CPT_E_NOSCENE, "no scene uploaded");
    if (c->rank.empty() && c->nt > 0) return fail(c, CPT_E_INVALID, "scene uploaded with CPT_SCENE_NO_TIE_RANK");
    std::memcpy(out, c->rank.data(), sizeof(int32_t) * c->rank.size());
    return CPT_OK;
}

int cpt_reference_tie_rank(const float* positions, const int32_t* indices, int32_t nt, int32_t* out) {
    if (nt < 0 || (nt > 0 && (!positions || !indices || !out))) return CPT_E_INVALID;
    std::vector<int32_t> r;
    reference_tie_rank(positions, indices, nt, r);
    if (nt > 0) std::memcpy(out, r.data(), sizeof(int32_t) * (size_t)nt);
    return CPT_OK;
}

static int default_grid(cpt_ctx* c, int B) { return c->cus * (2048 / B) ; }

static int trace_common(cpt_ctx* c, const float* rays, int32_t n, bool any, int32_t* tri_out, float* t_out,
                        uint8_t* occ_out) {
    if (!c || (n > 0 && !rays) || n < 0) return CPT_E_INVALID;
    if (!c->has_scene) return fail(c, CPT_E_NOSCENE, "no scene uploaded");
    if (n == 0) return CPT_OK;
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<float4> h(2 * (size_t)n);
    for (int i = 0; i < n; i++) {
        const float* r = rays + 7 * (size_t)i;
        h[2 * (size_t)i] = make_float4(r[0], r[1], r[2], r[6]);
        h[2 * (size_t)i + 1] = make_float4(r[3], r[4], r[5], 0.0f);
    }
    const int B = 128;
    int G = std::min(default_grid(c, B), (n + B - 1) / B);
    int ret = ensure_gstack(c, (size_t)G * B);
    if (ret) return ret;
    float4* d_rays = nullptr;
    void* d_out = nullptr;
    HIPCHK(c, hipMalloc((void**)&d_rays, h.size() * sizeof(float4)));
    size_t out_bytes = any ? (size_t)n : (size_t)n * sizeof(float4);
    if (hipMalloc(&d_out, out_bytes) != hipSuccess) {
        (void)hipFree(d_rays);
        return fail(c, CPT_E_HIP, "hipMalloc of %zu bytes failed", out_bytes);
    }
    int rc = CPT_OK;
    if (hipMemcpyAsync(d_rays, h.data(), h.size() * sizeof(float4), hipMemcpyHostToDevice, c->stream) != hipSuccess) {
        rc = fail(c, CPT_E_HIP, "ray upload failed");
    } else {
        DevScene sc = dev_scene(c);
        if (any)
            hipLaunchKernelGGL((k_trace_any<B>), dim3(G), dim3(B), 0, c->stream, sc, d_rays, n, (uint8_t*)d_out, c->gstack,
                               CPT_DEFAULT_REFILL);
        else
            hipLaunchKernelGGL((k_trace_closest<B>), dim3(G), dim3(B), 0, c->stream, sc, d_rays, n, (float4*)d_out, c->gstack,
                               CPT_DEFAULT_REFILL);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(c, CPT_E_HIP, "trace launch failed: %s", hipGetErrorString(e));
    }
    if (rc == CPT_OK) {
        if (any) {
            if (hipMemcpyAsync(occ_out, d_out, n, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
                rc = fail(c, CPT_E_HIP, "result download failed");
        } else {
            std::vector<float4> hh(n);
            if (hipMemcpyAsync(hh.data(), d_out, n * sizeof(float4), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess) {
                rc = fail(c, CPT_E_HIP, "result download failed");
            } else {
                for (int i = 0; i < n; i++) {
                    int32_t tid;
                    std::memcpy(&tid, &hh[i].w, sizeof(tid));
                    tri_out[i] = tid;
                    t_out[i] = tri_out[i] >= 0 ? hh[i].x : INFINITY;
                }
            }
        }
        if (rc == CPT_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, CPT_E_HIP, "trace failed");
    }
    (void)hipFree(d_rays);
    (void)hipFree(d_out);
    return rc;
}

int cpt_trace_closest(cpt_ctx* c, const float* rays, int32_t n, int32_t* tri_out, float* t_out) {
    if (!tri_out || !t_out) return CPT_E_INVALID;
    return trace_common(c, rays, n, false, tri_out, t_out, nullptr);
}

int cpt_trace_any(cpt_ctx* c, const float* rays, int32_t n, uint8_t* occ) {
    if (!occ) return CPT_E_INVALID;
    return trace_common(c, rays, n, true, nullptr, nullptr, occ);
}

int cpt_trace_closest_device(cpt_ctx* c, const void* d_rays, int32_t n, void* d_hits, void* stream) {
    if (!c || !d_rays || !d_hits || n < 0) return CPT_E_INVALID;
    if (!c->has_scene) return fail(c, CPT_E_NOSCENE, "no scene uploaded");
    if (n == 0) return CPT_OK;
    const int B = 128;
    int G = std::min(default_grid(c, B), (n + B - 1) / B);
    int ret = ensure_gstack(c, (size_t)G * B);
    if (ret) return ret;
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    hipLaunchKernelGGL((k_trace_closest<B>), dim3(G), dim3(B), 0, s, dev_scene(c), (const float4*)d_rays, n,
                       (float4*)d_hits, c->gstack, CPT_DEFAULT_REFILL);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, CPT_E_HIP, "trace launch failed: %s", hipGetErrorString(e));
    return CPT_OK;
}

int cpt_trace_closest_device_noalloc(cpt_ctx* c, const void* d_rays, int32_t n, void* d_hits, void* stream) {
    if (!c || !d_rays || !d_hits || n < 0) return CPT_E_INVALID;
    if (!c->has_scene) return fail(c, CPT_E_NOSCENE, "no scene uploaded");
    if (3 * c->bvh_depth + 1 > CPT_LDS_SS + CPT_GSTACK)  // upload refuses deeper trees; kept as the scratch's own bound
        return fail(c, CPT_E_LIMIT, "BVH depth %d needs more traversal stack than the scratch holds", c->bvh_depth);
    if (n == 0) return CPT_OK;
    // the grid never exceeds the threads the scratch covers; each wave walks its chunk of the batch
    const int G = std::min(NOALLOC_BLOCKS, (n + NOALLOC_B - 1) / NOALLOC_B);
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    hipLaunchKernelGGL((k_trace_closest<NOALLOC_B>), dim3(G), dim3(NOALLOC_B), 0, s, dev_scene(c), (const float4*)d_rays,
                       n, (float4*)d_hits, c->noalloc_scratch, CPT_DEFAULT_REFILL);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, CPT_E_HIP, "trace launch failed: %s", hipGetErrorString(e));
    return CPT_OK;
}

static int render_ppg(cpt_ctx* c, const cpt_render_params* p, cpt_render_stats* stats);
static int render_views_stream(cpt_ctx* c, int32_t n_views, const float* cameras, const int32_t* wh_spp,
                               int32_t max_depth, uint64_t seed, int32_t flags, float* films, double* view_ms,
                               cpt_render_stats* stats);
static int guide_train_device(cpt_ctx* c, int n, bool apply);

// Probability of sampling the guide after `step` Adam steps of the network: the
// configured fraction, reached linearly over the first fraction_ramp steps (0: at
// once).  Any value in [0, 1) keeps the one-sample MIS estimator unbiased.
static float guide_fraction_at(const cpt_guide_params& g, int step) {
    if (g.fraction_ramp <= 0 || step >= g.fraction_ramp) return g.guide_fraction;
    return g.guide_fraction * (float)step / (float)g.fraction_ramp;
}

// Pack views into passes of at most `cap` paths (see PassDesc).
static void plan_passes(const std::vector<ViewDesc>& vd, int64_t cap, std::vector<PassDesc>& passes,
                        std::vector<int>& last_pass_of_view) {
    passes.clear();
    last_pass_of_view.assign(vd.size(), 0);
    auto open_pass = [&]() {
        passes.emplace_back();
        PassDesc& pd = passes.back();
        std::memset(&pd, 0, sizeof(PassDesc));
        pd.pass_id = (uint32_t)(passes.size() - 1);
    };
    open_pass();
    auto add = [&](uint32_t v, uint32_t p0, uint32_t npix, uint32_t s0, uint32_t ns) {
        PassDesc& pd = passes.back();
        Segment& sg = pd.seg[pd.n_seg++];
        sg.view = v;
        sg.p0 = p0;
        sg.npix = npix;
        sg.s0 = s0;
        sg.ns = ns;
        sg.slot0 = pd.n_paths;
        sg.pix0 = pd.n_pix;
        sg.pad = 0;
        pd.n_paths += npix * ns;
        pd.n_pix += npix;
        last_pass_of_view[v] = (int)passes.size() - 1;
    };
    // views in ascending order of path count (shortest job first): small views
    // complete in the first passes instead of waiting behind large ones, which
    // minimises the mean completion time; each view's samples still run in
    // ascending order, so the films do not depend on the order
    std::vector<size_t> order(vd.size());
    for (size_t v = 0; v < vd.size(); v++) order[v] = v;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return vd[a].n_paths < vd[b].n_paths; });
    for (size_t v : order) {
        int64_t P = (int64_t)vd[v].W * vd[v].H;
        int64_t S2 = (int64_t)vd[v].spp;
        if (P <= cap) {
            int64_t s0 = 0;
            while (s0 < S2) {
                PassDesc* pd = &passes.back();
                int64_t room = cap - pd->n_paths;
                int64_t take = std::min<int64_t>(S2 - s0, room / P);
                if (take <= 0 || pd->n_seg == CPT_MAX_SEG) {
                    open_pass();
                    continue;
                }
                add((uint32_t)v, 0, (uint32_t)P, (uint32_t)s0, (uint32_t)take);
                s0 += take;
            }
        } else {
            for (int64_t s = 0; s < S2; s++) {
                int64_t p0 = 0;
                while (p0 < P) {
                    PassDesc* pd = &passes.back();
                    int64_t room = cap - pd->n_paths;
                    if (room <= 0 || pd->n_seg == CPT_MAX_SEG) {
                        open_pass();
                        continue;
                    }
                    int64_t n = std::min<int64_t>(room, P - p0);
                    add((uint32_t)v, (uint32_t)p0, (uint32_t)n, (uint32_t)s, 1);
                    p0 += n;
                }
            }
        }
    }
    if (passes.back().n_seg == 0) passes.pop_back();
}

// Flatten the host SD-tree into the device arrays (when it changed) and clear
// the splat accumulators.
static int sd_upload(cpt_ctx* c) {
    if (!c->sd_dirty) return CPT_OK;
    cpt::SdFlat& f = c->sd_flat;
    c->sd_tree->flatten(f);
    const size_t nd = f.dinfo.size() / 4, nr = f.rect.size() / 4, ns = f.slot_dtree.size();
    int r = CPT_OK;
    auto grow = [&](auto*& p, size_t& cap, size_t n) {
        if (r || (cap >= n && p)) return;
        drop_graph(c);
        const size_t want = std::max<size_t>(n + n / 2, 64);
        r = dalloc(c, p, want);
        if (!r) cap = want;
    };
    grow(c->sd_snode, c->sd_snode_cap, f.snode.size());
    grow(c->sd_dinfo, c->sd_dinfo_cap, nd);
    grow(c->sd_quad, c->sd_quad_cap, f.quad.size());
    grow(c->sd_rect, c->sd_rect_cap, nr);
    if (!r && (c->sd_slot_cap < ns || !c->sd_sum || !c->sd_cnt)) {
        drop_graph(c);
        const size_t want = std::max<size_t>(ns + ns / 2, 64);
        r = dalloc(c, c->sd_sum, want);
        if (!r) r = dalloc(c, c->sd_cnt, want);
        if (!r) c->sd_slot_cap = want;
    }
    if (r) return r;
    HIPCHK(c, hipMemcpyAsync(c->sd_snode, f.snode.data(), f.snode.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->sd_dinfo, f.dinfo.data(), f.dinfo.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->sd_quad, f.quad.data(), f.quad.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->sd_rect, f.rect.data(), f.rect.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->sd_sum, 0, ns * sizeof(float), c->stream));
    HIPCHK(c, hipMemsetAsync(c->sd_cnt, 0, ns * sizeof(uint32_t), c->stream));
    // the host copies must stay alive until the copies ran
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->sd_dirty = false;
    return CPT_OK;
}

// Render a batch of views into the concatenated FP64 films.
static int render_batch(cpt_ctx* c, std::vector<ViewDesc>& vd, int max_depth, uint64_t seed, int flags, int B, int G,
                        int SB, int refill, int64_t max_paths, double* view_ms, cpt_render_stats* stats) {
    if (max_depth <= 0 || max_depth > 255) return fail(c, CPT_E_INVALID, "max_depth must be in [1, 255]");
    if (refill < 1 || refill > 64) return fail(c, CPT_E_INVALID, "refill must be in [1, 64]");
    const int integrator = (flags & CPT_RENDER_LIGHT_TRACE) ? 4
                           : (flags & CPT_RENDER_PPG_STEP) ? 3
                           : (flags & CPT_RENDER_PURE_PT) ? 2
                           : (flags & CPT_RENDER_NEE_ONLY) ? 1
                                                            : 0;
    if (__builtin_popcount(flags & (CPT_RENDER_PURE_PT | CPT_RENDER_NEE_ONLY | CPT_RENDER_PPG_STEP |
                                    CPT_RENDER_LIGHT_TRACE)) > 1)
        return fail(c, CPT_E_INVALID,
                    "CPT_RENDER_NEE_ONLY, CPT_RENDER_PURE_PT, CPT_RENDER_PPG and CPT_RENDER_LIGHT_TRACE exclude each other");
    if (integrator == 4 && vd.size() != 1) return fail(c, CPT_E_INVALID, "the light tracer renders one view at a time");
    if (integrator && (flags & CPT_RENDER_GUIDED))
        return fail(c, CPT_E_INVALID, "the learned guide uses the MIS integrator");
    if (B != 64 && B != 128 && B != 256) return fail(c, CPT_E_INVALID, "block_size must be 64, 128 or 256");
    if (SB != 64 && SB != 128 && SB != 256 && SB != 512 && SB != 1024)
        return fail(c, CPT_E_INVALID, "shade_block must be 64, 128, 256, 512 or 1024");
    if (G < 0) return fail(c, CPT_E_INVALID, "grid_blocks must not be negative");
    int GE = 0, GS = 0;
    if (G == 0) {  // automatic
        G = default_grid(c, B);
        if (B == 64) resident_grids<64>(c->cus, GE, GS);
        else if (B == 256) resident_grids<256>(c->cus, GE, GS);
        else resident_grids<128>(c->cus, GE, GS);
    }
    unsigned long long film_px = 0;
    for (size_t v = 0; v < vd.size(); v++) {
        if (vd[v].W <= 0 || vd[v].H <= 0 || vd[v].S <= 0) return fail(c, CPT_E_INVALID, "view %zu: bad size or spp", v);
        if (vd[v].spp == 0) {  // stratified S x S
            if (vd[v].S > (1 << 10)) return fail(c, CPT_E_LIMIT, "view %zu too large", v);
            vd[v].spp = (unsigned)(vd[v].S * vd[v].S);
            vd[v].strat = 1;
        }
        int64_t P = (int64_t)vd[v].W * vd[v].H, S2 = (int64_t)vd[v].spp;
        if (P > (1 << 28) || S2 > (1 << 20) || P * S2 >= (1ll << 40)) return fail(c, CPT_E_LIMIT, "view %zu too large", v);
        vd[v].n_paths = (unsigned long long)(P * S2);
        vd[v].film_off = film_px;
        vd[v].path_begin = 0;
        film_px += (unsigned long long)P;
    }
    if (film_px >= (1ull << 31)) return fail(c, CPT_E_LIMIT, "films too large");
    HIPCHK(c, hipSetDevice(c->device));
    auto t0 = std::chrono::steady_clock::now();
    int64_t cap = std::max<int64_t>(max_paths, 1024);
    cap = std::min<int64_t>(cap, 1ll << 30);
    std::vector<PassDesc> passes;
    std::vector<int> last_pass;
    plan_passes(vd, cap, passes, last_pass);
    for (auto& pd : passes) pd.seed = seed;
    size_t slots = 0;
    for (auto& pd : passes) slots = std::max<size_t>(slots, pd.n_paths);
    int r = ensure_wavefront(c, slots);
    if (r) return r;
    r = ensure_gstack(c, (size_t)G * B);
    if (r) return r;
    r = ensure_film(c, film_px);
    if (r) return r;
    // adding to the film needs the previous render's film, views and sizes
    if (flags & CPT_RENDER_ACCUMULATE) {
        bool same = c->film_layout.size() == vd.size();
        for (size_t v = 0; same && v < vd.size(); v++)
            same = c->film_layout[v] == std::make_pair(vd[v].W, vd[v].H);
        if (!same)
            return fail(c, CPT_E_INVALID,
                        "CPT_RENDER_ACCUMULATE needs a film of the same views and sizes from a completed render");
    }
    const bool guided = (flags & CPT_RENDER_GUIDED) != 0;
    const bool train = guided && (flags & CPT_RENDER_GUIDE_TRAIN) != 0;
    size_t nrec = 0;
    if (guided) {
        if (!c->has_guide) return fail(c, CPT_E_INVALID, "guided render needs cpt_guide_init");
        const size_t lfloats = slots * 3 * (c->g_width == 128 ? 16 : 8);
        if (c->g_lobes_cap < lfloats || !c->g_lobes) {
            drop_graph(c);
            c->g_lobes_cap = 0;
            r = dalloc(c, c->g_lobes, lfloats);
            if (r) return r;
            c->g_lobes_cap = lfloats;
        }
        if (train) {
            nrec = std::min<size_t>(slots, (size_t)c->g_cfg.train_paths);
            size_t nr = nrec * (size_t)max_depth;
            if (c->g_rec_cap < nr || !c->g_rec || !c->g_rec_flag) {
                drop_graph(c);
                c->g_rec_cap = 0;
                r = dalloc(c, c->g_rec, nr * REC_FLOATS);
                if (!r) r = dalloc(c, c->g_rec_flag, nr);
                if (r) return r;
                c->g_rec_cap = nr;
            }
            if (c->g_samples_cap < nr || !c->g_samples) {
                drop_graph(c);
                c->g_samples_cap = 0;
                r = dalloc(c, c->g_samples, nr * guide::SAMPLE_FLOATS);
                if (r) return r;
                c->g_samples_cap = nr;
            }
        }
    }
    if (c->views_cap < vd.size() || !c->d_views) {
        drop_graph(c);
        c->views_cap = 0;
        r = dalloc(c, c->d_views, vd.size());
        if (r) return r;
        c->views_cap = vd.size();
    }
    if (c->tend_cap < passes.size() + 1 || !c->pass_tend) {
        drop_graph(c);
        c->tend_cap = 0;
        r = dalloc(c, c->pass_tend, passes.size() + 1);
        if (r) return r;
        c->tend_cap = passes.size() + 1;
    }
    // the previous render's copies from the pinned pass array must be done before it is rewritten
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->h_pass_cap < passes.size()) {
        if (c->h_pass) (void)hipHostFree(c->h_pass);
        c->h_pass = nullptr;
        c->h_pass_cap = 0;
        HIPCHK(c, hipHostMalloc((void**)&c->h_pass, passes.size() * sizeof(PassDesc), hipHostMallocDefault));
        c->h_pass_cap = passes.size();
    }
    std::memcpy(c->h_pass, passes.data(), passes.size() * sizeof(PassDesc));
    HIPCHK(c, hipMemcpyAsync(c->d_views, vd.data(), vd.size() * sizeof(ViewDesc), hipMemcpyHostToDevice, c->stream));
    c->film_layout.clear();  // until this render completes
    if (!(flags & CPT_RENDER_ACCUMULATE)) {
        HIPCHK(c, hipMemsetAsync(c->film, 0, 3 * sizeof(double) * film_px, c->stream));
        HIPCHK(c, hipMemsetAsync(c->film2, 0, 3 * sizeof(double) * film_px, c->stream));
    }
    if (integrator == 3) {
        if (!c->sd_tree) return fail(c, CPT_E_INVALID, "PPG step without a tree (cpt_ppg_reset)");
        r = sd_upload(c);
        if (r) return r;
        const size_t need = slots * (size_t)max_depth * 9;
        if (c->ppg_cap < need || !c->ppg_vtx || !c->ppg_nv) {
            drop_graph(c);
            c->ppg_cap = 0;
            dfree(c->ppg_nv);
            r = dalloc(c, c->ppg_vtx, need);
            if (!r) r = dalloc(c, c->ppg_nv, slots);
            if (r) return r;
            c->ppg_cap = need;
        }
    }
    WfArgs a = make_args(c, 0, 0, 1, max_depth, seed, 0);
    a.n_rec = (uint32_t)nrec;
    // rounded up, so the records reach the end of the pass
    a.rec_stride = nrec ? (uint32_t)((slots + nrec - 1) / nrec) : 1u;
    a.refill = (unsigned)refill;
    a.integrator = integrator;
    a.corrected = (flags & CPT_RENDER_CORRECTED) ? 1 : 0;
    LaunchCfg L{B, G, (flags & CPT_RENDER_COUNT) != 0};
    L.SB = SB;
    L.GE = GE;
    L.GS = GS;
    L.guided = guided;
    L.ppg = integrator == 3;
    L.lt = integrator == 4;
    L.infer_blocks = c->cus;
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->stream, c->stats, (int)ST_N);
    hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, c->stream, a, (int)passes.size());
    c->g_samples_total = 0;
    c->g_steps_total = 0;
    bool timing = (flags & CPT_RENDER_KERNEL_TIMING) != 0;
    bool use_graph = !(flags & (CPT_RENDER_NO_GRAPH | CPT_RENDER_KERNEL_TIMING));
    Timers tm;
    bool captured = false;
    if (use_graph) {
        char key[512];
        snprintf(key, sizeof(key),
                 "%d,%d,%d,%d,%d,%zu,%p,%p,%p,%p,%d,%zu,%u,%p,%p,%p,%d,%d,%d,%d,%zu,%p,%p,%p,%p,%p,%p,%p,%g,%g,%g,%g,%g,%g",
                 max_depth, B, G, SB, (int)L.count, c->cap, (void*)c->d_qnodes, (void*)c->film,
                 (void*)c->d_views, (void*)c->pass_tend, (int)guided, nrec, a.rec_stride, (void*)c->g_lobes, (void*)c->g_rec,
                 (void*)c->g_samples, a.guide_depth * 2 + a.guide_target, refill, integrator,
                 a.guide_lobes * 4 + a.corrected * 2 + a.guide_fp8, a.lobes_stride,
                 (void*)c->sd_snode, (void*)c->sd_dinfo, (void*)c->sd_quad, (void*)c->sd_rect, (void*)c->sd_sum,
                 (void*)c->ppg_vtx, (void*)c->ppg_nv, a.sdt.bmin[0], a.sdt.bmin[1], a.sdt.bmin[2], a.sdt.bmax[0],
                 a.sdt.bmax[1], a.sdt.bmax[2]);
        if (!c->gexec || c->graph_key != key) {
            drop_graph(c);
            hipGraph_t g;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int ir = issue_pass(c, L, a, c->stream, nullptr);
            hipError_t ec = hipStreamEndCapture(c->stream, &g);
            if (ir) return ir;
            if (ec != hipSuccess) return fail(c, CPT_E_HIP, "graph capture failed: %s", hipGetErrorString(ec));
            hipError_t ei = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) return fail(c, CPT_E_HIP, "graph instantiate failed: %s", hipGetErrorString(ei));
            c->graph_key = key;
            captured = true;
        }
    }
    for (size_t i = 0; i < passes.size(); i++) {
        // the guide fraction of this pass (the ramp follows the network's Adam
        // steps; training below synchronises the stream before the next pass)
        c->h_pass[i].guide_frac = guided ? guide_fraction_at(c->g_cfg, c->g_step) : 0.0f;
        HIPCHK(c, hipMemcpyAsync(c->d_pass, &c->h_pass[i], sizeof(PassDesc), hipMemcpyHostToDevice, c->stream));
        if (use_graph) {
            HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
        } else {
            r = issue_pass(c, L, a, c->stream, timing ? &tm : nullptr);
            if (r) return r;
        }
        if (L.ppg) hipLaunchKernelGGL(k_sd_splat, dim3(G), dim3(B), 0, c->stream, a, c->h_pass[i].n_paths);
        if (train) {
            if (timing) tm.mark(-1, c->stream);
            HIPCHK(c, hipMemsetAsync(c->g_count, 0, sizeof(uint32_t), c->stream));
            hipLaunchKernelGGL(k_guide_collect, dim3(G), dim3(B), 0, c->stream, a);
            HIPCHK(c, hipMemcpyAsync(c->h_count, c->g_count, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            int n = (int)c->h_count[0];
            if (n >= 64) {
                for (int k = 0; k < c->g_cfg.train_steps; k++) {
                    r = guide_train_device(c, n, true);
                    if (r) return r;
                    c->g_steps_total++;
                }
                c->g_samples_total += n;
            }
            if (timing) tm.mark(5, c->stream);
        }
    }
    unsigned long long st[ST_N];
    std::vector<unsigned long long> tend(passes.size() + 1);
    HIPCHK(c, hipMemcpyAsync(st, c->stats, sizeof(st), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(tend.data(), c->pass_tend, tend.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    auto t1 = std::chrono::steady_clock::now();
    if (view_ms)
        for (size_t v = 0; v < vd.size(); v++)
            view_ms[v] = (double)(tend[last_pass[v]] - tend[passes.size()]) / 100e3;  // 100 MHz clock
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
        stats->paths = (int64_t)st[ST_PATHS];
        stats->extension_rays = (int64_t)st[ST_EXT];
        stats->shadow_rays = (int64_t)st[ST_SHADOW];
        stats->node_visits = (int64_t)st[ST_NODES];
        stats->triangle_tests = (int64_t)st[ST_TRIS];
        stats->shadow_node_visits = (int64_t)st[ST_SNODES];
        stats->shadow_triangle_tests = (int64_t)st[ST_STRIS];
        stats->passes = (int32_t)passes.size();
        stats->paths_per_pass = (int32_t)slots;
        stats->block_size = B;
        stats->grid_blocks = G;
        stats->guide_samples = c->g_samples_total;
        stats->guide_steps = c->g_steps_total;
        stats->graph_captured = captured ? 1 : 0;
        for (size_t i = 1; i < tm.ev.size(); i++) {
            float ms = 0;
            if (tm.kind[i] >= 0 && hipEventElapsedTime(&ms, tm.ev[i - 1], tm.ev[i]) == hipSuccess) {
                switch (tm.kind[i]) {
                    case 0: stats->ms_extend += ms; stats->n_extend++; break;
                    case 1: stats->ms_shade += ms; stats->n_shade++; break;
                    case 2: stats->ms_shadow += ms; stats->n_shadow++; break;
                    case 4: stats->ms_guide += ms; stats->n_guide++; break;
                    case 5: stats->ms_train += ms; stats->n_train++; break;
                    default: stats->ms_other += ms; break;
                }
            }
        }
    }
    for (auto e : tm.ev) (void)hipEventDestroy(e);
    c->film_w = vd[0].W;
    c->film_h = vd[0].H;
    c->film_spp = vd[0].spp;
    for (auto& v : vd) c->film_layout.emplace_back(v.W, v.H);
    return CPT_OK;
}

int cpt_render(cpt_ctx* c, const cpt_render_params* p, cpt_render_stats* stats) {
    if (!c || !p) return CPT_E_INVALID;
    if (!c->has_scene) return fail(c, CPT_E_NOSCENE, "no scene uploaded");
    if (p->width <= 0 || p->height <= 0 || p->spp_axis <= 0)
        return fail(c, CPT_E_INVALID, "bad render size or spp");
    if (p->flags & CPT_RENDER_STREAM) {
        if (p->flags & CPT_RENDER_LIGHT_TRACE) return fail(c, CPT_E_INVALID, "the light tracer does not stream");
        if (p->flags & CPT_RENDER_ACCUMULATE)
            return fail(c, CPT_E_INVALID, "the streaming renderer starts from a cleared film (no CPT_RENDER_ACCUMULATE)");
        int32_t whs[3] = {p->width, p->height, p->spp_axis};
        return render_views_stream(c, 1, c->cam, whs, p->max_depth, p->seed, p->flags, nullptr, nullptr, stats);
    }
    if (p->flags & CPT_RENDER_PPG) return render_ppg(c, p, stats);
    std::vector<ViewDesc> vd(1);
    std::memcpy(vd[0].cam, c->cam, sizeof(c->cam));
    vd[0].W = p->width;
    vd[0].H = p->height;
    vd[0].S = p->spp_axis;
    vd[0].id = (unsigned)p->view;
    if (p->flags & CPT_RENDER_PPG_STEP) {  // spp_axis = samples per pixel, whole-pixel jitter
        vd[0].S = 1;
        vd[0].spp = (unsigned)p->spp_axis;
        vd[0].strat = 0;
    }
    // parameters left 0: the autotuned configuration for this key, else defaults
    const cpt_ctx::Tuned& T = c->tuned;
    const bool tuned = T.valid && T.scene_gen == c->scene_gen && T.W == p->width && T.H == p->height &&
                       T.depth == p->max_depth && T.guided == ((p->flags & CPT_RENDER_GUIDED) ? 1 : 0);
    int B = p->block_size ? p->block_size : (tuned ? T.B : 128);
    int G = p->grid_blocks ? p->grid_blocks : (tuned && !p->block_size ? T.G : 0);  // 0: resident grid
    int64_t cap = p->max_paths ? p->max_paths : (1 << 25);
    // PPG keeps max_depth vertices per path in flight
    if ((p->flags & CPT_RENDER_PPG_STEP) && !p->max_paths) cap = 1 << 22;
    // online training updates the guide between passes: one sample per pixel per pass
    if ((p->flags & CPT_RENDER_GUIDE_TRAIN) && !p->max_paths) cap = (int64_t)p->width * p->height;
    int SB = p->shade_block ? p->shade_block : (tuned ? T.SB : 256);
    int refill = p->refill ? p->refill : (tuned ? T.refill : (int)CPT_DEFAULT_REFILL);
    return render_batch(c, vd, p->max_depth, p->seed, p->flags, B, G, SB, refill, cap, nullptr, stats);
}

// PPG::render (ppg.h:145-205): iterations of 1, 2, 4, ... samples per pixel
// (the last one below spp_axis^2), each rendered with the SD-tree refined from
// the previous ones' splats; the film keeps the last iteration only.
static int render_ppg(cpt_ctx* c, const cpt_render_params* p, cpt_render_stats* stats) {
    if (p->flags & (CPT_RENDER_GUIDED | CPT_RENDER_STREAM | CPT_RENDER_ACCUMULATE | CPT_RENDER_PPG_STEP))
        return fail(c, CPT_E_INVALID, "CPT_RENDER_PPG runs its own schedule");
    std::vector<int> iters;
    const int64_t total = (int64_t)p->spp_axis * p->spp_axis;
    for (int64_t spp = 1;; spp *= 2) {
        iters.push_back((int)spp);
        if (spp * 2 >= total) break;
    }
    int r = cpt_ppg_reset(c);
    if (r) return r;
    cpt_render_stats acc{};
    for (size_t l = 0; l < iters.size(); l++) {
        cpt_render_params q = *p;
        q.flags = (p->flags & ~CPT_RENDER_PPG) | CPT_RENDER_PPG_STEP;
        q.spp_axis = iters[l];
        q.view = (int32_t)l;
        cpt_render_stats st{};
        r = cpt_render(c, &q, &st);
        if (r) return r;
        if (l + 1 < iters.size()) {
            r = cpt_ppg_update(c);
            if (r) return r;
        }
        acc.seconds += st.seconds;
        acc.paths += st.paths;
        acc.extension_rays += st.extension_rays;
        acc.passes += st.passes;
        acc.paths_per_pass = st.paths_per_pass;
        acc.block_size = st.block_size;
        acc.grid_blocks = st.grid_blocks;
    }
    if (stats) *stats = acc;
    return CPT_OK;
}

int cpt_ppg_reset(cpt_ctx* c) {
    if (!c) return CPT_E_INVALID;
    if (!c->has_scene) return fail(c, CPT_E_NOSCENE, "no scene uploaded");
    c->sd_tree.reset(new cpt::SdTree(c->tb_lo, c->tb_hi));
    c->sd_dirty = true;
    return CPT_OK;
}

int cpt_ppg_update(cpt_ctx* c) {
    if (!c) return CPT_E_INVALID;
    if (!c->sd_tree) return fail(c, CPT_E_INVALID, "no SD-tree (cpt_ppg_reset)");
    if (!c->sd_dirty) {  // the device holds splats made with the current flattening
        const size_t ns = c->sd_flat.slot_dtree.size();
        std::vector<float> sums(ns);
        std::vector<uint32_t> cnt(ns);
        HIPCHK(c, hipMemcpyAsync(sums.data(), c->sd_sum, ns * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(cnt.data(), c->sd_cnt, ns * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->sd_tree->add_splats(c->sd_flat, sums.data(), cnt.data());
    }
    c->sd_tree->update();
    c->sd_dirty = true;
    return CPT_OK;
}

int cpt_ppg_export(cpt_ctx* c, int32_t* sizes, int32_t* snode, int32_t* dinfo, int32_t* quad, float* rect, float* sums,
                   uint32_t* counts, float* bounds) {
    if (!c || !sizes) return CPT_E_INVALID;
    if (!c->sd_tree) return fail(c, CPT_E_INVALID, "no SD-tree (cpt_ppg_reset)");
    int r = sd_upload(c);  // the flattening the next step uses
    if (r) return r;
    const cpt::SdFlat& f = c->sd_flat;
    const size_t ns = f.slot_dtree.size();
    sizes[0] = (int32_t)f.snode.size();
    sizes[1] = (int32_t)(f.dinfo.size() / 4);
    sizes[2] = (int32_t)f.quad.size();
    sizes[3] = (int32_t)(f.rect.size() / 4);
    sizes[4] = (int32_t)ns;
    if (snode) std::memcpy(snode, f.snode.data(), f.snode.size() * 4);
    if (dinfo) std::memcpy(dinfo, f.dinfo.data(), f.dinfo.size() * 4);
    if (quad) std::memcpy(quad, f.quad.data(), f.quad.size() * 4);
    if (rect) std::memcpy(rect, f.rect.data(), f.rect.size() * 4);
    if (sums) HIPCHK(c, hipMemcpyAsync(sums, c->sd_sum, ns * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (counts) HIPCHK(c, hipMemcpyAsync(counts, c->sd_cnt, ns * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (bounds)
        for (int k = 0; k < 3; k++) {
            bounds[k] = c->sd_tree->bmin[k];
            bounds[3 + k] = c->sd_tree->bmax[k];
        }
    return CPT_OK;
}

int cpt_autotune(cpt_ctx* c, const cpt_render_params* p, cpt_tune_result* out) {
    if (!c || !p) return CPT_E_INVALID;
    if (!c->has_scene) return fail(c, CPT_E_NOSCENE, "no scene uploaded");
    if (p->width <= 0 || p->height <= 0) return fail(c, CPT_E_INVALID, "bad render size");
    if (p->flags & (CPT_RENDER_STREAM | CPT_RENDER_GUIDE_TRAIN))
        return fail(c, CPT_E_INVALID, "autotune covers pass renders without guide training");
    auto t0 = std::chrono::steady_clock::now();
    c->tuned.valid = false;
    // trial size: at least 3 x 2^23 paths (a 16-spp 1080p frame at 1080p, 4K and
    // 720p), so that the bounce tails weigh as in a frame-sized pass; smaller
    // trials favoured other block sizes than whole frames do
    const int64_t px = (int64_t)p->width * p->height;
    int S = 1;
    while (px * S * S < (3 << 23)) S++;
    const int flags = p->flags & (CPT_RENDER_GUIDED | CPT_RENDER_NO_GRAPH);
    const int64_t cap = p->max_paths ? p->max_paths : (1 << 25);
    int trials = 0;
    auto time_cfg = [&](int B, int G, int SB, int rf, double& sps) -> int {
        std::vector<ViewDesc> vd(1);
        std::memcpy(vd[0].cam, c->cam, sizeof(c->cam));
        vd[0].W = p->width;
        vd[0].H = p->height;
        vd[0].S = S;
        vd[0].id = (unsigned)p->view;
        cpt_render_stats st;
        // first call captures the graph of this configuration; the better of the next
        // two is kept (one render's timing varies by about as much as the 1% a
        // candidate must gain)
        int r = render_batch(c, vd, p->max_depth, p->seed, flags, B, G, SB, rf, cap, nullptr, &st);
        sps = 0;
        for (int k = 0; k < 2 && !r; k++) {
            r = render_batch(c, vd, p->max_depth, p->seed, flags, B, G, SB, rf, cap, nullptr, &st);
            if (!r) sps = std::max(sps, (double)st.paths / st.seconds);
        }
        if (r) return r;
        trials++;
        return CPT_OK;
    };
    int B = 128, G = 0, SB = 256, rf = (int)CPT_DEFAULT_REFILL;
    double best = 0;
    int r = time_cfg(B, G, SB, rf, best);
    if (r) return r;
    auto try_one = [&](int b, int g, int sb, int f) -> int {
        double sps = 0;
        int rr = time_cfg(b, g, sb, f, sps);
        if (rr) return rr;
        if (sps > best * 1.01) {  // keep the incumbent unless clearly faster
            best = sps;
            B = b;
            G = g;
            SB = sb;
            rf = f;
        }
        return CPT_OK;
    };
    for (int f : {16, 48, 64})
        if ((r = try_one(B, G, SB, f))) return r;
    for (int b : {64, 256})
        if ((r = try_one(b, G, SB, rf))) return r;
    for (int g : {default_grid(c, B), default_grid(c, B) / 2})
        if ((r = try_one(B, g, SB, rf))) return r;
    for (int sb : {128, 512})
        if ((r = try_one(B, G, sb, rf))) return r;
    c->tuned.valid = true;
    c->tuned.scene_gen = c->scene_gen;
    c->tuned.W = p->width;
    c->tuned.H = p->height;
    c->tuned.depth = p->max_depth;
    c->tuned.guided = (p->flags & CPT_RENDER_GUIDED) ? 1 : 0;
    c->tuned.B = B;
    c->tuned.G = G;
    c->tuned.SB = SB;
    c->tuned.refill = rf;
    if (out) {
        std::memset(out, 0, sizeof(*out));
        out->block_size = B;
        out->grid_blocks = G;
        out->shade_block = SB;
        out->refill = rf;
        out->trials = trials;
        out->samples_per_s = best;
        out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return CPT_OK;
}

int cpt_read_film(cpt_ctx* c, float* rgb, double* sums, double* sumsq) {
    if (!c) return CPT_E_INVALID;
    if (!c->film) return fail(c, CPT_E_INVALID, "nothing rendered yet");
    size_t P = (size_t)c->film_w * c->film_h;
    std::vector<double> h(3 * P);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->film, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (sumsq) HIPCHK(c, hipMemcpyAsync(sumsq, c->film2, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (sums) std::memcpy(sums, h.data(), h.size() * sizeof(double));
    if (rgb) {
        double inv = 1.0 / (double)c->film_spp;
        for (size_t i = 0; i < h.size(); i++) rgb[i] = (float)(h[i] * inv);
    }
    return CPT_OK;
}

static int render_views_stream(cpt_ctx* c, int32_t n_views, const float* cameras, const int32_t* wh_spp,
                               int32_t max_depth, uint64_t seed, int32_t flags, float* films, double* view_ms,
                               cpt_render_stats* stats) {
    if (!c || n_views <= 0 || !cameras || !wh_spp) return CPT_E_INVALID;
    if (!c->has_scene) return fail(c, CPT_E_NOSCENE, "no scene uploaded");
    if (max_depth <= 0 || max_depth > 255) return fail(c, CPT_E_INVALID, "bad max_depth");
    if (n_views > 65536) return fail(c, CPT_E_LIMIT, "at most 65536 views");
    HIPCHK(c, hipSetDevice(c->device));
    auto t0 = std::chrono::steady_clock::now();
    std::vector<ViewDesc> vd(n_views);
    unsigned long long total = 0, film_px = 0;
    for (int v = 0; v < n_views; v++) {
        int W = wh_spp[3 * v], H = wh_spp[3 * v + 1], S = wh_spp[3 * v + 2];
        if (W <= 0 || H <= 0 || S <= 0 || (int64_t)W * H > (1 << 26) || S > 1024)
            return fail(c, CPT_E_INVALID, "view %d: bad size %dx%d spp_axis %d", v, W, H, S);
        std::memcpy(vd[v].cam, cameras + 16 * (size_t)v, 16 * sizeof(float));
        vd[v].W = W;
        vd[v].H = H;
        vd[v].S = S;
        vd[v].id = (unsigned)v;
        vd[v].spp = (unsigned)(S * S);
        vd[v].strat = 1;
        vd[v].path_begin = total;
        vd[v].n_paths = (unsigned long long)W * H * S * S;
        vd[v].film_off = film_px;
        total += vd[v].n_paths;
        film_px += (unsigned long long)W * H;
    }
    if (film_px >= CPT_NO_PATH) return fail(c, CPT_E_LIMIT, "films too large");
    const int B = 128;
    int G = default_grid(c, B);
    size_t cap = (size_t)std::min<unsigned long long>(total, 1ull << 22);
    cap = std::max<size_t>(cap, 64);
    int r = ensure_wavefront(c, cap);
    if (r) return r;
    r = ensure_gstack(c, (size_t)G * B);
    if (r) return r;
    r = ensure_film(c, film_px);
    if (r) return r;
    c->film_layout.clear();  // the streaming renderer always starts from a cleared film
    ViewDesc* d_views = nullptr;
    unsigned long long* d_sc = nullptr;  // next_path, t_start, view_done[n], view_tend[n]
    HIPCHK(c, hipMalloc((void**)&d_views, sizeof(ViewDesc) * n_views));
    if (hipMalloc((void**)&d_sc, sizeof(unsigned long long) * (2 + 2 * (size_t)n_views)) != hipSuccess) {
        (void)hipFree(d_views);
        return fail(c, CPT_E_HIP, "hipMalloc failed");
    }
    unsigned long long* h_poll = nullptr;
    if (hipHostMalloc((void**)&h_poll, 4 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
        (void)hipFree(d_views);
        (void)hipFree(d_sc);
        return fail(c, CPT_E_HIP, "hipHostMalloc failed");
    }
    int rc = CPT_OK;
    hipGraphExec_t gx = nullptr;
    do {
        if (hipMemcpyAsync(d_views, vd.data(), sizeof(ViewDesc) * n_views, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipMemsetAsync(c->film, 0, 3 * sizeof(double) * film_px, c->stream) != hipSuccess ||
            hipMemsetAsync(c->film2, 0, 3 * sizeof(double) * film_px, c->stream) != hipSuccess) {
            rc = fail(c, CPT_E_HIP, "view upload failed");
            break;
        }
        StreamArgs sa;
        sa.views = d_views;
        sa.n_views = n_views;
        sa.total = total;
        sa.next_path = d_sc;
        sa.t_start = d_sc + 1;
        sa.view_done = d_sc + 2;
        sa.view_tend = d_sc + 2 + n_views;
        WfArgs a = make_args(c, 0, 0, 1, max_depth, seed, 0);
        a.integrator = (flags & CPT_RENDER_PURE_PT) ? 2 : (flags & CPT_RENDER_NEE_ONLY) ? 1 : 0;
        a.corrected = (flags & CPT_RENDER_CORRECTED) ? 1 : 0;
        LaunchCfg L{B, G, false, true};
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->stream, c->stats, (int)ST_N);
        hipLaunchKernelGGL(k_stream_init, dim3(G), dim3(B), 0, c->stream, a, sa, (unsigned)cap);
        hipLaunchKernelGGL((k_regen<B>), dim3(G), dim3(B), 0, c->stream, a, sa, 0);
        // one chunk = 8 bounce iterations (even, so the live queue is 0 again afterwards)
        const int K = 8;
        auto chunk = [&](hipStream_t st) {
            for (int it = 0; it < K; it++) {
                int cur = it & 1;
                L_extend(L, a, cur, st);
                L_shade(L, a, cur, max_depth, st);
                L_shadow(L, a, cur ^ 1, st);
                hipLaunchKernelGGL((k_regen<B>), dim3(G), dim3(B), 0, st, a, sa, cur ^ 1);
            }
        };
        bool use_graph = !(flags & CPT_RENDER_NO_GRAPH);
        if (use_graph) {
            hipGraph_t g;
            if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                rc = fail(c, CPT_E_HIP, "capture begin failed");
                break;
            }
            chunk(c->stream);
            hipError_t ec = hipStreamEndCapture(c->stream, &g);
            if (ec != hipSuccess) {
                rc = fail(c, CPT_E_HIP, "graph capture failed: %s", hipGetErrorString(ec));
                break;
            }
            hipError_t ei = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) {
                rc = fail(c, CPT_E_HIP, "graph instantiate failed: %s", hipGetErrorString(ei));
                break;
            }
        }
        // Every path ends within max_depth bounces after it starts and a new
        // path starts whenever one ends, so the loop is bounded by
        // (total / cap + 2) * max_depth iterations.
        unsigned long long max_chunks = (total / cap + 2) * (unsigned long long)max_depth / K + 2;
        bool finished = false;
        for (unsigned long long ch = 0; ch < max_chunks; ch++) {
            if (use_graph) {
                if (hipGraphLaunch(gx, c->stream) != hipSuccess) {
                    rc = fail(c, CPT_E_HIP, "graph launch failed");
                    break;
                }
            } else {
                chunk(c->stream);
            }
            if (hipMemcpyAsync(h_poll, d_sc, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipMemcpyAsync(h_poll + 1, c->counters, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess) {
                rc = fail(c, CPT_E_HIP, "progress poll failed");
                break;
            }
            const uint32_t* cnt = reinterpret_cast<const uint32_t*>(h_poll + 1);
            if (h_poll[0] >= total && cnt[0] == 0) {
                finished = true;
                break;
            }
        }
        if (rc) break;
        if (!finished) {
            rc = fail(c, CPT_E_INVALID, "multi-view render did not drain its queues");
            break;
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            rc = fail(c, CPT_E_HIP, "kernel failure: %s", hipGetErrorString(e));
            break;
        }
        std::vector<unsigned long long> hs(2 + 2 * (size_t)n_views);
        unsigned long long st[ST_N];
        std::vector<double> h(3 * (size_t)film_px);
        if (hipMemcpyAsync(hs.data(), d_sc, hs.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipMemcpyAsync(st, c->stats, sizeof(st), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipMemcpyAsync(h.data(), c->film, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            rc = fail(c, CPT_E_HIP, "result download failed");
            break;
        }
        for (int v = 0; v < n_views; v++) {
            if (hs[2 + v] != vd[v].n_paths) {
                rc = fail(c, CPT_E_INVALID, "view %d accumulated %llu of %llu paths", v, hs[2 + v], vd[v].n_paths);
                break;
            }
            if (view_ms) view_ms[v] = (double)(hs[2 + n_views + v] - hs[1]) / 100e3;  // 100 MHz clock
            if (films) {
                double inv = 1.0 / (double)vd[v].spp;
                for (size_t k = 0; k < 3 * (size_t)vd[v].W * vd[v].H; k++)
                    films[3 * vd[v].film_off + k] = (float)(h[3 * vd[v].film_off + k] * inv);
            }
        }
        if (rc) break;
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            stats->paths = (int64_t)total;
            stats->extension_rays = (int64_t)st[ST_EXT];
            stats->shadow_rays = (int64_t)st[ST_SHADOW];
            stats->paths_per_pass = (int32_t)cap;
            stats->block_size = B;
            stats->grid_blocks = G;
        }
    } while (false);
    (void)hipStreamSynchronize(c->stream);
    if (gx) (void)hipGraphExecDestroy(gx);
    (void)hipFree(d_views);
    (void)hipFree(d_sc);
    (void)hipHostFree(h_poll);
    // the film now holds the batch; a later cpt_read_film reads view 0's layout only
    c->film_w = vd[0].W;
    c->film_h = vd[0].H;
    c->film_spp = vd[0].spp;
    if (rc == CPT_OK)
        for (auto& v : vd) c->film_layout.emplace_back(v.W, v.H);
    return rc;
}

int cpt_render_views(cpt_ctx* c, int32_t n_views, const float* cameras, const int32_t* wh_spp, int32_t max_depth,
                     uint64_t seed, int32_t flags, float* films, double* view_ms, cpt_render_stats* stats) {
    if (!c || n_views <= 0 || !cameras || !wh_spp) return CPT_E_INVALID;
    if (!c->has_scene) return fail(c, CPT_E_NOSCENE, "no scene uploaded");
    if (n_views > (1 << 20)) return fail(c, CPT_E_LIMIT, "too many views");
    if (flags & CPT_RENDER_STREAM)
        return render_views_stream(c, n_views, cameras, wh_spp, max_depth, seed, flags, films, view_ms, stats);
    std::vector<ViewDesc> vd(n_views);
    for (int v = 0; v < n_views; v++) {
        std::memcpy(vd[v].cam, cameras + 16 * (size_t)v, 16 * sizeof(float));
        vd[v].W = wh_spp[3 * v];
        vd[v].H = wh_spp[3 * v + 1];
        vd[v].S = wh_spp[3 * v + 2];
        vd[v].id = (unsigned)v;
    }
    const int B = 128;
    int r = render_batch(c, vd, max_depth, seed, flags & ~CPT_RENDER_ACCUMULATE, B, 0, 256, (int)CPT_DEFAULT_REFILL, 1 << 25,
                         view_ms, stats);
    if (r) return r;
    if (films) {
        size_t px = 0;
        for (auto& v : vd) px += (size_t)v.W * v.H;
        std::vector<double> h(3 * px);
        HIPCHK(c, hipMemcpyAsync(h.data(), c->film, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (auto& v : vd) {
            double inv = 1.0 / (double)v.spp;
            for (size_t k = 0; k < 3 * (size_t)v.W * v.H; k++) films[3 * v.film_off + k] = (float)(h[3 * v.film_off + k] * inv);
        }
    }
    return CPT_OK;
}

// ------------------------------------------------------------------ learned guide
// two per CU (the training kernel's block uses 75 KB of LDS at width 64)
static const int GUIDE_TRAIN_BLOCKS = 512;
static const int GUIDE_DW_BLOCKS = 256;
static const int GUIDE_WSUM_BLOCKS = 128;

int cpt_guide_num_params(void) { return guide::NPARAM; }

}  // extern "C"

namespace {

// guide buffers of one network width (freed first when the width changes)
template <class N>
int guide_alloc(cpt_ctx* c) {
    if (c->g_width == N::kHid && c->g_params) return CPT_OK;
    drop_graph(c);
    dfree(c->g_params);
    dfree(c->g_m);
    dfree(c->g_v);
    dfree(c->g_grad);
    dfree(c->g_wt);
    dfree(c->g_wt8);
    dfree(c->g_partial);
    dfree(c->g_rows);
    c->g_rows_cap = 0;
    c->g_width = 0;
    int r = 0;
    r |= dalloc(c, c->g_params, N::kParams);
    r |= dalloc(c, c->g_m, N::kParams);
    r |= dalloc(c, c->g_v, N::kParams);
    r |= dalloc(c, c->g_grad, N::kParams);
    r |= dalloc(c, c->g_wt, N::kWt);
    r |= dalloc(c, c->g_wt8, N::kWq);
    r |= dalloc(c, c->g_partial, (size_t)GUIDE_DW_BLOCKS * N::kParams);
    if (r) return CPT_E_HIP;
    HIPCHK(c, hipFuncSetAttribute((const void*)N::mlp_forward, hipFuncAttributeMaxDynamicSharedMemorySize, N::kSmemFwd));
    HIPCHK(c, hipFuncSetAttribute((const void*)N::train_tile, hipFuncAttributeMaxDynamicSharedMemorySize, N::kSmemTrain));
    HIPCHK(c, hipFuncSetAttribute((const void*)N::mlp_forward8, hipFuncAttributeMaxDynamicSharedMemorySize, N::kSmemFwd8));
    HIPCHK(c, hipFuncSetAttribute((const void*)k_guide_infer<N::kHid>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  N::kSmemFwd));
    HIPCHK(c, hipFuncSetAttribute((const void*)k_guide_infer8<N>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  N::kSmemFwd8));
    HIPCHK(c, hipFuncSetAttribute((const void*)N::mlp_mixtures, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  N::kSmemFwd));
    HIPCHK(c, hipFuncSetAttribute((const void*)N::mlp_mixtures8, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  N::kSmemFwd8));
    if constexpr (N::kWave16) {
        HIPCHK(c, hipFuncSetAttribute((const void*)k_guide_infer16<N>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      N::kSmemInf16));
        HIPCHK(c, hipFuncSetAttribute((const void*)N::mlp_mixtures16, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      N::kSmemInf16));
        HIPCHK(c, hipFuncSetAttribute((const void*)N::mlp_forward16, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      N::kSmemInf16));
    }
    c->g_width = N::kHid;
    return CPT_OK;
}

// BF16 transposed copy and FP8 copy of the parameters at c->g_params
template <class N>
void guide_pack(cpt_ctx* c) {
    hipLaunchKernelGGL(N::pack_weights, dim3(64), dim3(256), 0, c->stream, c->g_params, c->g_wt);
    hipLaunchKernelGGL(N::pack_weights2, dim3(64), dim3(256), 0, c->stream, c->g_params, c->g_wt);
    hipLaunchKernelGGL(N::quant8, dim3(5), dim3(256), 0, c->stream, c->g_params, c->g_wt8, c->g_inv8);
}

// Glorot-uniform weights from a seeded splitmix64 stream; zero biases except the
// lobe means, which start at the 8 cube diagonals (a near-uniform mixture; with
// 16 lobes each diagonal twice)
template <class N>
int guide_init_params(cpt_ctx* c, uint64_t seed) {
    std::vector<float> h(N::kParams, 0.0f);
    uint64_t st = seed ^ 0x243F6A8885A308D3ull;
    auto rnd = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (float)((z >> 40) * (1.0 / 16777216.0)) * 2.0f - 1.0f;
    };
    const int ks[5] = {guide::K0, N::kHid, N::kHid, N::kHid, N::kHid};
    const int ns[5] = {N::kHid, N::kHid, N::kHid, N::kHid, N::kOut};
    for (int l = 0; l < 5; l++) {
        float a = std::sqrt(6.0f / (float)(ks[l] + ns[l]));
        if (l == 4) a *= 0.1f;
        for (int i = 0; i < ks[l] * ns[l]; i++) h[N::kW[l] + i] = a * rnd();
    }
    for (int k = 0; k < N::kLobes; k++) {
        h[N::kB4 + 5 * k + 2] = (k & 1) ? 1.0f : -1.0f;
        h[N::kB4 + 5 * k + 3] = (k & 2) ? 1.0f : -1.0f;
        h[N::kB4 + 5 * k + 4] = (k & 4) ? 1.0f : -1.0f;
    }
    HIPCHK(c, hipMemcpy(c->g_params, h.data(), N::kParams * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(c->g_m, 0, N::kParams * sizeof(float)));
    HIPCHK(c, hipMemset(c->g_v, 0, N::kParams * sizeof(float)));
    guide_pack<N>(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CPT_OK;
}

template <class N>
int guide_forward_impl(cpt_ctx* c, const float* d_in, int n, float* d_out) {
    int G = std::min(c->cus, (n + N::kTile - 1) / N::kTile);
    if (c->g_cfg.infer_fp8)
        hipLaunchKernelGGL(N::mlp_forward8, dim3(std::min(c->cus, (n + 16 * N::kWaves8 - 1) / (16 * N::kWaves8))),
                           dim3(64 * N::kWaves8), N::kSmemFwd8, c->stream, c->g_wt8, c->g_params, c->g_inv8, d_in, n,
                           d_out);
    else if constexpr (N::kWave16)
        hipLaunchKernelGGL(N::mlp_forward16, dim3(std::min(c->cus, (n + 16 * N::kWaves16 - 1) / (16 * N::kWaves16))),
                           dim3(64 * N::kWaves16), N::kSmemInf16, c->stream, c->g_wt, c->g_params, d_in, n, d_out);
    else
        hipLaunchKernelGGL(N::mlp_forward, dim3(G), dim3(256), N::kSmemFwd, c->stream, c->g_wt, c->g_params, d_in, n,
                           d_out);
    return hipGetLastError() == hipSuccess ? CPT_OK : CPT_E_HIP;
}

// One Adam step on device-resident samples (n x 11 floats at c->g_samples).
template <class N>
int guide_train_impl(cpt_ctx* c, int n, bool apply) {
    // per-sample rows, feature-major over n rounded up to a tile (padding samples zero)
    const int npad = (n + N::kTile - 1) / N::kTile * N::kTile;
    if (c->g_rows_cap < (size_t)npad || !c->g_rows) {
        c->g_rows_cap = 0;
        int r = dalloc(c, c->g_rows, (size_t)npad * N::kRow);
        if (r) return r;
        c->g_rows_cap = npad;
    }
    int per_block = ((npad + GUIDE_DW_BLOCKS - 1) / GUIDE_DW_BLOCKS + N::kChunk - 1) / N::kChunk * N::kChunk;
    hipLaunchKernelGGL(N::weight_partial, dim3(GUIDE_WSUM_BLOCKS), dim3(256), 0, c->stream, c->g_samples, n,
                       c->g_wpart);
    hipLaunchKernelGGL(N::weight_scale, dim3(1), dim3(256), 0, c->stream, c->g_wpart, GUIDE_WSUM_BLOCKS, c->g_wscale);
    hipLaunchKernelGGL(N::train_tile, dim3(GUIDE_TRAIN_BLOCKS), dim3(256), N::kSmemTrain, c->stream, c->g_wt,
                       c->g_params, c->g_samples, n, npad, c->g_wscale, c->g_rows, c->g_loss_part);
    hipLaunchKernelGGL(N::train_dw, dim3(GUIDE_DW_BLOCKS, N::kDwGroups), dim3(256), 0, c->stream, c->g_rows, npad,
                       per_block, c->g_partial);
    // one thread per parameter: each sums its GUIDE_DW_BLOCKS partials in block order
    hipLaunchKernelGGL(N::reduce_partials, dim3((N::kParams + 255) / 256), dim3(256), 0, c->stream, c->g_partial,
                       GUIDE_DW_BLOCKS,
                       N::kParams, N::kParams, c->g_grad);
    if (apply) {
        c->g_step++;
        hipLaunchKernelGGL(N::adam, dim3(64), dim3(256), 0, c->stream, c->g_params, c->g_m, c->g_v, c->g_grad,
                           c->g_step, c->g_cfg.lr, c->g_cfg.beta1, c->g_cfg.beta2, c->g_cfg.eps, c->g_wt);
        if (c->g_cfg.infer_fp8)
            hipLaunchKernelGGL(N::quant8, dim3(5), dim3(256), 0, c->stream, c->g_params, c->g_wt8, c->g_inv8);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, CPT_E_HIP, "guide training launch failed: %s", hipGetErrorString(e));
    return CPT_OK;
}

int guide_nparams(const cpt_ctx* c) { return c->g_width == 128 ? guide128::NPARAM : guide::NPARAM; }

}  // namespace

static int guide_train_device(cpt_ctx* c, int n, bool apply) {
    return c->g_width == 128 ? guide_train_impl<guide128::Net>(c, n, apply) : guide_train_impl<guide::Net>(c, n, apply);
}

extern "C" {

int cpt_guide_init(cpt_ctx* c, const cpt_guide_params* gp) {
    if (!c || !gp) return CPT_E_INVALID;
    if (gp->width != 0 && gp->width != 64 && gp->width != 128)
        return fail(c, CPT_E_INVALID, "guide width %d: 64 or 128", (int)gp->width);
    HIPCHK(c, hipSetDevice(c->device));
    // work of an earlier call (a render that failed part way) may still read the parameters
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const bool wide = gp->width == 128;
    int r = wide ? guide_alloc<guide128::Net>(c) : guide_alloc<guide::Net>(c);
    if (r) return r;
    if (!c->g_count || !c->h_count) {
        if (c->h_count) (void)hipHostFree(c->h_count);
        c->h_count = nullptr;
        r |= dalloc(c, c->g_inv8, 8);
        r |= dalloc(c, c->g_loss_part, GUIDE_TRAIN_BLOCKS);
        r |= dalloc(c, c->g_wscale, 1);
        r |= dalloc(c, c->g_wpart, GUIDE_WSUM_BLOCKS);
        r |= dalloc(c, c->g_count, 4);
        if (r) {
            dfree(c->g_count);  // retried as a whole by the next call
            return CPT_E_HIP;
        }
        hipError_t e = hipHostMalloc((void**)&c->h_count, 4 * sizeof(uint32_t), hipHostMallocDefault);
        if (e != hipSuccess) {
            c->h_count = nullptr;
            return fail(c, CPT_E_HIP, "hipHostMalloc failed: %s", hipGetErrorString(e));
        }
    }
    c->g_cfg = *gp;
    c->g_cfg.width = wide ? 128 : 64;
    if (c->g_cfg.lr <= 0) c->g_cfg.lr = 1e-3f;
    if (c->g_cfg.beta1 <= 0) c->g_cfg.beta1 = 0.9f;
    if (c->g_cfg.beta2 <= 0) c->g_cfg.beta2 = 0.99f;
    if (c->g_cfg.eps <= 0) c->g_cfg.eps = 1e-8f;
    if (c->g_cfg.guide_fraction <= 0 || c->g_cfg.guide_fraction >= 1) c->g_cfg.guide_fraction = 0.5f;
    if (c->g_cfg.train_paths <= 0) c->g_cfg.train_paths = 65536;
    if (c->g_cfg.train_steps <= 0) c->g_cfg.train_steps = 1;
    r = wide ? guide_init_params<guide128::Net>(c, gp->seed) : guide_init_params<guide::Net>(c, gp->seed);
    if (r) return r;
    c->g_step = 0;
    c->has_guide = true;
    return CPT_OK;
}

int cpt_guide_shape(cpt_ctx* c, int32_t* shape4) {
    if (!c || !shape4) return CPT_E_INVALID;
    const bool wide = c->has_guide && c->g_width == 128;
    shape4[0] = wide ? guide128::HID : guide::HID;
    shape4[1] = wide ? guide128::NLOBES : guide::NLOBES;
    shape4[2] = wide ? guide128::NOUT : guide::NOUT;
    shape4[3] = wide ? guide128::NPARAM : guide::NPARAM;
    return CPT_OK;
}

int cpt_guide_get_params(cpt_ctx* c, float* out) {
    if (!c || !out) return CPT_E_INVALID;
    if (!c->has_guide) return fail(c, CPT_E_INVALID, "no guide: call cpt_guide_init");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, c->g_params, guide_nparams(c) * sizeof(float), hipMemcpyDeviceToHost));
    return CPT_OK;
}

int cpt_guide_set_params(cpt_ctx* c, const float* in) {
    if (!c || !in) return CPT_E_INVALID;
    if (!c->has_guide) return fail(c, CPT_E_INVALID, "no guide: call cpt_guide_init");
    const int np = guide_nparams(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->g_params, in, np * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(c->g_m, 0, np * sizeof(float)));
    HIPCHK(c, hipMemset(c->g_v, 0, np * sizeof(float)));
    if (c->g_width == 128) guide_pack<guide128::Net>(c);
    else guide_pack<guide::Net>(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->g_step = 0;
    return CPT_OK;
}

int cpt_guide_forward(cpt_ctx* c, const float* inputs, int32_t n, float* outputs) {
    if (!c || n < 0 || (n > 0 && (!inputs || !outputs))) return CPT_E_INVALID;
    if (!c->has_guide) return fail(c, CPT_E_INVALID, "no guide: call cpt_guide_init");
    if (n == 0) return CPT_OK;
    const int nin = guide::NIN, nout = c->g_width == 128 ? guide128::NOUT : guide::NOUT;
    float *d_in = nullptr, *d_out = nullptr;
    HIPCHK(c, hipMalloc((void**)&d_in, sizeof(float) * nin * (size_t)n));
    if (hipMalloc((void**)&d_out, sizeof(float) * nout * (size_t)n) != hipSuccess) {
        (void)hipFree(d_in);
        return fail(c, CPT_E_HIP, "hipMalloc failed");
    }
    int rc = CPT_OK;
    if (hipMemcpyAsync(d_in, inputs, sizeof(float) * nin * (size_t)n, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = fail(c, CPT_E_HIP, "upload failed");
    if (!rc) {
        int lr = c->g_width == 128 ? guide_forward_impl<guide128::Net>(c, d_in, n, d_out)
                                   : guide_forward_impl<guide::Net>(c, d_in, n, d_out);
        if (lr != CPT_OK ||
            hipMemcpyAsync(outputs, d_out, sizeof(float) * nout * (size_t)n, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = fail(c, CPT_E_HIP, "guide forward failed");
    }
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    return rc;
}

int cpt_guide_mixtures(cpt_ctx* c, const float* inputs, int32_t n, float* mixtures) {
    if (!c || n < 0 || (n > 0 && (!inputs || !mixtures))) return CPT_E_INVALID;
    if (!c->has_guide) return fail(c, CPT_E_INVALID, "no guide: call cpt_guide_init");
    if (n == 0) return CPT_OK;
    HIPCHK(c, hipSetDevice(c->device));
    const bool wide = c->g_width == 128;
    const size_t nin = guide::NIN, nmix = 3 * (size_t)(wide ? guide128::NLOBES : guide::NLOBES);
    float *d_in = nullptr, *d_mix = nullptr;
    HIPCHK(c, hipMalloc((void**)&d_in, sizeof(float) * nin * (size_t)n));
    if (hipMalloc((void**)&d_mix, sizeof(float) * nmix * (size_t)n) != hipSuccess) {
        (void)hipFree(d_in);
        return fail(c, CPT_E_HIP, "hipMalloc failed");
    }
    int rc = CPT_OK;
    if (hipMemcpyAsync(d_in, inputs, sizeof(float) * nin * (size_t)n, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = fail(c, CPT_E_HIP, "upload failed");
    if (!rc) {
        const size_t stride = (size_t)n;
        if (c->g_cfg.infer_fp8) {
            const int w8 = wide ? guide128::Net::kWaves8 : guide::Net::kWaves8;
            const int G = std::min(c->cus, (n + 64 * w8 - 1) / (64 * w8));
            if (wide)
                hipLaunchKernelGGL(guide128::Net::mlp_mixtures8, dim3(G), dim3(64 * w8), guide128::Net::kSmemFwd8,
                                   c->stream, c->g_wt8, c->g_params, c->g_inv8, d_in, n, d_mix, stride);
            else
                hipLaunchKernelGGL(guide::Net::mlp_mixtures8, dim3(G), dim3(64 * w8), guide::Net::kSmemFwd8, c->stream,
                                   c->g_wt8, c->g_params, c->g_inv8, d_in, n, d_mix, stride);
        } else if (!wide) {
            const int w16 = guide::Net::kWaves16;
            const int G = std::min(c->cus, (n + 64 * w16 - 1) / (64 * w16));
            hipLaunchKernelGGL(guide::Net::mlp_mixtures16, dim3(G), dim3(64 * w16), guide::Net::kSmemInf16, c->stream,
                               c->g_wt, c->g_params, d_in, n, d_mix, stride);
        } else {
            const int G = std::min(c->cus, (n + guide::TILE - 1) / guide::TILE);
            if (wide)
                hipLaunchKernelGGL(guide128::Net::mlp_mixtures, dim3(G), dim3(256), guide128::Net::kSmemFwd, c->stream,
                                   c->g_wt, c->g_params, d_in, n, d_mix, stride);
            else
                hipLaunchKernelGGL(guide::Net::mlp_mixtures, dim3(G), dim3(256), guide::Net::kSmemFwd, c->stream,
                                   c->g_wt, c->g_params, d_in, n, d_mix, stride);
        }
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(mixtures, d_mix, sizeof(float) * nmix * (size_t)n, hipMemcpyDeviceToHost, c->stream) !=
                hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = fail(c, CPT_E_HIP, "guide mixtures failed");
    }
    (void)hipFree(d_in);
    (void)hipFree(d_mix);
    return rc;
}

int cpt_debug_guide_mixtures(cpt_ctx* c, int32_t n, float* mixtures) {
    if (!c || n < 0 || (n > 0 && !mixtures)) return CPT_E_INVALID;
    if (!c->has_guide || !c->g_lobes) return fail(c, CPT_E_INVALID, "no guided render to read back");
    const int L = c->g_width == 128 ? guide128::NLOBES : guide::NLOBES;
    const size_t stride = c->g_lobes_cap / (3 * (size_t)L);
    if ((size_t)n > stride) return fail(c, CPT_E_LIMIT, "only %zu queue entries are held", stride);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int v = 0; v < 3 * L; v++)
        HIPCHK(c, hipMemcpy(mixtures + (size_t)v * n, c->g_lobes + (size_t)v * stride, sizeof(float) * (size_t)n,
                            hipMemcpyDeviceToHost));
    return CPT_OK;
}

int cpt_guide_train_step(cpt_ctx* c, const float* samples, int32_t n, int32_t apply, float* loss, float* grad) {
    using namespace guide;
    if (!c || n <= 0 || !samples) return CPT_E_INVALID;
    if (!c->has_guide) return fail(c, CPT_E_INVALID, "no guide: call cpt_guide_init");
    HIPCHK(c, hipSetDevice(c->device));
    if (c->g_samples_cap < (size_t)n || !c->g_samples) {
        drop_graph(c);  // guided renders' graphs hold the sample buffer's address
        c->g_samples_cap = 0;
        int r = dalloc(c, c->g_samples, (size_t)n * SAMPLE_FLOATS);
        if (r) return r;
        c->g_samples_cap = n;
    }
    HIPCHK(c, hipMemcpyAsync(c->g_samples, samples, sizeof(float) * SAMPLE_FLOATS * (size_t)n, hipMemcpyHostToDevice, c->stream));
    int r = guide_train_device(c, n, apply != 0);
    if (r) return r;
    std::vector<float> lp(GUIDE_TRAIN_BLOCKS);
    HIPCHK(c, hipMemcpyAsync(lp.data(), c->g_loss_part, sizeof(float) * GUIDE_TRAIN_BLOCKS, hipMemcpyDeviceToHost, c->stream));
    if (grad) HIPCHK(c, hipMemcpyAsync(grad, c->g_grad, sizeof(float) * guide_nparams(c), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (loss) {
        double s = 0;
        for (float v : lp) s += v;
        *loss = (float)s;
    }
    return CPT_OK;
}

}  // extern "C"
