// CPU oracle for the cpp-pt PathtracerMIS hot path -- TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
// library; the product never does.  Every function restates the reference
// algorithm named in its comment (paths relative to the cpp-pt source tree)
// with the reference's float/double promotion rules, e.g. `v / a` on a Vec3 is
// v * float(1.0/double(a)) (math.h:103-105) and `x + EPS` with EPS = 1e-4 is a
// double add (math.h:11).  Built with -ffp-contract=off so that no FMA fuses
// what the reference rounds twice.
//
// libm note: cos/sin/sqrt are called in double, the way unqualified calls on
// float resolve with libstdc++ (::cos(double)).  The per-path bit pattern of
// sampling through sin/cos is therefore "parity unpinned" against an MSVC
// build of the reference (which would pick cosf); only statistical parity is
// claimed for full paths.
#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <vector>
#include <omp.h>

namespace {

const double EPSD = 1e-4;               // math.h:11
const float EPSF = (float)1e-4;         // EPS converted to float where a Vec3 is scaled by it

struct V3 {
    float x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(float a) : x(a), y(a), z(a) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
};
inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(V3 a, V3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 operator*(V3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
// math.h:103-105: operator/(T a) = (*this)*(1.0/a), 1.0/a in double, converted to float
inline V3 divs(V3 a, float s) { return a * (float)(1.0 / (double)s); }
inline V3 neg(V3 a) { return a * (float)(-1.0); }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float length(V3 a) { return (float)std::sqrt((double)dot(a, a)); }
inline V3 normalized(V3 a) { return divs(a, length(a)); }
inline V3 vmin(V3 a, V3 b) { return V3(a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z); }
inline V3 vmax(V3 a, V3 b) { return V3(a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z); }

// Mat3 built from three columns (math.h:238-248), Mat3*Vec3 (math.h:319-325)
struct M3 {
    float v[9];
    M3(V3 c0, V3 c1, V3 c2) {
        v[0] = c0.x; v[1] = c1.x; v[2] = c2.x;
        v[3] = c0.y; v[4] = c1.y; v[5] = c2.y;
        v[6] = c0.z; v[7] = c1.z; v[8] = c2.z;
    }
    M3() { for (float& f : v) f = 0; }
    V3 mul(V3 a) const {
        return V3(v[0] * a.x + v[1] * a.y + v[2] * a.z, v[3] * a.x + v[4] * a.y + v[5] * a.z,
                  v[6] * a.x + v[7] * a.y + v[8] * a.z);
    }
    M3 transpose() const {
        M3 m;
        m.v[0] = v[0]; m.v[1] = v[3]; m.v[2] = v[6];
        m.v[3] = v[1]; m.v[4] = v[4]; m.v[5] = v[7];
        m.v[6] = v[2]; m.v[7] = v[5]; m.v[8] = v[8];
        return m;
    }
};

// ---------------------------------------------------------------- samplers
struct Sampler {
    virtual float sample() = 0;
    virtual ~Sampler() {}
};

inline uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// PCG32 (O'Neill 2014, pcg32_srandom_r / pcg32_random_r, XSH-RR output),
// seeded per path: initstate = splitmix64(splitmix64(seed) ^ path_id), fixed increment.
struct PcgSampler : Sampler {
    uint64_t state, inc;
    PcgSampler(uint64_t seed, uint64_t path_id) {
        state = 0;
        inc = (0x5851F42D4C957F2Dull << 1) | 1u;
        next();
        state += splitmix64(splitmix64(seed) ^ path_id);
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31));
    }
    float sample() override { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

// sampler.h:13-23 RandomSampler: std::mt19937(seed) + uniform_real_distribution<float>(0,1)
struct MtSampler : Sampler {
    std::mt19937 gen;
    std::uniform_real_distribution<float> dist;
    MtSampler(int seed) : gen(seed), dist(0.0, 1.0) {}
    float sample() override { return dist(gen); }
};

// ---------------------------------------------------------------- scene
struct Material {
    int type;  // 0 diffuse, 1 glossy, 2 emissive, 3 dielectric
    V3 albedo;
    float rough;
    V3 emission;
    float ior;
    V3 nmap;
    bool has_nmap;
};

struct Camera {
    V3 o;
    M3 coords;
    float tan_fovy, aspect;
    int resX, resY;
};

struct RNode {
    V3 bmin, bmax;
    int left, right;  // child node indices ("right" = first visited, bvh.h:14-18,33)
    int first, count; // leaf range into leaf_prims (count >= 0 only for leaves)
    bool leaf;
};

struct Scene {
    std::vector<V3> P, N, B;
    std::vector<std::array<int, 3>> F;
    std::vector<int> tri_mat;
    std::vector<Material> mats;
    std::vector<int> lights;       // LightGroup order: triangle id, or -1-k for point light k
    std::vector<std::array<float, 6>> point_lights;  // position, radiance
    std::vector<int> light_of_tri; // -1 if not emissive
    Camera cam;
    V3 env;
    std::vector<RNode> nodes;
    std::vector<int> leaf_prims;
    std::vector<int> rank;
    int root = -1;
    // 0 PathtracerMIS (pathtracerMIS.h:29-130), 1 Pathtracer: NEE without MIS
    // (pathtracer.h:27-110), 2 PurePt: BSDF sampling only (pure_pt.h:27-67)
    int integrator = 0;
    // the reference's biased quirks switched off (the library's CPT_RENDER_CORRECTED):
    // MIS from the pre-roulette pdf, point lights weighted 1 in NEE, survival
    // probability clamped at 1, normal maps at every hit, light tracer aspect ratio
    bool corrected = false;
    uint32_t view = 0;  // view id in the per-path PCG stream id (multi-view batches)
    // image textures: per-vertex uv, images (w, h, channels, texel offset),
    // per material the image of its albedo, roughness, metal and normal slot (-1: constant)
    std::vector<std::array<float, 2>> UV;
    std::vector<std::array<int, 4>> images;
    std::vector<float> texels;
    std::vector<std::array<int, 4>> mat_tex;
};

struct Ray {
    V3 o, d;
    float max_t;
};

// ---------------------------------------------------------------- reference BVH
struct AABB {
    V3 mn, mx;
    AABB() : mn(std::numeric_limits<float>::infinity()), mx(-std::numeric_limits<float>::infinity()) {}
    void unite(const AABB& o) { mn = vmin(mn, o.mn); mx = vmax(mx, o.mx); }
    void unite(V3 p) { mn = vmin(mn, p); mx = vmax(mx, p); }
    V3 center() const { return (mn + mx) * (float)0.5; }  // aabb.h:62-64
    // aabb.h:77-82 (fabs returns double; the products are double)
    float area() const {
        return (float)(2.0 * (std::fabs((double)(mx.x - mn.x)) * (double)(mx.y - mn.y) +
                              std::fabs((double)(mx.z - mn.z)) * (double)(mx.y - mn.y) +
                              std::fabs((double)(mx.x - mn.x)) * (double)(mx.z - mn.z)));
    }
};

// math.h:699-705
inline void order2(float& mn, float& mx) {
    if (mx >= mn) return;
    float t = mn;
    mn = mx;
    mx = t;
}

// aabb.h:19-61, literal (including the idiosyncratic max3/min3 branches)
bool aabb_hit(V3 bmin, V3 bmax, const Ray& r) {
    V3 dmin = bmin - r.o, dmax = bmax - r.o;
    V3 nmin(dmin.x / r.d.x, dmin.y / r.d.y, dmin.z / r.d.z);
    V3 nmax(dmax.x / r.d.x, dmax.y / r.d.y, dmax.z / r.d.z);
    order2(nmin.x, nmax.x);
    order2(nmin.y, nmax.y);
    order2(nmin.z, nmax.z);
    float tmin = nmin.x;
    if (tmin < nmin.y) tmin = nmin.y;
    if (tmin < nmin.z && nmin.z > nmin.y) tmin = nmin.z;
    float tmax = nmax.x;
    if (tmax > nmax.y) tmax = nmax.y;
    if (tmax > nmax.z && nmax.z < nmax.y) tmax = nmax.z;
    if (tmax < tmin) return false;
    if (tmin > 0 && (double)tmin < (double)r.max_t + EPSD) return true;
    if (tmax > 0) return true;
    return false;
}

AABB tri_bounds(const Scene& s, int t) {  // triangle.h:181-195
    V3 a = s.P[s.F[t][0]], b = s.P[s.F[t][1]], c = s.P[s.F[t][2]];
    AABB bb;
    bb.mn = vmin(vmin(a, b), c);
    bb.mx = vmax(vmax(a, b), c);
    return bb;
}

// bvh.h:57-149 compute_split: 5 equal-width centroid bins, cost = left.area()+right.area()
std::pair<std::vector<int>, std::vector<int>> compute_split(const std::vector<AABB>& pb, int axis,
                                                            const std::vector<int>& to_split,
                                                            float& cost, const AABB& bounds) {
    const int NP = 5;
    std::vector<std::vector<int>> parts(NP);
    for (int i : to_split) {
        bool assigned = false;
        V3 c = pb[i].center();
        float cv = axis == 0 ? c.x : axis == 1 ? c.y : c.z;
        float lo = axis == 0 ? bounds.mn.x : axis == 1 ? bounds.mn.y : bounds.mn.z;
        float hi = axis == 0 ? bounds.mx.x : axis == 1 ? bounds.mx.y : bounds.mx.z;
        for (int j = 0; j < NP - 1; j++) {
            if (cv < lo + (hi - lo) * float(j + 1) / float(NP)) {
                parts[j].push_back(i);
                assigned = true;
                break;
            }
        }
        if (!assigned) parts[NP - 1].push_back(i);
    }
    std::vector<float> pc(NP - 1);
    for (int i = 0; i < NP - 1; i++) {
        AABB l, r;
        int lc = 0, rc = 0;
        for (int j = 0; j < NP; j++)
            for (int k : parts[j]) {
                if (j <= i) { lc++; l.unite(pb[k]); }
                else { rc++; r.unite(pb[k]); }
            }
        pc[i] = (lc == 0 || rc == 0) ? std::numeric_limits<float>::infinity() : l.area() + r.area();
    }
    int mi = 0;
    float mc = pc[0];
    for (int i = 1; i < NP - 1; i++)
        if (pc[i] < mc) { mc = pc[i]; mi = i; }
    cost = mc;
    std::pair<std::vector<int>, std::vector<int>> ret;
    for (int i = 0; i <= mi; i++) ret.first.insert(ret.first.end(), parts[i].begin(), parts[i].end());
    for (int i = mi + 1; i < NP; i++) ret.second.insert(ret.second.end(), parts[i].begin(), parts[i].end());
    return ret;
}

// bvh.h:151-202 build_node.  Deviation: a split that leaves one side empty
// (all centroids coincide) becomes a leaf instead of recursing forever
// (the reference's TODO at bvh.h:200).
int build_node(Scene& s, const std::vector<AABB>& pb, std::vector<int>& prims) {
    int id = (int)s.nodes.size();
    s.nodes.push_back(RNode());
    if (prims.size() < 4) {
        RNode& n = s.nodes[id];
        n.leaf = true;
        n.first = (int)s.leaf_prims.size();
        n.count = (int)prims.size();
        n.left = n.right = -1;
        for (int p : prims) s.leaf_prims.push_back(p);
        return id;
    }
    AABB bounds, cb;
    for (int p : prims) {
        bounds.unite(pb[p]);
        cb.unite(pb[p].center());
    }
    float c1 = 0, c2 = 0, c3 = 0;
    auto p1 = compute_split(pb, 0, prims, c1, cb);
    auto p2 = compute_split(pb, 1, prims, c2, cb);
    auto p3 = compute_split(pb, 2, prims, c3, cb);
    std::vector<int> left, right;
    if (c1 < c2) {
        if (c1 < c3) { left = p1.first; right = p1.second; }
        else if (c2 < c3) { left = p2.first; right = p2.second; }
        else { left = p3.first; right = p3.second; }
    } else if (c2 < c3) { left = p2.first; right = p2.second; }
    else { left = p3.first; right = p3.second; }
    if (left.empty() || right.empty()) {
        RNode& n = s.nodes[id];
        n.leaf = true;
        n.first = (int)s.leaf_prims.size();
        n.count = (int)prims.size();
        n.left = n.right = -1;
        for (int p : prims) s.leaf_prims.push_back(p);
        return id;
    }
    // bvh.h:201 + ctor bvh.h:18: member `right` holds build_node(left) and is visited first
    int a = build_node(s, pb, left);
    int b = build_node(s, pb, right);
    RNode& n = s.nodes[id];
    n.leaf = false;
    n.bmin = bounds.mn;
    n.bmax = bounds.mx;
    n.right = a;
    n.left = b;
    n.first = n.count = -1;
    return id;
}

void build_bvh(Scene& s) {
    int nt = (int)s.F.size();
    std::vector<AABB> pb(nt);
    for (int i = 0; i < nt; i++) pb[i] = tri_bounds(s, i);
    std::vector<int> prims(nt);
    for (int i = 0; i < nt; i++) prims[i] = i;
    s.nodes.reserve(2 * nt / 3 + 8);
    s.root = build_node(s, pb, prims);
    // DFS test order (right child first, leaf insertion order) == leaf_prims order
    s.rank.assign(nt, -1);
    for (int i = 0; i < (int)s.leaf_prims.size(); i++) s.rank[s.leaf_prims[i]] = i;
}

// ---------------------------------------------------------------- triangle
struct Hit {
    int tri = -1;
    float t = 0, beta = 0, gamma = 0;
};

// triangle.h:18-77 / 117-177: Cramer's rule on [b-a, c-a, -d].  Returns
// true and (t, beta, gamma) when t in [EPS, max_t] and barycentrics in [0,1].
// Deviation shared with the library: a ray lying in the triangle's plane
// (determinant 0, t = 0 * inf = NaN) misses.  The reference takes it as a hit
// whose NaN max_t then lets every later triangle through, so its answer would
// depend on the test order of the whole scene.
inline bool tri_test(const Scene& s, int id, const Ray& r, float& t, float& beta, float& gamma) {
    const auto& f = s.F[id];
    V3 a = s.P[f[0]], b = s.P[f[1]], c = s.P[f[2]];
    V3 v1 = b - a, v2 = c - a, v3 = neg(r.d), vs = r.o - a;
    // `-r.d` is the unary minus of math.h:95-97 (exact negation, same as *(-1))
    v3 = V3(-r.d.x, -r.d.y, -r.d.z);
    float inv = (float)(1.0 / (double)(v1.x * v2.y * v3.z + v2.x * v3.y * v1.z + v3.x * v1.y * v2.z -
                                       v3.x * v2.y * v1.z - v3.y * v2.z * v1.x - v3.z * v2.x * v1.y));
    t = (v1.x * v2.y * vs.z + v2.x * vs.y * v1.z + vs.x * v1.y * v2.z - vs.x * v2.y * v1.z -
         vs.y * v2.z * v1.x - vs.z * v2.x * v1.y) * inv;
    if (!((double)t >= EPSD && t <= r.max_t)) return false;
    beta = (vs.x * v2.y * v3.z + v2.x * v3.y * vs.z + v3.x * vs.y * v2.z - v3.x * v2.y * vs.z -
            v3.y * v2.z * vs.x - v3.z * v2.x * vs.y) * inv;
    gamma = (v1.x * vs.y * v3.z + vs.x * v3.y * v1.z + v3.x * v1.y * vs.z - v3.x * vs.y * v1.z -
             v3.y * vs.z * v1.x - v3.z * vs.x * v1.y) * inv;
    float alpha = (float)(1.0 - (double)(gamma + beta));
    if (alpha < 0 || alpha > 1 || beta < 0 || beta > 1 || gamma < 0 || gamma > 1) return false;
    return true;
}

// bvh.h:22-37 + simple_group.h:22-29 + primitive_leaf.h:27-34: unordered DFS,
// every candidate with t <= max_t overwrites the record (last tested wins ties).
void ref_closest(const Scene& s, int node, Ray& r, Hit& h, int64_t* tests) {
    const RNode& n = s.nodes[node];
    if (n.leaf) {
        for (int k = 0; k < n.count; k++) {
            int id = s.leaf_prims[n.first + k];
            float t, b, g;
            if (tests) (*tests)++;
            if (tri_test(s, id, r, t, b, g)) {
                r.max_t = t;
                h.tri = id;
                h.t = t;
                h.beta = b;
                h.gamma = g;
            }
        }
        return;
    }
    if (!aabb_hit(n.bmin, n.bmax, r)) return;
    ref_closest(s, n.right, r, h, tests);
    ref_closest(s, n.left, r, h, tests);
}

// bvh.h:38-42 + simple_group.h:30-39: short-circuit any hit
bool ref_any(const Scene& s, int node, Ray& r) {
    const RNode& n = s.nodes[node];
    if (n.leaf) {
        for (int k = 0; k < n.count; k++) {
            float t, b, g;
            if (tri_test(s, s.leaf_prims[n.first + k], r, t, b, g)) return true;
        }
        return false;
    }
    if (!aabb_hit(n.bmin, n.bmax, r)) return false;
    return ref_any(s, n.right, r) || ref_any(s, n.left, r);
}

// Brute force with the reference's tie rule made explicit: among triangles
// reaching the minimal t, the one the reference tests last (highest rank) wins.
void brute_closest(const Scene& s, Ray& r, Hit& h) {
    int nt = (int)s.F.size();
    float best = r.max_t;
    for (int id = 0; id < nt; id++) {
        float t, b, g;
        Ray rr = r;
        rr.max_t = best;
        if (!tri_test(s, id, rr, t, b, g)) continue;
        if (h.tri < 0 || t < best || (t == best && s.rank[id] > s.rank[h.tri])) {
            best = t;
            h.tri = id;
            h.t = t;
            h.beta = b;
            h.gamma = g;
        }
    }
    r.max_t = best;
}

bool brute_any(const Scene& s, Ray& r) {
    for (int id = 0; id < (int)s.F.size(); id++) {
        float t, b, g;
        if (tri_test(s, id, r, t, b, g)) return true;
    }
    return false;
}

// ---------------------------------------------------------------- shading record
struct Isect {
    V3 normal, g_normal, bitangent, tangent, hitpoint;
    float u = 0.0f, v = 0.0f;  // texture coordinates
    int tri = -1;
    bool computed_normal = false;
};

// triangle.h:79-113 hit record from (t, beta, gamma)
void fill_isect(const Scene& s, const Ray& r, const Hit& h, Isect& it) {
    const auto& f = s.F[h.tri];
    float beta = h.beta, gamma = h.gamma;
    float alpha = (float)(1.0 - (double)(gamma + beta));
    V3 a = s.P[f[0]], b = s.P[f[1]], c = s.P[f[2]];
    it.normal = normalized(s.N[f[0]] * alpha + s.N[f[1]] * beta + s.N[f[2]] * gamma);
    V3 bt = normalized(s.B[f[0]] * alpha + s.B[f[1]] * beta + s.B[f[2]] * gamma);
    it.g_normal = normalized(cross(b - a, c - a));
    it.bitangent = bt;
    it.tangent = normalized(cross(bt, it.normal));
    it.hitpoint = r.o + r.d * h.t;
    it.tri = h.tri;
    if (!s.UV.empty()) {  // triangle.h:90-95
        const auto &p = s.UV[f[0]], &q = s.UV[f[1]], &w = s.UV[f[2]];
        it.u = p[0] * alpha + q[0] * beta + w[0] * gamma;
        it.v = p[1] * alpha + q[1] * beta + w[1] * gamma;
    }
}

// Image<T>::get_pixel with the REPEAT border (rgb_image.h:211-299): an index at or
// past the end wraps with %, a negative one becomes i%n + n, which is n itself
// for negative multiples of n (the reference then reads the first texel of the
// next row; past the image's last texel this restatement takes the last one)
const float* image_texel(const Scene& s, int img, int i, int j) {
    const auto& d = s.images[img];
    const int w = d[0], h = d[1], ch = d[2];
    if (i >= w) i %= w;
    else if (i < 0) i = i % w + w;
    if (j >= h) j %= h;
    else if (j < 0) j = j % h + h;
    long long k = (long long)i + (long long)j * w;
    k = std::min(k, (long long)w * h - 1);
    return s.texels.data() + d[3] + k * ch;
}

// ImageTexture::sample, point filtering (image_texture.h:24-28): x in float,
// y = (1.0 - v) * height in double stored as float, pixel (floor x, floor y);
// a one-channel image broadcasts its value to the three components
V3 image_sample(const Scene& s, int img, float u, float v) {
    const float x = u * (float)s.images[img][0];
    const float y = (float)((1.0 - (double)v) * (double)s.images[img][1]);
    const float* t = image_texel(s, img, (int)std::floor(x), (int)std::floor(y));
    return s.images[img][2] == 3 ? V3(t[0], t[1], t[2]) : V3(t[0]);
}

// StandardMaterial::get_bxdf / get_normal_mapped inputs (standard_material.h:45-71)
// at the hit: albedo and normal map from their images, roughness from the first
// component of its image, and a metal image picks Lambertian (< 0.1) or GGX
Material hit_material(const Scene& s, const Isect& it) {
    const int id = s.tri_mat[it.tri];
    Material m = s.mats[id];
    if (s.mat_tex.empty()) return m;
    const auto& t = s.mat_tex[id];
    if (t[0] >= 0) m.albedo = image_sample(s, t[0], it.u, it.v);
    if (t[1] >= 0) m.rough = image_sample(s, t[1], it.u, it.v).x;
    if (t[2] >= 0) m.type = image_sample(s, t[2], it.u, it.v).x < 0.1 ? 0 : 1;
    if (t[3] >= 0) m.nmap = image_sample(s, t[3], it.u, it.v);
    return m;
}

// intersection.h:25-33 + standard_material.h:64-71 (normal map applied on the
// first hit only, because the Intersection object is reused across bounces)
void first_hit_frame(const Material& m, Isect& it) {
    if (it.computed_normal) return;
    if (m.has_nmap) {
        V3 nn = m.nmap;
        nn.x = (float)((double)nn.x * 2.0 - 1.0);
        nn.y = (float)((double)nn.y * 2.0 - 1.0);
        M3 mm(it.tangent, it.bitangent, it.normal);
        it.normal = normalized(mm.mul(nn));
    }
    it.bitangent = normalized(cross(it.normal, it.tangent));
    it.tangent = normalized(cross(it.bitangent, it.normal));
    it.computed_normal = true;
}

// bxdf.h:42-49
inline bool same_hemisphere(const Isect& i, V3 wo, V3 wi) {
    float sign = dot(wo, i.g_normal);
    return sign * dot(wi, i.normal) > 0.0 && sign * dot(wi, i.g_normal) > 0.0;
}

// math.h:647-653 (cos/sin in double: see file header)
V3 sample_hemisphere_cos(float r1, float r2) {
    float theta = (float)((double)r1 * 2.0 * M_PI);
    float sin_phi = (float)std::sqrt((double)r2);
    float cos_phi = (float)std::sqrt(1.0 - (double)(sin_phi * sin_phi));
    return V3((float)(std::cos((double)theta) * (double)sin_phi),
              (float)(std::sin((double)theta) * (double)sin_phi), cos_phi);
}

// math.h:663-688
void orthogonal(V3 v1, V3* x, V3* y, V3* z) {
    double ax = std::fabs((double)v1.x), ay = std::fabs((double)v1.y), az = std::fabs((double)v1.z);
    V3 aux;
    if (ax < ay) aux = ax < az ? V3(1, 0, 0) : V3(0, 0, 1);
    else aux = ay < az ? V3(0, 1, 0) : V3(0, 0, 1);
    V3 aux2 = cross(v1, aux);
    *x = normalized(v1);
    *y = normalized(cross(aux2, v1));
    *z = normalized(aux2);
}

inline V3 reflect(V3 wo, V3 wh) {  // math.h:735-737
    return normalized(wo * (float)(-1.0) + wh * dot(wh, wo) * (float)2.0);
}

// ---- GGX, microfacet_distribution.h (PBRT-3 restatement, quirks kept) ----
inline float cosTheta(V3 w) { return w.z; }
inline float cos2Theta(V3 w) { return w.z * w.z; }
inline float sinTheta(V3 w) { return (float)std::sqrt(1.0 - (double)cos2Theta(w)); }
inline float cosPhi(V3 w) {
    float st = sinTheta(w);
    if (st == 0.0) return 1.0f;
    return w.x / sinTheta(w);
}
inline float sinPhi(V3 w) {
    float st = sinTheta(w);
    if (st == 0.0) return 1.0f;
    return w.y / sinTheta(w);
}

struct Ggx {
    float ax, ay;
    // microfacet_distribution.h:13-58
    static void sample11(float cosT, float U1, float U2, float* sx, float* sy) {
        if ((double)cosT > .9999) {
            float r = (float)std::sqrt((double)(U1 / (1 - U1)));
            float phi = (float)(6.28318530718 * (double)U2);
            *sx = (float)((double)r * std::cos((double)phi));
            *sy = (float)((double)r * std::sin((double)phi));
            return;
        }
        float sinT = std::sqrt(std::max((float)0, (float)1 - cosT * cosT));
        float tanT = sinT / cosT;
        float a = (float)(1. / (double)tanT);
        float G1 = (float)(2. / (double)(1 + std::sqrt(1.f + 1.f / (a * a))));
        float A = 2 * U1 / G1 - 1;
        float tmp = 1.f / (A * A - 1.f);
        if ((double)tmp > 1e10) tmp = (float)1e10;
        float B = tanT;
        float D = std::sqrt(std::max(float(B * B * tmp * tmp - (A * A - B * B) * tmp), float(0)));
        float s1 = B * tmp - D;
        float s2 = B * tmp + D;
        *sx = (A < 0 || s2 > 1.f / tanT) ? s1 : s2;
        float S;
        if (U2 > 0.5f) { S = 1.f; U2 = 2.f * (U2 - .5f); }
        else { S = -1.f; U2 = 2.f * (.5f - U2); }
        float z = (U2 * (U2 * (U2 * 0.27385f - 0.73369f) + 0.46341f)) /
                  (U2 * (U2 * (U2 * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
        *sy = S * z * std::sqrt(1.f + *sx * *sx);
        if (std::isinf(*sy) || std::isnan(*sy)) *sy = 0.0f;
    }
    // microfacet_distribution.h:96-118
    V3 sample(V3 wi, float U1, float U2) const {
        V3 ws = normalized(V3(ax * wi.x, ay * wi.y, wi.z));
        float sx, sy;
        sample11(cosTheta(ws), U1, U2, &sx, &sy);
        float tmp = cosPhi(ws) * sx - sinPhi(ws) * sy;
        sy = sinPhi(ws) * sx + cosPhi(ws) * sy;
        sx = tmp;
        sx = ax * sx;
        sy = ay * sy;
        return normalized(V3(-sx, -sy, 1.0f));
    }
    // microfacet_distribution.h:121-130
    float d(V3 wh) const {
        float t2 = sinTheta(wh) / cosTheta(wh);
        t2 = t2 * t2;
        if (std::isinf(t2)) return 0.0f;
        float c4 = cos2Theta(wh) * cos2Theta(wh);
        float e = (cosPhi(wh) * cosPhi(wh) / (ax * ax) + sinPhi(wh) * sinPhi(wh) / (ay * ay)) * t2;
        return (float)(1.0 / (M_PI * (double)ax * (double)ay * (double)c4 * (double)(1 + e) * (double)(1 + e)));
    }
    // microfacet_distribution.h:137-145: |cos/sin| (cot) where PBRT has tan
    float lambda(V3 w) const {
        float at = std::fabs(cosTheta(w) / sinTheta(w));
        if (std::isinf(at)) return 0.0f;
        float alpha = std::sqrt(cosPhi(w) * cosPhi(w) * ax * ax + sinPhi(w) * sinPhi(w) * ay * ay);
        float a2t2 = (alpha * at) * (alpha * at);
        return (float)((-1.0 + std::sqrt(1.0 + (double)a2t2)) / 2.0);
    }
    float g(V3 wo, V3 wi) const { return (float)(1.0 / (1.0 + (double)lambda(wo) + (double)lambda(wi))); }
    float g1(V3 w) const { return (float)(1.0 / (1.0 + (double)lambda(w))); }
    // microfacet_distribution.h:152-165
    float pdf(V3 wo, V3 wh) const {
        float ad = d(wh);
        float ag1 = g1(wo);
        float adot = (float)std::fabs((double)dot(wo, wh));
        float aot = (float)std::fabs((double)cosTheta(wo));
        return ad * ag1 * adot / aot;
    }
};

inline float rough2alpha(float r) { return std::max(r * r, 0.01f); }  // glossy_bxdf.h:77-79

// glossy_bxdf.h:29-40 (no albedo, no Fresnel: eval is grey)
float glossy_eval(float rough, const Isect& it, V3 wo, V3 wi) {
    float al = rough2alpha(rough);
    M3 T = M3(it.tangent, it.bitangent, it.normal).transpose();
    V3 wol = normalized(T.mul(wo));
    V3 wil = normalized(T.mul(wi));
    V3 wh = normalized(wol + wil);
    Ggx md{al, al};
    float gd = md.g(wol, wil) * md.d(wh);
    return (float)((double)gd / (4.0 * (double)dot(wo, it.normal)));
}
// glossy_bxdf.h:55-72
float glossy_pdf(float rough, const Isect& it, V3 wo, V3 wi) {
    float al = rough2alpha(rough);
    M3 T = M3(it.tangent, it.bitangent, it.normal).transpose();
    V3 wol = normalized(T.mul(wo));
    V3 wh = normalized(T.mul(wo + wi));
    Ggx md{al, al};
    return (float)((double)md.pdf(wol, wh) / (4.0 * std::fabs((double)dot(wol, wh))));
}

struct DSample {
    V3 wi;
    float pdf;
    bool delta;
};

// ---------------------------------------------------------------- BxDF dispatch
DSample bsdf_sample(const Material& m, const Isect& it, V3 wo, Sampler& smp) {
    if (m.type == 0) {  // diffuse_bxdf.h:34-47
        float r1 = smp.sample();
        float r2 = smp.sample();
        V3 s = sample_hemisphere_cos(r1, r2);
        float p = (float)((double)s.z / M_PI);
        V3 n = it.normal;
        if ((double)dot(n, wo) < 0.0) n = n * (float)(-1.0);  // bxdf.h:52-57
        V3 x, y, z;
        orthogonal(n, &x, &y, &z);
        M3 coords(y, z, x);
        return {coords.mul(s), p, false};
    }
    if (m.type == 1) {  // glossy_bxdf.h:41-53
        if (m.rough < 0.01f) return {reflect(wo, it.normal), 1.0f, true};
        float al = rough2alpha(m.rough);
        M3 C(it.tangent, it.bitangent, it.normal);
        Ggx md{al, al};
        V3 wol = normalized(C.transpose().mul(wo));
        float u1 = smp.sample();
        float u2 = smp.sample();
        V3 whl = normalized(md.sample(wol, u1, u2));
        V3 wh = normalized(C.mul(whl));
        V3 dir = reflect(wo, wh);
        return {dir, glossy_pdf(m.rough, it, wo, dir), false};
    }
    // dielectric, refraction_bxdf.h:26-52
    V3 n = it.normal;
    float n2 = (float)(1.0 / (double)m.ior);
    if ((double)dot(n, wo) < 0.0) {
        n2 = (float)(1.0 / (double)n2);
        n = n * (float)(-1.0);
    }
    float cti = dot(wo, n);
    float ctt2 = (float)(1.0 - (double)(n2 * n2) * (1.0 - (double)(cti * cti)));
    V3 r = normalized(wo * (float)(-1.0) * n2 + n * (float)((double)(n2 * cti) - std::sqrt((double)ctt2)));
    if (ctt2 < 0) r = normalized(wo * (float)(-1.0) + it.normal * dot(it.normal, wo) * (float)2.0);
    return {r, 1.0f, true};
}

V3 bsdf_eval(const Material& m, const Isect& it, V3 wo, V3 wi) {
    if (m.type == 0)  // diffuse_bxdf.h:86-88: col*|n.wi| / M_PI (M_PI -> float, then 1.0/a)
        return divs(m.albedo * (float)std::fabs((double)dot(it.normal, wi)), (float)M_PI);
    if (m.type == 1) return V3(glossy_eval(m.rough, it, wo, wi));
    return V3(1.0f);  // refraction_bxdf.h:19-21
}

float bsdf_pdf(const Material& m, const Isect& it, V3 wo, V3 wi) {
    if (m.type == 0) return (float)(std::fabs((double)dot(it.normal, wi)) / M_PI);  // diffuse_bxdf.h:108-110
    if (m.type == 1) return glossy_pdf(m.rough, it, wo, wi);
    return 1.0f;
}

// bxdf.h:47-49; the dielectric (not loader-reachable in the reference) lets
// transmitted directions through, otherwise refraction would end every path.
inline bool bsdf_non_zero(const Material& m, const Isect& it, V3 wo, V3 wi) {
    if (m.type == 3) return true;
    return same_hemisphere(it, wo, wi);
}

// camera_perspective.h:30-39
Ray camera_ray(const Camera& c, float cx, float cy) {
    V3 d;
    d.z = -1.0f;
    d.y = cy * c.tan_fovy;
    d.x = cx * c.tan_fovy * c.aspect;
    d = normalized(c.coords.mul(d));
    return {c.o, d, std::numeric_limits<float>::infinity()};
}

inline float tri_area(const Scene& s, int id) {  // triangle.h:197-207
    const auto& f = s.F[id];
    V3 a = s.P[f[0]], b = s.P[f[1]], c = s.P[f[2]];
    return (float)((double)length(cross(b - a, c - a)) * 0.5);
}

struct LSample {
    float pdf;
    V3 position, intensity, normal;
    bool delta = false;
};

// shape_light.h:24-44 + triangle.h:209-269: uniform area sample of emitter
// triangle `id` seen from `from`; the group factor 1/NL is applied by the caller.
LSample tri_light_sample(const Scene& s, int id, float r1, float r2, const Isect& from) {
    const auto& f = s.F[id];
    float alpha = r1, beta = r2;
    if ((double)(alpha + beta) > 1.0) {
        alpha = (float)(1.0 - (double)alpha);
        beta = (float)(1.0 - (double)beta);
    }
    float gamma = (float)(1.0 - (double)alpha - (double)beta);
    V3 a = s.P[f[0]], b = s.P[f[1]], c = s.P[f[2]];
    LSample ls;
    ls.position = a * alpha + b * beta + c * gamma;
    ls.normal = normalized(s.N[f[0]] * alpha + s.N[f[1]] * beta + s.N[f[2]] * gamma);
    const Material& m = s.mats[s.tri_mat[id]];
    ls.intensity = m.emission * (float)std::fabs((double)dot(ls.normal, normalized(from.hitpoint - ls.position)));
    ls.pdf = (float)(1.0 / (double)tri_area(s, id));
    return ls;
}

// light_group.h:25-30: uniform pick int(u*N)%N over all lights, pdf *= 1.0/float(N)
// (a double factor applied to the float pdf).  A point light (point_light.h:20-28)
// draws no further numbers: pdf 1, its radiance, and the normal (1,0,0) that NEE's
// cosine term then uses.
LSample light_sample(const Scene& s, Sampler& smp, const Isect& from) {
    size_t NL = s.lights.size();
    int pick = (int)(int(smp.sample() * (float)NL) % NL);
    if (s.lights[pick] < 0) {
        const auto& p = s.point_lights[-1 - s.lights[pick]];
        LSample ls;
        ls.position = V3(p[0], p[1], p[2]);
        ls.intensity = V3(p[3], p[4], p[5]);
        ls.normal = V3(1.0f, 0.0f, 0.0f);
        ls.delta = true;
        ls.pdf = 1.0f;
        ls.pdf = (float)((double)ls.pdf * (1.0 / (double)(float)NL));
        return ls;
    }
    float r1 = smp.sample();
    float r2 = smp.sample();
    LSample ls = tri_light_sample(s, s.lights[pick], r1, r2, from);
    ls.pdf = (float)((double)ls.pdf * (1.0 / (double)(float)NL));
    return ls;
}

// light_group.h:43-55 + shape_light.h:46-48 (pdf of hitting emitter triangle `id`)
inline float light_pdf(const Scene& s, int id) {
    float pdf = (float)(1.0 / (double)(float)s.lights.size());
    pdf *= (float)(1.0 / (double)tri_area(s, id));
    return pdf;
}

// pathtracerMIS.h:24-27
inline float russian_roulette(V3 c) {
    return (float)(std::fabs((double)c.x * 0.2 + (double)c.y * 0.5 + (double)c.z * 0.3) * 0.5 + 0.4);
}
inline float rr_prob(const Scene& s, V3 c) {
    const float q = russian_roulette(c);
    return s.corrected ? std::min(q, 1.0f) : q;
}

struct Counts {
    int64_t ext = 0, shadow = 0;
};

bool intersect(const Scene& s, Ray& r, Hit& h) {
    h = Hit();
    ref_closest(s, s.root, r, h, nullptr);
    return h.tri >= 0;
}

// pathtracerMIS.h:29-130; with s.integrator 1 the NEE-only Pathtracer
// (pathtracer.h:27-110: emitters count only after delta/camera bounces, NEE
// unweighted) and with 2 PurePt (pure_pt.h:27-67: emitters always count, no NEE)
V3 integrate(const Scene& s, float cx, float cy, Sampler& smp, int max_depth, Counts* cnt) {
    Ray ray = camera_ray(s.cam, cx, cy);
    Isect it;
    V3 col(0.0f), mul(1.0f);
    bool sampled_delta = true;
    float p = 1.0f;
    for (int i = 0; i < max_depth; i++) {
        Hit h;
        if (cnt) cnt->ext++;
        if (!intersect(s, ray, h)) {
            col = col + mul * s.env;  // render_sky (0 in the reference; constant environment here)
            break;
        }
        fill_isect(s, ray, h, it);
        const Material m = hit_material(s, it);
        if (s.corrected) it.computed_normal = false;
        first_hit_frame(m, it);
        V3 wo = ray.d * (float)(-1.0);
        if (m.type == 2) {
            if (s.integrator == 2 || sampled_delta) {
                col = col + mul * m.emission;  // pure_pt.h:43-46, pathtracer.h:44-48
            } else if (s.integrator == 1) {
                // pathtracer.h:44-48: nothing after a non-delta bounce (NEE counted it)
            } else {
                float pbb = p;
                float r = ray.max_t;
                float cos_theta = (float)(std::fabs((double)dot(it.normal, ray.d)) + EPSD);
                float pbn = light_pdf(s, h.tri) * r * r / cos_theta;
                float w = pbb / (pbb + pbn);
                col = col + mul * m.emission * w;
            }
            break;
        }
        sampled_delta = false;
        DSample smpl = bsdf_sample(m, it, wo, smp);
        V3 sdir = smpl.wi;
        p = smpl.pdf;
        if (s.integrator == 2) goto no_nee;  // pure_pt.h: no light sampling
        if (!smpl.delta) {
            if (s.lights.empty()) goto no_nee;  // the reference divides by zero here
            {
            LSample ls = light_sample(s, smp, it);
            V3 sd = ls.position - it.hitpoint;
            float len = length(sd);
            sd = divs(sd, len);
            Ray sh{it.hitpoint + sd * EPSF, sd, (float)((double)len - 2.0 * EPSD)};
            if (bsdf_non_zero(m, it, wo, sh.d)) {
                if (cnt) cnt->shadow++;
                if (!ref_any(s, s.root, sh)) {
                    float r = len;
                    V3 ev = bsdf_eval(m, it, wo, sh.d);
                    if (s.integrator == 1) {
                        col = col + divs(mul * ev * ls.intensity, ls.pdf * r * r);  // pathtracer.h:70-76
                    } else {
                        float pnb = bsdf_pdf(m, it, wo, sd);
                        float ct = (float)(std::fabs((double)dot(ls.normal, sd)) + EPSD);
                        float pnn = ls.pdf * r * r / ct;
                        float w = pnn / (pnb + pnn);
                        if (s.corrected && ls.delta) w = 1.0f;
                        col = col + divs(mul * ev * ls.intensity, ls.pdf * r * r) * w;
                    }
                }
            }
            }
        } else {
            sampled_delta = true;
        }
    no_nee:
        if (!bsdf_non_zero(m, it, wo, sdir)) break;
        V3 ev = bsdf_eval(m, it, wo, sdir);
        const float p_sampled = p;
        if (i > 2) {
            float rr = rr_prob(s, ev);
            if (smp.sample() > rr) break;
            p = p * rr;
        }
        mul = divs(mul * ev, p);
        if (s.corrected) p = p_sampled;
        ray = Ray{it.hitpoint + sdir * EPSF, sdir, std::numeric_limits<float>::infinity()};
    }
    return col;
}

// ---------------------------------------------------------------- PPG (SD-tree guided arm)
// Flattened SD-tree (the layout of the library's cpt_ppg_export) and the
// reference's lookups on it: SDTree::get_leaf(Vec3) sdtree.h:345-396, DTree
// get_leaf/get_area 268-343, sample_direction 400-423, pdf 525-538, splat 514-524.
struct SdTreeView {
    const int32_t* snode;
    const int32_t* dinfo;
    const int32_t* quad;
    const float* rect;
    float lo[3], hi[3];
};

int sd_dtree_of(const SdTreeView& t, V3 p) {
    float lo[3] = {t.lo[0], t.lo[1], t.lo[2]}, hi[3] = {t.hi[0], t.hi[1], t.hi[2]};
    const float pc[3] = {p.x, p.y, p.z};
    int n = 0;
    for (int depth = 0;; depth++) {
        const int v = t.snode[n];
        if (v < 0) return -1 - v;
        const int ax = depth % 3;
        const float mid = (float)((double)(lo[ax] + hi[ax]) * 0.5);
        if (pc[ax] > mid) {
            lo[ax] = mid;
            n = v + 1;
        } else {
            hi[ax] = mid;
            n = v;
        }
    }
}

int sd_quad_of(const SdTreeView& t, int dt, float x, float y, float* ax, float* ay) {
    float mnx = (float)-M_PI, mny = -1.0f, mxx = (float)M_PI, mxy = 1.0f;
    int n = t.dinfo[4 * dt];
    while (true) {
        const float midx = (float)((double)(mnx + mxx) * 0.5), midy = (float)((double)(mny + mxy) * 0.5);
        const int v = t.quad[n];
        if (v < 0) {
            *ax = mxx - mnx;
            *ay = mxy - mny;
            return -1 - v;
        }
        int child;
        if (x < midx) {
            if (y < midy) { child = 0; mxx = midx; mxy = midy; }
            else { child = 2; mxx = midx; mny = midy; }
        } else {
            if (y < midy) { child = 1; mnx = midx; mxy = midy; }
            else { child = 3; mnx = midx; mny = midy; }
        }
        n = v + child;
    }
}

float sd_pdf_of(const SdTreeView& t, V3 p, V3 w) {
    const int dt = sd_dtree_of(t, p);
    const float phi = std::atan2(w.y, w.x);
    float ax, ay;
    sd_quad_of(t, dt, phi, w.z, &ax, &ay);
    return (float)((1.0 / (double)t.dinfo[4 * dt + 2]) * (1.0 / (double)(ax * ay)));
}

V3 sd_sample_of(const SdTreeView& t, V3 p, Sampler& smp, float* pdf) {
    const int dt = sd_dtree_of(t, p);
    const int sz = t.dinfo[4 * dt + 2];
    const int i = (int)(smp.sample() * (float)sz);
    const float u = smp.sample();
    const float v = smp.sample();
    const float* r = t.rect + 4 * (size_t)(t.dinfo[4 * dt + 1] + i % sz);
    const float area = r[2] * r[3];
    const float phi = r[0] + u * r[2];
    const float ct = r[1] + v * r[3];
    const float st = (float)std::sqrt(1.0 - (double)(ct * ct));
    *pdf = (float)(1.0 / (double)((float)sz * area));
    return V3(std::cos(phi) * st, std::sin(phi) * st, ct);
}

struct PpgVertex {
    V3 pos, dir, rad;
};

// PPG::integrate (ppg.h:41-143)
V3 ppg_integrate(const Scene& s, const SdTreeView& t, float cx, float cy, Sampler& smp, int max_depth,
                 std::vector<PpgVertex>& vs) {
    Ray ray = camera_ray(s.cam, cx, cy);
    Isect it;
    V3 col(0.0f), mul(1.0f);
    vs.clear();
    for (int i = 0; i < max_depth; i++) {
        Hit h;
        const bool hit = intersect(s, ray, h);
        if (i > 0) vs.push_back({ray.o, ray.d, V3(1.0f)});
        if (!hit) {
            col = col + mul * s.env;
            if (i > 0) vs.back().rad = s.env;
            break;
        }
        fill_isect(s, ray, h, it);
        const Material m = hit_material(s, it);
        if (s.corrected) it.computed_normal = false;
        first_hit_frame(m, it);
        V3 wo = ray.d * (float)(-1.0);
        if (m.type == 2) {
            if (i > 0) vs.back().rad = m.emission;
            col = col + mul * m.emission;
            break;
        }
        DSample smpl;
        float weight = 1.0f;
        if (smp.sample() > 0.5f) {
            smpl = bsdf_sample(m, it, wo, smp);
            weight = smpl.pdf / (smpl.pdf + sd_pdf_of(t, it.hitpoint, smpl.wi));
        } else {
            smpl.wi = sd_sample_of(t, it.hitpoint, smp, &smpl.pdf);
            smpl.delta = false;
            if (!bsdf_non_zero(m, it, wo, smpl.wi)) break;
            weight = smpl.pdf / (smpl.pdf + bsdf_pdf(m, it, wo, smpl.wi));
        }
        V3 sdir = smpl.wi;
        float p = (float)((double)smpl.pdf * 0.5);
        if (!bsdf_non_zero(m, it, wo, sdir)) break;
        V3 ev = bsdf_eval(m, it, wo, sdir);
        if (i > 2) {
            float rr = rr_prob(s, ev);
            if (smp.sample() > rr) break;
            p = p * rr;
        }
        V3 nm = ev * (weight / p);
        if (i > 0) vs.back().rad = nm;
        mul = mul * nm;
        ray = Ray{it.hitpoint + sdir * EPSF, sdir, std::numeric_limits<float>::infinity()};
    }
    return col;
}

// ---------------------------------------------------------------- light tracer
struct LightStart {
    bool ok = false;
    V3 o, d, rad;
    float pdf = 0;
};

// LightGroup::sample (light_group.h:33-39) over ShapeLight::sample
// (shape_light.h:50-68: Triangle::sample, triangle.h:209-269, then
// EmissiveBxDF::emit_sample, emissive_bxdf.h:36-55) or PointLight::sample
// (point_light.h:30-37 with sample_sphere, math.h:655-660, arguments drawn in
// source order).  ok = false for a direction that is not a number (most point
// light draws: acos of an argument below -1) or a scene without lights.
LightStart light_start(const Scene& s, Sampler& smp) {
    LightStart L;
    const size_t NL = s.lights.size();
    if (NL == 0) return L;
    const int pick = (int)(int(smp.sample() * (float)NL) % NL);
    const int id = s.lights[pick];
    if (id < 0) {
        const auto& pl = s.point_lights[-1 - id];
        const float r1 = smp.sample();
        const float r2 = smp.sample();
        const float theta = (float)(2.0 * M_PI * (double)r1);
        const float phi = (float)std::acos(1.0 - 2.0 * M_PI * (double)r2);
        L.d = V3((float)(std::sin((double)phi) * std::cos((double)theta)),
                 (float)(std::sin((double)phi) * std::sin((double)theta)), (float)std::cos((double)phi));
        L.o = V3(pl[0], pl[1], pl[2]);
        L.rad = V3(pl[3], pl[4], pl[5]);
        L.pdf = (float)(1.0 / (4.0 * M_PI * M_PI));
    } else {
        const auto& f = s.F[id];
        const float r1 = smp.sample();
        const float r2 = smp.sample();
        float alpha = r1, beta = r2;
        if ((double)(alpha + beta) > 1.0) {
            alpha = (float)(1.0 - (double)alpha);
            beta = (float)(1.0 - (double)beta);
        }
        const float gamma = (float)(1.0 - (double)alpha - (double)beta);
        const V3 a = s.P[f[0]], b = s.P[f[1]], c = s.P[f[2]];
        L.o = a * alpha + b * beta + c * gamma;
        const V3 n = normalized(s.N[f[0]] * alpha + s.N[f[1]] * beta + s.N[f[2]] * gamma);
        const V3 bt = normalized(s.B[f[0]] * alpha + s.B[f[1]] * beta + s.B[f[2]] * gamma);
        const V3 t = normalized(cross(bt, n));
        const float e1 = smp.sample();
        const float e2 = smp.sample();
        const V3 hs = sample_hemisphere_cos(e1, e2);
        const float p = (float)(0.5 * (double)hs.z / M_PI);
        const float side = smp.sample();
        V3 nn = n, tt = t;
        if ((double)side < 0.5) {
            nn = neg(n);
            tt = neg(t);
        }
        L.d = M3(tt, bt, nn).mul(hs);
        L.rad = s.mats[s.tri_mat[id]].emission * (float)std::fabs((double)dot(n, L.d));
        L.pdf = (float)((1.0 / (double)tri_area(s, id)) * (double)p);
    }
    L.pdf = (float)((double)L.pdf * (1.0 / (double)(float)NL));
    L.ok = !(std::isnan(L.d.x) || std::isnan(L.d.y) || std::isnan(L.d.z));
    return L;
}

// CameraPerspective::connect_light_path (camera_perspective.h:46-82)
struct CamConn {
    int i = -1, j = -1;
    float factor = 0;
};
CamConn camera_connect(const Camera& c, int W, int H, V3 p, bool with_aspect) {
    CamConn cc;
    V3 dir = c.coords.transpose().mul(p - c.o);
    const float ndl = length(dir);
    if ((double)ndl < EPSD || dir.z >= 0) return cc;
    const V3 nd = normalized(divs(dir, ndl));
    dir = divs(dir, dir.z) * (float)(-1.0);
    dir.x = dir.x / (c.tan_fovy * c.aspect);
    dir.y = dir.y / c.tan_fovy;
    const double fx = ((double)dir.x * 0.5 + 0.5) * (double)W;
    const double fy = ((double)dir.y * 0.5 + 0.5) * (double)H;
    if (!(fx >= 0.0 && fx < (double)W && fy >= 0.0 && fy < (double)H)) return cc;  // int() then bounds
    cc.i = (int)fx;
    cc.j = H - (int)fy - 1;
    double den = 4.0 * (double)std::fabs(nd.z * nd.z * nd.z) * (double)ndl * (double)ndl * (double)c.tan_fovy *
                 (double)c.tan_fovy;
    if (with_aspect) den *= (double)c.aspect;  // the corrected mode: image-plane area 4 tan^2 aspect
    cc.factor = (float)(1.0 / den);
    return cc;
}

// Lighttracer::integrate (lighttracer.h:27-110).  Splats `color * scale` into the
// FP64 film (scale = spp, so that film / spp is the reference's image).
void integrate_lt(const Scene& s, Sampler& smp, float factor, int W, int H, int max_depth, double scale, double* film,
                  Counts* cnt) {
    LightStart L = light_start(s, smp);
    if (!L.ok) return;
    Ray ray{L.o, L.d, std::numeric_limits<float>::infinity()};
    V3 mul = divs(L.rad, L.pdf);
    Isect it;
    for (int i = 0; i < max_depth; i++) {
        Hit h;
        if (cnt) cnt->ext++;
        if (!intersect(s, ray, h)) break;
        fill_isect(s, ray, h, it);
        const Material m = hit_material(s, it);
        if (s.corrected) it.computed_normal = false;
        first_hit_frame(m, it);
        if (m.type == 2) break;
        const V3 wo = neg(ray.d);
        DSample ds = bsdf_sample(m, it, wo, smp);
        const V3 eval = bsdf_eval(m, it, wo, ds.wi);
        float p = ds.pdf;
        if (!ds.delta) {
            CamConn cc = camera_connect(s.cam, W, H, it.hitpoint, s.corrected);
            const V3 to = s.cam.o - it.hitpoint;
            Ray sh{it.hitpoint, normalized(to), length(to)};
            if (bsdf_non_zero(m, it, wo, sh.d) && cc.i > -1) {
                if (cnt) cnt->shadow++;
                if (!ref_any(s, s.root, sh)) {
                    V3 color = mul * bsdf_eval(m, it, normalized(wo), normalized(sh.d)) * factor;
                    color = color * cc.factor;
                    double* f = film + 3 * ((size_t)cc.i + (size_t)cc.j * W);
                    f[0] += (double)color.x * scale;
                    f[1] += (double)color.y * scale;
                    f[2] += (double)color.z * scale;
                }
            }
        }
        if (i > 2) {
            float rr = rr_prob(s, eval);
            if (smp.sample() > rr) break;
            p = p * rr;
        }
        mul = divs(mul * eval, p);
        ray = Ray{it.hitpoint, ds.wi, std::numeric_limits<float>::infinity()};
    }
}

// pathtracerMIS.h:150-153 film coordinates of stratum (k,l) of pixel (i,j)
inline void film_coords(int i, int j, int k, int l, int S, int W, int H, float r1, float r2, float* cx,
                        float* cy) {
    float qx = (float(i) + float(k + r1) / S) / float(W);
    float qy = (float(j) + float(l + r2) / S) / float(H);
    *cx = (float)((double)qx * 2.0 - 1.0);
    *cy = (float)(-((double)qy * 2.0 - 1.0));
}

V3 ld(const float* p) { return V3(p[0], p[1], p[2]); }

}  // namespace

extern "C" {

void* orc_scene_create(int nv, const float* pos, const float* nrm, const float* bitan, int nt,
                       const int32_t* idx, const int32_t* tri_mat, int nm, const float* mats, int nl,
                       const int32_t* light_tris, const float* cam, const float* env) {
    Scene* s = new Scene();
    s->P.resize(nv);
    s->N.resize(nv);
    s->B.resize(nv);
    for (int i = 0; i < nv; i++) {
        s->P[i] = ld(pos + 3 * i);
        s->N[i] = ld(nrm + 3 * i);
        s->B[i] = ld(bitan + 3 * i);
    }
    s->F.resize(nt);
    s->tri_mat.resize(nt);
    for (int i = 0; i < nt; i++) {
        s->F[i] = {idx[3 * i], idx[3 * i + 1], idx[3 * i + 2]};
        s->tri_mat[i] = tri_mat[i];
    }
    s->mats.resize(nm);
    for (int i = 0; i < nm; i++) {
        const float* m = mats + 16 * i;
        Material& M = s->mats[i];
        M.type = (int)m[0];
        M.albedo = ld(m + 1);
        M.rough = m[4];
        M.emission = ld(m + 5);
        M.ior = m[8];
        M.nmap = ld(m + 9);
        M.has_nmap = m[12] != 0.0f;
    }
    s->lights.assign(light_tris, light_tris + nl);
    s->light_of_tri.assign(nt, -1);
    for (int i = 0; i < nl; i++)
        if (light_tris[i] >= 0) s->light_of_tri[light_tris[i]] = i;
    s->cam.o = ld(cam);
    s->cam.coords = M3(ld(cam + 3), ld(cam + 6), ld(cam + 9));
    s->cam.tan_fovy = cam[12];
    s->cam.aspect = cam[13];
    s->cam.resX = (int)cam[14];
    s->cam.resY = (int)cam[15];
    s->env = env ? ld(env) : V3(0.0f);
    if (nt > 0) build_bvh(*s);
    return s;
}

void orc_scene_set_textures(void* sp, const float* uvs, int n_images, const int32_t* images, int n_texels,
                            const float* texels, const int32_t* material_textures) {
    Scene* s = (Scene*)sp;
    s->UV.resize(s->P.size());
    for (size_t i = 0; i < s->UV.size(); i++) s->UV[i] = {uvs[2 * i], uvs[2 * i + 1]};
    s->images.resize(n_images);
    for (int i = 0; i < n_images; i++)
        s->images[i] = {images[4 * i], images[4 * i + 1], images[4 * i + 2], images[4 * i + 3]};
    s->texels.assign(texels, texels + n_texels);
    s->mat_tex.resize(s->mats.size());
    for (size_t i = 0; i < s->mat_tex.size(); i++)
        s->mat_tex[i] = {material_textures[4 * i], material_textures[4 * i + 1], material_textures[4 * i + 2],
                         material_textures[4 * i + 3]};
}

void orc_scene_set_point_lights(void* sp, int n, const float* rec) {
    Scene* s = (Scene*)sp;
    s->point_lights.resize(n);
    for (int k = 0; k < n; k++)
        for (int j = 0; j < 6; j++) s->point_lights[k][j] = rec[6 * k + j];
}

void orc_scene_destroy(void* s) { delete (Scene*)s; }

int orc_bvh_nodes(void* s) { return (int)((Scene*)s)->nodes.size(); }

void orc_bvh_rank(void* sp, int32_t* rank_out) {
    Scene* s = (Scene*)sp;
    for (size_t i = 0; i < s->rank.size(); i++) rank_out[i] = s->rank[i];
}

void orc_trace_closest(void* sp, const float* rays, int n, int mode, int32_t* tri, float* t, int nthreads) {
    const Scene* s = (const Scene*)sp;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; i++) {
        const float* r = rays + 7 * i;
        Ray ray{ld(r), ld(r + 3), r[6]};
        Hit h;
        if (s->root >= 0) {
            if (mode == 0) ref_closest(*s, s->root, ray, h, nullptr);
            else brute_closest(*s, ray, h);
        }
        tri[i] = h.tri;
        t[i] = h.tri >= 0 ? h.t : std::numeric_limits<float>::infinity();
    }
}

void orc_hit_material(void* sp, const float* rays, int n, float* out) {
    const Scene* s = (const Scene*)sp;
    for (int i = 0; i < n; i++) {
        const float* r = rays + 7 * i;
        Ray ray{ld(r), ld(r + 3), r[6]};
        Hit h;
        float* o = out + 10 * i;
        std::fill(o, o + 10, 0.0f);
        o[0] = -1.0f;
        if (s->root >= 0) ref_closest(*s, s->root, ray, h, nullptr);
        if (h.tri < 0) continue;
        Isect it;
        fill_isect(*s, ray, h, it);
        const Material m = hit_material(*s, it);
        const float v[10] = {(float)m.type, m.albedo.x, m.albedo.y, m.albedo.z, m.rough,
                             m.nmap.x, m.nmap.y, m.nmap.z, it.u, it.v};
        std::copy(v, v + 10, o);
    }
}

void orc_image_texel(void* sp, int img, int i, int j, float* out) {
    const Scene* s = (const Scene*)sp;
    const float* t = image_texel(*s, img, i, j);
    std::copy(t, t + s->images[img][2], out);
}

void orc_trace_any(void* sp, const float* rays, int n, int mode, uint8_t* occ, int nthreads) {
    const Scene* s = (const Scene*)sp;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; i++) {
        const float* r = rays + 7 * i;
        Ray ray{ld(r), ld(r + 3), r[6]};
        bool o = false;
        if (s->root >= 0) o = mode == 0 ? ref_any(*s, s->root, ray) : brute_any(*s, ray);
        occ[i] = o ? 1 : 0;
    }
}

void orc_camera_ray(void* sp, float cx, float cy, float* o3, float* d3) {
    Ray r = camera_ray(((Scene*)sp)->cam, cx, cy);
    o3[0] = r.o.x; o3[1] = r.o.y; o3[2] = r.o.z;
    d3[0] = r.d.x; d3[1] = r.d.y; d3[2] = r.d.z;
}

int orc_render(void* sp, int W, int H, int S, int max_depth, int sampler, uint64_t seed, int col0, int col1,
               int row0, int row1, double* sum_out, double* sumsq_out, float* film_out, int nthreads,
               int64_t* counts) {
    const Scene* s = (const Scene*)sp;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    int64_t paths = 0, ext = 0, shadow = 0;
    // pathtracerMIS.h:139-160: OpenMP over columns, one sampler per column
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : paths, ext, shadow)
    for (int i = col0; i < col1; i++) {
        // seed 0 is the reference's RandomSampler(i); other seeds shift the column seeds
        MtSampler mt(i + (int)(seed & 0x7fff) * W);
        Counts cnt;
        for (int j = row0; j < row1; j++) {
            V3 acc(0.0f);
            double ax = 0, ay = 0, az = 0, qx = 0, qy = 0, qz = 0;
            for (int k = 0; k < S; k++)
                for (int l = 0; l < S; l++) {
                    V3 c;
                    if (sampler == 1) {
                        float r1 = mt.sample();
                        float r2 = mt.sample();
                        float cx, cy;
                        film_coords(i, j, k, l, S, W, H, r1, r2, &cx, &cy);
                        c = integrate(*s, cx, cy, mt, max_depth, &cnt);
                    } else {
                        PcgSampler pcg(seed, ((uint64_t)s->view << 40) +
                                                 (uint64_t)(i + (int64_t)j * W) * (uint64_t)(S * S) + (uint64_t)(k * S + l));
                        float r1 = pcg.sample();
                        float r2 = pcg.sample();
                        float cx, cy;
                        film_coords(i, j, k, l, S, W, H, r1, r2, &cx, &cy);
                        c = integrate(*s, cx, cy, pcg, max_depth, &cnt);
                    }
                    acc = acc + c;
                    ax += c.x;
                    ay += c.y;
                    az += c.z;
                    qx += (double)c.x * (double)c.x;
                    qy += (double)c.y * (double)c.y;
                    qz += (double)c.z * (double)c.z;
                    paths++;
                }
            size_t pix = (size_t)i + (size_t)j * W;
            if (sum_out) {
                sum_out[3 * pix] = ax;
                sum_out[3 * pix + 1] = ay;
                sum_out[3 * pix + 2] = az;
            }
            if (sumsq_out) {
                sumsq_out[3 * pix] = qx;
                sumsq_out[3 * pix + 1] = qy;
                sumsq_out[3 * pix + 2] = qz;
            }
            if (film_out) {
                V3 f = divs(acc, float(S * S));
                film_out[3 * pix] = f.x;
                film_out[3 * pix + 1] = f.y;
                film_out[3 * pix + 2] = f.z;
            }
        }
        ext += cnt.ext;
        shadow += cnt.shadow;
    }
    if (counts) {
        counts[0] = paths;
        counts[1] = ext;
        counts[2] = shadow;
    }
    return 0;
}

int orc_ppg_render(void* sp, int W, int H, int spp, int max_depth, uint64_t seed, uint32_t view, const int32_t* snode,
                   const int32_t* dinfo, const int32_t* quad, const float* rect, const float* bounds, int n_slots,
                   double* sum_out, double* splat_sum, int64_t* splat_cnt, int nthreads) {
    const Scene* s = (const Scene*)sp;
    SdTreeView t{snode, dinfo, quad, rect, {bounds[0], bounds[1], bounds[2]}, {bounds[3], bounds[4], bounds[5]}};
    if (nthreads > 0) omp_set_num_threads(nthreads);
    for (int k = 0; k < n_slots; k++) {
        splat_sum[k] = 0.0;
        splat_cnt[k] = 0;
    }
#pragma omp parallel
    {
        std::vector<double> ls((size_t)n_slots, 0.0);
        std::vector<int64_t> lc((size_t)n_slots, 0);
        std::vector<PpgVertex> vs;
#pragma omp for schedule(dynamic, 1)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                const uint64_t pixel = (uint64_t)i + (uint64_t)j * W;
                double ax = 0, ay = 0, az = 0;
                for (int k = 0; k < spp; k++) {
                    PcgSampler pcg(seed, ((uint64_t)view << 40) + pixel * (uint64_t)spp + (uint64_t)k);
                    float r1 = pcg.sample();
                    float r2 = pcg.sample();
                    float cx, cy;
                    film_coords(i, j, 0, 0, 1, W, H, r1, r2, &cx, &cy);  // ppg.h:178-181
                    V3 c = ppg_integrate(*s, t, cx, cy, pcg, max_depth, vs);
                    ax += c.x;
                    ay += c.y;
                    az += c.z;
                    // ppg.h:130-141
                    if (!vs.empty() && (vs.back().rad.x > 0 || vs.back().rad.y > 0 || vs.back().rad.z > 0)) {
                        for (int v = (int)vs.size() - 1; v >= 0; v--) {
                            if (v < (int)vs.size() - 1) vs[v].rad = vs[v].rad * vs[v + 1].rad;
                            const V3 r = vs[v].rad;
                            const float f = std::max(r.x, std::max(r.y, r.z));
                            const int dt = sd_dtree_of(t, vs[v].pos);
                            float qx, qy;
                            const int slot = sd_quad_of(t, dt, std::atan2(vs[v].dir.y, vs[v].dir.x), vs[v].dir.z, &qx, &qy);
                            ls[slot] += f;
                            lc[slot] += 1;
                        }
                    }
                }
                sum_out[3 * pixel] = ax;
                sum_out[3 * pixel + 1] = ay;
                sum_out[3 * pixel + 2] = az;
            }
#pragma omp critical
        for (int k = 0; k < n_slots; k++) {
            splat_sum[k] += ls[k];
            splat_cnt[k] += lc[k];
        }
    }
    return 0;
}

void orc_set_integrator(void* sp, int integrator) { ((Scene*)sp)->integrator = integrator; }
void orc_set_view(void* sp, uint32_t view) { ((Scene*)sp)->view = view; }
void orc_set_corrected(void* sp, int corrected) { ((Scene*)sp)->corrected = corrected != 0; }

// Lighttracer::render (lighttracer.h:118-154): W*H*S*S light paths, one per camera
// sample, each with factor 1/(S*S).  sampler 0: per-path PCG stream (path id as
// for camera paths); 1: the reference's per-column mt19937 (RandomSampler(i)).
int orc_render_lt(void* sp, int W, int H, int S, int max_depth, int sampler, uint64_t seed, double* sum_out,
                  int nthreads, int64_t* counts) {
    const Scene* s = (const Scene*)sp;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const float factor = (float)(1.0 / (double)(S * S));
    const double scale = (double)(S * S);
    std::fill(sum_out, sum_out + 3 * (size_t)W * H, 0.0);
    int64_t paths = 0, ext = 0, shadow = 0;
#pragma omp parallel reduction(+ : paths, ext, shadow)
    {
        std::vector<double> film(3 * (size_t)W * H, 0.0);
        Counts cnt;
#pragma omp for schedule(dynamic, 1)
        for (int i = 0; i < W; i++) {
            MtSampler mt(i + (int)(seed & 0x7fff) * W);
            for (int j = 0; j < H; j++)
                for (int k = 0; k < S; k++)
                    for (int l = 0; l < S; l++) {
                        if (sampler == 1) {
                            integrate_lt(*s, mt, factor, W, H, max_depth, scale, film.data(), &cnt);
                        } else {
                            PcgSampler pcg(seed, (uint64_t)(i + (int64_t)j * W) * (uint64_t)(S * S) + (uint64_t)(k * S + l));
                            integrate_lt(*s, pcg, factor, W, H, max_depth, scale, film.data(), &cnt);
                        }
                        paths++;
                    }
        }
#pragma omp critical
        for (size_t q = 0; q < film.size(); q++) sum_out[q] += film[q];
        ext += cnt.ext;
        shadow += cnt.shadow;
    }
    if (counts) {
        counts[0] = paths;
        counts[1] = ext;
        counts[2] = shadow;
    }
    return 0;
}

void orc_path(void* sp, int W, int H, int S, int max_depth, uint64_t seed, int pixel, int sample, float* rgb) {
    const Scene* s = (const Scene*)sp;
    int i = pixel % W, j = pixel / W, k = sample / S, l = sample % S;
    PcgSampler pcg(seed, (uint64_t)pixel * (uint64_t)(S * S) + (uint64_t)sample);
    float r1 = pcg.sample();
    float r2 = pcg.sample();
    float cx, cy;
    film_coords(i, j, k, l, S, W, H, r1, r2, &cx, &cy);
    V3 c = integrate(*s, cx, cy, pcg, max_depth, nullptr);
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}

struct FixedSampler : Sampler {
    const float* u;
    int i = 0;
    explicit FixedSampler(const float* uu) : u(uu) {}
    float sample() override { return u[i++]; }
};

void orc_diffuse_sample(const float* n, const float* wo, float u1, float u2, float* wi, float* pdf) {
    Isect it;
    it.normal = ld(n);
    Material m{};
    m.type = 0;
    float us[2] = {u1, u2};
    FixedSampler fs(us);
    DSample d = bsdf_sample(m, it, ld(wo), fs);
    wi[0] = d.wi.x; wi[1] = d.wi.y; wi[2] = d.wi.z;
    *pdf = d.pdf;
}

static Isect frame_isect(const float* t, const float* b, const float* n) {
    Isect it;
    it.tangent = ld(t);
    it.bitangent = ld(b);
    it.normal = ld(n);
    it.g_normal = ld(n);
    return it;
}

void orc_glossy_sample(const float* t, const float* b, const float* n, float rough, const float* wo, float u1,
                       float u2, float* wi, float* pdf, int* delta) {
    Isect it = frame_isect(t, b, n);
    Material m{};
    m.type = 1;
    m.rough = rough;
    float us[2] = {u1, u2};
    FixedSampler fs(us);
    DSample d = bsdf_sample(m, it, ld(wo), fs);
    wi[0] = d.wi.x; wi[1] = d.wi.y; wi[2] = d.wi.z;
    *pdf = d.pdf;
    *delta = d.delta ? 1 : 0;
}

float orc_glossy_eval(const float* t, const float* b, const float* n, float rough, const float* wo, const float* wi) {
    return glossy_eval(rough, frame_isect(t, b, n), ld(wo), ld(wi));
}

float orc_glossy_pdf(const float* t, const float* b, const float* n, float rough, const float* wo, const float* wi) {
    return glossy_pdf(rough, frame_isect(t, b, n), ld(wo), ld(wi));
}

float orc_ggx_d(float alpha, const float* wh) { return Ggx{alpha, alpha}.d(ld(wh)); }
float orc_ggx_pdf(float alpha, const float* wo, const float* wh) { return Ggx{alpha, alpha}.pdf(ld(wo), ld(wh)); }
void orc_ggx_sample(float alpha, const float* wo, float u1, float u2, float* wh) {
    V3 h = Ggx{alpha, alpha}.sample(ld(wo), u1, u2);
    wh[0] = h.x; wh[1] = h.y; wh[2] = h.z;
}

void orc_triangle_sample(void* sp, int tri, float u1, float u2, float* pos, float* nrm, float* area) {
    const Scene* s = (const Scene*)sp;
    Isect from;
    from.hitpoint = V3(0.0f);
    LSample ls = tri_light_sample(*s, tri, u1, u2, from);
    pos[0] = ls.position.x; pos[1] = ls.position.y; pos[2] = ls.position.z;
    nrm[0] = ls.normal.x; nrm[1] = ls.normal.y; nrm[2] = ls.normal.z;
    *area = tri_area(*s, tri);
}

void orc_pcg_floats(uint64_t seed, uint64_t path_id, int n, float* out) {
    PcgSampler p(seed, path_id);
    for (int i = 0; i < n; i++) out[i] = p.sample();
}

void orc_mt_floats(uint32_t seed, int n, float* out) {
    MtSampler m((int)seed);
    for (int i = 0; i < n; i++) out[i] = m.sample();
}

// microfacet_distribution.h:137-150 (the Lambda with |cos/sin|) in the local frame
float orc_ggx_lambda(float alpha, const float* w) { return Ggx{alpha, alpha}.lambda(ld(w)); }
float orc_ggx_g1(float alpha, const float* w) { return Ggx{alpha, alpha}.g1(ld(w)); }

// refraction_bxdf.h:26-52 through the oracle's BSDF dispatch: shading normal n, ior
void orc_dielectric_sample(const float* n, const float* wo, float ior, float* wi) {
    Isect it;
    it.normal = ld(n);
    it.g_normal = ld(n);
    Material m{};
    m.type = 3;
    m.ior = ior;
    FixedSampler fs(nullptr);
    DSample d = bsdf_sample(m, it, ld(wo), fs);
    wi[0] = d.wi.x; wi[1] = d.wi.y; wi[2] = d.wi.z;
}

// SD-tree lookups on the flattened layout (sdtree.h:268-396, 400-423, 525-538):
// per point the directional tree it falls in, per (point, direction) the pdf,
// and per (point, 3 uniforms) a sampled direction with its pdf
static SdTreeView sd_view(const int32_t* snode, const int32_t* dinfo, const int32_t* quad, const float* rect,
                          const float* bounds) {
    SdTreeView t{snode, dinfo, quad, rect, {bounds[0], bounds[1], bounds[2]}, {bounds[3], bounds[4], bounds[5]}};
    return t;
}
void orc_sd_lookup(const int32_t* snode, const int32_t* dinfo, const int32_t* quad, const float* rect,
                   const float* bounds, int n, const float* pts, int32_t* dtree_out) {
    SdTreeView t = sd_view(snode, dinfo, quad, rect, bounds);
    for (int i = 0; i < n; i++) dtree_out[i] = sd_dtree_of(t, ld(pts + 3 * (size_t)i));
}
void orc_sd_pdf(const int32_t* snode, const int32_t* dinfo, const int32_t* quad, const float* rect,
                const float* bounds, int n, const float* pts, const float* dirs, float* pdf_out) {
    SdTreeView t = sd_view(snode, dinfo, quad, rect, bounds);
    for (int i = 0; i < n; i++) pdf_out[i] = sd_pdf_of(t, ld(pts + 3 * (size_t)i), ld(dirs + 3 * (size_t)i));
}
void orc_sd_sample(const int32_t* snode, const int32_t* dinfo, const int32_t* quad, const float* rect,
                   const float* bounds, int n, const float* pts, const float* u3, float* dirs_out, float* pdf_out) {
    SdTreeView t = sd_view(snode, dinfo, quad, rect, bounds);
    for (int i = 0; i < n; i++) {
        FixedSampler fs(u3 + 3 * (size_t)i);
        V3 d = sd_sample_of(t, ld(pts + 3 * (size_t)i), fs, pdf_out + i);
        dirs_out[3 * (size_t)i] = d.x;
        dirs_out[3 * (size_t)i + 1] = d.y;
        dirs_out[3 * (size_t)i + 2] = d.z;
    }
}

// microfacet_distribution.h:96-118 for n (u1, u2) pairs
void orc_ggx_sample_batch(float alpha, const float* wo, int n, const float* u2, float* wh) {
    Ggx g{alpha, alpha};
    V3 w = ld(wo);
    for (int i = 0; i < n; i++) {
        V3 h = g.sample(w, u2[2 * (size_t)i], u2[2 * (size_t)i + 1]);
        wh[3 * (size_t)i] = h.x;
        wh[3 * (size_t)i + 1] = h.y;
        wh[3 * (size_t)i + 2] = h.z;
    }
}

// pathtracerMIS.h:24-27 (corrected: clamped at 1)
float orc_russian_roulette(const float* c, int corrected) {
    Scene s;
    s.corrected = corrected != 0;
    return rr_prob(s, ld(c));
}

}  // extern "C"

// ============================================================================
// Guide MLP oracle (FP64): forward, loss gradient, backprop, Adam.  With
// bf16 != 0 the inputs of every product are rounded to BF16 where the GPU
// rounds them (features, weights, hidden activations, output and hidden
// gradients); otherwise everything is exact FP64 (used against finite
// differences).  Layout and loss as documented in include/cpt.h
// (cpt_guide_train_step): parameters W_l (K_l x N_l row-major) then b_l.
namespace {
namespace mlp {
// two network shapes: hidden width 64 with 8 lobes (48 outputs) and width 128
// with 16 lobes (96 outputs); the arrays below are sized for the larger one
const int NIN = 7, NB = 32, K0 = NIN * NB, HMAX = 128, NOUTMAX = 96, NLMAX = 16;
struct Shape {
    int H, NL, NOUT, NP;
    int KS[5], NS[5], WO[5], BO[5];
};
Shape make_shape(int width) {
    Shape S;
    S.H = width == 128 ? 128 : 64;
    S.NL = S.H == 128 ? 16 : 8;
    S.NOUT = (5 * S.NL + 15) / 16 * 16;
    int off = 0;
    for (int l = 0; l < 5; l++) {
        S.KS[l] = l == 0 ? K0 : S.H;
        S.NS[l] = l == 4 ? S.NOUT : S.H;
        S.WO[l] = off;
        off += S.KS[l] * S.NS[l];
        S.BO[l] = off;
        off += S.NS[l];
    }
    S.NP = off;
    return S;
}
Shape g_shape = make_shape(64);

inline double bf(double x, int on) {
    if (!on) return x;
    float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return f;
}

struct Acts {
    double a[5][K0];   // inputs of layer l (a[0] = features)
    double o[NOUTMAX];
};

void forward(const float* P, const float* in, int bfon, Acts& A) {
    const Shape& S = g_shape;
    const int *KS = S.KS, *NS = S.NS, *WO = S.WO, *BO = S.BO;
    for (int d = 0; d < NIN; d++)
        for (int b = 0; b < NB; b++) {
            double x = (double)in[d] * NB - (b + 0.5);
            A.a[0][d * NB + b] = bf(std::exp(-0.5 * x * x), bfon);
        }
    for (int l = 0; l < 5; l++) {
        for (int n = 0; n < NS[l]; n++) {
            double z = 0;
            for (int k = 0; k < KS[l]; k++) z += A.a[l][k] * bf(P[WO[l] + k * NS[l] + n], bfon);
            z += P[BO[l] + n];
            if (l < 4) A.a[l + 1][n] = bf(z > 0 ? z : 0.0, bfon);
            else A.o[n] = z;
        }
    }
}

struct Mix {
    double pi[NLMAX], kap[NLMAX], m[NLMAX][3], un[NLMAX], ls[NLMAX];
};

void decode(const double* o, Mix& M) {
    const int NL = g_shape.NL;
    double mx = -1e300;
    for (int k = 0; k < NL; k++) mx = std::max(mx, o[5 * k]);
    double s = 0;
    for (int k = 0; k < NL; k++) s += (M.pi[k] = std::exp(o[5 * k] - mx));
    for (int k = 0; k < NL; k++) {
        M.pi[k] /= s;
        M.ls[k] = o[5 * k + 1];
        M.kap[k] = std::exp(std::min(std::max(o[5 * k + 1], -3.0), 6.0));
        double x = o[5 * k + 2], y = o[5 * k + 3], z = o[5 * k + 4];
        double n = std::sqrt(x * x + y * y + z * z);
        M.un[k] = n;
        if (n < 1e-12) { x = 0; y = 0; z = 1; n = 1; }
        M.m[k][0] = x / n;
        M.m[k][1] = y / n;
        M.m[k][2] = z / n;
    }
}

inline double vmf(double kap, double c) {
    return kap / (2 * M_PI * (1 - std::exp(-2 * kap))) * std::exp(kap * (c - 1));
}

double mix_pdf(const Mix& M, const double* w) {
    const int NL = g_shape.NL;
    double p = 0;
    for (int k = 0; k < NL; k++) p += M.pi[k] * vmf(M.kap[k], M.m[k][0] * w[0] + M.m[k][1] * w[1] + M.m[k][2] * w[2]);
    return p;
}

// -wgt * log q(w) and its gradient w.r.t. the raw outputs (analytic)
double loss_grad(const double* o, const double* w, double wgt, double* g) {
    const int NL = g_shape.NL, NOUT = g_shape.NOUT;
    Mix M;
    decode(o, M);
    double v[NLMAX], q = 0;
    for (int k = 0; k < NL; k++) {
        v[k] = vmf(M.kap[k], M.m[k][0] * w[0] + M.m[k][1] * w[1] + M.m[k][2] * w[2]);
        q += M.pi[k] * v[k];
    }
    q = std::max(q, 1e-300);
    for (int k = 0; k < NL; k++) {
        double r = M.pi[k] * v[k] / q;
        g[5 * k] = -wgt * (r - M.pi[k]);
        double kap = M.kap[k], e2 = std::exp(-2 * kap);
        double c = M.m[k][0] * w[0] + M.m[k][1] * w[1] + M.m[k][2] * w[2];
        double dk = 1 / kap - 2 * e2 / (1 - e2) + (c - 1);
        g[5 * k + 1] = (M.ls[k] > -3 && M.ls[k] < 6) ? -wgt * r * dk * kap : 0.0;
        double f = M.un[k] < 1e-12 ? 0.0 : -wgt * r * kap / M.un[k];
        for (int j = 0; j < 3; j++) g[5 * k + 2 + j] = f * (w[j] - M.m[k][j] * c);
    }
    for (int k = 5 * NL; k < NOUT; k++) g[k] = 0;
    return -wgt * std::log(q);
}

}  // namespace mlp
}  // namespace

extern "C" {

int orc_mlp_num_params(void) { return mlp::g_shape.NP; }

// network shape of the orc_mlp_* functions: width 64 (8 lobes) or 128 (16 lobes)
void orc_mlp_set_width(int width) { mlp::g_shape = mlp::make_shape(width); }

void orc_mlp_forward(const float* P, const float* in, int n, int bfon, double* out) {
    const int NOUT = mlp::g_shape.NOUT;
#pragma omp parallel for
    for (int s = 0; s < n; s++) {
        mlp::Acts A;
        mlp::forward(P, in + (size_t)s * mlp::NIN, bfon, A);
        for (int k = 0; k < NOUT; k++) out[(size_t)s * NOUT + k] = A.o[k];
    }
}

double orc_mlp_mixture_pdf(const double* o, const double* w) {
    mlp::Mix M;
    mlp::decode(o, M);
    return mlp::mix_pdf(M, w);
}

void orc_mlp_mixture_pdf_batch(const double* o, const double* w, int n, double* out) {
    mlp::Mix M;
    mlp::decode(o, M);
    for (int i = 0; i < n; i++) out[i] = mlp::mix_pdf(M, w + 3 * (size_t)i);
}

// loss = sum_s w_s (-log q_s) / sum_s w_s and its gradient (NP doubles)
double orc_mlp_loss_grad(const float* P, const float* samples, int n, int bfon, double* grad) {
    using namespace mlp;
    const Shape& S = g_shape;
    const int *KS = S.KS, *NS = S.NS, *WO = S.WO, *BO = S.BO;
    const int NP = S.NP, NOUT = S.NOUT;
    double wsum = 0;
    for (int s = 0; s < n; s++) wsum += samples[(size_t)s * 11 + 10];
    // the GPU forms 1/sum in FP32 and multiplies per sample in FP32
    double ws = bfon ? (double)(1.0f / (float)wsum) : 1.0 / wsum;
    for (int p = 0; p < NP; p++) grad[p] = 0;
    double loss = 0;
    for (int s = 0; s < n; s++) {
        const float* sp = samples + (size_t)s * 11;
        Acts A;
        forward(P, sp, bfon, A);
        double w[3] = {sp[7], sp[8], sp[9]};
        double wgt = bfon ? (double)(sp[10] * (float)ws) : sp[10] * ws;
        double dz[5][K0];
        double g[NOUTMAX];
        loss += loss_grad(A.o, w, wgt, g);
        for (int k = 0; k < NOUT; k++) dz[4][k] = bf(g[k], bfon);
        for (int l = 4; l >= 1; l--) {
            // dz_{l-1} = (dz_l W_l^T) .* (a_l > 0)
            for (int k = 0; k < KS[l]; k++) {
                double acc = 0;
                for (int nn = 0; nn < NS[l]; nn++) acc += dz[l][nn] * bf(P[WO[l] + k * NS[l] + nn], bfon);
                dz[l - 1][k] = bf(A.a[l][k] > 0 ? acc : 0.0, bfon);
            }
        }
        for (int l = 0; l < 5; l++) {
            for (int k = 0; k < KS[l]; k++) {
                double a = A.a[l][k];
                if (a == 0) continue;
                for (int nn = 0; nn < NS[l]; nn++) grad[WO[l] + k * NS[l] + nn] += a * dz[l][nn];
            }
            for (int nn = 0; nn < NS[l]; nn++) grad[BO[l] + nn] += dz[l][nn];
        }
    }
    return loss;
}

// one Adam step in FP64 on FP32 parameters (in place)
void orc_mlp_adam(float* P, double* m, double* v, const double* g, int step, double lr, double b1, double b2,
                  double eps) {
    double c1 = 1 / (1 - std::pow(b1, step)), c2 = 1 / (1 - std::pow(b2, step));
    for (int p = 0; p < mlp::g_shape.NP; p++) {
        m[p] = b1 * m[p] + (1 - b1) * g[p];
        v[p] = b2 * v[p] + (1 - b2) * g[p] * g[p];
        P[p] = (float)(P[p] - lr * (m[p] * c1) / (std::sqrt(v[p] * c2) + eps));
    }
}

}  // extern "C"
