// 64-byte BVH4 nodes for traversal: each node's four child boxes as 8-bit planes
// relative to the low corner of their union, one power-of-two scale per axis.
// Low planes are rounded down and high planes up, so every quantized box contains
// its exact box and culling stays conservative.  A node step loads four 16-byte
// vectors instead of the seven of the float node (Node4, which the builds write and
// k_quantize_nodes converts).
#pragma once
#include <cstdint>

#include "bvh_build.h"

namespace cpt {

struct alignas(16) QNode {
    float ox, oy, oz;  // low corner of the children's union
    uint32_t exps;     // byte a (x, y, z): 127 + log2 of axis a's scale
    int32_t child[4];  // as Node4::child
    uint32_t q[6];     // axis a: q[2a] the low planes of children 0..3 (byte c), q[2a+1] the high planes
    uint32_t pad[2];
};
static_assert(sizeof(QNode) == 64, "QNode must be 64 bytes");

// Scale exponents stay within [QNODE_EMIN, QNODE_EMAX]: a scale times a ray's
// reciprocal direction (1 <= |1/d| <= 1e20) is then an exact normal float.
constexpr int QNODE_EMIN = -100, QNODE_EMAX = 60;
// extents beyond 255 x 2^QNODE_EMAX are refused at upload
constexpr double QNODE_MAX_EXTENT = 255.0 * 1152921504606846976.0;  // 255 x 2^60

// x = (p - o) / s for a power of two s; true when the double difference is exact
// (Knuth's TwoSum error of p + (-o) is 0, so x is exact too)
__host__ __device__ inline bool scaled_offset(float p, float o, double s, double& x) {
    const double a = (double)p, b = -(double)o;
    const double sum = a + b, bv = sum - a;
    const double err = (a - (sum - bv)) + (b - bv);
    x = sum / s;
    return err == 0.0;
}

__host__ __device__ inline void quantize_node(const Node4& n, QNode& qn) {
    const float* lo[3] = {n.lox, n.loy, n.loz};
    const float* hi[3] = {n.hix, n.hiy, n.hiz};
    float org[3];
    uint32_t exps = 0;
    for (int a = 0; a < 3; a++) {
        float o = __builtin_huge_valf(), h = -__builtin_huge_valf();
        for (int c = 0; c < 4; c++)
            if (n.child[c] != EMPTY_CHILD) {
                o = fminf(o, lo[a][c]);
                h = fmaxf(h, hi[a][c]);
            }
        if (!(o <= h)) o = h = 0.0f;  // no child
        // the smallest k with 255 x 2^k above the extent (with a relative margin)
        const double E = (double)h - (double)o;
        int k = QNODE_EMIN;
        if (E > 0.0) {
            int ex;
            frexp(E * (1.0 + 1e-6) / 255.0, &ex);  // E' / 255 = m 2^ex, m in [0.5, 1): 2^ex > E' / 255
            k = ex < QNODE_EMIN ? QNODE_EMIN : ex > QNODE_EMAX ? QNODE_EMAX : ex;
        }
        const double s = ldexp(1.0, k);
        org[a] = o;
        exps |= (uint32_t)(k + 127) << (8 * a);
        uint32_t ql = 0, qh = 0;
        for (int c = 0; c < 4; c++) {
            uint32_t bl = 255, bh = 0;  // an absent child: an empty box
            if (n.child[c] != EMPTY_CHILD) {
                // (plane - o) / s in units of the scale; where the double difference
                // is not exact (its TwoSum error is not 0), moved outward by far more
                // than that error before floor/ceil
                double xl, xh;
                const bool el = scaled_offset(lo[a][c], o, s, xl), eh = scaled_offset(hi[a][c], o, s, xh);
                double fl = floor(el ? xl : xl - 1e-7 * (1.0 + xl)), fh = ceil(eh ? xh : xh + 1e-7 * (1.0 + xh));
                fl = fl < 0.0 ? 0.0 : fl;  // lo >= o: plane 0 is o itself
                fh = fh > 255.0 ? 255.0 : fh < 0.0 ? 0.0 : fh;
                bl = (uint32_t)fl;
                bh = (uint32_t)fh;
            }
            ql |= bl << (8 * c);
            qh |= bh << (8 * c);
        }
        qn.q[2 * a] = ql;
        qn.q[2 * a + 1] = qh;
    }
    qn.ox = org[0];
    qn.oy = org[1];
    qn.oz = org[2];
    qn.exps = exps;
    for (int c = 0; c < 4; c++) qn.child[c] = n.child[c];
    qn.pad[0] = qn.pad[1] = 0;
}

}  // namespace cpt
