// Host BVH construction: binned SAH -> 4-wide BVH, and the reference tie rank.
#include "bvh_build.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace cpt {
namespace {

struct Box {
    float lo[3], hi[3];
    Box() {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::numeric_limits<float>::infinity();
            hi[k] = -std::numeric_limits<float>::infinity();
        }
    }
    void grow(const Box& b) {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void grow(const float* p) {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    float half_area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0) return 0.0f;
        return dx * dy + dy * dz + dz * dx;
    }
};

struct BNode2 {  // temporary binary node
    Box box;
    int left = -1, right = -1;  // children (internal)
    int first = 0, count = 0;   // leaf range into the index array
    bool leaf() const { return left < 0; }
};

struct Builder2 {
    const std::vector<Box>& tb;
    const std::vector<float>& cen;  // nt x 3
    std::vector<int>& idx;
    std::vector<BNode2> nodes;
    static const int NB = 32;

    Builder2(const std::vector<Box>& b, const std::vector<float>& c, std::vector<int>& i) : tb(b), cen(c), idx(i) {}

    int make_leaf(int first, int count, const Box& box) {
        BNode2 n;
        n.box = box;
        n.first = first;
        n.count = count;
        nodes.push_back(n);
        return (int)nodes.size() - 1;
    }

    int build(int first, int count) {
        Box box, cb;
        for (int i = first; i < first + count; i++) {
            box.grow(tb[idx[i]]);
            cb.grow(&cen[3 * idx[i]]);
        }
        if (count <= 2) return make_leaf(first, count, box);
        int best_axis = -1, best_split = -1;
        float best_cost = std::numeric_limits<float>::infinity();
        for (int axis = 0; axis < 3; axis++) {
            float ext = cb.hi[axis] - cb.lo[axis];
            if (!(ext > 0.0f)) continue;
            Box bins[NB];
            int cnt[NB] = {0};
            float scale = NB / ext;
            for (int i = first; i < first + count; i++) {
                int b = (int)((cen[3 * idx[i] + axis] - cb.lo[axis]) * scale);
                b = std::min(std::max(b, 0), NB - 1);
                cnt[b]++;
                bins[b].grow(tb[idx[i]]);
            }
            float right_area[NB];
            int right_cnt[NB];
            Box acc;
            int ac = 0;
            for (int b = NB - 1; b > 0; b--) {
                acc.grow(bins[b]);
                ac += cnt[b];
                right_area[b] = acc.half_area();
                right_cnt[b] = ac;
            }
            acc = Box();
            ac = 0;
            for (int b = 0; b < NB - 1; b++) {
                acc.grow(bins[b]);
                ac += cnt[b];
                if (ac == 0 || right_cnt[b + 1] == 0) continue;
                float cost = acc.half_area() * ac + right_area[b + 1] * right_cnt[b + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_split = b;
                }
            }
        }
        float leaf_cost = box.half_area() * count;
        // traversal cost ~ 1 triangle test per unit of area; force a split above LEAF_MAX
        if (count <= LEAF_MAX && (best_axis < 0 || best_cost + box.half_area() >= leaf_cost))
            return make_leaf(first, count, box);
        int mid;
        if (best_axis < 0) {
            // all centroids coincide: split the range in the middle
            mid = first + count / 2;
        } else {
            float ext = cb.hi[best_axis] - cb.lo[best_axis];
            float scale = NB / ext;
            auto it = std::partition(idx.begin() + first, idx.begin() + first + count, [&](int t) {
                int b = (int)((cen[3 * t + best_axis] - cb.lo[best_axis]) * scale);
                b = std::min(std::max(b, 0), NB - 1);
                return b <= best_split;
            });
            mid = (int)(it - idx.begin());
            if (mid == first || mid == first + count) mid = first + count / 2;
        }
        int id = (int)nodes.size();
        nodes.push_back(BNode2());
        nodes[id].box = box;
        int l = build(first, mid - first);
        int r = build(mid, first + count - mid);
        nodes[id].left = l;
        nodes[id].right = r;
        return id;
    }
};

}  // namespace

void build_bvh4(const float* P, const int32_t* I, int nt, const std::vector<int32_t>& rank, Bvh4& out) {
    out.nodes.clear();
    out.tris.clear();
    out.leaves = 0;
    out.max_depth = 0;
    std::vector<Box> tb(nt);
    std::vector<float> cen(3 * (size_t)nt);
    for (int t = 0; t < nt; t++) {
        for (int v = 0; v < 3; v++) tb[t].grow(P + 3 * (size_t)I[3 * (size_t)t + v]);
        for (int k = 0; k < 3; k++) cen[3 * (size_t)t + k] = 0.5f * (tb[t].lo[k] + tb[t].hi[k]);
    }
    float scene_max = 0.0f;
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++) scene_max = std::max(scene_max, std::max(std::fabs(tb[t].lo[k]), std::fabs(tb[t].hi[k])));
    std::vector<int> idx(nt);
    for (int t = 0; t < nt; t++) idx[t] = t;
    Builder2 b2(tb, cen, idx);
    b2.nodes.reserve(2 * (size_t)nt / 2 + 4);
    int root2 = nt > 0 ? b2.build(0, nt) : -1;

    // Collapse to 4-wide: repeatedly open the largest-area internal child.
    struct Item {
        int n2;    // binary node
        int n4;    // destination 4-wide node
        int depth;
    };
    std::vector<Item> work;
    out.nodes.push_back(Node4());
    auto init_node = [](Node4& n) {
        for (int c = 0; c < 4; c++) {
            n.lox[c] = n.loy[c] = n.loz[c] = 1e30f;
            n.hix[c] = n.hiy[c] = n.hiz[c] = 1e30f;  // degenerate far point; child is EMPTY
            n.child[c] = EMPTY_CHILD;
            n.pad[c] = 0;
        }
    };
    init_node(out.nodes[0]);
    if (root2 < 0) return;
    work.push_back({root2, 0, 1});
    out.tris.reserve(nt);
    while (!work.empty()) {
        Item it = work.back();
        work.pop_back();
        out.max_depth = std::max(out.max_depth, it.depth);
        int kids[4];
        int nk = 0;
        const BNode2& src = b2.nodes[it.n2];
        if (src.leaf()) {
            kids[nk++] = it.n2;  // root that is itself a leaf
        } else {
            kids[nk++] = src.left;
            kids[nk++] = src.right;
            while (nk < 4) {
                int bi = -1;
                float ba = -1.0f;
                for (int k = 0; k < nk; k++) {
                    const BNode2& c = b2.nodes[kids[k]];
                    if (!c.leaf() && c.box.half_area() > ba) {
                        ba = c.box.half_area();
                        bi = k;
                    }
                }
                if (bi < 0) break;
                int open = kids[bi];
                kids[bi] = b2.nodes[open].left;
                kids[nk++] = b2.nodes[open].right;
            }
        }
        Node4 node;
        init_node(node);
        int internal_slots[4];
        int ni = 0;
        for (int k = 0; k < nk; k++) {
            const BNode2& c = b2.nodes[kids[k]];
            // The reference's triangle test accepts hit points a few ulps outside
            // the triangle (rounded barycentrics, triangle.h:59-75), i.e. outside
            // its bounding box.  Child boxes are therefore widened by a few ulps
            // of their own coordinates plus a scene-scale term, so that
            // traversal never culls a leaf whose triangle the reference accepts.
            float lo[3], hi[3];
            for (int a = 0; a < 3; a++) {
                float m = std::max(std::fabs(c.box.lo[a]), std::fabs(c.box.hi[a]));
                float dl = m * 0x1p-20f + scene_max * 0x1p-22f;
                lo[a] = c.box.lo[a] - dl;
                hi[a] = c.box.hi[a] + dl;
            }
            node.lox[k] = lo[0];
            node.loy[k] = lo[1];
            node.loz[k] = lo[2];
            node.hix[k] = hi[0];
            node.hiy[k] = hi[1];
            node.hiz[k] = hi[2];
            if (c.leaf()) {
                int first = (int)out.tris.size();
                for (int j = 0; j < c.count; j++) {
                    int t = idx[c.first + j];
                    const float* a = P + 3 * (size_t)I[3 * (size_t)t];
                    const float* bb = P + 3 * (size_t)I[3 * (size_t)t + 1];
                    const float* cc = P + 3 * (size_t)I[3 * (size_t)t + 2];
                    TriRec r;
                    r.ax = a[0];
                    r.ay = a[1];
                    r.az = a[2];
                    r.id = t;
                    // b - a and c - a as the reference computes them (triangle.h:36-37)
                    r.e1x = bb[0] - a[0];
                    r.e1y = bb[1] - a[1];
                    r.e1z = bb[2] - a[2];
                    r.rank = rank.empty() ? -t : rank[t];
                    r.e2x = cc[0] - a[0];
                    r.e2y = cc[1] - a[1];
                    r.e2z = cc[2] - a[2];
                    r.pad = 0;
                    out.tris.push_back(r);
                }
                node.child[k] = LEAF_BIT | ((c.count - 1) << 27) | first;
                out.leaves++;
                out.max_depth = std::max(out.max_depth, it.depth + 1);
            } else {
                internal_slots[ni++] = k;
            }
        }
        // internal children get consecutive node ids
        int base = (int)out.nodes.size();
        for (int j = 0; j < ni; j++) {
            node.child[internal_slots[j]] = base + j;
            out.nodes.push_back(Node4());
        }
        out.nodes[it.n4] = node;
        for (int j = ni - 1; j >= 0; j--) {
            int k = internal_slots[j];
            work.push_back({kids[k], base + j, it.depth + 1});
        }
    }
}

// ---------------------------------------------------------------------------
// Reference tie rank.  Restates BVH::build_node / compute_split
// (src/primitive/bvh.h:57-202) with the same float/double arithmetic:
//  * centroid = (min+max)*0.5 of the triangle's vertex bounds (aabb.h:62-64)
//  * 5 bins per axis over the centroid bounds, bin j <=> c < lo + (hi-lo)*float(j+1)/float(5)
//  * cost(i) = area(bins<=i) + area(bins>i), area() in double (aabb.h:77-82), inf if one side empty
//  * axis choice c1<c2 ? (c1<c3 ? x : c2<c3 ? y : z) : (c2<c3 ? y : z)  (bvh.h:179-197)
//  * the chosen side lists are the bins concatenated in bin order (stable)
//  * leaf when fewer than 4 triangles; DFS visits the left list first (bvh.h:201,18,33)
// Deviation shared with the oracle: a split with an empty side becomes a leaf
// (the reference would recurse forever, bvh.h:200).
namespace {

struct RBox {
    float mn[3], mx[3];
    RBox() {
        for (int k = 0; k < 3; k++) {
            mn[k] = std::numeric_limits<float>::infinity();
            mx[k] = -std::numeric_limits<float>::infinity();
        }
    }
    void unite(const RBox& o) {
        for (int k = 0; k < 3; k++) {
            mn[k] = mn[k] < o.mn[k] ? mn[k] : o.mn[k];
            mx[k] = mx[k] > o.mx[k] ? mx[k] : o.mx[k];
        }
    }
    float area() const {
        double dx = std::fabs((double)(mx[0] - mn[0])), dz = std::fabs((double)(mx[2] - mn[2]));
        return (float)(2.0 * (dx * (double)(mx[1] - mn[1]) + dz * (double)(mx[1] - mn[1]) + dx * (double)(mx[2] - mn[2])));
    }
};

struct RankBuilder {
    const std::vector<RBox>& pb;
    const std::vector<float>& cen;
    std::vector<int32_t>& rank;
    std::vector<int> tmp;
    int32_t counter = 0;

    void build(std::vector<int>& prims, int first, int count) {
        if (count < 4) {
            for (int i = first; i < first + count; i++) rank[prims[i]] = counter++;
            return;
        }
        RBox cb;
        for (int i = first; i < first + count; i++) {
            const float* c = &cen[3 * (size_t)prims[i]];
            for (int k = 0; k < 3; k++) {
                cb.mn[k] = cb.mn[k] < c[k] ? cb.mn[k] : c[k];
                cb.mx[k] = cb.mx[k] > c[k] ? cb.mx[k] : c[k];
            }
        }
        float cost[3];
        int split[3];
        for (int axis = 0; axis < 3; axis++) {
            float thr[4];
            float lo = cb.mn[axis], hi = cb.mx[axis];
            for (int j = 0; j < 4; j++) thr[j] = lo + (hi - lo) * float(j + 1) / float(5);
            RBox bins[5];
            int cnt[5] = {0, 0, 0, 0, 0};
            for (int i = first; i < first + count; i++) {
                int p = prims[i];
                float cv = cen[3 * (size_t)p + axis];
                int b = 4;
                for (int j = 0; j < 4; j++)
                    if (cv < thr[j]) { b = j; break; }
                cnt[b]++;
                bins[b].unite(pb[p]);
            }
            float pc[4];
            for (int i = 0; i < 4; i++) {
                RBox l, r;
                int lc = 0, rc = 0;
                for (int j = 0; j < 5; j++) {
                    if (j <= i) { lc += cnt[j]; if (cnt[j]) l.unite(bins[j]); }
                    else { rc += cnt[j]; if (cnt[j]) r.unite(bins[j]); }
                }
                pc[i] = (lc == 0 || rc == 0) ? std::numeric_limits<float>::infinity() : l.area() + r.area();
            }
            int mi = 0;
            float mc = pc[0];
            for (int i = 1; i < 4; i++)
                if (pc[i] < mc) { mc = pc[i]; mi = i; }
            cost[axis] = mc;
            split[axis] = mi;
        }
        float c1 = cost[0], c2 = cost[1], c3 = cost[2];
        int axis;
        if (c1 < c2) axis = (c1 < c3) ? 0 : (c2 < c3 ? 1 : 2);
        else axis = (c2 < c3) ? 1 : 2;
        // stable partition by bin for the chosen axis
        float thr[4];
        float lo = cb.mn[axis], hi = cb.mx[axis];
        for (int j = 0; j < 4; j++) thr[j] = lo + (hi - lo) * float(j + 1) / float(5);
        tmp.resize(count);
        int nl = 0;
        int binc[5] = {0, 0, 0, 0, 0};
        std::vector<uint8_t> binof(count);
        for (int i = 0; i < count; i++) {
            float cv = cen[3 * (size_t)prims[first + i] + axis];
            int b = 4;
            for (int j = 0; j < 4; j++)
                if (cv < thr[j]) { b = j; break; }
            binof[i] = (uint8_t)b;
            binc[b]++;
        }
        int off[5];
        int acc = 0;
        for (int b = 0; b < 5; b++) { off[b] = acc; acc += binc[b]; }
        for (int b = 0; b <= split[axis]; b++) nl += binc[b];
        for (int i = 0; i < count; i++) tmp[off[binof[i]]++] = prims[first + i];
        std::copy(tmp.begin(), tmp.begin() + count, prims.begin() + first);
        int nr = count - nl;
        if (nl == 0 || nr == 0) {
            for (int i = first; i < first + count; i++) rank[prims[i]] = counter++;
            return;
        }
        build(prims, first, nl);
        build(prims, first + nl, nr);
    }
};

}  // namespace

void reference_tie_rank(const float* P, const int32_t* I, int nt, std::vector<int32_t>& rank) {
    std::vector<RBox> pb(nt);
    std::vector<float> cen(3 * (size_t)nt);
    for (int t = 0; t < nt; t++) {
        RBox b;
        for (int v = 0; v < 3; v++) {
            const float* p = P + 3 * (size_t)I[3 * (size_t)t + v];
            for (int k = 0; k < 3; k++) {
                // minF/maxF (math.h:197-205) of a, b, c: min(min(a,b),c)
                b.mn[k] = b.mn[k] < p[k] ? b.mn[k] : p[k];
                b.mx[k] = b.mx[k] > p[k] ? b.mx[k] : p[k];
            }
        }
        pb[t] = b;
        for (int k = 0; k < 3; k++) cen[3 * (size_t)t + k] = (b.mn[k] + b.mx[k]) * (float)0.5;
    }
    rank.assign(nt, 0);
    std::vector<int> prims(nt);
    for (int t = 0; t < nt; t++) prims[t] = t;
    RankBuilder rb{pb, cen, rank, {}, 0};
    if (nt > 0) rb.build(prims, 0, nt);
}

}  // namespace cpt
