// Host SD-tree: restates src/algorithms/sdtree.h (DTree::adapt_and_build_leaves
// 172-233, get_radiance_and_prune 104-121, SDTree::subdivide 459-504) with the
// same float/double roundings; the device side (sampling, pdf, splats) lives in
// sdtree_dev.h.
#include "sdtree.h"

#include <cmath>
#include <utility>

namespace cpt {

namespace {
// DTree::Leaf({-M_PI, -1}, {2.0*M_PI, 2}) (sdtree.h:118)
SdRect full_rect() { return SdRect{(float)-M_PI, -1.0f, (float)(2.0 * M_PI), 2.0f}; }

// quadrant i of l (sdtree.h:218-224): begin + size*0.5*(i%2, i/2) in double, size*0.5
SdRect child_rect(const SdRect& l, int i) {
    SdRect r;
    r.bx = (float)((double)l.bx + (double)l.sx * 0.5 * (double)(i % 2));
    r.by = (float)((double)l.by + (double)l.sy * 0.5 * (double)(i / 2));
    r.sx = (float)((double)l.sx * 0.5);
    r.sy = (float)((double)l.sy * 0.5);
    return r;
}
}  // namespace

SdDTree::SdDTree() : q(1), leaves(1, full_rect()) {}

// get_radiance_and_prune (sdtree.h:104-121): subtrees below the threshold
// collapse into one leaf holding their sum
float SdDTree::prune(int n, float threshold) {
    if (q[n].child < 0) return q[n].value;
    float rad = 0.0f;
    const int c0 = q[n].child;
    for (int i = 0; i < 4; i++) rad += prune(c0 + i, threshold);
    if (rad < threshold) {
        q[n].child = -1;
        q[n].value = rad;
    }
    return rad;
}

void SdDTree::adapt() {
    if (radiance == 0.0f) return;
    const float threshold = (float)((double)radiance * 0.01);
    leaves.clear();
    prune(0, threshold);
    // depth-first, last pushed first, as the reference's explicit stacks
    std::vector<std::pair<int, SdRect>> st;
    st.push_back({0, full_rect()});
    while (!st.empty()) {
        const int n = st.back().first;
        const SdRect l = st.back().second;
        st.pop_back();
        if (q[n].child < 0) {
            const float ratio = q[n].value / threshold;
            if (ratio > 1.0f) {
                const float quarter = (float)((double)q[n].value * 0.25);
                const int base = (int)q.size();
                q.resize(base + 4);
                for (int i = 0; i < 4; i++) q[base + i].value = quarter;
                q[n].child = base;
                q[n].value = 0.0f;
            } else {
                q[n].value = 0.0f;
                leaves.push_back(l);
                continue;
            }
        }
        const int c0 = q[n].child;
        for (int i = 0; i < 4; i++) st.push_back({c0 + i, child_rect(l, i)});
    }
    count = 0;
    radiance = 0.0f;
    compact();
}

// drop the nodes pruning left unreachable
void SdDTree::compact() {
    std::vector<SdQuad> out;
    out.reserve(q.size());
    out.push_back(q[0]);
    std::vector<std::pair<int, int>> st{{0, 0}};  // (old, new)
    while (!st.empty()) {
        auto [o, n] = st.back();
        st.pop_back();
        if (q[o].child < 0) continue;
        const int base = (int)out.size();
        for (int i = 0; i < 4; i++) out.push_back(q[q[o].child + i]);
        out[n].child = base;
        for (int i = 0; i < 4; i++) st.push_back({q[o].child + i, base + i});
    }
    q.swap(out);
}

SdTree::SdTree(const float lo[3], const float hi[3]) : s(1), d(1) {
    for (int k = 0; k < 3; k++) {
        bmin[k] = lo[k];
        bmax[k] = hi[k];
    }
    s[0].dtree = 0;
}

void SdTree::update() {
    subdivide(c, iteration);
    iteration++;
}

// SDTree::subdivide (sdtree.h:459-504): a leaf with more than c*int(2^(k/2))
// splats is split int(log2(count/threshold))+1 times, every descendant
// receiving a copy of its adapted directional tree; other leaves only adapt.
void SdTree::subdivide(int cc, int it) {
    const int threshold = cc * (int)std::pow(2, it * 0.5);
    std::vector<std::pair<int, int>> st{{0, 0}};  // (snode, remaining splits)
    while (!st.empty()) {
        const int n = st.back().first;
        const int sub = st.back().second;
        st.pop_back();
        if (s[n].child < 0) {
            const int di = s[n].dtree;
            if (d[di].count > threshold || sub) {
                int to_sub = sub;
                if (sub == 0) {
                    const float ratio = (float)d[di].count / (float)threshold;
                    to_sub = (int)std::log2(ratio) + 1;
                    d[di].adapt();
                }
                const int c0 = (int)s.size();
                const int d0 = (int)d.size();
                const SdDTree copy = d[di];
                d.push_back(copy);
                d.push_back(copy);
                s.resize(c0 + 2);
                s[c0].dtree = d0;
                s[c0 + 1].dtree = d0 + 1;
                s[n].child = c0;
                s[n].dtree = -1;
                d[di] = SdDTree();  // no longer referenced
                d[di].q.clear();
                d[di].leaves.clear();
                st.push_back({c0, to_sub - 1});
                st.push_back({c0 + 1, to_sub - 1});
            } else {
                d[di].adapt();
            }
        } else {
            st.push_back({s[n].child, 0});
            st.push_back({s[n].child + 1, 0});
        }
    }
    // renumber the directional trees still referenced
    std::vector<SdDTree> live;
    for (auto& sn : s)
        if (sn.child < 0) {
            live.push_back(std::move(d[sn.dtree]));
            sn.dtree = (int)live.size() - 1;
        }
    d.swap(live);
}

void SdTree::flatten(SdFlat& f) const {
    f.snode.assign(s.size(), 0);
    for (size_t i = 0; i < s.size(); i++) f.snode[i] = s[i].child >= 0 ? s[i].child : -1 - s[i].dtree;
    f.dinfo.assign(4 * d.size(), 0);
    f.quad.clear();
    f.rect.clear();
    f.slot_dtree.clear();
    f.slot_quad.clear();
    for (size_t t = 0; t < d.size(); t++) {
        const int qb = (int)f.quad.size();
        f.dinfo[4 * t] = qb;
        f.dinfo[4 * t + 1] = (int)(f.rect.size() / 4);
        f.dinfo[4 * t + 2] = (int)d[t].leaves.size();
        for (size_t k = 0; k < d[t].q.size(); k++) {
            const SdQuad& n = d[t].q[k];
            if (n.child >= 0) {
                f.quad.push_back(qb + n.child);
            } else {
                f.quad.push_back(-1 - (int)f.slot_dtree.size());
                f.slot_dtree.push_back((int)t);
                f.slot_quad.push_back((int)k);
            }
        }
        for (const SdRect& r : d[t].leaves) {
            f.rect.push_back(r.bx);
            f.rect.push_back(r.by);
            f.rect.push_back(r.sx);
            f.rect.push_back(r.sy);
        }
    }
}

void SdTree::add_splats(const SdFlat& f, const float* sums, const uint32_t* counts) {
    for (size_t k = 0; k < f.slot_dtree.size(); k++) {
        SdDTree& t = d[f.slot_dtree[k]];
        t.q[f.slot_quad[k]].value += sums[k];
        t.count += counts[k];
        t.radiance += sums[k];
    }
}

}  // namespace cpt
