"""Oracle restatement of the SD-tree update of the reference's practical path
guiding (src/algorithms/sdtree.h), on the flattened layout the library exports
(cpt.Context.ppg_export).  Test infrastructure only.

  DTree::get_radiance_and_prune       sdtree.h:104-121
  DTree::adapt_and_build_leaves       sdtree.h:172-233
  SDTree::subdivide                   sdtree.h:459-504

Float rounding follows the reference: quadrant rectangles are formed in double
and stored as float, the prune sums and thresholds are float.
"""
import math

import numpy as np

F = np.float32


def _f(x):
    return float(F(x))


FULL = (_f(-math.pi), -1.0, _f(2.0 * math.pi), 2.0)


def _child_rect(r, i):
    bx, by, sx, sy = r
    return (_f(bx + sx * 0.5 * (i % 2)), _f(by + sy * 0.5 * (i // 2)), _f(sx * 0.5), _f(sy * 0.5))


class DTree:
    """Quadtree as nested lists: a leaf is [value], an internal node [None, c0, c1, c2, c3]."""

    def __init__(self, root=None, leaves=None):
        self.root = root if root is not None else [0.0]
        self.leaves = leaves if leaves is not None else [FULL]
        self.count = 0
        self.radiance = 0.0

    def copy(self):
        def cp(n):
            return [n[0]] if len(n) == 1 else [None] + [cp(c) for c in n[1:]]
        d = DTree(cp(self.root), list(self.leaves))
        d.count, d.radiance = self.count, self.radiance
        return d

    @staticmethod
    def _prune(n, thr):
        if len(n) == 1:
            return n[0]
        rad = 0.0
        for c in n[1:]:
            rad = _f(rad + DTree._prune(c, thr))
        if rad < thr:
            del n[1:]
            n[0] = rad
        return rad

    def adapt(self):
        if self.radiance == 0.0:
            return
        thr = _f(self.radiance * 0.01)
        self.leaves = []
        self._prune(self.root, thr)
        st = [(self.root, FULL)]
        while st:
            n, r = st.pop()
            if len(n) == 1:
                ratio = _f(n[0] / thr) if thr != 0 else math.inf
                if ratio > 1.0:
                    q = _f(n[0] * 0.25)
                    n[:] = [None, [q], [q], [q], [q]]
                else:
                    n[0] = 0.0
                    self.leaves.append(r)
                    continue
            for i in range(4):
                st.append((n[1 + i], _child_rect(r, i)))
        self.count = 0
        self.radiance = 0.0


class SDTree:
    """Spatial binary tree: a leaf is ('L', DTree), an internal node ('N', lo, hi)."""

    def __init__(self, lo, hi):
        self.lo, self.hi = list(lo), list(hi)
        self.root = ["L", DTree()]

    def subdivide(self, c, iteration):
        threshold = c * int(math.pow(2, iteration * 0.5))
        st = [(self.root, 0)]
        while st:
            n, sub = st.pop()
            if n[0] == "L":
                d = n[1]
                if d.count > threshold or sub:
                    to_sub = sub
                    if sub == 0:
                        ratio = _f(F(d.count) / F(threshold))
                        to_sub = int(_f(math.log2(ratio))) + 1
                        d.adapt()
                    a, b = ["L", d.copy()], ["L", d.copy()]
                    n[:] = ["N", a, b]
                    st.append((a, to_sub - 1))
                    st.append((b, to_sub - 1))
                else:
                    d.adapt()
            else:
                st.append((n[1], 0))
                st.append((n[2], 0))

    @staticmethod
    def from_export(t):
        """Rebuild the tree from cpt.Context.ppg_export, with the exported splat
        sums and counts added to its quadtree leaves (the state an update starts from)."""
        sd = SDTree(t["bounds"][:3], t["bounds"][3:])
        snode, dinfo, quad, rect = t["snode"], t["dinfo"], t["quad"], t["rect"]
        sums, counts = t["sums"], t["counts"]

        def qnode(q):
            v = int(quad[q])
            if v < 0:
                return [float(sums[-1 - v])]
            return [None] + [qnode(v + i) for i in range(4)]

        def dtree(di):
            qb, lb, nl, _ = (int(x) for x in dinfo[di])
            d = DTree(qnode(qb), [tuple(float(x) for x in rect[lb + j]) for j in range(nl)])
            # the totals: splats of this dtree's leaves
            slots = []

            def collect(q):
                v = int(quad[q])
                if v < 0:
                    slots.append(-1 - v)
                else:
                    for i in range(4):
                        collect(v + i)
            collect(qb)
            slots.sort()  # the library adds them in accumulator order
            d.count = int(sum(int(counts[s]) for s in slots))
            rad = F(0.0)
            for s in slots:
                rad = F(rad + F(sums[s]))
            d.radiance = float(rad)
            return d

        def snode_of(n):
            v = int(snode[n])
            if v < 0:
                return ["L", dtree(-1 - v)]
            return ["N", snode_of(v), snode_of(v + 1)]
        sd.root = snode_of(0)
        return sd

    def structure(self):
        """Canonical description for comparisons: per spatial leaf in depth-first
        order (lower child first), its quadtree shape and sampling leaves."""
        out = []

        def qshape(n):
            return "L" if len(n) == 1 else "(" + "".join(qshape(c) for c in n[1:]) + ")"

        def walk(n):
            if n[0] == "L":
                out.append((qshape(n[1].root), tuple(n[1].leaves)))
            else:
                walk(n[1])
                walk(n[2])
        walk(self.root)
        return out


def flatten(sd):
    """The flattened layout of cpt.Context.ppg_export (without splat sums) for an
    SDTree of this module: snode (>= 0: children v, v+1; < 0: dtree -1-v), dinfo
    (quad base, leaf base, leaf count, 0 per dtree), quad (>= 0: first of four
    children, child i = x half i % 2, y half i // 2; < 0: accumulator -1-v), rect (the
    sampling leaves), bounds (lo, hi)."""
    snode, dinfo, quad, rect = [0], [], [], []
    slots = [0]

    def qfill(idx, n):
        if len(n) == 1:
            quad[idx] = -1 - slots[0]
            slots[0] += 1
            return
        base = len(quad)
        quad.extend([0] * 4)
        quad[idx] = base
        for i in range(4):
            qfill(base + i, n[1 + i])

    def sfill(idx, n):
        if n[0] == "L":
            d = n[1]
            snode[idx] = -1 - len(dinfo)
            qb = len(quad)
            quad.append(0)
            qfill(qb, d.root)
            dinfo.append((qb, len(rect), len(d.leaves), 0))
            rect.extend(d.leaves)
            return
        base = len(snode)
        snode.extend([0, 0])
        snode[idx] = base
        sfill(base, n[1])
        sfill(base + 1, n[2])

    sfill(0, sd.root)
    return {"snode": np.array(snode, np.int32), "dinfo": np.array(dinfo, np.int32).reshape(-1, 4),
            "quad": np.array(quad, np.int32), "rect": np.array(rect, np.float32).reshape(-1, 4),
            "bounds": np.array(list(sd.lo) + list(sd.hi), np.float32)}
