"""Pins of the SD-tree restatements (oracle/sdtree.py for the refinement the
library's host code is compared with; the oracle's flattened-tree lookups for
what the device samples) by properties of the algorithm of the reference's
practical path guiding (src/algorithms/sdtree.h):

  DTree::adapt_and_build_leaves (sdtree.h:172-233, prune 104-121): the sampling
    leaves partition the (phi, cos theta) domain; every leaf holds at most 1% of the
    tree's energy (splat mass, spread evenly over a leaf's area when it is split);
    every leaf's parent holds at least 1% (otherwise it would have stayed a leaf)
  pdf (sdtree.h:525-538) integrates to 1 over the sphere, and sample_direction
    (400-423) draws from it: E[f / pdf] = the integral of f
  SDTree::subdivide (sdtree.h:459-504): spatial leaves with more than
    12000 * int(2^(k/2)) splats split int(log2(count / threshold)) + 1 times, at
    midpoints along axis depth mod 3 (leaf counts worked out by hand)
"""
import math

import numpy as np
import pytest

import oracle
from oracle import sdtree

FULL = sdtree.FULL  # (-pi, -1, 2 pi, 2) in float


def _random_quadtree(rng, depth, p_split=0.55):
    if depth == 0 or rng.random() > p_split:
        v = 0.0 if rng.random() < 0.2 else float(np.float32(rng.exponential(1.0) ** 3))
        return [v]
    return [None] + [_random_quadtree(rng, depth - 1, p_split) for _ in range(4)]


def _leaves(node, rect=FULL):
    if len(node) == 1:
        yield rect, node[0]
        return
    for i in range(4):
        yield from _leaves(node[1 + i], sdtree._child_rect(rect, i))


def _energy(rect, orig):
    """Splat mass inside rect, each original leaf's mass spread evenly over its area."""
    x0, y0, w, h = rect
    e = 0.0
    for (a0, b0, aw, bh), v in orig:
        ox = max(0.0, min(x0 + w, a0 + aw) - max(x0, a0))
        oy = max(0.0, min(y0 + h, b0 + bh) - max(y0, b0))
        if ox > 0 and oy > 0:
            e += v * ox * oy / (aw * bh)
    return e


def _parent(rect):
    x0, y0, w, h = rect
    # cell indices at the leaf's level (rounded: the rectangles are float)
    ix, iy = round((x0 - FULL[0]) / w), round((y0 - FULL[1]) / h)
    return (FULL[0] + (ix // 2) * 2 * w, FULL[1] + (iy // 2) * 2 * h, 2 * w, 2 * h)


@pytest.mark.parametrize("seed", range(6))
def test_adapt_partitions_and_keeps_one_percent_rule(seed):
    rng = np.random.default_rng(seed)
    root = _random_quadtree(rng, 5)
    orig = list(_leaves(root))
    d = sdtree.DTree(root=root)
    total = np.float32(0.0)
    for _, v in orig:  # the library adds splats in float
        total = np.float32(total + np.float32(v))
    d.radiance = float(total)
    if d.radiance == 0.0:
        return
    d.adapt()
    thr = 0.01 * float(total)
    leaves = d.leaves
    # the sampling list is the new quadtree's leaves
    assert sorted(leaves) == sorted(r for r, _ in _leaves(d.root))
    # partition: areas add up to the domain's and random points fall in exactly one leaf
    assert abs(sum(w * h for _, _, w, h in leaves) - 4 * math.pi) < 1e-4
    pts = rng.random((4000, 2)) * [2 * math.pi, 2.0] + [-math.pi, -1.0]
    L = np.array(leaves, np.float64)
    inside = ((pts[:, None, 0] >= L[None, :, 0]) & (pts[:, None, 0] < L[None, :, 0] + L[None, :, 2]) &
              (pts[:, None, 1] >= L[None, :, 1]) & (pts[:, None, 1] < L[None, :, 1] + L[None, :, 3]))
    assert (inside.sum(axis=1) == 1).all()
    # the 1% rule, from the original splats alone
    for r in leaves:
        assert _energy(r, orig) <= thr * (1 + 1e-4) + 1e-12, r
        if r[2] < FULL[2] * 0.99:
            assert _energy(_parent(r), orig) >= thr * (1 - 1e-4), r
    print(f"seed {seed}: {len(orig)} splat leaves -> {len(leaves)} sampling leaves")


def _subdivided_tree(rng, counts_levels=3):
    """An SD-tree over [0,1]^3 split 2^levels ways, each directional tree adapted
    from random splats."""
    sd = sdtree.SDTree((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    d = sd.root[1]
    d.root = _random_quadtree(rng, 6, 0.6)
    tot = np.float32(0.0)
    for _, v in _leaves(d.root):
        tot = np.float32(tot + np.float32(v))
    d.radiance = float(tot)
    d.count = 12000 * (2 ** counts_levels) - 1  # int(log2(ratio)) + 1 = levels
    sd.subdivide(12000, 0)
    return sd


def test_pdf_integrates_to_one_and_sampling_matches():
    rng = np.random.default_rng(3)
    sd = _subdivided_tree(rng)
    t = sdtree.flatten(sd)
    k = 9  # dyadic grid finer than the deepest quadtree leaf (depth <= 7 here)
    depth = max(int(round(math.log2(FULL[2] / r[2]))) for r in t["rect"])
    assert depth <= k
    nx = ny = 2 ** k
    phi = (np.arange(nx) + 0.5) / nx * 2 * math.pi - math.pi
    ct = (np.arange(ny) + 0.5) / ny * 2 - 1
    P, C = np.meshgrid(phi, ct, indexing="ij")
    S = np.sqrt(1 - C * C)
    dirs = np.stack([np.cos(P) * S, np.sin(P) * S, C], -1).reshape(-1, 3)
    cell = (2 * math.pi / nx) * (2.0 / ny)
    for p in rng.random((5, 3)):
        pdf = oracle.sd_pdf(t, np.tile(p, (dirs.shape[0], 1)), dirs).astype(np.float64)
        # d omega = d phi d cos theta: the sum over cells is the integral
        assert abs(pdf.sum() * cell - 1) < 2e-3, pdf.sum() * cell
        # sampling: the estimator f / pdf of f = 1 + z^2 integrates to 4 pi + 4 pi / 3
        n = 200_000
        w, q = oracle.sd_sample(t, np.tile(p, (n, 1)), rng.random((n, 3)))
        assert np.allclose(q, oracle.sd_pdf(t, np.tile(p, (n, 1)), w), rtol=1e-5)
        est = (1 + w[:, 2].astype(np.float64) ** 2) / q
        want = 4 * math.pi * 4 / 3
        assert abs(est.mean() - want) < 4 * est.std() / math.sqrt(n) + 1e-3, (est.mean(), want)


@pytest.mark.parametrize("count,iteration,leaves", [
    (12000, 0, 1),      # not above the threshold 12000 * int(2^0) = 12000
    (12001, 0, 2),      # ratio 1.00008: int(log2) = 0, one split
    (24000, 0, 4),      # ratio 2: int(log2) = 1, two splits
    (100000, 0, 16),    # ratio 8.33: int(3.06) = 3, four splits
    (30000, 1, 4),      # threshold 12000 * int(2^0.5) = 12000, ratio 2.5: two splits
    (30000, 2, 2),      # threshold 12000 * int(2^1) = 24000, ratio 1.25: one split
    (24000, 2, 1),      # equal to the threshold: no split
    (200000, 3, 16),    # threshold 12000 * int(2^1.5) = 24000, ratio 8.33: four splits
    (200000, 4, 8),     # threshold 12000 * int(2^2) = 48000, ratio 4.17: three splits
])
def test_spatial_split_counts(count, iteration, leaves):
    sd = sdtree.SDTree((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    sd.root[1].count = count
    sd.root[1].radiance = 1.0
    sd.root[1].root = [1.0]
    sd.subdivide(12000, iteration)
    assert len(sd.structure()) == leaves


def test_spatial_splits_are_midpoints_along_depth_mod_3():
    """Split 16 ways (four levels): the directional tree a point falls in must be
    the one of its cell in a 4 x 2 x 2 grid (levels split x, y, z, x), and every
    cell has its own tree."""
    sd = sdtree.SDTree((-2.0, 0.0, 4.0), (2.0, 1.0, 6.0))
    sd.root[1].count = 100000
    sd.root[1].radiance = 1.0
    sd.root[1].root = [1.0]
    sd.subdivide(12000, 0)
    t = sdtree.flatten(sd)
    assert len(t["dinfo"]) == 16
    rng = np.random.default_rng(1)
    lo, hi = np.array([-2.0, 0.0, 4.0]), np.array([2.0, 1.0, 6.0])
    pts = lo + rng.random((5000, 3)) * (hi - lo)
    cell = np.floor((pts - lo) / (hi - lo) * [4, 2, 2]).astype(int)
    key = cell[:, 0] * 4 + cell[:, 1] * 2 + cell[:, 2]
    dt = oracle.sd_lookup(t, pts)
    for k in range(16):
        assert len(set(dt[key == k])) == 1
    assert len(set(dt)) == 16
