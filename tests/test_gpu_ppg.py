"""SD-tree guided arm (the reference's PPG renderer, src/renderer/ppg.h, with its
SD-tree, src/algorithms/sdtree.h) through the C ABI: one iteration matches the
oracle path by path for a given tree, the tree refinement matches the oracle's
restatement exactly, and the full schedule is unbiased."""
import numpy as np
import pytest

import cpt
import oracle
from oracle.sdtree import SDTree

pytestmark = pytest.mark.gpu


def _step(ctx, W, H, spp, seed, view, depth=12):
    ctx.render(W, H, spp, max_depth=depth, seed=seed, view=view, flags=cpt.CPT_RENDER_PPG_STEP)
    _, sums = ctx.read_film(sums=True)
    return sums


def _compare_step(ctx, scene, W, H, spp, seed, view, depth=12):
    tree = ctx.ppg_export()  # the tree this step uses
    g = _step(ctx, W, H, spp, seed, view, depth)
    after = ctx.ppg_export()
    r, ssum, scnt = oracle.OracleScene(scene).ppg_render(W, H, spp, depth, seed, view, tree)
    per_px = (np.abs(g - r) / np.maximum(np.abs(r), 1e-3)).max(axis=-1)
    # a path can leave the oracle's when a last-bit difference of the device's
    # cosf/sinf/atan2f flips a branch; everything else agrees to float rounding
    assert (per_px < 1e-4).mean() > 0.98, (per_px < 1e-4).mean()
    assert abs(g.sum() - r.sum()) / r.sum() < 2e-3
    c = after["counts"].astype(np.int64)
    assert abs(c.sum() - scnt.sum()) <= 0.01 * scnt.sum()
    assert (np.abs(c - scnt) <= np.maximum(2, 0.02 * scnt)).mean() > 0.99
    s = after["sums"].astype(np.float64)
    assert abs(s.sum() - ssum.sum()) / ssum.sum() < 1e-2
    return tree, after


def test_ppg_iterations_and_updates_match_oracle(ctx):
    W, H = 64, 64
    scene = cpt.scenes.box_room(res=(W, H))
    ctx.upload(scene)
    ctx.ppg_reset()
    t0, s0 = _compare_step(ctx, scene, W, H, 4, 5, 0)
    # the initial tree: one spatial leaf, one directional leaf over the sphere
    assert len(t0["snode"]) == 1 and len(t0["rect"]) == 1
    for it, (spp, seed) in enumerate([(8, 6), (8, 7)], start=1):
        want = SDTree.from_export(s0)
        want.subdivide(12000, it)
        ctx.ppg_update()
        got = ctx.ppg_export()
        got_tree = SDTree.from_export(got)
        assert got_tree.structure() == want.structure(), f"update {it}"
        n_leaves = [len(x[1]) for x in got_tree.structure()]
        print(f"update {it}: {len(got_tree.structure())} spatial leaves, directional leaves {min(n_leaves)}..{max(n_leaves)}")
        assert len(got["snode"]) > 1  # enough splats to split space
        _, s0 = _compare_step(ctx, scene, W, H, spp, seed, it)
    # after the second update the quadtrees follow the light (not uniform)
    assert max(n_leaves) > 256 or min(n_leaves) < 256
