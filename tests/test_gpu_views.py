"""Batched multi-view rendering (multi-view passes) through the C ABI."""
import numpy as np
import pytest

import cpt

pytestmark = pytest.mark.gpu


def _views(n, seed=0):
    rng = np.random.default_rng(seed)
    sizes = [(64, 64), (48, 36), (96, 72), (80, 45)]
    cams, whs = [], []
    for v in range(n):
        w, h = sizes[rng.integers(len(sizes))]
        spp = int(rng.choice([1, 2, 4]))
        ang = rng.uniform(-0.4, 0.4)
        o = (3.9 * np.sin(ang), rng.uniform(-0.3, 0.3), 3.9 * np.cos(ang))
        cams.append(cpt.scenes.camera_frame(o, (0, 0, 0), (0, 1, 0), 0.35411, w, h))
        whs.append((w, h, spp))
    return np.stack(cams), np.array(whs, np.int32)


def test_views_match_single_view_renders(ctx):
    s = cpt.scenes.box_room()
    ctx.upload(s)
    cams, whs = _views(6)
    imgs, ms, st = ctx.render_views(cams, whs, max_depth=16, seed=5)
    assert st["paths"] == int(np.sum(whs[:, 0] * whs[:, 1] * whs[:, 2] ** 2))
    assert np.all(ms > 0)
    for v in (0, 3, 5):
        w, h, spp = whs[v]
        # same per-path random streams (view id in the path id) -> same radiance
        # per path; only the FP64 summation order differs
        ctx.upload(s, camera=cams[v])
        ctx.render(int(w), int(h), int(spp), max_depth=16, seed=5, view=v)
        ref = ctx.read_film()
        assert np.allclose(imgs[v], ref, rtol=1e-5, atol=1e-7), v


def test_views_uneven_batch_completes(ctx):
    ctx.upload(cpt.scenes.box_room())
    cams, whs = _views(24, seed=3)
    imgs, ms, st = ctx.render_views(cams, whs, max_depth=8, seed=1)
    assert len(imgs) == 24 and all(np.isfinite(i).all() for i in imgs)
    assert st["extension_rays"] >= st["paths"]
    assert np.all(ms > 0) and np.percentile(ms, 99) <= st["seconds"] * 1e3 + 1.0
    # shortest job first: a view with fewer paths never completes after one with more
    work = whs[:, 0].astype(np.int64) * whs[:, 1] * whs[:, 2] ** 2
    late = [(a, b) for a in range(24) for b in range(24) if work[a] < work[b] and ms[a] > ms[b]]
    assert not late, late
