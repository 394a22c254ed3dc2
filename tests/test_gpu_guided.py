"""Guided path tracing (learned MLP guide, one-sample MIS with the BSDF, NEE kept)
through the C ABI: unbiased against plain PT-NEE-MIS, and better than it at equal
sample count on a hard-lighting interior once trained."""
import numpy as np
import pytest

import cpt

pytestmark = pytest.mark.gpu


def relmse(img, ref):
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2)))


def _z(gs, gq, n1, rs, rq, n2):
    m1, m2 = gs / n1, rs / n2
    v1 = np.maximum(gq / n1 - m1 ** 2, 0) * n1 / (n1 - 1)
    v2 = np.maximum(rq / n2 - m2 ** 2, 0) * n2 / (n2 - 1)
    se = np.sqrt(v1 / n1 + v2 / n2)
    ok = se > 0
    return (m1 - m2)[ok] / se[ok]


def test_guided_training_produces_samples_and_steps(ctx):
    s = cpt.scenes.box_room(res=(64, 64))
    ctx.upload(s)
    ctx.guide_init(seed=1, lr=3e-3, train_paths=4096)
    st = ctx.render(64, 64, 4, max_depth=12, seed=3, flags=cpt.CPT_RENDER_GUIDED | cpt.CPT_RENDER_GUIDE_TRAIN)
    print("guided stats", st["passes"], st["guide_steps"], st["guide_samples"])
    assert st["passes"] == 16 and st["guide_steps"] == 16 and st["guide_samples"] > 16 * 2000
    img = ctx.read_film()
    assert np.isfinite(img).all() and img.mean() > 0


def test_guided_is_unbiased(ctx):
    W = H = 48
    s = cpt.scenes.box_room(res=(W, H))
    ctx.upload(s)
    ctx.render(W, H, 16, max_depth=16, seed=5)
    _, ps, pq = ctx.read_film(sums=True, sumsq=True)
    ctx.guide_init(seed=2, lr=3e-3, train_paths=8192)
    ctx.render(W, H, 16, max_depth=16, seed=6, flags=cpt.CPT_RENDER_GUIDED | cpt.CPT_RENDER_GUIDE_TRAIN)
    _, gs, gq = ctx.read_film(sums=True, sumsq=True)
    z = _z(gs, gq, 256, ps, pq, 256)
    frac = (np.abs(z) > 3).mean()
    print(f"guided vs plain: 3-sigma exceedance {frac * 100:.2f}%, mean z {z.mean():+.4f}")
    assert frac < 0.01 and abs(z.mean()) < 0.06


def test_trained_guide_beats_plain_on_hard_lighting(ctx):
    W, H = 96, 54
    s = cpt.scenes.interior_scene(res=(W, H), n_boxes=3000)
    ctx.upload(s)
    ctx.render(W, H, 48, max_depth=12, seed=100)  # 2304 spp reference
    ref = ctx.read_film()
    ctx.render(W, H, 8, max_depth=12, seed=7)      # plain, 64 spp
    plain = ctx.read_film()
    ctx.guide_init(seed=3, lr=3e-3, train_paths=W * H, train_steps=2)
    st = ctx.render(W, H, 8, max_depth=12, seed=8, flags=cpt.CPT_RENDER_GUIDED | cpt.CPT_RENDER_GUIDE_TRAIN)
    online = ctx.read_film()
    ctx.render(W, H, 8, max_depth=12, seed=9, flags=cpt.CPT_RENDER_GUIDED)  # frozen, trained guide
    frozen = ctx.read_film()
    e_plain, e_online, e_frozen = relmse(plain, ref), relmse(online, ref), relmse(frozen, ref)
    print(f"relMSE @64spp: plain {e_plain:.4f}  guided online {e_online:.4f}  guided trained {e_frozen:.4f} "
          f"({st['guide_steps']} steps, {st['guide_samples']} samples)")
    assert e_frozen < 0.9 * e_plain
