"""Launch configuration: results do not depend on it, and cpt_autotune's choice is
what later renders use."""
import numpy as np
import pytest

import cpt

pytestmark = pytest.mark.gpu


def _render(ctx, W, H, **kw):
    st = ctx.render(W, H, 4, max_depth=8, seed=11, **kw)
    _, s, s2 = ctx.read_film(sums=True, sumsq=True)
    return st, s, s2


def test_film_independent_of_launch_configuration(ctx):
    W, H = 96, 64
    ctx.upload(cpt.scenes.box_room(res=(W, H)))
    _, s0, q0 = _render(ctx, W, H)
    configs = [dict(block_size=64, refill=1), dict(block_size=256, refill=64, shade_block=1024),
               dict(block_size=128, grid_blocks=7, shade_block=64), dict(refill=17, max_paths=5000),
               dict(block_size=64, grid_blocks=3000, shade_block=512, refill=40)]
    for cfg in configs:
        _, s, q = _render(ctx, W, H, **cfg)
        assert np.array_equal(s, s0) and np.array_equal(q, q0), cfg


def test_autotune_choice_is_used_and_exact(ctx):
    W, H = 256, 144
    ctx.upload(cpt.scenes.box_room(res=(W, H)))
    r = ctx.autotune(W, H, max_depth=8)
    print("autotune", r)
    assert r["trials"] == 10 and r["samples_per_s"] > 0
    assert r["block_size"] in (64, 128, 256) and r["shade_block"] in (128, 256, 512)
    assert r["refill"] in (16, 32, 48, 64) and r["grid_blocks"] >= 0
    st, s, q = _render(ctx, W, H)
    assert st["block_size"] == r["block_size"]
    _, s1, q1 = _render(ctx, W, H, block_size=r["block_size"], grid_blocks=r["grid_blocks"],
                        shade_block=r["shade_block"], refill=r["refill"])
    assert np.array_equal(s, s1) and np.array_equal(q, q1)
    # another size falls back to the defaults
    st2 = ctx.render(128, 72, 1, max_depth=8, seed=1)
    assert st2["block_size"] == 128
