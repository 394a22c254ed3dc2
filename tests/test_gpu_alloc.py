"""Calls after a failed device allocation: every buffer group is reallocated by
the next call (no stale capacity in front of a null or short buffer), and
CPT_RENDER_ACCUMULATE is refused when the film it would add to was reallocated
or has another layout.  Failures are injected from the host
(cpt_debug_fail_alloc), so no kernel ever sees a missing buffer."""
import numpy as np
import pytest

import cpt

pytestmark = pytest.mark.gpu

GUIDED_TRAIN = cpt.CPT_RENDER_GUIDED | cpt.CPT_RENDER_GUIDE_TRAIN


def _guided_ctx(scene):
    c = cpt.Context(0)
    c.upload(scene)
    c.guide_init(seed=4, lr=3e-3, train_paths=1024)
    return c


def test_guided_render_after_each_failed_allocation():
    """Fail the k-th device allocation of a larger guided training render (every
    allocation site of the render in turn: path state and queues, traversal
    stack, film, guide mixtures, training records and samples, view table, pass
    clock, training rows); the next, smaller render must run and give the film
    of a context that never failed."""
    scene = cpt.scenes.box_room(res=(32, 32))
    with _guided_ctx(scene) as c:
        c.render(32, 32, 1, max_depth=8, seed=5, flags=GUIDED_TRAIN)
        _, ref = c.read_film(sums=True)
    failed = 0
    for k in range(32):
        with _guided_ctx(scene) as c:
            print(f"case {k}", flush=True)
            c.debug_fail_alloc(k)
            try:
                c.render(48, 48, 1, max_depth=8, seed=5, flags=GUIDED_TRAIN)
            except cpt.CptError as e:
                assert e.code == cpt.CPT_E_HIP and "injected" in str(e)
                failed += 1
            c.debug_fail_alloc(-1)
            # the same guide state as the reference (a completed render has trained it)
            c.guide_init(seed=4, lr=3e-3, train_paths=1024)
            c.render(32, 32, 1, max_depth=8, seed=5, flags=GUIDED_TRAIN)
            _, got = c.read_film(sums=True)
            assert np.array_equal(got, ref), f"film differs after a failure at allocation {k}"
    print(f"{failed} allocation sites failed in turn")
    assert failed >= 20


@pytest.mark.parametrize("after", [0, 1])
def test_train_step_after_failed_allocation(after):
    """after 0: the sample buffer's reallocation fails; after 1: the row buffer's.
    A later call with fewer samples must reallocate and equal a fresh context."""
    rng = np.random.default_rng(3)
    s = np.zeros((4000, 11), np.float32)
    s[:, :7] = rng.random((4000, 7))
    d = rng.normal(size=(4000, 3))
    s[:, 7:10] = d / np.linalg.norm(d, axis=1, keepdims=True)
    s[:, 10] = rng.random(4000) + 0.1
    scene = cpt.scenes.box_room(res=(16, 16))
    with _guided_ctx(scene) as c:
        ref_loss, ref_grad = c.guide_train_step(s[:1000], apply=False, want_grad=True)
    with _guided_ctx(scene) as c:
        c.guide_train_step(s[:2000], apply=False)
        c.debug_fail_alloc(after)
        with pytest.raises(cpt.CptError) as ei:
            c.guide_train_step(s, apply=False)
        assert ei.value.code == cpt.CPT_E_HIP
        loss, grad = c.guide_train_step(s[:1000], apply=False, want_grad=True)
    assert loss == ref_loss and np.array_equal(grad, ref_grad)


def test_accumulate_refused_on_new_or_other_film():
    scene = cpt.scenes.box_room(res=(32, 32))
    with cpt.Context(0) as c:
        c.upload(scene)
        c.render(32, 32, 2, max_depth=8, seed=1)  # 4096 path slots, 1024-pixel film
        c.render(32, 32, 2, max_depth=8, seed=2, flags=cpt.CPT_RENDER_ACCUMULATE)
        with pytest.raises(cpt.CptError) as ei:  # another size
            c.render(16, 32, 2, max_depth=8, seed=3, flags=cpt.CPT_RENDER_ACCUMULATE)
        assert ei.value.code == cpt.CPT_E_INVALID
        # same path slots, larger film: the film's reallocation is the allocation that fails
        c.debug_fail_alloc(0)
        with pytest.raises(cpt.CptError) as ei:
            c.render(64, 64, 1, max_depth=8, seed=4)
        assert ei.value.code == cpt.CPT_E_HIP
        with pytest.raises(cpt.CptError) as ei:  # the film that held the 32x32 sums is gone
            c.render(32, 32, 2, max_depth=8, seed=5, flags=cpt.CPT_RENDER_ACCUMULATE)
        assert ei.value.code == cpt.CPT_E_INVALID
        c.render(64, 64, 1, max_depth=8, seed=5)
        _, a = c.read_film(sums=True)
    with cpt.Context(0) as fresh:
        fresh.upload(scene)
        fresh.render(64, 64, 1, max_depth=8, seed=5)
        _, b = fresh.read_film(sums=True)
    assert np.array_equal(a, b)
