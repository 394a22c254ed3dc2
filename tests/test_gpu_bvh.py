"""Device BVH build (PLOC clustering of Morton-sorted triangles collapsed to the
4-wide layout; CPT_SCENE_GPU_BVH), replacing the reference's BVH::build
(src/primitive/bvh.h:151-217): closest hits, any hits and whole renders are
identical to those over the host-built BVH, whose parity with the oracle the
closest-hit tests establish."""
import numpy as np
import pytest

import cpt
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mid_terrain():
    return cpt.scenes.terrain_scene(n_spheres=40, grid=(256, 128), subdiv=3)


def _hits(ctx, s, gpu_bvh, n=1 << 16, seed=6):
    ctx.upload(s, gpu_bvh=gpu_bvh)
    rays = cpt.raybatch.make_rays(s, n, seed=seed)
    srays = cpt.raybatch.make_rays(s, n // 2, seed=seed + 1, shadow=True)
    tri, t = ctx.trace_closest(rays)
    return tri, t, ctx.trace_any(srays), ctx.bvh_info()


@pytest.mark.parametrize("which", ["box", "terrain"])
def test_gpu_bvh_hits_equal_host_bvh(ctx, which, mid_terrain):
    s = cpt.scenes.box_room() if which == "box" else mid_terrain
    print(f"device BVH build: {which}, {s.n_triangles} triangles", flush=True)
    h_tri, h_t, h_any, hi = _hits(ctx, s, False)
    g_tri, g_t, g_any, gi = _hits(ctx, s, True)
    print(f"host build {hi['build_ms']:.1f} ms, {hi['nodes']} nodes, depth {hi['depth']}; "
          f"device build {gi['build_ms']:.1f} ms, {gi['nodes']} nodes, depth {gi['depth']}")
    assert gi["gpu_built"] and not hi["gpu_built"]
    assert gi["triangles"] == s.n_triangles and gi["nodes"] <= s.n_triangles
    assert np.array_equal(g_tri, h_tri)
    assert np.array_equal(g_t.view(np.int32), h_t.view(np.int32))
    assert np.array_equal(g_any, h_any)
    assert (h_tri >= 0).mean() > 0.3


def test_gpu_bvh_render_bit_identical(ctx):
    s = cpt.scenes.terrain_scene(res=(48, 32), n_spheres=30, grid=(64, 32), subdiv=2)
    films = []
    for gpu_bvh in (False, True):
        ctx.upload(s, gpu_bvh=gpu_bvh)
        ctx.render(48, 32, 2, max_depth=12, seed=3)
        films.append(ctx.read_film(sums=True)[1])
    assert np.array_equal(films[0], films[1])


def _soup(tris_xyz, res=(16, 16)):
    s = cpt.scenes.Scene()
    m = s.add_material(cpt.scenes.material(cpt.scenes.MAT_DIFFUSE, albedo=(0.5, 0.5, 0.5)))
    P = np.asarray(tris_xyz, np.float32).reshape(-1, 3)
    n = len(P) // 3
    N = np.tile(np.float32([0, 0, 1]), (len(P), 1))
    T = np.tile(np.float32([1, 0, 0]), (len(P), 1))
    s.add_mesh(P, N, T, np.arange(3 * n, dtype=np.int32).reshape(-1, 3), m)
    s.set_camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.5, *res)
    return s


@pytest.mark.parametrize("case", ["one", "two", "coincident", "line"])
def test_gpu_bvh_small_and_degenerate_inputs(ctx, case):
    """One and two triangles (the root is itself a leaf), 300 coincident triangles
    (every Morton key equal, every pair the same merged area: ties decide every
    clustering round), 500 triangles along a line (centroid extent in one axis)."""
    base = np.array([[-0.5, -0.5, 0], [0.5, -0.5, 0], [0, 0.5, 0]], np.float32)
    if case == "one":
        tris = base
    elif case == "two":
        tris = np.concatenate([base, base + [0.2, 0.1, 0.3]])
    elif case == "coincident":
        tris = np.tile(base, (300, 1))
    else:
        tris = np.concatenate([base * 0.01 + [x, 0, 0] for x in np.linspace(-1, 1, 500)])
    s = _soup(tris)
    print(f"degenerate case {case}: {s.n_triangles} triangles", flush=True)
    rays = cpt.raybatch.make_rays(s, 4096, seed=8)
    out = []
    for gpu_bvh in (False, True):
        ctx.upload(s, gpu_bvh=gpu_bvh)
        out.append(ctx.trace_closest(rays))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1].view(np.int32), out[1][1].view(np.int32))
    assert (out[0][0] >= 0).any()
    # and the reference's answer: among coincident triangles the last it tests
    r_tri, r_t = oracle.OracleScene(s).trace_closest(rays)
    assert np.array_equal(out[1][0], r_tri)
    hit = r_tri >= 0
    assert np.array_equal(out[1][1][hit].view(np.int32), r_t[hit].view(np.int32))
