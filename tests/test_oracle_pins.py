"""Pins of the CPU oracle against known answers, closed forms and brute force.

Each test pins one oracle function (oracle/oracle.cpp) independently of how the
oracle computes it.
"""
import math

import numpy as np
import pytest

import cpt
import oracle


# ---------------------------------------------------------------- random numbers
def _pcg32_ref(state, inc):
    """PCG32 XSH-RR, pcg32_random_r (O'Neill, pcg-random.org), pure Python."""
    M = 6364136223846793005
    mask = (1 << 64) - 1
    old = state
    state = (old * M + inc) & mask
    xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
    rot = old >> 59
    return state, ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF


def _pcg32_srandom(initstate, initseq):
    inc = ((initseq << 1) | 1) & ((1 << 64) - 1)
    state = 0
    state, _ = _pcg32_ref(state, inc)
    state = (state + initstate) & ((1 << 64) - 1)
    state, _ = _pcg32_ref(state, inc)
    return state, inc


def test_pcg32_python_reference_known_answer():
    # pcg32-global-demo output for pcg32_srandom(42, 54)
    state, inc = _pcg32_srandom(42, 54)
    out = []
    for _ in range(6):
        state, v = _pcg32_ref(state, inc)
        out.append(v)
    assert out == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def _splitmix64(x):
    mask = (1 << 64) - 1
    z = (x + 0x9E3779B97F4A7C15) & mask
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
    return z ^ (z >> 31)


@pytest.mark.parametrize("seed,path", [(1, 0), (1, 12345), (7, 2**39 + 5), (0xDEADBEEF, 99)])
def test_oracle_pcg_stream(seed, path):
    state, inc = _pcg32_srandom(_splitmix64(_splitmix64(seed) ^ path), 0x5851F42D4C957F2D)
    want = []
    for _ in range(16):
        state, v = _pcg32_ref(state, inc)
        want.append(np.float32(v >> 8) * np.float32(2.0**-24))
    got = oracle.pcg_floats(seed, path, 16)
    assert np.array_equal(got, np.array(want, np.float32))


def test_oracle_mt19937_known_answer():
    # std::mt19937 default seed 5489: first output 3499211612, 10000th 4123659995
    # (C++ standard [rand.predef]); uniform_real_distribution<float> divides the
    # float-converted 32-bit draw by 2^32 (one draw per float).
    f = oracle.mt_floats(5489, 10000)
    assert f[0] == np.float32(3499211612) * np.float32(2.0**-32)
    assert f[9999] == np.float32(4123659995) * np.float32(2.0**-32)


# ---------------------------------------------------------------- closest hit
@pytest.fixture(scope="module")
def small_terrain():
    return cpt.scenes.terrain_scene(n_spheres=6, grid=(48, 24), subdiv=2)


@pytest.mark.parametrize("which", ["box", "terrain"])
def test_reference_bvh_equals_brute_force(which, small_terrain):
    s = cpt.scenes.box_room() if which == "box" else small_terrain
    o = oracle.OracleScene(s)
    rays = cpt.raybatch.make_rays(s, 20000, seed=11)
    tr, tt = o.trace_closest(rays)
    tb, tbt = o.trace_closest(rays, brute=True)
    bad = (tr != tb) | ((tr >= 0) & (tt.view(np.int32) != tbt.view(np.int32)))
    # the reference's slab test can drop a box for rays lying in a box face plane;
    # such rays are rare and only among the adversarial 5%
    assert bad.mean() < 1e-3
    assert not bad[: len(rays) // 2].any()  # camera rays always agree
    sh = cpt.raybatch.make_rays(s, 20000, seed=12, shadow=True)
    assert (o.trace_any(sh) != o.trace_any(sh, brute=True)).mean() < 1e-3


def test_brute_force_closest_hit_by_hand():
    # a unit right triangle in z=0 and a ray straight down at (0.25, 0.25)
    s = cpt.scenes.Scene()
    m = s.add_material(cpt.scenes.material(cpt.scenes.MAT_DIFFUSE, albedo=(0.5, 0.5, 0.5)))
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, -1], [1, 0, -1], [0, 1, -1]], np.float32)
    N = np.tile([0, 0, 1], (6, 1))
    s.add_mesh(P, N, np.tile([1, 0, 0], (6, 1)), np.array([[0, 1, 2], [3, 4, 5]]), m)
    o = oracle.OracleScene(s)
    rays = np.array([[0.25, 0.25, 2.0, 0, 0, -1, np.inf],    # hits tri 0 at t=2
                     [0.25, 0.25, -0.5, 0, 0, -1, np.inf],   # hits tri 1 at t=0.5
                     [0.75, 0.75, 2.0, 0, 0, -1, np.inf],    # outside both
                     [0.25, 0.25, 0.00005, 0, 0, -1, np.inf],  # tri 0 at t < EPS: skipped -> tri 1
                     [0.25, 0.25, 2.0, 0, 0, -1, 1.5]],     # tmax before tri 0
                    np.float32)
    tri, t = o.trace_closest(rays, brute=True)
    assert tri.tolist() == [0, 1, -1, 1, -1]
    assert t[0] == 2.0 and t[1] == 0.5
    assert abs(t[3] - 1.00005) < 1e-6
    tri2, _ = o.trace_closest(rays)
    assert tri2.tolist() == tri.tolist()


def test_tie_rank_product_equals_oracle(small_terrain):
    for s in (cpt.scenes.box_room(), small_terrain):
        a = s.c_arrays()
        assert np.array_equal(cpt.reference_tie_rank(a["positions"], a["indices"]),
                              oracle.OracleScene(s).rank())


def test_tie_goes_to_highest_rank():
    # two coincident triangles: every hit ties; the reference keeps the one it tests last
    s = cpt.scenes.Scene()
    m = s.add_material(cpt.scenes.material(cpt.scenes.MAT_DIFFUSE))
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]] * 2, np.float32)
    s.add_mesh(P, np.tile([0, 0, 1], (6, 1)), np.tile([1, 0, 0], (6, 1)), np.array([[0, 1, 2], [3, 4, 5]]), m)
    o = oracle.OracleScene(s)
    rank = o.rank()
    rays = np.array([[0.2, 0.2, 1, 0, 0, -1, np.inf]], np.float32)
    assert o.trace_closest(rays)[0][0] == int(np.argmax(rank))
    assert o.trace_closest(rays, brute=True)[0][0] == int(np.argmax(rank))


# ---------------------------------------------------------------- BxDFs
def test_lambertian_sample_closed_form():
    rng = np.random.default_rng(0)
    n = np.array([0.0, 0.6, 0.8], np.float32)
    wo = np.array([0.0, 0.0, 1.0], np.float32)
    cosines = []
    for _ in range(4000):
        u1, u2 = rng.random(2).astype(np.float32)
        wi, pdf = oracle.diffuse_sample(n, wo, u1, u2)
        c = float(np.dot(wi, n))
        assert c >= -1e-6
        assert abs(pdf - c / math.pi) < 1e-5          # pdf = cos/pi
        assert abs(c - math.sqrt(1 - u2)) < 1e-5      # cos(theta) = sqrt(1-u2) by construction
        assert abs(np.linalg.norm(wi) - 1) < 1e-5
        cosines.append(c)
    assert abs(np.mean(cosines) - 2 / 3) < 0.01       # E[cos] under cos/pi = 2/3
    # normal flipped toward wo
    wi, _ = oracle.diffuse_sample(n, -wo, 0.3, 0.7)
    assert np.dot(wi, n) < 0


@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.5])
def test_ggx_d_normalisation(alpha):
    # integral of D(wh) cos(theta_h) over the hemisphere = 1
    nt, npf = 2000, 64
    th = (np.arange(nt) + 0.5) / nt * (math.pi / 2)
    ph = (np.arange(npf) + 0.5) / npf * 2 * math.pi
    total = 0.0
    for t in th:
        row = 0.0
        for p in ph[:8]:  # isotropic: D independent of phi
            wh = np.array([math.sin(t) * math.cos(p), math.sin(t) * math.sin(p), math.cos(t)], np.float32)
            row += oracle.ggx_d(alpha, wh)
        total += row / 8 * math.cos(t) * math.sin(t)
    total *= (math.pi / 2 / nt) * 2 * math.pi
    assert abs(total - 1.0) < 2e-3


def test_ggx_sample_normal_incidence_closed_form():
    # cos(theta_o) > .9999: slopes (r cos phi, r sin phi), r = sqrt(u1/(1-u1))
    # (microfacet_distribution.h:18-24).  At normal incidence sin(theta) = 0 and the
    # reference's cosPhi and sinPhi both return 1 (math.h:719-729), so the
    # "rotation" maps (x, y) -> (x - y, x + y) before unstretching.
    a = 0.3
    for u1, u2 in [(0.1, 0.2), (0.5, 0.9), (0.8, 0.4)]:
        r = math.sqrt(u1 / (1 - u1))
        phi = 2 * math.pi * u2
        sx, sy = r * math.cos(phi), r * math.sin(phi)
        want = np.array([-a * (sx - sy), -a * (sx + sy), 1.0])
        want /= np.linalg.norm(want)
        got = oracle.ggx_sample(a, np.array([0, 0, 1], np.float32), u1, u2)
        assert np.allclose(got, want, atol=2e-6)


# ---------------------------------------------------------------- lights
def test_triangle_sample_uniform_and_area():
    s = cpt.scenes.box_room()
    o = oracle.OracleScene(s)
    a = s.c_arrays()
    tri = int(a["light_tris"][0])
    A, B, C = (a["positions"][a["indices"][tri, k]].astype(np.float64) for k in range(3))
    rng = np.random.default_rng(1)
    pts = []
    for _ in range(20000):
        u1, u2 = rng.random(2).astype(np.float32)
        p, nrm, area = o.triangle_sample(tri, u1, u2)
        pts.append(p)
    pts = np.array(pts)
    assert abs(area - 0.5 * np.linalg.norm(np.cross(B - A, C - A))) < 1e-6
    assert np.allclose(pts.mean(0), (A + B + C) / 3, atol=3e-3)  # uniform density -> centroid
    # second moment of a uniform triangle: E[x x^T] = (A A^T + ... + (A+B+C)(A+B+C)^T)/12
    S = (np.outer(A, A) + np.outer(B, B) + np.outer(C, C) + np.outer(A + B + C, A + B + C)) / 12
    assert np.allclose(pts.T @ pts / len(pts), S, atol=5e-3)


# ---------------------------------------------------------------- integrator closed forms
def test_furnace_emissive_box_exact():
    le = (2.0, 3.0, 5.0)
    s = cpt.scenes.emissive_box(res=(16, 16), le=le)
    _, film, _ = oracle.OracleScene(s).render(16, 16, 2, sampler="pcg")
    assert np.array_equal(film, np.broadcast_to(np.array(le, np.float32), film.shape))


def _polygon_irradiance(poly, p, n):
    """Lambert's polygon formula: irradiance at p (normal n) from a unit-radiance polygon."""
    E = 0.0
    m = len(poly)
    for i in range(m):
        a = poly[i] - p
        b = poly[(i + 1) % m] - p
        a /= np.linalg.norm(a)
        b /= np.linalg.norm(b)
        c = np.cross(a, b)
        theta = math.acos(np.clip(np.dot(a, b), -1, 1))
        E += theta * np.dot(n, c / np.linalg.norm(c))
    return abs(E) / 2


def test_one_bounce_nee_mis_matches_lambert_polygon():
    size, height, albedo = 0.5, 1.0, 0.5
    s = cpt.scenes.single_quad_light(res=(8, 8), size=size, height=height, albedo=albedo)
    sums, film, counts = oracle.OracleScene(s).render(8, 8, 32, sampler="pcg", seed=3)
    h = size / 2
    poly = [np.array(v, float) for v in [(-h, height, -h), (h, height, -h), (h, height, h), (-h, height, h)]]
    want = albedo / math.pi * _polygon_irradiance(poly, np.zeros(3), np.array([0.0, 1.0, 0.0]))
    got = film[..., 0].mean()
    # 64 px x 1024 spp; per-path std is below the mean for this geometry
    assert abs(got - want) / want < 0.01, (got, want)


@pytest.mark.parametrize("integrator,spp_axis,tol", [(1, 32, 0.01), (2, 64, 0.03)])
def test_one_bounce_sibling_integrators_match_lambert_polygon(integrator, spp_axis, tol):
    """The same closed form through the reference's other two integrators: NEE
    without MIS (pathtracer.h: only the light sample counts after a diffuse
    bounce) and pure path tracing (pure_pt.h: only BSDF rays that hit the
    emitter count; noisier, so more samples and a 3% bound, about 4 sigma)."""
    size, height, albedo = 0.5, 1.0, 0.5
    s = cpt.scenes.single_quad_light(res=(8, 8), size=size, height=height, albedo=albedo)
    sums, film, counts = oracle.OracleScene(s).render(8, 8, spp_axis, sampler="pcg", seed=3, integrator=integrator)
    h = size / 2
    poly = [np.array(v, float) for v in [(-h, height, -h), (h, height, -h), (h, height, h), (-h, height, h)]]
    want = albedo / math.pi * _polygon_irradiance(poly, np.zeros(3), np.array([0.0, 1.0, 0.0]))
    got = sums[..., 0].sum() / (64 * spp_axis ** 2)
    if integrator == 2:
        assert counts[2] == 0  # no shadow rays
    assert abs(got - want) / want < tol, (got, want)


# ---------------------------------------------------------------- image textures
def _textured_quad(seed=0):
    """One quad z=0 over [-1,1]^2 facing +z, uv = ((x+1)*0.75, (y+1)*0.75) in [0, 1.5];
    albedo 5x3 RGB (sRGB), roughness 4x7 gray, metal 6x2 gray, normal map 3x5 RGB."""
    rng = np.random.default_rng(seed)
    s = cpt.scenes.Scene()
    imgs = dict(alb=rng.integers(0, 256, (3, 5, 3), dtype=np.uint8),
                rough=rng.integers(0, 256, (7, 4), dtype=np.uint8),
                metal=np.array([[0, 255, 20, 26, 255, 0], [255, 0, 25, 0, 30, 255]], np.uint8),
                nmap=rng.integers(0, 256, (5, 3, 3), dtype=np.uint8))
    ids = dict(alb=s.add_image(imgs["alb"], False, True), rough=s.add_image(imgs["rough"], True, False),
               metal=s.add_image(imgs["metal"], True, False), nmap=s.add_image(imgs["nmap"], False, False))
    m = s.add_material(cpt.scenes.material(cpt.scenes.MAT_DIFFUSE, albedo=(0.5, 0.5, 0.5), rough=0.2),
                       tex=(ids["alb"], ids["rough"], ids["metal"], ids["nmap"]))
    P = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
    N = np.tile(np.float32([0, 0, 1]), (4, 1))
    T = np.tile(np.float32([1, 0, 0]), (4, 1))
    F = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    s.add_mesh(P, N, T, F, m, uvs=(P[:, :2] + 1.0) * 0.75)
    s.set_camera((0, 0, 3), (0, 0, 0), (0, 1, 0), 0.5, 8, 8)
    return s, imgs, ids


def test_texture_lookup_closed_form():
    """ImageTexture::sample, point filter with the REPEAT border (image_texture.h:24-28,
    rgb_image.h:211-299), restated independently: for a ray hitting the quad at
    (x, y), u = (x+1)*0.75, v = (y+1)*0.75, texel (floor(u*w) mod w, floor((1-v)*h)
    mod h), row-major from the top.  Points within 1e-3 texel of a texel edge are
    skipped (the barycentric uv rounds differently from the affine one there)."""
    s, imgs, ids = _textured_quad()
    o = oracle.OracleScene(s)
    rng = np.random.default_rng(1)
    xy = rng.uniform(-0.999, 0.999, (4000, 2))
    rays = np.zeros((len(xy), 7), np.float32)
    rays[:, :2] = xy
    rays[:, 2] = 1.0
    rays[:, 5] = -1.0
    rays[:, 6] = 10.0
    got = o.hit_material(rays)
    assert (got[:, 0] >= 0).all()
    u = (rays[:, 0].astype(np.float64) + 1.0) * 0.75
    v = (rays[:, 1].astype(np.float64) + 1.0) * 0.75
    assert np.abs(got[:, 8] - u).max() < 1e-6 and np.abs(got[:, 9] - v).max() < 1e-6

    def texel(name, srgb, bw):
        px = imgs[name]
        h, w = px.shape[:2]
        fx, fy = u * w, (1.0 - v) * h
        ok = (np.abs(fx - np.round(fx)) > 1e-3) & (np.abs(fy - np.round(fy)) > 1e-3)
        i = np.floor(fx).astype(int) % w
        j = np.floor(fy).astype(int) % h
        lin = cpt.scenes.image_texels(px, bw, srgb)
        return lin[j, i], ok
    alb, ok_a = texel("alb", True, False)
    rough, ok_r = texel("rough", False, True)
    metal, ok_m = texel("metal", False, True)
    nmap, ok_n = texel("nmap", False, False)
    assert ok_a.mean() > 0.9
    assert np.array_equal(got[ok_a, 1:4], alb[ok_a])
    assert np.array_equal(got[ok_r, 4], rough[ok_r, 0])
    assert np.array_equal(got[ok_m, 0], np.where(metal[ok_m, 0] < 0.1, 0.0, 1.0))
    assert np.array_equal(got[ok_n, 5:8], nmap[ok_n])
    # both lobes occur (texels 0, 20, 25 -> diffuse; 26, 30, 255 -> GGX)
    assert set(got[:, 0].tolist()) == {0.0, 1.0}


def test_texture_repeat_border_quirk():
    """Image<T>::handle_border, REPEAT (rgb_image.h:216-258): i >= w wraps to i % w,
    i < 0 maps to i % w + w (C remainder), which is w itself for a negative
    multiple of w: get_pixel then reads pixels[w + j*w], the first texel of the
    next row."""
    s, imgs, ids = _textured_quad()
    o = oracle.OracleScene(s)
    lin = cpt.scenes.image_texels(imgs["alb"], False, True)  # (3, 5, 3)
    h, w = lin.shape[:2]
    assert np.array_equal(o.image_texel(ids["alb"], 7, 1), lin[1, 2])
    assert np.array_equal(o.image_texel(ids["alb"], -1, 0), lin[0, w - 1])
    assert np.array_equal(o.image_texel(ids["alb"], -7, -2), lin[1, 3])
    assert np.array_equal(o.image_texel(ids["alb"], -w, 0), lin[1, 0])   # the quirk
    assert np.array_equal(o.image_texel(ids["alb"], 2, 3 * h + 1), lin[1, 2])
    # past the last texel the reference reads beyond the image (undefined); both
    # this restatement and the device take the last texel
    assert np.array_equal(o.image_texel(ids["alb"], 2, -h), lin[h - 1, w - 1])
    assert np.array_equal(o.image_texel(ids["alb"], -w, h - 1), lin[h - 1, w - 1])


# ---------------------------------------------------------------- point lights
def _floor_under_point_light(res, pos, colour, intensity, albedo=0.5):
    s = cpt.scenes.Scene()
    m = s.add_material(cpt.scenes.material(cpt.scenes.MAT_DIFFUSE, albedo=(albedo,) * 3))
    P = np.array([[-4, 0, 4], [4, 0, 4], [4, 0, -4], [-4, 0, -4]], np.float32)
    N = np.tile(np.float32([0, 1, 0]), (4, 1))
    T = np.tile(np.float32([1, 0, 0]), (4, 1))
    s.add_mesh(P, N, T, np.array([[0, 1, 2], [0, 2, 3]], np.int32), m)
    s.add_point_light(pos, colour, intensity)
    s.set_camera((0.0, 3.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, -1.0), 0.3, *res)
    return s


def _floor_pixel_integral(s, W, H, n, radiance_at):
    """Pixel averages of radiance_at(points (..., 3)) over the floor y = 0 seen by the
    camera (camera_perspective.h:30-39), on an n x n midpoint grid per pixel."""
    cam = s.c_arrays()["camera"].astype(np.float64)
    o, X, Y, Z, tf, asp = cam[0:3], cam[3:6], cam[6:9], cam[9:12], cam[12], cam[13]
    sub = (np.arange(n) + 0.5) / n
    out = np.zeros((H, W) + np.shape(radiance_at(np.zeros((1, 3))))[1:])
    for j in range(H):
        for i in range(W):
            qx, qy = np.meshgrid((i + sub) / W, (j + sub) / H)
            cx, cy = 2 * qx - 1, -(2 * qy - 1)
            d = (cx * tf * asp)[..., None] * X + (cy * tf)[..., None] * Y - Z
            d /= np.linalg.norm(d, axis=-1, keepdims=True)
            p = o + d * (-o[1] / d[..., 1])[..., None]
            out[j, i] = np.mean(radiance_at(p.reshape(-1, 3)), axis=0)
    return out


def _point_light_floor(pos, albedo, mis):
    def f(p):
        sv = np.asarray(pos) - p
        r = np.linalg.norm(sv, axis=-1)
        sv /= r[..., None]
        cos = sv[..., 1]
        w = 1.0
        if mis:
            p_nee = r * r / (np.abs(sv[..., 0]) + 1e-4)
            w = p_nee / (p_nee + cos / math.pi)
        return albedo / math.pi * cos / (r * r) * w
    return f


@pytest.mark.parametrize("integrator", [0, 1])
def test_point_light_direct_lighting_closed_form(integrator):
    """A Lambertian floor lit by one point light, nothing else to hit: every path's
    value is the NEE term alone, albedo/pi cos(theta) * radiance / r^2 * w with the
    radiance colour * power / (4 pi) (point_light.h:20-28).  PathtracerMIS weighs it
    with w = p_nee / (p_nee + p_bsdf), p_nee = r^2 / (|s.x| + EPS) because the
    LightSample's normal is (1,0,0) (pathtracerMIS.h:95-100), p_bsdf = cos/pi;
    Pathtracer (NEE without MIS) uses w = 1.  The film is compared with the pixel
    integral of that closed form, evaluated on a 32x32 midpoint grid per pixel."""
    W = H = 24
    pos, col, power, albedo = (0.35, 0.5, -0.2), (1.0, 0.5, 0.25), 2.0, 0.5
    s = _floor_under_point_light((W, H), pos, col, power, albedo)
    _, film, _ = oracle.OracleScene(s).render(W, H, 32, max_depth=4, sampler="pcg", seed=5, integrator=integrator)
    want = _floor_pixel_integral(s, W, H, 32, _point_light_floor(pos, albedo, integrator == 0))
    for c in range(3):
        rad = col[c] * power / (4 * math.pi)
        rel = np.abs(film[..., c] - want * rad) / (want * rad)
        # jitter noise of 32x32 strata per pixel: measured 1.8e-4 max, 4e-5 mean
        assert rel.max() < 1e-3 and rel.mean() < 2e-4, (c, rel.max(), rel.mean())
    if integrator == 0:  # the weight is far from 1 on this geometry
        unweighted = _floor_pixel_integral(s, W, H, 32, _point_light_floor(pos, albedo, False))
        assert (1.0 - want / unweighted).max() > 0.2
        # the corrected mode weighs a point light 1 (no BSDF sample can reach it)
        _, fc, _ = oracle.OracleScene(s).render(W, H, 32, max_depth=4, sampler="pcg", seed=5, integrator=0,
                                                corrected=True)
        for c in range(3):
            rad = col[c] * power / (4 * math.pi)
            rel = np.abs(fc[..., c] - unweighted * rad) / (unweighted * rad)
            assert rel.max() < 1e-3 and rel.mean() < 2e-4


def _blocks(x, b=4):
    H, W = x.shape[:2]
    return x.reshape(H // b, b, W // b, b, *x.shape[2:]).mean(axis=(1, 3))


def test_light_tracer_point_light_closed_form():
    """Lighttracer (lighttracer.h:27-110) on the floor under a point light: the
    splatted image is an unbiased estimate of the same closed form (the pixel
    integral of albedo/pi cos(theta) * radiance / r^2).  It pins the light start
    (PointLight::sample: sample_sphere's acos leaves [-1, 1] for draws above 1/pi,
    which the 1/(4 pi^2) pdf exactly compensates), the camera connection
    (camera_perspective.h:46-82: pixel and importance factor 1/(4 cos^3 d^2 tan^2)
    for a square image) and the film scale.  Measured over three seeds: image
    mean within 0.07%, 4x4-pixel blocks within 1.6%."""
    W = H = 24
    pos, col, power, albedo = (0.35, 0.5, -0.2), (1.0, 0.5, 0.25), 2.0, 0.5
    s = _floor_under_point_light((W, H), pos, col, power, albedo)
    sums, counts = oracle.OracleScene(s).render_lt(W, H, 256, max_depth=4, seed=3)
    img = sums / 256 ** 2
    want = _floor_pixel_integral(s, W, H, 16, _point_light_floor(pos, albedo, False))
    want = want[..., None] * (np.array(col) * power / (4 * math.pi))
    assert abs(img.mean() / want.mean() - 1) < 3e-3
    assert np.abs(_blocks(img) / _blocks(want) - 1).max() < 0.04
    assert counts[0] == W * H * 256 ** 2


@pytest.mark.parametrize("corrected", [False, True])
def test_light_tracer_aspect_ratio(corrected):
    """A 2:1 image: the reference's importance factor leaves out the aspect ratio
    (camera_perspective.h:75-79, "todo: account for x"), so its light-traced image
    is `aspect` times the closed form; the corrected mode divides it out."""
    W, H = 32, 16
    pos, col, power, albedo = (0.35, 0.5, -0.2), (1.0, 0.5, 0.25), 2.0, 0.5
    s = _floor_under_point_light((W, H), pos, col, power, albedo)
    sums, _ = oracle.OracleScene(s).render_lt(W, H, 256, max_depth=4, seed=3, corrected=corrected)
    img = sums[..., 0] / 256 ** 2
    want = _floor_pixel_integral(s, W, H, 8, _point_light_floor(pos, albedo, False)) * col[0] * power / (4 * math.pi)
    ratio = img.mean() / want.mean()
    assert abs(ratio - (1.0 if corrected else W / H)) < 0.01 * (W / H)


def test_light_tracer_area_light_matches_lambert_polygon():
    """The light tracer under a square emitter (ShapeLight::sample, shape_light.h:50-68:
    a uniform point of an emitter triangle, EmissiveBxDF::emit_sample's cosine
    direction on a random side, pdf 0.5 cos/pi / area / lights): the image equals,
    per 4x4 block, albedo/pi * Le * Lambert's polygon irradiance at the floor points
    each pixel sees."""
    W = H = 16
    size, height, albedo = 0.5, 1.0, 0.5
    s = cpt.scenes.single_quad_light(res=(W, H), size=size, height=height, albedo=albedo)
    s.set_camera((0.0, 0.5, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, 1.0), 0.4, W, H)
    sums, _ = oracle.OracleScene(s).render_lt(W, H, 384, max_depth=3, seed=7)
    img = sums[..., 0] / 384 ** 2
    h = size / 2
    poly = [np.array(v, float) for v in [(-h, height, -h), (h, height, -h), (h, height, h), (-h, height, h)]]

    def lit(p):
        return np.array([albedo / math.pi * _polygon_irradiance(poly, q, np.array([0.0, 1.0, 0.0])) for q in p])
    want = _floor_pixel_integral(s, W, H, 4, lit)
    print(f"light tracer vs polygon formula: mean {img.mean() / want.mean() - 1:+.4f}, "
          f"worst block {np.abs(_blocks(img) / _blocks(want) - 1).max():.4f}")
    assert abs(img.mean() / want.mean() - 1) < 5e-3
    assert np.abs(_blocks(img) / _blocks(want) - 1).max() < 0.03


def test_ray_in_a_triangle_plane_misses():
    """A ray lying in a triangle's plane has determinant 0 and t = 0 * inf = NaN:
    the oracle (like the library) reports a miss, and the reference BVH agrees
    with brute force on coincident triangles hit by such rays."""
    s = cpt.scenes.Scene()
    m = s.add_material(cpt.scenes.material(cpt.scenes.MAT_DIFFUSE, albedo=(0.5,) * 3))
    P = np.tile(np.array([[-0.5, -0.5, 0], [0.5, -0.5, 0], [0, 0.5, 0]], np.float32), (40, 1))
    s.add_mesh(P, np.tile(np.float32([0, 0, 1]), (120, 1)), np.tile(np.float32([1, 0, 0]), (120, 1)),
               np.arange(120, dtype=np.int32).reshape(-1, 3), m)
    o = oracle.OracleScene(s)
    rays = np.array([[-2.0, 0.0, 0.0, 1.0, 0.0, 0.0, np.inf],    # in the plane, through the triangles
                     [0.0, -2.0, 0.0, 0.0, 1.0, 0.0, np.inf],
                     [0.1, 0.0, 2.0, 0.0, 0.0, -1.0, np.inf]], np.float32)  # a proper hit
    tri, t = o.trace_closest(rays)
    btri, _ = o.trace_closest(rays, brute=True)
    assert tri[0] == -1 and tri[1] == -1
    assert tri[2] == 39 and abs(t[2] - 2.0) < 1e-6   # the last of 40 coincident triangles
    assert np.array_equal(tri, btri)
